// fb_filter.hip -- filter_active_reads (src/graph/local_multik.rs:259-333) on
// the device.
//
// 1. The boundary set: canonical ntHash (k' <= 31) of the selected windows of
//    the unitig sequences, in an open-addressing u64 set in HBM (linear
//    probing from splitmix64(h); all-ones marks an empty slot and a separate
//    flag holds that one value when it is a member).
// 2. One thread per merged read or per pair hashes the windows at offsets 0,
//    (len-k')/2 and len-k' straight from the packed words and probes the set.
// 3. The kept reads are compacted into a new device read list in the
//    BitSeqArena::push_seq layout (reads contiguous inside each arena, arenas
//    word-aligned): a scan gives each kept read its place, and one thread per
//    output word assembles its 32 bases from the one or more source reads
//    that cover it, so no two threads write one word.
#include <algorithm>

#include "fb_device.hpp"
#include "fb_internal.hpp"

namespace fb {

using fbk::NtHash;
using fbk::splitmix64;

static __device__ __forceinline__ uint32_t ascii_code(uint8_t c) {
  // ASCII_TO_2BIT (bitseq.rs:67-81): A/C/G/T in either case, anything else A.
  switch (c) {
    case 'C': case 'c': return 1u;
    case 'G': case 'g': return 2u;
    case 'T': case 't': return 3u;
    default: return 0u;
  }
}

static __device__ __forceinline__ uint64_t set_slot(uint64_t h, uint32_t log2cap) {
  return splitmix64(h) >> (64u - log2cap);
}

static __device__ __forceinline__ bool set_has(const FaArgs& a, uint64_t h) {
  if (h == kEmpty) return a.set_flag[0] != 0u;
  const uint64_t mask = (1ull << a.log2cap) - 1ull;
  uint64_t s = set_slot(h, a.log2cap);
  for (;;) {
    const uint64_t v = a.set[s];
    if (v == h) return true;
    if (v == kEmpty) return false;
    s = (s + 1ull) & mask;
  }
}

__global__ void k_fa_set_clear(FaArgs a) {
  const uint64_t cap = 1ull << a.log2cap;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x)
    a.set[i] = kEmpty;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.set_flag[0] = 0u;
}

// Window w of the boundary list: range r holds windows rpre[r] .. rpre[r+1]-1,
// the first starting at useq[rstart[r]].
__global__ void k_fa_set_insert(FaArgs a) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.n_windows) return;
  uint32_t lo = 0, hi = a.n_ranges;  // largest r with rpre[r] <= w
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.rpre[mid] <= w) lo = mid;
    else hi = mid;
  }
  const uint8_t* p = a.useq + a.rstart[lo] + (w - a.rpre[lo]);
  NtHash h{0, 0};
  for (uint32_t i = 0; i < a.k; ++i) h.push(ascii_code(p[i]), a.k);
  const uint64_t key = h.canonical();
  if (key == kEmpty) {
    a.set_flag[0] = 1u;
    return;
  }
  const uint64_t mask = (1ull << a.log2cap) - 1ull;
  uint64_t s = set_slot(key, a.log2cap);
  for (;;) {
    const unsigned long long prev =
        atomicCAS((unsigned long long*)&a.set[s], (unsigned long long)kEmpty, (unsigned long long)key);
    if (prev == kEmpty || prev == key) return;
    s = (s + 1ull) & mask;
  }
}

// Canonical ntHash of the k' (<= 31) bases at global base g of the packed list.
static __device__ __forceinline__ uint64_t window_hash(const uint64_t* __restrict__ data, uint64_t g, uint32_t k) {
  const uint64_t w0 = data[g >> 5], w1 = data[(g >> 5) + 1];
  const uint32_t sh = (uint32_t)(g & 31u) << 1;
  const uint64_t bits = sh ? (w0 >> sh) | (w1 << (64u - sh)) : w0;
  NtHash h{0, 0};
  for (uint32_t i = 0; i < k; ++i) h.push((uint32_t)(bits >> (2u * i)) & 3u, k);
  return h.canonical();
}

// is_active (local_multik.rs:297-309).
static __device__ bool read_active(const FaArgs& a, uint64_t r) {
  const uint32_t len = a.reads.lens[r];
  if (len < a.k) return false;
  const uint64_t off = a.reads.offs[r];
  const uint32_t o[3] = {0u, (len - a.k) / 2u, len - a.k};
  for (int i = 0; i < 3; ++i)
    if (set_has(a, window_hash(a.reads.data, off + o[i], a.k))) return true;
  return false;
}

// keep[t]: t < n_merged is merged read t; otherwise pair t - n_merged (either
// mate active).  keep[n_merged + n_pairs] is left 0 for the scan's total.
__global__ void k_fa_keep(FaArgs a) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n = a.n_merged + a.n_pairs;
  if (t > n) return;
  if (t == n) {
    a.keep[t] = 0u;
    return;
  }
  bool on;
  if (t < a.n_merged) {
    on = read_active(a, a.first[0] + t);
  } else {
    const uint64_t p = t - a.n_merged;
    on = read_active(a, a.first[1] + p) || read_active(a, a.first[2] + p);
  }
  a.keep[t] = on ? 1u : 0u;
}

// Source index and length of every kept read, in output order: kept merged
// reads, then the kept pairs' forward mates, then their reverse mates.  The
// kept counts come from the scan (pos[n_merged], pos[n_merged + n_pairs]).
__global__ void k_fa_list(FaArgs a) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n = a.n_merged + a.n_pairs;
  if (t >= n || !a.keep[t]) return;
  const uint32_t nkm = a.pos[a.n_merged];
  const uint32_t nkp = a.pos[n] - nkm;
  if (t < a.n_merged) {
    const uint32_t j = a.pos[t];
    const uint64_t r = a.first[0] + t;
    a.src[j] = (uint32_t)r;
    a.olen[j] = a.reads.lens[r];
  } else {
    const uint64_t p = t - a.n_merged;
    const uint32_t j = a.pos[t] - nkm;
    const uint64_t r1 = a.first[1] + p, r2 = a.first[2] + p;
    a.src[nkm + j] = (uint32_t)r1;
    a.olen[nkm + j] = a.reads.lens[r1];
    a.src[nkm + nkp + j] = (uint32_t)r2;
    a.olen[nkm + nkp + j] = a.reads.lens[r2];
  }
}

static __device__ __forceinline__ uint32_t out_arena(const uint64_t* first, uint64_t j) {
  return j < first[1] ? 0u : (j < first[2] ? 1u : 2u);
}

// Output offsets (global base index of the new list, arenas word-aligned),
// lengths and the length histogram (counted per block in LDS for lengths
// below kLdsHist: the lengths of a read set take few values, and one global
// counter per length would serialise every thread on it).
constexpr uint32_t kLdsHist = 4096;
__global__ __launch_bounds__(256) void k_fa_offsets(FaArgs a) {
  __shared__ uint32_t h[kLdsHist];
  const uint32_t nh = a.hist_n < kLdsHist ? a.hist_n : kLdsHist;
  for (uint32_t i = threadIdx.x; i < nh; i += blockDim.x) h[i] = 0u;
  __syncthreads();
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < a.rfirst[3]) {
    const uint32_t ar = out_arena(a.rfirst, j);
    a.ooffs[j] = a.opos[j] - a.opos[a.rfirst[ar]] + a.word0[ar] * 32ull;
    const uint32_t len = a.olen[j];
    a.olens[j] = len;
    const uint32_t bin = len < a.hist_n ? len : a.hist_n - 1u;
    if (bin < nh) atomicAdd(&h[bin], 1u);
    else atomicAdd(&a.hist[bin], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nh; i += blockDim.x)
    if (h[i]) atomicAdd(&a.hist[i], h[i]);
}

// Output word w: bases 32w .. 32w+31 of the new list, taken from the kept
// reads of w's arena that cover them (several when reads are short).
__global__ void k_fa_gather(FaArgs a) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.word0[3]) return;
  const uint32_t ar = w < a.word0[1] ? 0u : (w < a.word0[2] ? 1u : 2u);
  uint64_t lo = a.rfirst[ar], hi = a.rfirst[ar + 1];  // a word of an arena has at least one read
  const uint64_t g0 = w * 32ull;
  while (hi - lo > 1) {  // largest j with ooffs[j] <= g0
    const uint64_t mid = (lo + hi) >> 1;
    if (a.ooffs[mid] <= g0) lo = mid;
    else hi = mid;
  }
  uint64_t out = 0;
  uint32_t filled = 0;
  uint64_t pos = g0;
  for (uint64_t j = lo; j < a.rfirst[ar + 1] && filled < 32u; ++j) {
    const uint64_t start = a.ooffs[j], end = start + a.olens[j];
    if (pos < start) pos = start;  // reads are contiguous: only the first iteration can start inside a read
    if (pos >= end) continue;
    const uint64_t rem = end - pos;
    const uint32_t take = rem < (uint64_t)(32u - filled) ? (uint32_t)rem : 32u - filled;
    const uint64_t g = a.reads.offs[a.src[j]] + (pos - start);
    const uint64_t w0 = a.reads.data[g >> 5], w1 = a.reads.data[(g >> 5) + 1];
    const uint32_t sh = (uint32_t)(g & 31u) << 1;
    uint64_t bits = sh ? (w0 >> sh) | (w1 << (64u - sh)) : w0;
    if (take < 32u) bits &= (1ull << (2u * take)) - 1ull;
    out |= bits << (2u * filled);
    filled += take;
    pos += take;
  }
  a.odata[w] = out;
}

static uint32_t fa_grid(uint64_t n, uint32_t block) {
  const uint64_t g = (n + block - 1) / block;
  return (uint32_t)(g ? g : 1);
}

hipError_t launch_fa_set_clear(const FaArgs& a, hipStream_t s) {
  const uint64_t cap = 1ull << a.log2cap;
  const uint32_t g = (uint32_t)std::min<uint64_t>(fa_grid(cap, 256), 65536);
  hipLaunchKernelGGL(k_fa_set_clear, dim3(g), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fa_set_insert(const FaArgs& a, hipStream_t s) {
  if (a.n_windows == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fa_set_insert, dim3(fa_grid(a.n_windows, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fa_keep(const FaArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_fa_keep, dim3(fa_grid(a.n_merged + a.n_pairs + 1, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fa_list(const FaArgs& a, hipStream_t s) {
  if (a.n_merged + a.n_pairs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fa_list, dim3(fa_grid(a.n_merged + a.n_pairs, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fa_offsets(const FaArgs& a, hipStream_t s) {
  if (a.rfirst[3] == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fa_offsets, dim3(fa_grid(a.rfirst[3], 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fa_gather(const FaArgs& a, hipStream_t s) {
  if (a.word0[3] == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fa_gather, dim3(fa_grid(a.word0[3], 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace fb
