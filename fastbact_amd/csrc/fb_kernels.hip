// fb_kernels.hip -- HIP kernels of the k-mer spectrum path (gfx950).
//
// Pass 1  (src/graph/build.rs:93-138): canonical k-mer -> to_u64 fold -> 4-row
//         Count-Min increments.  Counts are exact u32 cells; the reference's
//         saturating nibble / u16 counters are min(count, cap) of them, which
//         is order-independent (every update is a non-negative add).
// Pass 2  (build.rs:140-289): adjacent windows whose CMS estimates are both
//         >= t_solid add (canonical, edge) records to an open-addressing table.
// Pass 3  (build.rs:291-328): edge mask, cov = max(min(cov,65535)/2, 1), compaction.
// Estimator (src/ingest/estimator.rs:296-379) in first-insertion-time form.
#include <hipcub/hipcub.hpp>

#include "fb_device.hpp"
#include "fb_internal.hpp"
#include "fb_table.hpp"

namespace fb {
using namespace fbk;

constexpr int kBlock = 256;

static inline unsigned grid_for(uint64_t threads) {
  uint64_t g = (threads + kBlock - 1) / kBlock;
  return (unsigned)(g ? g : 1);
}

// Splits the `units` of read r evenly over `split` threads; returns false when
// this thread's share is empty.
__device__ __forceinline__ bool read_share(const DevReads& R, uint64_t gid, uint32_t split, uint32_t k,
                                           uint32_t extra, uint64_t& r, uint32_t& u0, uint32_t& u1) {
  r = gid / split;
  uint32_t part = (uint32_t)(gid - r * split);
  if (r >= R.n_reads) return false;
  uint32_t len = R.lens[r];
  if (len < k + extra) return false;
  uint32_t units = len - k + 1 - extra;
  u0 = (uint32_t)((uint64_t)part * units / split);
  u1 = (uint32_t)((uint64_t)(part + 1) * units / split);
  return u0 < u1;
}

// ---------------------------------------------------------------------------
// Pass 1
template <int L, bool CMS, bool HLL>
__device__ __forceinline__ void pass1_body(const Pass1Args& a, uint64_t gid) {
  uint64_t r;
  uint32_t w0, w1;
  const uint32_t k = a.k;
  if (!read_share(a.reads, gid, a.split, k, 0, r, w0, w1)) return;
  const uint64_t* __restrict__ data = a.reads.data;
  const uint64_t g = a.reads.offs[r] + w0;
  Kmer<L> win = window_bits<L>(data, g, k);
  Kmer<L> fwd = fwd_from_window<L>(win, k);
  Kmer<L> rev = rc_from_window<L>(win, k);
  BaseStream in;
  in.init(data, g + k);
  NtHash nh;
  BaseStream outs;
  if (HLL) {
    nh.f = 0;
    nh.r = 0;
    BaseStream s0;
    s0.init(data, g);
    for (uint32_t j = 0; j < k; ++j) nh.push(s0.next(), k);
    outs.init(data, g);
  }
  const uint32_t lw = a.cms.log2w;
  const uint64_t wmask = (1ull << lw) - 1ull;
  uint32_t* __restrict__ counts = a.cms.counts;
  for (uint32_t w = w0;;) {
    if (CMS) {
      const uint64_t key = kmer_lt<L>(rev, fwd) ? kmer_fold<L>(rev) : kmer_fold<L>(fwd);
#pragma unroll
      for (int row = 0; row < 4; ++row) {
        uint64_t col = splitmix64(key ^ cms_seed(row)) & wmask;
        atomicAdd(counts + (((uint64_t)row << lw) | col), a.weight);
      }
    }
    if (HLL) {
      uint32_t idx, rho;
      hll_index_rho(splitmix64(nh.canonical()), a.hll.p, idx, rho);
      if (rho > a.hll.regs[idx]) atomicMax(a.hll.regs + idx, rho);
    }
    if (++w >= w1) break;
    const uint32_t b = in.next();
    if (CMS) kmer_roll<L>(fwd, rev, b, k);
    if (HLL) nh.roll(outs.next(), b, k);
  }
}

template <int L, bool CMS, bool HLL>
__global__ __launch_bounds__(kBlock) void k_pass1(Pass1Args a) {
  pass1_body<L, CMS, HLL>(a, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// One launch over many isolates: block b belongs to isolate find_iso(b).
template <int L, bool CMS, bool HLL>
__global__ __launch_bounds__(kBlock) void k_pass1_m(const Pass1Args* __restrict__ args, const uint32_t* __restrict__ blk,
                                                    uint32_t n_iso) {
  const uint32_t iso = find_iso(blk, n_iso, blockIdx.x);
  pass1_body<L, CMS, HLL>(args[iso], (uint64_t)(blockIdx.x - blk[iso]) * blockDim.x + threadIdx.x);
}

hipError_t launch_pass1(const Pass1Args& a, hipStream_t s) {
  if (a.reads.n_reads == 0) return hipSuccess;
  const uint32_t L = (2 * a.k + 63) / 64;
  unsigned grid = grid_for(a.reads.n_reads * (uint64_t)a.split);
  const bool h = a.hll.regs != nullptr;
  const bool c = a.cms.counts != nullptr;
  if (!h && !c) return hipSuccess;
#define P1(LL)                                                                          \
  if (h && c) hipLaunchKernelGGL((k_pass1<LL, true, true>), dim3(grid), dim3(kBlock), 0, s, a); \
  else if (c) hipLaunchKernelGGL((k_pass1<LL, true, false>), dim3(grid), dim3(kBlock), 0, s, a); \
  else hipLaunchKernelGGL((k_pass1<LL, false, true>), dim3(grid), dim3(kBlock), 0, s, a)
  switch (L) {
    case 1: P1(1); break;
    case 2: P1(2); break;
    case 3: P1(3); break;
    default: P1(4); break;
  }
#undef P1
  return hipGetLastError();
}

hipError_t launch_pass1_multi(const MultiLaunch& m, uint32_t k, bool cms, bool hll, hipStream_t s) {
  if (!m.total_blocks) return hipSuccess;
  const uint32_t L = (2 * k + 63) / 64;
  const Pass1Args* a = static_cast<const Pass1Args*>(m.args);
#define P1M(LL)                                                                                                       \
  if (hll && cms) hipLaunchKernelGGL((k_pass1_m<LL, true, true>), dim3(m.total_blocks), dim3(kBlock), 0, s, a, m.blk, m.n_iso); \
  else if (cms) hipLaunchKernelGGL((k_pass1_m<LL, true, false>), dim3(m.total_blocks), dim3(kBlock), 0, s, a, m.blk, m.n_iso); \
  else hipLaunchKernelGGL((k_pass1_m<LL, false, true>), dim3(m.total_blocks), dim3(kBlock), 0, s, a, m.blk, m.n_iso)
  switch (L) {
    case 1: P1M(1); break;
    case 2: P1M(2); break;
    case 3: P1M(3); break;
    default: P1M(4); break;
  }
#undef P1M
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Threshold bitmap and reference-layout packing

__global__ __launch_bounds__(kBlock) void k_solid_bits(const uint32_t* __restrict__ counts, uint64_t n,
                                                       uint32_t t, uint32_t cap, uint32_t* __restrict__ bits) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool on = false;
  if (i < n) {
    uint32_t c = counts[i];
    c = c < cap ? c : cap;
    on = c >= t;
  }
  uint64_t m = __ballot(on);
  uint32_t lane = threadIdx.x & 63u;
  uint64_t wbase = (i - lane) >> 5;  // first u32 word of this wave
  uint64_t nwords = (n + 31) >> 5;
  if (lane == 0 && wbase < nwords) bits[wbase] = (uint32_t)m;
  if (lane == 0 && wbase + 1 < nwords) bits[wbase + 1] = (uint32_t)(m >> 32);
}

hipError_t launch_solid_bits(const uint32_t* counts, uint64_t n, uint32_t t, uint32_t cap, uint32_t* bits,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_solid_bits, dim3(grid_for(n)), dim3(kBlock), 0, s, counts, n, t, cap, bits);
  return hipGetLastError();
}

// Popcount of n words, one atomic per block.
__global__ __launch_bounds__(kBlock) void k_popcount(const uint32_t* __restrict__ w, uint64_t n,
                                                     unsigned long long* __restrict__ out) {
  __shared__ unsigned long long part[kBlock / 64];
  unsigned long long c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    c += (unsigned long long)__builtin_popcount(w[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < kBlock / 64; ++i) t += part[i];
    if (t) atomicAdd(out, t);
  }
}
// Histogram of min(count, 255) over n counters, added to hist[256].
__global__ __launch_bounds__(kBlock) void k_count_hist(const uint32_t* __restrict__ counts, uint64_t n,
                                                       uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) h[i] = 0;
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t c = counts[i];
    atomicAdd(h + (c < 255u ? c : 255u), 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x)
    if (h[i]) atomicAdd(hist + i, h[i]);
}
hipError_t launch_count_hist(const uint32_t* counts, uint64_t n, uint32_t* hist, hipStream_t s) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g > 1024) g = 1024;
  if (g == 0) g = 1;
  hipLaunchKernelGGL(k_count_hist, dim3((unsigned)g), dim3(kBlock), 0, s, counts, n, hist);
  return hipGetLastError();
}

hipError_t launch_popcount(const uint32_t* words, uint64_t n, unsigned long long* out, hipStream_t s) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g > 1024) g = 1024;
  if (g == 0) g = 1;
  hipLaunchKernelGGL(k_popcount, dim3((unsigned)g), dim3(kBlock), 0, s, words, n, out);
  return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_pack_nibble(const uint32_t* __restrict__ counts, uint64_t nbytes,
                                                        uint8_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbytes) return;
  uint32_t lo = counts[2 * i], hi = counts[2 * i + 1];
  lo = lo < 15u ? lo : 15u;
  hi = hi < 15u ? hi : 15u;
  out[i] = (uint8_t)(lo | (hi << 4));
}
hipError_t launch_pack_nibble(const uint32_t* counts, uint64_t n, uint8_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_nibble, dim3(grid_for(n / 2)), dim3(kBlock), 0, s, counts, n / 2, out);
  return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_pack_u16(const uint32_t* __restrict__ counts, uint64_t n,
                                                     uint16_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t c = counts[i];
  out[i] = (uint16_t)(c < 65535u ? c : 65535u);
}
hipError_t launch_pack_u16(const uint32_t* counts, uint64_t n, uint16_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_u16, dim3(grid_for(n)), dim3(kBlock), 0, s, counts, n, out);
  return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_pack_u8(const uint32_t* __restrict__ regs, uint64_t n,
                                                    uint8_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint8_t)regs[i];
}
hipError_t launch_pack_u8(const uint32_t* regs, uint64_t n, uint8_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_u8, dim3(grid_for(n)), dim3(kBlock), 0, s, regs, n, out);
  return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g == 0) g = 1;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)g), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}

// Plain copy of n 16-byte words: reads and writes 16 n bytes, nothing else.
__global__ __launch_bounds__(kBlock) void k_copy_calibration(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                             uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}
hipError_t launch_copy_calibration(const void* src, void* dst, uint64_t n16, hipStream_t s) {
  uint64_t g = (n16 + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g == 0) g = 1;
  hipLaunchKernelGGL(k_copy_calibration, dim3((unsigned)g), dim3(kBlock), 0, s, static_cast<const uint4*>(src),
                     static_cast<uint4*>(dst), n16);
  return hipGetLastError();
}

// Random u32 atomicAdd over a table of 2^log2n counters: update i goes to
// splitmix64(i ^ seed) mod 2^log2n, with (ret) or without a returned value.
// It measures the device's atomic rate at a given table size (L2-resident,
// infinity-cache-resident or HBM-resident), the figure the Count-Min and
// bin-cursor atomics of the passes are compared with.
// MODE 0: atomicAdd, 1: atomicAdd with its returned value used, 2: plain
// 4-byte loads at the same cells (the random-read rate), 3: plain 4-byte
// stores (the random-write rate).
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_atomic_rate(uint32_t* __restrict__ table, uint32_t log2n,
                                                        uint64_t n_updates, uint64_t seed, uint32_t* __restrict__ sink) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t mask = (1ull << log2n) - 1ull;
  uint32_t acc = 0;
  for (; i < n_updates; i += stride) {
    const uint64_t c = splitmix64(i ^ seed) & mask;
    if (MODE == 1) acc += atomicAdd(table + c, 1u);
    else if (MODE == 2) acc += table[c];
    else if (MODE == 3) table[c] = (uint32_t)i;
    else atomicAdd(table + c, 1u);
  }
  if ((MODE == 1 || MODE == 2) && acc == 0xFFFFFFFFu) sink[0] = acc;  // keeps the returned values live
}
hipError_t launch_atomic_rate(uint32_t* table, uint32_t log2n, uint64_t n_updates, int mode, uint32_t* sink,
                              hipStream_t s) {
  const uint64_t g = 256u * 64u;  // 64 blocks of 256 per CU: enough updates in flight
  const uint64_t seed = 0x6a09e667f3bcc909ull;
  if (mode == 1)
    hipLaunchKernelGGL(k_atomic_rate<1>, dim3((unsigned)g), dim3(kBlock), 0, s, table, log2n, n_updates, seed, sink);
  else if (mode == 2)
    hipLaunchKernelGGL(k_atomic_rate<2>, dim3((unsigned)g), dim3(kBlock), 0, s, table, log2n, n_updates, seed, sink);
  else if (mode == 3)
    hipLaunchKernelGGL(k_atomic_rate<3>, dim3((unsigned)g), dim3(kBlock), 0, s, table, log2n, n_updates, seed, sink);
  else
    hipLaunchKernelGGL(k_atomic_rate<0>, dim3((unsigned)g), dim3(kBlock), 0, s, table, log2n, n_updates, seed, sink);
  return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_table_clear(uint64_t* __restrict__ slots, uint64_t cap,
                                                        uint32_t stride, uint32_t kw) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t n = cap * stride;
  uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += step) {
    uint32_t w = (uint32_t)(i % stride);
    slots[i] = w < kw ? kEmpty : 0ull;
  }
}
hipError_t launch_table_clear(const DevSolidTable& t, hipStream_t s) {
  uint64_t g = (t.cap * t.stride + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_table_clear, dim3((unsigned)g), dim3(kBlock), 0, s, t.slots, t.cap, t.stride, t.kw);
  if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
  return launch_fill_u32(t.status, 2, 0, s);
}

// ---------------------------------------------------------------------------
// Solid table

// Packs the 2k-bit k-mer integer into KW words of 62 bits (bits 62j..62j+61
// in word j).  Two spare bits per word keep the all-ones empty marker out of
// reach of any key.
template <int L, int KW>
__device__ __forceinline__ void pack62(const Kmer<L>& x, uint64_t (&w)[KW]) {
#pragma unroll
  for (int j = 0; j < KW; ++j) {
    const int bit = 62 * j;
    const int limb = bit / 64, off = bit % 64;
    uint64_t v = limb < L ? (x.a[limb] >> off) : 0ull;
    if (off > 2 && limb + 1 < L) v |= x.a[limb + 1] << (64 - off);
    w[j] = v & 0x3fffffffffffffffull;
  }
}

__device__ __forceinline__ bool solid_lookup(const uint32_t* __restrict__ bits, uint32_t lw, uint64_t key) {
  const uint64_t wmask = (1ull << lw) - 1ull;
  bool ok = true;
#pragma unroll
  for (int row = 0; row < 4; ++row) {
    uint64_t i = ((uint64_t)row << lw) | (splitmix64(key ^ cms_seed(row)) & wmask);
    ok &= (bits[i >> 5] >> (i & 31u)) & 1u;
  }
  return ok;
}

__device__ __forceinline__ uint64_t slot_hash(uint64_t fold) { return splitmix64(fold ^ 0x9e3779b97f4a7c15ull); }

// One record batch for window x: cov += n*cov_w; edge counters below the edge
// gate get +edge_w (contig weights are always applied).
__device__ __forceinline__ void add_records(uint32_t* c, uint32_t nrec, bool left, uint32_t e_left, bool right,
                                            uint32_t e_right, uint32_t cov_w, uint32_t edge_w, uint32_t gate,
                                            bool always) {
  atomicAdd(c, nrec * cov_w);
  if (left && (always || c[2 + e_left] < gate)) atomicAdd(c + 2 + e_left, edge_w);
  if (right && (always || c[2 + e_right] < gate)) atomicAdd(c + 2 + e_right, edge_w);
}

template <int L, int KW>
__device__ __forceinline__ void pass2_body(const Pass2Args& a, uint64_t gid) {
  uint64_t r;
  uint32_t p0, p1;  // pairs [p0, p1): pair x joins windows x and x+1
  const uint32_t k = a.k;
  if (!read_share(a.reads, gid, a.split, k, 1, r, p0, p1)) return;
  const uint64_t* __restrict__ data = a.reads.data;
  const uint64_t g = a.reads.offs[r] + p0;
  const bool contig = a.contig_mode != 0;
  const uint32_t cov_w = contig ? 100u : 1u, edge_w = contig ? 1000u : 1u;
  Kmer<L> win = window_bits<L>(data, g, k);
  Kmer<L> fwd = fwd_from_window<L>(win, k);
  Kmer<L> rev = rc_from_window<L>(win, k);
  BaseStream in, first;
  in.init(data, g + k);
  first.init(data, g);
  bool cur_rc = kmer_lt<L>(rev, fwd);
  Kmer<L> cur = cur_rc ? rev : fwd;
  uint64_t cur_fold = kmer_fold<L>(cur);
  bool cur_solid = contig || solid_lookup(a.solid_bits, a.log2w, cur_fold);
  bool left = false;       // pair (x-1, x) is solid
  uint32_t left_rcp = 0;   // complement of the first base of window x-1
  for (uint32_t x = p0;; ++x) {
    bool right = false;
    uint32_t b = 0;
    bool nxt_rc = false, nxt_solid = false;
    Kmer<L> nxt;
    uint64_t nxt_fold = 0;
    if (x < p1) {
      b = in.next();
      kmer_roll<L>(fwd, rev, b, k);
      nxt_rc = kmer_lt<L>(rev, fwd);
      nxt = nxt_rc ? rev : fwd;
      nxt_fold = kmer_fold<L>(nxt);
      nxt_solid = contig || solid_lookup(a.solid_bits, a.log2w, nxt_fold);
      right = cur_solid && nxt_solid;
    }
    const uint32_t first_base = first.next();
    const uint32_t nrec = (uint32_t)left + (uint32_t)right;
    if (nrec) {
      uint64_t kwds[KW];
      pack62<L, KW>(cur, kwds);
      uint32_t* c = table_find_global<KW>(a.table, kwds, slot_hash(cur_fold));
      if (c) {
        // edge_v of pair (x-1, x) and edge_u of pair (x, x+1), build.rs:268-275
        const uint32_t e_left = cur_rc ? left_rcp : left_rcp + 4u;
        const uint32_t e_right = cur_rc ? b + 4u : b;
        add_records(c, nrec, left, e_left, right, e_right, cov_w, edge_w, a.min_trans_cov, contig);
      }
    }
    if (x >= p1) break;
    left = right;
    left_rcp = 3u - first_base;
    cur = nxt;
    cur_rc = nxt_rc;
    cur_fold = nxt_fold;
    cur_solid = nxt_solid;
  }
}

template <int L, int KW>
__global__ __launch_bounds__(kBlock) void k_pass2(Pass2Args a) {
  pass2_body<L, KW>(a, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

template <int L, int KW>
__global__ __launch_bounds__(kBlock) void k_pass2_m(const Pass2Args* __restrict__ args, const uint32_t* __restrict__ blk,
                                                    uint32_t n_iso) {
  const uint32_t iso = find_iso(blk, n_iso, blockIdx.x);
  pass2_body<L, KW>(args[iso], (uint64_t)(blockIdx.x - blk[iso]) * blockDim.x + threadIdx.x);
}

static uint32_t key_words(uint32_t k) { return (k + 30) / 31; }

hipError_t launch_pass2(const Pass2Args& a, hipStream_t s) {
  if (a.reads.n_reads == 0) return hipSuccess;
  const uint32_t L = (2 * a.k + 63) / 64;
  const uint32_t KW = key_words(a.k);
  if (KW != a.table.kw) return hipErrorInvalidValue;
  unsigned grid = grid_for(a.reads.n_reads * (uint64_t)a.split);
#define P2(LL, KK) hipLaunchKernelGGL((k_pass2<LL, KK>), dim3(grid), dim3(kBlock), 0, s, a)
  if (L == 1 && KW == 1) P2(1, 1);
  else if (L == 1) P2(1, 2);
  else if (L == 2 && KW == 2) P2(2, 2);
  else if (L == 2) P2(2, 3);
  else if (L == 3 && KW == 3) P2(3, 3);
  else if (L == 3) P2(3, 4);
  else if (KW == 4) P2(4, 4);
  else P2(4, 5);
#undef P2
  return hipGetLastError();
}

hipError_t launch_pass2_multi(const MultiLaunch& m, uint32_t k, hipStream_t s) {
  if (!m.total_blocks) return hipSuccess;
  const uint32_t L = (2 * k + 63) / 64;
  const uint32_t KW = key_words(k);
  const Pass2Args* a = static_cast<const Pass2Args*>(m.args);
#define P2M(LL, KK) hipLaunchKernelGGL((k_pass2_m<LL, KK>), dim3(m.total_blocks), dim3(kBlock), 0, s, a, m.blk, m.n_iso)
  if (L == 1 && KW == 1) P2M(1, 1);
  else if (L == 1) P2M(1, 2);
  else if (L == 2 && KW == 2) P2M(2, 2);
  else if (L == 2) P2M(2, 3);
  else if (L == 3 && KW == 3) P2M(3, 3);
  else if (L == 3) P2M(3, 4);
  else if (KW == 4) P2M(4, 4);
  else P2M(4, 5);
#undef P2M
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pass 3.  Three launches so no counter is contended: per-block counts of
// kept slots over contiguous slot ranges, one small exclusive scan, then the
// write pass (records leave in slot order).

constexpr uint32_t kP3Blocks = 2048;

template <int KW>
__device__ __forceinline__ bool slot_record(const uint64_t* __restrict__ slots, uint64_t i, uint32_t stride,
                                            uint32_t gate, uint64_t (&wk)[KW], uint32_t& mask, uint32_t& cov,
                                            uint64_t& raw_cov) {
  const uint64_t* s = slots + i * stride;
#pragma unroll
  for (int j = 0; j < KW; ++j) wk[j] = s[j];
  mask = 0;
  cov = 0;
  raw_cov = 0;
  if (wk[0] == kEmpty) return false;
  const uint32_t* c = reinterpret_cast<const uint32_t*>(s + KW);
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    uint32_t v = c[2 + e];
    v = v < 65535u ? v : 65535u;
    if (v >= gate) mask |= 1u << e;
  }
  raw_cov = c[0];
  const uint32_t cv = c[0] < 65535u ? c[0] : 65535u;
  cov = cv / 2u;
  if (cov < 1u) cov = 1u;
  return mask != 0u;
}

template <int KW>
__device__ __forceinline__ void pass3_count_body(const uint64_t* __restrict__ slots, uint64_t cap, uint32_t stride,
                                                 uint32_t gate, uint32_t* __restrict__ bcount,
                                                 unsigned long long* __restrict__ brec) {
  __shared__ uint32_t s_cnt[kBlock / 64];
  __shared__ unsigned long long s_rec[kBlock / 64];
  const uint64_t per = (cap + kP3Blocks - 1) / kP3Blocks;
  const uint64_t lo = (uint64_t)blockIdx.x * per, hi = lo + per < cap ? lo + per : cap;
  uint32_t n = 0;
  unsigned long long rec = 0;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    uint64_t wk[KW];
    uint32_t mask, cov;
    uint64_t raw;
    n += slot_record<KW>(slots, i, stride, gate, wk, mask, cov, raw) ? 1u : 0u;
    rec += raw;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    n += __shfl_xor(n, off);
    rec += __shfl_xor(rec, off);
  }
  if ((threadIdx.x & 63u) == 0) {
    s_cnt[threadIdx.x >> 6] = n;
    s_rec[threadIdx.x >> 6] = rec;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    unsigned long long r = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      t += s_cnt[w];
      r += s_rec[w];
    }
    bcount[blockIdx.x] = t;
    brec[blockIdx.x] = r;
  }
}

template <int KW>
__global__ __launch_bounds__(kBlock) void k_pass3_count(const uint64_t* __restrict__ slots, uint64_t cap,
                                                        uint32_t stride, uint32_t gate, uint32_t* __restrict__ bcount,
                                                        unsigned long long* __restrict__ brec) {
  pass3_count_body<KW>(slots, cap, stride, gate, bcount, brec);
}

// One block: exclusive scan of kP3Blocks counts in place; totals to out[0]
// (records kept) and rec_sum.
__device__ __forceinline__ void pass3_scan_body(uint32_t* __restrict__ bcount,
                                                const unsigned long long* __restrict__ brec,
                                                uint32_t* __restrict__ out_count,
                                                unsigned long long* __restrict__ rec_sum) {
  __shared__ uint32_t s[1024];
  __shared__ unsigned long long sr[1024];
  constexpr uint32_t per = kP3Blocks / 1024;
  uint32_t local[per];
  uint32_t tot = 0;
  unsigned long long rt = 0;
  for (uint32_t j = 0; j < per; ++j) {
    local[j] = bcount[threadIdx.x * per + j];
    tot += local[j];
    rt += brec[threadIdx.x * per + j];
  }
  s[threadIdx.x] = tot;
  sr[threadIdx.x] = rt;
  __syncthreads();
  // Hillis-Steele inclusive scan over 1024 partial sums
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    uint32_t v = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
    unsigned long long w = threadIdx.x >= off ? sr[threadIdx.x - off] : 0ull;
    __syncthreads();
    s[threadIdx.x] += v;
    sr[threadIdx.x] += w;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - tot;
  for (uint32_t j = 0; j < per; ++j) {
    bcount[threadIdx.x * per + j] = run;
    run += local[j];
  }
  if (threadIdx.x == 1023) {
    out_count[0] = s[1023];
    rec_sum[0] = sr[1023];
  }
}

__global__ __launch_bounds__(1024) void k_pass3_scan(uint32_t* __restrict__ bcount,
                                                     const unsigned long long* __restrict__ brec,
                                                     uint32_t* __restrict__ out_count,
                                                     unsigned long long* __restrict__ rec_sum) {
  pass3_scan_body(bcount, brec, out_count, rec_sum);
}

template <int KW>
__device__ __forceinline__ void pass3_write_body(const uint64_t* __restrict__ slots, uint64_t cap, uint32_t stride,
                                                 uint32_t nref, uint32_t gate, const uint32_t* __restrict__ boffset,
                                                 uint64_t* __restrict__ out_keys, uint16_t* __restrict__ out_cov,
                                                 uint8_t* __restrict__ out_mask, const uint32_t* __restrict__ out_base,
                                                 uint32_t out_cap) {
  __shared__ uint32_t s_w[kBlock / 64];
  const uint64_t per = (cap + kP3Blocks - 1) / kP3Blocks;
  const uint64_t lo = (uint64_t)blockIdx.x * per, hi = lo + per < cap ? lo + per : cap;
  uint32_t base = boffset[blockIdx.x] + (out_base ? *out_base : 0u);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  for (uint64_t c0 = lo; c0 < hi; c0 += blockDim.x) {
    const uint64_t i = c0 + threadIdx.x;
    uint64_t wk[KW];
    uint32_t mask = 0, cov = 0;
    uint64_t raw;
    const bool keep = i < hi && slot_record<KW>(slots, i, stride, gate, wk, mask, cov, raw);
    const uint64_t ballot = __ballot(keep);
    if (lane == 0) s_w[wave] = (uint32_t)__builtin_popcountll(ballot);
    __syncthreads();
    uint32_t before = 0, chunk = 0;
    for (uint32_t w = 0; w < kBlock / 64; ++w) {
      if (w < wave) before += s_w[w];
      chunk += s_w[w];
    }
    const uint32_t pos = base + before + (uint32_t)__builtin_popcountll(ballot & ((1ull << lane) - 1ull));
    if (keep && pos < out_cap) {
      uint64_t limb[4] = {0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < KW; ++j) {
        const int bit = 62 * j, li = bit / 64, off = bit % 64;
        if (li < 4) limb[li] |= wk[j] << off;
        if (off > 2 && li + 1 < 4) limb[li + 1] |= wk[j] >> (64 - off);
      }
      for (uint32_t w = 0; w < nref; ++w) out_keys[(uint64_t)pos * nref + w] = limb[nref - 1 - w];
      out_cov[pos] = (uint16_t)cov;
      out_mask[pos] = (uint8_t)mask;
    }
    base += chunk;
    __syncthreads();
  }
}

template <int KW>
__global__ __launch_bounds__(kBlock) void k_pass3_write(const uint64_t* __restrict__ slots, uint64_t cap,
                                                        uint32_t stride, uint32_t nref, uint32_t gate,
                                                        const uint32_t* __restrict__ boffset,
                                                        uint64_t* __restrict__ out_keys, uint16_t* __restrict__ out_cov,
                                                        uint8_t* __restrict__ out_mask, const uint32_t* __restrict__ out_base,
                                                        uint32_t out_cap) {
  pass3_write_body<KW>(slots, cap, stride, nref, gate, boffset, out_keys, out_cov, out_mask, out_base, out_cap);
}

// Pass 3 over many tables: blockIdx.y selects the table.
__device__ __forceinline__ void p3_scratch(void* scratch, uint32_t*& out_count, unsigned long long*& rec_sum,
                                           unsigned long long*& brec, uint32_t*& bcount) {
  out_count = static_cast<uint32_t*>(scratch);
  rec_sum = reinterpret_cast<unsigned long long*>(out_count + 2);
  brec = reinterpret_cast<unsigned long long*>(out_count + 4);
  bcount = reinterpret_cast<uint32_t*>(brec + kP3Blocks);
}
template <int KW>
__global__ __launch_bounds__(kBlock) void k_pass3_count_m(const P3Args* __restrict__ args) {
  const P3Args& a = args[blockIdx.y];
  uint32_t *oc, *bc;
  unsigned long long *rs, *br;
  p3_scratch(a.scratch, oc, rs, br, bc);
  pass3_count_body<KW>(a.t.slots, a.t.cap, a.t.stride, a.gate, bc, br);
}
__global__ __launch_bounds__(1024) void k_pass3_scan_m(const P3Args* __restrict__ args) {
  const P3Args& a = args[blockIdx.y];
  uint32_t *oc, *bc;
  unsigned long long *rs, *br;
  p3_scratch(a.scratch, oc, rs, br, bc);
  pass3_scan_body(bc, br, oc, rs);
}
template <int KW>
__global__ __launch_bounds__(kBlock) void k_pass3_write_m(const P3Args* __restrict__ args) {
  const P3Args& a = args[blockIdx.y];
  uint32_t *oc, *bc;
  unsigned long long *rs, *br;
  p3_scratch(a.scratch, oc, rs, br, bc);
  pass3_write_body<KW>(a.t.slots, a.t.cap, a.t.stride, a.nref, a.gate, bc, a.out_keys, a.out_cov, a.out_mask,
                       a.out_base, a.out_cap);
}

hipError_t launch_pass3_multi(const P3Args* args, uint32_t n, uint32_t kw, hipStream_t s) {
  if (!n) return hipSuccess;
#define P3M(KK)                                                                                   \
  hipLaunchKernelGGL((k_pass3_count_m<KK>), dim3(kP3Blocks, n), dim3(kBlock), 0, s, args);        \
  hipLaunchKernelGGL(k_pass3_scan_m, dim3(1, n), dim3(1024), 0, s, args);                         \
  hipLaunchKernelGGL((k_pass3_write_m<KK>), dim3(kP3Blocks, n), dim3(kBlock), 0, s, args)
  switch (kw) {
    case 1: P3M(1); break;
    case 2: P3M(2); break;
    case 3: P3M(3); break;
    case 4: P3M(4); break;
    default: P3M(5); break;
  }
#undef P3M
  return hipGetLastError();
}

// Fills many u32 ranges in one launch: blockIdx.y selects the range.
__global__ __launch_bounds__(kBlock) void k_fill_ranges(const FillRange* __restrict__ r) {
  const FillRange f = r[blockIdx.y];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < f.n; i += (uint64_t)gridDim.x * blockDim.x)
    f.p[i] = f.v;
}
hipError_t launch_fill_ranges(const FillRange* r, uint32_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_fill_ranges, dim3(2048, n), dim3(kBlock), 0, s, r);
  return hipGetLastError();
}

// Clears many solid tables in one launch: blockIdx.y selects the table.
__global__ __launch_bounds__(kBlock) void k_table_clear_m(const DevSolidTable* __restrict__ tabs) {
  const DevSolidTable& t = tabs[blockIdx.y];
  const uint64_t n = t.cap * t.stride;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    t.slots[i] = (uint32_t)(i % t.stride) < t.kw ? kEmpty : 0ull;
  if (blockIdx.x == 0 && threadIdx.x < 2) t.status[threadIdx.x] = 0u;
}
hipError_t launch_table_clear_multi(const DevSolidTable* tabs, uint32_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_table_clear_m, dim3(4096, n), dim3(kBlock), 0, s, tabs);
  return hipGetLastError();
}

size_t pass3_scratch_bytes() { return 16 + (size_t)kP3Blocks * 12 + 64; }

hipError_t launch_pass3(const DevSolidTable& t, uint32_t k, uint32_t nref, uint32_t gate, uint64_t* out_keys,
                        uint16_t* out_cov, uint8_t* out_mask, void* scratch, hipStream_t s, const uint32_t* out_base,
                        uint32_t out_cap) {
  (void)k;
  uint32_t* out_count = static_cast<uint32_t*>(scratch);
  unsigned long long* rec_sum = reinterpret_cast<unsigned long long*>(out_count + 2);
  unsigned long long* brec = reinterpret_cast<unsigned long long*>(out_count + 4);
  uint32_t* bcount = reinterpret_cast<uint32_t*>(brec + kP3Blocks);
#define P3(KK)                                                                                                  \
  hipLaunchKernelGGL((k_pass3_count<KK>), dim3(kP3Blocks), dim3(kBlock), 0, s, t.slots, t.cap, t.stride, gate, \
                     bcount, brec);                                                                             \
  hipLaunchKernelGGL(k_pass3_scan, dim3(1), dim3(1024), 0, s, bcount, brec, out_count, rec_sum);               \
  hipLaunchKernelGGL((k_pass3_write<KK>), dim3(kP3Blocks), dim3(kBlock), 0, s, t.slots, t.cap, t.stride, nref,  \
                     gate, bcount, out_keys, out_cov, out_mask, out_base, out_cap)
  switch (t.kw) {
    case 1: P3(1); break;
    case 2: P3(2); break;
    case 3: P3(3); break;
    case 4: P3(4); break;
    default: P3(5); break;
  }
#undef P3
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sorting the solid set by k-mer: stable LSD passes over the significant
// 64-bit limbs with a permutation, then one gather.

__global__ __launch_bounds__(kBlock) void k_iota(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}
__global__ __launch_bounds__(kBlock) void k_gather_limb(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ perm,
                                                        uint32_t n, uint32_t nref, uint32_t word,
                                                        uint64_t* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = keys[(uint64_t)perm[i] * nref + word];
}
// A solid k-mer's record side by side: its nref key words, then
// cov | mask << 16, so that the gather after the sort reads one run of
// bytes per record instead of one access per array.
__global__ __launch_bounds__(kBlock) void k_pack_records(const uint64_t* __restrict__ keys,
                                                         const uint16_t* __restrict__ cov,
                                                         const uint8_t* __restrict__ mask, uint32_t n,
                                                         uint32_t nref, uint64_t* __restrict__ rec) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t* r = rec + (uint64_t)i * (nref + 1);
  for (uint32_t w = 0; w < nref; ++w) r[w] = keys[(uint64_t)i * nref + w];
  r[nref] = (uint64_t)cov[i] | (uint64_t)mask[i] << 16;
}
__global__ __launch_bounds__(kBlock) void k_gather_records(const uint64_t* __restrict__ rec,
                                                           const uint32_t* __restrict__ perm, uint32_t n,
                                                           uint32_t nref, uint64_t* __restrict__ okeys,
                                                           uint16_t* __restrict__ ocov, uint8_t* __restrict__ omask) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t* r = rec + (uint64_t)perm[i] * (nref + 1);
  for (uint32_t w = 0; w < nref; ++w) okeys[(uint64_t)i * nref + w] = r[w];
  const uint64_t cm = r[nref];
  ocov[i] = (uint16_t)cm;
  omask[i] = (uint8_t)(cm >> 16);
}

// One-word keys sort with their record as the value: cov | mask << 16.
__global__ __launch_bounds__(kBlock) void k_pack_cm(const uint16_t* __restrict__ cov, const uint8_t* __restrict__ mask,
                                                    uint32_t n, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint32_t)cov[i] | (uint32_t)mask[i] << 16;
}
__global__ __launch_bounds__(kBlock) void k_unpack_cm(const uint64_t* __restrict__ skeys,
                                                      const uint32_t* __restrict__ cm, uint32_t n,
                                                      uint64_t* __restrict__ keys, uint16_t* __restrict__ cov,
                                                      uint8_t* __restrict__ mask) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (skeys != keys) keys[i] = skeys[i];
  const uint32_t v = cm[i];
  cov[i] = (uint16_t)v;
  mask[i] = (uint8_t)(v >> 16);
}

// Scratch layout: [limb keys x2][perm x2][packed records][tie run][cub temp].
// The first sort of keys of more than 32 bases orders them by their first
// 16 bases (kTopSortBits bits, 32-bit sort keys: four radix passes);
// keys sharing those bases form runs for the fix below.
constexpr uint32_t kTopSortBits = 32;
// The top kTopSortBits bits of every key: bits [2k - 32, 2k) of the
// integer (words most significant first); 2k > 64.
__global__ __launch_bounds__(kBlock) void k_top_digit(const uint64_t* __restrict__ keys, uint32_t n, uint32_t nref,
                                                      uint32_t k, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t sh = 2 * k - kTopSortBits, a = sh >> 6, off = sh & 63u;  // word a from the least significant
  const uint64_t* key = keys + (uint64_t)i * nref;
  uint64_t v = key[nref - 1 - a] >> off;
  if (off > 64u - kTopSortBits) v |= key[nref - 2 - a] << (64 - off);  // the digit spans two words
  out[i] = (uint32_t)v;
}
// Runs of equal values in the sorted v: the first element of a run of two
// or more walks it, up to kTieRunMax + 1 long.
constexpr uint32_t kTieRunMax = 64;
__device__ __forceinline__ uint32_t tie_run(const uint32_t* __restrict__ v, uint32_t n, uint32_t i) {
  if (i + 1 >= n || v[i] != v[i + 1] || (i > 0 && v[i - 1] == v[i])) return 0u;
  uint32_t j = i + 2;
  while (j < n && j - i <= kTieRunMax && v[j] == v[i]) ++j;
  return j - i;
}
// *maxrun = the longest run (capped at kTieRunMax + 1; *maxrun zeroed before).
__global__ __launch_bounds__(kBlock) void k_tie_runs(const uint32_t* __restrict__ v, uint32_t n,
                                                     uint32_t* __restrict__ maxrun) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t m = tie_run(v, n, i);
  if (m) atomicMax(maxrun, m);
}
// Sorts the permutation within every run (at most kTieRunMax long) by the
// whole key, words most significant first.
__global__ __launch_bounds__(kBlock) void k_fix_ties(const uint32_t* __restrict__ v, uint32_t n,
                                                     const uint64_t* __restrict__ keys, uint32_t nref,
                                                     uint32_t* __restrict__ perm) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t m = tie_run(v, n, i);
  auto less = [&](uint32_t a, uint32_t b) {
    for (uint32_t w = 0; w < nref; ++w) {
      const uint64_t x = keys[(uint64_t)a * nref + w], y = keys[(uint64_t)b * nref + w];
      if (x != y) return x < y;
    }
    return false;
  };
  for (uint32_t x = i + 1; x < i + m; ++x) {
    const uint32_t p = perm[x];
    uint32_t y = x;
    for (; y > i && less(p, perm[y - 1]); --y) perm[y] = perm[y - 1];
    perm[y] = p;
  }
}

hipError_t sort_solid(uint64_t* keys, uint16_t* cov, uint8_t* mask, uint32_t n, uint32_t nref, uint32_t k, void* tmp,
                      size_t* tmp_bytes, hipStream_t s, bool top_first) {
  size_t cub_bytes = 0;
  hipcub::DoubleBuffer<uint64_t> dk(nullptr, nullptr);
  hipcub::DoubleBuffer<uint32_t> dv(nullptr, nullptr);
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, cub_bytes, dk, dv, (int)(n ? n : 1), 0, 64, s);
  if (e != hipSuccess) return e;
  {
    size_t digit_bytes = 0;
    hipcub::DoubleBuffer<uint32_t> dd(nullptr, nullptr);
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, digit_bytes, dd, dv, (int)(n ? n : 1), 0, (int)kTopSortBits, s);
    if (e != hipSuccess) return e;
    if (digit_bytes > cub_bytes) cub_bytes = digit_bytes;
  }
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t nn = n ? n : 1;
  const size_t need = al(nn * 8) * 2 + al(nn * 4) * 2 + al(nn * (nref + 1) * 8) + al(4) + al(cub_bytes);
  if (keys == nullptr) {
    *tmp_bytes = need;
    return hipSuccess;
  }
  if (*tmp_bytes < need) return hipErrorInvalidValue;
  if (n <= 1) return hipSuccess;
  char* p = (char*)tmp;
  uint64_t* k0 = (uint64_t*)p; p += al(nn * 8);
  uint64_t* k1 = (uint64_t*)p; p += al(nn * 8);
  uint32_t* v0 = (uint32_t*)p; p += al(nn * 4);
  uint32_t* v1 = (uint32_t*)p; p += al(nn * 4);
  uint64_t* rec = (uint64_t*)p; p += al(nn * (nref + 1) * 8);
  uint32_t* ties = (uint32_t*)p; p += al(4);
  void* ctmp = p;
  const unsigned grid = grid_for(n);
  // one-word keys (k <= 32) sort directly with their record as the value:
  // for them that is faster than a digit sort followed by a gather
  const bool digits = top_first && nref > 1;
  if (nref == 1) {
    // one word a key (k <= 32): its cov and mask travel with it as the
    // sort's value, so no permutation and no gather
    hipLaunchKernelGGL(k_pack_cm, dim3(grid), dim3(kBlock), 0, s, cov, mask, n, v0);
    hipcub::DoubleBuffer<uint64_t> bk(keys, k0);
    hipcub::DoubleBuffer<uint32_t> bv(v0, v1);
    size_t cb = cub_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(ctmp, cb, bk, bv, (int)n, 0, (int)(2 * k), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_unpack_cm, dim3(grid), dim3(kBlock), 0, s, bk.Current(), bv.Current(), n, keys, cov, mask);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_iota, dim3(grid), dim3(kBlock), 0, s, v0, n);
  // Significant limbs: L = ceil(2k/64); the top limb has 2k - 64(L-1) bits.
  const uint32_t L = (2 * k + 63) / 64;
  uint32_t* perm = v0;
  uint32_t* perm_alt = v1;
  uint32_t first = 0;  // limbs to sort, from the least significant
  if (digits) {
    // One sort by the top kTopSortBits bits gives the full order when no
    // two keys share them; short runs of keys that do are then sorted in
    // place, and the limb-by-limb sort runs only for a run longer than
    // kTieRunMax.
    uint32_t* d0 = reinterpret_cast<uint32_t*>(k0);
    hipLaunchKernelGGL(k_top_digit, dim3(grid), dim3(kBlock), 0, s, keys, n, nref, k, d0);
    hipcub::DoubleBuffer<uint32_t> bk(d0, reinterpret_cast<uint32_t*>(k1));
    hipcub::DoubleBuffer<uint32_t> bv(v0, v1);
    size_t cb = cub_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(ctmp, cb, bk, bv, (int)n, 0, (int)kTopSortBits, s);
    if (e != hipSuccess) return e;
    if ((e = launch_fill_u32(ties, 1, 0u, s))) return e;
    hipLaunchKernelGGL(k_tie_runs, dim3(grid), dim3(kBlock), 0, s, bk.Current(), n, ties);
    uint32_t longest = kTieRunMax + 1;
    if ((e = hipMemcpyAsync(&longest, ties, 4, hipMemcpyDeviceToHost, s))) return e;
    if ((e = hipStreamSynchronize(s))) return e;
    if (longest > kTieRunMax) {
      hipLaunchKernelGGL(k_iota, dim3(grid), dim3(kBlock), 0, s, v0, n);
    } else {
      perm = bv.Current();
      first = L;
      if (longest)
        hipLaunchKernelGGL(k_fix_ties, dim3(grid), dim3(kBlock), 0, s, bk.Current(), n, keys, nref, perm);
    }
  }
  for (uint32_t limb = first; limb < L; ++limb) {
    const uint32_t word = nref - 1 - limb;  // reference words are most significant first
    hipLaunchKernelGGL(k_gather_limb, dim3(grid), dim3(kBlock), 0, s, keys, perm, n, nref, word, k0);
    const int end_bit = (limb == L - 1) ? (int)(2 * k - 64 * (L - 1)) : 64;
    hipcub::DoubleBuffer<uint64_t> bk(k0, k1);
    hipcub::DoubleBuffer<uint32_t> bv(perm, perm_alt);
    size_t cb = cub_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(ctmp, cb, bk, bv, (int)n, 0, end_bit, s);
    if (e != hipSuccess) return e;
    if (bv.Current() != perm) {
      perm_alt = perm;
      perm = bv.Current();
    }
  }
  hipLaunchKernelGGL(k_pack_records, dim3(grid), dim3(kBlock), 0, s, keys, cov, mask, n, nref, rec);
  hipLaunchKernelGGL(k_gather_records, dim3(grid), dim3(kBlock), 0, s, rec, perm, n, nref, keys, cov, mask);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Explicit-key sketch operations (CountMinSketch::add / estimate, HyperLogLog::add_hash)

__global__ __launch_bounds__(kBlock) void k_cms_add_keys(DevCms c, const uint64_t* __restrict__ keys, uint64_t n,
                                                         uint32_t weight) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t key = keys[i];
  const uint64_t wmask = (1ull << c.log2w) - 1ull;
#pragma unroll
  for (int row = 0; row < 4; ++row)
    atomicAdd(c.counts + (((uint64_t)row << c.log2w) | (splitmix64(key ^ cms_seed(row)) & wmask)), weight);
}
hipError_t launch_cms_add_keys(DevCms c, const uint64_t* keys, uint64_t n, uint32_t weight, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_cms_add_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, c, keys, n, weight);
  return hipGetLastError();
}

// sketch.rs:227-246: min over rows of the saturating counters.
__global__ __launch_bounds__(kBlock) void k_cms_query_keys(DevCms c, const uint64_t* __restrict__ keys, uint64_t n,
                                                           uint32_t cap, uint32_t* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t key = keys[i];
  const uint64_t wmask = (1ull << c.log2w) - 1ull;
  uint32_t mn = cap;
#pragma unroll
  for (int row = 0; row < 4; ++row) {
    uint32_t v = c.counts[((uint64_t)row << c.log2w) | (splitmix64(key ^ cms_seed(row)) & wmask)];
    mn = v < mn ? v : mn;
  }
  out[i] = mn;
}
hipError_t launch_cms_query_keys(DevCms c, const uint64_t* keys, uint64_t n, uint32_t cap, uint32_t* out,
                                 hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_cms_query_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, c, keys, n, cap, out);
  return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_hll_add_hashes(DevHll h, const uint64_t* __restrict__ hashes, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t idx, rho;
  hll_index_rho(hashes[i], h.p, idx, rho);
  if (rho > h.regs[idx]) atomicMax(h.regs + idx, rho);
}
hipError_t launch_hll_add_hashes(DevHll h, const uint64_t* hashes, uint64_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_hll_add_hashes, dim3(grid_for(n)), dim3(kBlock), 0, s, h, hashes, n);
  return hipGetLastError();
}

// add_canonical_hash (sketch.rs:66-70): add_hash of the splitmix64-mixed hash.
__global__ __launch_bounds__(kBlock) void k_hll_add_canonical(DevHll h, const uint64_t* __restrict__ hashes,
                                                              uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t idx, rho;
  hll_index_rho(splitmix64(hashes[i]), h.p, idx, rho);
  if (rho > h.regs[idx]) atomicMax(h.regs + idx, rho);
}
hipError_t launch_hll_add_canonical(DevHll h, const uint64_t* hashes, uint64_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_hll_add_canonical, dim3(grid_for(n)), dim3(kBlock), 0, s, h, hashes, n);
  return hipGetLastError();
}

// HyperLogLog::merge (sketch.rs:73-80): register-wise max.
__global__ __launch_bounds__(kBlock) void k_hll_merge(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                                      uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && src[i] > dst[i]) dst[i] = src[i];
}
hipError_t launch_hll_merge(uint32_t* dst, const uint32_t* src, uint64_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_hll_merge, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
  return hipGetLastError();
}

// add_and_estimate (sketch.rs:248-253) over a batch in index order.  For row
// `row`: cells[i] = key i's cell, idx[i] = i, to be sorted by cell (a stable
// sort keeps the indices of one cell increasing).
__global__ __launch_bounds__(kBlock) void k_cms_row_cells(DevCms c, const uint64_t* __restrict__ keys, uint64_t n,
                                                          uint32_t row, uint32_t* __restrict__ cells,
                                                          uint32_t* __restrict__ idx) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t wmask = (1ull << c.log2w) - 1ull;
  cells[i] = (uint32_t)(((uint64_t)row << c.log2w) | (splitmix64(keys[i] ^ cms_seed((int)row)) & wmask));
  idx[i] = (uint32_t)i;
}
// Sorted position p holds the rank-th add to its cell within the batch (rank
// = p minus the first position of its cell's run), so right after that add
// the cell reads counts[cell] + rank + 1; the estimate is the minimum over
// the rows, clamped at cap.  Row 0 writes out[], later rows lower it.
__global__ __launch_bounds__(kBlock) void k_cms_row_estimate(const uint32_t* __restrict__ counts,
                                                             const uint32_t* __restrict__ cells,
                                                             const uint32_t* __restrict__ idx, uint64_t n, uint32_t row,
                                                             uint32_t cap, uint32_t* __restrict__ out) {
  uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t cell = cells[p];
  uint64_t lo = 0, hi = p;  // first position whose cell equals cell
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (cells[mid] < cell) lo = mid + 1;
    else hi = mid;
  }
  const uint64_t v = (uint64_t)counts[cell] + (p - lo) + 1u;
  const uint32_t e = v < cap ? (uint32_t)v : cap;
  const uint32_t i = idx[p];
  if (row == 0 || e < out[i]) out[i] = e;
}
hipError_t launch_cms_row_cells(DevCms c, const uint64_t* keys, uint64_t n, uint32_t row, uint32_t* cells,
                                uint32_t* idx, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_cms_row_cells, dim3(grid_for(n)), dim3(kBlock), 0, s, c, keys, n, row, cells, idx);
  return hipGetLastError();
}
hipError_t launch_cms_row_estimate(const uint32_t* counts, const uint32_t* cells, const uint32_t* idx, uint64_t n,
                                   uint32_t row, uint32_t cap, uint32_t* out, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_cms_row_estimate, dim3(grid_for(n)), dim3(kBlock), 0, s, counts, cells, idx, n, row, cap, out);
  return hipGetLastError();
}
hipError_t sort_u32_pairs(const uint32_t* kin, uint32_t* kout, const uint32_t* vin, uint32_t* vout, uint64_t n,
                          uint32_t end_bit, void* tmp, size_t* tmp_bytes, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortPairs(tmp, *tmp_bytes, kin, kout, vin, vout, (int)n, 0, (int)end_bit, s);
}

// RollingNtHash (kmer.rs:433-477) over an ASCII sequence: thread t takes
// windows [64t, 64t + 64), pushing its first window's k bases (new) and
// rolling from there (roll); out[i] = canonical hash of window i.
__device__ __forceinline__ uint32_t ascii_code(uint8_t c) {
  return (c == 'C' || c == 'c') ? 1u : (c == 'G' || c == 'g') ? 2u : (c == 'T' || c == 't') ? 3u : 0u;
}
__global__ __launch_bounds__(kBlock) void k_nthash_windows(const uint8_t* __restrict__ seq, uint64_t n_windows,
                                                           uint32_t k, uint64_t* __restrict__ out) {
  const uint64_t w0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 64u;
  if (w0 >= n_windows) return;
  const uint64_t w1 = w0 + 64u < n_windows ? w0 + 64u : n_windows;
  NtHash h{0, 0};
  for (uint32_t j = 0; j < k; ++j) h.push(ascii_code(seq[w0 + j]), k);
  out[w0] = h.canonical();
  for (uint64_t w = w0 + 1; w < w1; ++w) {
    h.roll(ascii_code(seq[w - 1]), ascii_code(seq[w + k - 1]), k);
    out[w] = h.canonical();
  }
}
hipError_t launch_nthash_windows(const uint8_t* seq, uint64_t n_windows, uint32_t k, uint64_t* out, hipStream_t s) {
  if (!n_windows) return hipSuccess;
  const uint64_t threads = (n_windows + 63) / 64;
  hipLaunchKernelGGL(k_nthash_windows, dim3(grid_for(threads)), dim3(kBlock), 0, s, seq, n_windows, k, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Estimator.  Occurrence t of a k-mer sets cms "c >= 2" at insertion time iff
// every row's cell already had an insertion before t, i.e. max over rows of
// the cell's first insertion index is < t.  First insertion indices are an
// atomicMin per cell; HLL registers are recovered at any pair prefix from the
// first pair index at which each (register, rank) appeared.

constexpr uint32_t kEstK = 21;
constexpr uint32_t kEstLog2W = 23;

template <bool SOLID>
__global__ __launch_bounds__(kBlock) void k_est(EstArgs a) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n_reads) return;
  const uint32_t len = a.reads.lens[r];
  if (len < kEstK) return;
  const uint64_t* __restrict__ data = a.reads.data;
  const uint64_t g = a.reads.offs[r];
  const uint32_t pair = r >> 1;
  uint32_t t = a.occ_base[r];
  NtHash nh;
  nh.f = 0;
  nh.r = 0;
  BaseStream in, outs;
  in.init(data, g);
  for (uint32_t j = 0; j < kEstK; ++j) nh.push(in.next(), kEstK);
  outs.init(data, g);
  const uint64_t wmask = (1ull << kEstLog2W) - 1ull;
  const uint32_t nwin = len - kEstK + 1;
  for (uint32_t w = 0;; ++w, ++t) {
    const uint64_t raw = nh.canonical();
    uint32_t idx, rho;
    hll_index_rho(splitmix64(raw), 14, idx, rho);
    if (!SOLID) {
#pragma unroll
      for (int row = 0; row < 4; ++row)
        atomicMin(a.first_ins + (((uint64_t)row << kEstLog2W) | (splitmix64(raw ^ cms_seed(row)) & wmask)), t);
      uint32_t* f = a.first_total + idx * 64u + rho;
      if (pair < *f) atomicMin(f, pair);
    } else {
      uint32_t mx = 0;
#pragma unroll
      for (int row = 0; row < 4; ++row) {
        uint32_t v = a.first_ins[((uint64_t)row << kEstLog2W) | (splitmix64(raw ^ cms_seed(row)) & wmask)];
        mx = v > mx ? v : mx;
      }
      if (mx < t) {
        uint32_t* f = a.first_solid + idx * 64u + rho;
        if (pair < *f) atomicMin(f, pair);
      }
    }
    if (w + 1 >= nwin) break;
    nh.roll(outs.next(), in.next(), kEstK);
  }
}

hipError_t launch_est_first_insert(const EstArgs& a, hipStream_t s) {
  if (!a.n_reads) return hipSuccess;
  hipLaunchKernelGGL((k_est<false>), dim3(grid_for(a.n_reads)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_est_solid(const EstArgs& a, hipStream_t s) {
  if (!a.n_reads) return hipSuccess;
  hipLaunchKernelGGL((k_est<true>), dim3(grid_for(a.n_reads)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace fb
