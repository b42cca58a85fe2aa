// fb_polish.hip -- in-memory consensus polishing (src/polish/consensus.rs:64-271)
// on the device.
//
// 1. Seeds: the canonical ntHash (k 31) of every 25th window of the
//    output-eligible unitigs goes into an open-addressing table in HBM
//    (k_pol_sample); one thread per window of those unitigs then counts the
//    sampled hashes over all their windows (k_pol_count).  Seeds counted once
//    are the unique ones the reference keeps.
// 2. k_pol_reads: one thread per read (merged reads, then pe_forward) tries
//    its windows 0, 30, 60, .. against the unique seeds, checks the contig
//    k-mer forward or as the exact reverse complement, and votes its bases
//    into u32 pileup counters with atomics (+1 each, so a final clamp at
//    65535 equals the reference's u16 saturating adds in any order).
// 3. k_pol_consensus: one thread per unitig base takes the dominant base when
//    the (u16, wrapping) coverage is >= 5 and the dominant count is > 75%.
#include "fb_device.hpp"
#include "fb_internal.hpp"

namespace fb {

using namespace fbk;

constexpr int kPolBlock = 256;
constexpr uint32_t kPolK = 31;

static __device__ __forceinline__ uint32_t pol_code(uint8_t c) {
  switch (c) {
    case 'C': case 'c': return 1u;
    case 'G': case 'g': return 2u;
    case 'T': case 't': return 3u;
    default: return 0u;
  }
}
static __device__ __forceinline__ uint8_t pol_up(uint8_t c) { return (c >= 'a' && c <= 'z') ? (uint8_t)(c - 32) : c; }

static __device__ __forceinline__ uint64_t pol_slot(uint64_t h, uint32_t log2cap) {
  return splitmix64(h ^ 0xbb67ae8584caa73bull) >> (64u - log2cap);
}

// ntHash canonical of the 31 ASCII bases at p.
static __device__ __forceinline__ uint64_t ascii_hash(const uint8_t* p) {
  NtHash h{0, 0};
  for (uint32_t i = 0; i < kPolK; ++i) h.push(pol_code(p[i]), kPolK);
  return h.canonical();
}

// Window w of the range list: unitig rng_u[r], first window rng_p0[r], step rng_step.
static __device__ __forceinline__ void pol_window(const PolArgs& a, uint64_t w, uint32_t& u, uint32_t& p) {
  uint32_t lo = 0, hi = a.n_ranges;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.rpre[mid] <= w) lo = mid;
    else hi = mid;
  }
  u = a.rng_u[lo];
  p = (uint32_t)(w - a.rpre[lo]) * a.rng_step;
}

__global__ void k_pol_clear(PolArgs a) {
  const uint64_t cap = 1ull << a.log2cap;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
    a.keys[i] = kEmpty;
    a.cnt[i] = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x < 2) a.special[threadIdx.x] = 0u;  // [0] sampled, [1] count
}

// Sampled seeds (rng_step 25): claim the key and keep one (unitig, position).
__global__ __launch_bounds__(kPolBlock) void k_pol_sample(PolArgs a) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.n_windows) return;
  uint32_t u, p;
  pol_window(a, w, u, p);
  const uint64_t h = ascii_hash(a.seq + a.soff[u] + p);
  const uint64_t val = ((uint64_t)u << 32) | p;
  if (h == kEmpty) {
    a.special[0] = 1u;
    *a.special_val = val;
    return;
  }
  const uint64_t mask = (1ull << a.log2cap) - 1ull;
  for (uint64_t s = pol_slot(h, a.log2cap);; s = (s + 1) & mask) {
    const unsigned long long prev =
        atomicCAS((unsigned long long*)&a.keys[s], (unsigned long long)kEmpty, (unsigned long long)h);
    if (prev == kEmpty || prev == h) {
      a.val[s] = val;
      return;
    }
  }
}

// Every window (rng_step 1): count the sampled hashes it meets.
__global__ __launch_bounds__(kPolBlock) void k_pol_count(PolArgs a) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.n_windows) return;
  uint32_t u, p;
  pol_window(a, w, u, p);
  const uint64_t h = ascii_hash(a.seq + a.soff[u] + p);
  if (h == kEmpty) {
    if (a.special[0]) atomicAdd(a.special + 1, 1u);
    return;
  }
  const uint64_t mask = (1ull << a.log2cap) - 1ull;
  for (uint64_t s = pol_slot(h, a.log2cap);; s = (s + 1) & mask) {
    const uint64_t v = a.keys[s];
    if (v == h) {
      atomicAdd(a.cnt + s, 1u);
      return;
    }
    if (v == kEmpty) return;
  }
}

static __device__ __forceinline__ bool seed_of(const PolArgs& a, uint64_t h, uint64_t& val) {
  if (h == kEmpty) {
    if (!a.special[0] || a.special[1] != 1u) return false;
    val = *a.special_val;
    return true;
  }
  const uint64_t mask = (1ull << a.log2cap) - 1ull;
  for (uint64_t s = pol_slot(h, a.log2cap);; s = (s + 1) & mask) {
    const uint64_t v = a.keys[s];
    if (v == h) {
      if (a.cnt[s] != 1u) return false;
      val = a.val[s];
      return true;
    }
    if (v == kEmpty) return false;
  }
}

// The 31 bases at global base g of the packed reads, 2 bits each.
static __device__ __forceinline__ uint64_t read_bits(const uint64_t* __restrict__ data, uint64_t g) {
  const uint64_t w0 = data[g >> 5], w1 = data[(g >> 5) + 1];
  const uint32_t sh = (uint32_t)(g & 31u) << 1;
  return sh ? (w0 >> sh) | (w1 << (64u - sh)) : w0;
}

// Stats sums per block: [0] fwd reads, [1] rev reads, [2] fwd bases, [3] rev bases, [4] rejected.
__global__ __launch_bounds__(kPolBlock) void k_pol_reads(PolArgs a) {
  __shared__ unsigned long long s_sum[5];
  if (threadIdx.x < 5) s_sum[threadIdx.x] = 0ull;
  __syncthreads();
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < a.n_reads) {
    const uint64_t r = t < a.n_first ? a.first0 + t : a.first1 + (t - a.n_first);
    const uint32_t len = a.reads.lens[r];
    const uint64_t off = a.reads.offs[r];
    bool had = false;
    int orient = -1;
    uint32_t cid = 0, cpos = 0, rpos = 0;
    for (uint32_t p = 0; len >= kPolK && p + kPolK <= len && orient < 0; p += 30) {
      const uint64_t bits = read_bits(a.reads.data, off + p);
      NtHash h{0, 0};
      for (uint32_t i = 0; i < kPolK; ++i) h.push((uint32_t)(bits >> (2u * i)) & 3u, kPolK);
      uint64_t v = 0;
      if (!seed_of(a, h.canonical(), v)) continue;
      had = true;
      const uint32_t u = (uint32_t)(v >> 32), cp = (uint32_t)v;
      const uint8_t* anc = a.seq + a.soff[u] + cp;
      bool fwd = true, rev = true;
      for (uint32_t j = 0; j < kPolK; ++j) {
        const uint8_t c = pol_up(anc[j]);
        fwd = fwd && c == (uint8_t)"ACGT"[(bits >> (2u * j)) & 3u];
        rev = rev && c == (uint8_t)"TGCA"[(bits >> (2u * (kPolK - 1u - j))) & 3u];
      }
      if (fwd || rev) {
        orient = fwd ? 0 : 1;
        cid = u;
        cpos = cp;
        rpos = p;
      }
    }
    if (orient < 0) {
      if (had) atomicAdd(&s_sum[4], 1ull);
    } else {
      uint32_t* pile = a.pile + a.poff[cid] * 4ull;
      const int64_t plen = a.ulen[cid];
      uint64_t voted = 0;
      uint64_t g = off;
      uint64_t word = a.reads.data[g >> 5] >> ((g & 31u) << 1);
      for (uint32_t idx = 0; idx < len; ++idx, ++g) {
        if ((g & 31u) == 0 && idx) word = a.reads.data[g >> 5];
        const uint32_t b = (uint32_t)word & 3u;
        word >>= 2;
        const int64_t tp = orient == 0 ? (int64_t)cpos - rpos + idx : (int64_t)cpos + kPolK - 1 + rpos - idx;
        if (tp < 0 || tp >= plen) continue;
        atomicAdd(pile + (uint64_t)tp * 4u + (orient == 0 ? b : 3u - b), 1u);
        ++voted;
      }
      if (voted) {
        atomicAdd(&s_sum[orient], 1ull);
        atomicAdd(&s_sum[2 + orient], (unsigned long long)voted);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < 5 && s_sum[threadIdx.x]) atomicAdd(a.stats + threadIdx.x, s_sum[threadIdx.x]);
}

// One thread per base of an eligible unitig (pileup index i).
__global__ __launch_bounds__(kPolBlock) void k_pol_consensus(PolArgs a) {
  __shared__ unsigned long long s_mod;
  if (threadIdx.x == 0) s_mod = 0ull;
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n_pile) {
    uint32_t lo = 0, hi = a.n_elig;  // largest e with poff[elig[e]] <= i
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (a.poff[a.elig[mid]] <= i) lo = mid;
      else hi = mid;
    }
    const uint32_t u = a.elig[lo];
    const uint64_t p = i - a.poff[u];
    uint32_t c[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) c[b] = min(a.pile[i * 4 + b], 65535u);
    const uint32_t tot = (c[0] + c[1] + c[2] + c[3]) & 0xFFFFu;  // the reference sums u16s
    if (tot >= 5u) {
      uint32_t mi = 0;
      for (uint32_t b = 1; b < 4; ++b)
        if (c[b] > c[mi]) mi = b;
      const uint8_t nb = (uint8_t)"ACGT"[mi];
      uint8_t* out = a.out + a.soff[u] + p;
      if ((double)c[mi] / (double)tot > 0.75 && pol_up(*out) != nb) {
        *out = nb;
        atomicAdd(&s_mod, 1ull);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_mod) atomicAdd(a.stats + 5, s_mod);
}

static uint32_t pol_grid(uint64_t n) {
  const uint64_t g = (n + kPolBlock - 1) / kPolBlock;
  return (uint32_t)(g ? g : 1);
}

hipError_t launch_pol_clear(const PolArgs& a, hipStream_t s) {
  const uint32_t g = (uint32_t)std::min<uint64_t>(pol_grid(1ull << a.log2cap), 65536);
  hipLaunchKernelGGL(k_pol_clear, dim3(g), dim3(kPolBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_pol_sample(const PolArgs& a, hipStream_t s) {
  if (!a.n_windows) return hipSuccess;
  hipLaunchKernelGGL(k_pol_sample, dim3(pol_grid(a.n_windows)), dim3(kPolBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_pol_count(const PolArgs& a, hipStream_t s) {
  if (!a.n_windows) return hipSuccess;
  hipLaunchKernelGGL(k_pol_count, dim3(pol_grid(a.n_windows)), dim3(kPolBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_pol_reads(const PolArgs& a, hipStream_t s) {
  if (!a.n_reads) return hipSuccess;
  hipLaunchKernelGGL(k_pol_reads, dim3(pol_grid(a.n_reads)), dim3(kPolBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_pol_consensus(const PolArgs& a, hipStream_t s) {
  if (!a.n_pile) return hipSuccess;
  hipLaunchKernelGGL(k_pol_consensus, dim3(pol_grid(a.n_pile)), dim3(kPolBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace fb
