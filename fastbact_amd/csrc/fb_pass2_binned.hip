// fb_pass2_binned.hip -- pass 2 and pass 3 with LDS-privatised solid tables.
//
// Records are grouped by a hash of their canonical k-mer into bins small
// enough that one workgroup holds the bin's distinct keys in LDS
// (src/graph/build.rs:140-328: the accumulation is a set of saturating adds per
// key, so any grouping of the records gives the same result).
//   Phase A (k_p2_scatter): each thread walks its windows twice.  The first
//     walk gates windows against the Count-Min bitmap, keeps the gate bits in
//     a register and counts the workgroup's records per bin in LDS; one global
//     atomic per bin reserves a run.  The second walk writes one record per
//     window that has any (the key's 62-bit words and an edge byte).  A run
//     that would pass its bin's region sets an overflow flag and writes
//     nothing; the host then runs the global-table pass instead.
//   Phase B (k_p2_bins): one workgroup per bin builds the bin's table in LDS
//     (write-once keys claimed word by word with compare-and-swap, as the
//     global table does), then emits the solid k-mers (pass 3) directly.  A
//     key that finds the LDS table full goes to the global table; each key
//     is in exactly one of the two, so pass 3 over the global table appends
//     the rest.
#include "fb_device.hpp"
#include "fb_internal.hpp"
#include "fb_table.hpp"

namespace fb {
using namespace fbk;

constexpr int kP2ScatterBlock = 512;
constexpr int kP2BinBlock = 1024;

__device__ __forceinline__ uint64_t slot_hash2(uint64_t fold) { return splitmix64(fold ^ 0x9e3779b97f4a7c15ull); }

template <int L, int KW>
__device__ __forceinline__ void pack62b(const Kmer<L>& x, uint64_t (&w)[KW]) {
#pragma unroll
  for (int j = 0; j < KW; ++j) {
    const int bit = 62 * j;
    const int limb = bit / 64, off = bit % 64;
    uint64_t v = limb < L ? (x.a[limb] >> off) : 0ull;
    if (off > 2 && limb + 1 < L) v |= x.a[limb + 1] << (64 - off);
    w[j] = v & 0x3fffffffffffffffull;
  }
}

// One thread's walk over windows p0..p1 of read r.  cb(x, key, fold, left,
// e_left, right, e_right) is called for every window with a record; gates
// holds (PASS 0) or supplies (PASS 1) the gate bit of window p0+i at bit i.
template <int L, int PASS, class CB>
__device__ __forceinline__ void p2_walk(const P2BinArgs& a, uint64_t r, uint32_t p0, uint32_t p1, uint64_t& gates,
                                        CB&& cb) {
  const uint32_t k = a.k;
  const uint64_t* __restrict__ data = a.reads.data;
  const uint64_t g = a.reads.offs[r] + p0;
  Kmer<L> win = window_bits<L>(data, g, k);
  Kmer<L> fwd = fwd_from_window<L>(win, k);
  Kmer<L> rev = rc_from_window<L>(win, k);
  BaseStream in, first;
  in.init(data, g + k);
  first.init(data, g);
  bool cur_rc = kmer_lt<L>(rev, fwd);
  Kmer<L> cur = cur_rc ? rev : fwd;
  uint64_t cur_fold = kmer_fold<L>(cur);
  auto gate_of = [&](uint32_t i, uint64_t fold) -> bool {
    if (PASS == 0) {
      const uint64_t wmask = (1ull << a.log2w) - 1ull;
      bool ok = true;
#pragma unroll
      for (int row = 0; row < 4; ++row) {
        const uint64_t c = ((uint64_t)row << a.log2w) | (splitmix64(fold ^ cms_seed(row)) & wmask);
        ok &= (a.solid_bits[c >> 5] >> (c & 31u)) & 1u;
      }
      if (ok) gates |= 1ull << i;
      return ok;
    }
    return (gates >> i) & 1ull;
  };
  bool cur_solid = gate_of(0, cur_fold);
  bool left = false;
  uint32_t left_rcp = 0;
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
      nxt_solid = gate_of(x + 1 - p0, nxt_fold);
      right = cur_solid && nxt_solid;
    }
    const uint32_t first_base = first.next();
    if (left || right) {
      // edge_v of pair (x-1, x) and edge_u of pair (x, x+1), build.rs:268-275
      const uint32_t e_left = cur_rc ? left_rcp : left_rcp + 4u;
      const uint32_t e_right = cur_rc ? b + 4u : b;
      cb(cur, cur_fold, left, e_left, right, e_right);
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

__device__ __forceinline__ bool p2_share(const DevReads& R, uint64_t gid, uint32_t split, uint32_t k, uint64_t& r,
                                         uint32_t& p0, uint32_t& p1) {
  r = gid / split;
  const uint32_t part = (uint32_t)(gid - r * split);
  if (r >= R.n_reads) return false;
  const uint32_t len = R.lens[r];
  if (len < k + 1) return false;
  const uint32_t units = len - k;  // pairs
  p0 = (uint32_t)((uint64_t)part * units / split);
  p1 = (uint32_t)((uint64_t)(part + 1) * units / split);
  return p0 < p1;
}

template <int L, int KW>
__device__ __forceinline__ void p2_scatter_body(const P2BinArgs& a, uint64_t gid, uint32_t* sh) {
  uint32_t* hist = sh;
  uint32_t* cur = sh + a.nbins;
  for (uint32_t i = threadIdx.x; i < a.nbins; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  uint64_t r = 0;
  uint32_t p0 = 0, p1 = 0;
  const bool valid = p2_share(a.reads, gid, a.split, a.k, r, p0, p1);
  const uint32_t shift = 64u - a.log2_nbins;
  uint64_t gates = 0;
  if (valid)
    p2_walk<L, 0>(a, r, p0, p1, gates, [&](const Kmer<L>&, uint64_t fold, bool, uint32_t, bool, uint32_t) {
      const uint32_t bin = a.log2_nbins ? (uint32_t)(slot_hash2(fold) >> shift) : 0u;
      atomicAdd(hist + bin, 1u);
    });
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < a.nbins; b += blockDim.x) {
    const uint32_t c = hist[b];
    cur[b] = c ? atomicAdd(a.cursor + b, c) : 0u;
  }
  __syncthreads();
  if (valid) {
    const uint64_t cap = a.bin_cap;
    p2_walk<L, 1>(a, r, p0, p1, gates,
                  [&](const Kmer<L>& key, uint64_t fold, bool left, uint32_t el, bool right, uint32_t er) {
                    const uint32_t bin = a.log2_nbins ? (uint32_t)(slot_hash2(fold) >> shift) : 0u;
                    const uint32_t pos = atomicAdd(cur + bin, 1u);
                    if (pos >= cap) {
                      a.status[0] = 1u;  // region full: the host re-runs this pass on the global table
                      return;
                    }
                    uint64_t w[KW];
                    pack62b<L, KW>(key, w);
#pragma unroll
                    for (int j = 0; j < KW; ++j) a.rec_keys[((uint64_t)bin * KW + j) * cap + pos] = w[j];
                    a.rec_info[(uint64_t)bin * cap + pos] =
                        (uint8_t)((left ? 1u : 0u) | (el << 1) | (right ? 16u : 0u) | (er << 5));
                  });
  }
}

template <int L, int KW>
__global__ __launch_bounds__(kP2ScatterBlock) void k_p2_scatter(P2BinArgs a) {
  extern __shared__ uint32_t sh[];
  p2_scatter_body<L, KW>(a, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, sh);
}

template <int L, int KW>
__global__ __launch_bounds__(kP2ScatterBlock) void k_p2_scatter_m(const P2BinArgs* __restrict__ args,
                                                                  const uint32_t* __restrict__ blk, uint32_t n_iso) {
  extern __shared__ uint32_t sh[];
  const uint32_t iso = find_iso(blk, n_iso, blockIdx.x);
  p2_scatter_body<L, KW>(args[iso], (uint64_t)(blockIdx.x - blk[iso]) * blockDim.x + threadIdx.x, sh);
}

// LDS hash of a packed key (placement inside a bin; independent of the bin bits).
template <int KW>
__device__ __forceinline__ uint32_t lds_slot(const uint64_t (&w)[KW], uint32_t cap) {
  uint64_t h = w[0];
#pragma unroll
  for (int j = 1; j < KW; ++j) h = splitmix64(h) ^ w[j];
  h = splitmix64(h ^ 0x51ed270b27a3d9c5ull);
  return (uint32_t)(((h & 0xffffffffull) * (uint64_t)cap) >> 32);
}

// Finds or claims the LDS entry of key; -1 when the table is full.
template <int KW>
__device__ __forceinline__ int lds_find(unsigned long long* keys, uint32_t cap, const uint64_t (&w)[KW]) {
  uint32_t s = lds_slot<KW>(w, cap);
  for (uint32_t probe = 0; probe < cap; ++probe) {
    bool match = true;
#pragma unroll
    for (int j = 0; j < KW; ++j) {
      if (!match) break;
      unsigned long long* p = keys + (uint64_t)j * cap + s;
      unsigned long long v = *reinterpret_cast<volatile unsigned long long*>(p);
      if (v == kEmpty) {
        const unsigned long long old = atomicCAS(p, (unsigned long long)kEmpty, (unsigned long long)w[j]);
        v = old == kEmpty ? (unsigned long long)w[j] : old;
      }
      match = (v == w[j]);
    }
    if (match) return (int)s;
    if (++s == cap) s = 0;
  }
  return -1;
}

template <int KW>
__device__ __forceinline__ void p2_bins_body(const P2BinArgs& a, uint32_t bin, unsigned char* smem) {
  const uint32_t cap = a.lds_cap;
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(smem);      // KW * cap
  uint32_t* cov = reinterpret_cast<uint32_t*>(keys + (uint64_t)KW * cap);       // cap
  uint32_t* edges = cov + cap;                                                  // 4 * cap (u16 pairs)
  __shared__ uint32_t s_wave[kP2BinBlock / 64];
  __shared__ uint32_t s_base;
  __shared__ unsigned long long s_stat[2];
  for (uint32_t i = threadIdx.x; i < KW * cap; i += blockDim.x) keys[i] = kEmpty;
  for (uint32_t i = threadIdx.x; i < 5 * cap; i += blockDim.x) cov[i] = 0u;
  if (threadIdx.x < 2) s_stat[threadIdx.x] = 0ull;
  __syncthreads();
  const uint32_t filled = a.cursor[bin];
  const uint64_t rcap = a.bin_cap;
  const uint32_t n = filled < rcap ? filled : (uint32_t)rcap;
  const uint32_t gate = a.min_trans_cov;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    uint64_t w[KW];
#pragma unroll
    for (int j = 0; j < KW; ++j) w[j] = a.rec_keys[((uint64_t)bin * KW + j) * rcap + i];
    const uint32_t info = a.rec_info[(uint64_t)bin * rcap + i];
    const bool left = info & 1u, right = (info >> 4) & 1u;
    const uint32_t el = (info >> 1) & 7u, er = (info >> 5) & 7u;
    const int e = lds_find<KW>(keys, cap, w);
    if (e >= 0) {
      atomicAdd(cov + e, (uint32_t)left + (uint32_t)right);
      // u16 edge counters, two per word; an increment happens only while the
      // counter is below the gate, so it stays below gate + blockDim < 2^16
      if (left) {
        uint32_t* ew = edges + 4u * (uint32_t)e + (el >> 1);
        const uint32_t sh = (el & 1u) * 16u;
        if (((*reinterpret_cast<volatile uint32_t*>(ew) >> sh) & 0xffffu) < gate) atomicAdd(ew, 1u << sh);
      }
      if (right) {
        uint32_t* ew = edges + 4u * (uint32_t)e + (er >> 1);
        const uint32_t sh = (er & 1u) * 16u;
        if (((*reinterpret_cast<volatile uint32_t*>(ew) >> sh) & 0xffffu) < gate) atomicAdd(ew, 1u << sh);
      }
    } else {
      // LDS table full: this key lives in the global table for the whole pass
      uint32_t* c = table_find_global<KW>(a.table, w, slot_hash_words<KW>(w));
      if (c) {
        atomicAdd(c, (uint32_t)left + (uint32_t)right);
        if (left && c[2 + el] < gate) atomicAdd(c + 2 + el, 1u);
        if (right && c[2 + er] < gate) atomicAdd(c + 2 + er, 1u);
      }
    }
  }
  __syncthreads();
  // pass 3 on the LDS table: count, one atomic per block, write
  const uint32_t per = (cap + blockDim.x - 1) / blockDim.x;
  const uint32_t lo = threadIdx.x * per, hi = lo + per < cap ? lo + per : cap;
  auto keep_of = [&](uint32_t e, uint32_t& mask, uint32_t& cv) -> bool {
    if (keys[e] == kEmpty) return false;
    mask = 0;
#pragma unroll
    for (uint32_t q = 0; q < 8; ++q) {
      const uint32_t v = (edges[4u * e + (q >> 1)] >> ((q & 1u) * 16u)) & 0xffffu;
      if (v >= gate) mask |= 1u << q;
    }
    const uint32_t c = cov[e] < 65535u ? cov[e] : 65535u;
    cv = c / 2u;
    if (cv < 1u) cv = 1u;
    return mask != 0u;
  };
  uint32_t mine = 0, held = 0;
  unsigned long long raw = 0;
  for (uint32_t e = lo; e < hi; ++e) {
    uint32_t m, cv;
    mine += keep_of(e, m, cv) ? 1u : 0u;
    if (keys[e] != kEmpty) {
      ++held;
      raw += cov[e];
    }
  }
  if (held) {
    atomicAdd(s_stat, raw);
    atomicAdd(s_stat + 1, (unsigned long long)held);
  }
  // block exclusive scan of `mine`
  uint32_t incl = mine;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t v = __shfl_up(incl, off);
    if ((threadIdx.x & 63u) >= (uint32_t)off) incl += v;
  }
  if ((threadIdx.x & 63u) == 63u) s_wave[threadIdx.x >> 6] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
      const uint32_t v = s_wave[w];
      s_wave[w] = t;
      t += v;
    }
    s_base = t ? atomicAdd(a.out_count, t) : 0u;
    if (a.stats && s_stat[1]) {
      atomicAdd(a.stats, s_stat[0]);
      atomicAdd(a.stats + 1, s_stat[1]);
    }
  }
  __syncthreads();
  uint32_t pos = s_base + s_wave[threadIdx.x >> 6] + incl - mine;
  const uint32_t nref = a.nref;
  for (uint32_t e = lo; e < hi; ++e) {
    uint32_t m, cv;
    if (!keep_of(e, m, cv)) continue;
    uint64_t wk[KW];
#pragma unroll
    for (int j = 0; j < KW; ++j) wk[j] = keys[(uint64_t)j * cap + e];
    uint64_t limb[4] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < KW; ++j) {
      const int bit = 62 * j, li = bit / 64, off = bit % 64;
      if (li < 4) limb[li] |= wk[j] << off;
      if (off > 2 && li + 1 < 4) limb[li + 1] |= wk[j] >> (64 - off);
    }
    for (uint32_t q = 0; q < nref; ++q) a.out_keys[(uint64_t)pos * nref + q] = limb[nref - 1 - q];
    a.out_cov[pos] = (uint16_t)cv;
    a.out_mask[pos] = (uint8_t)m;
    ++pos;
  }
}

template <int KW>
__global__ __launch_bounds__(kP2BinBlock) void k_p2_bins(P2BinArgs a) {
  extern __shared__ unsigned char smem[];
  p2_bins_body<KW>(a, blockIdx.x, smem);
}

template <int KW>
__global__ __launch_bounds__(kP2BinBlock) void k_p2_bins_m(const P2BinArgs* __restrict__ args,
                                                           const uint32_t* __restrict__ blk, uint32_t n_iso) {
  extern __shared__ unsigned char smem[];
  const uint32_t iso = find_iso(blk, n_iso, blockIdx.x);
  p2_bins_body<KW>(args[iso], blockIdx.x - blk[iso], smem);
}

// LDS the bin kernels may use: the full 160 KB once the attribute is granted,
// else the 64 KB every workgroup has.
static uint32_t g_lds_bytes = 0;

// Dynamic LDS requested beyond the 64 KB default: the workgroup limit less
// 1 KB for the kernels' static LDS.
constexpr int kDynLds = (int)kP2LdsBytes - 1024;

static hipError_t set_attr(const void* f) {
  return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kDynLds);
}

template <int KW>
static hipError_t set_bins_attr() {
  hipError_t e = set_attr(reinterpret_cast<const void*>(&k_p2_bins<KW>));
  if (e == hipSuccess) e = set_attr(reinterpret_cast<const void*>(&k_p2_bins_m<KW>));
  return e;
}
template <int L, int KW>
static hipError_t set_scatter_attr() {
  hipError_t e = set_attr(reinterpret_cast<const void*>(&k_p2_scatter<L, KW>));
  if (e == hipSuccess) e = set_attr(reinterpret_cast<const void*>(&k_p2_scatter_m<L, KW>));
  return e;
}

static uint32_t lds_budget() {
  if (!g_lds_bytes) {
    hipError_t e = set_bins_attr<1>();
    if (e == hipSuccess) e = set_bins_attr<2>();
    if (e == hipSuccess) e = set_bins_attr<3>();
    if (e == hipSuccess) e = set_bins_attr<4>();
    if (e == hipSuccess) e = set_bins_attr<5>();
    if (e == hipSuccess) e = set_scatter_attr<1, 1>();
    if (e == hipSuccess) e = set_scatter_attr<1, 2>();
    if (e == hipSuccess) e = set_scatter_attr<2, 2>();
    if (e == hipSuccess) e = set_scatter_attr<2, 3>();
    if (e == hipSuccess) e = set_scatter_attr<3, 3>();
    if (e == hipSuccess) e = set_scatter_attr<3, 4>();
    if (e == hipSuccess) e = set_scatter_attr<4, 4>();
    if (e == hipSuccess) e = set_scatter_attr<4, 5>();
    if (e != hipSuccess) (void)hipGetLastError();
    g_lds_bytes = e == hipSuccess ? kP2LdsBytes : 65536u;
  }
  return g_lds_bytes;
}

uint32_t p2_max_bins() { return (lds_budget() - 1024u) / 8u >= 16384u ? 16384u : 8192u; }

uint32_t p2_lds_cap(uint32_t kw) {
  // keys + cov + 8 u16 edges per entry within the workgroup's LDS
  const uint32_t entry = kw * 8u + 4u + 16u;
  return (lds_budget() - 1024u) / entry;
}
size_t p2_lds_bytes(uint32_t kw, uint32_t cap) { return (size_t)cap * (kw * 8u + 4u + 16u); }

hipError_t launch_p2_binned(const P2BinArgs& a, hipStream_t s) {
  const uint32_t L = (2 * a.k + 63) / 64;
  const uint32_t KW = (a.k + 30) / 31;
  const size_t scat_lds = (size_t)a.nbins * 8;
  if (a.reads.n_reads) {
    const uint64_t threads = a.reads.n_reads * (uint64_t)a.split;
    const unsigned grid = (unsigned)((threads + kP2ScatterBlock - 1) / kP2ScatterBlock);
#define SC(LL, KK) hipLaunchKernelGGL((k_p2_scatter<LL, KK>), dim3(grid), dim3(kP2ScatterBlock), scat_lds, s, a)
    if (L == 1 && KW == 1) SC(1, 1);
    else if (L == 1) SC(1, 2);
    else if (L == 2 && KW == 2) SC(2, 2);
    else if (L == 2) SC(2, 3);
    else if (L == 3 && KW == 3) SC(3, 3);
    else if (L == 3) SC(3, 4);
    else if (KW == 4) SC(4, 4);
    else SC(4, 5);
#undef SC
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const size_t lds = p2_lds_bytes(KW, a.lds_cap);
  if (lds + 1024 > lds_budget() || scat_lds + 1024 > lds_budget()) return hipErrorInvalidConfiguration;
#define BN(KK) hipLaunchKernelGGL((k_p2_bins<KK>), dim3(a.nbins), dim3(kP2BinBlock), lds, s, a)
  switch (KW) {
    case 1: BN(1); break;
    case 2: BN(2); break;
    case 3: BN(3); break;
    case 4: BN(4); break;
    default: BN(5); break;
  }
#undef BN
  return hipGetLastError();
}

hipError_t launch_p2_binned_multi(const MultiLaunch& scatter, const MultiLaunch& bins, uint32_t k,
                                  uint32_t max_nbins, uint32_t lds_cap, hipStream_t s) {
  const uint32_t L = (2 * k + 63) / 64;
  const uint32_t KW = (k + 30) / 31;
  const P2BinArgs* a = static_cast<const P2BinArgs*>(scatter.args);
  const size_t scat_lds = (size_t)max_nbins * 8;
  if (scatter.total_blocks) {
#define SCM(LL, KK)                                                                                              \
  hipLaunchKernelGGL((k_p2_scatter_m<LL, KK>), dim3(scatter.total_blocks), dim3(kP2ScatterBlock), scat_lds, s, a, \
                     scatter.blk, scatter.n_iso)
    if (L == 1 && KW == 1) SCM(1, 1);
    else if (L == 1) SCM(1, 2);
    else if (L == 2 && KW == 2) SCM(2, 2);
    else if (L == 2) SCM(2, 3);
    else if (L == 3 && KW == 3) SCM(3, 3);
    else if (L == 3) SCM(3, 4);
    else if (KW == 4) SCM(4, 4);
    else SCM(4, 5);
#undef SCM
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (!bins.total_blocks) return hipSuccess;
  const size_t lds = p2_lds_bytes(KW, lds_cap);
  const P2BinArgs* ab = static_cast<const P2BinArgs*>(bins.args);
#define BNM(KK) \
  hipLaunchKernelGGL((k_p2_bins_m<KK>), dim3(bins.total_blocks), dim3(kP2BinBlock), lds, s, ab, bins.blk, bins.n_iso)
  switch (KW) {
    case 1: BNM(1); break;
    case 2: BNM(2); break;
    case 3: BNM(3); break;
    case 4: BNM(4); break;
    default: BNM(5); break;
  }
#undef BNM
  return hipGetLastError();
}

}  // namespace fb
