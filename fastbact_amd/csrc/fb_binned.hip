// fb_binned.hip -- pass 1 with LDS-privatised Count-Min slices.
//
// The Count-Min table (4 x width cells) is cut into bins of 2^sb cells, each
// small enough to count in one workgroup's LDS.
//   Phase A (k_bin_scatter): each workgroup walks its reads twice.  The first
//     walk counts the workgroup's updates per bin in LDS (and feeds the HLL).
//     One global atomic per bin then reserves a run in that bin's region,
//     in the slice that belongs to the workgroup's XCD.  The second walk goes
//     in rounds of kRound windows per thread: the round's cells are counted
//     per bin, sorted by bin into an LDS stage (a counting sort), and written
//     out in stage order, so a wave's store covers a few contiguous runs
//     instead of 64 scattered lines.  An update that finds its slice full is
//     added to the table directly with a global atomic, so any input
//     (low-complexity reads included) stays exact.
//   Phase B (k_bin_count): one workgroup per bin counts the offsets of the
//     bin's slices with LDS atomics, adds them to the table (the bin's cells belong to this
//     workgroup alone) and writes the pass-2 threshold bits of those cells.
// Every update is +1, so the final counters equal the direct-atomic ones.
#include "fb_device.hpp"
#include "fb_internal.hpp"

namespace fb {
using namespace fbk;

constexpr int kScatterBlock = (int)kP1BlockSize;
constexpr int kCountBlock = 1024;
constexpr int kRound = 7;                                     // windows per thread per staging round
constexpr uint32_t kStage = 4u * kRound * kP1BlockSize;  // staged cells per round
static_assert(kSlices == kBinSlices, "slice count of device and host code");

__device__ __forceinline__ bool share_of(const DevReads& R, uint64_t gid, uint32_t split, uint32_t k, uint64_t& r,
                                         uint32_t& w0, uint32_t& w1) {
  r = gid / split;
  uint32_t part = (uint32_t)(gid - r * split);
  if (r >= R.n_reads) return false;
  uint32_t len = R.lens[r];
  if (len < k) return false;
  uint32_t units = len - k + 1;
  w0 = (uint32_t)((uint64_t)part * units / split);
  w1 = (uint32_t)((uint64_t)(part + 1) * units / split);
  return w0 < w1;
}

// Calls f(cell) for the 4 cells of every window of this thread's share and,
// when HLL, updates the HLL registers.
template <int L, bool HLL, class F>
__device__ __forceinline__ void walk_cells(const BinArgs& a, uint64_t r, uint32_t w0, uint32_t w1, F&& f) {
  const uint32_t k = a.k;
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
  const uint32_t lw = a.log2w;
  const uint64_t wmask = (1ull << lw) - 1ull;
  for (uint32_t w = w0;;) {
    const uint64_t key = kmer_lt<L>(rev, fwd) ? kmer_fold<L>(rev) : kmer_fold<L>(fwd);
#pragma unroll
    for (int row = 0; row < 4; ++row) f(((uint64_t)row << lw) | (splitmix64(key ^ cms_seed(row)) & wmask));
    if (HLL) {
      uint32_t idx, rho;
      hll_index_rho(splitmix64(nh.canonical()), a.hll.p, idx, rho);
      if (rho > a.hll.regs[idx]) atomicMax(a.hll.regs + idx, rho);
    }
    if (++w >= w1) break;
    const uint32_t b = in.next();
    kmer_roll<L>(fwd, rev, b, k);
    if (HLL) nh.roll(outs.next(), b, k);
  }
}

// The windows of a thread's shares gid0, gid0 + blockDim, ... (chunks of
// them), one share at a time.
template <int L>
struct CellWalker {
  uint64_t gid;     // next share to open
  uint32_t left;    // shares not yet opened
  uint32_t w, w1;   // next window of the open share and its end; w == w1: none open
  Kmer<L> fwd, rev;
  BaseStream in;

  __device__ __forceinline__ void init(uint64_t gid0, uint32_t chunks) {
    gid = gid0;
    left = chunks;
    w = w1 = 0;
  }
  // Opens the next share with windows, if any.
  __device__ __forceinline__ void open(const BinArgs& a) {
    while (left) {
      uint64_t r = 0;
      uint32_t w0 = 0, e = 0;
      const bool ok = share_of(a.reads, gid, a.split, a.k, r, w0, e);
      gid += blockDim.x;
      --left;
      if (ok) {
        const uint64_t g = a.reads.offs[r] + w0;
        const Kmer<L> win = window_bits<L>(a.reads.data, g, a.k);
        fwd = fwd_from_window<L>(win, a.k);
        rev = rc_from_window<L>(win, a.k);
        in.init(a.reads.data, g + a.k);
        w = w0;
        w1 = e;
        return;
      }
    }
  }
  __device__ __forceinline__ bool has() const { return w < w1; }
  // Canonical fold of the next window of the open share (has() must hold).
  __device__ __forceinline__ uint64_t take(const BinArgs& a) {
    const uint64_t key = kmer_lt<L>(rev, fwd) ? kmer_fold<L>(rev) : kmer_fold<L>(fwd);
    if (++w < w1) kmer_roll<L>(fwd, rev, in.next(), a.k);
    return key;
  }
};

__device__ __forceinline__ uint64_t cell_of(const BinArgs& a, uint64_t key, int row) {
  return ((uint64_t)row << a.log2w) | (splitmix64(key ^ cms_seed(row)) & ((1ull << a.log2w) - 1ull));
}

// Exclusive scan of v[0..n) in place; one block.
__device__ __forceinline__ void block_exclusive_scan(uint32_t* v, uint32_t n, uint32_t* s_wave) {
  const uint32_t per = (n + blockDim.x - 1) / blockDim.x;
  const uint32_t lo = threadIdx.x * per < n ? threadIdx.x * per : n;
  const uint32_t hi = lo + per < n ? lo + per : n;
  uint32_t sum = 0;
  for (uint32_t i = lo; i < hi; ++i) sum += v[i];
  uint32_t incl = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(incl, off);
    if ((threadIdx.x & 63u) >= (uint32_t)off) incl += t;
  }
  if ((threadIdx.x & 63u) == 63u) s_wave[threadIdx.x >> 6] = incl;
  __syncthreads();
  uint32_t run = incl - sum;
  for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w) run += s_wave[w];
  for (uint32_t i = lo; i < hi; ++i) {
    const uint32_t t = v[i];
    v[i] = run;
    run += t;
  }
}

// Block blk covers thread shares [blk * chunks * blockDim, (blk + 1) * chunks * blockDim).
// LDS: gcur[nbins] (the block's next position per bin), rcur[nbins] (round
// counts, then run ends), stage[kStage].
template <int L, bool HLL>
__device__ __forceinline__ void bin_scatter_body(const BinArgs& a, uint64_t blk, uint32_t* sh) {
  __shared__ uint32_t s_wave[kScatterBlock / 64];
  const uint32_t nb = a.nbins;
  uint32_t* gcur = sh;
  uint32_t* rcur = sh + nb;
  uint32_t* stage = sh + 2 * nb;
  for (uint32_t i = threadIdx.x; i < 2 * nb; i += blockDim.x) sh[i] = 0;
  __syncthreads();
  const uint64_t gid0 = blk * a.chunks * blockDim.x + threadIdx.x;
  const uint32_t sb = a.sb;
  for (uint32_t ch = 0; ch < a.chunks; ++ch) {
    uint64_t r = 0;
    uint32_t w0 = 0, w1 = 0;
    if (share_of(a.reads, gid0 + (uint64_t)ch * blockDim.x, a.split, a.k, r, w0, w1))
      walk_cells<L, HLL>(a, r, w0, w1, [&](uint64_t cell) { atomicAdd(gcur + (cell >> sb), 1u); });
  }
  __syncthreads();
  const uint32_t sl = xcc_slice();
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    const uint32_t c = gcur[b];
    gcur[b] = c ? atomicAdd(a.cursor + (uint64_t)b * kSlices + sl, c) : 0u;
  }
  // (the first round's barrier orders these writes before their use)
  const uint64_t cap = a.cap;
  const uint32_t smask = (1u << sb) - 1u;
  uint16_t* __restrict__ out = a.data + sl * cap;
  uint32_t* __restrict__ counts = a.counts;
  CellWalker<L> wk;
  wk.init(gid0, a.chunks);
  for (;;) {
    // a round takes up to kRound windows of one share; its cells stay in registers
    if (!wk.has()) wk.open(a);
    uint32_t cells[4 * kRound];
    uint32_t n = 0;
#pragma unroll
    for (int j = 0; j < kRound; ++j) {
      if (wk.has()) {
        const uint64_t key = wk.take(a);
#pragma unroll
        for (int row = 0; row < 4; ++row) {
          cells[4 * j + row] = (uint32_t)cell_of(a, key, row);
          atomicAdd(rcur + (cells[4 * j + row] >> sb), 1u);
        }
        n = j + 1;
      }
    }
    if (!__syncthreads_or(n)) break;
    block_exclusive_scan(rcur, nb, s_wave);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRound; ++j) {
      if ((uint32_t)j < n) {
#pragma unroll
        for (int row = 0; row < 4; ++row) {
          const uint32_t cell = cells[4 * j + row];
          stage[atomicAdd(rcur + (cell >> sb), 1u)] = cell;
        }
      }
    }
    __syncthreads();
    // rcur[b] is now the end of bin b's run in the stage
    const uint32_t total = rcur[nb - 1];
    for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
      const uint32_t cell = stage[i];
      const uint32_t b = cell >> sb;
      const uint32_t pos = gcur[b] + i - (b ? rcur[b - 1] : 0u);
      if (pos < cap) out[(uint64_t)b * (kSlices * cap) + pos] = (uint16_t)(cell & smask);
      else atomicAdd(counts + cell, 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) gcur[b] += rcur[b] - (b ? rcur[b - 1] : 0u);
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) rcur[b] = 0;
    __syncthreads();
  }
}

template <int L, bool HLL>
__global__ __launch_bounds__(kScatterBlock) void k_bin_scatter(BinArgs a) {
  extern __shared__ uint32_t sh[];
  bin_scatter_body<L, HLL>(a, blockIdx.x, sh);
}

template <int L, bool HLL>
__global__ __launch_bounds__(kScatterBlock) void k_bin_scatter_m(const BinArgs* __restrict__ args,
                                                                 const uint32_t* __restrict__ blk, uint32_t n_iso) {
  extern __shared__ uint32_t sh[];
  const uint32_t iso = find_iso(blk, n_iso, blockIdx.x);
  bin_scatter_body<L, HLL>(args[iso], blockIdx.x - blk[iso], sh);
}

__device__ __forceinline__ void bin_count_body(const BinArgs& a, uint32_t b, uint32_t* lds) {
  const uint32_t S = 1u << a.sb;
  const uint64_t first = (uint64_t)b << a.sb;
  const uint32_t ncells = (uint32_t)((a.n_cells - first) < S ? (a.n_cells - first) : S);
  for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) lds[i] = 0;
  __syncthreads();
  for (uint32_t sl = 0; sl < kSlices; ++sl) {
    const uint32_t filled = a.cursor[(uint64_t)b * kSlices + sl];
    const uint32_t n = filled < a.cap ? filled : (uint32_t)a.cap;
    const uint16_t* __restrict__ d = a.data + ((uint64_t)b * kSlices + sl) * a.cap;  // cap is a multiple of 4
    const uint32_t n4 = n & ~3u;
    for (uint32_t i = threadIdx.x * 4u; i < n4; i += blockDim.x * 4u) {
      const uint2 v = *reinterpret_cast<const uint2*>(d + i);
      atomicAdd(lds + (v.x & 0xFFFFu), 1u);
      atomicAdd(lds + (v.x >> 16), 1u);
      atomicAdd(lds + (v.y & 0xFFFFu), 1u);
      atomicAdd(lds + (v.y >> 16), 1u);
    }
    for (uint32_t i = n4 + threadIdx.x; i < n; i += blockDim.x) atomicAdd(lds + d[i], 1u);
  }
  __syncthreads();
  // add into the table and emit the threshold bits, 64 cells per wave step
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  for (uint32_t base = wave * 64u; base < ncells; base += nwaves * 64u) {
    const uint32_t i = base + lane;
    bool on = false;
    if (i < ncells) {
      const uint64_t cell = first + i;
      const uint32_t v = a.counts[cell] + lds[i];
      a.counts[cell] = v;
      const uint32_t c = v < a.cap_value ? v : a.cap_value;
      on = c >= a.t_solid;
    }
    const uint64_t m = __ballot(on);
    if (a.bits && lane == 0) {
      const uint64_t word = (first + base) >> 5;
      const uint64_t nwords = (a.n_cells + 31) >> 5;
      if (word < nwords) a.bits[word] = (uint32_t)m;
      if (word + 1 < nwords) a.bits[word + 1] = (uint32_t)(m >> 32);
    }
  }
}

__global__ __launch_bounds__(kCountBlock) void k_bin_count(BinArgs a) {
  extern __shared__ uint32_t lds[];
  bin_count_body(a, blockIdx.x, lds);
}

__global__ __launch_bounds__(kCountBlock) void k_bin_count_m(const BinArgs* __restrict__ args,
                                                             const uint32_t* __restrict__ blk, uint32_t n_iso) {
  extern __shared__ uint32_t lds[];
  const uint32_t iso = find_iso(blk, n_iso, blockIdx.x);
  bin_count_body(args[iso], blockIdx.x - blk[iso], lds);
}

size_t bin_scatter_lds(uint32_t nbins) { return ((size_t)nbins * 2 + kStage) * sizeof(uint32_t); }

// Dynamic LDS beyond the 64 KB default for the scatter (bins + stage) and
// count (2^sb counters) kernels; the limit less 1 KB of static LDS.
static hipError_t pass1_lds_attrs() {
  static hipError_t done = hipErrorNotReady;
  if (done == hipErrorNotReady) {
    const int bytes = (int)kP2LdsBytes - 1024;
    const void* f[] = {reinterpret_cast<const void*>(&k_bin_scatter<1, true>),
                       reinterpret_cast<const void*>(&k_bin_scatter<1, false>),
                       reinterpret_cast<const void*>(&k_bin_scatter<2, true>),
                       reinterpret_cast<const void*>(&k_bin_scatter<2, false>),
                       reinterpret_cast<const void*>(&k_bin_scatter<3, true>),
                       reinterpret_cast<const void*>(&k_bin_scatter<3, false>),
                       reinterpret_cast<const void*>(&k_bin_scatter<4, true>),
                       reinterpret_cast<const void*>(&k_bin_scatter<4, false>),
                       reinterpret_cast<const void*>(&k_bin_scatter_m<1, true>),
                       reinterpret_cast<const void*>(&k_bin_scatter_m<1, false>),
                       reinterpret_cast<const void*>(&k_bin_scatter_m<2, true>),
                       reinterpret_cast<const void*>(&k_bin_scatter_m<2, false>),
                       reinterpret_cast<const void*>(&k_bin_scatter_m<3, true>),
                       reinterpret_cast<const void*>(&k_bin_scatter_m<3, false>),
                       reinterpret_cast<const void*>(&k_bin_scatter_m<4, true>),
                       reinterpret_cast<const void*>(&k_bin_scatter_m<4, false>),
                       reinterpret_cast<const void*>(&k_bin_count),
                       reinterpret_cast<const void*>(&k_bin_count_m)};
    hipError_t e = hipSuccess;
    for (const void* fn : f)
      if (e == hipSuccess) e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    done = e;
  }
  return done;
}

hipError_t check_bin_pass1(uint32_t nbins, uint32_t sb) {
  if (hipError_t e = pass1_lds_attrs()) return e;
  if (bin_scatter_lds(nbins) + 1024 > kP2LdsBytes || ((size_t)4 << sb) + 1024 > kP2LdsBytes)
    return hipErrorInvalidConfiguration;
  return hipSuccess;
}

hipError_t launch_bin_pass1(const BinArgs& a, hipStream_t s, hipEvent_t mid) {
  if (a.nbins == 0 || (a.nbins << a.sb) < a.n_cells) return hipErrorInvalidValue;
  if (hipError_t e = check_bin_pass1(a.nbins, a.sb)) return e;
  if (a.reads.n_reads) {
    const uint32_t L = (2 * a.k + 63) / 64;
    const unsigned grid = (unsigned)scatter_blocks(a.reads.n_reads * (uint64_t)a.split, a.chunks, kP1BlockSize);
    const size_t lds = bin_scatter_lds(a.nbins);
    const bool h = a.hll.regs != nullptr;
#define BS(LL)                                                                                              \
  if (h) hipLaunchKernelGGL((k_bin_scatter<LL, true>), dim3(grid), dim3(kScatterBlock), lds, s, a);        \
  else hipLaunchKernelGGL((k_bin_scatter<LL, false>), dim3(grid), dim3(kScatterBlock), lds, s, a)
    switch (L) {
      case 1: BS(1); break;
      case 2: BS(2); break;
      case 3: BS(3); break;
      default: BS(4); break;
    }
#undef BS
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (mid)
    if (hipError_t e = hipEventRecord(mid, s)) return e;
  hipLaunchKernelGGL(k_bin_count, dim3(a.nbins), dim3(kCountBlock), (size_t)(1u << a.sb) * 4, s, a);
  return hipGetLastError();
}

hipError_t launch_bin_scatter_multi(const MultiLaunch& m, uint32_t k, bool hll, uint32_t max_nbins, hipStream_t s) {
  if (!m.total_blocks) return hipSuccess;
  if (hipError_t e = check_bin_pass1(max_nbins, 6)) return e;  // count LDS checked by its launcher
  const uint32_t L = (2 * k + 63) / 64;
  const BinArgs* a = static_cast<const BinArgs*>(m.args);
  const size_t lds = bin_scatter_lds(max_nbins);
#define BSM(LL)                                                                                                   \
  if (hll) hipLaunchKernelGGL((k_bin_scatter_m<LL, true>), dim3(m.total_blocks), dim3(kScatterBlock), lds, s, a, m.blk, \
                              m.n_iso);                                                                           \
  else hipLaunchKernelGGL((k_bin_scatter_m<LL, false>), dim3(m.total_blocks), dim3(kScatterBlock), lds, s, a, m.blk, m.n_iso)
  switch (L) {
    case 1: BSM(1); break;
    case 2: BSM(2); break;
    case 3: BSM(3); break;
    default: BSM(4); break;
  }
#undef BSM
  return hipGetLastError();
}

hipError_t launch_bin_count_multi(const MultiLaunch& m, uint32_t max_sb, hipStream_t s) {
  if (!m.total_blocks) return hipSuccess;
  if (hipError_t e = check_bin_pass1(1, max_sb)) return e;
  hipLaunchKernelGGL(k_bin_count_m, dim3(m.total_blocks), dim3(kCountBlock), (size_t)(1u << max_sb) * 4, s,
                     static_cast<const BinArgs*>(m.args), m.blk, m.n_iso);
  return hipGetLastError();
}

}  // namespace fb
