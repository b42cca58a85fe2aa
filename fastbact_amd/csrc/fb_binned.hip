// fb_binned.hip -- pass 1 with LDS-privatised Count-Min slices.
//
// The Count-Min table (4 x width cells) is cut into bins of 2^sb cells, each
// small enough to count in one workgroup's LDS.
//   Phase A (k_bin_scatter): each workgroup walks its reads once, in rounds
//     of kRound windows per thread.  A round's cells are counted per bin,
//     sorted by bin into an LDS stage (a counting sort) and written out in
//     stage order, so a wave's store covers a few contiguous runs instead of
//     64 scattered lines.  Each bin's offsets go into chunks of kChunk
//     entries (one 128-byte line) that the workgroup reserves in the bin's
//     region with one global atomic per chunk, in the slice that belongs to
//     the workgroup's XCD; a workgroup fills its own lines only, and marks
//     the unused tail of its last chunk per bin with kNoCell.  An update that
//     finds its slice full is added to the table directly with a global
//     atomic, so any input (low-complexity reads included) stays exact.
//     With a resident grid (or the batch's per-isolate grids) and
//     kBufMinBins to kBufMaxBins bins, phase A is
//     k_bin_buffer instead: every bin has a 64-byte buffer in LDS, and a
//     buffer is written out only when full, so each store to a bin region
//     covers 64 whole bytes (no line is written back partly filled).
//   Phase B (k_bin_count): one workgroup per bin counts the offsets of the
//     bin's slices with LDS atomics, adds them to the table (the bin's cells
//     belong to this workgroup alone) and writes the pass-2 threshold bits of
//     those cells.
// Every update is +1, so the final counters equal the direct-atomic ones.
#include "fb_device.hpp"
#include "fb_internal.hpp"

namespace fb {
using namespace fbk;

constexpr int kScatterBlock = (int)kP1BlockSize;
constexpr int kCountBlock = 1024;
constexpr int kRound = 6;                                // windows per thread per staging round
constexpr uint32_t kStage = 4u * kRound * kP1BlockSize;  // staged cells per round
constexpr uint32_t kChunk = kP1Chunk;                    // offsets per reserved chunk (one line)
constexpr uint16_t kNoCell = 0xFFFFu;                    // unused slot (bins hold at most 2^15 cells)
static_assert(kSlices == kBinSlices, "slice count of device and host code");
static_assert(kBinLog2Cells <= 15, "kNoCell must not be a cell offset");

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

// The windows of a thread's shares gid0, gid0 + step, ... (count of them),
// one share at a time.  With HLL every window also feeds the HLL
// registers (add_hash of splitmix64(canonical ntHash)).
template <int L, bool HLL>
struct CellWalker {
  uint64_t gid;     // next share to open
  uint64_t step;    // distance between the thread's shares
  uint32_t left;    // shares not yet opened
  uint32_t w, w1;   // next window of the open share and its end; w == w1: none open
  Kmer<L> fwd, rev;
  BaseStream in;
  NtHash nh;        // HLL only

  __device__ __forceinline__ void init(uint64_t gid0, uint64_t stride, uint32_t count) {
    gid = gid0;
    step = stride;
    left = count;
    w = w1 = 0;
  }
  // Opens the next share with windows, if any (nt: as for take()).
  __device__ __forceinline__ void open(const BinArgs& a, const NtRollTab& nt) {
    while (left) {
      uint64_t r = 0;
      uint32_t w0 = 0, e = 0;
      const bool ok = share_of(a.reads, gid, a.split, a.k, r, w0, e);
      gid += step;
      --left;
      if (ok) {
        const uint64_t g = a.reads.offs[r] + w0;
        const Kmer<L> win = window_bits<L>(a.reads.data, g, a.k);
        fwd = fwd_from_window<L>(win, a.k);
        rev = rc_from_window<L>(win, a.k);
        in.init(a.reads.data, g + a.k);
        if (HLL) {
          nh.f = 0;
          nh.r = 0;
          BaseStream s0;
          s0.init(a.reads.data, g);
          for (uint32_t j = 0; j < a.k; ++j) nt_push(nh, s0.next(), nt);
        }
        w = w0;
        w1 = e;
        return;
      }
    }
  }
  __device__ __forceinline__ bool has() const { return w < w1; }
  // Canonical fold of the next window of the open share (has() must hold);
  // nt holds the ntHash roll terms of this k (in LDS).
  __device__ __forceinline__ uint64_t take(const BinArgs& a, const NtRollTab& nt) {
    const uint64_t key = kmer_lt<L>(rev, fwd) ? kmer_fold<L>(rev) : kmer_fold<L>(fwd);
    if (HLL) {
      uint32_t idx, rho;
      hll_index_rho(splitmix64(nh.canonical()), a.hll.p, idx, rho);
      if (rho > a.hll.regs[idx]) atomicMax(a.hll.regs + idx, rho);
    }
    if (++w < w1) {
      const uint32_t b = in.next();
      // the base leaving the window is the forward k-mer's first (most
      // significant) base, in its top limb since 2k > 64 (L - 1)
      const uint32_t out = (uint32_t)(fwd.a[L - 1] >> ((2u * a.k - 2u) & 63u)) & 3u;
      kmer_roll<L>(fwd, rev, b, a.k);
      if (HLL) nt_roll(nh, out, b, nt);
    }
    return key;
  }
};

__device__ __forceinline__ uint64_t cell_of(const BinArgs& a, uint64_t key, int row) {
  return ((uint64_t)row << a.log2w) | (splitmix64(key ^ cms_seed(row)) & ((1ull << a.log2w) - 1ull));
}

// A bin's current chunk in one LDS word: next free entry (region position,
// clamped to kPosMax, which is past every region's capacity) << 6 | entries
// left in the chunk (0..63).
constexpr uint32_t kPosMax = (1u << 26) - 1u;
static_assert(kChunk == 64, "entries left in a chunk take 6 bits");
__device__ __forceinline__ uint32_t chunk_word(uint64_t pos, uint32_t left) {
  return ((pos < kPosMax ? (uint32_t)pos : kPosMax) << 6) | left;
}

// Exclusive scan of the round counts v[0..n) in place (one block).  Before
// the scan each thread reserves, for the bins it scans, the chunks a count
// needs beyond what is left of the bin's current chunk: nbase[b] is the
// first entry of the new chunks.
__device__ __forceinline__ void scan_and_reserve(uint32_t* v, uint32_t n, uint32_t* s_wave, const uint32_t* cur,
                                                 uint32_t* nbase, uint32_t* __restrict__ cursor, uint32_t sl) {
  const uint32_t per = (n + blockDim.x - 1) / blockDim.x;
  const uint32_t lo = threadIdx.x * per < n ? threadIdx.x * per : n;
  const uint32_t hi = lo + per < n ? lo + per : n;
  uint32_t sum = 0;
  for (uint32_t i = lo; i < hi; ++i) {
    const uint32_t need = v[i];
    const uint32_t lft = cur[i] & 63u;
    sum += need;
    if (need > lft) {
      const uint32_t more = (need - lft + kChunk - 1) / kChunk * kChunk;
      nbase[i] = atomicAdd(cursor + (uint64_t)i * kSlices + sl, more);
    }
  }
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

// Block blk covers thread shares [blk * chunks * blockDim, (blk + 1) * chunks * blockDim),
// or, with a.grid, every a.grid-th block-sized group of shares from blk on.
// LDS: rcur[nbins] (round counts, then run ends), cur[nbins] (the bin's
// current chunk, chunk_word), nbase[nbins] (first entry of the round's new
// chunks), stage[kStage].
template <int L, bool HLL>
__device__ __forceinline__ void bin_scatter_body(const BinArgs& a, uint64_t blk, uint32_t* sh) {
  __shared__ uint32_t s_wave[kScatterBlock / 64];
  const uint32_t nb = a.nbins;
  uint32_t* rcur = sh;
  uint32_t* cur = sh + nb;
  uint32_t* nbase = sh + 2 * nb;
  uint32_t* stage = sh + 3 * nb;
  for (uint32_t i = threadIdx.x; i < 2 * nb; i += blockDim.x) sh[i] = 0;  // rcur, cur (no chunk yet)
  uint64_t gid0, stride;
  uint32_t count;
  if (a.grid) {
    const uint64_t total = a.reads.n_reads * (uint64_t)a.split;
    gid0 = blk * blockDim.x + threadIdx.x;
    stride = (uint64_t)a.grid * blockDim.x;
    count = gid0 < total ? (uint32_t)((total - 1 - gid0) / stride + 1) : 0u;
  } else {
    gid0 = blk * a.chunks * blockDim.x + threadIdx.x;
    stride = blockDim.x;
    count = a.chunks;
  }
  const uint32_t sb = a.sb;
  const uint32_t sl = xcc_slice();
  const uint64_t cap = a.cap;
  const uint32_t smask = (1u << sb) - 1u;
  uint16_t* __restrict__ out = a.data + sl * cap;
  uint32_t* __restrict__ counts = a.counts;
  __shared__ NtRollTab s_nt;
  if (HLL && threadIdx.x < 4) s_nt.fill(threadIdx.x, a.k);
  CellWalker<L, HLL> wk;
  wk.init(gid0, stride, count);
  __syncthreads();
  for (;;) {
    // a round takes up to kRound windows of one share; its cells stay in registers
    if (!wk.has()) wk.open(a, s_nt);
    uint32_t cells[4 * kRound];
    uint32_t n = 0;
#pragma unroll
    for (int j = 0; j < kRound; ++j) {
      if (wk.has()) {
        const uint64_t key = wk.take(a, s_nt);
#pragma unroll
        for (int row = 0; row < 4; ++row) {
          cells[4 * j + row] = (uint32_t)cell_of(a, key, row);
          atomicAdd(rcur + (cells[4 * j + row] >> sb), 1u);
        }
        n = j + 1;
      }
    }
    if (!__syncthreads_or(n)) break;
    scan_and_reserve(rcur, nb, s_wave, cur, nbase, a.cursor, sl);
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
    // rcur[b] is now the end of bin b's run in the stage; rank q of the run
    // goes to the current chunk while it lasts, then to the new chunks
    const uint32_t total = rcur[nb - 1];
    for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
      const uint32_t cell = stage[i];
      const uint32_t b = cell >> sb;
      const uint32_t q = i - (b ? rcur[b - 1] : 0u);
      const uint32_t cw = cur[b], lft = cw & 63u;
      const uint64_t pos = q < lft ? (uint64_t)(cw >> 6) + q : (uint64_t)nbase[b] + (q - lft);
      if (pos < cap) out[(uint64_t)b * (kSlices * cap) + pos] = (uint16_t)(cell & smask);
      else atomicAdd(counts + cell, 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
      const uint32_t need = rcur[b] - (b ? rcur[b - 1] : 0u);
      const uint32_t cw = cur[b], lft = cw & 63u;
      if (need <= lft) {
        cur[b] = chunk_word((uint64_t)(cw >> 6) + need, lft - need);
      } else {
        const uint32_t rest = need - lft;
        const uint32_t more = (rest + kChunk - 1) / kChunk * kChunk;
        cur[b] = chunk_word((uint64_t)nbase[b] + rest, more - rest);
      }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) rcur[b] = 0;
    __syncthreads();
  }
  // mark what is left of each bin's last chunk (a wave per bin)
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  for (uint32_t b = wave; b < nb; b += blockDim.x >> 6) {
    const uint32_t cw = cur[b], lft = cw & 63u;
    for (uint32_t j = lane; j < lft; j += 64u) {
      const uint64_t pos = (uint64_t)(cw >> 6) + j;
      if (pos < cap) out[(uint64_t)b * (kSlices * cap) + pos] = kNoCell;
    }
  }
}

template <int L, bool HLL>
__global__ __launch_bounds__(kScatterBlock, 4) void k_bin_scatter(BinArgs a) {
  extern __shared__ uint32_t sh[];
  bin_scatter_body<L, HLL>(a, blockIdx.x, sh);
}

template <int L, bool HLL>
__global__ __launch_bounds__(kScatterBlock, 4) void k_bin_scatter_m(const BinArgs* __restrict__ args,
                                                                 const uint32_t* __restrict__ blk, uint32_t n_iso) {
  extern __shared__ uint32_t sh[];
  const uint32_t iso = find_iso(blk, n_iso, blockIdx.x);
  bin_scatter_body<L, HLL>(args[iso], blockIdx.x - blk[iso], sh);
}

// Phase A with a buffer per bin (from kBufMinBins to kBufMaxBins bins).
// Each bin holds kBufE offsets (64 bytes) in LDS.  A thread appends its
// round's cells to their bins' buffers with one LDS atomic each.  After a
// barrier a thread per bin lists the full buffers (reserving a new chunk of
// 64 offsets, one line, with one global atomic where the bin has none
// left), and after another barrier 8 lanes per full bin write the buffer into
// the bin's current chunk (8 lanes of 8 bytes a bin).  Cells that found
// their buffer full try again after the flush.  Every store to the bin
// regions thus writes 64 whole bytes, once; a line is written by one
// workgroup only, in two halves.
// After kBufTries placement steps in one round, the cells still waiting are
// added to the table with global atomics (heavily repeated k-mers: exact,
// and no workgroup waits on a bin for long).
// LDS: cnt[nbins] (entries in the buffer; may pass kBufE while cells are
// placed), cur[nbins] (half_word of the current chunk), full[nbins],
// buf[nbins * kBufE] (u16).
constexpr int kBufBlock = (int)kP1BufBlock;
constexpr uint32_t kBufE = 32;
constexpr uint32_t kBufMaxBins = kP1BufMaxBins;
constexpr uint32_t kBufMinBins = kP1BufMinBins;
constexpr int kBufTries = 4;
static_assert(2 * kBufE == kChunk, "a chunk takes two buffer flushes");

// A bin's current chunk in the buffered kernel: its next free entry (a
// multiple of kBufE) / kBufE << 2 | the halves of the chunk still free (0..2).
__device__ __forceinline__ uint32_t half_word(uint32_t pos, uint32_t halves) { return (pos / kBufE) << 2 | halves; }
__device__ __forceinline__ uint32_t half_pos(uint32_t w) { return (w >> 2) * kBufE; }

// With fewer bins the staged scatter's open lines fit the caches better and
// its two blocks per CU hide each other's barriers: it is the faster form
// there (1024 bins: 18.3 against 18.0 Gbases/s for the whole build).
bool bin_buffered(uint32_t nbins) { return nbins >= kBufMinBins && nbins <= kBufMaxBins; }

size_t bin_buffer_lds(uint32_t nbins) { return (size_t)nbins * (3 * sizeof(uint32_t) + kBufE * sizeof(uint16_t)); }

template <int L, bool HLL>
__device__ __forceinline__ void bin_buffer_body(const BinArgs& a, uint64_t blk, uint32_t* sh) {
  // windows per thread per round: as many as the walk state leaves room
  // for in registers (8 for k <= 64, 7 for k <= 96, then 6)
  constexpr int kBufRound = L <= 2 ? 8 : (L == 3 ? 7 : 6);
  __shared__ uint32_t s_nfull;
  const uint32_t nb = a.nbins;
  uint32_t* cnt = sh;
  uint32_t* cur = sh + nb;
  uint32_t* full = sh + 2 * nb;
  uint16_t* buf = reinterpret_cast<uint16_t*>(sh + 3 * nb);
  for (uint32_t i = threadIdx.x; i < 2 * nb; i += blockDim.x) sh[i] = 0;  // cnt, cur (no chunk yet)

  const uint64_t total = a.reads.n_reads * (uint64_t)a.split;
  const uint64_t gid0 = blk * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)a.grid * blockDim.x;
  const uint32_t count = gid0 < total ? (uint32_t)((total - 1 - gid0) / stride + 1) : 0u;
  const uint32_t sb = a.sb;
  const uint32_t sl = xcc_slice();
  const uint64_t cap = a.cap;
  const uint32_t smask = (1u << sb) - 1u;
  uint16_t* __restrict__ out = a.data + sl * cap;
  uint32_t* __restrict__ cursor = a.cursor + sl;
  const uint32_t h = threadIdx.x & 7u;  // a flushing lane's 4 offsets: entries 4h .. 4h+3
  __shared__ NtRollTab s_nt;
  if (HLL && threadIdx.x < 4) s_nt.fill(threadIdx.x, a.k);
  CellWalker<L, HLL> wk;
  wk.init(gid0, stride, count);
  __syncthreads();
  for (;;) {
    if (!wk.has()) wk.open(a, s_nt);
    uint32_t cells[4 * kBufRound] = {};  // zero for windows not taken: the asm below reads every cell
    uint32_t pend = 0;
#pragma unroll
    for (int j = 0; j < kBufRound; ++j) {
      if (wk.has()) {
        const uint64_t key = wk.take(a, s_nt);
#pragma unroll
        for (int row = 0; row < 4; ++row) cells[4 * j + row] = (uint32_t)cell_of(a, key, row);
        pend |= 0xFu << (4 * j);
      }
    }
    if (!__syncthreads_or(pend)) break;
    for (int tries = 0;; ++tries) {
      // values derived from the cells are recomputed in every step rather
      // than kept live across the loop (which would spill)
#pragma unroll
      for (int i = 0; i < 4 * kBufRound; ++i) asm volatile("" : "+v"(cells[i]));
      if (threadIdx.x == 0) s_nfull = 0;  // every thread read it before the last barrier
      if (tries == kBufTries) {
#pragma unroll
        for (int i = 0; i < 4 * kBufRound; ++i)
          if ((pend >> i) & 1u) atomicAdd(a.counts + cells[i], 1u);
        pend = 0;
      }
#pragma unroll
      for (int i = 0; i < 4 * kBufRound; ++i) {
        if ((pend >> i) & 1u) {
          const uint32_t cell = cells[i], b = cell >> sb;
          const uint32_t slot = atomicAdd(cnt + b, 1u);
          if (slot < kBufE) {
            buf[b * kBufE + slot] = (uint16_t)(cell & smask);
            pend &= ~(1u << i);
          }
        }
      }
      __syncthreads();
      // the full bins, a thread per bin: a new chunk where none is left (all
      // reserved at once), then onto the list
      for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
        if (cnt[b] < kBufE) continue;
        if ((cur[b] & 3u) == 0) cur[b] = half_word(atomicAdd(cursor + (uint64_t)b * kSlices, kChunk), 2u);
        full[atomicAdd(&s_nfull, 1u)] = b;
      }
      __syncthreads();
      const uint32_t nf = s_nfull;
      // 8 lanes a bin, 8 bytes each.  A half chunk starts at a multiple of
      // kBufE and cap is a multiple of kChunk, so it lies wholly below cap
      // or wholly past it.
      for (uint32_t f = threadIdx.x >> 3; f < nf; f += blockDim.x >> 3) {
        const uint32_t b = full[f];
        const uint32_t cw = cur[b];
        const uint64_t pos = half_pos(cw);
        const uint64_t e4 = reinterpret_cast<const uint64_t*>(buf + b * kBufE)[h];
        if (pos < cap) {
          reinterpret_cast<uint64_t*>(out + (uint64_t)b * (kSlices * cap) + pos)[h] = e4;
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) atomicAdd(a.counts + ((uint64_t)b << sb) + ((e4 >> (16 * q)) & 0xFFFFu), 1u);
        }
        if (h == 0) {
          cur[b] = cw + (1u << 2) - 1u;  // next half: position + kBufE, one half fewer
          cnt[b] = 0;
        }
      }
      if (!__syncthreads_or(pend)) break;
    }
  }
  // each bin's buffer (fewer than kBufE entries: a full one was flushed in
  // its step), then kNoCell to the end of its chunk (a wave per bin)
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x)
    if (cnt[b] && (cur[b] & 3u) == 0) cur[b] = half_word(atomicAdd(cursor + (uint64_t)b * kSlices, kChunk), 2u);
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  for (uint32_t b = wave; b < nb; b += blockDim.x >> 6) {
    const uint32_t c = cnt[b];
    const uint32_t cw = cur[b];
    const uint32_t left = (cw & 3u) * kBufE;
    const uint64_t pos = half_pos(cw);
    if (lane < left) {
      const uint16_t e = lane < c ? buf[b * kBufE + lane] : kNoCell;
      if (pos + lane < cap) out[(uint64_t)b * (kSlices * cap) + pos + lane] = e;
      else if (e != kNoCell) atomicAdd(a.counts + ((uint64_t)b << sb) + e, 1u);
    }
  }
}

template <int L, bool HLL>
__global__ __launch_bounds__(kBufBlock) void k_bin_buffer(BinArgs a) {
  extern __shared__ uint32_t sh[];
  bin_buffer_body<L, HLL>(a, blockIdx.x, sh);
}

template <int L, bool HLL>
__global__ __launch_bounds__(kBufBlock) void k_bin_buffer_m(const BinArgs* __restrict__ args,
                                                           const uint32_t* __restrict__ blk, uint32_t n_iso) {
  extern __shared__ uint32_t sh[];
  const uint32_t iso = find_iso(blk, n_iso, blockIdx.x);
  bin_buffer_body<L, HLL>(args[iso], blockIdx.x - blk[iso], sh);
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
    // cap is a multiple of 64, so every region starts on a 128-byte line.
    // The count is bound by load latency, so a thread has eight 16-byte
    // loads (64 offsets) in flight before it counts them: load j of a step
    // takes group j * blockDim + thread of 8 offsets (a wave's loads are
    // contiguous), its index clamped to the last whole group (always inside
    // the region) and its offsets skipped when past it.
    const uint16_t* __restrict__ d = a.data + ((uint64_t)b * kSlices + sl) * a.cap;
    const uint32_t n8 = n & ~7u;
    for (uint32_t base = 0; base < n8; base += blockDim.x * 64u) {
      uint4 v[8];
      uint32_t ok = 0;  // loads inside the region
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t q = base + ((uint32_t)j * blockDim.x + threadIdx.x) * 8u;
        v[j] = *reinterpret_cast<const uint4*>(d + (q < n8 ? q : n8 - 8u));
        ok |= (q < n8 ? 1u : 0u) << j;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (!((ok >> j) & 1u)) continue;
        const uint32_t w[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const uint32_t lo = w[t] & 0xFFFFu, hi = w[t] >> 16;
          if (lo != kNoCell) atomicAdd(lds + lo, 1u);
          if (hi != kNoCell) atomicAdd(lds + hi, 1u);
        }
      }
    }
    for (uint32_t i = n8 + threadIdx.x; i < n; i += blockDim.x)
      if (d[i] != kNoCell) atomicAdd(lds + d[i], 1u);
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

size_t bin_scatter_lds(uint32_t nbins) { return ((size_t)nbins * 3 + kStage) * sizeof(uint32_t); }

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
                       reinterpret_cast<const void*>(&k_bin_buffer<1, true>),
                       reinterpret_cast<const void*>(&k_bin_buffer<1, false>),
                       reinterpret_cast<const void*>(&k_bin_buffer<2, true>),
                       reinterpret_cast<const void*>(&k_bin_buffer<2, false>),
                       reinterpret_cast<const void*>(&k_bin_buffer<3, true>),
                       reinterpret_cast<const void*>(&k_bin_buffer<3, false>),
                       reinterpret_cast<const void*>(&k_bin_buffer<4, true>),
                       reinterpret_cast<const void*>(&k_bin_buffer<4, false>),
                       reinterpret_cast<const void*>(&k_bin_buffer_m<1, true>),
                       reinterpret_cast<const void*>(&k_bin_buffer_m<1, false>),
                       reinterpret_cast<const void*>(&k_bin_buffer_m<2, true>),
                       reinterpret_cast<const void*>(&k_bin_buffer_m<2, false>),
                       reinterpret_cast<const void*>(&k_bin_buffer_m<3, true>),
                       reinterpret_cast<const void*>(&k_bin_buffer_m<3, false>),
                       reinterpret_cast<const void*>(&k_bin_buffer_m<4, true>),
                       reinterpret_cast<const void*>(&k_bin_buffer_m<4, false>),
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

uint32_t bin_scatter_resident(uint32_t k, bool hll, uint32_t nbins) {
  if (pass1_lds_attrs() != hipSuccess) return 1;
  const uint32_t L = (2 * k + 63) / 64;
  const bool buffered = bin_buffered(nbins);
  const void* f = nullptr;
#define BR(LL)                                                                                 \
  if (buffered) f = hll ? reinterpret_cast<const void*>(&k_bin_buffer<LL, true>)              \
                        : reinterpret_cast<const void*>(&k_bin_buffer<LL, false>);            \
  else f = hll ? reinterpret_cast<const void*>(&k_bin_scatter<LL, true>)                     \
               : reinterpret_cast<const void*>(&k_bin_scatter<LL, false>)
  switch (L) {
    case 1: BR(1); break;
    case 2: BR(2); break;
    case 3: BR(3); break;
    default: BR(4); break;
  }
#undef BR
  int per = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, buffered ? kBufBlock : kScatterBlock,
                                                   buffered ? bin_buffer_lds(nbins) : bin_scatter_lds(nbins)) !=
          hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 1;
  return (uint32_t)(per > 0 ? per : 1) * (uint32_t)(cus > 0 ? cus : 1);
}

hipError_t launch_bin_pass1(const BinArgs& a, hipStream_t s, hipEvent_t mid) {
  if (a.nbins == 0 || (a.nbins << a.sb) < a.n_cells) return hipErrorInvalidValue;
  if (a.cap % kChunk) return hipErrorInvalidValue;  // regions start on lines (16-byte loads in the count kernel)
  if (hipError_t e = check_bin_pass1(a.nbins, a.sb)) return e;
  if (a.reads.n_reads && a.grid && bin_buffered(a.nbins)) {
    if (bin_buffer_lds(a.nbins) + 1024 > kP2LdsBytes) return hipErrorInvalidConfiguration;
    const uint32_t L = (2 * a.k + 63) / 64;
    const size_t lds = bin_buffer_lds(a.nbins);
    const bool h = a.hll.regs != nullptr;
#define BB(LL)                                                                                           \
  if (h) hipLaunchKernelGGL((k_bin_buffer<LL, true>), dim3(a.grid), dim3(kBufBlock), lds, s, a);        \
  else hipLaunchKernelGGL((k_bin_buffer<LL, false>), dim3(a.grid), dim3(kBufBlock), lds, s, a)
    switch (L) {
      case 1: BB(1); break;
      case 2: BB(2); break;
      case 3: BB(3); break;
      default: BB(4); break;
    }
#undef BB
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  } else if (a.reads.n_reads) {
    const uint32_t L = (2 * a.k + 63) / 64;
    const unsigned grid =
        a.grid ? a.grid : (unsigned)scatter_blocks(a.reads.n_reads * (uint64_t)a.split, a.chunks, kP1BlockSize);
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

hipError_t launch_bin_scatter_multi(const MultiLaunch& m, uint32_t k, bool hll, uint32_t max_nbins, bool buffered,
                                    hipStream_t s) {
  if (!m.total_blocks) return hipSuccess;
  if (hipError_t e = check_bin_pass1(max_nbins, 6)) return e;  // count LDS checked by its launcher
  const uint32_t L = (2 * k + 63) / 64;
  const BinArgs* a = static_cast<const BinArgs*>(m.args);
  if (buffered) {
    if (max_nbins > kBufMaxBins || bin_buffer_lds(max_nbins) + 1024 > kP2LdsBytes) return hipErrorInvalidConfiguration;
    const size_t lds = bin_buffer_lds(max_nbins);
    const dim3 g(m.total_blocks), bl(kBufBlock);
#define BBM(LL)                                                                                    \
  if (hll) hipLaunchKernelGGL((k_bin_buffer_m<LL, true>), g, bl, lds, s, a, m.blk, m.n_iso);      \
  else hipLaunchKernelGGL((k_bin_buffer_m<LL, false>), g, bl, lds, s, a, m.blk, m.n_iso)
    switch (L) {
      case 1: BBM(1); break;
      case 2: BBM(2); break;
      case 3: BBM(3); break;
      default: BBM(4); break;
    }
#undef BBM
    return hipGetLastError();
  }
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
