// fb_pass2_binned.hip -- pass 2 and pass 3 with LDS-privatised solid tables.
//
// Records are grouped by a hash of their canonical k-mer into bins small
// enough that one workgroup holds the bin's distinct keys in LDS
// (src/graph/build.rs:140-328: the accumulation is a set of saturating adds per
// key, so any grouping of the records gives the same result).
//   Gate (k_p2_gate, one launch per Count-Min row): each thread walks its
//     windows and clears, in a 64-bit mask, the window bits whose cell in
//     this row is below the threshold.  One row's threshold bits (width/8
//     bytes: 2 MB at width 2^24) stay resident in every XCD's L2 for a whole
//     launch; the four rows together do not fit one L2, and probing them all
//     per window is bound by the misses.  The launch of the last row
//     (k_p2_count, with the block layout of the scatter kernel) also counts
//     the records per (bin, slice) in the same walk (a window's record is
//     settled one window behind its gate probe); a one-block scan
//     (k_p2_scan) turns the counts into exact region starts, so no region
//     can overflow whatever the skew of the k-mer spectrum.
//   Phase A (k_p2_scatter): block b writes into the region of (bin, slice),
//     slice = b mod 8.  Workgroups are handed to the 8 XCDs in turn, so each
//     region is filled from one XCD and its partly written lines stay in one
//     L2 until they are whole.  The count kernel stores every block's
//     per-bin record counts; a column scan over the blocks of each slice
//     (k_p2_block_offsets) turns them into each block's run starts, so the
//     scatter walks its windows once, writing one record per window that has
//     any: the canonical k-mer's 2k bits and the edge byte, packed densely
//     into rw 64-bit words.
//   Phase B (k_p2_bins): one workgroup per bin builds the bin's table in LDS
//     (write-once keys claimed word by word with compare-and-swap, as the
//     global table does), then emits the solid k-mers (pass 3) directly.  A
//     key that finds no free LDS entry near its home slot goes to the global
//     table; each key is in exactly one of the two, so pass 3 over the global
//     table appends the rest.
//
// The ungated form (P2BinArgs::ungated) drops the per-window gate and its
// walks.  build.rs:253-289 pushes the pair (u, v) when est(u) >= t and
// est(v) >= t, and pass 3 keeps a key when one of its edge counters reaches
// m = min_trans_cov >= t.  Each edge slot e of a canonical k-mer u names one
// neighbour v_e, and every pair occurrence counted on slot e is a distinct
// window of u and a distinct window of v_e.  So when a slot has occ >= t
// pair occurrences, u and v_e both have >= t windows, and the Count-Min
// estimate (never below the count, while the counters' cap is >= t) passes
// both: the slot's gated count equals occ.  When occ < t the slot cannot
// reach m either way.  Hence the mask, and which keys are kept, do not
// depend on the gate at all.  Only cov does: cov(u) is the sum of the
// gated counts over u's slots, and a slot with 0 < occ < t contributes
// only if est(v_e) >= t.  The records therefore cover every window pair,
// and the bin kernel, once per kept key and low slot, rebuilds v_e from
// the key and the slot and probes its four Count-Min rows; failing slots
// leave cov.  (A kept key whose own estimate fails would need every slot
// dropped; it can only be a prior contig's key, whose ten adds per window
// pass any gate t <= 10, the limit the host enforces with prior contigs.)
// Ungated records include every error k-mer, so a bin kernel first counts
// its records into an LDS filter: nf bit levels (level l set at a position
// when l + 1 records hashed there), two positions per key.  A key whose
// two positions do not both reach level nf - 1 has fewer than nf <= m
// records, so no slot of it can reach m: its records are skipped, and the
// tables hold the keys that can be kept (plus the filter's few false
// positives, which pass 3 drops by their mask).
#include "fb_device.hpp"
#include "fb_internal.hpp"
#include "fb_table.hpp"

namespace fb {
using namespace fbk;

constexpr int kP2GateBlock = (int)kP2GateBlockSize;
constexpr int kP2ScatterBlock = (int)kScatterBlockSize;
constexpr int kP2BinBlock = 1024;

// Bin hash of a k-mer's fold: its top bits pick the pass-2 round and bin.
// It only spreads keys over bins (results never depend on it), so one
// xorshift and one multiply do: over the canonical k-mers of a random 5 Mbp
// genome (k = 21, 55, 77) the bins' key counts vary as evenly as under
// splitmix64 (coefficient of variation 0.020 at 2048 bins, the Poisson
// level), and the bin-table slot hash stays as uniform within a bin.
__device__ __forceinline__ uint64_t bin_hash(uint64_t fold) { return (fold ^ (fold >> 29)) * 0x2545f4914f6cdd1dull; }

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

// Record layout: bits [0, 2k) hold the k-mer integer (limbs least significant
// first), bits [2k, 2k+8) the edge byte, and with prior contigs bit 2k+8 the
// contig flag; rw = ceil((2k+8+cflag)/64) words.
template <int L>
__device__ __forceinline__ void pack_record(const Kmer<L>& key, uint32_t info, uint32_t k, uint64_t (&w)[L + 1]) {
  const uint32_t q = (2u * k) >> 6, s = (2u * k) & 63u;
#pragma unroll
  for (int j = 0; j <= L; ++j) {
    uint64_t v = j < L ? key.a[j] : 0ull;
    if ((uint32_t)j == q) v |= (uint64_t)info << s;
    if (s > 55u && (uint32_t)j == q + 1u) v |= (uint64_t)info >> (64u - s);
    w[j] = v;
  }
}
// Record pos of rw words at dst = recs + pos * rw.  The lanes of a store
// write to scattered lines, so the number of stores per lane, not their
// width, sets the cost: 16-byte stores wherever they stay 16-byte aligned
// (every word pair for even rw; for odd rw the pairs from word 0 or from
// word 1, by the parity of pos) and one 8-byte store for the word left over.
template <int L>
__device__ __forceinline__ void store_record(uint64_t* dst, const uint64_t (&w)[L + 1], uint32_t rw, uint64_t pos) {
  if (rw == 1u) {
    dst[0] = w[0];
    return;
  }
  if (!(rw & 1u)) {
#pragma unroll
    for (int j = 0; j + 1 <= L; j += 2)
      if ((uint32_t)j < rw) *reinterpret_cast<ulonglong2*>(dst + j) = make_ulonglong2(w[j], w[j + 1]);
    return;
  }
  const uint32_t odd = (uint32_t)(pos & 1u);
  uint64_t last = w[0];
#pragma unroll
  for (int j = 0; j + 2 <= L; j += 2) {
    if ((uint32_t)j + 2u < rw)
      *reinterpret_cast<ulonglong2*>(dst + j + odd) =
          odd ? make_ulonglong2(w[j + 1], w[j + 2]) : make_ulonglong2(w[j], w[j + 1]);
    if ((uint32_t)j + 2u == rw - 1u) last = w[j + 2];
  }
  if (odd) last = w[0];
  dst[odd ? 0u : rw - 1u] = last;
}
template <int L>
__device__ __forceinline__ void load_record(const uint64_t* __restrict__ src, uint32_t rw, uint64_t (&w)[L + 1]) {
#pragma unroll
  for (int j = 0; j <= L; ++j) w[j] = (uint32_t)j < rw ? src[j] : 0ull;
}
template <int L>
__device__ __forceinline__ void decode_record(const uint64_t (&w)[L + 1], uint32_t k, Kmer<L>& key, uint32_t& info) {
  const uint32_t q = (2u * k) >> 6, s = (2u * k) & 63u;
  uint64_t lo = 0, hi = 0;
#pragma unroll
  for (int j = 0; j <= L; ++j) {
    if ((uint32_t)j == q) lo = w[j];
    if ((uint32_t)j == q + 1u) hi = w[j];
  }
  uint64_t v = lo >> s;
  if (s > 55u) v |= hi << (64u - s);
  info = (uint32_t)v & 0x1ffu;
#pragma unroll
  for (int j = 0; j < L; ++j) key.a[j] = w[j];
  key.a[L - 1] &= top_mask<L>(k);
}

// A fresh read of a workgroup-shared word that other threads update.  The
// pointer must point into LDS: it is cast to the LDS address space, since a
// volatile read through a generic pointer compiles to a flat load, which
// also waits for every global load in flight.
template <class T>
__device__ __forceinline__ T lds_vload(const T* p) {
  typedef __attribute__((address_space(3))) const volatile T* lds_ptr;
  return *(lds_ptr)(p);
}

// Two consecutive 64-bit LDS words from a 16-byte aligned p, fresh, in one
// 16-byte read.
__device__ __forceinline__ void lds_vload_bucket(const unsigned long long* p, unsigned long long (&v)[2]) {
  typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
  typedef __attribute__((address_space(3))) const volatile u64x2* lds_ptr;
  const u64x2 a = *(lds_ptr)(p);
  v[0] = a.x;
  v[1] = a.y;
}

// Bits of every window of the share p0..p1 (p1 - p0 <= 63 pairs).
__device__ __forceinline__ uint64_t all_windows(uint32_t p0, uint32_t p1) {
  return (p1 - p0 >= 63u) ? ~0ull : (2ull << (p1 - p0)) - 1ull;
}

// Reverse complement of the 2k-bit k-mer y: the complement with its 2-bit
// groups reversed within 2k bits.
template <int L>
__device__ __forceinline__ Kmer<L> kmer_revcomp(const Kmer<L>& y, uint32_t k) {
  Kmer<L> c;
#pragma unroll
  for (int i = 0; i < L; ++i) c.a[i] = ~y.a[i];
  c.a[L - 1] &= top_mask<L>(k);
  return fwd_from_window<L>(c, k);
}

// Whether the neighbour of canonical k-mer x through edge slot q passes all
// four Count-Min rows of the gate bitmap.  Slot q < 4 appends base q in x's
// orientation (edge_u of a forward-canonical u, edge_v of a reverse-canonical
// v); slot 4 + b prepends base 3 - b (build.rs:268-275).
template <int L>
__device__ __forceinline__ bool neighbour_passes(const P2BinArgs& a, const Kmer<L>& x, uint32_t q) {
  const uint32_t k = a.k;
  Kmer<L> y;
  if (q < 4u) {
#pragma unroll
    for (int i = L - 1; i > 0; --i) y.a[i] = (x.a[i] << 2) | (x.a[i - 1] >> 62);
    y.a[0] = (x.a[0] << 2) | q;
    y.a[L - 1] &= top_mask<L>(k);
  } else {
#pragma unroll
    for (int i = 0; i < L - 1; ++i) y.a[i] = (x.a[i] >> 2) | (x.a[i + 1] << 62);
    y.a[L - 1] = x.a[L - 1] >> 2;
    const uint32_t pos = 2u * (k - 1u);
    const uint64_t b = (uint64_t)(7u - q) << (pos & 63u);
#pragma unroll
    for (int i = 0; i < L; ++i)
      if ((pos >> 6) == (uint32_t)i) y.a[i] |= b;
  }
  const Kmer<L> r = kmer_revcomp<L>(y, k);
  const uint64_t fold = kmer_lt<L>(r, y) ? kmer_fold<L>(r) : kmer_fold<L>(y);
  const uint64_t wmask = (1ull << a.log2w) - 1ull;
#pragma unroll
  for (int row = 0; row < 4; ++row) {
    const uint64_t c = ((uint64_t)row << a.log2w) | (splitmix64(fold ^ cms_seed(row)) & wmask);
    if (!((a.solid_bits[c >> 5] >> (c & 31u)) & 1u)) return false;
  }
  return true;
}

// Key of 62-bit words wk as limbs.
template <int L, int KW>
__device__ __forceinline__ Kmer<L> unpack62b(const uint64_t (&wk)[KW]) {
  uint64_t limb[4] = {0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < KW; ++j) {
    const int bit = 62 * j, li = bit / 64, off = bit % 64;
    if (li < 4) limb[li] |= wk[j] << off;
    if (off > 2 && li + 1 < 4) limb[li + 1] |= wk[j] >> (64 - off);
  }
  Kmer<L> x;
#pragma unroll
  for (int i = 0; i < L; ++i) x.a[i] = limb[i];
  return x;
}

// What an ungated key's cov loses to the gate: the counts of its slots with
// 0 < count < t whose neighbour fails the gate (edge(q): the slot's count,
// exact below the edge gate m >= t; prior-contig slots hold >= 1000).
template <int L, class Edge>
__device__ __forceinline__ uint32_t gate_loss(const P2BinArgs& a, const Kmer<L>& x, Edge&& edge) {
  uint32_t loss = 0;
#pragma unroll
  for (uint32_t q = 0; q < 8; ++q) {
    const uint32_t v = edge(q);
    if (v != 0u && v < a.t_solid && !neighbour_passes<L>(a, x, q)) loss += v;
  }
  return loss;
}

// LDS record filter of the ungated bin kernel: kP2FiltWords words per level.
constexpr uint32_t kFiltLog2Bits = 16;
static_assert((1u << kFiltLog2Bits) == 32u * kP2FiltWords, "filter level size");
// The two positions of a key of fold f: bits of one multiply of the premixed fold
// (as the bin hash, with another multiplier).  They only decide which keys
// the filter lets through, never a result; on real k-mers (k = 21, 55, 77,
// the benchmark's bins and error rates) this lets through as few error
// keys as a multiply chain over all the key's limbs (0.2-0.5%).
__device__ __forceinline__ void filt_positions(uint64_t f, uint32_t& p0, uint32_t& p1) {
  const uint64_t h = (f ^ (f >> 29)) * 0xbf58476d1ce4e5b9ull;
  p0 = (uint32_t)(h >> (64u - kFiltLog2Bits));
  p1 = (uint32_t)(h >> 24) & ((1u << kFiltLog2Bits) - 1u);
}
// One record at both positions: it sets the lowest level not yet set there
// (an atomicOr that finds the bit set moves on to the next level), so after
// all records level l is set where at least l + 1 records landed.  A prior
// contig's record sets the top level: its key (and any key sharing both
// positions) always passes.
__device__ __forceinline__ void filt_add(uint32_t* f, uint32_t nf, uint64_t fold, bool contig) {
  uint32_t pos[2];
  filt_positions(fold, pos[0], pos[1]);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const uint32_t wd = pos[j] >> 5, bit = 1u << (pos[j] & 31u);
    uint32_t* top = f + (nf - 1u) * kP2FiltWords + wd;
    if (lds_vload(top) & bit) continue;
    if (contig) {
      atomicOr(top, bit);
      continue;
    }
    for (uint32_t l = 0; l < nf; ++l)
      if (!(atomicOr(f + l * kP2FiltWords + wd, bit) & bit)) break;
  }
}
__device__ __forceinline__ bool filt_pass(const uint32_t* f, uint32_t nf, uint64_t fold) {
  uint32_t p0, p1;
  filt_positions(fold, p0, p1);
  const uint32_t* top = f + (nf - 1u) * kP2FiltWords;
  return ((lds_vload(top + (p0 >> 5)) >> (p0 & 31u)) & (lds_vload(top + (p1 >> 5)) >> (p1 & 31u)) & 1u) != 0u;
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

// Share gid of the pass-2 share space: the reads' shares (n_reads * split),
// then the prior contigs' (contigs.n_reads * csplit).  src is the list the
// share's sequence r belongs to.
__device__ __forceinline__ bool p2_share_any(const P2BinArgs& a, uint64_t gid, const DevReads*& src, uint64_t& r,
                                             uint32_t& p0, uint32_t& p1, bool& contig) {
  const uint64_t nr = a.reads.n_reads * (uint64_t)a.split;
  contig = gid >= nr;
  src = contig ? &a.contigs : &a.reads;
  return contig ? p2_share(a.contigs, gid - nr, a.csplit, a.k, r, p0, p1) : p2_share(a.reads, gid, a.split, a.k, r, p0, p1);
}

// Gate of one Count-Min row over the windows p0..p1 of read r (p1 - p0 <=
// 63): clears bit i of m when window p0+i's cell in this row is below the
// threshold.
template <int L>
__device__ __forceinline__ uint64_t gate_row(const P2BinArgs& a, uint64_t r, uint32_t p0, uint32_t p1, uint64_t m,
                                             uint32_t row) {
  const uint32_t k = a.k;
  const uint64_t* __restrict__ data = a.reads.data;
  const uint64_t g = a.reads.offs[r] + p0;
  Kmer<L> win = window_bits<L>(data, g, k);
  Kmer<L> fwd = fwd_from_window<L>(win, k);
  Kmer<L> rev = rc_from_window<L>(win, k);
  ShortStream in;  // p1 - p0 <= 63 bases
  in.init(data, g + k);
  const uint64_t wmask = (1ull << a.log2w) - 1ull;
  const uint64_t rbase = (uint64_t)row << a.log2w;
  const uint64_t seed = cms_seed((int)row);
  const uint32_t last = p1 - p0;
  for (uint32_t i = 0;; ++i) {
    if ((m >> i) & 1ull) {
      const uint64_t fold = kmer_lt<L>(rev, fwd) ? kmer_fold<L>(rev) : kmer_fold<L>(fwd);
      const uint64_t c = rbase | (splitmix64(fold ^ seed) & wmask);
      if (!((a.solid_bits[c >> 5] >> (c & 31u)) & 1u)) m &= ~(1ull << i);
    }
    if (i == last) break;
    kmer_roll<L>(fwd, rev, in.next(), k);
  }
  return m;
}

// Rows 0..2: row 0 starts from all ones, rows 1 and 2 narrow the stored mask.
template <int L>
__device__ __forceinline__ void p2_gate_body(const P2BinArgs& a, uint64_t gid, uint32_t row) {
  uint64_t r = 0;
  uint32_t p0 = 0, p1 = 0;
  if (a.ungated || !p2_share(a.reads, gid, a.split, a.k, r, p0, p1)) return;
  const uint64_t m = row ? a.gate_masks[gid] : ~0ull;
  a.gate_masks[gid] = m ? gate_row<L>(a, r, p0, p1, m, row) : 0ull;
}

template <int L>
__global__ __launch_bounds__(kP2GateBlock) void k_p2_gate(P2BinArgs a, uint32_t row) {
  p2_gate_body<L>(a, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, row);
}

template <int L>
__global__ __launch_bounds__(kP2GateBlock) void k_p2_gate_m(const P2BinArgs* __restrict__ args,
                                                            const uint32_t* __restrict__ blk, uint32_t n_iso,
                                                            uint32_t row) {
  const uint32_t iso = find_iso(blk, n_iso, blockIdx.x);
  p2_gate_body<L>(args[iso], (uint64_t)(blockIdx.x - blk[iso]) * blockDim.x + threadIdx.x, row);
}

// One thread's walk over windows p0..p1 of read r with the gate bits of
// window p0+i at bit i of gates.  cb(key, fold, left, e_left, right, e_right)
// is called for every window with a record (build.rs:253-285).
template <int L, class CB>
__device__ __forceinline__ void p2_walk(const P2BinArgs& a, const DevReads& R, uint64_t r, uint32_t p0, uint32_t p1,
                                        uint64_t gates, CB&& cb) {
  const uint32_t k = a.k;
  const uint64_t* __restrict__ data = R.data;
  const uint64_t g = R.offs[r] + p0;
  Kmer<L> win = window_bits<L>(data, g, k);
  Kmer<L> fwd = fwd_from_window<L>(win, k);
  Kmer<L> rev = rc_from_window<L>(win, k);
  ShortStream in;  // p1 - p0 <= 63 bases
  in.init(data, g + k);
  bool cur_rc = kmer_lt<L>(rev, fwd);
  Kmer<L> cur = cur_rc ? rev : fwd;
  bool cur_solid = gates & 1ull;
  bool left = false;
  uint32_t left_rcp = 0;
  for (uint32_t x = p0;; ++x) {
    bool right = false;
    uint32_t b = 0;
    bool nxt_rc = false, nxt_solid = false;
    Kmer<L> nxt;
    // window x's first base: the forward k-mer's most significant
    const uint32_t first_base = (uint32_t)(fwd.a[L - 1] >> ((2u * k - 2u) & 63u)) & 3u;
    if (x < p1) {
      b = in.next();
      kmer_roll<L>(fwd, rev, b, k);
      nxt_rc = kmer_lt<L>(rev, fwd);
      nxt = nxt_rc ? rev : fwd;
      nxt_solid = (gates >> (x + 1 - p0)) & 1ull;
      right = cur_solid && nxt_solid;
    }
    if (left || right) {
      // edge_v of pair (x-1, x) and edge_u of pair (x, x+1), build.rs:268-275
      const uint32_t e_left = cur_rc ? left_rcp : left_rcp + 4u;
      const uint32_t e_right = cur_rc ? b + 4u : b;
      cb(cur, kmer_fold<L>(cur), left, e_left, right, e_right);
    }
    if (x >= p1) break;
    left = right;
    left_rcp = 3u - first_base;
    cur = nxt;
    cur_rc = nxt_rc;
    cur_solid = nxt_solid;
  }
}

// Bin of a key within its round: the top log2_rounds bits of the bin hash
// pick the round, the next log2_nbins bits the bin.
__device__ __forceinline__ uint32_t p2_bin_of(const P2BinArgs& a, uint64_t fold) {
  return a.log2_nbins ? (uint32_t)((bin_hash(fold) << a.log2_rounds) >> (64u - a.log2_nbins)) : 0u;
}
__device__ __forceinline__ bool p2_in_round(const P2BinArgs& a, uint64_t fold, uint32_t round) {
  return a.log2_rounds == 0 || (uint32_t)(bin_hash(fold) >> (64u - a.log2_rounds)) == round;
}

// Row 3 of the gate and the record count of round 0 in one walk over the
// windows p0..p1 of read r (gate bits m of rows 0..2): window i's row-3 probe
// settles its gate bit, and with it the pair (i-1, i), so window i-1's
// record (it has one when either of its pairs joins two gate-passing
// windows, as p2_walk decides) is counted one window behind.  Returns the
// final mask.
template <int L>
__device__ __forceinline__ uint64_t gate3_count(const P2BinArgs& a, uint64_t r, uint32_t p0, uint32_t p1, uint64_t m,
                                                uint32_t* hist) {
  const uint32_t k = a.k;
  const uint64_t* __restrict__ data = a.reads.data;
  const uint64_t g = a.reads.offs[r] + p0;
  Kmer<L> win = window_bits<L>(data, g, k);
  Kmer<L> fwd = fwd_from_window<L>(win, k);
  Kmer<L> rev = rc_from_window<L>(win, k);
  ShortStream in;  // p1 - p0 <= 63 bases
  in.init(data, g + k);
  const uint64_t wmask = (1ull << a.log2w) - 1ull;
  const uint64_t rbase = 3ull << a.log2w;
  const uint64_t seed = cms_seed(3);
  const uint32_t last = p1 - p0;
  uint64_t prev_fold = 0;
  bool prev_solid = false, prev_left = false;
  auto count = [&](uint64_t fold) {
    if (p2_in_round(a, fold, 0)) atomicAdd(hist + p2_bin_of(a, fold), 1u);
  };
  for (uint32_t i = 0;; ++i) {
    const uint64_t fold = kmer_lt<L>(rev, fwd) ? kmer_fold<L>(rev) : kmer_fold<L>(fwd);
    bool solid = (m >> i) & 1ull;
    if (solid) {
      const uint64_t c = rbase | (splitmix64(fold ^ seed) & wmask);
      if (!((a.solid_bits[c >> 5] >> (c & 31u)) & 1u)) {
        m &= ~(1ull << i);
        solid = false;
      }
    }
    if (i) {
      const bool right = prev_solid && solid;  // pair (i-1, i)
      if (prev_left || right) count(prev_fold);
      prev_left = right;
    }
    prev_solid = solid;
    prev_fold = fold;
    if (i == last) break;
    kmer_roll<L>(fwd, rev, in.next(), k);
  }
  if (prev_left) count(prev_fold);  // window p1: its left pair only
  return m;
}

// Round 0 applies row 3 of the gate and stores the final mask; every round
// then counts its records of the final mask per bin for slice sl.  Block blk
// covers the thread shares of scatter block blk:
// [blk * chunks * blockDim, (blk + 1) * chunks * blockDim).
template <int L>
__device__ __forceinline__ void p2_count_body(const P2BinArgs& a, uint64_t blk, uint32_t* hist, uint32_t round) {
  if ((round >> a.log2_rounds) != 0u) return;  // this isolate has fewer rounds
  for (uint32_t i = threadIdx.x; i < a.nbins; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const uint64_t gid0 = blk * a.chunks * blockDim.x + threadIdx.x;
  for (uint32_t ch = 0; ch < a.chunks; ++ch) {
    const uint64_t gid = gid0 + (uint64_t)ch * blockDim.x;
    uint64_t r = 0;
    uint32_t p0 = 0, p1 = 0;
    const DevReads* src = nullptr;
    bool contig = false;
    if (!p2_share_any(a, gid, src, r, p0, p1, contig)) continue;
    uint64_t m;
    if (a.ungated) {  // every window pair
      m = all_windows(p0, p1);
    } else if (contig) {  // every window of a prior contig (build.rs:198-239)
      m = all_windows(p0, p1);
      if (round == 0) a.gate_masks[gid] = m;
    } else {
      m = a.gate_masks[gid];
      if (round == 0) {
        if (m) m = gate3_count<L>(a, r, p0, p1, m, hist);
        a.gate_masks[gid] = m;
        continue;
      }
    }
    if (m)
      p2_walk<L>(a, *src, r, p0, p1, m, [&](const Kmer<L>&, uint64_t fold, bool, uint32_t, bool, uint32_t) {
        if (p2_in_round(a, fold, round)) atomicAdd(hist + p2_bin_of(a, fold), 1u);
      });
  }
  __syncthreads();
  const uint32_t sl = (uint32_t)(blk & (kSlices - 1u));
  uint32_t* __restrict__ row = a.blk_hist + blk * a.nbins;
  for (uint32_t b = threadIdx.x; b < a.nbins; b += blockDim.x) {
    const uint32_t c = hist[b];
    row[b] = c;
    if (c) atomicAdd(a.runs + (uint64_t)b * kSlices + sl, c);
  }
}

template <int L>
__global__ __launch_bounds__(kP2ScatterBlock) void k_p2_count(P2BinArgs a) {
  extern __shared__ uint32_t sh[];
  p2_count_body<L>(a, blockIdx.x, sh, a.round);
}

template <int L>
__global__ __launch_bounds__(kP2ScatterBlock) void k_p2_count_m(const P2BinArgs* __restrict__ args,
                                                                const uint32_t* __restrict__ blk, uint32_t n_iso,
                                                                uint32_t round) {
  extern __shared__ uint32_t sh[];
  const uint32_t iso = find_iso(blk, n_iso, blockIdx.x);
  p2_count_body<L>(args[iso], blockIdx.x - blk[iso], sh, round);
}

// Exclusive scan of runs[0..n) in place, runs[n] = total (also added to
// *acc, the records of all rounds, when acc is set); one block.
__device__ __forceinline__ void p2_scan_body(uint32_t* runs, uint32_t n, uint32_t* acc) {
  __shared__ uint32_t s_wave[16];
  const uint32_t per = (n + blockDim.x - 1) / blockDim.x;
  const uint32_t lo = threadIdx.x * per < n ? threadIdx.x * per : n;
  const uint32_t hi = lo + per < n ? lo + per : n;
  uint32_t sum = 0;
  for (uint32_t i = lo; i < hi; ++i) sum += runs[i];
  uint32_t incl = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t v = __shfl_up(incl, off);
    if ((threadIdx.x & 63u) >= (uint32_t)off) incl += v;
  }
  if ((threadIdx.x & 63u) == 63u) s_wave[threadIdx.x >> 6] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w) before += s_wave[w];
  uint32_t run = before + incl - sum;
  for (uint32_t i = lo; i < hi; ++i) {
    const uint32_t v = runs[i];
    runs[i] = run;
    run += v;
  }
  if (threadIdx.x == blockDim.x - 1) {
    runs[n] = run;
    if (acc) atomicAdd(acc, run);
  }
}

constexpr int kP2ScanBlock = 1024;
__global__ __launch_bounds__(kP2ScanBlock) void k_p2_scan(uint32_t* runs, uint32_t n, uint32_t* acc) {
  p2_scan_body(runs, n, acc);
}

__global__ __launch_bounds__(kP2ScanBlock) void k_p2_scan_m(const P2BinArgs* __restrict__ args, uint32_t n_iso) {
  const P2BinArgs& a = args[blockIdx.x];
  if (blockIdx.x < n_iso && a.runs) p2_scan_body(a.runs, a.nbins * kSlices, a.status ? a.status + 1 : nullptr);
}

// Thread i = sl * nbins + b: the blocks of slice sl (sl, sl + 8, ...) take
// consecutive runs of bin b's region in block order.
constexpr int kP2OffBlock = 256;

// LDS requested by the scatter kernel: its bin cursors need nbins * 4 bytes,
// but it asks for 96 KB so that one block runs per CU.  Every running block
// keeps a partly written line open in each bin; with 2-4 blocks per CU those
// open lines outgrow the caches and the record stores slow down by a third
// (measured at 8192 bins: 6.6 vs 10.8 ms for 433M windows).
constexpr size_t kP2ScatterLds = 96u * 1024u;
inline size_t p2_scatter_lds(uint32_t nbins) {
  return (size_t)nbins * 4 > kP2ScatterLds ? (size_t)nbins * 4 : kP2ScatterLds;
}
__device__ __forceinline__ void p2_offsets_body(const P2BinArgs& a, uint64_t i) {
  if (i >= (uint64_t)a.nbins * kSlices) return;
  const uint32_t sl = (uint32_t)(i / a.nbins), b = (uint32_t)(i % a.nbins);
  uint32_t run = a.runs[(uint64_t)b * kSlices + sl];
  uint32_t* __restrict__ h = a.blk_hist + b;
  const uint64_t stride = (uint64_t)kSlices * a.nbins;
  uint64_t blk = sl;
  for (; blk + 3 * kSlices < a.n_blocks; blk += 4 * kSlices) {
    const uint32_t t0 = h[blk * a.nbins], t1 = h[blk * a.nbins + stride];
    const uint32_t t2 = h[blk * a.nbins + 2 * stride], t3 = h[blk * a.nbins + 3 * stride];
    h[blk * a.nbins] = run;
    h[blk * a.nbins + stride] = run + t0;
    h[blk * a.nbins + 2 * stride] = run + t0 + t1;
    h[blk * a.nbins + 3 * stride] = run + t0 + t1 + t2;
    run += t0 + t1 + t2 + t3;
  }
  for (; blk < a.n_blocks; blk += kSlices) {
    const uint32_t t = h[blk * a.nbins];
    h[blk * a.nbins] = run;
    run += t;
  }
}

__global__ __launch_bounds__(kP2OffBlock) void k_p2_block_offsets(P2BinArgs a) {
  p2_offsets_body(a, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// blockIdx.y = isolate
__global__ __launch_bounds__(kP2OffBlock) void k_p2_block_offsets_m(const P2BinArgs* __restrict__ args) {
  p2_offsets_body(args[blockIdx.y], (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Block blk covers the same thread shares as count block blk and writes
// into slice blk mod kSlices; workgroups go to the XCDs in turn, so the
// slice is the XCD's.
template <int L, int KW>
__device__ __forceinline__ void p2_scatter_body(const P2BinArgs& a, uint64_t blk, uint32_t* cur, uint32_t round) {
  if ((round >> a.log2_rounds) != 0u) return;
  // absolute start of this block's run in every bin's (bin, slice) region
  const uint32_t* __restrict__ row = a.blk_hist + blk * a.nbins;
  for (uint32_t b = threadIdx.x; b < a.nbins; b += blockDim.x) cur[b] = row[b];
  __syncthreads();
  const uint64_t gid0 = blk * a.chunks * blockDim.x + threadIdx.x;
  const uint64_t cap = a.rec_cap;
  const uint32_t rw = a.rw, k = a.k;
  for (uint32_t ch = 0; ch < a.chunks; ++ch) {
    const uint64_t gid = gid0 + (uint64_t)ch * blockDim.x;
    uint64_t r = 0;
    uint32_t p0 = 0, p1 = 0;
    const DevReads* src = nullptr;
    bool contig = false;
    if (!p2_share_any(a, gid, src, r, p0, p1, contig)) continue;
    const uint64_t gates = a.ungated ? all_windows(p0, p1) : a.gate_masks[gid];
    if (!gates) continue;
    const uint32_t cbit = contig ? 256u : 0u;
    p2_walk<L>(a, *src, r, p0, p1, gates,
               [&](const Kmer<L>& key, uint64_t fold, bool left, uint32_t el, bool right, uint32_t er) {
                 if (!p2_in_round(a, fold, round)) return;
                 const uint32_t pos = atomicAdd(cur + p2_bin_of(a, fold), 1u);
                 if (pos >= cap) {
                   a.status[0] = 1u;  // runs disagree with the counts: the host falls back
                   return;
                 }
                 uint64_t w[L + 1];
                 pack_record<L>(key, (left ? 1u : 0u) | (el << 1) | (right ? 16u : 0u) | (er << 5) | cbit, k, w);
                 store_record<L>(a.recs + (uint64_t)pos * rw, w, rw, pos);
               });
  }
}

template <int L, int KW>
__global__ __launch_bounds__(kP2ScatterBlock) void k_p2_scatter(P2BinArgs a) {
  extern __shared__ uint32_t sh[];
  p2_scatter_body<L, KW>(a, blockIdx.x, sh, a.round);
}

template <int L, int KW>
__global__ __launch_bounds__(kP2ScatterBlock) void k_p2_scatter_m(const P2BinArgs* __restrict__ args,
                                                                  const uint32_t* __restrict__ blk, uint32_t n_iso,
                                                                  uint32_t round) {
  extern __shared__ uint32_t sh[];
  const uint32_t iso = find_iso(blk, n_iso, blockIdx.x);
  p2_scatter_body<L, KW>(args[iso], blockIdx.x - blk[iso], sh, round);
}

// The windows of a thread's shares (as p2_walk, one window per take), so
// that a thread can stop after a few windows and resume after a barrier.
// A window's record is settled before the window is rolled on: its right
// pair depends only on the gate bits, and its right edge on the next base.
template <int L>
struct RecWalker {
  uint64_t gid, step;  // next share to open and the distance between a thread's shares
  uint32_t left;       // shares not yet opened
  uint32_t i, last;    // window i of the open share (0..last); i > last: none open
  uint64_t gates;      // gate bit of window i at bit i
  uint32_t cbit;       // 256 in a prior contig's records
  bool lpair;          // the pair (i-1, i) joins two gate-passing windows
  uint32_t lrcp;       // 3 - first base of window i-1
  Kmer<L> fwd, rev;
  ShortStream in;  // at most 63 bases after a share's first window

  __device__ __forceinline__ void init(uint64_t gid0, uint64_t stride, uint32_t count) {
    gid = gid0;
    step = stride;
    left = count;
    i = 1;
    last = 0;
  }
  __device__ __forceinline__ bool has() const { return i <= last; }
  __device__ __forceinline__ bool alive() const { return i <= last || left != 0u; }
  __device__ __forceinline__ void open(const P2BinArgs& a) {
    while (left) {
      uint64_t r = 0;
      uint32_t p0 = 0, p1 = 0;
      const DevReads* src = nullptr;
      bool contig = false;
      const uint64_t g = gid;
      const bool ok = p2_share_any(a, g, src, r, p0, p1, contig);
      gid += step;
      --left;
      if (!ok) continue;
      gates = a.ungated || contig ? all_windows(p0, p1) : a.gate_masks[g];
      if (!gates) continue;
      const uint64_t w0 = src->offs[r] + p0;
      const Kmer<L> win = window_bits<L>(src->data, w0, a.k);
      fwd = fwd_from_window<L>(win, a.k);
      rev = rc_from_window<L>(win, a.k);
      in.init(src->data, w0 + a.k);
      i = 0;
      last = p1 - p0;
      lpair = false;
      lrcp = 0;
      cbit = contig ? 256u : 0u;
      return;
    }
  }
  // Window i (has() must hold): true when it has a record of this round, in
  // w (pack_record) with its bin; then moves on to window i + 1.
  __device__ __forceinline__ bool take(const P2BinArgs& a, uint32_t round, uint64_t (&w)[L + 1], uint32_t& bin) {
    const uint32_t k = a.k;
    const bool rc = kmer_lt<L>(rev, fwd);
    const bool more = i < last;
    const uint32_t b = more ? in.next() : 0u;
    const bool rpair = more && ((gates >> i) & (gates >> (i + 1u)) & 1ull);
    bool rec = lpair || rpair;
    if (rec) {
      const Kmer<L> key = rc ? rev : fwd;
      const uint64_t fold = kmer_fold<L>(key);
      rec = p2_in_round(a, fold, round);
      if (rec) {
        // edge_v of pair (i-1, i) and edge_u of pair (i, i+1), build.rs:268-275
        const uint32_t el = rc ? lrcp : lrcp + 4u, er = rc ? b + 4u : b;
        pack_record<L>(key, (lpair ? 1u : 0u) | (el << 1) | (rpair ? 16u : 0u) | (er << 5) | cbit, k, w);
        bin = p2_bin_of(a, fold);
      }
    }
    if (more) {
      // the base leaving the window: the forward k-mer's most significant
      lrcp = 3u - ((uint32_t)(fwd.a[L - 1] >> ((2u * k - 2u) & 63u)) & 3u);
      kmer_roll<L>(fwd, rev, b, k);
      lpair = rpair;
    }
    ++i;
    return rec;
  }
};

// Phase A through per-bin LDS buffers (at most kP2BufMaxBins bins, L <= 3):
// every bin holds kP2BufWords words (64 bytes) of records in LDS, and a
// buffer goes out when it holds the records up to the next multiple of E =
// kP2BufWords / rw positions of the block's run in the bin (E records once
// the run is aligned), written by adjacent lanes: one aligned 64-byte store
// (for rw <= 2) covers what took E scattered stores.  The block's records
// per bin are those of the count block with the same index (the same
// shares), so each flush goes to the next positions of the block's exact
// run.  A thread walks kRecRound windows a round and keeps their records in
// registers; it appends each with one LDS atomic on the bin's count, and
// the append that fills a buffer lists the bin.  After a barrier groups of 8
// lanes write the listed buffers out; records that found their buffer full
// try again after the flush, and after kP2BufTries steps in a round the
// rest are stored directly (one store each) at the run's next positions.
// What is left in the buffers goes out at the end.
// LDS: buf[nbins * kP2BufWords] (u64), cnt[nbins], pos[nbins] (next
// position of the block's run), full[nbins].
constexpr int kP2BufTries = 4;
size_t p2_scatter_buf_lds(uint32_t nbins) { return (size_t)nbins * (kP2BufWords * 8u + 12u); }

template <int L>
__device__ __forceinline__ void p2_scatter_buf_body(const P2BinArgs& a, uint64_t blk, uint64_t* smem,
                                                    uint32_t round) {
  if ((round >> a.log2_rounds) != 0u) return;
  constexpr int kRecRound = L == 1 ? 8 : (L == 2 ? 6 : 4);
  // Per step parity: buffers listed full, and whether any thread still holds
  // records / has windows left.  A step reads its parity's words after its
  // first barrier and clears the other parity's, which the step before read
  // before its second barrier and the step after writes after this step's
  // second barrier.
  __shared__ uint32_t s_nfull[2], s_pend[2], s_live[2];
  const uint32_t nb = a.nbins, rw = a.rw, E = kP2BufWords / rw;  // E: 8, 4 or 2
  uint64_t* buf = smem;
  uint32_t* cnt = reinterpret_cast<uint32_t*>(buf + (uint64_t)nb * kP2BufWords);
  uint32_t* pos = cnt + nb;
  uint32_t* full = pos + nb;
  const uint32_t* __restrict__ row = a.blk_hist + blk * nb;
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    cnt[b] = 0u;
    pos[b] = row[b];
  }
  if (threadIdx.x < 2) s_nfull[threadIdx.x] = s_pend[threadIdx.x] = s_live[threadIdx.x] = 0u;
  const uint64_t cap = a.rec_cap;
  uint64_t* __restrict__ recs = a.recs;
  const uint32_t h = threadIdx.x & 7u;
  RecWalker<L> wk;
  wk.init(blk * a.chunks * blockDim.x + threadIdx.x, blockDim.x, a.chunks);
  uint32_t par = 0;
  __syncthreads();
  for (bool live = true; live;) {
    if (!wk.has()) wk.open(a);
    uint64_t w[kRecRound][L + 1];
    uint32_t bin[kRecRound];
    uint32_t pend = 0;
#pragma unroll
    for (int j = 0; j < kRecRound; ++j)
      if (wk.has() && wk.take(a, round, w[j], bin[j])) pend |= 1u << j;
    if (wk.alive()) s_live[par] = 1u;
    for (int tries = 0;; ++tries) {
#pragma unroll
      for (int j = 0; j < kRecRound; ++j) {
        if (!((pend >> j) & 1u)) continue;
        const uint32_t b = bin[j];
        if (tries == kP2BufTries) {
          // no thread appends in this step: the run's next position directly
          const uint32_t p = atomicAdd(pos + b, 1u);
          if (p < cap) store_record<L>(recs + (uint64_t)p * rw, w[j], rw, p);
          else a.status[0] = 1u;
          pend &= ~(1u << j);
          continue;
        }
        // records up to the run's next multiple of E (pos is fixed in this phase)
        const uint32_t lim = E - (pos[b] & (E - 1u));
        const uint32_t slot = atomicAdd(cnt + b, 1u);
        if (slot < lim) {
          uint64_t* d = buf + (uint64_t)b * kP2BufWords + slot * rw;
#pragma unroll
          for (int q = 0; q <= L; ++q)
            if ((uint32_t)q < rw) d[q] = w[j][q];
          if (slot == lim - 1u) full[atomicAdd(s_nfull + par, 1u)] = b;
          pend &= ~(1u << j);
        }
      }
      if (pend) s_pend[par] = 1u;
      __syncthreads();
      const uint32_t nfull = s_nfull[par];
      const bool again = s_pend[par] != 0u;
      live = s_live[par] != 0u || again;
      if (threadIdx.x == 0) s_nfull[par ^ 1u] = s_pend[par ^ 1u] = s_live[par ^ 1u] = 0u;
      for (uint32_t f = threadIdx.x >> 3; f < nfull; f += blockDim.x >> 3) {
        const uint32_t b = full[f];
        const uint32_t p = pos[b], lim = E - (p & (E - 1u));
        if (h < lim * rw) {
          if ((uint64_t)p + lim <= cap) recs[(uint64_t)p * rw + h] = buf[(uint64_t)b * kP2BufWords + h];
          else a.status[0] = 1u;
        }
        if (h == 0) {
          pos[b] = p + lim;
          cnt[b] = 0u;
        }
      }
      par ^= 1u;
      __syncthreads();
      if (!again) break;
      // a thread still holding records keeps its walk where it is; live
      // stays set for the next step
      if (wk.alive()) s_live[par] = 1u;
    }
  }
  // every buffer holds fewer records than its limit (a full one went out in its step)
  for (uint32_t b = threadIdx.x >> 3; b < nb; b += blockDim.x >> 3) {
    const uint32_t c = cnt[b], p = pos[b];
    if (h < c * rw) {
      if ((uint64_t)p + c <= cap) recs[(uint64_t)p * rw + h] = buf[(uint64_t)b * kP2BufWords + h];
      else a.status[0] = 1u;
    }
  }
}

template <int L>
__global__ __launch_bounds__(kP2ScatterBlock) void k_p2_scatter_buf(P2BinArgs a) {
  extern __shared__ uint64_t sh64[];
  p2_scatter_buf_body<L>(a, blockIdx.x, sh64, a.round);
}

// A batch's scatter: an isolate's workgroups take the buffered body when it
// has at most kP2BufMaxBins bins and the direct one otherwise (the choice is
// the workgroup's, so its barriers stay uniform); both write the runs of the
// count walk.
template <int L, int KW>
__global__ __launch_bounds__(kP2ScatterBlock) void k_p2_scatter_buf_m(const P2BinArgs* __restrict__ args,
                                                                     const uint32_t* __restrict__ blk, uint32_t n_iso,
                                                                     uint32_t round) {
  extern __shared__ uint64_t sh64[];
  const uint32_t iso = find_iso(blk, n_iso, blockIdx.x);
  const P2BinArgs& a = args[iso];
  if (a.nbins <= kP2BufMaxBins)
    p2_scatter_buf_body<L>(a, blockIdx.x - blk[iso], sh64, round);
  else
    p2_scatter_body<L, KW>(a, blockIdx.x - blk[iso], reinterpret_cast<uint32_t*>(sh64), round);
}

// Phase A without the count walk (P2BinArgs::region_cap != 0; the ungated
// form, one round, the buffered scatter's bins): a resident grid walks all
// shares, and a workgroup writes its records of bin b into the region of
// (b, its XCD's slice), region_cap records, in chunks of kP2ChunkUnits
// flush units (E records each) that it reserves with one global atomic on
// runs[b * kSlices + slice].  A bin's buffer goes out when it holds a whole
// unit, so every store is one aligned unit.  A step: appends (the one that
// fills a buffer lists the bin), barrier, a new chunk for each listed bin
// that has none left (a thread per bin, so the reservations' round trips
// overlap), barrier, the flushes (8 lanes a bin), barrier.  At the end each
// bin's partial buffer and the rest of its chunk go out with all-zero
// records after it: the bin kernel skips them (a record always has a left
// or a right pair).  A chunk past region_cap raises status[0] and its
// records are dropped; the host then runs the pass again with exact runs.
// LDS: buf[nbins * kP2BufWords] (u64), cl[nbins] (records in the buffer,
// low 16 bits (may pass E while records are placed); units left in the
// current chunk, high 16 bits), pos[nbins] (record index of the next unit),
// full[nbins].
constexpr uint32_t kNoPos = 0xFFFFFFFFu;
template <int L>
__device__ __forceinline__ void p2_scatter_chunk_body(const P2BinArgs& a, uint32_t blk, uint32_t grid,
                                                      uint64_t* smem) {
  constexpr int kRecRound = L == 1 ? 8 : (L == 2 ? 6 : 4);
  __shared__ uint32_t s_nfull[2], s_pend[2], s_live[2];
  __shared__ uint32_t s_recs;
  const uint32_t nb = a.nbins, rw = a.rw, E = kP2BufWords / rw;  // E: 8, 4 or 2
  const uint32_t CR = kP2ChunkUnits * E;                           // records a chunk
  uint64_t* buf = smem;
  uint32_t* cl = reinterpret_cast<uint32_t*>(buf + (uint64_t)nb * kP2BufWords);
  uint32_t* pos = cl + nb;
  uint32_t* full = pos + nb;
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    cl[b] = 0u;
    pos[b] = kNoPos;
  }
  if (threadIdx.x < 2) s_nfull[threadIdx.x] = s_pend[threadIdx.x] = s_live[threadIdx.x] = 0u;
  if (threadIdx.x == 0) s_recs = 0u;
  const uint32_t sl = xcc_slice();
  const uint64_t rcap = a.region_cap;
  uint64_t* __restrict__ recs = a.recs;
  uint32_t* __restrict__ cursor = a.runs + sl;
  const uint32_t h = threadIdx.x & 7u;
  const uint64_t total = a.reads.n_reads * (uint64_t)a.split + a.contigs.n_reads * (uint64_t)a.csplit;
  const uint64_t gid0 = (uint64_t)blk * blockDim.x + threadIdx.x, stride = (uint64_t)grid * blockDim.x;
  RecWalker<L> wk;
  wk.init(gid0, stride, gid0 < total ? (uint32_t)((total - 1 - gid0) / stride + 1) : 0u);
  uint32_t par = 0, mine = 0;
  // a new chunk for bin b (its region in this XCD's slice)
  auto reserve = [&](uint32_t b) {
    const uint32_t q = atomicAdd(cursor + (uint64_t)b * kSlices, CR);
    if ((uint64_t)q + CR <= rcap) {
      pos[b] = (uint32_t)(((uint64_t)b * kSlices + sl) * rcap + q);
    } else {
      pos[b] = kNoPos;  // region full: the host runs the pass again with exact runs
      a.status[0] = 1u;
    }
    cl[b] = (cl[b] & 0xFFFFu) | (kP2ChunkUnits << 16);
  };
  __syncthreads();
  for (bool live = true; live;) {
    if (!wk.has()) wk.open(a);
    uint64_t w[kRecRound][L + 1];
    uint32_t bin[kRecRound];
    uint32_t pend = 0;
#pragma unroll
    for (int j = 0; j < kRecRound; ++j)
      if (wk.has() && wk.take(a, 0u, w[j], bin[j])) pend |= 1u << j;
    mine += __builtin_popcount(pend);
    if (wk.alive()) s_live[par] = 1u;
    for (;;) {
#pragma unroll
      for (int j = 0; j < kRecRound; ++j) {
        if (!((pend >> j) & 1u)) continue;
        const uint32_t b = bin[j];
        const uint32_t slot = atomicAdd(cl + b, 1u) & 0xFFFFu;
        if (slot < E) {
          uint64_t* d = buf + (uint64_t)b * kP2BufWords + slot * rw;
#pragma unroll
          for (int q = 0; q <= L; ++q)
            if ((uint32_t)q < rw) d[q] = w[j][q];
          if (slot == E - 1u) full[atomicAdd(s_nfull + par, 1u)] = b;
          pend &= ~(1u << j);
        }
      }
      if (pend) s_pend[par] = 1u;
      __syncthreads();
      const uint32_t nfull = s_nfull[par];
      const bool again = s_pend[par] != 0u;
      live = s_live[par] != 0u || again;
      if (threadIdx.x == 0) s_nfull[par ^ 1u] = s_pend[par ^ 1u] = s_live[par ^ 1u] = 0u;
      for (uint32_t f = threadIdx.x; f < nfull; f += blockDim.x) {
        const uint32_t b = full[f];
        if ((cl[b] >> 16) == 0u) reserve(b);
      }
      __syncthreads();
      for (uint32_t f = threadIdx.x >> 3; f < nfull; f += blockDim.x >> 3) {
        const uint32_t b = full[f];
        const uint32_t p = pos[b];
        if (h < E * rw && p != kNoPos) recs[(uint64_t)p * rw + h] = buf[(uint64_t)b * kP2BufWords + h];
        if (h == 0) {
          if (p != kNoPos) pos[b] = p + E;
          cl[b] = ((cl[b] >> 16) - 1u) << 16;  // one unit fewer, buffer empty
        }
      }
      par ^= 1u;
      __syncthreads();
      if (!again) break;
      if (wk.alive()) s_live[par] = 1u;
    }
  }
  // every buffer holds fewer than E records: a bin with records and no unit
  // left takes a chunk, then its buffer and zeros fill its current chunk
  if (mine) atomicAdd(&s_recs, mine);
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x)
    if ((cl[b] & 0xFFFFu) && (cl[b] >> 16) == 0u) reserve(b);
  __syncthreads();
  if (threadIdx.x == 0 && s_recs) atomicAdd(a.status + 1, s_recs);
  for (uint32_t b = threadIdx.x >> 3; b < nb; b += blockDim.x >> 3) {
    const uint32_t c = cl[b] & 0xFFFFu, units = cl[b] >> 16, p = pos[b];
    if (p == kNoPos || h >= E * rw) continue;
    for (uint32_t u = 0; u < units; ++u) {
      const uint64_t v = (u == 0 && h < c * rw) ? buf[(uint64_t)b * kP2BufWords + h] : 0ull;
      recs[((uint64_t)p + (uint64_t)u * E) * rw + h] = v;
    }
  }
}

template <int L>
__global__ __launch_bounds__(kP2ScatterBlock) void k_p2_scatter_chunk(P2BinArgs a) {
  extern __shared__ uint64_t sh64[];
  p2_scatter_chunk_body<L>(a, blockIdx.x, gridDim.x, sh64);
}

uint32_t p2_chunk_grid() {
  static uint32_t g = 0;
  if (!g) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess || cus <= 0) {
      (void)hipGetLastError();
      cus = 256;
    }
    g = (uint32_t)cus;
  }
  return g;
}

// Whether phase A of nbins bins per round and k uses the buffered scatter.
bool p2_scatter_buffered(uint32_t k, uint32_t nbins) { return (2u * k + 63u) / 64u <= 3u && nbins <= kP2BufMaxBins; }

// Home slot of a key of fold f inside its bin's table.  It sets only where
// a key is placed (probe lengths), never a result, so one multiply of the
// premixed fold and a fold of the product's halves suffice; its multiplier
// and shift differ from the bin hash's and the filter's, and over the
// canonical k-mers of a random genome (k = 21, 55, 77) the keys of one bin
// spread over the table's buckets as evenly as random keys do.
__device__ __forceinline__ uint32_t lds_slot(uint64_t f, uint32_t cap) {
  const uint64_t h = (f ^ (f >> 31)) * 0xc2b2ae3d27d4eb4full;
  const uint32_t x = (uint32_t)(h >> 32) ^ (uint32_t)h;
  return (uint32_t)(((uint64_t)x * (uint64_t)cap) >> 32);
}

// Finds or claims the LDS entry of key within kLdsMaxProbe slots of its
// home bucket (the bucket of slot home); -1 when all of them hold other keys.  Slots are probed in
// aligned buckets of kLdsBucket: one 16-byte read brings a bucket's first
// key words.  Buckets of two were measured faster than of four (bin kernel
// 4.61/5.11/3.99 -> 4.41/4.94/3.84 ms at k = 21/55/77) and of eight (5.6 /
// 6.0 / 4.9 ms): the reads of a probe cost more than the extra round trips
// of the few lanes that go past their home bucket.  Entries
// are written once, so a key that finds its probe window full finds it
// full at every later lookup too and lives in the global table for the
// whole pass, and a key claimed in the window is found there by every
// later lookup.  The bound keeps a crowded bin from scanning its whole
// table for every new key.  cap is a multiple of kLdsBucket.
constexpr uint32_t kLdsMaxProbe = 128;
constexpr uint32_t kLdsBucket = 2;
template <int KW>
__device__ __forceinline__ int lds_find(unsigned long long* keys, uint32_t cap, const uint64_t (&w)[KW], uint32_t home) {
  uint32_t s = home & ~(kLdsBucket - 1u);
  const uint32_t buckets = (cap < kLdsMaxProbe ? cap : kLdsMaxProbe) / kLdsBucket;
  for (uint32_t b = 0; b < buckets; ++b) {
    unsigned long long v0[kLdsBucket];
    lds_vload_bucket(keys + s, v0);
#pragma unroll
    for (uint32_t i = 0; i < kLdsBucket; ++i) {
      // another key's slot: its first word is written once and never changes
      if (v0[i] != w[0] && v0[i] != kEmpty) continue;
      // the key's slot or a free one: settled word by word, a free word
      // claimed by compare-and-swap (which returns the winner's word)
      bool match = true;
#pragma unroll
      for (int j = 0; j < KW; ++j) {
        if (!match) break;
        unsigned long long* p = keys + (uint64_t)j * cap + s + i;
        unsigned long long v = j == 0 ? v0[i] : lds_vload(p);
        if (v == kEmpty) {
          const unsigned long long old = atomicCAS(p, (unsigned long long)kEmpty, (unsigned long long)w[j]);
          v = old == kEmpty ? (unsigned long long)w[j] : old;
        }
        match = (v == w[j]);
      }
      if (match) return (int)(s + i);
    }
    s += kLdsBucket;
    if (s >= cap) s = 0;
  }
  return -1;
}

// f(record) for records [0, n) of src, thread t taking t, t + blockDim, ...
// The loop is software-pipelined: the loads of a thread's next group of
// records go out before it works on the current group, so the load latency
// (the loop's bound: its LDS work is short) overlaps that work.  Loads are
// unconditional (indices clamped to the last record, L + 1 words each), so
// the compiler waits for exactly the group it works on; src must be
// readable one word past record n - 1 (the record buffers carry a pad).
template <int L, class F>
__device__ __forceinline__ void for_records(const uint64_t* __restrict__ src, uint32_t n, uint32_t rw, F&& f) {
  constexpr int U = L <= 2 ? 4 : 2;  // records per group (registers)
  if (n == 0) return;
  const uint32_t step = blockDim.x;
  auto load_group = [&](uint32_t i0, uint64_t (&r)[U][L + 1]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      uint32_t i = i0 + (uint32_t)u * step;
      i = i < n ? i : n - 1u;
      const uint64_t* p = src + (uint64_t)i * rw;
#pragma unroll
      for (int j = 0; j <= L; ++j) r[u][j] = p[j];
    }
  };
  uint64_t nx[U][L + 1];
  load_group(threadIdx.x, nx);
  for (uint32_t i0 = threadIdx.x; i0 < n; i0 += U * step) {
    uint64_t cur[U][L + 1];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int j = 0; j <= L; ++j) cur[u][j] = nx[u][j];
    load_group(i0 + U * step, nx);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (rw == (uint32_t)L) cur[u][L] = 0ull;  // the word past a record of L words
      if (i0 + (uint32_t)u * step < n) f(cur[u]);
    }
  }
}

// One record's sides into entry e of the LDS table, or into the global table
// when e < 0 (the key's probe window in LDS is full).
template <int KW>
__device__ __forceinline__ void p2_table_add(const P2BinArgs& a, uint32_t* cov, uint32_t* edges, int e,
                                             const uint64_t (&w)[KW], uint32_t info) {
  const uint32_t gate = a.min_trans_cov;
  const bool left = info & 1u, right = (info >> 4) & 1u;
  const uint32_t el = (info >> 1) & 7u, er = (info >> 5) & 7u;
  // a prior contig's side adds cov 100 and edge 1000 (build.rs:219-231)
  const bool contig = (info >> 8) & 1u;
  const uint32_t cw = contig ? 100u : 1u;
  if (e >= 0) {
    atomicAdd(cov + e, cw * ((uint32_t)left + (uint32_t)right));
    // u8 edge counters, four per word.  Only `value >= gate` (gate <= 255)
    // and values below the gate are ever read, so a read's side adds 1 by
    // compare and swap while its byte is below the gate (exact however the
    // adds interleave, and a byte below 255 cannot carry into the next); a
    // prior contig's side sets its byte to 255, at or past any gate.  Only a
    // key's first adds on a slot reach the compare and swap: later ones
    // find the gate reached and only read.
    for (int side = 0; side < 2; ++side) {
      if (!(side ? right : left)) continue;
      const uint32_t ed = side ? er : el;
      uint32_t* ew = edges + 2u * (uint32_t)e + (ed >> 2);
      const uint32_t sh = (ed & 3u) * 8u;
      if (contig) {
        atomicOr(ew, 0xffu << sh);
      } else {
        uint32_t old = lds_vload(ew);
        while (((old >> sh) & 0xffu) < gate) {
          const uint32_t seen = old;
          old = atomicCAS(ew, seen, seen + (1u << sh));
          if (old == seen) break;
        }
      }
    }
  } else {
    // no LDS entry: this key lives in the global table for the whole pass
    uint32_t* c = table_find_global<KW>(a.table, w, slot_hash_words<KW>(w));
    if (c) {
      atomicAdd(c, cw * ((uint32_t)left + (uint32_t)right));
      if (contig) {
        if (left) atomicAdd(c + 2 + el, 1000u);
        if (right) atomicAdd(c + 2 + er, 1000u);
      } else {
        if (left && c[2 + el] < gate) atomicAdd(c + 2 + el, 1u);
        if (right && c[2 + er] < gate) atomicAdd(c + 2 + er, 1u);
      }
    }
  }
}

// With the record filter (nf levels) a bin's records are read twice: the
// first pass counts every record into the filter, the second tables the
// records whose key passes it (and every prior contig's record).  Both
// passes keep the lanes of a wave on the same short path; a single pass
// that tables keys as their filter count completes was measured slower, its
// per-record branches (new keys, parking) being taken by some lane of
// nearly every wave.
template <int L, int KW>
__device__ __forceinline__ void p2_bins_body(const P2BinArgs& a, uint32_t bin, unsigned char* smem, uint32_t round) {
  if ((round >> a.log2_rounds) != 0u) return;  // this isolate has fewer rounds
  const uint32_t cap = a.lds_cap;
  const uint32_t nf = a.filt;                                                   // filter levels (0: none)
  uint32_t* filt = reinterpret_cast<uint32_t*>(smem);                           // nf * kP2FiltWords
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(filt + nf * kP2FiltWords);  // KW * cap
  uint32_t* cov = reinterpret_cast<uint32_t*>(keys + (uint64_t)KW * cap);       // cap
  uint32_t* edges = cov + cap;                                                  // 2 * cap (u8 quads)
  __shared__ uint32_t s_wave[kP2BinBlock / 64];
  __shared__ uint32_t s_base;
  __shared__ unsigned long long s_stat[2];
  for (uint32_t i = threadIdx.x; i < nf * kP2FiltWords; i += blockDim.x) filt[i] = 0u;
  for (uint32_t i = threadIdx.x; i < KW * cap; i += blockDim.x) keys[i] = kEmpty;
  for (uint32_t i = threadIdx.x; i < 3 * cap; i += blockDim.x) cov[i] = 0u;  // cov and edges
  if (threadIdx.x < 2) s_stat[threadIdx.x] = 0ull;
  __syncthreads();
  const uint32_t rw = a.rw;
  const uint32_t gate = a.min_trans_cov;
  // f(record) over the bin's records: with exact runs the bin's kSlices
  // regions are adjacent, [start of its first slice, start of the next bin's
  // first slice); with chunked regions each slice's region holds its
  // reserved chunks, whose all-zero tail records are skipped
  auto each_record = [&](auto&& f) {
    if (a.region_cap) {
      for (uint32_t sl = 0; sl < kSlices; ++sl) {
        const uint64_t r = (uint64_t)bin * kSlices + sl;
        const uint64_t fill = a.runs[r];
        const uint32_t n = (uint32_t)(fill < a.region_cap ? fill : a.region_cap);
        for_records<L>(a.recs + r * a.region_cap * rw, n, rw, [&](const uint64_t (&cur)[L + 1]) {
          Kmer<L> key;
          uint32_t info;
          decode_record<L>(cur, a.k, key, info);
          if (info & 0x11u) f(key, info);  // a left or a right pair: not a tail record
        });
      }
    } else {
      const uint64_t e0 = a.runs[(uint64_t)bin * kSlices];
      uint64_t e1 = a.runs[(uint64_t)(bin + 1) * kSlices];
      if (e1 > a.rec_cap) e1 = a.rec_cap;
      const uint32_t n = e1 > e0 ? (uint32_t)(e1 - e0) : 0u;
      for_records<L>(a.recs + e0 * rw, n, rw, [&](const uint64_t (&cur)[L + 1]) {
        Kmer<L> key;
        uint32_t info;
        decode_record<L>(cur, a.k, key, info);
        f(key, info);
      });
    }
  };
  if (nf) {
    each_record([&](const Kmer<L>& key, uint32_t info) { filt_add(filt, nf, kmer_fold<L>(key), (info >> 8) & 1u); });
    __syncthreads();
  }
  each_record([&](const Kmer<L>& key, uint32_t info) {
    const uint64_t fold = kmer_fold<L>(key);
    if (nf && !((info >> 8) & 1u) && !filt_pass(filt, nf, fold)) return;  // fewer than nf records
    uint64_t w[KW];
    pack62b<L, KW>(key, w);
    p2_table_add<KW>(a, cov, edges, lds_find<KW>(keys, cap, w, lds_slot(fold, cap)), w, info);
  });
  __syncthreads();
  // pass 3 on the LDS table: count, one atomic per block, write
  const uint32_t per = (cap + blockDim.x - 1) / blockDim.x;
  const uint32_t lo = threadIdx.x * per, hi = lo + per < cap ? lo + per : cap;
  auto keep_of = [&](uint32_t e, uint32_t& mask, uint32_t& cv) -> bool {
    if (keys[e] == kEmpty) return false;
    mask = 0;
#pragma unroll
    for (uint32_t q = 0; q < 8; ++q) {
      const uint32_t v = (edges[2u * e + (q >> 2)] >> ((q & 3u) * 8u)) & 0xffu;
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
    if (keys[e] == kEmpty) continue;
    ++held;
    raw += cov[e];
    uint32_t m, cv;
    if (!keep_of(e, m, cv)) continue;
    if (a.ungated) {  // the gate's share of cov, once per kept key (this thread's entries only)
      uint64_t wk[KW];
#pragma unroll
      for (int j = 0; j < KW; ++j) wk[j] = keys[(uint64_t)j * cap + e];
      cov[e] -= gate_loss<L>(a, unpack62b<L, KW>(wk),
                             [&](uint32_t q) { return (edges[2u * e + (q >> 2)] >> ((q & 3u) * 8u)) & 0xffu; });
    }
    ++mine;
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
    if (pos >= a.out_cap) break;  // the host sees out_count > out_cap and re-runs the isolate
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

template <int L, int KW>
__global__ __launch_bounds__(kP2BinBlock) void k_p2_bins(P2BinArgs a) {
  extern __shared__ unsigned char smem[];
  p2_bins_body<L, KW>(a, blockIdx.x, smem, a.round);
}

template <int L, int KW>
__global__ __launch_bounds__(kP2BinBlock) void k_p2_bins_m(const P2BinArgs* __restrict__ args,
                                                           const uint32_t* __restrict__ blk, uint32_t n_iso,
                                                           uint32_t round) {
  extern __shared__ unsigned char smem[];
  const uint32_t iso = find_iso(blk, n_iso, blockIdx.x);
  p2_bins_body<L, KW>(args[iso], blockIdx.x - blk[iso], smem, round);
}

// The ungated form's correction of the global (spill) table before its pass
// 3: every kept key's cov loses its slots whose neighbour fails the gate.
template <int L, int KW>
__device__ __forceinline__ void p2_spill_gate_body(const P2BinArgs& a, uint64_t i0, uint64_t step) {
  const DevSolidTable& t = a.table;
  for (uint64_t i = i0; i < t.cap; i += step) {
    uint64_t* s = t.slots + i * t.stride;
    uint64_t wk[KW];
#pragma unroll
    for (int j = 0; j < KW; ++j) wk[j] = s[j];
    if (wk[0] == kEmpty) continue;
    uint32_t* c = reinterpret_cast<uint32_t*>(s + KW);  // cov, pad, 8 edges
    bool keep = false;
#pragma unroll
    for (int q = 0; q < 8; ++q) keep |= c[2 + q] >= a.min_trans_cov;
    if (!keep) continue;
    c[0] -= gate_loss<L>(a, unpack62b<L, KW>(wk), [&](uint32_t q) { return c[2 + q]; });
  }
}
template <int L, int KW>
__global__ __launch_bounds__(256) void k_p2_spill_gate(P2BinArgs a) {
  p2_spill_gate_body<L, KW>(a, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, (uint64_t)gridDim.x * blockDim.x);
}
template <int L, int KW>
__global__ __launch_bounds__(256) void k_p2_spill_gate_m(const P2BinArgs* __restrict__ args) {
  const P2BinArgs& a = args[blockIdx.y];
  if (a.ungated)
    p2_spill_gate_body<L, KW>(a, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, (uint64_t)gridDim.x * blockDim.x);
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

template <int L, int KW>
static hipError_t set_pair_attr() {
  hipError_t e = set_attr(reinterpret_cast<const void*>(&k_p2_count<L>));
  if (e == hipSuccess) e = set_attr(reinterpret_cast<const void*>(&k_p2_count_m<L>));
  if (e == hipSuccess) e = set_attr(reinterpret_cast<const void*>(&k_p2_bins<L, KW>));
  if (e == hipSuccess) e = set_attr(reinterpret_cast<const void*>(&k_p2_bins_m<L, KW>));
  if (e == hipSuccess) e = set_attr(reinterpret_cast<const void*>(&k_p2_scatter<L, KW>));
  if (e == hipSuccess) e = set_attr(reinterpret_cast<const void*>(&k_p2_scatter_m<L, KW>));
  if constexpr (L <= 3) {
    if (e == hipSuccess) e = set_attr(reinterpret_cast<const void*>(&k_p2_scatter_buf<L>));
    if (e == hipSuccess) e = set_attr(reinterpret_cast<const void*>(&k_p2_scatter_buf_m<L, KW>));
    if (e == hipSuccess) e = set_attr(reinterpret_cast<const void*>(&k_p2_scatter_chunk<L>));
  }
  return e;
}

static uint32_t lds_budget() {
  if (!g_lds_bytes) {
    hipError_t e = set_pair_attr<1, 1>();
    if (e == hipSuccess) e = set_pair_attr<1, 2>();
    if (e == hipSuccess) e = set_pair_attr<2, 2>();
    if (e == hipSuccess) e = set_pair_attr<2, 3>();
    if (e == hipSuccess) e = set_pair_attr<3, 3>();
    if (e == hipSuccess) e = set_pair_attr<3, 4>();
    if (e == hipSuccess) e = set_pair_attr<4, 4>();
    if (e == hipSuccess) e = set_pair_attr<4, 5>();
    if (e != hipSuccess) (void)hipGetLastError();
    g_lds_bytes = e == hipSuccess ? kP2LdsBytes : 65536u;
  }
  return g_lds_bytes;
}

uint32_t p2_max_bins() { return (lds_budget() - 1024u) / 4u >= 16384u ? 16384u : 8192u; }

uint32_t p2_lds_cap(uint32_t kw, uint32_t filt) {
  // keys + cov + 8 u8 edges per entry within the workgroup's LDS, after the filter
  const uint32_t entry = p2_entry_bytes(kw);
  return ((lds_budget() - 1024u - filt * kP2FiltWords * 4u) / entry) & ~(kLdsBucket - 1u);
}
size_t p2_lds_bytes(uint32_t kw, uint32_t cap, uint32_t filt) {
  return (size_t)filt * kP2FiltWords * 4u + (size_t)cap * p2_entry_bytes(kw);
}

// Calls F<L, KW>() for the template pair of k.
template <class F>
static hipError_t dispatch_lkw(uint32_t k, F&& f) {
  const uint32_t L = (2 * k + 63) / 64;
  const uint32_t KW = (k + 30) / 31;
  if (L == 1 && KW == 1) return f.template operator()<1, 1>();
  if (L == 1) return f.template operator()<1, 2>();
  if (L == 2 && KW == 2) return f.template operator()<2, 2>();
  if (L == 2) return f.template operator()<2, 3>();
  if (L == 3 && KW == 3) return f.template operator()<3, 3>();
  if (L == 3) return f.template operator()<3, 4>();
  if (KW == 4) return f.template operator()<4, 4>();
  return f.template operator()<4, 5>();
}

struct P2Launch {
  const P2BinArgs& a;
  uint64_t threads;       // all shares: reads, then prior contigs
  uint64_t read_threads;  // the reads' shares (the gate kernels' grid)
  size_t scat_lds, lds;
  bool buffered;
  hipStream_t s;
  hipEvent_t gated, scattered;
  template <int L, int KW>
  hipError_t operator()() {
    if (read_threads && !a.ungated) {
      const unsigned ggrid = (unsigned)((read_threads + kP2GateBlock - 1) / kP2GateBlock);
      for (uint32_t row = 0; row < 3; ++row)
        hipLaunchKernelGGL((k_p2_gate<L>), dim3(ggrid), dim3(kP2GateBlock), 0, s, a, row);
    }
    const uint32_t rounds = 1u << a.log2_rounds;
    for (uint32_t round = 0; round < rounds; ++round) {
      P2BinArgs ar = a;
      ar.round = round;
      if (hipError_t e = launch_fill_u32(a.runs, (uint64_t)a.nbins * kSlices + 1, 0u, s)) return e;
      if constexpr (L <= 3) {
        if (a.region_cap) {  // no count walk: chunks reserved as the records go out
          if (gated)
            if (hipError_t e = hipEventRecord(gated, s)) return e;
          if (threads)
            hipLaunchKernelGGL((k_p2_scatter_chunk<L>), dim3(p2_chunk_grid()), dim3(kP2ScatterBlock), scat_lds, s,
                               ar);
          if (hipError_t e = hipGetLastError()) return e;
          if (scattered)
            if (hipError_t e = hipEventRecord(scattered, s)) return e;
          hipLaunchKernelGGL((k_p2_bins<L, KW>), dim3(a.nbins), dim3(kP2BinBlock), lds, s, ar);
          if (hipError_t e = hipGetLastError()) return e;
          continue;
        }
      }
      if (threads)
        hipLaunchKernelGGL((k_p2_count<L>), dim3(scatter_blocks(threads, a.chunks)), dim3(kP2ScatterBlock),
                           (size_t)a.nbins * 4, s, ar);
      hipLaunchKernelGGL(k_p2_scan, dim3(1), dim3(kP2ScanBlock), 0, s, a.runs, a.nbins * kSlices,
                         a.status ? a.status + 1 : nullptr);
      if (threads) {
        const unsigned ogrid = (unsigned)(((uint64_t)a.nbins * kSlices + kP2OffBlock - 1) / kP2OffBlock);
        hipLaunchKernelGGL(k_p2_block_offsets, dim3(ogrid), dim3(kP2OffBlock), 0, s, ar);
      }
      if (gated && round == 0)
        if (hipError_t e = hipEventRecord(gated, s)) return e;
      if (threads) {
        const unsigned grid = (unsigned)scatter_blocks(threads, a.chunks);
        if constexpr (L <= 3) {
          if (buffered)
            hipLaunchKernelGGL((k_p2_scatter_buf<L>), dim3(grid), dim3(kP2ScatterBlock), scat_lds, s, ar);
          else
            hipLaunchKernelGGL((k_p2_scatter<L, KW>), dim3(grid), dim3(kP2ScatterBlock), scat_lds, s, ar);
        } else {
          hipLaunchKernelGGL((k_p2_scatter<L, KW>), dim3(grid), dim3(kP2ScatterBlock), scat_lds, s, ar);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
      }
      if (scattered && round == 0)
        if (hipError_t e = hipEventRecord(scattered, s)) return e;
      hipLaunchKernelGGL((k_p2_bins<L, KW>), dim3(a.nbins), dim3(kP2BinBlock), lds, s, ar);
      if (hipError_t e = hipGetLastError()) return e;
    }
    if (a.ungated) {
      const uint64_t g = (a.table.cap + 255) / 256;
      hipLaunchKernelGGL((k_p2_spill_gate<L, KW>), dim3((unsigned)(g < 4096 ? (g ? g : 1) : 4096)), dim3(256), 0, s,
                         a);
      if (hipError_t e = hipGetLastError()) return e;
    }
    return hipSuccess;
  }
};

hipError_t launch_p2_binned(const P2BinArgs& a, hipStream_t s, hipEvent_t gated, hipEvent_t scattered) {
  const uint32_t KW = (a.k + 30) / 31;
  const bool buffered = p2_scatter_buffered(a.k, a.nbins);
  const size_t scat_lds = buffered ? p2_scatter_buf_lds(a.nbins) : p2_scatter_lds(a.nbins);
  const size_t lds = p2_lds_bytes(KW, a.lds_cap, a.filt);
  if (lds + 1024 > lds_budget() || scat_lds + 1024 > lds_budget()) return hipErrorInvalidConfiguration;
  if (a.lds_cap % kLdsBucket || a.lds_cap == 0) return hipErrorInvalidValue;
  if (a.cflag > 1u || (a.contigs.n_reads && !a.cflag)) return hipErrorInvalidValue;
  if (a.filt > kP2MaxFilt || (a.ungated && (!a.solid_bits || a.t_solid > a.min_trans_cov)))
    return hipErrorInvalidValue;
  if (a.rw != (2 * a.k + 8 + a.cflag + 63) / 64) return hipErrorInvalidValue;
  const uint64_t read_threads = a.reads.n_reads * (uint64_t)a.split;
  const uint64_t threads = read_threads + a.contigs.n_reads * (uint64_t)a.csplit;
  if (a.n_blocks != scatter_blocks(threads, a.chunks)) return hipErrorInvalidValue;
  if (a.log2_rounds > kP2MaxLog2Rounds || a.nbins != (1u << a.log2_nbins)) return hipErrorInvalidValue;
  if (a.region_cap) {
    const uint32_t unit = kP2BufWords / a.rw;
    if (!buffered || !a.ungated || a.log2_rounds || a.region_cap % (kP2ChunkUnits * unit) ||
        (uint64_t)a.nbins * kSlices * a.region_cap >= (1ull << 32))
      return hipErrorInvalidValue;
  }
  return dispatch_lkw(a.k, P2Launch{a, threads, read_threads, scat_lds, lds, buffered, s, gated, scattered});
}

struct P2LaunchMulti {
  const MultiLaunch &gate, &scatter, &bins;
  const FillRange* runs_fill;
  uint32_t n_fill, max_rounds;
  size_t count_lds, scat_lds, lds;
  bool buffered;
  hipStream_t s;
  template <int L, int KW>
  hipError_t operator()() {
    const P2BinArgs* ga = static_cast<const P2BinArgs*>(gate.args);
    if (gate.total_blocks) {
      for (uint32_t row = 0; row < 3; ++row)
        hipLaunchKernelGGL((k_p2_gate_m<L>), dim3(gate.total_blocks), dim3(kP2GateBlock), 0, s, ga, gate.blk,
                           gate.n_iso, row);
    }
    for (uint32_t round = 0; round < max_rounds; ++round) {
      if (hipError_t e = launch_fill_ranges(runs_fill, n_fill, s)) return e;
      if (scatter.total_blocks)
        hipLaunchKernelGGL((k_p2_count_m<L>), dim3(scatter.total_blocks), dim3(kP2ScatterBlock), count_lds, s,
                           static_cast<const P2BinArgs*>(scatter.args), scatter.blk, scatter.n_iso, round);
      if (gate.n_iso) {
        hipLaunchKernelGGL(k_p2_scan_m, dim3(gate.n_iso), dim3(kP2ScanBlock), 0, s, ga, gate.n_iso);
        const unsigned ogrid = (unsigned)((count_lds / 4 * kSlices + kP2OffBlock - 1) / kP2OffBlock);
        hipLaunchKernelGGL(k_p2_block_offsets_m, dim3(ogrid, gate.n_iso), dim3(kP2OffBlock), 0, s, ga);
      }
      if (scatter.total_blocks) {
        const P2BinArgs* sa = static_cast<const P2BinArgs*>(scatter.args);
        if constexpr (L <= 3) {
          if (buffered)
            hipLaunchKernelGGL((k_p2_scatter_buf_m<L, KW>), dim3(scatter.total_blocks), dim3(kP2ScatterBlock),
                               scat_lds, s, sa, scatter.blk, scatter.n_iso, round);
          else
            hipLaunchKernelGGL((k_p2_scatter_m<L, KW>), dim3(scatter.total_blocks), dim3(kP2ScatterBlock), scat_lds,
                               s, sa, scatter.blk, scatter.n_iso, round);
        } else {
          hipLaunchKernelGGL((k_p2_scatter_m<L, KW>), dim3(scatter.total_blocks), dim3(kP2ScatterBlock), scat_lds, s,
                             sa, scatter.blk, scatter.n_iso, round);
        }
      }
      if (hipError_t e = hipGetLastError()) return e;
      if (bins.total_blocks)
        hipLaunchKernelGGL((k_p2_bins_m<L, KW>), dim3(bins.total_blocks), dim3(kP2BinBlock), lds, s,
                           static_cast<const P2BinArgs*>(bins.args), bins.blk, bins.n_iso, round);
      if (hipError_t e = hipGetLastError()) return e;
    }
    if (bins.n_iso) {  // the spill tables' gate correction (each ungated isolate's)
      hipLaunchKernelGGL((k_p2_spill_gate_m<L, KW>), dim3(256, bins.n_iso), dim3(256), 0, s,
                         static_cast<const P2BinArgs*>(bins.args));
      if (hipError_t e = hipGetLastError()) return e;
    }
    return hipSuccess;
  }
};

hipError_t launch_p2_binned_multi(const MultiLaunch& gate, const MultiLaunch& scatter, const MultiLaunch& bins,
                                  const FillRange* runs_fill, uint32_t n_fill, uint32_t k, uint32_t max_nbins,
                                  uint32_t max_rounds, uint32_t lds_cap, uint32_t filt, hipStream_t s) {
  const uint32_t KW = (k + 30) / 31;
  // every isolate of a batch may take the buffered body when the keys fit
  // it (an isolate with more bins than kP2BufMaxBins takes the direct one)
  const bool buffered = p2_scatter_buffered(k, 1u);
  size_t scat_lds = p2_scatter_lds(max_nbins);
  if (buffered) {
    const size_t buf_lds = p2_scatter_buf_lds(max_nbins < kP2BufMaxBins ? max_nbins : kP2BufMaxBins);
    if (buf_lds > scat_lds) scat_lds = buf_lds;
  }
  const size_t lds = p2_lds_bytes(KW, lds_cap, filt);
  if (lds + 1024 > lds_budget() || scat_lds + 1024 > lds_budget()) return hipErrorInvalidConfiguration;
  if (max_rounds == 0 || max_rounds > (1u << kP2MaxLog2Rounds) || filt > kP2MaxFilt) return hipErrorInvalidValue;
  if (lds_cap % kLdsBucket || lds_cap == 0) return hipErrorInvalidValue;
  return dispatch_lkw(k, P2LaunchMulti{gate, scatter, bins, runs_fill, n_fill, max_rounds, (size_t)max_nbins * 4,
                                       scat_lds, lds, buffered, s});
}

}  // namespace fb
