// fb_device.hpp -- device primitives of the k-mer spectrum path for gfx950.
//
// K-mer integers live in L 64-bit limbs, least significant limb first
// (a[0] = bits 0..63).  L = ceil(2k/64), so k<=32 -> 1 limb, k<=64 -> 2,
// k<=96 -> 3, k<=127 -> 4.  The integer value is the reference's KmerInt value
// (src/common/kmer.rs:18-335: first base most significant), so ordering and the
// to_u64 fold (XOR of 64-bit limbs, kmer.rs:115-117,192-194,332-334) agree with
// u64 / u128 / [u64;4] whatever limb count the reference type uses.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbk {

// src/common/sketch.rs:13-18
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// src/common/sketch.rs:122-127
__device__ __forceinline__ uint64_t cms_seed(int r) {
  return r == 0 ? 0x243f6a8885a308d3ull
       : r == 1 ? 0x13198a2e03707344ull
       : r == 2 ? 0xa4093822299f31d0ull
                : 0x082efa98ec4e6c89ull;
}

// ntHash seeds, src/common/kmer.rs:385-391 (A, C, G, T).  The reverse table
// is the forward table indexed by the complement (3 - b).
__device__ __forceinline__ uint64_t nt_fwd(uint32_t b) {
  const uint64_t A = 0x3c8bfbb395c60474ull, C = 0x3193c18562a02b4cull;
  const uint64_t G = 0x20323ee046061444ull, T = 0x295c504fac61002eull;
  uint64_t lo = (b & 1) ? C : A;
  uint64_t hi = (b & 1) ? T : G;
  return (b & 2) ? hi : lo;
}
__device__ __forceinline__ uint64_t nt_rev(uint32_t b) { return nt_fwd(3u - b); }

__device__ __forceinline__ uint64_t rol64(uint64_t x, uint32_t r) {
  r &= 63u;
  return r ? (x << r) | (x >> (64u - r)) : x;
}
__device__ __forceinline__ uint64_t ror64(uint64_t x, uint32_t r) {
  r &= 63u;
  return r ? (x >> r) | (x << (64u - r)) : x;
}

// Reverses the order of the 32 2-bit groups of a word.
__device__ __forceinline__ uint64_t rev2(uint64_t x) {
  x = __builtin_bitreverse64(x);
  return ((x >> 1) & 0x5555555555555555ull) | ((x & 0x5555555555555555ull) << 1);
}

template <int L>
struct Kmer {
  uint64_t a[L];
};

template <int L>
__device__ __forceinline__ bool kmer_lt(const Kmer<L>& x, const Kmer<L>& y) {
#pragma unroll
  for (int i = L - 1; i >= 0; --i)
    if (x.a[i] != y.a[i]) return x.a[i] < y.a[i];
  return false;
}
template <int L>
__device__ __forceinline__ uint64_t kmer_fold(const Kmer<L>& x) {
  uint64_t f = x.a[0];
#pragma unroll
  for (int i = 1; i < L; ++i) f ^= x.a[i];
  return f;
}

// Mask of the valid bits of the top limb for a 2k-bit value in L limbs.
template <int L>
__device__ __forceinline__ uint64_t top_mask(uint32_t k) {
  uint32_t bits = 2u * k - 64u * (L - 1);  // 1..64
  return bits >= 64u ? ~0ull : ((1ull << bits) - 1ull);
}

// The 2k arena-order bits of the window that starts at base g of the packed
// stream (base g+j at bits 2j).  `data` must be readable for L+1 words past
// word g>>5 (the device arena carries that padding).
template <int L>
__device__ __forceinline__ Kmer<L> window_bits(const uint64_t* __restrict__ data, uint64_t g, uint32_t k) {
  const uint64_t* p = data + (g >> 5);
  uint32_t sh = (uint32_t)(g & 31u) * 2u;
  Kmer<L> w;
  uint64_t cur = p[0];
#pragma unroll
  for (int i = 0; i < L; ++i) {
    uint64_t nxt = p[i + 1];
    w.a[i] = sh ? (cur >> sh) | (nxt << (64u - sh)) : cur;
    cur = nxt;
  }
  w.a[L - 1] &= top_mask<L>(k);
  return w;
}

// Forward k-mer (first base most significant) from arena-order window bits:
// the 2-bit groups reversed within 2k bits.
template <int L>
__device__ __forceinline__ Kmer<L> fwd_from_window(const Kmer<L>& w, uint32_t k) {
  Kmer<L> r;
#pragma unroll
  for (int i = 0; i < L; ++i) r.a[i] = rev2(w.a[L - 1 - i]);
  uint32_t s = 64u * L - 2u * k;  // 0..63 by the choice of L
  if (s) {
#pragma unroll
    for (int i = 0; i < L; ++i) {
      uint64_t hi = (i + 1 < L) ? r.a[i + 1] : 0ull;
      r.a[i] = (r.a[i] >> s) | (hi << (64u - s));
    }
  }
  return r;
}

// Reverse-complement k-mer from arena-order window bits: base j of the window
// complemented lands at bits 2j of the RC integer, i.e. the NOT of the window.
template <int L>
__device__ __forceinline__ Kmer<L> rc_from_window(const Kmer<L>& w, uint32_t k) {
  Kmer<L> r;
#pragma unroll
  for (int i = 0; i < L; ++i) r.a[i] = ~w.a[i];
  r.a[L - 1] &= top_mask<L>(k);
  return r;
}

// roll_next (kmer.rs:104-112, 181-189, 307-329): shift in base b at the 3'
// end of fwd, its complement at the 5' end of rev.
template <int L>
__device__ __forceinline__ void kmer_roll(Kmer<L>& fwd, Kmer<L>& rev, uint32_t b, uint32_t k) {
#pragma unroll
  for (int i = L - 1; i > 0; --i) fwd.a[i] = (fwd.a[i] << 2) | (fwd.a[i - 1] >> 62);
  fwd.a[0] = (fwd.a[0] << 2) | b;
  fwd.a[L - 1] &= top_mask<L>(k);
#pragma unroll
  for (int i = 0; i < L - 1; ++i) rev.a[i] = (rev.a[i] >> 2) | (rev.a[i + 1] << 62);
  rev.a[L - 1] >>= 2;
  uint32_t pos = 2u * (k - 1u);
  uint64_t cb = (uint64_t)(3u - b) << (pos & 63u);
#pragma unroll
  for (int i = 0; i < L; ++i)
    if ((pos >> 6) == (uint32_t)i) rev.a[i] |= cb;
}

// Base at global position g of the packed stream.
__device__ __forceinline__ uint32_t base_at(const uint64_t* __restrict__ data, uint64_t g) {
  return (uint32_t)(data[g >> 5] >> ((g & 31u) * 2u)) & 3u;
}

// Streams the bases of the packed arena from a start position.  The three
// words from the current base on (at least 65 bases) are loaded together,
// and again after every 64 bases.  With a load at each word boundary, some
// lane of a wave would cross one at almost every base and the whole wave
// would wait for its load; lanes that walk shares of equal length (reads of
// one length) instead reload at the same base, once per 64.  It reads up to
// two words past the last base streamed: every device arena carries
// kArenaPadWords zero words after its data.
struct BaseStream {
  const uint64_t* p;  // the word holding the next base
  uint64_t word, nxt, nxt2;
  uint32_t sh, left;  // bit offset of the next base in *p; bases before the reload
  __device__ __forceinline__ void load() {
    word = p[0] >> sh;
    nxt = p[1];
    nxt2 = p[2];
    left = 64u;
  }
  __device__ __forceinline__ void init(const uint64_t* __restrict__ data, uint64_t g) {
    p = data + (g >> 5);
    sh = (uint32_t)(g & 31u) * 2u;
    load();
  }
  __device__ __forceinline__ uint32_t next() {
    const uint32_t b = (uint32_t)word & 3u;
    sh += 2u;
    const bool turn = sh == 64u;
    word = turn ? nxt : word >> 2;
    nxt = turn ? nxt2 : nxt;
    p += turn ? 1 : 0;
    sh = turn ? 0u : sh;
    if (--left == 0u) load();
    return b;
  }
};

// At most 64 bases from a start position (a pass-2 share's bases after its
// first window).  The three words they can span are loaded at init, so
// next() never waits on memory: with a word load at the word boundary, as
// BaseStream does, some lane of a wave crosses a word at almost every base
// and the whole wave waits for its load.  It reads up to two words past
// the last base streamed: every device arena carries kArenaPadWords zero
// words after its data.
struct ShortStream {
  uint64_t word, nxt, nxt2;
  uint32_t sh;
  __device__ __forceinline__ void init(const uint64_t* __restrict__ data, uint64_t g) {
    const uint64_t* p = data + (g >> 5);
    sh = (uint32_t)(g & 31u) * 2u;
    word = p[0] >> sh;
    nxt = p[1];
    nxt2 = p[2];
  }
  __device__ __forceinline__ uint32_t next() {
    const uint32_t b = (uint32_t)word & 3u;
    sh += 2u;
    const bool turn = sh == 64u;
    word = turn ? nxt : word >> 2;
    nxt = turn ? nxt2 : nxt;
    sh = turn ? 0u : sh;
    return b;
  }
};

// Rolling ntHash state (kmer.rs:425-477).
struct NtHash {
  uint64_t f, r;
  // Pushes base b into a window that is still filling (no base leaves):
  // after k pushes f/r equal nthash_fwd/nthash_rev of the k bases.
  __device__ __forceinline__ void push(uint32_t b, uint32_t k) {
    f = rol64(f, 1) ^ nt_fwd(b);
    r = ror64(r, 1) ^ rol64(nt_rev(b), (k - 1u) & 63u);
  }
  // RollingNtHash::roll: out leaves the 5' end, in enters the 3' end.
  __device__ __forceinline__ void roll(uint32_t out, uint32_t in, uint32_t k) {
    f = rol64(f, 1) ^ rol64(nt_fwd(out), k & 63u) ^ nt_fwd(in);
    r = ror64(r, 1) ^ ror64(nt_rev(out), 1) ^ rol64(nt_rev(in), (k - 1u) & 63u);
  }
  __device__ __forceinline__ uint64_t canonical() const { return f < r ? f : r; }
};

// The four per-base terms of NtHash::roll for a fixed k, looked up (from
// LDS) instead of selected among the seeds and rotated at every window.
struct NtRollTab {
  uint64_t out_f[4];  // rol(nt_fwd(b), k)
  uint64_t in_f[4];   // nt_fwd(b)
  uint64_t out_r[4];  // ror(nt_rev(b), 1)
  uint64_t in_r[4];   // rol(nt_rev(b), k - 1)
  // Entry b; threads 0..3 of a block fill the table together.
  __device__ __forceinline__ void fill(uint32_t b, uint32_t k) {
    out_f[b] = rol64(nt_fwd(b), k & 63u);
    in_f[b] = nt_fwd(b);
    out_r[b] = ror64(nt_rev(b), 1);
    in_r[b] = rol64(nt_rev(b), (k - 1u) & 63u);
  }
};
// NtHash::push with the table's "in" terms (they are the same two terms).
__device__ __forceinline__ void nt_push(NtHash& h, uint32_t b, const NtRollTab& t) {
  h.f = rol64(h.f, 1) ^ t.in_f[b];
  h.r = ror64(h.r, 1) ^ t.in_r[b];
}
__device__ __forceinline__ void nt_roll(NtHash& h, uint32_t out, uint32_t in, const NtRollTab& t) {
  h.f = rol64(h.f, 1) ^ t.out_f[out] ^ t.in_f[in];
  h.r = ror64(h.r, 1) ^ t.out_r[out] ^ t.in_r[in];
}

// HyperLogLog register index and rank (sketch.rs:49-63) at precision p.
__device__ __forceinline__ void hll_index_rho(uint64_t h, uint32_t p, uint32_t& idx, uint32_t& rho) {
  idx = (uint32_t)(h >> (64u - p));
  uint64_t w = h & ((1ull << (64u - p)) - 1ull);
  rho = w ? (uint32_t)__builtin_clzll(w) - p + 1u : 64u - p + 1u;
}

// Slices per bin region: one per XCD (accelerator complex die) of the MI355X.
constexpr uint32_t kSlices = 8;

// XCD the calling wave runs on (hardware register XCC_ID, bits 3:0), folded
// into 0..kSlices-1.  All waves of a workgroup share one CU and so one XCD.
// A workgroup writes its share of a bin into its XCD's slice, so every line
// that is still being filled is held in one L2 only.  Results never depend
// on the value: any slice choice is exact, only the write traffic changes.
__device__ __forceinline__ uint32_t xcc_slice() {
  return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (kSlices - 1u);
}

// Isolate of block b in a multi-isolate launch: the largest i with blk[i] <= b
// (blk holds n+1 block prefixes, blk[0] = 0).
__device__ __forceinline__ uint32_t find_iso(const uint32_t* __restrict__ blk, uint32_t n, uint32_t b) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (blk[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

}  // namespace fbk
