// fb_compact.hip -- unitig compaction of a sorted solid set on the device
// (src/graph/build.rs:330-597).
//
// The reference walks from every "boundary" directed k-mer (not exactly one
// in-edge and one out-edge in its mask) along each unvisited out-edge,
// through 1-in/1-out k-mers, until it reaches a boundary k-mer, a k-mer
// outside the set, or an edge it marked before; each step also marks the
// reverse-strand edge, so a path and its reverse complement are walked once.
// Here every directed k-mer d = 2i + o (o = 1: reverse complement of k-mer i)
// knows its successor; walks become list ranking:
//   k_cg_links   successor and out-base of 1-in/1-out directed k-mers (binary
//                search of the next k-mer in the sorted set), next node of
//                every out-edge of boundary k-mers;
//   k_cg_pred    predecessor = twin(successor(twin)); a predecessor that is
//                the node's own twin (a palindromic (k+1)-mer) ends the walk
//                there, and the nodes after such an end are walked by no one;
//   k_cg_jump    pointer jumping towards the path head: rank and start edge of
//                every 1-in/1-out node, cycles left unresolved, with their
//                smallest node as label;
//   k_cg_paths   per start edge: cov sum and last node; the path is walked
//                when its start edge precedes the start edge of its reverse
//                complement in ascending (k-mer, strand, base) order, which
//                is the order the walks are taken in here;
//   k_cg_write   the bases of every walked path and circular component, then
//                k_cg_flip turns a linear unitig into its reverse complement
//                when that is smaller.
#include <hipcub/hipcub.hpp>

#include "fb_device.hpp"
#include "fb_internal.hpp"

namespace fb {
using namespace fbk;

constexpr int kCgBlock = 256;
constexpr uint32_t kNone = 0xFFFFFFFFu;

// node kinds
constexpr uint8_t kBoundary = 0, kInternal = 1;
// jump states of a directed node
constexpr uint32_t kStHead = 0xFFFFFFF0u;  // up = head reached (head[] holds the start edge)
constexpr uint32_t kStDead = 0xFFFFFFF1u;  // walked by no one
constexpr uint32_t kStOff = 0xFFFFFFF2u;   // not a counted 1-in/1-out node

static inline unsigned cg_grid(uint64_t n) {
  uint64_t g = (n + kCgBlock - 1) / kCgBlock;
  return (unsigned)(g < 1 ? 1 : (g > (1u << 24) ? (1u << 24) : g));
}

// Reference key words (most significant first) -> limbs (least significant first).
template <int L>
__device__ __forceinline__ Kmer<L> cg_load(const uint64_t* __restrict__ keys, uint64_t i, uint32_t nref) {
  Kmer<L> x;
#pragma unroll
  for (int j = 0; j < L; ++j) x.a[j] = keys[i * nref + (nref - 1 - j)];
  return x;
}

// Reverse complement of a k-mer integer: the 2-bit groups reversed within 2k
// bits, complemented.
template <int L>
__device__ __forceinline__ Kmer<L> cg_rc(const Kmer<L>& x, uint32_t k) {
  Kmer<L> w = fwd_from_window<L>(x, k);  // group reversal within 2k bits (an involution)
#pragma unroll
  for (int j = 0; j < L; ++j) w.a[j] = ~w.a[j];
  w.a[L - 1] &= top_mask<L>(k);
  return w;
}

template <int L>
__device__ __forceinline__ uint32_t cg_base(const Kmer<L>& x, uint32_t pos_from_lsb) {
  return (uint32_t)(x.a[pos_from_lsb >> 5] >> ((pos_from_lsb & 31u) * 2u)) & 3u;
}

// Index of the canonical k-mer c in the sorted keys, or kNone.
template <int L>
__device__ __forceinline__ uint32_t cg_find(const uint64_t* __restrict__ keys, uint32_t n, uint32_t nref,
                                            const Kmer<L>& c) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    const Kmer<L> m = cg_load<L>(keys, mid, nref);
    if (kmer_lt<L>(m, c)) lo = mid + 1;
    else hi = mid;
  }
  if (lo < n) {
    const Kmer<L> m = cg_load<L>(keys, lo, nref);
    bool eq = true;
#pragma unroll
    for (int j = 0; j < L; ++j) eq = eq && m.a[j] == c.a[j];
    if (eq) return lo;
  }
  return kNone;
}

// Per k-mer i, both strands o: kind, successor / out-base (1-in/1-out), the
// first and last base of the strand's k-mer, the next node of each out-edge
// of a boundary strand (bnext[d*4+b]), and the start-edge count of d.
template <int L>
__global__ __launch_bounds__(kCgBlock) void k_cg_links(CgArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t k = a.k;
  const Kmer<L> xf = cg_load<L>(a.keys, i, a.nref);
  const Kmer<L> xr = cg_rc<L>(xf, k);
  const uint32_t m = a.mask[i];
#pragma unroll
  for (uint32_t o = 0; o < 2; ++o) {
    const uint32_t d = 2u * (uint32_t)i + o;
    const Kmer<L> X = o ? xr : xf;
    const Kmer<L> XR = o ? xf : xr;
    const uint32_t out = o ? (m >> 4) : (m & 15u);
    const uint32_t in = o ? (m & 15u) : (m >> 4);
    const bool internal = __popc(in) == 1 && __popc(out) == 1;
    a.kind[d] = internal ? kInternal : kBoundary;
    a.fbase[d] = (uint8_t)cg_base<L>(X, k - 1u);
    a.lbase[d] = (uint8_t)cg_base<L>(X, 0u);
    a.nstart[d] = internal ? 0u : (uint32_t)__popc(out);
    uint32_t succ = kNone;
    for (uint32_t b = 0; b < 4; ++b) {
      if (!((out >> b) & 1u)) {
        if (!internal) a.bnext[4u * d + b] = kNone;
        continue;
      }
      Kmer<L> Y = X, YR = XR;
      kmer_roll<L>(Y, YR, b, k);
      const bool rc = kmer_lt<L>(YR, Y);
      const uint32_t j = cg_find<L>(a.keys, a.n, a.nref, rc ? YR : Y);
      const uint32_t nd = j == kNone ? kNone : 2u * j + (rc ? 1u : 0u);
      if (internal) {
        succ = nd;
        a.obase[d] = (uint8_t)b;
      } else {
        a.bnext[4u * d + b] = nd;
      }
    }
    a.succ[d] = succ;
  }
}

// Predecessors and the initial jump state of every directed node.
__global__ __launch_bounds__(kCgBlock) void k_cg_pred(CgArgs a) {
  const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= 2ull * a.n) return;
  uint32_t up = kStOff, head = kNone;
  if (a.kind[d] == kInternal) {
    const uint32_t ts = a.succ[d ^ 1u];  // the twin is 1-in/1-out too
    const uint32_t p = ts == kNone ? kNone : (ts ^ 1u);
    if (p == (uint32_t)(d ^ 1u)) {
      up = kStOff;  // a hairpin end: the walk stops on it, it is not counted
    } else if (p == kNone) {
      up = kStDead;  // no predecessor in the set: reached by no walk
    } else if (a.kind[p] == kBoundary) {
      // start edge (p, base leading into d) = (p, last base of d's strand)
      const uint32_t b = a.lbase[d];
      const uint32_t mp = a.mask[p >> 1];
      const uint32_t outp = (p & 1u) ? (mp >> 4) : (mp & 15u);
      if ((outp >> b) & 1u) {
        up = kStHead;
        head = a.sbase[p] + (uint32_t)__popc(outp & ((1u << b) - 1u));
      } else {
        up = kStDead;  // masks that disagree across an edge: no walk enters d
      }
    } else {
      // p is 1-in/1-out; if p is a hairpin end, d lies beyond it
      const uint32_t tp = a.succ[p ^ 1u];
      up = (tp != kNone && (tp ^ 1u) == (p ^ 1u)) ? kStDead : p;
    }
  }
  a.js[0][d] = make_uint4(up, head, (up == kStOff || up == kStDead) ? 0u : 1u, (uint32_t)d);
}

// One pointer-jumping round r: up/head/rank/lab from buffer s into buffer t.
// A round that turns no node terminal and lowers no label ends the jump:
// round r turns terminal the nodes 2^r < D <= 2^(r+1) links behind a head
// (D = 1 starts terminal), and gives a cycle's smallest node as label to the
// cycle nodes 2^r <= D < 2^(r+1) links after it, so a round with neither
// means every path is resolved and every cycle labelled.
// The later rounds return at once; both buffers then agree on everything the
// later kernels read (the state of path nodes, and lab and up < kStHead of
// cycle nodes, whose up and rank keep moving).
__global__ __launch_bounds__(kCgBlock) void k_cg_jump(CgArgs a, uint32_t s, uint32_t t, uint32_t* flag, uint32_t r) {
  const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= 2ull * a.n) return;
  if (r > 0 && !flag[r - 1]) return;
  const uint4 me = a.js[s][d];
  const uint32_t u = me.x;
  uint32_t nu = u, nh = me.y, nr = me.z, nl = me.w;
  if (u < kStHead) {
    const uint4 o = a.js[s][u];
    nl = min(nl, o.w);
    if (o.x == kStDead) {
      nu = kStDead;
    } else {
      nu = o.x;
      nr += o.z;
      nh = o.y;
    }
  }
  a.js[t][d] = make_uint4(nu, nh, nr, nl);
  if ((u < kStHead && nu >= kStHead) || nl != me.w) flag[r] = 1u;
}

// Start edges: the start node's cov and itself as the last node so far.
__global__ __launch_bounds__(kCgBlock) void k_cg_start_init(CgArgs a) {
  const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= 2ull * a.n || a.kind[d] != kBoundary) return;
  const uint32_t m = a.mask[d >> 1];
  const uint32_t out = (d & 1u) ? (m >> 4) : (m & 15u);
  uint32_t e = a.sbase[d];
  for (uint32_t b = 0; b < 4; ++b) {
    if (!((out >> b) & 1u)) continue;
    a.csum[e] = a.cov[d >> 1];
    a.last[e] = (uint64_t)d;  // rank 0
    ++e;
  }
}

// Counted 1-in/1-out nodes add their cov to their path and bid for its last
// node; cycle nodes add to their cycle's label.  A graph of few long unitigs
// sends most nodes of a block to the same few paths, so each block first
// combines its nodes per path (or cycle) in an LDS table and then adds one
// total per entry; a node that finds no free LDS entry adds its own.
constexpr int kAccBlock = 1024;
constexpr uint32_t kAccSlots = 2048;
constexpr uint32_t kCycleTag = 0x80000000u;  // marks a cycle label in the LDS table

__device__ __forceinline__ void acc_global(const CgArgs& a, uint32_t tag, unsigned long long v, unsigned long long m) {
  if (tag & kCycleTag) {
    const uint32_t c = tag & ~kCycleTag;
    atomicAdd((unsigned long long*)a.ccov + c, v);
    atomicAdd(a.clen + c, (uint32_t)m);
  } else {
    atomicAdd((unsigned long long*)a.csum + tag, v);
    atomicMax((unsigned long long*)a.last + tag, m);
  }
}

__global__ __launch_bounds__(kAccBlock) void k_cg_accumulate(CgArgs a, uint32_t s) {
  __shared__ uint32_t key[kAccSlots];
  __shared__ unsigned long long sum[kAccSlots];  // cov total
  __shared__ unsigned long long agg[kAccSlots];  // paths: max (rank << 32 | node); cycles: node count
  for (uint32_t i = threadIdx.x; i < kAccSlots; i += blockDim.x) {
    key[i] = kNone;
    sum[i] = 0ull;
    agg[i] = 0ull;
  }
  __syncthreads();
  const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d < 2ull * a.n) {
    const uint4 me = a.js[s][d];
    const uint32_t u = me.x;
    uint32_t tag = kNone;
    unsigned long long v = 0, m = 0;
    if (u == kStHead) {
      tag = me.y;
      v = a.cov[d >> 1];
      m = ((unsigned long long)me.z << 32) | (unsigned long long)d;
    } else if (u < kStHead) {  // on a cycle: lab = smallest node of the cycle
      tag = me.w | kCycleTag;
      v = a.cov[d >> 1];
      m = 1ull;
    }
    if (tag != kNone) {
      uint32_t slot = (tag * 0x9E3779B1u) >> (32 - 11);  // 11 = log2(kAccSlots)
      bool done = false;
      for (int probe = 0; probe < 16 && !done; ++probe, slot = (slot + 1) & (kAccSlots - 1)) {
        uint32_t k = key[slot];
        if (k == kNone) {
          const uint32_t old = atomicCAS(&key[slot], kNone, tag);
          k = old == kNone ? tag : old;
        }
        if (k == tag) {
          atomicAdd(&sum[slot], v);
          if (tag & kCycleTag) atomicAdd(&agg[slot], m);
          else atomicMax(&agg[slot], m);
          done = true;
        }
      }
      if (!done) acc_global(a, tag, v, m);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kAccSlots; i += blockDim.x)
    if (key[i] != kNone) acc_global(a, key[i], sum[i], agg[i]);
}

// Which start edges are walked, and the length of their unitigs.
__global__ __launch_bounds__(kCgBlock) void k_cg_paths(CgArgs a) {
  const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= 2ull * a.n || a.kind[d] != kBoundary) return;
  const uint32_t m = a.mask[d >> 1];
  const uint32_t out = (d & 1u) ? (m >> 4) : (m & 15u);
  uint32_t e = a.sbase[d];
  for (uint32_t b = 0; b < 4; ++b) {
    if (!((out >> b) & 1u)) continue;
    const uint64_t lst = a.last[e];
    const uint32_t last = (uint32_t)lst, rk = (uint32_t)(lst >> 32);
    const uint32_t end = rk == 0 ? a.bnext[4u * d + b] : a.succ[last];
    bool walked = true;
    if (end != kNone && a.kind[end] == kBoundary) {
      // the reverse complement starts at twin(end) with the complement of
      // the first base of the last counted node
      const uint32_t tb = 3u - a.fbase[last];
      const uint32_t te = end ^ 1u;
      const uint32_t mt = a.mask[te >> 1];
      const uint32_t outt = (te & 1u) ? (mt >> 4) : (mt & 15u);
      const uint32_t twin_edge = a.sbase[te] + (uint32_t)__popc(outt & ((1u << tb) - 1u));
      walked = e <= twin_edge;
    }
    a.walked[e] = walked ? 1u : 0u;
    a.slen[e] = walked ? a.k + 1u + rk : 0u;
    ++e;
  }
}

// Circular components walked: the cycle labelled by its smallest node s,
// when s is a forward strand and the cycle does not hold s's reverse strand.
__global__ __launch_bounds__(kCgBlock) void k_cg_cycles(CgArgs a, uint32_t s) {
  const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= 2ull * a.n) return;
  uint32_t len = 0;
  const uint4 me = a.js[s][d];
  if (me.x < kStHead && me.w == (uint32_t)d && !(d & 1u)) {
    const uint4 tw = a.js[s][d ^ 1u];
    const bool self_twin = tw.x < kStHead && tw.w == (uint32_t)d;
    if (!self_twin) len = a.clen[d];
  }
  a.cyc_len[d] = len;  // bases of the circular unitig (one per node), 0 = none
}

// Second jump over cycles (into buffer t): the rank of each node after its
// cycle's start.
__global__ __launch_bounds__(kCgBlock) void k_cg_cycle_init(CgArgs a, uint32_t s, uint32_t t) {
  const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= 2ull * a.n) return;
  const uint4 me = a.js[s][d];
  const uint32_t u = me.x;
  uint32_t nu = kStOff, nr = 0;
  if (u < kStHead) {
    const uint32_t c = me.w;
    if (a.cyc_len[c]) {
      if (c == (uint32_t)d) {
        nu = kStHead;
      } else {
        // the original predecessor: twin(successor(twin))
        nu = a.succ[d ^ 1u] ^ 1u;
        nr = 1;
      }
    }
  }
  a.js[t][d] = make_uint4(nu, kNone, nr, (uint32_t)d);
}

// The strand of k-mer d written as ASCII at out[0..cnt).
template <int L>
__device__ __forceinline__ void cg_write_kmer(const CgArgs& a, uint32_t d, uint8_t* out, uint32_t cnt) {
  const Kmer<L> xf = cg_load<L>(a.keys, d >> 1, a.nref);
  const Kmer<L> X = (d & 1u) ? cg_rc<L>(xf, a.k) : xf;
  for (uint32_t j = 0; j < cnt; ++j) out[j] = "ACGT"[cg_base<L>(X, a.k - 1u - j)];
}

// Bases of walked paths: start k-mers and their first step (one thread per
// boundary node), counted nodes' out-bases (one per node), cycle starts and
// cycle nodes' out-bases.
template <int L>
__global__ __launch_bounds__(kCgBlock) void k_cg_write(CgArgs a, uint32_t sp, uint32_t sc) {
  const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= 2ull * a.n) return;
  const uint32_t k = a.k;
  if (a.kind[d] == kBoundary) {
    const uint32_t m = a.mask[d >> 1];
    const uint32_t out = (d & 1u) ? (m >> 4) : (m & 15u);
    uint32_t e = a.sbase[d];
    for (uint32_t b = 0; b < 4; ++b) {
      if (!((out >> b) & 1u)) continue;
      if (a.walked[e]) {
        uint8_t* o = a.seq + a.soff[e];
        cg_write_kmer<L>(a, (uint32_t)d, o, k);
        o[k] = "ACGT"[b];
      }
      ++e;
    }
    return;
  }
  // counted node of a path (sp: the buffer of the first jump)
  const uint4 jp = a.js[sp][d];
  if (jp.x == kStHead) {
    const uint32_t e = jp.y;
    if (a.walked[e]) a.seq[a.soff[e] + k + jp.z] = "ACGT"[a.obase[d]];
    return;
  }
  // node of a walked cycle (sc: the buffer of the cycle jump)
  const uint4 jc = a.js[sc][d];
  if (jc.x == kStHead) {
    const uint32_t c = jp.w;  // cycle start (label of the first jump)
    const uint32_t L_ = a.cyc_len[c];
    const uint64_t off = a.circ_base + a.coff[c];
    const uint32_t r = jc.z;
    if (r == 0) cg_write_kmer<L>(a, (uint32_t)d, a.seq + off, k < L_ ? k : L_);
    if (k + r < L_) a.seq[off + k + r] = "ACGT"[a.obase[d]];
  }
}

// Linear unitigs whose reverse complement is smaller are flipped (one thread
// per unitig decides, one per base pair flips).
__global__ __launch_bounds__(kCgBlock) void k_cg_decide_flip(const uint8_t* __restrict__ seq,
                                                           const uint64_t* __restrict__ off,
                                                           const uint32_t* __restrict__ len, uint32_t n_lin,
                                                           uint8_t* __restrict__ flip) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n_lin) return;
  const uint8_t* s = seq + off[u];
  const uint32_t L = len[u];
  uint8_t f = 0;
  for (uint32_t j = 0; j < L; ++j) {
    const uint8_t c = s[L - 1 - j];
    const uint8_t rc = c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A';
    if (rc != s[j]) {
      f = rc < s[j] ? 1 : 0;
      break;
    }
  }
  flip[u] = f;
}

__global__ __launch_bounds__(kCgBlock) void k_cg_flip(uint8_t* __restrict__ seq, const uint64_t* __restrict__ off,
                                                    const uint32_t* __restrict__ len, uint32_t n_lin,
                                                    const uint8_t* __restrict__ flip, uint64_t lin_bases) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= lin_bases) return;
  // unitig of base p
  uint32_t lo = 0, hi = n_lin;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= p) lo = mid;
    else hi = mid;
  }
  if (!flip[lo]) return;
  const uint64_t j = p - off[lo];
  const uint32_t L = len[lo];
  if (2 * j + 1 > L) return;  // the first half (and the middle) swaps
  uint8_t* s = seq + off[lo];
  auto comp = [](uint8_t c) -> uint8_t { return c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A'; };
  const uint8_t x = s[j], y = s[L - 1 - j];
  s[j] = comp(y);
  s[L - 1 - j] = comp(x);
}

// 1 for every walked cycle label.
__global__ __launch_bounds__(kCgBlock) void k_cg_cycle_flags(const uint32_t* __restrict__ cyc_len, uint32_t* flag,
                                                           uint64_t nodes) {
  const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d < nodes) flag[d] = cyc_len[d] ? 1u : 0u;
}

// The unitig table: linear unitigs in start-edge order, then circular ones
// in the order of their start k-mer.
__global__ __launch_bounds__(kCgBlock) void k_cg_table_lin(CgArgs a) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.n_se || !a.walked[e]) return;
  const uint32_t u = a.uid[e];
  const uint32_t points = a.slen[e] - a.k;
  a.u_off[u] = a.soff[e];
  a.u_len[u] = a.slen[e];
  const float c = (float)a.csum[e] / (float)points;
  a.u_cov[u] = c >= 1.0f ? c : 1.0f;
  a.u_circ[u] = 0;
}
__global__ __launch_bounds__(kCgBlock) void k_cg_table_circ(CgArgs a) {
  const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= 2ull * a.n || !a.cyc_len[d]) return;
  const uint32_t u = a.n_lin + a.cid[d];
  a.u_off[u] = a.circ_base + a.coff[d];
  a.u_len[u] = a.cyc_len[d];
  const float c = (float)a.ccov[d] / (float)a.clen[d];
  a.u_cov[u] = c >= 1.0f ? c : 1.0f;
  a.u_circ[u] = 1;
}

// ---------------------------------------------------------------------------
// launchers

template <class F>
static hipError_t cg_dispatch(uint32_t k, F&& f) {
  const uint32_t L = (2 * k + 63) / 64;
  switch (L) {
    case 1: return f.template operator()<1>();
    case 2: return f.template operator()<2>();
    case 3: return f.template operator()<3>();
    default: return f.template operator()<4>();
  }
}

struct CgLinks {
  const CgArgs& a;
  hipStream_t s;
  template <int L>
  hipError_t operator()() {
    hipLaunchKernelGGL((k_cg_links<L>), dim3(cg_grid(a.n)), dim3(kCgBlock), 0, s, a);
    return hipGetLastError();
  }
};
struct CgWrite {
  const CgArgs& a;
  uint32_t sp, sc;
  hipStream_t s;
  template <int L>
  hipError_t operator()() {
    hipLaunchKernelGGL((k_cg_write<L>), dim3(cg_grid(2ull * a.n)), dim3(kCgBlock), 0, s, a, sp, sc);
    return hipGetLastError();
  }
};

hipError_t launch_cg_links(const CgArgs& a, hipStream_t s) { return cg_dispatch(a.k, CgLinks{a, s}); }

hipError_t launch_cg_pred(const CgArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_pred, dim3(cg_grid(2ull * a.n)), dim3(kCgBlock), 0, s, a);
  return hipGetLastError();
}

uint32_t cg_jump_rounds(uint64_t nodes) {
  uint32_t r = 1;
  while ((1ull << r) < nodes + 1) ++r;
  return r + 1;
}

// Pointer jumping between buffers s and t (starting in s); returns the
// buffer holding the result.  Every chain of at most 2n links resolves.
hipError_t launch_cg_jump(const CgArgs& a, uint32_t s, uint32_t t, uint32_t* out_s, hipStream_t st) {
  const uint32_t rounds = cg_jump_rounds(2ull * a.n);
  if (rounds > 64) return hipErrorInvalidValue;
  uint32_t* flag = a.jflag + (s == 0 ? 0 : 64);  // the path jump starts in buffer 0, the cycle jump in 2
  if (hipError_t e = launch_fill_u32(flag, 64, 0u, st)) return e;
  for (uint32_t r = 0; r < rounds; ++r) {
    hipLaunchKernelGGL(k_cg_jump, dim3(cg_grid(2ull * a.n)), dim3(kCgBlock), 0, st, a, s, t, flag, r);
    const uint32_t x = s;
    s = t;
    t = x;
  }
  *out_s = s;
  return hipGetLastError();
}

hipError_t launch_cg_accumulate(const CgArgs& a, uint32_t s, hipStream_t st) {
  hipLaunchKernelGGL(k_cg_start_init, dim3(cg_grid(2ull * a.n)), dim3(kCgBlock), 0, st, a);
  hipLaunchKernelGGL(k_cg_accumulate, dim3((unsigned)((2ull * a.n + kAccBlock - 1) / kAccBlock)), dim3(kAccBlock), 0,
                     st, a, s);
  hipLaunchKernelGGL(k_cg_paths, dim3(cg_grid(2ull * a.n)), dim3(kCgBlock), 0, st, a);
  hipLaunchKernelGGL(k_cg_cycles, dim3(cg_grid(2ull * a.n)), dim3(kCgBlock), 0, st, a, s);
  return hipGetLastError();
}

// Cycle ranks in buffers 2/3 (the path ranks stay in buffer s).
hipError_t launch_cg_cycle_ranks(const CgArgs& a, uint32_t s, uint32_t* out_s, hipStream_t st) {
  hipLaunchKernelGGL(k_cg_cycle_init, dim3(cg_grid(2ull * a.n)), dim3(kCgBlock), 0, st, a, s, 2u);
  return launch_cg_jump(a, 2u, 3u, out_s, st);
}

hipError_t launch_cg_cycle_flags(const CgArgs& a, uint32_t* flag, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_cycle_flags, dim3(cg_grid(2ull * a.n)), dim3(kCgBlock), 0, s, a.cyc_len, flag, 2ull * a.n);
  return hipGetLastError();
}

hipError_t launch_cg_table(const CgArgs& a, hipStream_t s) {
  if (a.n_se) hipLaunchKernelGGL(k_cg_table_lin, dim3(cg_grid(a.n_se)), dim3(kCgBlock), 0, s, a);
  hipLaunchKernelGGL(k_cg_table_circ, dim3(cg_grid(2ull * a.n)), dim3(kCgBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_cg_write(const CgArgs& a, uint32_t sp, uint32_t sc, hipStream_t s) {
  return cg_dispatch(a.k, CgWrite{a, sp, sc, s});
}

hipError_t launch_cg_flip(uint8_t* seq, const uint64_t* off, const uint32_t* len, uint32_t n_lin, uint8_t* flip,
                          uint64_t lin_bases, hipStream_t s) {
  if (!n_lin) return hipSuccess;
  hipLaunchKernelGGL(k_cg_decide_flip, dim3(cg_grid(n_lin)), dim3(kCgBlock), 0, s, seq, off, len, n_lin, flip);
  hipLaunchKernelGGL(k_cg_flip, dim3(cg_grid(lin_bases)), dim3(kCgBlock), 0, s, seq, off, len, n_lin, flip,
                     lin_bases);
  return hipGetLastError();
}

// ---- overlap edges (build.rs:552-594) -------------------------------------
// Strand d = 2u + o of unitig u (o = 1: its reverse complement).  The
// ov-prefix of strand 2u is u's first ov bases, of 2u + 1 the reverse
// complement of u's last ov bases.  End e = 2u + side: side 0 is u's last ov
// bases, side 1 the reverse complement of its first ov bases.  For ov = k,
// then k - 1, end (u, side) meets every strand whose ov-prefix equals it, in
// ascending strand order (the order of build.rs's sorted prefix list for
// equal prefixes); a strand of u itself only when u is circular.  A
// circular unitig's forward end starts with a self edge of overlap k - 1
// (support 10), and its own k - 1 match adds 10 to that edge instead of a
// second one (UnitigNode::add_edge).  Prefixes are found by a 64-bit hash of
// their bases, sorted with the strand ids, and confirmed base by base.

__device__ __forceinline__ uint32_t cg_code(uint8_t c) { return c == 'A' ? 0u : c == 'C' ? 1u : c == 'G' ? 2u : 3u; }

// Base j of strand d's ov-prefix.
__device__ __forceinline__ uint32_t cg_pre_base(const CgEdgeArgs& a, uint32_t d, uint32_t j) {
  const uint32_t u = d >> 1, L = a.len[u];
  const uint8_t* s = a.seq + a.off[u];
  return (d & 1u) ? 3u - cg_code(s[L - 1u - j]) : cg_code(s[j]);
}
// Base j of end e's ov bases.
__device__ __forceinline__ uint32_t cg_end_base(const CgEdgeArgs& a, uint32_t e, uint32_t j) {
  const uint32_t u = e >> 1, L = a.len[u];
  const uint8_t* s = a.seq + a.off[u];
  return (e & 1u) ? 3u - cg_code(s[a.ov - 1u - j]) : cg_code(s[L - a.ov + j]);
}
template <class B>
__device__ __forceinline__ uint64_t cg_hash(uint32_t ov, B&& base) {
  uint64_t h = 0x6a09e667f3bcc909ull, w = 0;
  for (uint32_t j = 0; j < ov; ++j) {
    w = (w << 2) | base(j);
    if ((j & 31u) == 31u || j + 1u == ov) {
      h = splitmix64(h ^ w);
      w = 0;
    }
  }
  return h;
}

__global__ __launch_bounds__(kCgBlock) void k_cg_prefix(CgEdgeArgs a) {
  const uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= 2ull * a.U) return;
  const uint32_t dd = (uint32_t)d;
  a.hkey[d] = a.len[dd >> 1] >= a.ov ? cg_hash(a.ov, [&](uint32_t j) { return cg_pre_base(a, dd, j); }) : ~0ull;
  a.hval[d] = dd;
}

// Visits the strands whose ov-prefix equals end e's, ascending: f(d).
template <class F>
__device__ __forceinline__ void cg_matches(const CgEdgeArgs& a, uint32_t e, F&& f) {
  const uint32_t u = e >> 1;
  if (a.len[u] < a.ov) return;
  const uint64_t h = cg_hash(a.ov, [&](uint32_t j) { return cg_end_base(a, e, j); });
  uint32_t lo = 0, hi = 2u * a.U;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.skey[mid] < h) lo = mid + 1;
    else hi = mid;
  }
  for (uint32_t i = lo; i < 2u * a.U && a.skey[i] == h; ++i) {
    const uint32_t d = a.sval[i];
    if (a.len[d >> 1] < a.ov) continue;
    if ((d >> 1) == u && !a.circ[u]) continue;
    bool eq = true;
    for (uint32_t j = 0; j < a.ov && eq; ++j) eq = cg_pre_base(a, d, j) == cg_end_base(a, e, j);
    if (eq) f(d);
  }
}

// The self edge of a circular unitig's forward end absorbs its k - 1 match.
__device__ __forceinline__ bool cg_merged(const CgEdgeArgs& a, uint32_t e, uint32_t d) {
  return a.merge && (e & 1u) == 0u && a.circ[e >> 1] && d == e;
}

// cnt[e] += the end's matches of this pass (the merged one not counted; it
// sets bit 31 instead).  With first, cnt[e] starts at the self edge.
__global__ __launch_bounds__(kCgBlock) void k_cg_end_count(CgEdgeArgs a, uint32_t first) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2ull * a.U) return;
  const uint32_t ee = (uint32_t)e;
  uint32_t c = first ? (((ee & 1u) == 0u && a.circ[ee >> 1]) ? 1u : 0u) : a.cnt[e];
  cg_matches(a, ee, [&](uint32_t d) {
    if (cg_merged(a, ee, d)) c |= 0x80000000u;
    else ++c;
  });
  a.cnt[e] = c;
}

// Writes end e's edges from base[e]: its self edge in the first pass, then
// this pass's matches after the first pass's (pos[e] counts written edges).
__global__ __launch_bounds__(kCgBlock) void k_cg_end_write(CgEdgeArgs a, uint32_t first) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2ull * a.U) return;
  const uint32_t ee = (uint32_t)e;
  uint32_t p = first ? 0u : a.pos[e];
  const uint64_t b = a.base[e];
  if (first && (ee & 1u) == 0u && a.circ[ee >> 1]) {
    a.e_src[b] = ee;
    a.e_dst[b] = ee;
    a.e_ov[b] = (uint8_t)(a.k - 1u);
    a.e_sup[b] = (a.cnt[e] & 0x80000000u) ? 20 : 10;
    p = 1;
  }
  cg_matches(a, ee, [&](uint32_t d) {
    if (cg_merged(a, ee, d)) return;
    a.e_src[b + p] = ee;
    a.e_dst[b + p] = d;
    a.e_ov[b + p] = (uint8_t)a.ov;
    a.e_sup[b + p] = 10;
    ++p;
  });
  a.pos[e] = p;
}

// tot[e] = end e's edge count (cnt without its flag).
__global__ __launch_bounds__(kCgBlock) void k_cg_end_total(const uint32_t* cnt, uint32_t* tot, uint32_t n) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) tot[e] = cnt[e] & 0x7FFFFFFFu;
}

hipError_t launch_cg_prefix(const CgEdgeArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_prefix, dim3(cg_grid(2ull * a.U)), dim3(kCgBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_cg_end_count(const CgEdgeArgs& a, bool first, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_end_count, dim3(cg_grid(2ull * a.U)), dim3(kCgBlock), 0, s, a, first ? 1u : 0u);
  return hipGetLastError();
}
hipError_t launch_cg_end_total(const uint32_t* cnt, uint32_t* tot, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_end_total, dim3(cg_grid(n)), dim3(kCgBlock), 0, s, cnt, tot, n);
  return hipGetLastError();
}
hipError_t launch_cg_end_write(const CgEdgeArgs& a, bool first, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_end_write, dim3(cg_grid(2ull * a.U)), dim3(kCgBlock), 0, s, a, first ? 1u : 0u);
  return hipGetLastError();
}
// Stable radix sort of (u64 key, u32 value) pairs; tmp == nullptr sizes *tmp_bytes.
hipError_t cg_sort_pairs_u64(const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout, uint64_t n,
                             void* tmp, size_t* tmp_bytes, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortPairs(tmp, *tmp_bytes, kin, kout, vin, vout, (int)n, 0, 64, s);
}

// Exclusive prefix sums (device-wide).
hipError_t cg_exclusive_sum_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* tmp, size_t* tmp_bytes,
                                hipStream_t s) {
  return hipcub::DeviceScan::ExclusiveSum(tmp, *tmp_bytes, in, out, (int)n, s);
}
hipError_t cg_exclusive_sum_u64(const uint32_t* in, uint64_t* out, uint64_t n, void* tmp, size_t* tmp_bytes,
                                hipStream_t s) {
  return hipcub::DeviceScan::ExclusiveScan(tmp, *tmp_bytes, in, out, hipcub::Sum(), (uint64_t)0, (int)n, s);
}

}  // namespace fb
