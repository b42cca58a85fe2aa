// fb_mapper.hip -- the PE mapper's terminal k-mer index and paired mapping
// (src/resolve/mapper.rs:89-279) on the device.
//
// 1. k_map_insert: one thread per indexed window (every window of a unitig
//    with <= 2000, else its first and last 1000) takes the canonical k-mer
//    from the 2-bit packed unitig, folds it (KmerInt::to_u64) and counts it in
//    an open-addressing table in HBM (key, count, DirUnitigId << 32 | pos).
//    Only keys counted once are used, so the value of a repeated key never
//    matters and its writes may race.
// 2. k_map_pairs: one thread per pair looks up the windows at 0, 20, .., 100
//    of both mates (the first unique hit wins), then records an insert size
//    (both mates on one unitig of >= 1500 bp, inward-facing, span within
//    [max(read lengths, 100), 1500]) or a link with its exit / entry
//    distances (different unitigs).
// 3. Two scans compact the links (two evidence records per pair) and the
//    insert sizes in pair order.
#include "fb_device.hpp"
#include "fb_internal.hpp"

namespace fb {

using namespace fbk;

constexpr int kMapBlock = 256;

static __device__ __forceinline__ uint64_t map_slot(uint64_t key, uint32_t log2cap) {
  return splitmix64(key ^ 0x6a09e667f3bcc909ull) >> (64u - log2cap);
}

// Canonical k-mer of the window at global base g: (to_u64, is_rc).
template <int L>
static __device__ __forceinline__ uint64_t window_key(const uint64_t* __restrict__ data, uint64_t g, uint32_t k,
                                                      bool& is_rc) {
  const Kmer<L> win = window_bits<L>(data, g, k);
  const Kmer<L> fwd = fwd_from_window<L>(win, k);
  const Kmer<L> rev = rc_from_window<L>(win, k);
  is_rc = kmer_lt<L>(rev, fwd);
  return is_rc ? kmer_fold<L>(rev) : kmer_fold<L>(fwd);
}

template <int L>
__global__ __launch_bounds__(kMapBlock) void k_map_insert(MapArgs a) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.n_windows) return;
  uint32_t lo = 0, hi = a.n_ranges;  // largest r with rpre[r] <= w
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.rpre[mid] <= w) lo = mid;
    else hi = mid;
  }
  const uint32_t u = a.rng_u[lo];
  const uint32_t p = a.rng_p0[lo] + (uint32_t)(w - a.rpre[lo]);
  bool is_rc = false;
  const uint64_t key = window_key<L>(a.useq.data, a.useq.offs[u] + p, a.k, is_rc);
  const uint64_t val = ((uint64_t)((u << 1) | (is_rc ? 1u : 0u)) << 32) | p;
  if (key == kEmpty) {  // the one key the empty marker cannot hold
    atomicAdd(a.special_cnt, 1u);
    *a.special_val = val;
    return;
  }
  const uint64_t mask = (1ull << a.log2cap) - 1ull;
  for (uint64_t s = map_slot(key, a.log2cap);; s = (s + 1) & mask) {
    const unsigned long long prev =
        atomicCAS((unsigned long long*)&a.keys[s], (unsigned long long)kEmpty, (unsigned long long)key);
    if (prev == kEmpty || prev == key) {
      atomicAdd(a.cnt + s, 1u);
      a.val[s] = val;
      return;
    }
  }
}

// The unique entry of key, or false.
static __device__ __forceinline__ bool map_find(const MapArgs& a, uint64_t key, uint64_t& val) {
  if (key == kEmpty) {
    if (*a.special_cnt != 1u) return false;
    val = *a.special_val;
    return true;
  }
  const uint64_t mask = (1ull << a.log2cap) - 1ull;
  for (uint64_t s = map_slot(key, a.log2cap);; s = (s + 1) & mask) {
    const uint64_t v = a.keys[s];
    if (v == key) {
      if (a.cnt[s] != 1u) return false;
      val = a.val[s];
      return true;
    }
    if (v == kEmpty) return false;
  }
}

// map_read (mapper.rs:150-162): directed unitig, position and step of the
// first unique window among steps 0, 20, .., 100.
template <int L>
static __device__ bool map_read(const MapArgs& a, uint64_t r, uint32_t len, uint32_t& dir, uint32_t& pos,
                                uint32_t& step) {
  for (uint32_t st = 0; st <= 100; st += 20) {
    if (st + a.k > len) continue;
    bool is_rc = false;
    const uint64_t key = window_key<L>(a.reads.data, a.reads.offs[r] + st, a.k, is_rc);
    uint64_t v = 0;
    if (map_find(a, key, v)) {
      const uint32_t d = (uint32_t)(v >> 32);
      dir = is_rc ? (d ^ 1u) : d;
      pos = (uint32_t)v;
      step = st;
      return true;
    }
  }
  return false;
}

template <int L>
__global__ __launch_bounds__(kMapBlock) void k_map_pairs(MapArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > a.n_pairs) return;
  uint32_t is_size = 0, is_link = 0;
  if (i < a.n_pairs) {
    const uint64_t q1 = a.first_f + i, q2 = a.first_r + i;
    const uint32_t l1 = a.reads.lens[q1], l2 = a.reads.lens[q2];
    const int64_t k = a.k;
    uint32_t u1, p1, s1, u2, p2, s2;
    if (l1 >= a.k && l2 >= a.k && map_read<L>(a, q1, l1, u1, p1, s1) && map_read<L>(a, q2, l2, u2, p2, s2)) {
      const uint32_t u2f = u2 ^ 1u;
      if ((u1 >> 1) == (u2 >> 1)) {
        if (a.ulen[u1 >> 1] >= 1500u) {
          int64_t span = -1;
          if (!(u1 & 1u) && (u2 & 1u)) span = ((int64_t)p2 + k + s2) - ((int64_t)p1 - s1);
          else if ((u1 & 1u) && !(u2 & 1u)) span = ((int64_t)p1 + k + s1) - ((int64_t)p2 - s2);
          const int64_t min_len = max((int64_t)max(l1, l2), (int64_t)100);
          if (span >= min_len && span <= 1500) {
            is_size = 1;
            a.span[i] = (uint32_t)span;
          }
        }
      } else {
        const int64_t n1 = a.ulen[u1 >> 1], n2 = a.ulen[u2 >> 1];
        int64_t d1 = (u1 & 1u) ? (int64_t)p1 + k + s1 : n1 - ((int64_t)p1 - s1);
        int64_t d2 = (u2f & 1u) ? n2 - ((int64_t)p2 - s2) : (int64_t)p2 + k + s2;
        d1 = d1 < 0 ? 0 : d1;
        d2 = d2 < 0 ? 0 : d2;
        if (d1 <= 0xFFFFFFFFll && d2 <= 0xFFFFFFFFll && l1 <= 65535u && l2 <= 65535u) {
          is_link = 1;
          a.lu1[i] = u1;
          a.lu2[i] = u2f;
          a.ld1[i] = (uint32_t)d1;
          a.ld2[i] = (uint32_t)d2;
        }
      }
    }
  }
  a.f_size[i] = is_size;
  a.f_link[i] = is_link;
}

// Compaction: pair i's link records go to 2 * pos_link[i] and the next slot,
// its size to pos_size[i].
__global__ __launch_bounds__(kMapBlock) void k_map_compact(MapArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_pairs) return;
  if (a.f_size[i]) a.sizes[a.pos_size[i]] = a.span[i];
  if (a.f_link[i]) {
    const uint32_t l1 = a.reads.lens[a.first_f + i], l2 = a.reads.lens[a.first_r + i];
    const uint32_t u1 = a.lu1[i], u2f = a.lu2[i], d1 = a.ld1[i], d2 = a.ld2[i];
    PairLinkRec* o = a.links + 2ull * a.pos_link[i];
    o[0] = PairLinkRec{u1, u2f, (uint32_t)i, d1, d2, (uint16_t)l1, (uint16_t)l2};
    o[1] = PairLinkRec{u2f ^ 1u, u1 ^ 1u, (uint32_t)i, d2, d1, (uint16_t)l2, (uint16_t)l1};
  }
}

__global__ void k_map_clear(MapArgs a) {
  const uint64_t cap = 1ull << a.log2cap;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
    a.keys[i] = kEmpty;
    a.cnt[i] = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.special_cnt = 0u;
}

static uint32_t map_grid(uint64_t n) {
  const uint64_t g = (n + kMapBlock - 1) / kMapBlock;
  return (uint32_t)(g ? g : 1);
}

hipError_t launch_map_clear(const MapArgs& a, hipStream_t s) {
  const uint32_t g = (uint32_t)std::min<uint64_t>(map_grid(1ull << a.log2cap), 65536);
  hipLaunchKernelGGL(k_map_clear, dim3(g), dim3(kMapBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_map_insert(const MapArgs& a, hipStream_t s) {
  if (!a.n_windows) return hipSuccess;
  const uint32_t L = (2 * a.k + 63) / 64;
  const dim3 g(map_grid(a.n_windows));
  switch (L) {
    case 1: hipLaunchKernelGGL(k_map_insert<1>, g, dim3(kMapBlock), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_map_insert<2>, g, dim3(kMapBlock), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_map_insert<3>, g, dim3(kMapBlock), 0, s, a); break;
    default: hipLaunchKernelGGL(k_map_insert<4>, g, dim3(kMapBlock), 0, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_map_pairs(const MapArgs& a, hipStream_t s) {
  const uint32_t L = (2 * a.k + 63) / 64;
  const dim3 g(map_grid(a.n_pairs + 1));
  switch (L) {
    case 1: hipLaunchKernelGGL(k_map_pairs<1>, g, dim3(kMapBlock), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_map_pairs<2>, g, dim3(kMapBlock), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_map_pairs<3>, g, dim3(kMapBlock), 0, s, a); break;
    default: hipLaunchKernelGGL(k_map_pairs<4>, g, dim3(kMapBlock), 0, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_map_compact(const MapArgs& a, hipStream_t s) {
  if (!a.n_pairs) return hipSuccess;
  hipLaunchKernelGGL(k_map_compact, dim3(map_grid(a.n_pairs)), dim3(kMapBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace fb
