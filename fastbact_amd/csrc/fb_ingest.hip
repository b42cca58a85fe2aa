// fb_ingest.hip -- the ingest record loop (src/ingest/trim_merge.rs:17-417)
// on the device, from parsed FASTQ/FASTA records to a device read bank.
//
// 1. k_ing_pairs: one wave per pair applies the subsample keep-test
//    (splitmix64 of the 1-based pair index), 3' adapter clipping, the
//    sliding-window quality trim and the overlap merge, and records the
//    outcome (dropped / merged / unmerged), the trimmed lengths and the
//    overlap.  Counters of IngestStats are summed per block, one atomic each.
// 2. Two scans place the merged reads and the unmerged pairs; k_ing_list
//    writes each output read's pair and length (merged reads, then the
//    forward mates, then the reverse mates, in pair order), a scan of the
//    lengths gives the BitSeqArena offsets.
// 3. k_ing_gather: one thread per output word computes its 32 bases from
//    the records: trimmed mates read through, merged reads as the R1 prefix,
//    the overlap consensus and the reverse-complemented R2 suffix.
#include "fb_device.hpp"
#include "fb_internal.hpp"

namespace fb {

using fbk::splitmix64;

static __device__ __forceinline__ uint8_t up_ascii(uint8_t c) { return (c >= 'a' && c <= 'z') ? (uint8_t)(c - 32) : c; }

static __device__ __forceinline__ uint32_t code_of(uint8_t c) {
  // ASCII_TO_2BIT (bitseq.rs:67-81): A/C/G/T in either case, anything else A.
  switch (c) {
    case 'C': case 'c': return 1u;
    case 'G': case 'g': return 2u;
    case 'T': case 't': return 3u;
    default: return 0u;
  }
}

// The reverse complement byte of the merge (trim_merge.rs:150-158): anything
// but ACGT becomes 'A'.
static __device__ __forceinline__ uint8_t rc_ascii(uint8_t c) {
  switch (c) {
    case 'A': case 'a': return 'T';
    case 'C': case 'c': return 'G';
    case 'G': case 'g': return 'C';
    case 'T': case 't': return 'A';
    default: return 'A';
  }
}

static __device__ __forceinline__ bool close10(const uint8_t* w, const char* ad) {
  int mm = 0;
  for (int j = 0; j < 10; ++j)
    if (up_ascii(w[j]) != (uint8_t)ad[j] && ++mm > 1) return false;
  return true;
}

// The record loop runs one wave per pair: the 64 lanes take consecutive
// positions of each scan (adapter starts, quality-trim cuts, overlap bases),
// so every load of a step is coalesced and a step ends with one ballot.
// The three functions below are called by all lanes of the wave with the
// same arguments and return the same value on every lane.

// trim_adapter (trim_merge.rs:17-75): the first position, among the last 100
// that leave 10 bases, where TruSeq or Nextera (behind A/G) or the Nextera
// reverse complement (behind C/T) match their first 10 bases with at most one
// mismatch.
static __device__ uint32_t wave_trim_adapter(const uint8_t* s, uint32_t len, uint32_t lane) {
  if (len < 10) return len;
  for (uint32_t base = len > 100 ? len - 100 : 0; base + 10 <= len; base += 64) {
    const uint32_t i = base + lane;
    bool hit = false;
    if (i + 10 <= len) {
      const uint8_t b0 = up_ascii(s[i]), b1 = up_ascii(s[i + 1]);
      const bool agat = b0 == 'A' || b1 == 'G';
      const bool ctgt = b0 == 'C' || b1 == 'T';
      hit = (agat && (close10(s + i, "AGATCGGAAG") || close10(s + i, "AGATGTGTAT"))) ||
            (ctgt && close10(s + i, "CTGTCTCTTA"));
    }
    const uint64_t m = __ballot(hit);
    if (m) return base + (uint32_t)__ffsll((long long)m) - 1u;
  }
  return len;
}

static __device__ __forceinline__ uint32_t phred(uint8_t q) { return q > 33 ? q - 33u : 0u; }

// trim_quality (trim_merge.rs:77-113) on a read whose qualities are as long
// as its bases: the kept length, or -1.  The reference keeps the whole read
// when its last 4 qualities average min_qual (sum >= 4 min_qual, the same
// test as sum / 4 >= min_qual), else drops bases until the window ending at
// the cut passes, stopping at cut 3: the cut is len - j for the first j whose
// window passes.
static __device__ int64_t wave_trim_quality(const uint8_t* q, uint32_t len, uint32_t min_qual, uint32_t min_len,
                                            uint32_t lane) {
  if (len < min_len) return -1;
  uint32_t cut = len;
  if (len >= 4) {
    cut = 3;
    for (uint32_t j0 = 0; j0 + 4 <= len; j0 += 64) {
      const uint32_t j = j0 + lane;
      bool pass = false;
      if (j + 4 <= len) {
        const uint32_t c = len - j;
        pass = (phred(q[c - 4]) + phred(q[c - 3]) + phred(q[c - 2]) + phred(q[c - 1])) / 4u >= min_qual;
      }
      const uint64_t m = __ballot(pass);
      if (m) {
        cut = len - (j0 + (uint32_t)__ffsll((long long)m) - 1u);
        break;
      }
    }
  }
  return cut < min_len ? -1 : (int64_t)cut;
}

// merge_read_pair's overlap search (trim_merge.rs:127-196) with min overlap
// 15 and 10% mismatches: the best overlap or 0.  rc2[i] = rc_ascii(r2[l2-1-i]).
// A step counts 64 positions of an overlap; counting stops once the mismatches
// pass the limit, where the reference's stop too (an overlap over the limit
// is not used, so its exact count does not matter).  Scores d - 2.5 m are kept
// doubled (2d - 5m) in integers, which orders them as the f64 scores do.
static __device__ uint32_t wave_merge(const uint8_t* r1, uint32_t l1, const uint8_t* r2, uint32_t l2,
                                      uint32_t lane) {
  const uint32_t min_overlap = 15;
  if (l1 < min_overlap || l2 < min_overlap) return 0;
  const uint32_t max_d = l1 < l2 ? l1 : l2;
  uint32_t best_d = 0;
  int64_t best2 = INT64_MIN;
  for (uint32_t d = max_d; d >= min_overlap; --d) {
    const uint8_t* s1 = r1 + (l1 - d);
    const uint64_t max_allowed = (uint64_t)((double)d * 0.10);
    uint64_t mm = 0;
    for (uint32_t i0 = 0; i0 < d && mm <= max_allowed; i0 += 64) {
      const uint32_t i = i0 + lane;
      const bool mis = i < d && up_ascii(s1[i]) != rc_ascii(r2[l2 - 1 - i]);
      mm += (uint64_t)__popcll(__ballot(mis));
    }
    if (mm <= max_allowed) {
      const int64_t sc2 = 2 * (int64_t)d - 5 * (int64_t)mm;
      if (sc2 > best2) {
        best2 = sc2;
        best_d = d;
      }
    }
    if (best2 != INT64_MIN && 2 * (int64_t)d <= best2) break;
  }
  return best_d >= min_overlap ? best_d : 0u;
}

// The same overlap search with every overlap counted at once, for mates of
// at most kPlaneBases bases.  The trimmed R1 and the reverse complement of
// the trimmed R2 are held as bit planes in LDS (bit j of word w is base
// 64 w + j): R1's two code bits and an ACGT flag, rc(R2)'s two code bits
// (rc_ascii gives ACGT only).  Lane t takes overlaps d = 15 + t, 79 + t, ..;
// the mismatches of overlap d are the set bits of
// ~ok1 | (lo1 ^ lo2) | (hi1 ^ hi2) over R1 from l1 - d against rc(R2) from
// 0, 64 positions per step.  The best overlap maximises 2d - 5m among those
// within the limit, ties to the longer overlap: the overlap the descending
// search keeps (it replaces its best only on a strictly higher score).
constexpr uint32_t kPlaneBases = 1024;
constexpr uint32_t kPlaneWords = kPlaneBases / 64 + 1;  // one zero word past the end for the funnel shifts
struct MergePlanes {
  uint64_t lo1[kPlaneWords], hi1[kPlaneWords], ok1[kPlaneWords], lo2[kPlaneWords], hi2[kPlaneWords];
};

static __device__ __forceinline__ uint64_t plane_bits(const uint64_t* p, uint32_t pos) {
  const uint32_t w = pos >> 6, s = pos & 63u;
  return s ? (p[w] >> s) | (p[w + 1] << (64u - s)) : p[w];
}

static __device__ uint32_t wave_merge_planes(const uint8_t* r1, uint32_t l1, const uint8_t* r2, uint32_t l2,
                                             uint32_t lane, MergePlanes& P) {
  const uint32_t min_overlap = 15;
  if (l1 < min_overlap || l2 < min_overlap) return 0;
  const uint32_t n1 = (l1 + 63) / 64, n2 = (l2 + 63) / 64;
  for (uint32_t w = 0; w < n1; ++w) {
    const uint32_t i = 64 * w + lane;
    const uint8_t u = i < l1 ? up_ascii(r1[i]) : (uint8_t)0;
    const uint32_t c = code_of(u);
    const uint64_t lo = __ballot(c & 1u), hi = __ballot(c >> 1), ok = __ballot(u == 'A' || u == 'C' || u == 'G' || u == 'T');
    if (lane == 0) {
      P.lo1[w] = lo;
      P.hi1[w] = hi;
      P.ok1[w] = ok;
    }
  }
  for (uint32_t w = 0; w < n2; ++w) {
    const uint32_t i = 64 * w + lane;
    const uint32_t c = i < l2 ? code_of(rc_ascii(r2[l2 - 1 - i])) : 0u;
    const uint64_t lo = __ballot(c & 1u), hi = __ballot(c >> 1);
    if (lane == 0) {
      P.lo2[w] = lo;
      P.hi2[w] = hi;
    }
  }
  if (lane == 0) {
    P.lo1[n1] = P.hi1[n1] = P.ok1[n1] = 0ull;
    P.lo2[n2] = P.hi2[n2] = 0ull;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const uint32_t max_d = l1 < l2 ? l1 : l2;
  uint64_t best = 0;  // (2d - 5m) << 32 | d; every score within the limit is positive
  for (uint32_t d0 = min_overlap; d0 <= max_d; d0 += 64) {
    const uint32_t d = d0 + lane;
    if (d > max_d) continue;
    const uint32_t o = l1 - d;
    uint32_t mm = 0;
    for (uint32_t q = 0; 64 * q < d; ++q) {
      const uint32_t p = o + 64 * q;
      uint64_t m = ~plane_bits(P.ok1, p) | (plane_bits(P.lo1, p) ^ P.lo2[q]) | (plane_bits(P.hi1, p) ^ P.hi2[q]);
      const uint32_t n = d - 64 * q;
      if (n < 64) m &= (1ull << n) - 1ull;
      mm += (uint32_t)__popcll(m);
    }
    const uint32_t max_allowed = (uint32_t)((double)d * 0.10);
    if (mm <= max_allowed) {
      const uint64_t key = (uint64_t)(2 * d - 5 * mm) << 32 | d;
      best = key > best ? key : best;
    }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const uint32_t lo = __shfl_xor((uint32_t)best, off), hi = __shfl_xor((uint32_t)(best >> 32), off);
    const uint64_t o = (uint64_t)hi << 32 | lo;
    best = o > best ? o : best;
  }
  __builtin_amdgcn_wave_barrier();  // the planes are rewritten for the next read only after every lane is done
  return (uint32_t)best;
}

// Block sums of the stats counters: [0] retained pairs, [1] bases after
// trimming, [2] merged pairs, [3] merged bases, [4] unmerged pairs.
constexpr int kIngBlock = 256;
constexpr int kIngWaves = kIngBlock / 64;  // pairs per block in k_ing_pairs

__global__ __launch_bounds__(kIngBlock) void k_ing_pairs(IngArgs a) {
  __shared__ unsigned long long s_sum[5];
  __shared__ MergePlanes s_planes[kIngWaves];
  if (threadIdx.x < 5) s_sum[threadIdx.x] = 0ull;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t i = (uint64_t)blockIdx.x * kIngWaves + (threadIdx.x >> 6);
  if (i < a.n) {
    const uint8_t* s1 = a.r1.seq + a.r1.seq_off[i];
    const uint8_t* s2 = a.r2.seq + a.r2.seq_off[i];
    const uint8_t* q1 = a.r1.qual + a.r1.qual_off[i];
    const uint8_t* q2 = a.r2.qual + a.r2.qual_off[i];
    uint32_t l1 = a.r1.seq_len[i], l2 = a.r2.seq_len[i];
    bool keep = !a.subsample || splitmix64((i + 1) ^ a.seed) <= a.keep_threshold;
    if (keep && a.enable_trim) {
      const uint32_t c1 = wave_trim_adapter(s1, l1, lane), c2 = wave_trim_adapter(s2, l2, lane);
      if (a.r1.qual_len[i] < c1 || a.r2.qual_len[i] < c2) {
        if (lane == 0) a.status[0] = 1u;  // the reference's quality slice would panic: the host reports an error
        keep = false;
      } else {
        const int64_t x1 = wave_trim_quality(q1, c1, a.min_qual, a.min_len, lane);
        const int64_t x2 = wave_trim_quality(q2, c2, a.min_qual, a.min_len, lane);
        keep = x1 >= 0 && x2 >= 0;
        l1 = (uint32_t)x1;
        l2 = (uint32_t)x2;
      }
    } else if (keep) {
      keep = l1 >= a.min_len && l2 >= a.min_len;
    }
    uint32_t st = 0, t1 = 0, t2 = 0, d = 0;
    if (keep) {
      t1 = l1;
      t2 = l2;
      if (a.enable_merge)
        d = l1 <= kPlaneBases && l2 <= kPlaneBases ? wave_merge_planes(s1, l1, s2, l2, lane, s_planes[threadIdx.x >> 6])
                                                   : wave_merge(s1, l1, s2, l2, lane);
      st = d ? 1u : 2u;
    }
    if (lane == 0) {
      if (keep) {
        atomicAdd(&s_sum[0], 1ull);
        atomicAdd(&s_sum[1], (unsigned long long)t1 + t2);
        if (d) {
          atomicAdd(&s_sum[2], 1ull);
          atomicAdd(&s_sum[3], (unsigned long long)(t1 + t2 - d));
        } else {
          atomicAdd(&s_sum[4], 1ull);
        }
      }
      a.merged[i] = st == 1u ? 1u : 0u;
      a.unmerged[i] = st == 2u ? 1u : 0u;
      a.t1[i] = t1;
      a.t2[i] = t2;
      a.ov[i] = d;
    }
  } else if (i == a.n && lane == 0) {  // the scans' totals
    a.merged[i] = 0u;
    a.unmerged[i] = 0u;
  }
  __syncthreads();
  if (threadIdx.x < 5 && s_sum[threadIdx.x]) atomicAdd(a.stats + threadIdx.x, s_sum[threadIdx.x]);
}

// Output read j: pair and length.  Merged reads take places [0, nm), forward
// mates [nm, nm + nu), reverse mates [nm + nu, nm + 2 nu).
__global__ __launch_bounds__(kIngBlock) void k_ing_list(IngArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t nm = a.pos_m[a.n], nu = a.pos_u[a.n];
  if (a.merged[i]) {
    const uint32_t j = a.pos_m[i];
    a.src[j] = (uint32_t)i;
    a.olen[j] = a.t1[i] + a.t2[i] - a.ov[i];
  } else if (a.unmerged[i]) {
    const uint32_t j = a.pos_u[i];
    a.src[nm + j] = (uint32_t)i;
    a.olen[nm + j] = a.t1[i];
    a.src[nm + nu + j] = (uint32_t)i;
    a.olen[nm + nu + j] = a.t2[i];
  }
}

static __device__ __forceinline__ uint32_t arena_of(const uint64_t* first, uint64_t j) {
  return j < first[1] ? 0u : (j < first[2] ? 1u : 2u);
}

// Output offsets (global base index of the new list, arenas word-aligned),
// lengths and the length histogram (counted per block in LDS for lengths
// below kLdsHist: the lengths of a read set take few values, and one global
// counter per length would serialise every thread on it).
constexpr uint32_t kLdsHist = 4096;
__global__ __launch_bounds__(kIngBlock) void k_ing_offsets(IngArgs a) {
  __shared__ uint32_t h[kLdsHist];
  const uint32_t nh = a.hist_n < kLdsHist ? a.hist_n : kLdsHist;
  for (uint32_t i = threadIdx.x; i < nh; i += blockDim.x) h[i] = 0u;
  __syncthreads();
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < a.rfirst[3]) {
    const uint32_t ar = arena_of(a.rfirst, j);
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

// Base p of output read j (of arena ar) as a 2-bit code.
static __device__ __forceinline__ uint32_t out_base(const IngArgs& a, uint32_t ar, uint32_t i, uint32_t p) {
  const uint8_t* s1 = a.r1.seq + a.r1.seq_off[i];
  const uint8_t* s2 = a.r2.seq + a.r2.seq_off[i];
  if (ar == 1) return code_of(s1[p]);
  if (ar == 2) return code_of(s2[p]);
  // merged (trim_merge.rs:198-224)
  const uint32_t l1 = a.t1[i], l2 = a.t2[i], d = a.ov[i];
  if (p < l1 - d) return code_of(s1[p]);
  const uint32_t k = p - (l1 - d);  // index into rc2
  const uint8_t b2 = rc_ascii(s2[l2 - 1 - k]);
  if (k >= d) return code_of(b2);
  const uint8_t b1 = s1[p];
  if (up_ascii(b1) == b2) return code_of(b1);
  const uint32_t ql1 = a.r1.qual_len[i], ql2 = a.r2.qual_len[i];
  const uint8_t qa = p < ql1 ? a.r1.qual[a.r1.qual_off[i] + p] : (uint8_t)'I';
  const uint8_t qb = (l2 - 1 - k) < ql2 ? a.r2.qual[a.r2.qual_off[i] + (l2 - 1 - k)] : (uint8_t)'I';
  return code_of(qa >= qb ? b1 : b2);
}

__global__ __launch_bounds__(kIngBlock) void k_ing_gather(IngArgs a) {
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
    if (pos < start) pos = start;
    const uint32_t i = a.src[j];
    for (; pos < end && filled < 32u; ++pos, ++filled)
      out |= (uint64_t)out_base(a, ar, i, (uint32_t)(pos - start)) << (2u * filled);
  }
  a.odata[w] = out;
}

static uint32_t ing_grid(uint64_t n) {
  const uint64_t g = (n + kIngBlock - 1) / kIngBlock;
  return (uint32_t)(g ? g : 1);
}

hipError_t launch_ing_pairs(const IngArgs& a, hipStream_t s) {
  const uint64_t blocks = (a.n + 1 + kIngWaves - 1) / kIngWaves;  // one wave per pair, and one for the totals
  hipLaunchKernelGGL(k_ing_pairs, dim3((uint32_t)blocks), dim3(kIngBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_ing_list(const IngArgs& a, hipStream_t s) {
  if (!a.n) return hipSuccess;
  hipLaunchKernelGGL(k_ing_list, dim3(ing_grid(a.n)), dim3(kIngBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_ing_offsets(const IngArgs& a, hipStream_t s) {
  if (!a.rfirst[3]) return hipSuccess;
  hipLaunchKernelGGL(k_ing_offsets, dim3(ing_grid(a.rfirst[3])), dim3(kIngBlock), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_ing_gather(const IngArgs& a, hipStream_t s) {
  if (!a.word0[3]) return hipSuccess;
  hipLaunchKernelGGL(k_ing_gather, dim3(ing_grid(a.word0[3])), dim3(kIngBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace fb
