// fb_host.cpp -- host-side parts of the path: the HyperLogLog f64 estimator,
// the estimator's Poisson/confidence tail, read statistics, the k ladder, the
// 2-bit packer and the seeded synthetic isolate generator.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fastbact_gpu.h"
#include "fastbact_synth.h"
#include "fb_host.hpp"

namespace {

inline uint64_t sm64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// ASCII -> 2-bit code: C/c 1, G/g 2, T/t 3, A and every other byte 0
// (bitseq.rs:67-81), as a table so that packing does not branch per base.
struct CodeTable {
  uint8_t v[256];
  constexpr CodeTable() : v() {
    v['C'] = v['c'] = 1;
    v['G'] = v['g'] = 2;
    v['T'] = v['t'] = 3;
  }
};
constexpr CodeTable kCodes;
inline uint8_t code_of(uint8_t c) { return kCodes.v[c]; }

}  // namespace

extern "C" uint64_t fb_splitmix64(uint64_t z) { return sm64(z); }

// ---------------------------------------------------------------------------
// Admission by footprint (serving)

namespace fb {

int Admission::submit(uint64_t id, uint64_t bytes) {
  if (bytes > budget_) return FB_ERR_OOM;
  queue_.emplace_back(id, bytes);
  return FB_OK;
}

void Admission::admit(std::vector<uint64_t>& out) {
  while (!queue_.empty() && inflight_.size() < max_inflight_ && used_ + queue_.front().second <= budget_) {
    const auto q = queue_.front();
    queue_.pop_front();
    inflight_[q.first] = q.second;
    used_ += q.second;
    out.push_back(q.first);
  }
}

int Admission::retire(uint64_t id) {
  auto it = inflight_.find(id);
  if (it == inflight_.end()) return FB_ERR_INVALID;
  used_ -= it->second;
  inflight_.erase(it);
  return FB_OK;
}

}  // namespace fb

struct fb_admission {
  fb::Admission a;
};

extern "C" int fb_admission_create(uint64_t budget, uint32_t max_inflight, fb_admission** out) {
  if (!out || max_inflight == 0) return FB_ERR_INVALID;
  *out = new fb_admission{fb::Admission(budget, max_inflight)};
  return FB_OK;
}
extern "C" void fb_admission_destroy(fb_admission* q) { delete q; }
extern "C" int fb_admission_submit(fb_admission* q, uint64_t id, uint64_t bytes) {
  return q ? q->a.submit(id, bytes) : FB_ERR_INVALID;
}
extern "C" uint32_t fb_admission_admit(fb_admission* q, uint64_t* ids, uint32_t max_ids) {
  if (!q) return 0;
  std::vector<uint64_t> got;
  q->a.admit(got);
  const uint32_t n = (uint32_t)got.size();
  for (uint32_t i = 0; i < n && i < max_ids; ++i) ids[i] = got[i];
  return n;
}
extern "C" int fb_admission_retire(fb_admission* q, uint64_t id) { return q ? q->a.retire(id) : FB_ERR_INVALID; }
extern "C" uint64_t fb_admission_used(const fb_admission* q) { return q ? q->a.used() : 0; }
extern "C" uint32_t fb_admission_queued(const fb_admission* q) { return q ? (uint32_t)q->a.queued() : 0; }
extern "C" uint32_t fb_admission_inflight(const fb_admission* q) { return q ? (uint32_t)q->a.inflight() : 0; }

extern "C" const char* fb_version(void) { return "fastbact-mi355x 0.1.0 (gfx950)"; }

// sketch.rs:165-168
extern "C" uint64_t fb_cms_width_for_genome_size(uint64_t g) {
  uint64_t t = g * 16, p = 1;
  while (p < t) p <<= 1;
  return std::min<uint64_t>(std::max<uint64_t>(p, 1ull << 21), 1ull << 27);
}

// sketch.rs:83-114 with m = 2^p.  The sum runs in register order;
// 2.0f64.powi(-v) is an exact power of two.
extern "C" double fb_hll_estimate_registers(const uint8_t* regs, int p) {
  const size_t mi = (size_t)1 << p;
  const double m = (double)mi;
  const double alpha_m = 0.7213 / (1.0 + 1.079 / m);
  double sum = 0.0;
  size_t zeros = 0;
  for (size_t i = 0; i < mi; ++i) {
    sum += ldexp(1.0, -(int)regs[i]);
    if (regs[i] == 0) ++zeros;
  }
  const double raw = alpha_m * m * m / sum;
  if (raw <= 2.5 * m) {
    if (zeros > 0) return m * log(m / (double)zeros);
    return raw;
  }
  if (raw <= (1.0 / 30.0) * 18446744073709551616.0) return raw;
  const double two64 = 18446744073709551616.0;
  return -two64 * log(1.0 - raw / two64);
}

namespace fb {

// Rust's `f64 as u64`: saturating, NaN -> 0.
uint64_t f64_to_u64(double x) {
  if (!(x > 0.0)) return 0;
  if (x >= 18446744073709551616.0) return UINT64_MAX;
  return (uint64_t)x;
}

static uint64_t sat_sub(uint64_t a, uint64_t b) { return a > b ? a - b : 0; }

bool checkpoint_stop(uint64_t u1, uint64_t u2, uint64_t m_sample) {
  // estimator.rs:366-375
  uint64_t e1 = sat_sub(u1, u2);
  uint64_t m_tr = sat_sub(m_sample, e1);
  double lam = u2 > 0 ? (double)m_tr / (double)u2 : 0.0;
  return lam >= 15.0;
}

// estimator.rs:385-439
void estimate_finish(uint64_t u_ge_1, uint64_t u_ge_2, uint64_t m_sample, const fb_count_stats* st,
                     double target_cov, fb_genome_estimate* o) {
  const double k = 21.0;
  uint64_t e_1 = sat_sub(u_ge_1, u_ge_2);
  uint64_t m_true = sat_sub(m_sample, e_1);
  double lambda_s = u_ge_2 > 0 ? (double)m_true / (double)u_ge_2 : 0.0;
  double p_ge_2 = 1.0 - (1.0 + lambda_s) * exp(-lambda_s);
  uint64_t est_g;
  if (p_ge_2 > 0.05) est_g = f64_to_u64(round((double)u_ge_2 / p_ge_2));
  else est_g = std::max<uint64_t>(u_ge_2, 10000);
  if (est_g < 10000) est_g = 10000;
  double estimated_cov = (double)st->total_bases / (double)est_g;
  double d_k = estimated_cov * ((st->mean_read_len - k + 1.0) / st->mean_read_len);
  double solid_ratio = u_ge_1 > 0 ? (double)u_ge_2 / (double)u_ge_1 : 0.0;
  int conf;
  if (d_k < 5.0 || solid_ratio < 0.05 || u_ge_2 < 1000) conf = 2;
  else if (d_k < 20.0 || solid_ratio < 0.15) conf = 1;
  else conf = 0;
  double sub = 1.0;
  if (target_cov > 0.0 && estimated_cov > 1.5 * target_cov) sub = std::min(std::max(target_cov / estimated_cov, 0.01), 1.0);
  o->genome_size = est_g;
  o->exact_read_pairs = st->exact_read_pairs;
  o->total_bases = st->total_bases;
  o->median_read_len = st->median_read_len;
  o->max_read_len = st->max_read_len;
  o->estimated_cov = estimated_cov;
  o->subsample_factor = sub;
  o->n_kmers = fb_compute_dynamic_kmers(st->median_read_len, -1, o->kmers);
  o->confidence = conf;
  o->u_ge_1 = u_ge_1;
  o->u_ge_2 = u_ge_2;
  o->sample_kmer_depth = lambda_s;
  o->m_sample = m_sample;
}

}  // namespace fb

// main.rs:66-185 with --enable-auto-subsample and without --no-subsample.
extern "C" double fb_auto_subsample_factor(const fb_genome_estimate* est, double depth, double threshold_ratio,
                                           uint64_t genome_size, int32_t* reason) {
  int32_t why = 2;
  double factor = 1.0;
  double cov = 0.0;
  bool have = false;
  if (genome_size > 0 && est) {
    cov = (double)est->total_bases / (double)genome_size;
    have = true;
  } else if (est && est->confidence == 0) {
    cov = est->estimated_cov;
    have = true;
  } else {
    why = 3;
  }
  if (have && depth > 0.0 && cov > threshold_ratio * depth) {
    factor = std::min(1.0, std::max(0.01, depth / cov));
    why = genome_size > 0 ? 0 : 1;
  }
  if (reason) *reason = why;
  return factor;
}

// estimator.rs:104-161
extern "C" uint32_t fb_compute_dynamic_kmers(uint64_t median_len, int64_t target_max_k, uint32_t* out) {
  uint64_t min_k = median_len < 75 ? 21 : 31;
  uint64_t effective_len = (median_len * 16) / 10;
  uint64_t auto_max = fb::f64_to_u64((double)effective_len * 0.85);
  uint64_t max_k = target_max_k >= 0 ? (uint64_t)target_max_k : auto_max;
  max_k = std::min<uint64_t>(std::max(max_k, min_k), 127);
  if (max_k % 2 == 0) max_k -= 1;
  std::vector<uint64_t> ks;
  if (max_k <= min_k) {
    ks.push_back(min_k);
  } else if (median_len >= 100) {
    const uint64_t mid_k = 63;
    if (max_k > mid_k + 10) ks = {min_k, mid_k, max_k};
    else ks = {min_k, max_k};
  } else {
    uint64_t step = (max_k - min_k) / 2;
    if (step % 2 == 1) step += 1;
    if (step < 6) step = 6;
    for (uint64_t k = min_k; k <= max_k; k += step) ks.push_back(k);
    if (ks.back() != max_k && max_k > ks.back()) {
      if (max_k - ks.back() >= 10) ks.push_back(max_k);
      else ks.back() = max_k;
    }
    ks.erase(std::unique(ks.begin(), ks.end()), ks.end());
  }
  uint32_t n = (uint32_t)std::min<size_t>(ks.size(), 8);
  for (uint32_t i = 0; i < n; ++i) out[i] = (uint32_t)ks[i];
  return n;
}

// estimator.rs:167-283, the part determined by the read lengths.
extern "C" int fb_count_stats_from_lengths(const uint32_t* l1, const uint32_t* l2, uint64_t n, fb_count_stats* o) {
  if (!o) return FB_ERR_INVALID;
  memset(o, 0, sizeof *o);
  if (n == 0) return FB_ERR_EMPTY;
  if (!l1 || !l2) return FB_ERR_INVALID;
  std::vector<uint64_t> hist(4096, 0), over;
  uint64_t mn = UINT64_MAX, mx = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t a = l1[i], b = l2[i];
    o->total_bases_r1 += a;
    o->total_bases_r2 += b;
    mn = std::min(mn, std::min(a, b));
    mx = std::max(mx, std::max(a, b));
    if (a < 4096) ++hist[a]; else over.push_back(a);
    if (b < 4096) ++hist[b]; else over.push_back(b);
  }
  o->exact_read_pairs = n;
  o->total_bases = o->total_bases_r1 + o->total_bases_r2;
  o->min_read_len = mn;
  o->max_read_len = mx;
  const uint64_t total_reads = n * 2;
  o->mean_read_len = (double)o->total_bases / (double)total_reads;
  const uint64_t target = (total_reads + 1) / 2;
  uint64_t cum = 0, med = 0;
  for (uint64_t len = 0; len < 4096; ++len) {
    cum += hist[len];
    if (cum >= target) {
      med = len;
      break;
    }
  }
  if (med == 0 && !over.empty()) {
    std::sort(over.begin(), over.end());
    med = over[std::min(over.size() / 2, over.size() - 1)];
  }
  o->median_read_len = med;
  return FB_OK;
}

// BitSeqArena::push_seq (bitseq.rs:195-222) for a batch of sequences.
extern "C" int fb_pack_ascii(const uint8_t* ascii, const uint32_t* lengths, uint64_t n, uint64_t* data,
                             uint64_t* offsets) {
  if ((n && (!lengths || !offsets)) || (!data && n)) return FB_ERR_INVALID;
  uint64_t pos = 0;
  for (uint64_t i = 0; i < n; ++i) {
    offsets[i] = pos;
    const uint32_t len = lengths[i];
    uint64_t j = 0;
    // whole words where aligned
    while (j < len) {
      const uint64_t g = pos + j;
      const uint32_t sh = (uint32_t)(g & 31) * 2;
      if (sh == 0 && j + 32 <= len) {
        uint64_t w = 0;
        for (int b = 0; b < 32; ++b) w |= (uint64_t)code_of(ascii[j + b]) << (2 * b);
        data[g >> 5] = w;
        j += 32;
        // then whole words for as long as they last
        while (j + 32 <= len) {
          uint64_t x = 0;
          const uint8_t* a = ascii + j;
          for (int b = 0; b < 32; ++b) x |= (uint64_t)kCodes.v[a[b]] << (2 * b);
          data[(pos + j) >> 5] = x;
          j += 32;
        }
      } else {
        data[g >> 5] |= (uint64_t)code_of(ascii[j]) << sh;
        ++j;
      }
    }
    ascii += len;
    pos += len;
  }
  return FB_OK;
}

// ---------------------------------------------------------------------------
// Synthetic isolates

namespace {

// xoshiro256** seeded through splitmix64 (published by Blackman and Vigna).
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
      z += 0x9e3779b97f4a7c15ull;
      s[i] = sm64(z);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
  // Box-Muller
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
};

uint8_t random_base(Rng& r, double gc) {
  const uint64_t x = r.next();
  const double u = (double)(x >> 11) * (1.0 / 9007199254740992.0);
  const bool strong = u < gc;
  const bool second = (x & 1) != 0;
  return strong ? (second ? 'G' : 'C') : (second ? 'T' : 'A');
}

uint8_t complement(uint8_t c) {
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    default: return 'A';
  }
}

}  // namespace

extern "C" void fb_synth_params_default(fb_synth_params* p) {
  memset(p, 0, sizeof *p);
  p->seed = 0xFB01;
  p->genome_len = 1000000;
  p->gc = 0.5;
  p->coverage = 30.0;
  p->read_len = 150;
  p->insert_mean = 350.0;
  p->insert_sd = 30.0;
  p->sub_rate = 0.005;
  p->indel_rate = 0.0;
  p->tandem_unit_min = 100;
  p->tandem_unit_max = 500;
  p->tandem_copies_min = 2;
  p->tandem_copies_max = 6;
}

extern "C" int fb_synth_genome(const fb_synth_params* p, uint8_t* out) {
  if (!p || !out) return FB_ERR_INVALID;
  const uint64_t G = p->genome_len;
  Rng r(p->seed ^ 0x67656e6f6d65ull);
  for (uint64_t i = 0; i < G; ++i) out[i] = random_base(r, p->gc);
  // identical operon copies at random non-overlapping-by-construction slots
  if (p->operon_copies && p->operon_len && (uint64_t)p->operon_len * p->operon_copies < G) {
    std::vector<uint8_t> op(p->operon_len);
    for (auto& c : op) c = random_base(r, p->gc);
    const uint64_t slot = G / p->operon_copies;
    for (uint32_t c = 0; c < p->operon_copies; ++c) {
      uint64_t room = slot > p->operon_len ? slot - p->operon_len : 0;
      uint64_t pos = (uint64_t)c * slot + r.below(room + 1);
      if (pos + p->operon_len <= G) memcpy(out + pos, op.data(), p->operon_len);
    }
  }
  // tandem arrays: a random unit repeated back to back
  for (uint32_t t = 0; t < p->tandem_arrays; ++t) {
    const uint32_t umin = std::max<uint32_t>(1, p->tandem_unit_min);
    const uint32_t umax = std::max(umin, p->tandem_unit_max);
    const uint32_t cmin = std::max<uint32_t>(1, p->tandem_copies_min);
    const uint32_t cmax = std::max(cmin, p->tandem_copies_max);
    const uint32_t ulen = umin + (uint32_t)r.below(umax - umin + 1);
    const uint32_t copies = cmin + (uint32_t)r.below(cmax - cmin + 1);
    std::vector<uint8_t> unit(ulen);
    for (auto& c : unit) c = random_base(r, p->gc);
    const uint64_t span = (uint64_t)ulen * copies;
    if (span >= G) continue;
    const uint64_t pos = r.below(G - span);
    for (uint32_t c = 0; c < copies; ++c) memcpy(out + pos + (uint64_t)c * ulen, unit.data(), ulen);
  }
  return FB_OK;
}

extern "C" uint64_t fb_synth_n_pairs(const fb_synth_params* p) {
  if (!p || p->read_len == 0) return 0;
  return (uint64_t)llround(p->coverage * (double)p->genome_len / (2.0 * (double)p->read_len));
}

// Copies read_len bases of the template starting at `start` (forward or
// reverse-complement strand) with substitutions and single-base indels.
static void emit_read(Rng& r, const uint8_t* g, uint64_t glen, uint64_t start, bool rc, uint32_t len, double sub,
                      double indel, uint8_t* out) {
  // rc: read runs leftwards from `start` on the reverse strand
  int64_t pos = (int64_t)start;
  const int64_t step = rc ? -1 : 1;
  uint32_t i = 0;
  while (i < len) {
    if (pos < 0 || pos >= (int64_t)glen) {
      out[i++] = 'A';
      pos += step;
      continue;
    }
    uint8_t c = rc ? complement(g[pos]) : g[pos];
    if (indel > 0.0 && r.uniform() < indel) {
      if (r.next() & 1) {  // insertion of a random base
        out[i++] = "ACGT"[r.next() & 3];
        continue;
      }
      pos += step;  // deletion: skip one template base
      continue;
    }
    if (sub > 0.0 && r.uniform() < sub) {
      const char* alt = "ACGT";
      uint8_t nc;
      do {
        nc = (uint8_t)alt[r.next() & 3];
      } while (nc == c);
      c = nc;
    }
    out[i++] = c;
    pos += step;
  }
}

extern "C" int fb_synth_reads(const fb_synth_params* p, const uint8_t* genome, uint64_t glen, uint64_t n_pairs,
                              uint8_t* r1, uint8_t* r2) {
  if (!p || !genome || (n_pairs && (!r1 || !r2))) return FB_ERR_INVALID;
  const uint32_t L = p->read_len;
  if (L == 0 || glen < L) return FB_ERR_INVALID;
  Rng r(p->seed ^ 0x7265616473ull);
  for (uint64_t i = 0; i < n_pairs; ++i) {
    double ins = p->insert_mean + p->insert_sd * r.normal();
    uint64_t insert = (uint64_t)std::max<double>(L, std::min<double>((double)glen, ins));
    uint64_t fstart = r.below(glen - insert + 1);
    const bool flip = (r.next() & 1) != 0;
    uint8_t* a = r1 + i * L;
    uint8_t* b = r2 + i * L;
    // forward fragment: R1 from fstart forwards, R2 from the fragment end backwards (RC)
    if (!flip) {
      emit_read(r, genome, glen, fstart, false, L, p->sub_rate, p->indel_rate, a);
      emit_read(r, genome, glen, fstart + insert - 1, true, L, p->sub_rate, p->indel_rate, b);
    } else {
      emit_read(r, genome, glen, fstart + insert - 1, true, L, p->sub_rate, p->indel_rate, a);
      emit_read(r, genome, glen, fstart, false, L, p->sub_rate, p->indel_rate, b);
    }
  }
  return FB_OK;
}
