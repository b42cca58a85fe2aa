/*
 * fb_oracle.c -- CPU restatement of fastBact's k-mer spectrum path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline
 * ("port" kind).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg load it.  The GPU product never links, loads or calls it.
 *
 * Citations are to the fastBact source tree (src/...).  How each part is
 * pinned (tests/test_oracle_*.py):
 *   splitmix64            closed form vs an independent Python big-int restatement
 *   rev_comp_word         golden vectors of bitseq.rs:434-446 + brute force on strings
 *   KmerInt (1/2/4 words) from_ascii/to_ascii round trip, rev_comp vs string RC,
 *                         rolling == fresh, fwd/RC canonical agreement for
 *                         k = 21/55/99/127 (kmer.rs:543-651), fold vs big-int
 *   ntHash                closed-form definition in Python, roll == fresh (kmer.rs:653-664)
 *   CountMinSketch        0/5/15 and 4x3=12 (sketch.rs:304-342), clamped per-cell
 *                         histogram computed by brute force
 *   HyperLogLog           per-hash register closed form, p16 -> p14 fold, error
 *                         gates of sketch.rs:261-407
 *   estimator             count/ladder known answers (estimator.rs:463-546), Poisson tail
 *                         recomputed in Python, order-dependent hll_solid vs a
 *                         sequential Python emulation on small inputs
 *   solid set             brute force with exact counts at a collision-free width,
 *                         and the 10 kb circular-repeat case of build.rs:604-687
 *                         (all genome k-mers solid, errors filtered, 3 unitigs)
 *   f64 HLL estimate      "parity unpinned" against the Rust binary (cannot be run
 *                         here); sequential sum in register order with glibc log/exp.
 */
#define _GNU_SOURCE
#include "fb_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define FORCE_INLINE static inline __attribute__((always_inline))

/* ------------------------------------------------------------------------- */
/* sketch.rs:13-18                                                           */
uint64_t orc_splitmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
FORCE_INLINE uint64_t sm64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* bitseq.rs:67-81: A/a=0 C/c=1 G/g=2 T/t=3, everything else 0 */
static uint8_t A2B[256];
static void init_a2b(void) __attribute__((constructor));
static void init_a2b(void) {
    memset(A2B, 0, sizeof A2B);
    A2B['C'] = A2B['c'] = 1;
    A2B['G'] = A2B['g'] = 2;
    A2B['T'] = A2B['t'] = 3;
}
uint8_t orc_ascii_to_2bit(uint8_t c) { return A2B[c]; }
static const uint8_t TWO_BIT_TO_ASCII[4] = {'A', 'C', 'G', 'T'};

/* bitseq.rs:120-138 */
FORCE_INLINE uint64_t rcw(uint64_t word, uint32_t nb) {
    if (nb == 0) return 0;
    uint64_t w = __builtin_bswap64(word);
    w = ((w & 0x0f0f0f0f0f0f0f0fULL) << 4) | ((w & 0xf0f0f0f0f0f0f0f0ULL) >> 4);
    w = ((w & 0x3333333333333333ULL) << 2) | ((w & 0xccccccccccccccccULL) >> 2);
    w = ~w;
    if (nb < 32) {
        uint32_t shift = (32 - nb) * 2;
        uint64_t mask = (1ULL << (nb * 2)) - 1;
        return (w >> shift) & mask;
    }
    return w;
}
uint64_t orc_rev_comp_word(uint64_t word, uint32_t nb) { return rcw(word, nb); }

/* bitseq.rs:195-222 */
void orc_arena_pack(const uint8_t* ascii, const uint32_t* lengths, uint64_t n, uint64_t* data,
                    uint64_t* offsets) {
    uint64_t pos = 0;
    for (uint64_t i = 0; i < n; i++) {
        offsets[i] = pos;
        for (uint32_t j = 0; j < lengths[i]; j++) {
            uint64_t g = pos + j;
            data[g >> 5] |= (uint64_t)A2B[*ascii++] << ((g & 31) << 1);
        }
        pos += lengths[i];
    }
}
/* bitseq.rs:286-299 */
void orc_arena_extract_ascii(const orc_arena* a, uint64_t id, uint8_t* out) {
    uint64_t start = a->offsets[id];
    uint32_t len = a->lengths[id];
    for (uint32_t i = 0; i < len; i++) {
        uint64_t g = start + i;
        out[i] = TWO_BIT_TO_ASCII[(a->data[g >> 5] >> ((g & 31) << 1)) & 3];
    }
}

/* ------------------------------------------------------------------------- */
/* KmerInt.  km_t holds nw words, most significant first.                     */
typedef struct { uint64_t w[4]; } km_t;

int orc_kmer_words_for_k(uint32_t k) { return k <= 31 ? 1 : (k <= 63 ? 2 : 4); }

FORCE_INLINE u128 k2_get(km_t x) { return ((u128)x.w[0] << 64) | x.w[1]; }
FORCE_INLINE km_t k2_put(u128 v) { km_t r = {{(uint64_t)(v >> 64), (uint64_t)v, 0, 0}}; return r; }

/* kmer.rs:73-81, 132-140, 210-222 */
FORCE_INLINE km_t km_from_ascii(const uint8_t* s, uint32_t k, int nw) {
    km_t r = {{0, 0, 0, 0}};
    if (nw == 1) {
        uint64_t v = 0;
        for (uint32_t i = 0; i < k; i++) v = (v << 2) | A2B[s[i]];
        r.w[0] = v;
    } else if (nw == 2) {
        u128 v = 0;
        for (uint32_t i = 0; i < k; i++) v = (v << 2) | A2B[s[i]];
        r = k2_put(v);
    } else {
        for (uint32_t i = 0; i < k; i++) {
            uint64_t code = A2B[s[i]];
            r.w[0] = (r.w[0] << 2) | (r.w[1] >> 62);
            r.w[1] = (r.w[1] << 2) | (r.w[2] >> 62);
            r.w[2] = (r.w[2] << 2) | (r.w[3] >> 62);
            r.w[3] = (r.w[3] << 2) | code;
        }
    }
    return r;
}

/* kmer.rs:245-257, 291-302: mask a 4-word value to 2k bits */
FORCE_INLINE void k4_mask(uint64_t* w, uint32_t k) {
    uint32_t total_bits = 2 * k;
    for (int j = 0; j < 4; j++) {
        uint32_t word_min_bit = (uint32_t)(3 - j) * 64;
        if (total_bits <= word_min_bit) {
            w[j] = 0;
        } else if (total_bits < word_min_bit + 64) {
            uint32_t bits = total_bits - word_min_bit;
            w[j] &= (1ULL << bits) - 1;
        }
    }
}

/* kmer.rs:99-101, 158-178, 260-305 */
FORCE_INLINE km_t km_rev_comp(km_t x, uint32_t k, int nw) {
    km_t r = {{0, 0, 0, 0}};
    if (nw == 1) {
        r.w[0] = rcw(x.w[0], k);
    } else if (nw == 2) {
        if (k == 0) return r;
        uint64_t hi = x.w[0], lo = x.w[1];
        u128 full = ((u128)rcw(lo, 32) << 64) | (u128)rcw(hi, 32);
        if (k < 64) {
            uint32_t shift = (64 - k) * 2;
            u128 mask = (((u128)1) << (2 * k)) - 1;
            full = (full >> shift) & mask;
        }
        r = k2_put(full);
    } else {
        if (k == 0) return r;
        uint64_t full[4] = {rcw(x.w[3], 32), rcw(x.w[2], 32), rcw(x.w[1], 32), rcw(x.w[0], 32)};
        if (k == 128) {
            memcpy(r.w, full, sizeof full);
            return r;
        }
        uint32_t s = (128 - k) * 2;
        uint32_t w_shift = s / 64, b_shift = s % 64;
        uint64_t sh[4] = {0, 0, 0, 0};
        for (uint32_t i = 0; i < 4; i++) {
            if (i + w_shift < 4) {
                sh[i + w_shift] |= full[i] >> b_shift;
                if (b_shift > 0 && i + w_shift + 1 < 4) sh[i + w_shift + 1] |= full[i] << (64 - b_shift);
            }
        }
        k4_mask(sh, k);
        memcpy(r.w, sh, sizeof sh);
    }
    return r;
}

/* kmer.rs:104-112, 181-189, 307-329 */
FORCE_INLINE void km_roll(km_t* fwd, km_t* rev, uint8_t base, uint32_t k, int nw) {
    uint64_t b = base & 3;
    uint64_t cb = 3 - b; /* COMPLEMENT_2BIT */
    if (nw == 1) {
        uint64_t mask = k >= 32 ? ~0ULL : ((1ULL << (2 * k)) - 1);
        fwd->w[0] = ((fwd->w[0] << 2) | b) & mask;
        rev->w[0] = (rev->w[0] >> 2) | (cb << (2 * (k - 1)));
    } else if (nw == 2) {
        u128 mask = k >= 64 ? ~(u128)0 : ((((u128)1) << (2 * k)) - 1);
        u128 f = ((k2_get(*fwd) << 2) | b) & mask;
        u128 r = (k2_get(*rev) >> 2) | ((u128)cb << (2 * (k - 1)));
        *fwd = k2_put(f);
        *rev = k2_put(r);
    } else {
        uint64_t* f = fwd->w;
        uint64_t nf[4] = {(f[0] << 2) | (f[1] >> 62), (f[1] << 2) | (f[2] >> 62),
                          (f[2] << 2) | (f[3] >> 62), (f[3] << 2) | b};
        k4_mask(nf, k);
        memcpy(f, nf, sizeof nf);
        uint64_t* r = rev->w;
        uint64_t nr[4] = {r[0] >> 2, (r[1] >> 2) | (r[0] << 62), (r[2] >> 2) | (r[1] << 62),
                          (r[3] >> 2) | (r[2] << 62)};
        uint32_t bit_pos = 2 * (k - 1);
        nr[3 - bit_pos / 64] |= cb << (bit_pos % 64);
        memcpy(r, nr, sizeof nr);
    }
}

/* derived Ord on u64 / u128 / [u64;4]: lexicographic over MSB-first words */
FORCE_INLINE int km_lt(const km_t* a, const km_t* b, int nw) {
    for (int i = 0; i < nw; i++)
        if (a->w[i] != b->w[i]) return a->w[i] < b->w[i];
    return 0;
}
FORCE_INLINE int km_eq(const km_t* a, const km_t* b, int nw) {
    for (int i = 0; i < nw; i++)
        if (a->w[i] != b->w[i]) return 0;
    return 1;
}
/* kmer.rs:115-117, 192-194, 332-334 */
FORCE_INLINE uint64_t km_fold(const km_t* x, int nw) {
    uint64_t f = x->w[0];
    for (int i = 1; i < nw; i++) f ^= x->w[i];
    return f;
}

static km_t km_load(const uint64_t* p, int nw) {
    km_t r = {{0, 0, 0, 0}};
    memcpy(r.w, p, (size_t)nw * 8);
    return r;
}
void orc_kmer_from_ascii(const uint8_t* ascii, uint32_t k, int nw, uint64_t* out) {
    km_t r = km_from_ascii(ascii, k, nw);
    memcpy(out, r.w, (size_t)nw * 8);
}
void orc_kmer_rev_comp(const uint64_t* x, uint32_t k, int nw, uint64_t* out) {
    km_t r = km_rev_comp(km_load(x, nw), k, nw);
    memcpy(out, r.w, (size_t)nw * 8);
}
void orc_kmer_roll_next(const uint64_t* fwd, const uint64_t* rev, uint8_t base, uint32_t k, int nw,
                        uint64_t* out_fwd, uint64_t* out_rev) {
    km_t f = km_load(fwd, nw), r = km_load(rev, nw);
    km_roll(&f, &r, base, k, nw);
    memcpy(out_fwd, f.w, (size_t)nw * 8);
    memcpy(out_rev, r.w, (size_t)nw * 8);
}
uint64_t orc_kmer_to_u64(const uint64_t* x, int nw) {
    km_t v = km_load(x, nw);
    return km_fold(&v, nw);
}
void orc_kmer_canonical_windows(const uint8_t* s, uint32_t len, uint32_t k, int nw, uint64_t* out,
                                uint8_t* is_rc) {
    if (len < k || k == 0) return;
    km_t fwd = km_from_ascii(s, k, nw);
    km_t rev = km_rev_comp(fwd, k, nw);
    for (uint32_t i = k;; i++) {
        int rc = km_lt(&rev, &fwd, nw);
        const km_t* c = rc ? &rev : &fwd;
        uint64_t w = i - k;
        memcpy(out + w * nw, c->w, (size_t)nw * 8);
        is_rc[w] = (uint8_t)rc;
        if (i >= len) break;
        km_roll(&fwd, &rev, A2B[s[i]], k, nw);
    }
}

/* ------------------------------------------------------------------------- */
/* ntHash, kmer.rs:385-477                                                    */
static const uint64_t NT_FWD[4] = {0x3c8bfbb395c60474ULL, 0x3193c18562a02b4cULL,
                                   0x20323ee046061444ULL, 0x295c504fac61002eULL};
static const uint64_t NT_REV[4] = {0x295c504fac61002eULL, 0x20323ee046061444ULL,
                                   0x3193c18562a02b4cULL, 0x3c8bfbb395c60474ULL};
FORCE_INLINE uint64_t rol64(uint64_t x, uint32_t r) { r &= 63; return r ? (x << r) | (x >> (64 - r)) : x; }
FORCE_INLINE uint64_t ror64(uint64_t x, uint32_t r) { r &= 63; return r ? (x >> r) | (x << (64 - r)) : x; }

uint64_t orc_nthash_fwd(const uint8_t* s, uint32_t k) {
    uint64_t h = 0;
    for (uint32_t i = 0; i < k; i++) h ^= rol64(NT_FWD[A2B[s[i]]], (k - 1 - i) % 64);
    return h;
}
uint64_t orc_nthash_rev(const uint8_t* s, uint32_t k) {
    uint64_t h = 0;
    for (uint32_t i = 0; i < k; i++) h ^= rol64(NT_REV[A2B[s[i]]], i % 64);
    return h;
}
uint64_t orc_nthash_canonical(const uint8_t* s, uint32_t k) {
    uint64_t f = orc_nthash_fwd(s, k), r = orc_nthash_rev(s, k);
    return f < r ? f : r;
}
typedef struct { uint32_t k; uint64_t f, r; } nth_t;
FORCE_INLINE uint64_t nth_roll(nth_t* h, uint8_t out_c, uint8_t in_c) {
    uint32_t kmod = h->k % 64, km1 = (h->k - 1) % 64;
    h->f = rol64(h->f, 1) ^ rol64(NT_FWD[A2B[out_c]], kmod) ^ NT_FWD[A2B[in_c]];
    h->r = ror64(h->r, 1) ^ ror64(NT_REV[A2B[out_c]], 1) ^ rol64(NT_REV[A2B[in_c]], km1);
    return h->f < h->r ? h->f : h->r;
}
void orc_nthash_rolling(const uint8_t* s, uint32_t len, uint32_t k, uint64_t* out) {
    if (len < k || k == 0) return;
    nth_t h = {k, orc_nthash_fwd(s, k), orc_nthash_rev(s, k)};
    out[0] = h.f < h.r ? h.f : h.r;
    for (uint32_t i = k; i < len; i++) out[i - k + 1] = nth_roll(&h, s[i - k], s[i]);
}

/* ------------------------------------------------------------------------- */
/* CountMinSketch, sketch.rs:121-254                                          */
static const uint64_t CMS_SEEDS[4] = {0x243f6a8885a308d3ULL, 0x13198a2e03707344ULL,
                                      0xa4093822299f31d0ULL, 0x082efa98ec4e6c89ULL};
struct orc_cms {
    uint64_t width, mask;
    int bits;
    uint8_t* nib;
    uint16_t* u16;
};

orc_cms* orc_cms_new(uint64_t width, int bits) {
    if (width == 0) width = 1;
    if (width & (width - 1)) { /* next_power_of_two (sketch.rs:141-143) */
        uint64_t p = 1;
        while (p < width) p <<= 1;
        width = p;
    }
    orc_cms* c = (orc_cms*)calloc(1, sizeof *c);
    c->width = width;
    c->mask = width - 1;
    c->bits = bits;
    if (bits == 4) c->nib = (uint8_t*)calloc(width * 4 / 2 + 1, 1);
    else c->u16 = (uint16_t*)calloc(width * 4, 2);
    return c;
}
void orc_cms_free(orc_cms* c) {
    if (!c) return;
    free(c->nib);
    free(c->u16);
    free(c);
}
uint64_t orc_cms_width(const orc_cms* c) { return c->width; }
size_t orc_cms_table_bytes(const orc_cms* c) { return c->bits == 4 ? c->width * 2 : c->width * 8; }
const void* orc_cms_table(const orc_cms* c) { return c->bits == 4 ? (const void*)c->nib : (const void*)c->u16; }
uint64_t orc_cms_width_for_genome_size(uint64_t g) {
    uint64_t t = g * 16, p = 1;
    while (p < t) p <<= 1; /* next_power_of_two; 0 -> 1 */
    if (p < (1ULL << 21)) p = 1ULL << 21;
    if (p > (1ULL << 27)) p = 1ULL << 27;
    return p;
}

/* sketch.rs:191-223; atomic variant so the same code serves the threaded baseline */
FORCE_INLINE void cms_add(orc_cms* c, uint64_t key) {
    for (int r = 0; r < 4; r++) {
        uint64_t col = sm64(key ^ CMS_SEEDS[r]) & c->mask;
        uint64_t idx = (uint64_t)r * c->width + col;
        if (c->bits == 4) {
            uint8_t* p = &c->nib[idx >> 1];
            int hi = (int)(idx & 1);
            uint8_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
            for (;;) {
                uint8_t cnt = hi ? (cur >> 4) & 0x0F : cur & 0x0F;
                if (cnt >= 15) break;
                uint8_t nc = cnt + 1;
                uint8_t nv = hi ? (uint8_t)((cur & 0x0F) | (nc << 4)) : (uint8_t)((cur & 0xF0) | nc);
                if (__atomic_compare_exchange_n(p, &cur, nv, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) break;
            }
        } else {
            uint16_t* p = &c->u16[idx];
            uint16_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
            for (;;) {
                if (cur == 0xFFFF) break;
                if (__atomic_compare_exchange_n(p, &cur, (uint16_t)(cur + 1), 1, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED))
                    break;
            }
        }
    }
}
/* sketch.rs:227-246 (min starts at the counter maximum, early out at 0) */
FORCE_INLINE uint32_t cms_est(const orc_cms* c, uint64_t key) {
    uint32_t mn = c->bits == 4 ? 15u : 65535u;
    for (int r = 0; r < 4; r++) {
        uint64_t col = sm64(key ^ CMS_SEEDS[r]) & c->mask;
        uint64_t idx = (uint64_t)r * c->width + col;
        uint32_t cnt;
        if (c->bits == 4) {
            uint8_t cur = __atomic_load_n(&c->nib[idx >> 1], __ATOMIC_RELAXED);
            cnt = (idx & 1) ? (cur >> 4) & 0x0F : cur & 0x0F;
        } else {
            cnt = __atomic_load_n(&c->u16[idx], __ATOMIC_RELAXED);
        }
        if (cnt < mn) mn = cnt;
        if (mn == 0) return 0;
    }
    return mn;
}
void orc_cms_add(orc_cms* c, uint64_t key) { cms_add(c, key); }
uint32_t orc_cms_estimate(const orc_cms* c, uint64_t key) { return cms_est(c, key); }
uint32_t orc_cms_add_and_estimate(orc_cms* c, uint64_t key) {
    cms_add(c, key);
    return cms_est(c, key);
}

/* ------------------------------------------------------------------------- */
/* HyperLogLog, sketch.rs:24-115, generalised to precision p                  */
FORCE_INLINE void hll_add(uint8_t* regs, int p, uint64_t h) {
    uint64_t idx = h >> (64 - p);
    uint64_t w = h & ((1ULL << (64 - p)) - 1);
    uint8_t rho = w == 0 ? (uint8_t)(64 - p + 1) : (uint8_t)(__builtin_clzll(w) - p + 1);
    if (rho > regs[idx]) regs[idx] = rho;
}
void orc_hll_add_hash(uint8_t* regs, int p, uint64_t h) { hll_add(regs, p, h); }
void orc_hll_merge(uint8_t* regs, const uint8_t* other, int p) {
    size_t m = (size_t)1 << p;
    for (size_t i = 0; i < m; i++)
        if (other[i] > regs[i]) regs[i] = other[i];
}
double orc_hll_estimate(const uint8_t* regs, int p) {
    size_t mi = (size_t)1 << p;
    double m = (double)mi;
    double alpha_m = 0.7213 / (1.0 + 1.079 / m);
    double sum = 0.0;
    size_t zeros = 0;
    for (size_t i = 0; i < mi; i++) {
        sum += ldexp(1.0, -(int)regs[i]); /* 2.0f64.powi(-v) is exact */
        if (regs[i] == 0) zeros++;
    }
    double raw = alpha_m * m * m / sum;
    if (raw <= 2.5 * m) {
        if (zeros > 0) return m * log(m / (double)zeros);
        return raw;
    } else if (raw <= (1.0 / 30.0) * 18446744073709551616.0) {
        return raw;
    } else {
        double two64 = 18446744073709551616.0;
        return -two64 * log(1.0 - raw / two64);
    }
}

/* ------------------------------------------------------------------------- */
/* Estimator                                                                  */
uint64_t orc_f64_to_u64(double x) {
    if (!(x > 0.0)) return 0; /* negative and NaN */
    if (x >= 18446744073709551616.0) return UINT64_MAX;
    return (uint64_t)x;
}
static uint64_t sat_sub(uint64_t a, uint64_t b) { return a > b ? a - b : 0; }

/* estimator.rs:104-161 */
uint32_t orc_compute_dynamic_kmers(uint64_t median_len, int64_t target_max_k, uint32_t* out) {
    uint64_t min_k = median_len < 75 ? 21 : 31;
    uint64_t effective_len = (median_len * 16) / 10;
    uint64_t auto_max = orc_f64_to_u64((double)effective_len * 0.85);
    uint64_t max_k = target_max_k >= 0 ? (uint64_t)target_max_k : auto_max;
    if (max_k < min_k) max_k = min_k;
    if (max_k > 127) max_k = 127;
    if (max_k % 2 == 0) max_k -= 1;
    uint32_t n = 0;
    if (max_k <= min_k) {
        out[n++] = (uint32_t)min_k;
        return n;
    }
    if (median_len >= 100) {
        uint64_t mid_k = 63;
        if (max_k > mid_k + 10) {
            out[n++] = (uint32_t)min_k;
            out[n++] = (uint32_t)mid_k;
            out[n++] = (uint32_t)max_k;
        } else {
            out[n++] = (uint32_t)min_k;
            out[n++] = (uint32_t)max_k;
        }
        return n;
    }
    uint64_t step = (max_k - min_k) / 2;
    if (step % 2 == 1) step += 1;
    if (step < 6) step = 6;
    for (uint64_t k = min_k; k <= max_k && n < 8; k += step) out[n++] = (uint32_t)k;
    if (out[n - 1] != max_k && max_k > out[n - 1]) {
        if (max_k - out[n - 1] >= 10) out[n++] = (uint32_t)max_k;
        else out[n - 1] = (uint32_t)max_k;
    }
    /* dedup of consecutive equal values */
    uint32_t m = 0;
    for (uint32_t i = 0; i < n; i++)
        if (m == 0 || out[m - 1] != out[i]) out[m++] = out[i];
    return m;
}

/* estimator.rs:167-283 restricted to what the lengths determine */
int orc_count_stats_from_lengths(const uint32_t* l1, const uint32_t* l2, uint64_t n,
                                 orc_count_stats* o) {
    memset(o, 0, sizeof *o);
    if (n == 0) return -1;
    uint64_t* hist = (uint64_t*)calloc(4096, sizeof(uint64_t));
    uint64_t* over = (uint64_t*)malloc(sizeof(uint64_t) * 2 * n);
    uint64_t n_over = 0;
    uint64_t mn = UINT64_MAX, mx = 0;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t a = l1[i], b = l2[i];
        o->total_bases_r1 += a;
        o->total_bases_r2 += b;
        if (a < mn) mn = a;
        if (b < mn) mn = b;
        if (a > mx) mx = a;
        if (b > mx) mx = b;
        if (a < 4096) hist[a]++; else over[n_over++] = a;
        if (b < 4096) hist[b]++; else over[n_over++] = b;
    }
    o->exact_read_pairs = n;
    o->total_bases = o->total_bases_r1 + o->total_bases_r2;
    o->min_read_len = mn;
    o->max_read_len = mx;
    uint64_t total_reads = n * 2;
    o->mean_read_len = (double)o->total_bases / (double)total_reads;
    uint64_t target = (total_reads + 1) / 2, cum = 0, med = 0;
    for (uint64_t len = 0; len < 4096; len++) {
        cum += hist[len];
        if (cum >= target) {
            med = len;
            break;
        }
    }
    if (med == 0 && n_over > 0) {
        /* sort_unstable then the middle element */
        for (uint64_t i = 1; i < n_over; i++) {
            uint64_t v = over[i], j = i;
            while (j > 0 && over[j - 1] > v) {
                over[j] = over[j - 1];
                j--;
            }
            over[j] = v;
        }
        uint64_t idx = n_over / 2;
        if (idx > n_over - 1) idx = n_over - 1;
        med = over[idx];
    }
    o->median_read_len = med;
    free(hist);
    free(over);
    return 0;
}

/* The sequential sketch loop body for one read (estimator.rs:319-339). */
static void est_read(const orc_arena* a, uint64_t id, uint8_t* buf, orc_cms* cms, uint8_t* ht,
                     uint8_t* hs, uint64_t* m_sample) {
    const uint32_t k = 21;
    uint32_t len = a->lengths[id];
    if (len < k) return;
    orc_arena_extract_ascii(a, id, buf);
    *m_sample += len - k + 1;
    nth_t h = {k, orc_nthash_fwd(buf, k), orc_nthash_rev(buf, k)};
    uint64_t raw = h.f < h.r ? h.f : h.r;
    for (uint32_t i = k;; i++) {
        uint32_t c = orc_cms_add_and_estimate(cms, raw);
        uint64_t mixed = sm64(raw);
        hll_add(ht, 14, mixed);
        if (c >= 2) hll_add(hs, 14, mixed);
        if (i >= len) break;
        raw = nth_roll(&h, buf[i - k], buf[i]);
    }
}

static uint64_t arena_max_len(const orc_arena* a) {
    uint64_t m = 0;
    for (uint64_t i = 0; i < a->n_reads; i++)
        if (a->lengths[i] > m) m = a->lengths[i];
    return m;
}

void orc_estimate_finish(uint64_t u_ge_1, uint64_t u_ge_2, uint64_t m_sample,
                         const orc_count_stats* st, double target_cov, orc_genome_estimate* o) {
    const double k = 21.0;
    uint64_t e_1 = sat_sub(u_ge_1, u_ge_2);
    uint64_t m_true = sat_sub(m_sample, e_1);
    double lambda_s = u_ge_2 > 0 ? (double)m_true / (double)u_ge_2 : 0.0;
    double p_ge_2 = 1.0 - (1.0 + lambda_s) * exp(-lambda_s);
    uint64_t est_g;
    if (p_ge_2 > 0.05) est_g = orc_f64_to_u64(round((double)u_ge_2 / p_ge_2));
    else est_g = u_ge_2 > 10000 ? u_ge_2 : 10000;
    if (est_g < 10000) est_g = 10000;
    double estimated_cov = (double)st->total_bases / (double)est_g;
    double d_k = estimated_cov * ((st->mean_read_len - k + 1.0) / st->mean_read_len);
    double solid_ratio = u_ge_1 > 0 ? (double)u_ge_2 / (double)u_ge_1 : 0.0;
    int conf;
    if (d_k < 5.0 || solid_ratio < 0.05 || u_ge_2 < 1000) conf = 2;
    else if (d_k < 20.0 || solid_ratio < 0.15) conf = 1;
    else conf = 0;
    double sub = 1.0;
    if (target_cov > 0.0 && estimated_cov > 1.5 * target_cov) {
        sub = target_cov / estimated_cov;
        if (sub < 0.01) sub = 0.01;
        if (sub > 1.0) sub = 1.0;
    }
    o->genome_size = est_g;
    o->exact_read_pairs = st->exact_read_pairs;
    o->total_bases = st->total_bases;
    o->median_read_len = st->median_read_len;
    o->max_read_len = st->max_read_len;
    o->estimated_cov = estimated_cov;
    o->subsample_factor = sub;
    o->n_kmers = orc_compute_dynamic_kmers(st->median_read_len, -1, o->kmers);
    o->confidence = conf;
    o->u_ge_1 = u_ge_1;
    o->u_ge_2 = u_ge_2;
    o->sample_kmer_depth = lambda_s;
    o->m_sample = m_sample;
}

/* estimator.rs:286-440 */
int orc_estimate_genome(const orc_arena* r1, const orc_arena* r2, const orc_count_stats* st,
                        uint64_t max_pairs, double target_cov, orc_genome_estimate* o) {
    memset(o, 0, sizeof *o);
    uint64_t avail = r1->n_reads < r2->n_reads ? r1->n_reads : r2->n_reads;
    uint64_t target = max_pairs;
    if (target > 150000) target = 150000;
    if (target > st->exact_read_pairs) target = st->exact_read_pairs;
    uint8_t* ht = (uint8_t*)calloc(1 << 14, 1);
    uint8_t* hs = (uint8_t*)calloc(1 << 14, 1);
    orc_cms* cms = orc_cms_new(1ULL << 23, 4);
    uint64_t ml = arena_max_len(r1), ml2 = arena_max_len(r2);
    uint8_t* buf = (uint8_t*)malloc((ml > ml2 ? ml : ml2) + 1);
    uint64_t m_sample = 0, pc = 0;
    while (pc < target) {
        if (pc >= avail) break; /* `_ => break` when a reader runs dry */
        est_read(r1, pc, buf, cms, ht, hs, &m_sample);
        est_read(r2, pc, buf, cms, ht, hs, &m_sample);
        pc++;
        if (pc >= 5000 && pc % 2000 == 0) {
            uint64_t u1 = orc_f64_to_u64(round(orc_hll_estimate(ht, 14)));
            uint64_t u2 = orc_f64_to_u64(round(orc_hll_estimate(hs, 14)));
            uint64_t e1 = sat_sub(u1, u2);
            uint64_t m_tr = sat_sub(m_sample, e1);
            double lam = u2 > 0 ? (double)m_tr / (double)u2 : 0.0;
            if (lam >= 15.0) break;
        }
    }
    int rc = 0;
    if (pc == 0) {
        rc = -1;
    } else {
        uint64_t u1 = orc_f64_to_u64(round(orc_hll_estimate(ht, 14)));
        uint64_t u2 = orc_f64_to_u64(round(orc_hll_estimate(hs, 14)));
        orc_estimate_finish(u1, u2, m_sample, st, target_cov, o);
        o->pairs_sampled = pc;
    }
    free(buf);
    orc_cms_free(cms);
    free(ht);
    free(hs);
    return rc;
}

void orc_estimator_sketch(const orc_arena* r1, const orc_arena* r2, uint64_t n_pairs, uint8_t* ht,
                          uint8_t* hs, uint64_t* m_sample) {
    orc_cms* cms = orc_cms_new(1ULL << 23, 4);
    uint64_t ml = arena_max_len(r1), ml2 = arena_max_len(r2);
    uint8_t* buf = (uint8_t*)malloc((ml > ml2 ? ml : ml2) + 1);
    *m_sample = 0;
    for (uint64_t pc = 0; pc < n_pairs; pc++) {
        est_read(r1, pc, buf, cms, ht, hs, m_sample);
        est_read(r2, pc, buf, cms, ht, hs, m_sample);
    }
    free(buf);
    orc_cms_free(cms);
}

/* ------------------------------------------------------------------------- */
/* Solid-set accumulator: open addressing over keys of nw words.             */
typedef struct {
    uint16_t cov;
    uint16_t edge[8];
} acc_t;

typedef struct {
    uint64_t cap, n;
    int nw;
    uint64_t* keys;
    acc_t* vals;
    uint8_t* used;
} accmap_t;

static void am_init(accmap_t* m, int nw, uint64_t cap) {
    uint64_t c = 1024;
    while (c < cap) c <<= 1;
    m->cap = c;
    m->n = 0;
    m->nw = nw;
    m->keys = (uint64_t*)malloc(c * nw * 8);
    m->vals = (acc_t*)calloc(c, sizeof(acc_t));
    m->used = (uint8_t*)calloc(c, 1);
}
static void am_free(accmap_t* m) {
    free(m->keys);
    free(m->vals);
    free(m->used);
    memset(m, 0, sizeof *m);
}
FORCE_INLINE uint64_t am_slot_hash(const km_t* key, int nw) { return sm64(km_fold(key, nw) ^ 0x9e3779b97f4a7c15ULL); }
static acc_t* am_entry(accmap_t* m, const km_t* key);
static void am_grow(accmap_t* m) {
    accmap_t n2;
    am_init(&n2, m->nw, m->cap * 2);
    for (uint64_t i = 0; i < m->cap; i++) {
        if (!m->used[i]) continue;
        km_t k = km_load(m->keys + i * m->nw, m->nw);
        *am_entry(&n2, &k) = m->vals[i];
    }
    am_free(m);
    *m = n2;
}
static acc_t* am_entry(accmap_t* m, const km_t* key) {
    if ((m->n + 1) * 2 > m->cap) am_grow(m);
    const int nw = m->nw;
    uint64_t s = am_slot_hash(key, nw) & (m->cap - 1);
    for (;;) {
        if (!m->used[s]) {
            m->used[s] = 1;
            memcpy(m->keys + s * nw, key->w, (size_t)nw * 8);
            m->n++;
            return &m->vals[s];
        }
        if (memcmp(m->keys + s * nw, key->w, (size_t)nw * 8) == 0) return &m->vals[s];
        s = (s + 1) & (m->cap - 1);
    }
}
FORCE_INLINE uint16_t sat_add16(uint16_t a, uint32_t b) {
    uint32_t s = (uint32_t)a + b;
    return s > 65535 ? 65535 : (uint16_t)s;
}
/* build.rs:183-187, 218-231 */
FORCE_INLINE void acc_push(accmap_t* m, const km_t* key, uint8_t edge, uint32_t cov_w, uint32_t edge_w) {
    acc_t* a = am_entry(m, key);
    a->cov = sat_add16(a->cov, cov_w);
    a->edge[edge] = sat_add16(a->edge[edge], edge_w);
}

void orc_build_params_default(orc_build_params* p, int target, uint64_t width) {
    p->width = width;
    p->counter_bits = 4;
    p->t_solid = target >= 0 ? target : 2;
    p->min_trans_cov = target >= 0 ? target : 3;
    p->hll_p = 0;
}

static uint64_t arenas_max_len(const orc_arena* arenas, int n) {
    uint64_t m = 0;
    for (int i = 0; i < n; i++) {
        uint64_t x = arena_max_len(&arenas[i]);
        if (x > m) m = x;
    }
    return m;
}

/* Pass 1 for one read (build.rs:97-113) plus the per-pass HLL over
 * splitmix64(canonical ntHash) of the same windows. */
FORCE_INLINE void pass1_read(const uint8_t* seq, uint32_t len, uint32_t k, int nw, orc_cms* cms,
                             uint8_t* hll, int hp) {
    km_t fwd = km_from_ascii(seq, k, nw);
    km_t rev = km_rev_comp(fwd, k, nw);
    const km_t* can = km_lt(&rev, &fwd, nw) ? &rev : &fwd;
    cms_add(cms, km_fold(can, nw));
    for (uint32_t i = k; i < len; i++) {
        km_roll(&fwd, &rev, A2B[seq[i]], k, nw);
        can = km_lt(&rev, &fwd, nw) ? &rev : &fwd;
        cms_add(cms, km_fold(can, nw));
    }
    if (hll) {
        nth_t h = {k, orc_nthash_fwd(seq, k), orc_nthash_rev(seq, k)};
        hll_add(hll, hp, sm64(h.f < h.r ? h.f : h.r));
        for (uint32_t i = k; i < len; i++) hll_add(hll, hp, sm64(nth_roll(&h, seq[i - k], seq[i])));
    }
}

/* Contig weight in pass 1 (build.rs:119-138): every k-mer added 10 times. */
static void pass1_contigs(const uint8_t* contigs, const uint32_t* lens, uint64_t n, uint32_t k, int nw,
                          orc_cms* cms) {
    const uint8_t* c = contigs;
    for (uint64_t ci = 0; ci < n; ci++) {
        uint32_t len = lens[ci];
        if (len >= k) {
            km_t fwd = km_from_ascii(c, k, nw);
            km_t rev = km_rev_comp(fwd, k, nw);
            const km_t* can = km_lt(&rev, &fwd, nw) ? &rev : &fwd;
            for (int t = 0; t < 10; t++) cms_add(cms, km_fold(can, nw));
            for (uint32_t i = k; i < len; i++) {
                km_roll(&fwd, &rev, A2B[c[i]], k, nw);
                can = km_lt(&rev, &fwd, nw) ? &rev : &fwd;
                for (int t = 0; t < 10; t++) cms_add(cms, km_fold(can, nw));
            }
        }
        c += len;
    }
}

/* Pass 2 contig backbone (build.rs:203-240) */
static void pass2_contigs(const uint8_t* contigs, const uint32_t* lens, uint64_t n, uint32_t k, int nw,
                          void (*push)(void*, const km_t*, uint8_t, uint32_t, uint32_t), void* ctx) {
    const uint8_t* c = contigs;
    for (uint64_t ci = 0; ci < n; ci++) {
        uint32_t len = lens[ci];
        if (len >= k + 1) {
            km_t uf = km_from_ascii(c, k, nw);
            km_t ur = km_rev_comp(uf, k, nw);
            int u_rc = km_lt(&ur, &uf, nw);
            for (uint32_t i = 0; i + k + 1 <= len; i++) {
                uint8_t base = A2B[c[i + k]];
                km_t vf = uf, vr = ur;
                km_roll(&vf, &vr, base, k, nw);
                int v_rc = km_lt(&vr, &vf, nw);
                uint8_t edge_u = !u_rc ? base : (uint8_t)(base + 4);
                push(ctx, u_rc ? &ur : &uf, edge_u, 100, 1000);
                uint8_t rc_p = (uint8_t)(3 - A2B[c[i]]);
                uint8_t edge_v = !v_rc ? (uint8_t)(rc_p + 4) : rc_p;
                push(ctx, v_rc ? &vr : &vf, edge_v, 100, 1000);
                uf = vf;
                ur = vr;
                u_rc = v_rc;
            }
        }
        c += len;
    }
}
static void push_single(void* ctx, const km_t* key, uint8_t e, uint32_t cw, uint32_t ew) {
    acc_push((accmap_t*)ctx, key, e, cw, ew);
}

/* Pass 2 for one read (build.rs:253-285); emit(u, edge) for both records. */
#define PASS2_READ(seq, len, k, nw, cms, t_solid, EMIT)                                   \
    do {                                                                                 \
        km_t uf = km_from_ascii(seq, k, nw);                                             \
        km_t ur = km_rev_comp(uf, k, nw);                                                \
        int u_rc = km_lt(&ur, &uf, nw);                                                  \
        uint32_t u_cnt = cms_est(cms, km_fold(u_rc ? &ur : &uf, nw));                    \
        for (uint32_t i = 0; i + k + 1 <= (len); i++) {                                  \
            uint8_t base = A2B[seq[i + k]];                                              \
            km_t vf = uf, vr = ur;                                                       \
            km_roll(&vf, &vr, base, k, nw);                                              \
            int v_rc = km_lt(&vr, &vf, nw);                                              \
            uint32_t v_cnt = cms_est(cms, km_fold(v_rc ? &vr : &vf, nw));                \
            if (u_cnt >= (uint32_t)(t_solid) && v_cnt >= (uint32_t)(t_solid)) {          \
                uint8_t edge_u = !u_rc ? base : (uint8_t)(base + 4);                     \
                EMIT(u_rc ? &ur : &uf, edge_u);                                          \
                uint8_t rc_p = (uint8_t)(3 - A2B[seq[i]]);                               \
                uint8_t edge_v = !v_rc ? (uint8_t)(rc_p + 4) : rc_p;                     \
                EMIT(v_rc ? &vr : &vf, edge_v);                                          \
            }                                                                            \
            uf = vf;                                                                     \
            ur = vr;                                                                     \
            u_rc = v_rc;                                                                 \
            u_cnt = v_cnt;                                                               \
        }                                                                                \
    } while (0)

/* Pass 3 (build.rs:302-328) into a record list. */
typedef struct {
    uint64_t n, cap;
    int nw;
    uint64_t* keys;
    uint16_t* cov;
    uint8_t* mask;
} rec_list;
static void rl_push(rec_list* r, const uint64_t* key, uint16_t cov, uint8_t mask) {
    if (r->n == r->cap) {
        r->cap = r->cap ? r->cap * 2 : 1024;
        r->keys = (uint64_t*)realloc(r->keys, r->cap * r->nw * 8);
        r->cov = (uint16_t*)realloc(r->cov, r->cap * 2);
        r->mask = (uint8_t*)realloc(r->mask, r->cap);
    }
    memcpy(r->keys + r->n * r->nw, key, (size_t)r->nw * 8);
    r->cov[r->n] = cov;
    r->mask[r->n] = mask;
    r->n++;
}
static void pass3_map(const accmap_t* m, int min_trans_cov, rec_list* out) {
    for (uint64_t i = 0; i < m->cap; i++) {
        if (!m->used[i]) continue;
        const acc_t* a = &m->vals[i];
        uint8_t mask = 0;
        for (int e = 0; e < 8; e++)
            if (a->edge[e] >= (uint16_t)min_trans_cov) mask |= (uint8_t)(1u << e);
        if (mask != 0) {
            uint16_t cov = (uint16_t)(a->cov / 2);
            if (cov < 1) cov = 1;
            rl_push(out, m->keys + i * m->nw, cov, mask);
        }
    }
}

/* Sort the record list by k-mer (parity form). */
static int memcmp_words(const uint64_t* x, const uint64_t* y, int n) {
    for (int w = 0; w < n; w++)
        if (x[w] != y[w]) return x[w] < y[w] ? -1 : 1;
    return 0;
}
typedef struct { const uint64_t* keys; int nw; } sort_ctx;
static int cmp_idx_r(const void* a, const void* b, void* cp) {
    const sort_ctx* c = (const sort_ctx*)cp;
    uint64_t i = *(const uint64_t*)a, j = *(const uint64_t*)b;
    return memcmp_words(c->keys + i * c->nw, c->keys + j * c->nw, c->nw);
}
static void rl_to_sorted_set(rec_list* r, orc_solid_set* out) {
    uint64_t n = r->n;
    int nw = r->nw;
    uint64_t* idx = (uint64_t*)malloc((n ? n : 1) * 8);
    for (uint64_t i = 0; i < n; i++) idx[i] = i;
    sort_ctx sc = {r->keys, nw};
    qsort_r(idx, n, 8, cmp_idx_r, &sc);
    out->n = n;
    out->nw = nw;
    out->keys = (uint64_t*)malloc((n ? n : 1) * nw * 8);
    out->cov = (uint16_t*)malloc((n ? n : 1) * 2);
    out->mask = (uint8_t*)malloc(n ? n : 1);
    for (uint64_t i = 0; i < n; i++) {
        memcpy(out->keys + i * nw, r->keys + idx[i] * nw, (size_t)nw * 8);
        out->cov[i] = r->cov[idx[i]];
        out->mask[i] = r->mask[idx[i]];
    }
    free(idx);
    free(r->keys);
    free(r->cov);
    free(r->mask);
    r->keys = NULL;
    r->cov = NULL;
    r->mask = NULL;
    r->n = r->cap = 0;
}

void orc_solid_free(orc_solid_set* s) {
    free(s->keys);
    free(s->cov);
    free(s->mask);
    memset(s, 0, sizeof *s);
}

static int check_k(uint32_t k) { return k >= 1 && k <= 127; }

int orc_build_solid(const orc_arena* arenas, int n_arenas, const uint8_t* contigs,
                    const uint32_t* contig_lens, uint64_t n_contigs, uint32_t k,
                    const orc_build_params* p, orc_cms** cms_out, uint8_t* hll_out, orc_solid_set* out) {
    if (!check_k(k)) return -1;
    const int nw = orc_kmer_words_for_k(k);
    orc_cms* cms = orc_cms_new(p->width, p->counter_bits);
    uint8_t* buf = (uint8_t*)malloc(arenas_max_len(arenas, n_arenas) + 1);
    if (p->hll_p <= 0) hll_out = NULL;
    /* pass 1 */
    for (int a = 0; a < n_arenas; a++) {
        const orc_arena* ar = &arenas[a];
        for (uint64_t id = 0; id < ar->n_reads; id++) {
            if (ar->lengths[id] < k) continue;
            orc_arena_extract_ascii(ar, id, buf);
            switch (nw) {
            case 1: pass1_read(buf, ar->lengths[id], k, 1, cms, hll_out, p->hll_p); break;
            case 2: pass1_read(buf, ar->lengths[id], k, 2, cms, hll_out, p->hll_p); break;
            default: pass1_read(buf, ar->lengths[id], k, 4, cms, hll_out, p->hll_p); break;
            }
        }
    }
    if (n_contigs) pass1_contigs(contigs, contig_lens, n_contigs, k, nw, cms);
    /* pass 2 */
    accmap_t m;
    am_init(&m, nw, 1 << 16);
    if (n_contigs) pass2_contigs(contigs, contig_lens, n_contigs, k, nw, push_single, &m);
#define EMIT_SINGLE(key, e) acc_push(&m, (key), (e), 1, 1)
    for (int a = 0; a < n_arenas; a++) {
        const orc_arena* ar = &arenas[a];
        for (uint64_t id = 0; id < ar->n_reads; id++) {
            uint32_t len = ar->lengths[id];
            if (len < k + 1) continue;
            orc_arena_extract_ascii(ar, id, buf);
            switch (nw) {
            case 1: PASS2_READ(buf, len, k, 1, cms, p->t_solid, EMIT_SINGLE); break;
            case 2: PASS2_READ(buf, len, k, 2, cms, p->t_solid, EMIT_SINGLE); break;
            default: PASS2_READ(buf, len, k, 4, cms, p->t_solid, EMIT_SINGLE); break;
            }
        }
    }
#undef EMIT_SINGLE
    /* pass 3 */
    rec_list rl = {0, 0, nw, NULL, NULL, NULL};
    pass3_map(&m, p->min_trans_cov, &rl);
    am_free(&m);
    rl_to_sorted_set(&rl, out);
    free(buf);
    if (cms_out) *cms_out = cms;
    else orc_cms_free(cms);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Reference-shaped threaded form (build.rs:93-116, 143-200, 243-289, 303-328) */
#define NUM_SHARDS 64
#define TBUF_CAP 2048
#define P2_CHUNK 4096
#define P1_CHUNK 256

typedef struct {
    pthread_mutex_t mu;
    accmap_t map;
} shard_t;

typedef struct {
    const orc_arena* arenas;
    int n_arenas;
    uint32_t k;
    int nw;
    const orc_build_params* p;
    orc_cms* cms;
    shard_t* shards;
    uint64_t maxlen;
    _Atomic uint64_t next;
    pthread_barrier_t bar;
    int nthreads;
    uint8_t** thread_hll;
    rec_list* shard_out;
    _Atomic int next_shard;
} mt_ctx;

typedef struct {
    km_t key[TBUF_CAP];
    uint8_t edge[TBUF_CAP];
    int n;
} tbuf_t;

/* ThreadBuffer::flush_shard (build.rs:177-189) */
static void tb_flush(mt_ctx* c, tbuf_t* tb, int s) {
    if (tb->n == 0) return;
    pthread_mutex_lock(&c->shards[s].mu);
    for (int i = 0; i < tb->n; i++) acc_push(&c->shards[s].map, &tb->key[i], tb->edge[i], 1, 1);
    pthread_mutex_unlock(&c->shards[s].mu);
    tb->n = 0;
}

typedef struct {
    mt_ctx* c;
    int tid;
} mt_arg;

/* All threads drain one arena, then meet; thread 0 resets the cursor. */
static void mt_arena_done(mt_ctx* c, int tid) {
    pthread_barrier_wait(&c->bar);
    if (tid == 0) atomic_store(&c->next, 0);
    pthread_barrier_wait(&c->bar);
}

static void* mt_worker(void* argp) {
    mt_arg* arg = (mt_arg*)argp;
    mt_ctx* c = arg->c;
    const uint32_t k = c->k;
    const int nw = c->nw;
    uint8_t* buf = (uint8_t*)malloc(c->maxlen + 1);
    uint8_t* hll = c->thread_hll ? c->thread_hll[arg->tid] : NULL;
    /* pass 1: each arena in turn, dynamic chunks of reads */
    for (int a = 0; a < c->n_arenas; a++) {
        const orc_arena* ar = &c->arenas[a];
        for (;;) {
            uint64_t start = atomic_fetch_add(&c->next, P1_CHUNK);
            if (start >= ar->n_reads) break;
            uint64_t end = start + P1_CHUNK < ar->n_reads ? start + P1_CHUNK : ar->n_reads;
            for (uint64_t id = start; id < end; id++) {
                if (ar->lengths[id] < k) continue;
                orc_arena_extract_ascii(ar, id, buf);
                switch (nw) {
                case 1: pass1_read(buf, ar->lengths[id], k, 1, c->cms, hll, c->p->hll_p); break;
                case 2: pass1_read(buf, ar->lengths[id], k, 2, c->cms, hll, c->p->hll_p); break;
                default: pass1_read(buf, ar->lengths[id], k, 4, c->cms, hll, c->p->hll_p); break;
                }
            }
        }
        mt_arena_done(c, arg->tid);
    }
    /* pass 2: 4096-read chunks, one thread buffer per chunk, flush_all at chunk end */
    tbuf_t* tbs = (tbuf_t*)malloc(sizeof(tbuf_t) * NUM_SHARDS);
    for (int a = 0; a < c->n_arenas; a++) {
        const orc_arena* ar = &c->arenas[a];
        for (;;) {
            uint64_t start = atomic_fetch_add(&c->next, P2_CHUNK);
            if (start >= ar->n_reads) break;
            uint64_t end = start + P2_CHUNK < ar->n_reads ? start + P2_CHUNK : ar->n_reads;
            for (int s = 0; s < NUM_SHARDS; s++) tbs[s].n = 0;
#define EMIT_MT(keyp, e)                                                      \
    do {                                                                      \
        const km_t* kp_ = (keyp);                                             \
        int s_ = (int)(sm64(km_fold(kp_, nw)) & (NUM_SHARDS - 1));            \
        tbs[s_].key[tbs[s_].n] = *kp_;                                        \
        tbs[s_].edge[tbs[s_].n] = (e);                                        \
        if (++tbs[s_].n >= TBUF_CAP) tb_flush(c, &tbs[s_], s_);               \
    } while (0)
            for (uint64_t id = start; id < end; id++) {
                uint32_t len = ar->lengths[id];
                if (len < k + 1) continue;
                orc_arena_extract_ascii(ar, id, buf);
                switch (nw) {
                case 1: PASS2_READ(buf, len, k, 1, c->cms, c->p->t_solid, EMIT_MT); break;
                case 2: PASS2_READ(buf, len, k, 2, c->cms, c->p->t_solid, EMIT_MT); break;
                default: PASS2_READ(buf, len, k, 4, c->cms, c->p->t_solid, EMIT_MT); break;
                }
            }
#undef EMIT_MT
            for (int s = 0; s < NUM_SHARDS; s++) tb_flush(c, &tbs[s], s);
        }
        mt_arena_done(c, arg->tid);
    }
    free(tbs);
    /* pass 3: per shard, each shard's records sorted; merged by the caller */
    for (;;) {
        int s = atomic_fetch_add(&c->next_shard, 1);
        if (s >= NUM_SHARDS) break;
        rec_list* r = &c->shard_out[s];
        pass3_map(&c->shards[s].map, c->p->min_trans_cov, r);
        am_free(&c->shards[s].map);
        orc_solid_set tmp;
        rl_to_sorted_set(r, &tmp);
        r->n = tmp.n;
        r->cap = tmp.n;
        r->keys = tmp.keys;
        r->cov = tmp.cov;
        r->mask = tmp.mask;
    }
    free(buf);
    return NULL;
}

static void heap_sift(int* heap, int hn, int j, const rec_list* sh, const uint64_t* pos, int nw) {
    for (;;) {
        int l = 2 * j + 1, r = l + 1, m = j;
        if (l < hn && memcmp_words(sh[heap[l]].keys + pos[heap[l]] * nw,
                                   sh[heap[m]].keys + pos[heap[m]] * nw, nw) < 0)
            m = l;
        if (r < hn && memcmp_words(sh[heap[r]].keys + pos[heap[r]] * nw,
                                   sh[heap[m]].keys + pos[heap[m]] * nw, nw) < 0)
            m = r;
        if (m == j) break;
        int t = heap[j];
        heap[j] = heap[m];
        heap[m] = t;
        j = m;
    }
}

/* k-way merge of the sorted shard lists into one sorted set */
static void merge_sorted_shards(const rec_list* sh, int nsh, int nw, orc_solid_set* out) {
    uint64_t total = 0;
    for (int s = 0; s < nsh; s++) total += sh[s].n;
    out->n = total;
    out->nw = nw;
    out->keys = (uint64_t*)malloc((total ? total : 1) * nw * 8);
    out->cov = (uint16_t*)malloc((total ? total : 1) * 2);
    out->mask = (uint8_t*)malloc(total ? total : 1);
    int heap[NUM_SHARDS];
    uint64_t pos[NUM_SHARDS];
    int hn = 0;
    for (int s = 0; s < nsh; s++) {
        pos[s] = 0;
        if (sh[s].n) heap[hn++] = s;
    }
    for (int i = hn / 2 - 1; i >= 0; i--) heap_sift(heap, hn, i, sh, pos, nw);
    uint64_t o = 0;
    while (hn > 0) {
        int s = heap[0];
        memcpy(out->keys + o * nw, sh[s].keys + pos[s] * nw, (size_t)nw * 8);
        out->cov[o] = sh[s].cov[pos[s]];
        out->mask[o] = sh[s].mask[pos[s]];
        o++;
        pos[s]++;
        if (pos[s] >= sh[s].n) heap[0] = heap[--hn];
        heap_sift(heap, hn, 0, sh, pos, nw);
    }
}

int orc_build_solid_mt(const orc_arena* arenas, int n_arenas, const uint8_t* contigs,
                       const uint32_t* contig_lens, uint64_t n_contigs, uint32_t k,
                       const orc_build_params* p, int threads, orc_cms** cms_out, uint8_t* hll_out,
                       orc_solid_set* out) {
    if (!check_k(k)) return -1;
    if (threads < 1) threads = 1;
    mt_ctx c;
    memset(&c, 0, sizeof c);
    c.arenas = arenas;
    c.n_arenas = n_arenas;
    c.k = k;
    c.nw = orc_kmer_words_for_k(k);
    c.p = p;
    c.cms = orc_cms_new(p->width, p->counter_bits);
    c.maxlen = arenas_max_len(arenas, n_arenas);
    c.nthreads = threads;
    c.shards = (shard_t*)calloc(NUM_SHARDS, sizeof(shard_t));
    c.shard_out = (rec_list*)calloc(NUM_SHARDS, sizeof(rec_list));
    for (int s = 0; s < NUM_SHARDS; s++) {
        pthread_mutex_init(&c.shards[s].mu, NULL);
        am_init(&c.shards[s].map, c.nw, 1 << 12);
        c.shard_out[s].nw = c.nw;
    }
    if (hll_out && p->hll_p > 0) {
        c.thread_hll = (uint8_t**)calloc(threads, sizeof(uint8_t*));
        for (int t = 0; t < threads; t++) c.thread_hll[t] = (uint8_t*)calloc((size_t)1 << p->hll_p, 1);
    }
    atomic_store(&c.next, 0);
    atomic_store(&c.next_shard, 0);
    pthread_barrier_init(&c.bar, NULL, (unsigned)threads);
    /* Contigs are serial in the reference (build.rs:119,203).  Their pass-1
     * adds only raise CMS counters and their pass-2 records do not read the
     * CMS, so both can run before the threaded read passes: saturating adds
     * commute, and the read passes see the same final CMS. */
    if (n_contigs) {
        pass1_contigs(contigs, contig_lens, n_contigs, k, c.nw, c.cms);
        accmap_t tmp;
        am_init(&tmp, c.nw, 1 << 12);
        pass2_contigs(contigs, contig_lens, n_contigs, k, c.nw, push_single, &tmp);
        for (uint64_t i = 0; i < tmp.cap; i++) {
            if (!tmp.used[i]) continue;
            km_t key = km_load(tmp.keys + i * c.nw, c.nw);
            int s = (int)(sm64(km_fold(&key, c.nw)) & (NUM_SHARDS - 1));
            acc_t* a = am_entry(&c.shards[s].map, &key);
            a->cov = sat_add16(a->cov, tmp.vals[i].cov);
            for (int e = 0; e < 8; e++) a->edge[e] = sat_add16(a->edge[e], tmp.vals[i].edge[e]);
        }
        am_free(&tmp);
    }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
    mt_arg* args = (mt_arg*)malloc(sizeof(mt_arg) * threads);
    for (int t = 0; t < threads; t++) {
        args[t].c = &c;
        args[t].tid = t;
        pthread_create(&th[t], NULL, mt_worker, &args[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    pthread_barrier_destroy(&c.bar);
    merge_sorted_shards(c.shard_out, NUM_SHARDS, c.nw, out);
    for (int s = 0; s < NUM_SHARDS; s++) {
        free(c.shard_out[s].keys);
        free(c.shard_out[s].cov);
        free(c.shard_out[s].mask);
        pthread_mutex_destroy(&c.shards[s].mu);
    }
    if (c.thread_hll) {
        memset(hll_out, 0, (size_t)1 << p->hll_p);
        for (int t = 0; t < threads; t++) {
            orc_hll_merge(hll_out, c.thread_hll[t], p->hll_p);
            free(c.thread_hll[t]);
        }
        free(c.thread_hll);
    }
    free(c.shards);
    free(c.shard_out);
    free(th);
    free(args);
    if (cms_out) *cms_out = c.cms;
    else orc_cms_free(c.cms);
    return 0;
}

/* ==========================================================================
 * Unitig compaction: src/graph/build.rs:330-597, on a sorted solid set.
 * ========================================================================== */

typedef struct {
    uint32_t target;   /* DirUnitigId */
    uint8_t ov;
    uint16_t support;
} cg_edge;
typedef struct {
    cg_edge* e;
    uint32_t n, cap;
} cg_elist;

typedef struct {
    const uint64_t* keys;
    const uint16_t* cov;
    const uint8_t* mask;
    uint8_t* visited;
    uint64_t n;
    uint32_t k;
    int nw;
    /* unitigs */
    uint8_t* seq;
    uint64_t seq_len, seq_cap;
    uint64_t* off;
    uint32_t* len;
    float* ucov;
    uint8_t* circ;
    cg_elist* efwd;
    cg_elist* erev;
    uint64_t nu, ucap;
} cg_t;

/* the k-mer's index in the sorted set, or -1 (hash-map lookup of the reference) */
static int64_t cg_find(const cg_t* g, const km_t* x) {
    uint64_t lo = 0, hi = g->n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        const km_t m = km_load(g->keys + mid * (uint64_t)g->nw, g->nw);
        if (km_lt(&m, x, g->nw)) lo = mid + 1;
        else hi = mid;
    }
    if (lo < g->n) {
        const km_t m = km_load(g->keys + lo * (uint64_t)g->nw, g->nw);
        if (km_eq(&m, x, g->nw)) return (int64_t)lo;
    }
    return -1;
}

/* base i (0 = first, most significant) of a k-mer: KmerInt::to_ascii */
static uint8_t km_base(const km_t* x, uint32_t k, int nw, uint32_t i) {
    const uint32_t bit = 2 * (k - 1 - i); /* from the least significant end */
    const uint32_t word = (uint32_t)nw - 1 - bit / 64;
    return (uint8_t)((x->w[word] >> (bit % 64)) & 3);
}
static const uint8_t CG_ASCII[4] = {'A', 'C', 'G', 'T'};
static uint8_t cg_comp(uint8_t c) {
    switch (c) {
        case 'A': return 'T';
        case 'C': return 'G';
        case 'G': return 'C';
        default: return 'A';
    }
}

static void cg_push_bytes(cg_t* g, const uint8_t* s, uint64_t n) {
    if (g->seq_len + n > g->seq_cap) {
        g->seq_cap = (g->seq_len + n) * 2 + 1024;
        g->seq = (uint8_t*)realloc(g->seq, g->seq_cap);
    }
    memcpy(g->seq + g->seq_len, s, n);
    g->seq_len += n;
}

static uint64_t cg_new_unitig(cg_t* g, const uint8_t* s, uint64_t n, float cov, int circular) {
    if (g->nu == g->ucap) {
        g->ucap = g->ucap ? g->ucap * 2 : 1024;
        g->off = (uint64_t*)realloc(g->off, g->ucap * 8);
        g->len = (uint32_t*)realloc(g->len, g->ucap * 4);
        g->ucov = (float*)realloc(g->ucov, g->ucap * 4);
        g->circ = (uint8_t*)realloc(g->circ, g->ucap);
        g->efwd = (cg_elist*)realloc(g->efwd, g->ucap * sizeof(cg_elist));
        g->erev = (cg_elist*)realloc(g->erev, g->ucap * sizeof(cg_elist));
    }
    const uint64_t id = g->nu++;
    g->off[id] = g->seq_len;
    g->len[id] = (uint32_t)n;
    g->ucov[id] = cov;
    g->circ[id] = (uint8_t)circular;
    memset(&g->efwd[id], 0, sizeof(cg_elist));
    memset(&g->erev[id], 0, sizeof(cg_elist));
    cg_push_bytes(g, s, n);
    return id;
}

/* UnitigNode::add_edge (cdbg.rs): an edge to the same target and overlap adds
 * its support (saturating u16). */
static void cg_add_edge(cg_elist* l, uint32_t target, uint8_t ov, uint16_t support) {
    for (uint32_t i = 0; i < l->n; i++) {
        if (l->e[i].target == target && l->e[i].ov == ov) {
            const uint32_t s = (uint32_t)l->e[i].support + support;
            l->e[i].support = (uint16_t)(s > 65535 ? 65535 : s);
            return;
        }
    }
    if (l->n == l->cap) {
        l->cap = l->cap ? l->cap * 2 : 4;
        l->e = (cg_edge*)realloc(l->e, l->cap * sizeof(cg_edge));
    }
    l->e[l->n].target = target;
    l->e[l->n].ov = ov;
    l->e[l->n].support = support;
    l->n++;
}

/* the walk's oriented (fwd, rev) pair of a directed k-mer */
static void cg_oriented(const cg_t* g, uint64_t i, int is_rc, km_t* fwd, km_t* rev) {
    const km_t x = km_load(g->keys + i * (uint64_t)g->nw, g->nw);
    const km_t r = km_rev_comp(x, g->k, g->nw);
    if (!is_rc) {
        *fwd = x;
        *rev = r;
    } else {
        *fwd = r;
        *rev = x;
    }
}

/* One step of the walk from the oriented (fwd, rev) by base b: the next
 * directed k-mer and its index (-1 outside the set); the twin edge on the
 * next k-mer is marked visited (build.rs:377-390, 425-432). */
static int64_t cg_step(cg_t* g, const km_t* fwd, const km_t* rev, uint8_t b, int* next_rc) {
    km_t nf = *fwd, nr = *rev;
    km_roll(&nf, &nr, b, g->k, g->nw);
    const int is_rc = km_lt(&nr, &nf, g->nw);
    const km_t can = is_rc ? nr : nf;
    const uint8_t rc_first = (uint8_t)(3 - km_base(fwd, g->k, g->nw, 0));
    const int twin_edge = !is_rc ? rc_first + 4 : rc_first;
    const int64_t j = cg_find(g, &can);
    if (j >= 0) g->visited[j] |= (uint8_t)(1u << twin_edge);
    *next_rc = is_rc;
    return j;
}

/* The walk loop over 1-in/1-out k-mers (build.rs:393-439 and 486-523).
 * stop_at >= 0: also ends when the canonical k-mer returns to stop_at. */
static int64_t cg_walk(cg_t* g, int64_t cur, int cur_rc, int64_t stop_at, uint8_t** path, uint64_t* plen,
                       uint64_t* pcap, uint64_t* total_cov, uint64_t* points, int* out_rc) {
    while (cur >= 0 && cur != stop_at) {
        const uint8_t m = g->mask[cur];
        const uint8_t out_m = !cur_rc ? (m & 0x0F) : ((m >> 4) & 0x0F);
        const uint8_t in_m = !cur_rc ? ((m >> 4) & 0x0F) : (m & 0x0F);
        if (__builtin_popcount(in_m) != 1 || __builtin_popcount(out_m) != 1) break;
        const uint8_t nb = (uint8_t)__builtin_ctz(out_m);
        const int edge_bit = !cur_rc ? nb : nb + 4;
        if (g->visited[cur] & (1u << edge_bit)) break;
        g->visited[cur] |= (uint8_t)(1u << edge_bit);
        km_t cf, cr;
        cg_oriented(g, (uint64_t)cur, cur_rc, &cf, &cr);
        int nrc = 0;
        const int64_t nxt = cg_step(g, &cf, &cr, nb, &nrc);
        if (*plen == *pcap) {
            *pcap *= 2;
            *path = (uint8_t*)realloc(*path, *pcap);
        }
        (*path)[(*plen)++] = CG_ASCII[nb];
        *total_cov += g->cov[cur];
        *points += 1;
        cur = nxt;
        cur_rc = nrc;
        if (cur < 0) break;
    }
    *out_rc = cur_rc;
    return cur;
}

int orc_compact(const uint64_t* keys, const uint16_t* cov, const uint8_t* mask, uint64_t n, uint32_t k,
                orc_graph* out) {
    memset(out, 0, sizeof *out);
    if (!check_k(k)) return -1;
    cg_t g;
    memset(&g, 0, sizeof g);
    g.keys = keys;
    g.cov = cov;
    g.mask = mask;
    g.n = n;
    g.k = k;
    g.nw = orc_kmer_words_for_k(k);
    g.visited = (uint8_t*)calloc(n ? n : 1, 1);
    uint64_t pcap = 1024, plen = 0;
    uint8_t* path = (uint8_t*)malloc(pcap);
    uint8_t* rcbuf = NULL;
    uint64_t rccap = 0;

    /* 4. walks from boundary directed k-mers (build.rs:332-449) */
    for (uint64_t i = 0; i < n; i++) {
        for (int urc = 0; urc < 2; urc++) {
            const uint8_t m0 = mask[i];
            const int fo = __builtin_popcount(m0 & 0x0F), fi = __builtin_popcount((m0 >> 4) & 0x0F);
            const int boundary = !urc ? (fi != 1 || fo != 1) : (fo != 1 || fi != 1);
            if (!boundary) continue;
            for (uint8_t b = 0; b < 4; b++) {
                const uint8_t m = mask[i];
                const uint8_t u_out = !urc ? (m & 0x0F) : ((m >> 4) & 0x0F);
                const uint8_t u_vis = !urc ? (g.visited[i] & 0x0F) : ((g.visited[i] >> 4) & 0x0F);
                if (!(u_out & (1u << b)) || (u_vis & (1u << b))) continue;
                g.visited[i] |= (uint8_t)(1u << (!urc ? b : b + 4));
                km_t fwd, rev;
                cg_oriented(&g, i, urc, &fwd, &rev);
                plen = 0;
                for (uint32_t j = 0; j < k; j++) path[plen++] = CG_ASCII[km_base(&fwd, k, g.nw, j)];
                int crc = 0;
                int64_t cur = cg_step(&g, &fwd, &rev, b, &crc);
                if (plen == pcap) {
                    pcap *= 2;
                    path = (uint8_t*)realloc(path, pcap);
                }
                path[plen++] = CG_ASCII[b];
                uint64_t total = cov[i], points = 1;
                int end_rc = 0;
                (void)cg_walk(&g, cur, crc, -1, &path, &plen, &pcap, &total, &points, &end_rc);
                if (plen > rccap) {
                    rccap = plen * 2;
                    rcbuf = (uint8_t*)realloc(rcbuf, rccap);
                }
                for (uint64_t j = 0; j < plen; j++) rcbuf[j] = cg_comp(path[plen - 1 - j]);
                const uint8_t* fin = memcmp(rcbuf, path, plen) < 0 ? rcbuf : path;
                float c = (float)total / (float)points;
                if (!(c >= 1.0f)) c = 1.0f;
                cg_new_unitig(&g, fin, plen, c, 0);
            }
        }
    }

    /* 5. isolated circular components (build.rs:452-549) */
    for (uint64_t i = 0; i < n; i++) {
        const uint8_t m = mask[i];
        const uint8_t fo = m & 0x0F, fi = (m >> 4) & 0x0F;
        if (__builtin_popcount(fo) != 1 || __builtin_popcount(fi) != 1) continue;
        const uint8_t b = (uint8_t)__builtin_ctz(fo);
        if (g.visited[i] & (1u << b)) continue;
        g.visited[i] |= (uint8_t)(1u << b);
        km_t fwd, rev;
        cg_oriented(&g, i, 0, &fwd, &rev);
        plen = 0;
        for (uint32_t j = 0; j < k; j++) path[plen++] = CG_ASCII[km_base(&fwd, k, g.nw, j)];
        int crc = 0;
        const int64_t cur = cg_step(&g, &fwd, &rev, b, &crc);
        if (plen == pcap) {
            pcap *= 2;
            path = (uint8_t*)realloc(path, pcap);
        }
        path[plen++] = CG_ASCII[b];
        uint64_t total = cov[i], points = 1;
        int end_rc = 0;
        const int64_t end = cg_walk(&g, cur, crc, (int64_t)i, &path, &plen, &pcap, &total, &points, &end_rc);
        if (end == (int64_t)i && plen > k) {
            if (memcmp(path + plen - k, path, k) == 0) plen -= k;
            else if (plen > k - 1 && memcmp(path + plen - (k - 1), path, k - 1) == 0) plen -= k - 1;
            float c = (float)total / (float)points;
            if (!(c >= 1.0f)) c = 1.0f;
            const uint64_t id = cg_new_unitig(&g, path, plen, c, 1);
            cg_add_edge(&g.efwd[id], (uint32_t)(id << 1), (uint8_t)(k - 1), 10);
        }
    }

    /* 6. edges by exact k and k-1 overlaps (build.rs:552-594) */
    const uint64_t nu = g.nu;
    for (int pass = 0; pass < 2; pass++) {
        const uint32_t ov = pass == 0 ? k : k - 1;
        if (ov == 0) continue;
        /* prefix map: (prefix, DirUnitigId) sorted by prefix then insertion order */
        uint64_t np = 0;
        uint32_t* ent = (uint32_t*)malloc((nu ? nu : 1) * 2 * sizeof(uint32_t));
        uint8_t* rcs = (uint8_t*)malloc((size_t)ov * (nu ? nu : 1));
        for (uint64_t id = 0; id < nu; id++) {
            if (g.len[id] < ov) continue;
            const uint8_t* s = g.seq + g.off[id];
            const uint32_t L = g.len[id];
            for (uint32_t j = 0; j < ov; j++) rcs[id * ov + j] = cg_comp(s[L - 1 - j]);
            ent[np++] = (uint32_t)(id << 1);
            ent[np++] = (uint32_t)(id << 1 | 1);
        }
        /* insertion sort is too slow for large graphs: sort by (prefix, dir id) */
        typedef struct { const uint8_t* p; uint32_t d; } pe_t;
        pe_t* pe = (pe_t*)malloc((np ? np : 1) * sizeof(pe_t));
        for (uint64_t q = 0; q < np; q++) {
            const uint32_t d = ent[q], id = d >> 1;
            pe[q].p = (d & 1) ? rcs + (uint64_t)id * ov : g.seq + g.off[id];
            pe[q].d = d;
        }
        /* merge sort on (prefix bytes, d) */
        for (uint64_t width = 1; width < np; width *= 2) {
            pe_t* tmp = (pe_t*)malloc(np * sizeof(pe_t));
            for (uint64_t lo = 0; lo < np; lo += 2 * width) {
                uint64_t mid = lo + width < np ? lo + width : np, hi = lo + 2 * width < np ? lo + 2 * width : np;
                uint64_t a = lo, bb = mid, o = lo;
                while (a < mid && bb < hi) {
                    const int c = memcmp(pe[a].p, pe[bb].p, ov);
                    if (c < 0 || (c == 0 && pe[a].d <= pe[bb].d)) tmp[o++] = pe[a++];
                    else tmp[o++] = pe[bb++];
                }
                while (a < mid) tmp[o++] = pe[a++];
                while (bb < hi) tmp[o++] = pe[bb++];
            }
            memcpy(pe, tmp, np * sizeof(pe_t));
            free(tmp);
        }
        for (uint64_t id = 0; id < nu; id++) {
            const uint32_t L = g.len[id];
            if (L < ov) continue;
            const uint8_t* s = g.seq + g.off[id];
            /* fwd suffix = s[L-ov..]; rc suffix = rc(s)[L-ov..] = rc of s[0..ov) */
            uint8_t* rcsuf = (uint8_t*)malloc(ov);
            for (uint32_t j = 0; j < ov; j++) rcsuf[j] = cg_comp(s[ov - 1 - j]);
            for (int side = 0; side < 2; side++) {
                const uint8_t* suf = side == 0 ? s + (L - ov) : rcsuf;
                uint64_t lo = 0, hi = np;
                while (lo < hi) {
                    const uint64_t mid = (lo + hi) >> 1;
                    if (memcmp(pe[mid].p, suf, ov) < 0) lo = mid + 1;
                    else hi = mid;
                }
                for (uint64_t q = lo; q < np && memcmp(pe[q].p, suf, ov) == 0; q++) {
                    const uint32_t t = pe[q].d;
                    if ((t >> 1) != id || g.circ[id])
                        cg_add_edge(side == 0 ? &g.efwd[id] : &g.erev[id], t, (uint8_t)ov, 10);
                }
            }
            free(rcsuf);
        }
        free(pe);
        free(rcs);
        free(ent);
    }

    /* flatten */
    out->n_unitigs = nu;
    out->seq = g.seq ? g.seq : (uint8_t*)malloc(1);
    out->offsets = g.off;
    out->lengths = g.len;
    out->cov = g.ucov;
    out->circular = g.circ;
    uint64_t ne = 0;
    for (uint64_t id = 0; id < nu; id++) ne += g.efwd[id].n + g.erev[id].n;
    out->n_edges = ne;
    out->e_src = (uint32_t*)malloc((ne ? ne : 1) * 4);
    out->e_dst = (uint32_t*)malloc((ne ? ne : 1) * 4);
    out->e_ov = (uint8_t*)malloc(ne ? ne : 1);
    out->e_support = (uint16_t*)malloc((ne ? ne : 1) * 2);
    uint64_t q = 0;
    for (uint64_t id = 0; id < nu; id++) {
        for (int side = 0; side < 2; side++) {
            const cg_elist* l = side == 0 ? &g.efwd[id] : &g.erev[id];
            for (uint32_t j = 0; j < l->n; j++, q++) {
                out->e_src[q] = (uint32_t)(id << 1 | (uint32_t)side);
                out->e_dst[q] = l->e[j].target;
                out->e_ov[q] = l->e[j].ov;
                out->e_support[q] = l->e[j].support;
            }
            free(l->e);
        }
    }
    free(g.efwd);
    free(g.erev);
    free(g.visited);
    free(path);
    free(rcbuf);
    return 0;
}

void orc_graph_free(orc_graph* g) {
    if (!g) return;
    free(g->seq);
    free(g->offsets);
    free(g->lengths);
    free(g->cov);
    free(g->circular);
    free(g->e_src);
    free(g->e_dst);
    free(g->e_ov);
    free(g->e_support);
    memset(g, 0, sizeof *g);
}

/* ------------------------------------------------------------------------- */
/* filter_active_reads, local_multik.rs:259-333.  The boundary set is a sorted
 * array of ntHash values (the reference uses an FxHashSet; membership is all
 * that matters). */
static int u64_cmp(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}
static int set_has(const uint64_t* s, uint64_t n, uint64_t h) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (s[mid] < h) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && s[lo] == h;
}
/* is_active closure, local_multik.rs:297-309 */
static int read_active(const uint8_t* seq, uint32_t len, uint32_t k, const uint64_t* set, uint64_t n) {
    if (len < k) return 0;
    uint32_t offs[3] = {0, (len - k) / 2, len - k};
    for (int i = 0; i < 3; i++)
        if (set_has(set, n, orc_nthash_canonical(seq + offs[i], k))) return 1;
    return 0;
}
int orc_filter_active_reads(const orc_graph_view* g, const orc_arena arenas[3], uint8_t* keep_merged,
                            uint8_t* keep_pairs) {
    const uint32_t k = g->k < 31 ? g->k : 31; /* local_multik.rs:261 */
    const uint64_t win = 350;                 /* boundary_window */
    if (k == 0) return -1;
    uint64_t cap = 0;
    for (uint64_t i = 0; i < g->n_nodes; i++)
        if (g->lengths[i] >= k) cap += g->lengths[i] - k + 1;
    uint64_t* set = (uint64_t*)malloc((cap ? cap : 1) * sizeof(uint64_t));
    if (!set) return -1;
    uint64_t n = 0;
    for (uint64_t i = 0; i < g->n_nodes; i++) {
        const uint8_t fl = g->flags ? g->flags[i] : 0;
        if (fl & 1) continue; /* is_deleted */
        const uint64_t L = g->lengths[i];
        if (L < k) continue;
        const uint8_t* s = g->seq + g->offsets[i];
        const uint32_t nf = g->n_fwd ? g->n_fwd[i] : 0, nr = g->n_rev ? g->n_rev[i] : 0;
        if (L < 2000 || (fl & 2) || nf > 1 || nr > 1) {
            for (uint64_t p = 0; p + k <= L; p++) set[n++] = orc_nthash_canonical(s + p, k);
        } else {
            uint64_t left_end = win < L - k + 1 ? win : L - k + 1;
            for (uint64_t p = 0; p < left_end; p++) set[n++] = orc_nthash_canonical(s + p, k);
            uint64_t rs = L > win ? L - win : 0;
            if (rs > L - k) rs = L - k;
            for (uint64_t p = rs; p <= L - k; p++) set[n++] = orc_nthash_canonical(s + p, k);
        }
    }
    qsort(set, n, sizeof(uint64_t), u64_cmp);
    uint64_t u = 0;
    for (uint64_t i = 0; i < n; i++)
        if (u == 0 || set[i] != set[u - 1]) set[u++] = set[i];
    uint32_t maxlen = 0;
    for (int a = 0; a < 3; a++)
        for (uint64_t i = 0; i < arenas[a].n_reads; i++)
            if (arenas[a].lengths[i] > maxlen) maxlen = arenas[a].lengths[i];
    uint8_t* b1 = (uint8_t*)malloc(maxlen + 1);
    uint8_t* b2 = (uint8_t*)malloc(maxlen + 1);
    if (!b1 || !b2) {
        free(set);
        free(b1);
        free(b2);
        return -1;
    }
    for (uint64_t i = 0; i < arenas[0].n_reads; i++) {
        orc_arena_extract_ascii(&arenas[0], i, b1);
        keep_merged[i] = (uint8_t)read_active(b1, arenas[0].lengths[i], k, set, u);
    }
    const uint64_t np = arenas[1].n_reads < arenas[2].n_reads ? arenas[1].n_reads : arenas[2].n_reads;
    for (uint64_t i = 0; i < np; i++) {
        orc_arena_extract_ascii(&arenas[1], i, b1);
        orc_arena_extract_ascii(&arenas[2], i, b2);
        keep_pairs[i] = (uint8_t)(read_active(b1, arenas[1].lengths[i], k, set, u) ||
                                  read_active(b2, arenas[2].lengths[i], k, set, u));
    }
    free(set);
    free(b1);
    free(b2);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Ingest, trim_merge.rs:17-417                                               */
static uint8_t up_ascii(uint8_t c) { return (c >= 'a' && c <= 'z') ? (uint8_t)(c - 32) : c; }
static int eq_nocase(uint8_t a, uint8_t b) { return up_ascii(a) == up_ascii(b); }
static const uint8_t AD_TRUSEQ[] = "AGATCGGAAGAGC";
static const uint8_t AD_NEXTERA[] = "AGATGTGTATAAGAGACAG";
static const uint8_t AD_NEXTERA_RC[] = "CTGTCTCTTATACACATCT";

/* at most one mismatch over the first 10 bases of the adapter */
static int adapter_at(const uint8_t* w, const uint8_t* ad) {
    int mm = 0;
    for (int j = 0; j < 10; j++)
        if (!eq_nocase(w[j], ad[j]) && ++mm > 1) return 0;
    return 1;
}
/* trim_merge.rs:17-75 */
uint32_t orc_trim_adapter(const uint8_t* seq, uint32_t len) {
    if (len < 10) return len;
    uint32_t cut = len;
    const uint32_t start = len > 100 ? len - 100 : 0;
    for (uint32_t i = start; i + 10 <= len; i++) {
        const uint8_t b0 = up_ascii(seq[i]), b1 = up_ascii(seq[i + 1]);
        const int agat = (b0 == 'A' && b1 == 'G') || b0 == 'A' || b1 == 'G';
        const int ctgt = (b0 == 'C' && b1 == 'T') || b0 == 'C' || b1 == 'T';
        if (!agat && !ctgt) continue;
        if (agat && (adapter_at(seq + i, AD_TRUSEQ) || adapter_at(seq + i, AD_NEXTERA)))
            if (i < cut) cut = i;
        if (ctgt && cut > i && adapter_at(seq + i, AD_NEXTERA_RC)) cut = i;
    }
    return cut;
}
/* trim_merge.rs:77-113 */
int64_t orc_trim_quality(const uint8_t* qual, uint32_t len, uint32_t qual_len, uint8_t min_qual, uint32_t min_len) {
    if (len < min_len || qual_len != len) return -1;
    uint32_t cut = len;
    if (qual_len >= 4) {
        uint64_t sum = 0;
        for (uint32_t j = cut - 4; j < cut; j++) sum += qual[j] > 33 ? qual[j] - 33 : 0;
        if (sum >= (uint64_t)min_qual * 4) return len;
        while (cut >= 4) {
            sum = 0;
            for (uint32_t j = cut - 4; j < cut; j++) sum += qual[j] > 33 ? qual[j] - 33 : 0;
            if (sum / 4 >= min_qual) break;
            cut--;
        }
    }
    return cut < min_len ? -1 : (int64_t)cut;
}
static uint8_t rc_ascii(uint8_t c) {
    switch (c) {
        case 'A': case 'a': return 'T';
        case 'C': case 'c': return 'G';
        case 'G': case 'g': return 'C';
        case 'T': case 't': return 'A';
        default: return 'A';
    }
}
/* trim_merge.rs:127-231 */
uint32_t orc_merge_read_pair(const uint8_t* r1, const uint8_t* q1, uint32_t l1, uint32_t ql1, const uint8_t* r2,
                             const uint8_t* q2, uint32_t l2, uint32_t ql2, uint32_t min_overlap,
                             double max_mismatch_ratio, uint8_t* out, uint32_t* mismatches) {
    if (l1 < min_overlap || l2 < min_overlap) return 0;
    uint8_t* rc2 = (uint8_t*)malloc(l2 ? l2 : 1);
    for (uint32_t i = 0; i < l2; i++) rc2[i] = rc_ascii(r2[l2 - 1 - i]);
    const uint32_t max_d = l1 < l2 ? l1 : l2;
    uint32_t best_d = 0, best_m = 0;
    double best_score = -INFINITY;
    for (uint32_t d = max_d; d >= min_overlap && d > 0; d--) {
        const uint8_t* s1 = r1 + (l1 - d);
        const uint64_t max_allowed = (uint64_t)((double)d * max_mismatch_ratio);
        uint64_t mm = 0;
        for (uint32_t i = 0; i < d; i++)
            if (!eq_nocase(s1[i], rc2[i]) && ++mm > max_allowed) break;
        if (mm <= max_allowed) {
            const double score = (double)d - 2.5 * (double)mm;
            if (score > best_score) {
                best_score = score;
                best_d = d;
                best_m = (uint32_t)mm;
            }
        }
        if ((double)d <= best_score) break;
    }
    if (best_d < min_overlap || best_d == 0) {
        free(rc2);
        return 0;
    }
    uint32_t o = 0;
    for (uint32_t i = 0; i < l1 - best_d; i++) out[o++] = r1[i];
    for (uint32_t i = 0; i < best_d; i++) {
        const uint8_t b1 = r1[l1 - best_d + i], b2 = rc2[i];
        const uint32_t i1 = l1 - best_d + i;
        const uint8_t qa = i1 < ql1 ? q1[i1] : 'I';
        const uint8_t qb = (i < l2 && l2 - 1 - i < ql2) ? q2[l2 - 1 - i] : 'I';
        if (eq_nocase(b1, b2) || qa >= qb) out[o++] = up_ascii(b1);
        else out[o++] = up_ascii(b2);
    }
    for (uint32_t i = best_d; i < l2; i++) out[o++] = rc2[i];
    if (mismatches) *mismatches = best_m;
    free(rc2);
    return best_d;
}

typedef struct { uint8_t* p; uint64_t n, cap; } bytes_t;
typedef struct { uint32_t* p; uint64_t n, cap; } u32s_t;
static void bytes_push(bytes_t* b, const uint8_t* s, uint64_t n) {
    if (b->n + n > b->cap) {
        b->cap = (b->n + n) * 2 + 1024;
        b->p = (uint8_t*)realloc(b->p, b->cap);
    }
    memcpy(b->p + b->n, s, n);
    b->n += n;
}
static void u32_push(u32s_t* v, uint32_t x) {
    if (v->n == v->cap) {
        v->cap = v->cap * 2 + 1024;
        v->p = (uint32_t*)realloc(v->p, v->cap * 4);
    }
    v->p[v->n++] = x;
}
/* trim_merge.rs:299-405 */
int orc_ingest_records(const orc_records* r1, const orc_records* r2, const orc_ingest_params* p, orc_ingest_out* out) {
    memset(out, 0, sizeof *out);
    if (r1->n != r2->n) return -1;
    bytes_t seq[3] = {{0}};
    u32s_t len[3] = {{0}};
    const int subsample = p->subsample_factor < 0.999;
    const double thr_f = p->subsample_factor * 18446744073709551616.0;
    const uint64_t keep_thr = thr_f <= 0.0 ? 0 : (thr_f >= 18446744073709551616.0 ? UINT64_MAX : (uint64_t)thr_f);
    uint8_t* mbuf = NULL;
    uint64_t mcap = 0;
    int rc = 0;
    for (uint64_t i = 0; i < r1->n; i++) {
        const uint64_t pair_idx = i + 1;
        const uint8_t* s1 = r1->seq + r1->seq_off[i];
        const uint8_t* s2 = r2->seq + r2->seq_off[i];
        const uint8_t* q1 = r1->qual + r1->qual_off[i];
        const uint8_t* q2 = r2->qual + r2->qual_off[i];
        uint32_t l1 = r1->seq_len[i], l2 = r2->seq_len[i];
        uint32_t ql1 = r1->qual_len[i], ql2 = r2->qual_len[i];
        out->stats.input_read_pairs++;
        out->stats.bases_before_trim += (uint64_t)l1 + l2;
        if (subsample && orc_splitmix64(pair_idx ^ p->seed) > keep_thr) continue;
        if (p->enable_trim) {
            const uint32_t a1 = orc_trim_adapter(s1, l1), a2 = orc_trim_adapter(s2, l2);
            if (ql1 < a1 || ql2 < a2) {
                rc = -3;
                break;
            }
            const int64_t t1 = orc_trim_quality(q1, a1, a1, p->min_qual, p->min_len);
            const int64_t t2 = orc_trim_quality(q2, a2, a2, p->min_qual, p->min_len);
            if (t1 < 0 || t2 < 0) continue;
            l1 = (uint32_t)t1;
            l2 = (uint32_t)t2;
            ql1 = l1;
            ql2 = l2;
        } else if (l1 < p->min_len || l2 < p->min_len) {
            continue;
        }
        out->stats.retained_read_pairs++;
        out->stats.bases_after_trim += (uint64_t)l1 + l2;
        uint32_t d = 0;
        if (p->enable_merge) {
            if ((uint64_t)l1 + l2 > mcap) {
                mcap = ((uint64_t)l1 + l2) * 2;
                mbuf = (uint8_t*)realloc(mbuf, mcap);
            }
            d = orc_merge_read_pair(s1, q1, l1, ql1, s2, q2, l2, ql2, 15, 0.10, mbuf, NULL);
        }
        if (d) {
            const uint32_t ml = l1 + l2 - d;
            out->stats.merged_pairs_count++;
            out->stats.merged_bases += ml;
            bytes_push(&seq[0], mbuf, ml);
            u32_push(&len[0], ml);
        } else {
            out->stats.unmerged_pairs_count++;
            bytes_push(&seq[1], s1, l1);
            u32_push(&len[1], l1);
            bytes_push(&seq[2], s2, l2);
            u32_push(&len[2], l2);
        }
    }
    free(mbuf);
    for (int a = 0; a < 3; a++) {
        out->seq[a] = seq[a].p;
        out->len[a] = len[a].p;
        out->n[a] = len[a].n;
    }
    if (rc) return rc;
    if (out->n[0] + out->n[1] + out->n[2] == 0) return -2;
    return 0;
}
void orc_ingest_free(orc_ingest_out* out) {
    for (int a = 0; a < 3; a++) {
        free(out->seq[a]);
        free(out->len[a]);
    }
    memset(out, 0, sizeof *out);
}

/* ------------------------------------------------------------------------- */
/* PE mapper, mapper.rs:89-279.  The index is a sorted array of (key, value);
 * the reference's FxHashMap only answers membership and the unique value. */
typedef struct { uint64_t key; uint32_t dir, pos; } tk_t;
static int tk_cmp(const void* a, const void* b) {
    const tk_t *x = (const tk_t*)a, *y = (const tk_t*)b;
    return x->key < y->key ? -1 : x->key > y->key;
}
static const tk_t* tk_find(const tk_t* t, uint64_t n, uint64_t key) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (t[mid].key < key) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && t[lo].key == key ? t + lo : NULL;
}
/* canonical k-mer of ascii[0..k): to_u64 of the canonical and whether it is the rc */
static uint64_t can_key(const uint8_t* ascii, uint32_t k, int nw, int* is_rc) {
    km_t f = km_from_ascii(ascii, k, nw);
    km_t r = km_rev_comp(f, k, nw);
    *is_rc = km_lt(&r, &f, nw);
    return *is_rc ? km_fold(&r, nw) : km_fold(&f, nw);
}
static int d_cmp(const void* a, const void* b) {
    const double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : x > y;
}
static int u32_cmp(const void* a, const void* b) {
    const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : x > y;
}
void orc_insert_dist_of(const uint32_t* sizes, uint64_t n, orc_insert_dist* o) {
    if (n < 20) {
        o->median = 450.0;
        o->mad = 50.0;
        o->min_insert = 250;
        o->max_insert = 700;
        return;
    }
    uint32_t* s = (uint32_t*)malloc(n * 4);
    double* d = (double*)malloc(n * sizeof(double));
    memcpy(s, sizes, n * 4);
    qsort(s, n, 4, u32_cmp);
    const double med = (double)s[n / 2];
    for (uint64_t i = 0; i < n; i++) d[i] = fabs((double)s[i] - med);
    qsort(d, n, sizeof(double), d_cmp);
    double mad = d[n / 2];
    if (mad < 20.0) mad = 20.0;
    double lo = med - 3.5 * mad;
    if (lo < 50.0) lo = 50.0;
    o->median = med;
    o->mad = mad;
    o->min_insert = (uint64_t)lo;
    o->max_insert = (uint64_t)(med + 3.5 * mad);
    free(s);
    free(d);
}
/* map_read closure, mapper.rs:150-162 */
static int map_read(const uint8_t* r, uint32_t len, uint32_t k, int nw, const tk_t* idx, uint64_t n,
                    uint32_t* dir, uint32_t* pos, uint32_t* step) {
    for (uint32_t st = 0; st <= 100; st += 20) {
        if (st + k > len) continue;
        int is_rc = 0;
        const tk_t* h = tk_find(idx, n, can_key(r + st, k, nw, &is_rc));
        if (h) {
            *dir = is_rc ? (h->dir ^ 1u) : h->dir;
            *pos = h->pos;
            *step = st;
            return 1;
        }
    }
    return 0;
}
int orc_map_paired_reads(const orc_graph_view* g, const orc_arena* pf, const orc_arena* pr, uint32_t k,
                         orc_pair_map* out) {
    memset(out, 0, sizeof *out);
    if (k < 1 || k > 127) return -1;
    const int nw = orc_kmer_words_for_k(k);
    /* index_terminal_kmers, mapper.rs:89-135 */
    uint64_t cap = 0;
    for (uint64_t i = 0; i < g->n_nodes; i++)
        if (!(g->flags && (g->flags[i] & 1)) && g->lengths[i] >= k) {
            const uint64_t t = g->lengths[i] - k + 1;
            cap += t <= 2000 ? t : 2000;
        }
    tk_t* all = (tk_t*)malloc((cap ? cap : 1) * sizeof(tk_t));
    uint64_t n = 0;
    for (uint64_t i = 0; i < g->n_nodes; i++) {
        if ((g->flags && (g->flags[i] & 1)) || g->lengths[i] < k) continue;
        const uint8_t* s = g->seq + g->offsets[i];
        const uint64_t t = g->lengths[i] - k + 1;
        for (uint64_t p = 0; p < t; p++) {
            if (t > 2000 && p >= 1000 && p < t - 1000) continue;
            int is_rc = 0;
            all[n].key = can_key(s + p, k, nw, &is_rc);
            all[n].dir = ((uint32_t)i << 1) | (uint32_t)is_rc;
            all[n].pos = (uint32_t)p;
            n++;
        }
    }
    qsort(all, n, sizeof(tk_t), tk_cmp);
    uint64_t u = 0;
    for (uint64_t i = 0; i < n;) {
        uint64_t j = i + 1;
        while (j < n && all[j].key == all[i].key) j++;
        if (j - i == 1) all[u++] = all[i];
        i = j;
    }
    /* map_paired_reads, mapper.rs:138-279 */
    const uint64_t np = pf->n_reads < pr->n_reads ? pf->n_reads : pr->n_reads;
    uint32_t maxlen = 1;
    for (uint64_t i = 0; i < np; i++) {
        if (pf->lengths[i] > maxlen) maxlen = pf->lengths[i];
        if (pr->lengths[i] > maxlen) maxlen = pr->lengths[i];
    }
    uint8_t* b1 = (uint8_t*)malloc(maxlen);
    uint8_t* b2 = (uint8_t*)malloc(maxlen);
    uint64_t lcap = 1024, scap = 1024;
    out->links = (orc_pair_link*)malloc(lcap * sizeof(orc_pair_link));
    out->sizes = (uint32_t*)malloc(scap * 4);
    for (uint64_t i = 0; i < np; i++) {
        const uint32_t l1 = pf->lengths[i], l2 = pr->lengths[i];
        if (l1 < k || l2 < k) continue;
        orc_arena_extract_ascii(pf, i, b1);
        orc_arena_extract_ascii(pr, i, b2);
        uint32_t u1, p1, s1, u2, p2, s2;
        if (!map_read(b1, l1, k, nw, all, u, &u1, &p1, &s1) || !map_read(b2, l2, k, nw, all, u, &u2, &p2, &s2))
            continue;
        const uint32_t u2f = u2 ^ 1u;
        if ((u1 >> 1) == (u2 >> 1)) {
            const uint32_t node_len = g->lengths[u1 >> 1];
            if (node_len < 1500) continue;
            int64_t span = -1;
            if (!(u1 & 1) && (u2 & 1)) span = ((int64_t)p2 + k + s2) - ((int64_t)p1 - s1);
            else if ((u1 & 1) && !(u2 & 1)) span = ((int64_t)p1 + k + s1) - ((int64_t)p2 - s2);
            int64_t min_len = l1 > l2 ? l1 : l2;
            if (min_len < 100) min_len = 100;
            if (span >= min_len && span <= 1500) {
                if (out->n_sizes == scap) out->sizes = (uint32_t*)realloc(out->sizes, (scap *= 2) * 4);
                out->sizes[out->n_sizes++] = (uint32_t)span;
            }
        } else {
            const int64_t n1 = g->lengths[u1 >> 1], n2 = g->lengths[u2 >> 1];
            int64_t d1 = (u1 & 1) ? (int64_t)p1 + k + s1 : n1 - ((int64_t)p1 - s1);
            int64_t d2 = (u2f & 1) ? n2 - ((int64_t)p2 - s2) : (int64_t)p2 + k + s2;
            if (d1 < 0) d1 = 0;
            if (d2 < 0) d2 = 0;
            if (d1 > UINT32_MAX || d2 > UINT32_MAX || l1 > 65535 || l2 > 65535) continue;
            if (out->n_links + 2 > lcap) out->links = (orc_pair_link*)realloc(out->links, (lcap *= 2) * sizeof(orc_pair_link));
            orc_pair_link a = {u1, u2f, (uint32_t)i, (uint32_t)d1, (uint32_t)d2, (uint16_t)l1, (uint16_t)l2};
            orc_pair_link b = {u2f ^ 1u, u1 ^ 1u, (uint32_t)i, (uint32_t)d2, (uint32_t)d1, (uint16_t)l2, (uint16_t)l1};
            out->links[out->n_links++] = a;
            out->links[out->n_links++] = b;
        }
    }
    orc_insert_dist_of(out->sizes, out->n_sizes, &out->dist);
    free(all);
    free(b1);
    free(b2);
    return 0;
}
void orc_pair_map_free(orc_pair_map* m) {
    free(m->links);
    free(m->sizes);
    memset(m, 0, sizeof *m);
}

/* ------------------------------------------------------------------------- */
/* Consensus polishing, consensus.rs:64-271.  Seeds are a sorted array of
 * (hash, node, pos, count); the reference's FxHashMaps answer the same
 * membership and unique-value queries. */
typedef struct { uint64_t h; uint32_t id, pos; uint32_t cnt; } seed_t;
static int seed_cmp(const void* a, const void* b) {
    const seed_t *x = (const seed_t*)a, *y = (const seed_t*)b;
    return x->h < y->h ? -1 : x->h > y->h;
}
static seed_t* seed_find(seed_t* t, uint64_t n, uint64_t h) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (t[mid].h < h) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && t[lo].h == h ? t + lo : NULL;
}
static int rc_exact(const uint8_t* anchor, const uint8_t* rk, uint32_t k) {
    for (uint32_t j = 0; j < k; j++) {
        uint8_t c;
        switch (rk[k - 1 - j]) {
            case 'A': case 'a': c = 'T'; break;
            case 'T': case 't': c = 'A'; break;
            case 'C': case 'c': c = 'G'; break;
            case 'G': case 'g': c = 'C'; break;
            default: return 0;
        }
        if (!eq_nocase(anchor[j], c)) return 0;
    }
    return 1;
}
int orc_polish_contigs(const orc_graph_view* g, const orc_arena* reads, int n_arenas, uint32_t min_output_len,
                       orc_polish_out* out) {
    memset(out, 0, sizeof *out);
    const uint32_t k = 31;
    const uint64_t U = g->n_nodes;
    uint64_t total = 0, nsamp = 0;
    for (uint64_t i = 0; i < U; i++) total += g->lengths[i];
#define ELIG(i) (!(g->flags && (g->flags[i] & 1)) && g->lengths[i] >= min_output_len)
    for (uint64_t i = 0; i < U; i++)
        if (ELIG(i) && g->lengths[i] >= k) nsamp += (g->lengths[i] - k) / 25 + 1;
    seed_t* sd = (seed_t*)malloc((nsamp ? nsamp : 1) * sizeof(seed_t));
    uint64_t n = 0;
    for (uint64_t i = 0; i < U; i++) {
        if (!ELIG(i) || g->lengths[i] < k) continue;
        const uint8_t* s = g->seq + g->offsets[i];
        for (uint32_t p = 0; p + k <= g->lengths[i]; p += 25) {
            sd[n].h = orc_nthash_canonical(s + p, k);
            sd[n].id = (uint32_t)i;
            sd[n].pos = p;
            sd[n].cnt = 0;
            n++;
        }
    }
    /* insert order: the last sampled (id, pos) of a hash is the stored one */
    for (uint64_t i = 0; i < n; i++) sd[i].cnt = (uint32_t)i;
    qsort(sd, n, sizeof(seed_t), seed_cmp);
    uint64_t u = 0;
    for (uint64_t i = 0; i < n;) {
        uint64_t j = i, last = i;
        while (j < n && sd[j].h == sd[i].h) {
            if (sd[j].cnt > sd[last].cnt) last = j;
            j++;
        }
        sd[u] = sd[last];
        sd[u].cnt = 0;
        u++;
        i = j;
    }
    for (uint64_t i = 0; i < U; i++) {
        if (!ELIG(i) || g->lengths[i] < k) continue;
        const uint8_t* s = g->seq + g->offsets[i];
        for (uint32_t p = 0; p + k <= g->lengths[i]; p++) {
            seed_t* e = seed_find(sd, u, orc_nthash_canonical(s + p, k));
            if (e && e->cnt < UINT32_MAX) e->cnt++;
        }
    }
    uint64_t v = 0;
    for (uint64_t i = 0; i < u; i++)
        if (sd[i].cnt == 1) sd[v++] = sd[i];
    /* pileups */
    uint64_t* poff = (uint64_t*)calloc(U + 1, 8);
    for (uint64_t i = 0; i < U; i++) poff[i + 1] = poff[i] + (ELIG(i) ? g->lengths[i] : 0);
    uint16_t* pile = (uint16_t*)calloc(poff[U] * 4 + 4, 2);
    uint32_t maxlen = 1;
    for (int a = 0; a < n_arenas; a++)
        for (uint64_t i = 0; i < reads[a].n_reads; i++)
            if (reads[a].lengths[i] > maxlen) maxlen = reads[a].lengths[i];
    uint8_t* r = (uint8_t*)malloc(maxlen);
    for (int a = 0; a < n_arenas; a++)
        for (uint64_t ri = 0; ri < reads[a].n_reads; ri++) {
            const uint32_t len = reads[a].lengths[ri];
            if (len < k) continue;
            orc_arena_extract_ascii(&reads[a], ri, r);
            int had = 0, orient = -1;
            uint32_t cid = 0, cpos = 0, rpos = 0;
            for (uint32_t p = 0; p + k <= len; p += 30) {
                const seed_t* e = seed_find(sd, v, orc_nthash_canonical(r + p, k));
                if (!e) continue;
                had = 1;
                const uint8_t* anc = g->seq + g->offsets[e->id] + e->pos;
                int fwd = 1;
                for (uint32_t j = 0; j < k && fwd; j++) fwd = eq_nocase(anc[j], r[p + j]);
                if (fwd || rc_exact(anc, r + p, k)) {
                    orient = fwd ? 0 : 1;
                    cid = e->id;
                    cpos = e->pos;
                    rpos = p;
                    break;
                }
            }
            if (orient < 0) {
                if (had) out->stats.rejected_anchors++;
                continue;
            }
            uint16_t* pl = pile + poff[cid] * 4;
            const int64_t plen = g->lengths[cid];
            int voted = 0;
            for (uint32_t idx = 0; idx < len; idx++) {
                const int64_t t = orient == 0 ? (int64_t)cpos - rpos + idx : (int64_t)cpos + k - 1 + rpos - idx;
                if (t < 0 || t >= plen) continue;
                uint32_t code = A2B[r[idx]];
                if (orient == 1) code = 3 - code;
                if (pl[t * 4 + code] < 65535) pl[t * 4 + code]++;
                if (orient == 0) out->stats.fwd_pileup_bases++;
                else out->stats.rev_pileup_bases++;
                voted = 1;
            }
            if (voted) {
                if (orient == 0) out->stats.fwd_mapped_reads++;
                else out->stats.rev_mapped_reads++;
            }
        }
    /* consensus */
    out->seq = (uint8_t*)malloc(total ? total : 1);
    uint64_t o = 0;
    for (uint64_t i = 0; i < U; i++) {
        const uint8_t* s = g->seq + g->offsets[i];
        memcpy(out->seq + o, s, g->lengths[i]);
        if (ELIG(i)) {
            const uint16_t* pl = pile + poff[i] * 4;
            for (uint32_t p = 0; p < g->lengths[i]; p++) {
                const uint16_t* c = pl + (uint64_t)p * 4;
                const uint16_t tot = (uint16_t)(c[0] + c[1] + c[2] + c[3]); /* u16 sum wraps */
                if (tot < 5) continue;
                uint32_t mi = 0;
                for (uint32_t b = 1; b < 4; b++)
                    if (c[b] > c[mi]) mi = b;
                if ((double)c[mi] / (double)tot > 0.75) {
                    const uint8_t nb = TWO_BIT_TO_ASCII[mi];
                    if (!eq_nocase(out->seq[o + p], nb)) {
                        out->seq[o + p] = nb;
                        out->stats.bases_modified++;
                    }
                }
            }
        }
        o += g->lengths[i];
    }
#undef ELIG
    free(sd);
    free(poff);
    free(pile);
    free(r);
    return 0;
}
void orc_polish_free(orc_polish_out* o) {
    free(o->seq);
    memset(o, 0, sizeof *o);
}
