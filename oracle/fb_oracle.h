/*
 * fb_oracle.h -- CPU restatement of fastBact's k-mer spectrum path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the timed
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg load it.  The product (fastbact_amd, libfastbact_gpu.so) never
 * links, loads or calls it.
 *
 * Every function restates the reference algorithm it names (paths relative
 * to the fastBact source tree).  Pinning of each function is listed in the
 * header comment of fb_oracle.c.
 */
#ifndef FB_ORACLE_H
#define FB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/common/sketch.rs:13-18 */
uint64_t orc_splitmix64(uint64_t z);
/* src/common/bitseq.rs:120-138 */
uint64_t orc_rev_comp_word(uint64_t word, uint32_t num_bases);
/* src/common/bitseq.rs:67-81 */
uint8_t orc_ascii_to_2bit(uint8_t c);

/* One BitSeqArena (src/common/bitseq.rs:145-149): data packs base g of the
 * concatenated stream at bits 2*(g&31) of data[g>>5]. */
typedef struct {
    const uint64_t* data;
    const uint64_t* offsets;
    const uint32_t* lengths;
    uint64_t n_reads;
} orc_arena;

/* Packs ASCII sequences into an arena exactly as BitSeqArena::push_seq
 * (src/common/bitseq.rs:195-222).  data must hold ceil(total/32) zeroed words. */
void orc_arena_pack(const uint8_t* ascii_concat, const uint32_t* lengths, uint64_t n,
                    uint64_t* data, uint64_t* offsets);
/* BitSeqArena::extract_ascii_into (src/common/bitseq.rs:286-299). */
void orc_arena_extract_ascii(const orc_arena* a, uint64_t id, uint8_t* out);

/* ---- KmerInt (src/common/kmer.rs:18-335) -----------------------------------
 * nw selects the reference type: 1 = u64 (k<=31), 2 = u128 (k<=63),
 * 4 = [u64;4] (k<=127).  Words are most-significant first: nw=2 stores
 * {hi, lo}; nw=4 stores words[0..4] as the reference does. */
int  orc_kmer_words_for_k(uint32_t k);
void orc_kmer_from_ascii(const uint8_t* ascii, uint32_t k, int nw, uint64_t* out);
void orc_kmer_rev_comp(const uint64_t* x, uint32_t k, int nw, uint64_t* out);
void orc_kmer_roll_next(const uint64_t* fwd, const uint64_t* rev, uint8_t base, uint32_t k,
                        int nw, uint64_t* out_fwd, uint64_t* out_rev);
uint64_t orc_kmer_to_u64(const uint64_t* x, int nw);
/* Canonical k-mers of every window of an ASCII sequence by the reference's
 * rolling loop (src/graph/build.rs:100-112): out[w*nw ..], is_rc[w]. */
void orc_kmer_canonical_windows(const uint8_t* ascii, uint32_t len, uint32_t k, int nw,
                                uint64_t* out, uint8_t* is_rc);

/* ---- ntHash (src/common/kmer.rs:385-477) ---------------------------------- */
uint64_t orc_nthash_fwd(const uint8_t* ascii, uint32_t k);
uint64_t orc_nthash_rev(const uint8_t* ascii, uint32_t k);
uint64_t orc_nthash_canonical(const uint8_t* ascii, uint32_t k);
/* Rolling canonical ntHash over every window (RollingNtHash::new + roll). */
void orc_nthash_rolling(const uint8_t* ascii, uint32_t len, uint32_t k, uint64_t* out);

/* ---- CountMinSketch (src/common/sketch.rs:121-254) -------------------------
 * counter_bits 4 is the reference layout (two nibbles per byte, low nibble =
 * even counter); counter_bits 16 is the same sketch with saturating u16
 * counters.  Table layout: counter idx = row*width + col. */
typedef struct orc_cms orc_cms;
orc_cms* orc_cms_new(uint64_t width, int counter_bits);
void     orc_cms_free(orc_cms* c);
uint64_t orc_cms_width(const orc_cms* c);
void     orc_cms_add(orc_cms* c, uint64_t key);
uint32_t orc_cms_estimate(const orc_cms* c, uint64_t key);
uint32_t orc_cms_add_and_estimate(orc_cms* c, uint64_t key);
/* Raw table bytes: width*4/2 (nibble) or width*4*2 (u16). */
size_t   orc_cms_table_bytes(const orc_cms* c);
const void* orc_cms_table(const orc_cms* c);
/* sketch.rs:165-168 */
uint64_t orc_cms_width_for_genome_size(uint64_t g);

/* ---- HyperLogLog (src/common/sketch.rs:24-115), precision p ---------------- */
void   orc_hll_add_hash(uint8_t* regs, int p, uint64_t h);
void   orc_hll_merge(uint8_t* regs, const uint8_t* other, int p);
double orc_hll_estimate(const uint8_t* regs, int p);

/* ---- Genome estimator (src/ingest/estimator.rs:104-161,167-283,286-440) --- */
typedef struct {
    uint64_t exact_read_pairs;
    uint64_t total_bases_r1, total_bases_r2, total_bases;
    uint64_t min_read_len, max_read_len;
    double   mean_read_len;
    uint64_t median_read_len;
} orc_count_stats;

typedef struct {
    uint64_t genome_size;
    uint64_t exact_read_pairs;
    uint64_t total_bases;
    uint64_t median_read_len;
    uint64_t max_read_len;
    double   estimated_cov;
    double   subsample_factor;
    uint32_t kmers[8];
    uint32_t n_kmers;
    int32_t  confidence;        /* 0 High, 1 LowCoverageWarning, 2 Unreliable */
    uint64_t u_ge_1, u_ge_2;
    double   sample_kmer_depth;
    uint64_t pairs_sampled;     /* pairs consumed before the loop ended */
    uint64_t m_sample;
} orc_genome_estimate;

/* Length statistics of count_fastx_exact from per-read lengths.  Returns 0,
 * or -1 when n_pairs == 0 (the reference's "0 valid read pairs" error). */
int orc_count_stats_from_lengths(const uint32_t* len_r1, const uint32_t* len_r2, uint64_t n_pairs,
                                 orc_count_stats* out);
/* compute_dynamic_kmers; target_max_k < 0 means None.  Returns count. */
uint32_t orc_compute_dynamic_kmers(uint64_t median_len, int64_t target_max_k, uint32_t* out8);
/* estimate_genome over the first pairs of two arenas (R1[i] pairs with R2[i]).
 * Returns 0, or -1 when no pair is available. */
int orc_estimate_genome(const orc_arena* r1, const orc_arena* r2, const orc_count_stats* stats,
                        uint64_t max_pairs, double target_cov, orc_genome_estimate* out);
/* Checkpoint-free view of the estimator's sketch loop over exactly n_pairs
 * pairs: final registers of hll_total and hll_solid (p=14). */
void orc_estimator_sketch(const orc_arena* r1, const orc_arena* r2, uint64_t n_pairs,
                          uint8_t* hll_total, uint8_t* hll_solid, uint64_t* m_sample);
/* The Poisson/confidence tail of estimate_genome (estimator.rs:385-439). */
void orc_estimate_finish(uint64_t u_ge_1, uint64_t u_ge_2, uint64_t m_sample,
                         const orc_count_stats* stats, double target_cov, orc_genome_estimate* out);
/* Rust's `f64 as u64` (saturating, NaN -> 0). */
uint64_t orc_f64_to_u64(double x);

/* ---- cDBG build passes 1-3 (src/graph/build.rs:67-328) --------------------
 * Output record of the solid set: the canonical k-mer (nw words, MSB first)
 * followed by cov (u16) and mask (u8); the parity form is sorted by k-mer. */
typedef struct {
    uint64_t width;              /* CMS row width (power of two) */
    int      counter_bits;       /* 4 (reference) or 16 */
    int      t_solid;            /* CMS gate for both ends of a pair */
    int      min_trans_cov;      /* edge gate */
    int      hll_p;              /* 0 = no per-pass HLL */
} orc_build_params;

typedef struct {
    uint64_t n;                  /* solid k-mers */
    int      nw;
    uint64_t* keys;              /* n*nw words */
    uint16_t* cov;
    uint8_t*  mask;
} orc_solid_set;

/* Reference defaults: t_solid = target.unwrap_or(2), min_trans_cov =
 * target.unwrap_or(3) (build.rs:141,302).  target < 0 means None. */
void orc_build_params_default(orc_build_params* p, int target_solid_cov, uint64_t width);

/* contigs: ASCII concatenation + lengths (build.rs:119-138,203-240). */
int orc_build_solid(const orc_arena* arenas, int n_arenas,
                    const uint8_t* contigs, const uint32_t* contig_lens, uint64_t n_contigs,
                    uint32_t k, const orc_build_params* p,
                    orc_cms** cms_out, uint8_t* hll_out, orc_solid_set* out);
/* Reference-shaped multi-threaded form of the same passes: atomic CAS CMS,
 * 64 mutex-sharded maps fed by 2048-entry thread buffers, 4096-read chunks. */
int orc_build_solid_mt(const orc_arena* arenas, int n_arenas,
                       const uint8_t* contigs, const uint32_t* contig_lens, uint64_t n_contigs,
                       uint32_t k, const orc_build_params* p, int threads,
                       orc_cms** cms_out, uint8_t* hll_out, orc_solid_set* out);
void orc_solid_free(orc_solid_set* s);

/* ---- Unitig compaction (src/graph/build.rs:330-597) -------------------------
 * Input: a solid set sorted by k-mer (keys nw words MSB first, cov, mask).
 * The reference walks its hash-map shards in their iteration order, which
 * decides the direction each linear unitig is walked in (its cov skips the
 * node it ends on), where a circular unitig starts, and the node numbering.
 * This restatement takes the k-mers in ascending order: boundary directed
 * k-mers forward then reverse per k-mer, cycles by their smallest k-mer.
 * Unitig i: seq[offsets[i] .. + lengths[i]) (ASCII), cov (f32), circular.
 * Edges in each node's insertion order, forward side then reverse side:
 * src = id << 1 | side, dst = DirUnitigId (id << 1 | is_rc), ov, support. */
typedef struct {
    uint64_t n_unitigs;
    uint8_t* seq;
    uint64_t* offsets;
    uint32_t* lengths;
    float* cov;
    uint8_t* circular;
    uint64_t n_edges;
    uint32_t* e_src;
    uint32_t* e_dst;
    uint8_t* e_ov;
    uint16_t* e_support;
} orc_graph;
int orc_compact(const uint64_t* keys, const uint16_t* cov, const uint8_t* mask, uint64_t n, uint32_t k,
                orc_graph* out);
void orc_graph_free(orc_graph* g);

/* ---- filter_active_reads (src/graph/local_multik.rs:259-333) ----------------
 * A CdbgGraph given as arrays: node i has ASCII seq[offsets[i] .. + lengths[i]),
 * flags bit 0 = is_deleted, bit 1 = is_circular, and edges_fwd.len() /
 * edges_rev.len() in n_fwd / n_rev.  The boundary set holds canonical ntHash
 * values at k' = min(k, 31) of every window of short (< 2000 bp), circular or
 * branching unitigs and of the 350-window ends of the others; a read is active
 * when its window at offset 0, (len-k')/2 or len-k' is in the set.  keep_merged
 * gets one byte per read of arenas[0]; keep_pairs one byte per pair
 * i < min(n(arenas[1]), n(arenas[2])), set when either mate is active. */
typedef struct {
    uint32_t k;
    uint64_t n_nodes;
    const uint8_t* seq;
    const uint64_t* offsets;
    const uint32_t* lengths;
    const uint8_t* flags;
    const uint32_t* n_fwd;
    const uint32_t* n_rev;
} orc_graph_view;
int orc_filter_active_reads(const orc_graph_view* g, const orc_arena arenas[3], uint8_t* keep_merged,
                            uint8_t* keep_pairs);

/* ---- Ingest: trimming, merging, subsampling (src/ingest/trim_merge.rs:17-417)
 * Records are given as parsed FASTQ/FASTA: seq and qual bytes (qual_len 0
 * for FASTA). */
/* trim_adapter (trim_merge.rs:17-75): the kept length of seq. */
uint32_t orc_trim_adapter(const uint8_t* seq, uint32_t len);
/* trim_quality (trim_merge.rs:77-113): the kept length, or -1 for None. */
int64_t orc_trim_quality(const uint8_t* qual, uint32_t len, uint32_t qual_len, uint8_t min_qual, uint32_t min_len);
/* merge_read_pair (trim_merge.rs:127-231): the overlap of the merge, or 0 when
 * the pair stays unmerged; out receives l1 + l2 - overlap bytes. */
uint32_t orc_merge_read_pair(const uint8_t* r1, const uint8_t* q1, uint32_t l1, uint32_t ql1, const uint8_t* r2,
                             const uint8_t* q2, uint32_t l2, uint32_t ql2, uint32_t min_overlap,
                             double max_mismatch_ratio, uint8_t* out, uint32_t* mismatches);
typedef struct {
    const uint8_t* seq;
    const uint64_t* seq_off;
    const uint32_t* seq_len;
    const uint8_t* qual;
    const uint64_t* qual_off;
    const uint32_t* qual_len;
    uint64_t n;
} orc_records;
typedef struct {
    double subsample_factor;
    uint32_t min_len;
    uint8_t min_qual;
    int enable_trim;
    int enable_merge;
    uint64_t seed;
} orc_ingest_params;
/* IngestStats (trim_merge.rs:256-265) */
typedef struct {
    uint64_t input_read_pairs, retained_read_pairs, bases_before_trim, bases_after_trim;
    uint64_t merged_pairs_count, merged_bases, unmerged_pairs_count;
} orc_ingest_stats;
/* The three read lists of the bank as ASCII: list a holds n[a] reads of
 * lengths len[a], concatenated in seq[a] (malloc'd; orc_ingest_free). */
typedef struct {
    uint8_t* seq[3];
    uint32_t* len[3];
    uint64_t n[3];
    orc_ingest_stats stats;
} orc_ingest_out;
/* ingest_reads_ext_with_seed's record loop (trim_merge.rs:299-405).
 * 0 ok; -1 mate counts differ; -2 no read survived; -3 a quality string is
 * shorter than its adapter-trimmed read (the reference's slice panics). */
int orc_ingest_records(const orc_records* r1, const orc_records* r2, const orc_ingest_params* p, orc_ingest_out* out);
void orc_ingest_free(orc_ingest_out* out);

/* ---- PE mapper (src/resolve/mapper.rs:89-279) -------------------------------
 * index_terminal_kmers over the graph's active unitigs (flags bit 0 deleted;
 * every window when a unitig has <= 2000, else the first and last 1000),
 * keeping keys (KmerInt::to_u64 of the canonical k-mer) seen once; then
 * map_paired_reads over pe_forward / pe_reverse.  Evidence records come in
 * pair order, two per linked pair: (u1, twin(u2)) then (u2, twin(u1)) as
 * DirUnitigIds; insert sizes in pair order. */
typedef struct {
    uint32_t u1, u2, pair_id, dist_u1_exit, dist_u2_entry;
    uint16_t r1_len, r2_len;
} orc_pair_link;
typedef struct {
    double median, mad;
    uint64_t min_insert, max_insert;
} orc_insert_dist;
typedef struct {
    uint64_t n_links, n_sizes;
    orc_pair_link* links;
    uint32_t* sizes;
    orc_insert_dist dist;
} orc_pair_map;
int orc_map_paired_reads(const orc_graph_view* g, const orc_arena* pe_forward, const orc_arena* pe_reverse,
                         uint32_t k, orc_pair_map* out);
void orc_pair_map_free(orc_pair_map* m);
/* InsertSizeDist from sizes (mapper.rs:255-276); fewer than 20: the default. */
void orc_insert_dist_of(const uint32_t* sizes, uint64_t n, orc_insert_dist* out);

/* ---- Consensus polishing (src/polish/consensus.rs:64-271) -------------------
 * Seeds: canonical ntHash (k 31) of every 25th window of the output-eligible
 * unitigs (not deleted, >= min_output_len), kept when the hash occurs once
 * over all their windows.  Reads (the bank's merged reads, then pe_forward)
 * anchor at the first seed among windows 0, 30, 60, .. that equals the
 * contig k-mer forward or as its reverse complement; each anchored read votes
 * its bases into u16 saturating pileups; a position with coverage >= 5 takes
 * the dominant base when it has > 75%.  out->seq receives every node's
 * sequence after polishing (the node's own bytes when unchanged). */
typedef struct {
    uint64_t bases_modified, fwd_mapped_reads, rev_mapped_reads, fwd_pileup_bases, rev_pileup_bases,
        rejected_anchors;
} orc_polish_stats;
typedef struct {
    uint8_t* seq; /* concatenated in node order, lengths unchanged */
    orc_polish_stats stats;
} orc_polish_out;
int orc_polish_contigs(const orc_graph_view* g, const orc_arena* reads, int n_arenas, uint32_t min_output_len,
                       orc_polish_out* out);
void orc_polish_free(orc_polish_out* o);

#ifdef __cplusplus
}
#endif
#endif
