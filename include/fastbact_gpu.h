/*
 * fastbact_gpu.h -- C ABI of the MI355X (gfx950) k-mer spectrum path.
 *
 * A drop-in for fastBact's k-mer spectrum hot path: rolling canonical k-mers
 * over 2-bit packed reads, the Count-Min Sketch, HyperLogLog registers, the
 * genome estimator's sketch loop and the solid-k-mer set that feeds unitig
 * compaction.  Plain pointers and sizes only.  Every call returns an int
 * status (FB_OK or a negative FB_ERR_*); fb_last_error() names the cause.  No
 * exception crosses the boundary.
 *
 * Threading: a context owns one HIP stream.  Calls on one context must not
 * overlap (lock externally when sharing a context between host threads).
 * Objects created from a context are used only with that context.
 *
 * Reference interfaces replaced (paths in the fastBact source tree):
 *   fb_bank_*           ReadBank / BitSeqArena           src/common/bitseq.rs:145-399
 *   fb_cms_*            CountMinSketch                   src/common/sketch.rs:121-254
 *   fb_hll_*            HyperLogLog                      src/common/sketch.rs:24-115
 *   fb_build_solid      build_cdbg_with_contigs_ext<K>   src/graph/build.rs:67-328
 *                       (passes 1-3; unitig compaction, build.rs:330-597, is downstream)
 *   fb_compact_*        unitig compaction (steps 4-6)    src/graph/build.rs:330-597
 *   fb_filter_active_reads  filter_active_reads          src/graph/local_multik.rs:259-333
 *   fb_upgrade_multik   upgrade_multik_ext<K2>           src/graph/local_multik.rs:38-147
 *   fb_ingest_*         ingest_reads_ext_with_seed       src/ingest/trim_merge.rs:17-417
 *   fb_map_paired_reads map_paired_reads<K>              src/resolve/mapper.rs:89-279
 *   fb_polish_contigs   polish_contigs                   src/polish/consensus.rs:64-271
 *   fb_fastx_*          open_fastx + record parsing      src/ingest/trim_merge.rs:233-252
 *   fb_estimate_genome  estimate_genome                  src/ingest/estimator.rs:286-440
 *   fb_count_stats_from_lengths  count_fastx_exact (length part)  src/ingest/estimator.rs:167-283
 *   fb_compute_dynamic_kmers                             src/ingest/estimator.rs:104-161
 */
#ifndef FASTBACT_GPU_H
#define FASTBACT_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FB_OK 0
#define FB_ERR_INVALID -1   /* bad argument (k out of range, null pointer, ...)    */
#define FB_ERR_OOM -2       /* device allocation failed or HBM budget exceeded      */
#define FB_ERR_OVERFLOW -3  /* solid table could not hold the keys (after growth)   */
#define FB_ERR_HIP -4       /* HIP runtime error                                    */
#define FB_ERR_EMPTY -5     /* no read pairs: "Input FASTQ files contain 0 valid read pairs" */
#define FB_ERR_NODEV -6     /* no usable GPU                                        */
#define FB_ERR_BUSY -7      /* serving queue full, or the server stopped            */
#define FB_ERR_TIMEOUT -8   /* fb_server_wait: the request has not finished yet     */

typedef struct fb_ctx fb_ctx;
typedef struct fb_bank fb_bank;
typedef struct fb_cms fb_cms;
typedef struct fb_hll fb_hll;
typedef struct fb_solid fb_solid;

/* One BitSeqArena (bitseq.rs:145-149): base g of the concatenated stream at
 * bits 2*(g&31) of data[g>>5]; reads contiguous, offsets[i] = base index. */
typedef struct {
  const uint64_t* data;
  uint64_t n_words;
  const uint64_t* offsets;
  const uint32_t* lengths;
  uint64_t n_reads;
} fb_arena_view;

/* ---- context -------------------------------------------------------------
 * hbm_budget_bytes caps this context's device allocations (0 = no cap). */
int fb_ctx_create(int device, uint64_t hbm_budget_bytes, fb_ctx** out);
void fb_ctx_destroy(fb_ctx* ctx);
const char* fb_last_error(const fb_ctx* ctx);
int fb_ctx_synchronize(fb_ctx* ctx);
/* Device bytes currently held by the context's objects and caches. */
uint64_t fb_ctx_device_bytes(const fb_ctx* ctx);
/* Returns the context's cached free buffers to the device (objects keep theirs). */
int fb_ctx_trim(fb_ctx* ctx);
/* Diagnostics, no counterpart in the reference: one plain device-to-device
 * copy of `bytes` (kernel k_copy_calibration reads and writes bytes rounded
 * up to 16), for scaling hardware traffic counters against a known size. */
int fb_diag_copy(fb_ctx* ctx, uint64_t bytes);
/* Diagnostics, no counterpart in the reference: the device's rate of random
 * 4-byte accesses over a table of 2^log2_counters u32 (10..30), n_updates
 * accesses at splitmix64-spread cells.  mode 0: atomicAdd, 1: atomicAdd with
 * its returned value used, 2: plain loads, 3: plain stores.  *ms = the time of one launch
 * (kernel k_atomic_rate) after a warm-up launch.  Small tables stay in L2,
 * mid-size ones in the infinity cache, large ones reach HBM. */
int fb_diag_atomic_rate(fb_ctx* ctx, uint32_t log2_counters, uint64_t n_updates, int mode, float* ms);
/* Free and total device memory of `device` in bytes. */
int fb_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes);
/* Library version string. */
const char* fb_version(void);

/* ---- ReadBank (bitseq.rs:350-399) -----------------------------------------
 * arenas[0] = merged_reads, [1] = pe_forward, [2] = pe_reverse; any may be
 * empty (n_reads = 0).  Host buffers are copied; the caller keeps ownership. */
int fb_bank_upload(fb_ctx* ctx, const fb_arena_view arenas[3], fb_bank** out);
void fb_bank_destroy(fb_bank* bank);
/* ReadBank::total_bases / total_reads */
uint64_t fb_bank_total_bases(const fb_bank* bank);
uint64_t fb_bank_total_reads(const fb_bank* bank);
/* Sum of lengths of reads with len >= k (the bases a k pass counts). */
uint64_t fb_bank_bases_for_k(const fb_bank* bank, uint32_t k);
/* Arena a (0 merged, 1 pe_forward, 2 pe_reverse): BitSeqArena::len(),
 * total_bases() (bitseq.rs:185-192) and data words. */
int fb_bank_arena_info(const fb_bank* bank, int arena, uint64_t* n_reads, uint64_t* total_bases, uint64_t* n_words);
/* Copies arena a back in the BitSeqArena layout (arena-local offsets);
 * sizes from fb_bank_arena_info, any pointer may be null. */
int fb_bank_download_arena(const fb_bank* bank, int arena, uint64_t* data, uint64_t* offsets, uint32_t* lengths);

/* ---- CountMinSketch (sketch.rs:121-254) -----------------------------------
 * width is rounded up to a power of two (sketch.rs:141-143).  counter_bits 4
 * is the reference's nibble sketch; 16 is the same sketch with saturating
 * u16 counters.  Counts are kept exact on the device; reads of the sketch
 * see min(count, 15) or min(count, 65535). */
int fb_cms_create(fb_ctx* ctx, uint64_t width, int counter_bits, fb_cms** out);
void fb_cms_destroy(fb_cms* cms);
uint64_t fb_cms_width(const fb_cms* cms);
/* CountMinSketch::width_for_genome_size (sketch.rs:165-168) */
uint64_t fb_cms_width_for_genome_size(uint64_t genome_size);
int fb_cms_clear(fb_cms* cms);
/* CountMinSketch::add for each key (host array). */
int fb_cms_add_keys(fb_cms* cms, const uint64_t* keys, uint64_t n);
/* CountMinSketch::estimate for each key. */
int fb_cms_query(fb_cms* cms, const uint64_t* keys, uint64_t n, uint32_t* out);
/* CountMinSketch::add_and_estimate (sketch.rs:248-253) for keys[0..n) in
 * order: out[i] is what add_and_estimate(keys[i]) returns when the keys are
 * added one at a time, earlier keys of the batch included (each row's cell
 * reads its count before the call plus the adds of keys 0..i landing in
 * it).  On the device: per row a stable sort of the batch by cell gives each
 * add its rank in its cell. */
int fb_cms_add_and_estimate(fb_cms* cms, const uint64_t* keys, uint64_t n, uint32_t* out);
/* fb_cms_add_and_estimate without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, hosts that forbid it).
 * Every buffer of the call (keys, estimates, the per-row cells and indices,
 * their sorted copies and the sort's space) lives in a static 16 MiB device
 * array that each GPU holds from its first context until the process ends.
 * At most 262144 keys per call (FB_ERR_INVALID beyond).  All calls on one
 * GPU share that array: calls that overlap on one GPU race, and the array
 * keeps the last call's contents.  Results equal fb_cms_add_and_estimate's. */
int fb_cms_add_and_estimate_static(fb_cms* cms, const uint64_t* keys, uint64_t n, uint32_t* out);
/* Pass 1 of build_cdbg_with_contigs_ext (build.rs:93-116): every canonical
 * k-mer of every read with len >= k, key = KmerInt::to_u64. */
int fb_cms_add_bank(fb_cms* cms, const fb_bank* bank, uint32_t k);
/* Table in the reference layout: nibble -> width*4/2 bytes (counter idx =
 * row*width+col, byte idx>>1, high nibble for odd idx); u16 -> width*4 u16. */
uint64_t fb_cms_table_bytes(const fb_cms* cms);
int fb_cms_download(fb_cms* cms, void* out, uint64_t bytes);

/* ---- HyperLogLog (sketch.rs:24-115), precision 4..18 ---------------------- */
int fb_hll_create(fb_ctx* ctx, int p, fb_hll** out);
void fb_hll_destroy(fb_hll* hll);
int fb_hll_clear(fb_hll* hll);
/* HyperLogLog::add_hash for each hash. */
int fb_hll_add_hashes(fb_hll* hll, const uint64_t* hashes, uint64_t n);
/* HyperLogLog::add_canonical_hash (sketch.rs:66-70) for each hash:
 * add_hash(splitmix64(h)). */
int fb_hll_add_canonical_hashes(fb_hll* hll, const uint64_t* hashes, uint64_t n);
/* HyperLogLog::merge (sketch.rs:73-80): dst's registers become the
 * register-wise max of dst and src.  FB_ERR_INVALID when the precisions
 * differ (the reference asserts equal register counts). */
int fb_hll_merge(fb_hll* dst, const fb_hll* src);
/* add_hash(splitmix64(canonical ntHash_k)) over every window of every read
 * with len >= k (HyperLogLog::add_canonical_hash, sketch.rs:68-70). */
int fb_hll_add_bank(fb_hll* hll, const fb_bank* bank, uint32_t k);
int fb_hll_registers(fb_hll* hll, uint8_t* out /* 2^p bytes */);
/* HyperLogLog::estimate (sketch.rs:83-114), generalised to m = 2^p. */
int fb_hll_estimate(fb_hll* hll, double* out);

/* ---- RollingNtHash (kmer.rs:433-477) ---------------------------------------
 * out[i] = canonical() of window i of the ASCII sequence (len - k + 1
 * windows; none when len < k): RollingNtHash::new on the first k bases,
 * then roll(out_base, in_base) once per window.  Non-ACGT bytes hash as A
 * (ASCII_TO_2BIT).  On the device, a thread per 64 windows. */
int fb_nthash_windows(fb_ctx* ctx, const uint8_t* seq, uint64_t len, uint32_t k, uint64_t* out);
/* The same estimator on a host register array. */
double fb_hll_estimate_registers(const uint8_t* regs, int p);

/* ---- Solid k-mer set: build_cdbg_with_contigs_ext passes 1-3 --------------- */
typedef struct {
  uint32_t k;                 /* 1..127 */
  int32_t target_solid_cov;   /* Option<u8>: < 0 = None (t_solid 2, edge gate 3) */
  uint64_t genome_size;       /* Option<u64>: 0 = None (bank bases / 60, build.rs:74-83) */
  uint64_t cms_width;         /* 0 = CountMinSketch::for_genome_size(est_g); else explicit */
  int32_t counter_bits;       /* 4 (reference nibble) or 16 (saturating u16) */
  int32_t hll_p;              /* 0 = no per-pass HLL; else precision of the pass HLL */
  uint32_t sort;              /* 1 = records sorted by k-mer (parity form) */
  uint32_t split;             /* threads per read; 0 = tuned default */
  uint32_t placement;         /* sketch and solid-table placement: 0 auto, 1 global tables, 2 LDS bins,
                                 3 LDS bins with the pass-2 gate applied per window (2 applies it per
                                 kept key wherever the counters' cap allows: the same result) */
  uint32_t bins2;             /* pass-2 LDS bins over all rounds (power of two <= 2^20); 0 = sized from the sketch */
  uint32_t chunks1;           /* thread shares per pass-1 scatter thread; 0 = tuned default */
  uint32_t chunks2;           /* thread shares per pass-2 count/scatter thread; 0 = tuned default */
} fb_build_params;

void fb_build_params_default(fb_build_params* p, uint32_t k);

/* contigs: prior unitig sequences as ASCII (concatenated) + lengths; may be
 * null/0.  Weights follow build.rs:119-138 (x10 CMS adds) and 203-240
 * (+100 cov, +1000 edge).  The result owns the pass's CMS, HLL and solid
 * set on the device. */
int fb_build_solid(fb_ctx* ctx, const fb_bank* bank, const uint8_t* contigs, const uint32_t* contig_lens,
                   uint64_t n_contigs, const fb_build_params* params, fb_solid** out);
void fb_solid_destroy(fb_solid* s);
uint64_t fb_solid_count(const fb_solid* s);
/* Words per key: 1 (k<=31, u64), 2 (k<=63, u128 {hi,lo}), 4 (k<=127, [u64;4]),
 * most significant first as the reference stores them. */
int fb_solid_key_words(const fb_solid* s);
int fb_solid_download(const fb_solid* s, uint64_t* keys, uint16_t* cov, uint8_t* mask);
/* The pass's sketches: table bytes as fb_cms_download, registers as fb_hll_registers. */
uint64_t fb_solid_cms_bytes(const fb_solid* s);
int fb_solid_cms_download(const fb_solid* s, void* out, uint64_t bytes);
int fb_solid_hll_registers(const fb_solid* s, uint8_t* out);
/* Timing of the build (HIP events on the context stream around the kernels),
 * milliseconds: [0] pass 1 kernel (CMS + HLL), [1] pass 2 kernel, [2] pass 3
 * + sort, [3] whole build. */
int fb_solid_timings(const fb_solid* s, float out[4]);
/* Kernel-level breakdown of the LDS-binned route, milliseconds (0 where the
 * build took another route): [0] pass-1 bin scatter, [1] pass-1 bin count
 * and gate bits, [2] pass-2 gate rows, record count and run offsets, [3]
 * pass-2 record scatter, [4] pass-2 bin tables with the fused pass 3.  No
 * counterpart in the reference (diagnostics of this implementation). */
int fb_solid_kernel_timings(const fb_solid* s, float out[5]);
/* Work counts of the build: [0] k-mer windows of pass 1, [1] pass-2 records
 * (sum of raw cov), [2] keys in the solid table, [3] table slots, [4] pass-2
 * route (0 global table, 1 LDS bins, 2 LDS bins abandoned for a full record
 * buffer, 3 abandoned for a full spill table, 4 abandoned for a full output
 * buffer; the abandoned routes finish on the global table), [5] pass-2 LDS bins over all
 * rounds, [6] keys per LDS bin table, [7] keys estimated from the pass-1
 * gate bits, [8] records the LDS route wrote (one per window with an edge,
 * 0 on another route), [9] reads of length >= k, [10] pass-2 rounds (the
 * LDS route splits the bins into rounds of at most 16384; 0 on another
 * route). */
int fb_solid_stats(const fb_solid* s, uint64_t out[11]);
/* Where the pass-2 Count-Min gate was applied: 0 per window pair (the
 * global route, or the LDS route with placement 3), 1 once per kept key and
 * low edge (the ungated LDS route: records of every window pair, the mask
 * and kept keys do not depend on the gate, cov's low edges are probed per
 * key).  Both give the reference's set.  Diagnostics, no counterpart in
 * the reference. */
int fb_solid_pass2_gate(const fb_solid* s);
/* Record layout of the LDS route's pass 2: 0 exact runs (a count walk
 * before the scatter), 1 chunked regions (no count walk: the ungated form
 * with the buffered scatter), 2 chunked regions that filled up, then exact
 * runs.  Every layout gives the same set.  Diagnostics, no counterpart in
 * the reference. */
int fb_solid_pass2_layout(const fb_solid* s);

/* ---- Unitig compaction: build_cdbg_with_contigs_ext step 4-6 (build.rs:330-597)
 * Walks from boundary directed k-mers through 1-in/1-out k-mers, isolated
 * cycles, then overlap edges (k and k-1).  The reference walks its hash-map
 * shards in their iteration order; here the k-mers are taken in ascending
 * order (boundary strands forward then reverse, bases 0..3; cycles by their
 * smallest k-mer), which fixes the direction each linear unitig is walked in
 * (its cov leaves out the node it ends on), where a circular unitig starts,
 * and the node numbering: linear unitigs in walk order, then circular ones.
 * The walks run on the device as list ranking, and so do the overlap edges
 * (every strand end against the hashed, sorted strand prefixes). */
typedef struct fb_graph fb_graph;
/* The solid set of a build (built with params sort = 1). */
int fb_compact_solid(const fb_solid* s, fb_graph** out);
/* A solid set given on the host: keys strictly ascending in the
 * fb_solid_download layout, with cov and mask per key. */
int fb_compact_arrays(fb_ctx* ctx, const uint64_t* keys, const uint16_t* cov, const uint8_t* mask, uint64_t n,
                      uint32_t k, fb_graph** out);
void fb_graph_destroy(fb_graph* g);
uint64_t fb_graph_unitig_count(const fb_graph* g);
uint64_t fb_graph_base_count(const fb_graph* g);
uint64_t fb_graph_edge_count(const fb_graph* g);
/* Unitig i: seq[offsets[i] .. + lengths[i]) as ASCII (CdbgGraph::get_unitig_seq),
 * UnitigNode::kmer_cov and is_circular. */
int fb_graph_unitigs(const fb_graph* g, uint8_t* seq, uint64_t* offsets, uint32_t* lengths, float* cov,
                     uint8_t* circular);
/* Edges of every node in insertion order, forward side then reverse side:
 * src = id << 1 | side, dst = DirUnitigId (id << 1 | is_rc), overlap, support. */
int fb_graph_edges(const fb_graph* g, uint32_t* src, uint32_t* dst, uint8_t* ov, uint16_t* support);
/* Milliseconds: [0] device walks, [1] downloads (sequences and edges), [2] device overlap edges. */
int fb_graph_timings(const fb_graph* g, float out[3]);

/* ---- Multi-k flow: filter_active_reads and upgrade_multik_ext ---------------
 * (src/graph/local_multik.rs:38-147,259-333)
 * A CdbgGraph as arrays, so that any graph (this library's compaction, or a
 * caller's simplified graph with deleted nodes) can drive the flow.  Node i:
 * ASCII seq[offsets[i] .. + lengths[i]), kmer_cov, flags (bit 0 is_deleted,
 * bit 1 is_circular), edges_fwd.len() and edges_rev.len().  cov, flags and
 * the edge counts may be null (read as 0). */
typedef struct {
  uint32_t k;
  uint64_t n_nodes;
  const uint8_t* seq;
  const uint64_t* offsets;
  const uint32_t* lengths;
  const float* cov;
  const uint8_t* flags;
  const uint32_t* n_edges_fwd;
  const uint32_t* n_edges_rev;
} fb_graph_view;
/* A view of a compacted graph (valid while g lives). */
int fb_graph_view_of(const fb_graph* g, fb_graph_view* out);
/* filter_active_reads: the reads of `bank` near unitig ends, branches,
 * short or circular unitigs, as a new device bank (pushed in order, as
 * BitSeqArena::push_seq packs them).  The boundary set holds the canonical
 * ntHash (k' = min(k, 31)) of every window of unitigs < 2000 bp, circular or
 * with > 1 edge on a side, and of the 350 windows at each end of the others;
 * a read is active when its window at 0, (len-k')/2 or len-k' is in the set;
 * a pair (i < min(pe_forward, pe_reverse reads)) is kept when either mate is.
 * keep_merged (one byte per merged read) and keep_pairs (one per pair) may
 * be null. */
int fb_filter_active_reads(fb_ctx* ctx, const fb_bank* bank, const fb_graph_view* g, fb_bank** out,
                           uint8_t* keep_merged, uint8_t* keep_pairs);

#define FB_MULTIK_K2_LE_PREV_K 1          /* early_stop "K2_LE_PREV_K" */
#define FB_MULTIK_NO_BRANCHING_REGIONS 2  /* early_stop "NO_BRANCHING_REGIONS" */
#define FB_GENOME_EXPLICIT_PASSED 1       /* genome_size_source "EXPLICIT_PASSED" */
#define FB_GENOME_FALLBACK_READ_BASES_DIV_60 2
/* MultiKUpgradeResult (local_multik.rs:12-23) plus the filter's counts. */
typedef struct {
  int32_t executed;
  int32_t early_stop;          /* 0 none, else FB_MULTIK_* */
  int32_t genome_size_source;  /* 0 skipped, else FB_GENOME_* */
  int32_t reserved;
  uint64_t genome_size_used;
  uint64_t cms_width;
  uint64_t cms_allocated_bytes;  /* CountMinSketch::allocated_bytes_for_genome_size (nibbles) */
  uint64_t active_unitigs_before;
  uint64_t active_edges_before;
  uint64_t active_unitigs_after;
  uint64_t active_edges_after;
  uint64_t kept_merged;           /* merged reads kept by the filter */
  uint64_t kept_pairs;            /* pairs kept by the filter */
  uint64_t prior_contigs;         /* active linear unitigs passed as contigs */
  uint64_t transferred_circular;  /* old circular unitigs added to the new graph */
} fb_multik_result;
/* upgrade_multik_ext<K2>: on a graph with branching, builds the solid set at
 * k2 from the filtered bank plus every active linear unitig as a prior contig
 * (x10 CMS adds, +100 cov, +1000 edges), target_solid_cov Some(4), genome
 * size max(genome_size, 50000) or, with genome_size 0, max(bank bases / 60,
 * 50000); compacts it; then appends each old circular unitig that no new
 * circular unitig matches within k2 + 10 bases.  simplify_graph (before and
 * after the transfer in the reference) is not part of this library: the
 * returned graph is the compaction itself.  params (may be null: the
 * reference's nibble sketch) sets counter_bits, cms_width, hll_p and the
 * launch shape; k, target, genome size and sort are set here.  On an early
 * stop (res->executed = 0) nothing is built and *out_graph is null.
 * out_solid and out_active (may be null) receive the k2 solid set and the
 * filtered bank, both on the device. */
int fb_upgrade_multik(fb_ctx* ctx, const fb_bank* bank, const fb_graph_view* prev, uint32_t k2, uint64_t genome_size,
                      const fb_build_params* params, fb_multik_result* res, fb_graph** out_graph, fb_solid** out_solid,
                      fb_bank** out_active);

/* ---- PE mapper (src/resolve/mapper.rs:89-279) --------------------------------
 * map_paired_reads<K>: index_terminal_kmers over the graph's active unitigs
 * (every window of a unitig with <= 2000, else its first and last 1000;
 * keys seen once), then each pair of pe_forward / pe_reverse mapped by the
 * first unique window among steps 0, 20, .., 100 of each mate.  The
 * reference's PairLinkMatrix is a hash map from (u1, u2) to a vector of
 * ReadPairEvidence; here it is the flat list of evidence records in pair
 * order (two per linked pair: (u1, twin(u2)) then (u2, twin(u1)), as
 * DirUnitigIds id << 1 | is_rc), which regroups into the same map. */
typedef struct {
  uint32_t u1, u2;            /* DirUnitigIds of the link */
  uint32_t pair_id;           /* ReadPairEvidence */
  uint32_t dist_u1_exit;
  uint32_t dist_u2_entry;
  uint16_t r1_len, r2_len;
} fb_pair_link;
/* InsertSizeDist (mapper.rs:12-30,255-276) and the sizes it came from. */
typedef struct {
  double median, mad;
  uint64_t min_insert, max_insert;
  uint64_t n_sizes;
} fb_insert_size_dist;
typedef struct fb_pair_map fb_pair_map;
int fb_map_paired_reads(fb_ctx* ctx, const fb_graph_view* g, const fb_bank* bank, uint32_t k, fb_pair_map** out);
uint64_t fb_pair_map_link_count(const fb_pair_map* m);
int fb_pair_map_links(const fb_pair_map* m, fb_pair_link* out);
/* Intra-unitig insert sizes in pair order. */
int fb_pair_map_sizes(const fb_pair_map* m, uint32_t* out);
int fb_pair_map_dist(const fb_pair_map* m, fb_insert_size_dist* out);
void fb_pair_map_destroy(fb_pair_map* m);

/* ---- Consensus polishing (src/polish/consensus.rs:64-271) -------------------
 * polish_contigs on the graph's output-eligible nodes (not deleted, at least
 * min_output_len bases) with the bank's merged reads and then pe_forward
 * reads, as main.rs:466-470 passes them: unique 31-mer ntHash seeds every 25
 * bp, read anchors at windows 0, 30, 60, .., saturating u16 pileups, and the
 * dominant base where coverage >= 5 and it has > 75%.  out_seq receives every
 * node's sequence in node order (the sum of g->lengths bytes; nodes that are
 * not eligible are copied unchanged). */
typedef struct {
  uint64_t bases_modified;
  uint64_t fwd_mapped_reads;
  uint64_t rev_mapped_reads;
  uint64_t fwd_pileup_bases;
  uint64_t rev_pileup_bases;
  uint64_t rejected_anchors;
} fb_polish_stats;
int fb_polish_contigs(fb_ctx* ctx, const fb_graph_view* g, const fb_bank* bank, uint32_t min_output_len,
                      uint8_t* out_seq, fb_polish_stats* stats);

/* ---- Ingest front end (src/ingest/trim_merge.rs:17-417) ---------------------
 * Parsed records: record i has seq[seq_offsets[i] .. + seq_lengths[i]) and
 * qual[qual_offsets[i] .. + qual_lengths[i]) (qual_lengths 0 for FASTA). */
typedef struct {
  const uint8_t* seq;
  const uint64_t* seq_offsets;
  const uint32_t* seq_lengths;
  const uint8_t* qual;
  const uint64_t* qual_offsets;
  const uint32_t* qual_lengths;
  uint64_t n;
} fb_records_view;

/* open_fastx + needletail's record reader (trim_merge.rs:233-252): FASTQ
 * (4-line records) or FASTA (multi-line), gzip when the name ends in ".gz".
 * Errors: FB_ERR_INVALID with fb_fastx_last_error() (per thread). */
typedef struct fb_fastx fb_fastx;
int fb_fastx_read(const char* path, fb_fastx** out);
const char* fb_fastx_last_error(void);
uint64_t fb_fastx_count(const fb_fastx* x);
/* Pointers into x (valid while x lives). */
int fb_fastx_view(const fb_fastx* x, fb_records_view* out);
void fb_fastx_destroy(fb_fastx* x);

/* ingest_reads_ext_with_seed's arguments; main.rs:215-224 passes min_len 30,
 * min_qual 20 (fb_ingest_params_default). */
typedef struct {
  double subsample_factor; /* < 0.999: keep pair i (1-based) iff splitmix64(i ^ seed) <= factor * 2^64 */
  uint32_t min_len;
  uint32_t min_qual;
  int32_t enable_trim;
  int32_t enable_merge;
  uint64_t seed;
} fb_ingest_params;
void fb_ingest_params_default(fb_ingest_params* p);
/* IngestStats (trim_merge.rs:256-265), then timings with no counterpart in
 * the reference (milliseconds: parsing on the host, the device steps). */
typedef struct {
  uint64_t input_read_pairs;
  uint64_t retained_read_pairs;
  uint64_t bases_before_trim;
  uint64_t bases_after_trim;
  uint64_t merged_pairs_count;
  uint64_t merged_bases;
  uint64_t unmerged_pairs_count;
  float parse_ms;
  float device_ms;
} fb_ingest_stats;
/* The record loop on the device: subsample keep-test, trim_adapter,
 * trim_quality, merge_read_pair (min overlap 15, 10% mismatches), then the
 * reads pushed into a new device bank as the reference pushes them (merged
 * reads, and the unmerged mates in pe_forward / pe_reverse, in pair order).
 * Errors as the reference's: FB_ERR_INVALID for unequal mate counts (and,
 * where the reference's slice would panic, a quality string shorter than its
 * adapter-clipped read), FB_ERR_EMPTY when no read survives. */
int fb_ingest_records(fb_ctx* ctx, const fb_records_view* r1, const fb_records_view* r2, const fb_ingest_params* p,
                      fb_bank** out, fb_ingest_stats* stats);
/* ingest_reads_ext_with_seed: both files parsed (on two host threads), then
 * fb_ingest_records. */
int fb_ingest_files(fb_ctx* ctx, const char* r1_path, const char* r2_path, const fb_ingest_params* p, fb_bank** out,
                    fb_ingest_stats* stats);

/* ---- Batched form: many isolates per launch --------------------------------
 * The reference builds one isolate per process run (main.rs:262).  A batch
 * holds n isolates' banks and per-isolate sketches and tables; one
 * fb_batch_run executes passes 1-3 for every isolate with one launch per
 * kernel (blocks map to isolates through a prefix table, no tile crosses
 * isolates).  A run is two HIP graphs, captured on the first run at a k and
 * replayed while the launch shapes hold: pass 1 with the gate occupancy of
 * every isolate, then (each isolate's pass 2 sized from its occupancy) passes
 * 2 and 3 on LDS bins.  An isolate whose LDS route overflows re-runs pass 2
 * alone on a global table.  Per-isolate results equal fb_build_solid's on
 * the same isolate.  genome_sizes[i] sizes isolate i's CMS when
 * params->cms_width is 0 (build.rs:74-84); params->k is ignored. */
typedef struct fb_batch fb_batch;
int fb_batch_create(fb_ctx* ctx, const fb_bank* const* banks, const uint64_t* genome_sizes, uint32_t n,
                    const fb_build_params* params, fb_batch** out);
void fb_batch_destroy(fb_batch* b);
/* Device bytes one isolate of this shape needs at k (sketches, table,
 * outputs, scratch) -- the admission footprint of the serving loop. */
uint64_t fb_batch_footprint(uint64_t genome_size, uint64_t bases, uint64_t reads, const fb_build_params* params,
                            uint32_t k);
/* Output sizing, host arithmetic only (no device; no counterpart in the
 * reference).  An isolate of the batched and served forms keeps its solid
 * output buffers across the k ladder while an entry widens with k (1, 2 or 4
 * key words), so buffers are sized and checked in bytes at the k of each
 * launch.  fb_solid_out_bytes: out[0] keys = cap * nref * 8, out[1] cov =
 * cap * 2, out[2] mask = cap.  fb_solid_out_fits: 1 when buffers of have[0..2]
 * bytes hold cap entries of nref words; the library refuses (FB_ERR_INVALID)
 * every launch that writes solid outputs unless this holds. */
void fb_solid_out_bytes(uint64_t cap, uint32_t nref, uint64_t out[3]);
int fb_solid_out_fits(uint64_t cap, uint32_t nref, const uint64_t have[3]);
/* Pass-2 plan of the LDS route for an isolate with est_keys estimated keys
 * (the sizing fb_batch_footprint uses; a workgroup granted 160 KB of LDS):
 * bins, rounds, record/spill/output capacities in entries. */
typedef struct {
  uint64_t nb2, nbr, spill_cap, out_cap, rec_cap, n_blocks2, threads2;
  uint32_t log2_rounds, split2, chunks2, lds_cap, ok;
} fb_pass2_plan;
int fb_plan_pass2(double est_keys, uint64_t windows, uint64_t n_reads, uint32_t max_len, uint32_t k,
                  fb_pass2_plan* out);
int fb_batch_run(fb_batch* b, uint32_t k, int32_t target_solid_cov);
uint32_t fb_batch_size(const fb_batch* b);
uint64_t fb_batch_count(const fb_batch* b, uint32_t iso);
int fb_batch_key_words(const fb_batch* b);
int fb_batch_download(const fb_batch* b, uint32_t iso, uint64_t* keys, uint16_t* cov, uint8_t* mask);
uint64_t fb_batch_cms_bytes(const fb_batch* b, uint32_t iso);
int fb_batch_cms_download(const fb_batch* b, uint32_t iso, void* out, uint64_t bytes);
int fb_batch_hll_registers(const fb_batch* b, uint32_t iso, uint8_t* out);
/* Last run, milliseconds: [0] graphs (passes 1-3, all isolates), [1]
 * global-table re-runs, [2] sort, [3] whole run. */
int fb_batch_timings(const fb_batch* b, float out[4]);
/* Pass-2 route of every isolate in the last run (n entries, codes as
 * fb_solid_stats[4]).  Diagnostics, no counterpart in the reference. */
int fb_batch_routes(const fb_batch* b, uint32_t* out);

/* ---- Autotuner (no counterpart in the reference) ----------------------------
 * Picks the launch shapes of a batch at k by timing: pass-1 placement (LDS
 * bins or global atomics), threads per read, and thread shares per scatter
 * thread of passes 1 and 2, by coordinate descent from the defaults.  Fields
 * the batch's params set explicitly are not tuned.  Every candidate must give
 * the same solid-set sizes as the defaults (FB_ERR_INVALID otherwise).  The
 * choice is kept for the batch at k and, for the process, for every batch of
 * the same shape (k-mer width, counter bits, sketch width, bases per
 * isolate and isolate count, each rounded to a power of two); batches and
 * serving iterations of that shape use it unless their params say otherwise. */
typedef struct {
  uint32_t placement, split, chunks1, chunks2;  /* the choice */
  float best_ms, default_ms;                    /* whole-run times of the choice and of the defaults */
  uint32_t trials;                              /* candidates timed */
} fb_tune_result;
int fb_batch_autotune(fb_batch* b, uint32_t k, int32_t target_solid_cov, fb_tune_result* out);
/* The process's tuned choices as text (one line per shape), read back by load. */
int fb_autotune_save(const char* path);
int fb_autotune_load(const char* path);
void fb_autotune_clear(void);

/* ---- Genome estimator (estimator.rs) -------------------------------------- */
typedef struct {
  uint64_t exact_read_pairs;
  uint64_t total_bases_r1, total_bases_r2, total_bases;
  uint64_t min_read_len, max_read_len;
  double mean_read_len;
  uint64_t median_read_len;
} fb_count_stats;

typedef struct {
  uint64_t genome_size;
  uint64_t exact_read_pairs;
  uint64_t total_bases;
  uint64_t median_read_len;
  uint64_t max_read_len;
  double estimated_cov;
  double subsample_factor;
  uint32_t kmers[8];
  uint32_t n_kmers;
  int32_t confidence;  /* 0 HIGH, 1 LOW_COVERAGE_WARNING, 2 UNRELIABLE */
  uint64_t u_ge_1, u_ge_2;
  double sample_kmer_depth;
  uint64_t pairs_sampled;
  uint64_t m_sample;
} fb_genome_estimate;

/* count_fastx_exact's statistics from per-read lengths of the pairs.
 * FB_ERR_EMPTY when n_pairs == 0. */
int fb_count_stats_from_lengths(const uint32_t* len_r1, const uint32_t* len_r2, uint64_t n_pairs, fb_count_stats* out);
/* count_fastx_exact (estimator.rs:167-283) on two FASTQ/FASTA (.gz) files:
 * the pairing and read-ID prefix checks (extract_read_id_prefix,
 * estimator.rs:89-99) and the length statistics.  Errors: FB_ERR_INVALID or
 * FB_ERR_EMPTY with fb_fastx_last_error(). */
int fb_count_fastx(const char* r1_path, const char* r2_path, fb_count_stats* out);
/* estimate_genome (estimator.rs:286-440) on two read files. */
int fb_estimate_genome_files(fb_ctx* ctx, const char* r1_path, const char* r2_path, uint64_t max_pairs,
                             double target_cov, fb_genome_estimate* out);
/* compute_dynamic_kmers; target_max_k < 0 = None.  Returns the count (<= 8). */
uint32_t fb_compute_dynamic_kmers(uint64_t median_len, int64_t target_max_k, uint32_t* out8);
/* estimate_genome over R1[i]/R2[i] pairs (ASCII->2-bit is lossless for the
 * estimator: N maps to A before hashing, bitseq.rs:67-81). */
int fb_estimate_genome(fb_ctx* ctx, const fb_arena_view* r1, const fb_arena_view* r2, const fb_count_stats* stats,
                       uint64_t max_pairs, double target_cov, fb_genome_estimate* out);
/* The pipeline's subsampling policy with auto-subsampling enabled
 * (main.rs:66-185; host arithmetic only): coverage = total_bases /
 * genome_size when genome_size > 0, else the estimate's estimated_cov when its
 * confidence is HIGH (LOW_COVERAGE_WARNING and UNRELIABLE keep every read);
 * factor = clamp(depth / coverage, 0.01, 1) when depth > 0 and coverage >
 * threshold_ratio * depth (the hidden --subsample-threshold-ratio, 1.5 by
 * default), else 1.  The factor feeds fb_ingest_params.subsample_factor.
 * *reason (may be null): 0 subsampled from the genome size, 1 subsampled
 * from the estimate, 2 kept: within the trigger, 3 kept: estimate not HIGH. */
double fb_auto_subsample_factor(const fb_genome_estimate* est, double depth, double threshold_ratio,
                                uint64_t genome_size, int32_t* reason);

/* ---- Serving: a request loop with continuous batching ----------------------
 * No counterpart in the reference, which assembles one isolate per process
 * run (main.rs:37-564); each request runs the reference flow of one isolate:
 * estimate_genome (estimator.rs:286-440) and then build_cdbg_with_contigs_ext
 * passes 1-3 (build.rs:67-328) at every k of the ladder.
 *
 * A server owns a worker thread and its own context on `device`.  Requests
 * (one isolate's R1/R2 read pairs) wait in arrival order and are admitted
 * while their device footprint fits the budget.  Every iteration runs one
 * batched pass for each ladder position in-flight requests are at, so
 * requests that arrive while others are on the GPU join the next iteration
 * and share its launches; a request that finishes leaves at once and frees its
 * memory.  Sketches and buffers of a request stay on the device across its
 * passes.  Each result equals the one the request gets alone (fb_build_solid
 * at the request's genome size).  submit / wait / download may be called from
 * any host thread. */
typedef struct fb_server fb_server;
typedef struct {
  uint32_t n_passes;            /* 1..4 passes of the k ladder */
  uint32_t k[4];
  int32_t target_solid_cov[4];  /* Option<u8> per pass: < 0 = None */
  uint64_t cms_width;           /* 0 = CountMinSketch::for_genome_size(G) */
  int32_t counter_bits;         /* 4 or 16 */
  int32_t hll_p;                /* per-pass HLL precision, 0 = none */
  uint32_t estimate;            /* 1 = estimate_genome first; its G sizes the sketch when no size is given */
  uint64_t est_max_pairs;       /* estimate_genome max_pairs */
  double target_cov;            /* estimate_genome target coverage */
  uint64_t hbm_budget;          /* admission budget in bytes; 0 = 80% of the free device memory */
  uint32_t max_inflight;        /* requests on the device at once; 0 = 64 */
  uint32_t max_queue;           /* waiting requests; 0 = 4096, FB_ERR_BUSY beyond */
  uint32_t keep_sketches;       /* 1 = keep every pass's CMS table and HLL registers with its result */
  uint64_t ctx_budget;          /* device bytes the server's context may hold; 0 = unlimited.  A
                                   request whose upload or pass needs more fails with FB_ERR_OOM.
                                   The context keeps at most a fifth of the device memory free at
                                   creation in cached free buffers, the longest-unused freed first */
} fb_serve_params;

typedef struct {
  int32_t status;               /* FB_OK or the request's error */
  char error[160];
  uint32_t n_passes;
  uint64_t genome_size;         /* sizes the request's sketch (build.rs:74-84) */
  uint32_t has_estimate;
  fb_genome_estimate estimate;
  uint64_t n_solid[4];          /* solid k-mers per pass */
  uint32_t route[4];            /* pass-2 route per pass (fb_solid_stats[4] codes) */
  uint64_t bases;
  uint64_t footprint;           /* device bytes reserved at admission */
  double t_submit, t_admit, t_done;  /* seconds since the server started */
} fb_serve_result;

typedef struct {
  uint64_t submitted, completed, failed, queued, inflight;
  uint64_t iterations;          /* worker iterations that ran passes */
  uint64_t pass_runs;           /* request-passes run */
  uint64_t max_batch;           /* most requests in flight in one iteration */
  uint64_t max_groups;          /* most ladder positions in one iteration */
  uint64_t budget, inflight_bytes, device_bytes;
  double busy_s;                /* time the worker spent running passes */
  double start_s;               /* worker time admitting requests: upload, estimate, sketches */
  double pass_s;                /* worker time in the batched passes (device work and its sizing) */
  double fetch_s;               /* worker time copying results to the host */
} fb_serve_stats;

void fb_serve_params_default(fb_serve_params* p);
int fb_server_create(int device, const fb_serve_params* p, fb_server** out);
/* Finishes the requests in flight, refuses the queued ones (FB_ERR_BUSY) and
 * wakes their waiters, waits until every fb_server_wait in progress has
 * returned, then frees everything.  No other call may start once destroy has. */
void fb_server_destroy(fb_server* s);
const char* fb_server_last_error(const fb_server* s);
/* Copies the read pairs (R1 -> pe_forward, R2 -> pe_reverse) and queues the
 * request.  genome_size 0 = from the estimate (or bases / 60 without one).
 * FB_ERR_BUSY when the queue is full, FB_ERR_OOM when the request alone
 * exceeds the budget. */
int fb_server_submit(fb_server* s, const fb_arena_view* r1, const fb_arena_view* r2, uint64_t genome_size,
                     uint64_t* ticket);
/* Waits up to timeout_s (< 0: no limit) for the request; FB_ERR_TIMEOUT (out
 * untouched) when it has not finished, else its status with out filled. */
int fb_server_wait(fb_server* s, uint64_t ticket, double timeout_s, fb_serve_result* out);
/* Solid set of a finished request's pass, sorted, in the fb_solid_download layout. */
int fb_server_download(fb_server* s, uint64_t ticket, uint32_t pass, uint64_t* keys, uint16_t* cov, uint8_t* mask);
/* The pass's CMS table (fb_cms_download layout) and HLL registers; needs keep_sketches. */
int fb_server_sketches(fb_server* s, uint64_t ticket, uint32_t pass, void* cms, uint64_t cms_bytes, uint8_t* hll);
/* Drops a finished (done or failed) request's results.  The worker touches
 * no request once it has published it as finished, so this may follow
 * fb_server_wait at once, whatever its status. */
int fb_server_release(fb_server* s, uint64_t ticket);
int fb_server_stats(fb_server* s, fb_serve_stats* out);

/* The serving admission policy alone (host only, no device): requests wait in
 * arrival order; the head is admitted while its bytes fit next to the
 * admitted ones and fewer than max_inflight are in; it is never overtaken.
 * submit refuses (FB_ERR_OOM) a request larger than the budget. */
typedef struct fb_admission fb_admission;
int fb_admission_create(uint64_t budget, uint32_t max_inflight, fb_admission** out);
void fb_admission_destroy(fb_admission* q);
int fb_admission_submit(fb_admission* q, uint64_t id, uint64_t bytes);
/* Admits what fits; writes up to max_ids admitted ids in order and returns how many were admitted. */
uint32_t fb_admission_admit(fb_admission* q, uint64_t* ids, uint32_t max_ids);
int fb_admission_retire(fb_admission* q, uint64_t id);
uint64_t fb_admission_used(const fb_admission* q);
uint32_t fb_admission_queued(const fb_admission* q);
uint32_t fb_admission_inflight(const fb_admission* q);

/* ---- Host helpers ---------------------------------------------------------- */
/* BitSeqArena::push_seq for many sequences: data must hold ceil(total/32)
 * zeroed words; offsets receives n entries. */
int fb_pack_ascii(const uint8_t* ascii, const uint32_t* lengths, uint64_t n, uint64_t* data, uint64_t* offsets);
/* splitmix64 (sketch.rs:13-18) */
uint64_t fb_splitmix64(uint64_t z);

#ifdef __cplusplus
}
#endif
#endif
