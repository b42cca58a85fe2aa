// fb_internal.hpp -- launchers shared between the kernels and the runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fb {

constexpr uint64_t kEmpty = ~0ull;     // empty key word of the solid table
constexpr int kArenaPadWords = 8;      // zero words after every device arena
constexpr uint32_t kBinSlices = 8;     // XCD slices of every bin region (one per XCD)

// One device read list: bases packed LSB-first, 32 per word (the reference's
// BitSeqArena layout, src/common/bitseq.rs:145-149).  Several host arenas are
// concatenated word-aligned; offs[] are global base indices.
struct DevReads {
  const uint64_t* data = nullptr;
  const uint64_t* offs = nullptr;
  const uint32_t* lens = nullptr;
  uint64_t n_reads = 0;
};

// Count-Min table: rows*width u32 exact counts (index = row*width + col).
struct DevCms {
  uint32_t* counts = nullptr;
  uint32_t log2w = 0;
};

struct DevHll {
  uint32_t* regs = nullptr;  // 2^p u32 registers
  uint32_t p = 0;
};

// Open-addressing solid table.  Slot = kw key words (k-mer packed 31 bases
// per 62-bit word, least significant word first; all-ones = empty), then u32
// cov, u32 pad, u32 edge[8].
struct DevSolidTable {
  uint64_t* slots = nullptr;
  uint64_t cap = 0;           // power of two
  uint32_t kw = 1;            // key words
  uint32_t stride = 0;        // u64 words per slot = kw + 5
  uint32_t max_probe = 0;
  uint32_t* status = nullptr; // [0] = overflow flag, [1] = claimed slots
};

struct Pass1Args {
  DevReads reads;
  uint32_t k;
  uint32_t split;        // threads per read
  uint32_t weight;       // 1 for reads, 10 for prior contigs (build.rs:124,133)
  DevCms cms;
  DevHll hll;            // regs == nullptr: no HLL
};

struct Pass2Args {
  DevReads reads;
  uint32_t k;
  uint32_t split;
  const uint32_t* solid_bits;  // rows*width bits: count >= t_solid
  uint32_t log2w;
  DevSolidTable table;
  uint32_t min_trans_cov;      // edge counters stop being incremented at this value
  uint32_t contig_mode;        // 1: contig backbone (+100 / +1000, no CMS gate)
};

// Pass 1 (build.rs:93-138) with the per-pass HLL over splitmix64(ntHash).
hipError_t launch_pass1(const Pass1Args& a, hipStream_t s);

// A launch over many isolates: args is a device array of per-isolate
// argument structs, blk a device array of n_iso+1 block prefixes; block b
// works for the isolate i with blk[i] <= b < blk[i+1], so no tile crosses
// isolates.
struct MultiLaunch {
  const void* args = nullptr;
  const uint32_t* blk = nullptr;
  uint32_t n_iso = 0;
  uint32_t total_blocks = 0;
};
constexpr uint32_t kPassBlock = 256;   // block size of k_pass1 / k_pass2
hipError_t launch_pass1_multi(const MultiLaunch& m, uint32_t k, bool cms, bool hll, hipStream_t s);

// Pass 1 with LDS-privatised Count-Min slices (fb_binned.hip).
struct BinArgs {
  DevReads reads;
  uint32_t k = 0;
  uint32_t split = 1;
  uint32_t chunks = 1;         // thread shares per scatter-block thread (blocks = threads / (block * chunks))
  uint32_t grid = 0;           // nonzero: a grid of this many blocks strides over all thread shares
  uint32_t log2w = 0;
  uint64_t n_cells = 0;        // 4 * width
  uint32_t* counts = nullptr;  // exact u32 cells
  DevHll hll;
  uint32_t sb = 15;            // log2 cells per bin (<= 15: 4 << sb bytes of LDS counters)
  uint32_t nbins = 0;
  uint64_t cap = 0;            // offsets per (bin, XCD slice) region, multiple of 4
  uint16_t* data = nullptr;    // nbins * kBinSlices * cap cell offsets
  uint32_t* cursor = nullptr;  // nbins * kBinSlices fill counters, zeroed before phase A
  uint32_t* bits = nullptr;    // optional threshold bits of count >= t_solid
  uint32_t t_solid = 0;
  uint32_t cap_value = 65535;  // counter saturation (15 or 65535)
};
size_t bin_scatter_lds(uint32_t nbins);
// Whether pass 1 with nbins bins uses the per-bin buffer kernel.
bool bin_buffered(uint32_t nbins);
// Pass-1 scatter blocks resident on the whole device at once (the grid of
// the striding form).
uint32_t bin_scatter_resident(uint32_t k, bool hll, uint32_t nbins);
// mid (optional) is recorded between the scatter and count kernels.
hipError_t launch_bin_pass1(const BinArgs& a, hipStream_t s, hipEvent_t mid = nullptr);
constexpr uint32_t kScatterBlockSize = 1024;  // block size of the pass-2 count and scatter kernels
constexpr uint32_t kP1BlockSize = 512;       // block size of the pass-1 bin scatter kernel
constexpr uint32_t kBinLog2Cells = 15;       // pass-1 bin: 2^15 Count-Min cells (128 KB of LDS counters)
constexpr uint32_t kP1Chunk = 64;            // pass-1 offsets per reserved chunk: one 128-byte line
constexpr uint32_t kP1BufBlock = 1024;       // block size of the pass-1 per-bin buffer kernel
constexpr uint32_t kP1BufMaxBins = 2048;     // most bins the per-bin buffer kernel holds in LDS
constexpr uint32_t kP1BufMinBins = 1025;     // fewest bins it is used for
constexpr uint32_t kP2GateBlockSize = 256;   // block size of the pass-2 gate kernel
// Thread shares per scatter-block thread: enough blocks to fill the chip a few
// times over, each flushing its bin histogram once for more work.
inline uint32_t scatter_chunks(uint64_t threads, uint32_t block = kScatterBlockSize) {
  const uint64_t c = threads / ((uint64_t)block * 2048u);
  return c < 1 ? 1u : (c > 16 ? 16u : (uint32_t)c);
}
// Pass 1: about 1024 workgroups (two rounds of two per CU), since each
// leaves a partly used line per bin behind.
inline uint32_t p1_chunks(uint64_t threads) {
  const uint64_t c = threads / ((uint64_t)kP1BlockSize * 1024u);
  return c < 1 ? 1u : (c > 64 ? 64u : (uint32_t)c);
}
inline uint64_t scatter_blocks(uint64_t threads, uint32_t chunks, uint32_t block = kScatterBlockSize) {
  const uint64_t per = (uint64_t)block * chunks;
  return (threads + per - 1) / per;
}
// buffered: the per-isolate BinArgs carry grid = the isolate's block count
// and every isolate has at most kP1BufMaxBins bins (the per-bin LDS buffer
// kernel); otherwise the staged scatter over chunked blocks.
hipError_t launch_bin_scatter_multi(const MultiLaunch& m, uint32_t k, bool hll, uint32_t max_nbins, bool buffered,
                                    hipStream_t s);
// m.blk counts bins: one block per bin.
hipError_t launch_bin_count_multi(const MultiLaunch& m, uint32_t max_sb, hipStream_t s);
// Bits of count >= t (saturating-counter semantics: cap 15 or 65535).
hipError_t launch_solid_bits(const uint32_t* counts, uint64_t n_counters, uint32_t t, uint32_t cap,
                             uint32_t* bits, hipStream_t s);
// Sum of set bits of n u32 words added to *out (zeroed by the caller).
hipError_t launch_popcount(const uint32_t* words, uint64_t n, unsigned long long* out, hipStream_t s);
// Histogram of min(count, 255) over n counters added to hist[256] (zeroed by the caller).
hipError_t launch_count_hist(const uint32_t* counts, uint64_t n, uint32_t* hist, hipStream_t s);
// Keys with at least f windows expected from the row-0 count histogram (see
// fb_runtime.hip): the pass-2 tables' load in the ungated form.
double est_kept_keys(const uint32_t* hist, uint64_t width, uint64_t windows, uint32_t t, uint32_t f);
// Reference-layout packing of the exact counts.
hipError_t launch_pack_nibble(const uint32_t* counts, uint64_t n_counters, uint8_t* out, hipStream_t s);
hipError_t launch_pack_u16(const uint32_t* counts, uint64_t n_counters, uint16_t* out, hipStream_t s);
hipError_t launch_pack_u8(const uint32_t* regs, uint64_t n, uint8_t* out, hipStream_t s);
// Table initialisation (all key words empty, counters zero).
hipError_t launch_table_clear(const DevSolidTable& t, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
hipError_t launch_copy_calibration(const void* src, void* dst, uint64_t n16, hipStream_t s);
hipError_t launch_atomic_rate(uint32_t* table, uint32_t log2n, uint64_t n_updates, int mode, uint32_t* sink,
                              hipStream_t s);
// Pass 2 (build.rs:140-289).
hipError_t launch_pass2(const Pass2Args& a, hipStream_t s);
hipError_t launch_pass2_multi(const MultiLaunch& m, uint32_t k, hipStream_t s);
// Pass 3 (build.rs:291-328): compaction into reference key words.
// out_keys: n_ref_words words per record, most significant first.
// scratch (pass3_scratch_bytes()): u32 [0] records kept, u64 at byte 8 the sum
// of raw cov (records added to the table), then per-block partials.
size_t pass3_scratch_bytes();
// out_base (nullable): device u32 added to every output position; positions
// at or past out_cap are not written (the count in scratch still has them).
hipError_t launch_pass3(const DevSolidTable& t, uint32_t k, uint32_t n_ref_words, uint32_t min_trans_cov,
                        uint64_t* out_keys, uint16_t* out_cov, uint8_t* out_mask, void* scratch, hipStream_t s,
                        const uint32_t* out_base = nullptr, uint32_t out_cap = 0xFFFFFFFFu);
// Sort the compacted records by k-mer (integer order).  Uses scratch from
// the caller: tmp_bytes queried with keys == nullptr.  top_first: for keys
// over 64 bits, first one sort by their first 16 bases, a check
// (synchronising the stream) for runs of keys that share them, an in-place
// sort of runs of at most 64, else the limb-by-limb sort.
hipError_t sort_solid(uint64_t* keys, uint16_t* cov, uint8_t* mask, uint32_t n, uint32_t n_ref_words,
                      uint32_t k, void* tmp, size_t* tmp_bytes, hipStream_t s, bool top_first = false);

// Pass 2 + 3 with LDS-privatised solid tables (fb_pass2_binned.hip).
constexpr uint32_t kP2LdsBytes = 160u * 1024u;  // workgroup LDS of the bin kernel
struct P2BinArgs {
  DevReads reads;
  uint32_t k = 0;
  uint32_t split = 1;            // pairs per thread must stay <= 63
  uint32_t chunks = 1;           // thread shares per count/scatter-block thread
  const uint32_t* solid_bits = nullptr;
  uint32_t log2w = 0;
  uint32_t nbins = 1;             // bins per round
  uint32_t log2_nbins = 0;
  // Keys are split over 2^log2_rounds rounds by the top bits of their bin
  // hash; each round counts, scatters and tables only its own records, so the
  // distinct keys of one round fit the LDS tables of nbins bins.
  uint32_t log2_rounds = 0;
  uint32_t round = 0;            // round of this launch sequence
  // Nonzero: the scatter walks without a count (ungated, one round, the
  // buffered scatter's bins): the records of bin b from XCD slice sl go to
  // region b * kBinSlices + sl of region_cap records (a multiple of
  // kP2ChunkUnits flush units), runs[b * kBinSlices + sl] counts the records
  // reserved there in whole chunks (tails all-zero records), and status[0]
  // is raised when a region fills up.
  uint64_t region_cap = 0;
  uint64_t* gate_masks = nullptr;  // n_reads * split, gate bit of window p0+i at bit i
  uint32_t rw = 1;               // 64-bit words per record: ceil((2k + 8 + cflag) / 64)
  // Prior contigs (the multi-k upgrade): their windows follow the reads' in
  // the share space, csplit shares of at most 63 pairs per contig, ungated;
  // their records carry a ninth info bit (cflag = 1) and add the contig
  // weights (cov +100 a side, edge +1000) in the bin kernel.
  DevReads contigs;
  uint32_t csplit = 1;
  uint32_t cflag = 0;
  // nbins * kBinSlices + 1 run counters, zeroed before every round: record
  // counts per (bin, slice), then (after the scan) region starts; the last
  // entry holds the round's record total.
  uint32_t* runs = nullptr;
  // n_blocks * nbins: each scatter block's record count per bin, then (after
  // the column scan) the absolute start of its run.
  uint32_t* blk_hist = nullptr;
  uint32_t n_blocks = 0;         // scatter blocks: scatter_blocks(n_reads * split, chunks)
  uint64_t* recs = nullptr;      // records of rw words, one per window at most
  uint64_t rec_cap = 0;          // records the buffer holds
  uint32_t* status = nullptr;    // zeroed: [0] record buffer overflow (never for exact runs), [1] records of all rounds
  uint32_t lds_cap = 0;          // LDS table entries
  uint32_t min_trans_cov = 0;
  // Ungated form: records of every window pair, the Count-Min gate applied
  // per kept key in the bin kernel (needs the counters' cap >= t_solid, and
  // t_solid <= 10 with prior contigs); filt levels of the LDS record filter
  // (0: none, else min(min_trans_cov, kP2MaxFilt)).
  uint32_t ungated = 0;
  uint32_t t_solid = 0;
  uint32_t filt = 0;
  DevSolidTable table;           // keys that find the LDS table full
  uint32_t nref = 1;
  uint64_t* out_keys = nullptr;
  uint16_t* out_cov = nullptr;
  uint8_t* out_mask = nullptr;
  uint32_t* out_count = nullptr;  // zeroed; records emitted from LDS (counted also past out_cap)
  uint32_t out_cap = 0xFFFFFFFFu; // output entries; positions past it are not written
  unsigned long long* stats = nullptr;  // zeroed; [0] records (sum of raw cov), [1] keys held in LDS
};
constexpr uint32_t kP2FiltWords = 2048;  // u32 words per level of the ungated bin kernel's record filter
constexpr uint32_t kP2MaxFilt = 4;       // filter levels: records counted up to 4 per position
// Target fill of the ungated form's LDS tables at the kept-key estimate
// (which carries a 1.1 margin; the filter's false positives add a few %).
constexpr double kP2UngatedLoad = 0.65;
constexpr uint32_t kP2BufWords = 8;      // words of a bin's record buffer in the buffered pass-2 scatter
constexpr uint32_t kP2BufMaxBins = 2048; // most bins a round of the buffered scatter holds in LDS
constexpr uint32_t kP2ChunkUnits = 8;    // flush units (kP2BufWords words each) a chunked-scatter reservation takes
// Workgroups of the chunked pass-2 scatter: one per CU (its LDS takes one).
uint32_t p2_chunk_grid();
// Whether phase A of pass 2 (k, nbins bins a round) writes records through
// per-bin LDS buffers; else one store per record.
bool p2_scatter_buffered(uint32_t k, uint32_t nbins);
// LDS bytes of one bin-table entry: kw key words, u32 cov, 8 u8 edge counters.
constexpr uint32_t p2_entry_bytes(uint32_t kw) { return kw * 8u + 4u + 8u; }
uint32_t p2_lds_cap(uint32_t kw, uint32_t filt = 0);
uint32_t p2_max_bins();  // largest pass-2 bin count the scatter kernel's LDS allows
size_t p2_lds_bytes(uint32_t kw, uint32_t cap, uint32_t filt = 0);
// gated (optional) is recorded after the gate, count, scan and offset kernels,
// scattered after the scatter kernel.
hipError_t launch_p2_binned(const P2BinArgs& a, hipStream_t s, hipEvent_t gated = nullptr,
                            hipEvent_t scattered = nullptr);
// gate: one block per kP2GateBlock threads of each isolate; scatter (the
// count and scatter kernels): one block per kScatterBlockSize * chunks
// threads; bins: one block per bin of a round.  runs_fill clears every
// isolate's run counters before each of the max_rounds rounds (an isolate
// with fewer rounds skips the later ones).
struct FillRange;
// lds_cap, filt: the isolates' LDS table entries and record-filter levels
// (one value for all of them); the spill tables of ungated isolates get
// their gate correction after the last round.
hipError_t launch_p2_binned_multi(const MultiLaunch& gate, const MultiLaunch& scatter, const MultiLaunch& bins,
                                  const FillRange* runs_fill, uint32_t n_fill, uint32_t k, uint32_t max_nbins,
                                  uint32_t max_rounds, uint32_t lds_cap, uint32_t filt, hipStream_t s);
constexpr uint32_t kP2MaxLog2Rounds = 6;  // at most 64 pass-2 rounds

// Multi-table forms for the batched pass (blockIdx.y = isolate).
struct P3Args {
  DevSolidTable t;
  uint32_t nref = 1;
  uint32_t gate = 0;
  uint64_t* out_keys = nullptr;
  uint16_t* out_cov = nullptr;
  uint8_t* out_mask = nullptr;
  void* scratch = nullptr;  // pass3_scratch_bytes()
  const uint32_t* out_base = nullptr;
  uint32_t out_cap = 0xFFFFFFFFu;
};
hipError_t launch_pass3_multi(const P3Args* args, uint32_t n, uint32_t kw, hipStream_t s);
struct FillRange {
  uint32_t* p;
  uint64_t n;
  uint32_t v;
};
hipError_t launch_fill_ranges(const FillRange* r, uint32_t n, hipStream_t s);
hipError_t launch_table_clear_multi(const DevSolidTable* tabs, uint32_t n, hipStream_t s);

// CountMinSketch::add / estimate on explicit 64-bit keys.
hipError_t launch_cms_add_keys(DevCms c, const uint64_t* keys, uint64_t n, uint32_t weight, hipStream_t s);
hipError_t launch_cms_query_keys(DevCms c, const uint64_t* keys, uint64_t n, uint32_t cap, uint32_t* out,
                                 hipStream_t s);
hipError_t launch_hll_add_hashes(DevHll h, const uint64_t* hashes, uint64_t n, hipStream_t s);
hipError_t launch_hll_add_canonical(DevHll h, const uint64_t* hashes, uint64_t n, hipStream_t s);
hipError_t launch_hll_merge(uint32_t* dst, const uint32_t* src, uint64_t n, hipStream_t s);
hipError_t launch_cms_row_cells(DevCms c, const uint64_t* keys, uint64_t n, uint32_t row, uint32_t* cells,
                                uint32_t* idx, hipStream_t s);
hipError_t launch_cms_row_estimate(const uint32_t* counts, const uint32_t* cells, const uint32_t* idx, uint64_t n,
                                   uint32_t row, uint32_t cap, uint32_t* out, hipStream_t s);
// Stable radix sort of (u32 key, u32 value) pairs on key bits [0, end_bit);
// tmp == nullptr sizes *tmp_bytes.
hipError_t sort_u32_pairs(const uint32_t* kin, uint32_t* kout, const uint32_t* vin, uint32_t* vout, uint64_t n,
                          uint32_t end_bit, void* tmp, size_t* tmp_bytes, hipStream_t s);
hipError_t launch_nthash_windows(const uint8_t* seq, uint64_t n_windows, uint32_t k, uint64_t* out, hipStream_t s);

// Estimator (estimator.rs:296-379), first-insertion-time form.  Reads are
// R1[0], R2[0], R1[1], R2[1], ... in the order the reference visits them.
struct EstArgs {
  DevReads reads;               // interleaved sample reads
  const uint32_t* occ_base;     // per read: global index of its first window
  uint32_t n_reads;
  uint32_t* first_ins;          // 4 * 2^23 cells, initialised to ~0
  uint32_t* first_total;        // 2^14 * 64 (register, rho) -> first pair index
  uint32_t* first_solid;
};
hipError_t launch_est_first_insert(const EstArgs& a, hipStream_t s);
hipError_t launch_est_solid(const EstArgs& a, hipStream_t s);

// Unitig compaction of a sorted solid set (fb_compact.hip, build.rs:330-597).
// Directed k-mer d = 2i + o (o = 1: reverse strand of k-mer i); start edges
// are the out-edges of boundary strands, numbered in (d, base) order.
struct CgArgs {
  const uint64_t* keys = nullptr;  // n * nref words, most significant first, sorted
  const uint16_t* cov = nullptr;
  const uint8_t* mask = nullptr;
  uint32_t n = 0, k = 0, nref = 1;
  uint8_t *kind = nullptr, *fbase = nullptr, *lbase = nullptr, *obase = nullptr;  // 2n each
  uint32_t *nstart = nullptr, *sbase = nullptr, *succ = nullptr;                // 2n each
  uint32_t* bnext = nullptr;                                                    // 8n
  // jump buffers 0/1 (paths) and 2/3 (cycles), 2n each: per directed node
  // x = up (next node towards the head, or a state), y = head (start edge),
  // z = rank, w = lab (smallest node seen), one 16-byte gather per jump
  uint4* js[4] = {};
  // 2 x 64 progress flags, zeroed before each jump (paths, then cycles):
  // flag r is set when round r turned a node terminal or lowered a label
  uint32_t* jflag = nullptr;
  uint64_t n_se = 0;       // start edges
  unsigned long long* csum = nullptr;  // n_se: cov sum of the path's counted nodes
  unsigned long long* last = nullptr;  // n_se: rank << 32 | last counted node
  uint32_t *walked = nullptr, *slen = nullptr, *uid = nullptr;  // n_se
  uint64_t* soff = nullptr;                                      // n_se
  unsigned long long* ccov = nullptr;  // 2n: cov sum per cycle label
  uint32_t *clen = nullptr, *cyc_len = nullptr, *cid = nullptr;  // 2n
  uint64_t* coff = nullptr;                                      // 2n, relative to circ_base
  uint64_t circ_base = 0;                                        // bases of the linear unitigs
  uint8_t* seq = nullptr;
  uint32_t n_lin = 0;
  uint64_t* u_off = nullptr;
  uint32_t* u_len = nullptr;
  float* u_cov = nullptr;
  uint8_t* u_circ = nullptr;
};
hipError_t launch_cg_links(const CgArgs& a, hipStream_t s);
hipError_t launch_cg_pred(const CgArgs& a, hipStream_t s);
uint32_t cg_jump_rounds(uint64_t nodes);
hipError_t launch_cg_jump(const CgArgs& a, uint32_t s, uint32_t t, uint32_t* out_s, hipStream_t st);
hipError_t launch_cg_accumulate(const CgArgs& a, uint32_t s, hipStream_t st);
hipError_t launch_cg_cycle_ranks(const CgArgs& a, uint32_t s, uint32_t* out_s, hipStream_t st);
hipError_t launch_cg_cycle_flags(const CgArgs& a, uint32_t* flag, hipStream_t s);
hipError_t launch_cg_table(const CgArgs& a, hipStream_t s);
hipError_t launch_cg_write(const CgArgs& a, uint32_t sp, uint32_t sc, hipStream_t s);
// Overlap edges of the compacted unitigs (fb_compact.hip): per pass (ov = k,
// then k - 1) every strand's ov-prefix hash (k_cg_prefix), sorted with the
// strand ids (skey/sval), then per end e = 2u + side its edge count and,
// after a scan into base, its edges.
struct CgEdgeArgs {
  const uint8_t* seq = nullptr;
  const uint64_t* off = nullptr;
  const uint32_t* len = nullptr;
  const uint8_t* circ = nullptr;
  uint32_t U = 0, k = 0, ov = 0;
  uint32_t merge = 0;  // this pass is ov = k - 1: circular self edges absorb their match
  uint64_t* hkey = nullptr;
  uint32_t* hval = nullptr;
  const uint64_t* skey = nullptr;
  const uint32_t* sval = nullptr;
  uint32_t* cnt = nullptr;   // per end: edges (bit 31: the self edge absorbed a match)
  uint32_t* pos = nullptr;   // per end: edges written so far
  const uint32_t* base = nullptr;
  uint32_t* e_src = nullptr;
  uint32_t* e_dst = nullptr;
  uint8_t* e_ov = nullptr;
  uint16_t* e_sup = nullptr;
};
hipError_t launch_cg_prefix(const CgEdgeArgs& a, hipStream_t s);
hipError_t launch_cg_end_count(const CgEdgeArgs& a, bool first, hipStream_t s);
hipError_t launch_cg_end_total(const uint32_t* cnt, uint32_t* tot, uint32_t n, hipStream_t s);
hipError_t launch_cg_end_write(const CgEdgeArgs& a, bool first, hipStream_t s);
hipError_t cg_sort_pairs_u64(const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout, uint64_t n,
                             void* tmp, size_t* tmp_bytes, hipStream_t s);
hipError_t launch_cg_flip(uint8_t* seq, const uint64_t* off, const uint32_t* len, uint32_t n_lin, uint8_t* flip,
                          uint64_t lin_bases, hipStream_t s);
// filter_active_reads (local_multik.rs:259-333), fb_filter.hip.
struct FaArgs {
  // boundary set: 2^log2cap u64 slots (kEmpty = empty) + set_flag[0] for the key kEmpty
  uint64_t* set = nullptr;
  uint32_t* set_flag = nullptr;
  uint32_t log2cap = 0;
  uint32_t k = 0;  // k' = min(graph k, 31)
  // unitig windows: range r holds windows rpre[r] .. rpre[r+1]-1 starting at useq[rstart[r]]
  const uint8_t* useq = nullptr;
  const uint64_t* rstart = nullptr;
  const uint64_t* rpre = nullptr;
  uint32_t n_ranges = 0;
  uint64_t n_windows = 0;
  // source reads: merged reads, pe_forward, pe_reverse start at first[0..2]
  DevReads reads;
  uint64_t first[3] = {0, 0, 0};
  uint64_t n_merged = 0, n_pairs = 0;
  uint32_t* keep = nullptr;  // n_merged + n_pairs + 1
  uint32_t* pos = nullptr;   // exclusive scan of keep
  uint32_t* src = nullptr;   // kept reads' source index, output order
  uint32_t* olen = nullptr;  // their lengths (+ one trailing 0)
  uint64_t* opos = nullptr;  // exclusive scan of olen
  // output list: arena a holds reads rfirst[a] .. rfirst[a+1]-1 and words word0[a] .. word0[a+1]-1
  uint64_t rfirst[4] = {0, 0, 0, 0};
  uint64_t word0[4] = {0, 0, 0, 0};
  uint64_t* odata = nullptr;
  uint64_t* ooffs = nullptr;
  uint32_t* olens = nullptr;
  uint32_t* hist = nullptr;  // read count per length (hist_n bins, last open)
  uint32_t hist_n = 1;
};
hipError_t launch_fa_set_clear(const FaArgs& a, hipStream_t s);
hipError_t launch_fa_set_insert(const FaArgs& a, hipStream_t s);
hipError_t launch_fa_keep(const FaArgs& a, hipStream_t s);
hipError_t launch_fa_list(const FaArgs& a, hipStream_t s);
hipError_t launch_fa_offsets(const FaArgs& a, hipStream_t s);
hipError_t launch_fa_gather(const FaArgs& a, hipStream_t s);

// Ingest record loop (trim_merge.rs:17-417), fb_ingest.hip.
struct DevRecords {
  const uint8_t* seq = nullptr;
  const uint64_t* seq_off = nullptr;
  const uint32_t* seq_len = nullptr;
  const uint8_t* qual = nullptr;
  const uint64_t* qual_off = nullptr;
  const uint32_t* qual_len = nullptr;
};
struct IngArgs {
  DevRecords r1, r2;
  uint64_t n = 0;  // pairs
  uint32_t subsample = 0;
  uint64_t keep_threshold = 0, seed = 0;
  uint32_t enable_trim = 1, enable_merge = 1, min_len = 0, min_qual = 0;
  uint32_t* status = nullptr;            // [0] = 1: a quality string shorter than its clipped read
  unsigned long long* stats = nullptr;   // retained, bases after trim, merged, merged bases, unmerged
  uint32_t *merged = nullptr, *unmerged = nullptr;  // n + 1 flags
  uint32_t *pos_m = nullptr, *pos_u = nullptr;      // their exclusive scans
  uint32_t *t1 = nullptr, *t2 = nullptr, *ov = nullptr;
  uint32_t* src = nullptr;   // pair of every output read
  uint32_t* olen = nullptr;  // its length (+ one trailing 0)
  uint64_t* opos = nullptr;  // exclusive scan of olen
  uint64_t rfirst[4] = {0, 0, 0, 0};
  uint64_t word0[4] = {0, 0, 0, 0};
  uint64_t* odata = nullptr;
  uint64_t* ooffs = nullptr;
  uint32_t* olens = nullptr;
  uint32_t* hist = nullptr;
  uint32_t hist_n = 1;
};
hipError_t launch_ing_pairs(const IngArgs& a, hipStream_t s);
hipError_t launch_ing_list(const IngArgs& a, hipStream_t s);
hipError_t launch_ing_offsets(const IngArgs& a, hipStream_t s);
hipError_t launch_ing_gather(const IngArgs& a, hipStream_t s);

// PE mapper (mapper.rs:89-279), fb_mapper.hip.
struct PairLinkRec {  // the layout of fb_pair_link
  uint32_t u1, u2, pair_id, dist_u1_exit, dist_u2_entry;
  uint16_t r1_len, r2_len;
};
struct MapArgs {
  uint32_t k = 0;
  // unitigs as a packed read list (one read per unitig) and their lengths
  DevReads useq;
  const uint32_t* ulen = nullptr;
  // indexed windows: range r holds windows rpre[r] .. rpre[r+1]-1 of unitig rng_u[r] from rng_p0[r]
  const uint32_t* rng_u = nullptr;
  const uint32_t* rng_p0 = nullptr;
  const uint64_t* rpre = nullptr;
  uint32_t n_ranges = 0;
  uint64_t n_windows = 0;
  // index table: 2^log2cap slots; the key kEmpty is counted in special_*
  uint64_t* keys = nullptr;
  uint32_t* cnt = nullptr;
  uint64_t* val = nullptr;
  uint32_t log2cap = 0;
  uint32_t* special_cnt = nullptr;
  uint64_t* special_val = nullptr;
  // pairs: pe_forward read first_f + i and pe_reverse read first_r + i
  DevReads reads;
  uint64_t first_f = 0, first_r = 0, n_pairs = 0;
  uint32_t *f_size = nullptr, *f_link = nullptr;      // n_pairs + 1 flags
  uint32_t *pos_size = nullptr, *pos_link = nullptr;  // their scans
  uint32_t *span = nullptr, *lu1 = nullptr, *lu2 = nullptr, *ld1 = nullptr, *ld2 = nullptr;
  uint32_t* sizes = nullptr;
  PairLinkRec* links = nullptr;
};
hipError_t launch_map_clear(const MapArgs& a, hipStream_t s);
hipError_t launch_map_insert(const MapArgs& a, hipStream_t s);
hipError_t launch_map_pairs(const MapArgs& a, hipStream_t s);
hipError_t launch_map_compact(const MapArgs& a, hipStream_t s);

// Consensus polishing (consensus.rs:64-271), fb_polish.hip.
struct PolArgs {
  // node sequences (ASCII, concatenated; soff per node) and lengths
  const uint8_t* seq = nullptr;
  const uint64_t* soff = nullptr;
  const uint32_t* ulen = nullptr;
  // windows: range r covers windows rpre[r] .. rpre[r+1]-1 of unitig rng_u[r], every rng_step bases from 0
  const uint32_t* rng_u = nullptr;
  const uint64_t* rpre = nullptr;
  uint32_t n_ranges = 0, rng_step = 1;
  uint64_t n_windows = 0;
  // seed table; the hash kEmpty lives in special (sampled flag, count) and special_val
  uint64_t* keys = nullptr;
  uint32_t* cnt = nullptr;
  uint64_t* val = nullptr;
  uint32_t log2cap = 0;
  uint32_t* special = nullptr;
  uint64_t* special_val = nullptr;
  // reads: n_first of them from first0, the rest from first1
  DevReads reads;
  uint64_t n_reads = 0, n_first = 0, first0 = 0, first1 = 0;
  // pileups of the eligible nodes: 4 u32 per base from poff[node]
  uint32_t* pile = nullptr;
  const uint64_t* poff = nullptr;
  const uint32_t* elig = nullptr;  // eligible node ids, ascending
  uint32_t n_elig = 0;
  uint64_t n_pile = 0;
  uint8_t* out = nullptr;            // polished sequences, laid out as seq
  unsigned long long* stats = nullptr;  // fwd reads, rev reads, fwd bases, rev bases, rejected, modified
};
hipError_t launch_pol_clear(const PolArgs& a, hipStream_t s);
hipError_t launch_pol_sample(const PolArgs& a, hipStream_t s);
hipError_t launch_pol_count(const PolArgs& a, hipStream_t s);
hipError_t launch_pol_reads(const PolArgs& a, hipStream_t s);
hipError_t launch_pol_consensus(const PolArgs& a, hipStream_t s);

hipError_t cg_exclusive_sum_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* tmp, size_t* tmp_bytes,
                                hipStream_t s);
hipError_t cg_exclusive_sum_u64(const uint32_t* in, uint64_t* out, uint64_t n, void* tmp, size_t* tmp_bytes,
                                hipStream_t s);

}  // namespace fb
