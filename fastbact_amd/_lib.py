"""ctypes binding of libfastbact_gpu.so (include/fastbact_gpu.h, include/fastbact_synth.h).

The library is built in-tree by ``python -m fastbact_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfastbact_gpu.so"

FB_OK = 0
FB_ERR_INVALID = -1
FB_ERR_OOM = -2
FB_ERR_OVERFLOW = -3
FB_ERR_HIP = -4
FB_ERR_EMPTY = -5
FB_ERR_NODEV = -6
FB_ERR_BUSY = -7
FB_ERR_TIMEOUT = -8

ERR_NAMES = {
    FB_ERR_INVALID: "FB_ERR_INVALID",
    FB_ERR_OOM: "FB_ERR_OOM",
    FB_ERR_OVERFLOW: "FB_ERR_OVERFLOW",
    FB_ERR_HIP: "FB_ERR_HIP",
    FB_ERR_EMPTY: "FB_ERR_EMPTY",
    FB_ERR_NODEV: "FB_ERR_NODEV",
    FB_ERR_BUSY: "FB_ERR_BUSY",
    FB_ERR_TIMEOUT: "FB_ERR_TIMEOUT",
}

u8p = C.POINTER(C.c_uint8)
u16p = C.POINTER(C.c_uint16)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
vp = C.c_void_p


class ArenaView(C.Structure):
    _fields_ = [
        ("data", u64p),
        ("n_words", C.c_uint64),
        ("offsets", u64p),
        ("lengths", u32p),
        ("n_reads", C.c_uint64),
    ]


class GraphView(C.Structure):
    _fields_ = [("k", C.c_uint32), ("n_nodes", C.c_uint64), ("seq", C.POINTER(C.c_uint8)),
                ("offsets", C.POINTER(C.c_uint64)), ("lengths", C.POINTER(C.c_uint32)),
                ("cov", C.POINTER(C.c_float)), ("flags", C.POINTER(C.c_uint8)),
                ("n_edges_fwd", C.POINTER(C.c_uint32)), ("n_edges_rev", C.POINTER(C.c_uint32))]


class MultiKResult(C.Structure):
    _fields_ = [("executed", C.c_int32), ("early_stop", C.c_int32), ("genome_size_source", C.c_int32),
                ("reserved", C.c_int32), ("genome_size_used", C.c_uint64), ("cms_width", C.c_uint64),
                ("cms_allocated_bytes", C.c_uint64), ("active_unitigs_before", C.c_uint64),
                ("active_edges_before", C.c_uint64), ("active_unitigs_after", C.c_uint64),
                ("active_edges_after", C.c_uint64), ("kept_merged", C.c_uint64), ("kept_pairs", C.c_uint64),
                ("prior_contigs", C.c_uint64), ("transferred_circular", C.c_uint64)]


class RecordsView(C.Structure):
    _fields_ = [("seq", C.POINTER(C.c_uint8)), ("seq_offsets", C.POINTER(C.c_uint64)),
                ("seq_lengths", C.POINTER(C.c_uint32)), ("qual", C.POINTER(C.c_uint8)),
                ("qual_offsets", C.POINTER(C.c_uint64)), ("qual_lengths", C.POINTER(C.c_uint32)), ("n", C.c_uint64)]


class IngestParams(C.Structure):
    _fields_ = [("subsample_factor", C.c_double), ("min_len", C.c_uint32), ("min_qual", C.c_uint32),
                ("enable_trim", C.c_int32), ("enable_merge", C.c_int32), ("seed", C.c_uint64)]


class IngestStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("input_read_pairs", "retained_read_pairs", "bases_before_trim",
                                           "bases_after_trim", "merged_pairs_count", "merged_bases",
                                           "unmerged_pairs_count")] + [("parse_ms", C.c_float), ("device_ms", C.c_float)]


class PairLink(C.Structure):
    _fields_ = [("u1", C.c_uint32), ("u2", C.c_uint32), ("pair_id", C.c_uint32), ("dist_u1_exit", C.c_uint32),
                ("dist_u2_entry", C.c_uint32), ("r1_len", C.c_uint16), ("r2_len", C.c_uint16)]


class InsertSizeDist(C.Structure):
    _fields_ = [("median", C.c_double), ("mad", C.c_double), ("min_insert", C.c_uint64), ("max_insert", C.c_uint64),
                ("n_sizes", C.c_uint64)]


class PolishStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("bases_modified", "fwd_mapped_reads", "rev_mapped_reads",
                                           "fwd_pileup_bases", "rev_pileup_bases", "rejected_anchors")]


class BuildParams(C.Structure):
    _fields_ = [
        ("k", C.c_uint32),
        ("target_solid_cov", C.c_int32),
        ("genome_size", C.c_uint64),
        ("cms_width", C.c_uint64),
        ("counter_bits", C.c_int32),
        ("hll_p", C.c_int32),
        ("sort", C.c_uint32),
        ("split", C.c_uint32),
        ("placement", C.c_uint32),
        ("bins2", C.c_uint32),
        ("chunks1", C.c_uint32),
        ("chunks2", C.c_uint32),
    ]


class CountStats(C.Structure):
    _fields_ = [
        ("exact_read_pairs", C.c_uint64),
        ("total_bases_r1", C.c_uint64),
        ("total_bases_r2", C.c_uint64),
        ("total_bases", C.c_uint64),
        ("min_read_len", C.c_uint64),
        ("max_read_len", C.c_uint64),
        ("mean_read_len", C.c_double),
        ("median_read_len", C.c_uint64),
    ]


class GenomeEstimateC(C.Structure):
    _fields_ = [
        ("genome_size", C.c_uint64),
        ("exact_read_pairs", C.c_uint64),
        ("total_bases", C.c_uint64),
        ("median_read_len", C.c_uint64),
        ("max_read_len", C.c_uint64),
        ("estimated_cov", C.c_double),
        ("subsample_factor", C.c_double),
        ("kmers", C.c_uint32 * 8),
        ("n_kmers", C.c_uint32),
        ("confidence", C.c_int32),
        ("u_ge_1", C.c_uint64),
        ("u_ge_2", C.c_uint64),
        ("sample_kmer_depth", C.c_double),
        ("pairs_sampled", C.c_uint64),
        ("m_sample", C.c_uint64),
    ]


class SynthParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("genome_len", C.c_uint64),
        ("gc", C.c_double),
        ("operon_copies", C.c_uint32),
        ("operon_len", C.c_uint32),
        ("tandem_arrays", C.c_uint32),
        ("tandem_unit_min", C.c_uint32),
        ("tandem_unit_max", C.c_uint32),
        ("tandem_copies_min", C.c_uint32),
        ("tandem_copies_max", C.c_uint32),
        ("coverage", C.c_double),
        ("read_len", C.c_uint32),
        ("insert_mean", C.c_double),
        ("insert_sd", C.c_double),
        ("sub_rate", C.c_double),
        ("indel_rate", C.c_double),
    ]


class TuneResult(C.Structure):
    _fields_ = [
        ("placement", C.c_uint32),
        ("split", C.c_uint32),
        ("chunks1", C.c_uint32),
        ("chunks2", C.c_uint32),
        ("best_ms", C.c_float),
        ("default_ms", C.c_float),
        ("trials", C.c_uint32),
    ]


class ServeParams(C.Structure):
    _fields_ = [
        ("n_passes", C.c_uint32),
        ("k", C.c_uint32 * 4),
        ("target_solid_cov", C.c_int32 * 4),
        ("cms_width", C.c_uint64),
        ("counter_bits", C.c_int32),
        ("hll_p", C.c_int32),
        ("estimate", C.c_uint32),
        ("est_max_pairs", C.c_uint64),
        ("target_cov", C.c_double),
        ("hbm_budget", C.c_uint64),
        ("max_inflight", C.c_uint32),
        ("max_queue", C.c_uint32),
        ("keep_sketches", C.c_uint32),
        ("ctx_budget", C.c_uint64),
    ]


class Pass2Plan(C.Structure):
    _fields_ = [
        ("nb2", C.c_uint64),
        ("nbr", C.c_uint64),
        ("spill_cap", C.c_uint64),
        ("out_cap", C.c_uint64),
        ("rec_cap", C.c_uint64),
        ("n_blocks2", C.c_uint64),
        ("threads2", C.c_uint64),
        ("log2_rounds", C.c_uint32),
        ("split2", C.c_uint32),
        ("chunks2", C.c_uint32),
        ("lds_cap", C.c_uint32),
        ("ok", C.c_uint32),
    ]


class ServeResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("error", C.c_char * 160),
        ("n_passes", C.c_uint32),
        ("genome_size", C.c_uint64),
        ("has_estimate", C.c_uint32),
        ("estimate", GenomeEstimateC),
        ("n_solid", C.c_uint64 * 4),
        ("route", C.c_uint32 * 4),
        ("bases", C.c_uint64),
        ("footprint", C.c_uint64),
        ("t_submit", C.c_double),
        ("t_admit", C.c_double),
        ("t_done", C.c_double),
    ]


class ServeStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("failed", C.c_uint64),
        ("queued", C.c_uint64),
        ("inflight", C.c_uint64),
        ("iterations", C.c_uint64),
        ("pass_runs", C.c_uint64),
        ("max_batch", C.c_uint64),
        ("max_groups", C.c_uint64),
        ("budget", C.c_uint64),
        ("inflight_bytes", C.c_uint64),
        ("device_bytes", C.c_uint64),
        ("busy_s", C.c_double),
        ("start_s", C.c_double),
        ("pass_s", C.c_double),
        ("fetch_s", C.c_double),
    ]


# (name, restype, argtypes): every symbol include/fastbact_gpu.h and
# include/fastbact_synth.h declare.
SIGNATURES = [
    ("fb_ctx_create", C.c_int, [C.c_int, C.c_uint64, C.POINTER(vp)]),
    ("fb_ctx_destroy", None, [vp]),
    ("fb_last_error", C.c_char_p, [vp]),
    ("fb_ctx_synchronize", C.c_int, [vp]),
    ("fb_ctx_device_bytes", C.c_uint64, [vp]),
    ("fb_ctx_trim", C.c_int, [vp]),
    ("fb_version", C.c_char_p, []),
    ("fb_device_memory", C.c_int, [C.c_int, u64p, u64p]),
    ("fb_bank_upload", C.c_int, [vp, C.POINTER(ArenaView), C.POINTER(vp)]),
    ("fb_bank_destroy", None, [vp]),
    ("fb_bank_total_bases", C.c_uint64, [vp]),
    ("fb_bank_total_reads", C.c_uint64, [vp]),
    ("fb_bank_bases_for_k", C.c_uint64, [vp, C.c_uint32]),
    ("fb_bank_arena_info", C.c_int, [vp, C.c_int, u64p, u64p, u64p]),
    ("fb_bank_download_arena", C.c_int, [vp, C.c_int, u64p, u64p, u32p]),
    ("fb_cms_create", C.c_int, [vp, C.c_uint64, C.c_int, C.POINTER(vp)]),
    ("fb_cms_destroy", None, [vp]),
    ("fb_cms_width", C.c_uint64, [vp]),
    ("fb_cms_width_for_genome_size", C.c_uint64, [C.c_uint64]),
    ("fb_cms_clear", C.c_int, [vp]),
    ("fb_cms_add_keys", C.c_int, [vp, u64p, C.c_uint64]),
    ("fb_cms_query", C.c_int, [vp, u64p, C.c_uint64, u32p]),
    ("fb_cms_add_bank", C.c_int, [vp, vp, C.c_uint32]),
    ("fb_cms_table_bytes", C.c_uint64, [vp]),
    ("fb_cms_download", C.c_int, [vp, vp, C.c_uint64]),
    ("fb_hll_create", C.c_int, [vp, C.c_int, C.POINTER(vp)]),
    ("fb_hll_destroy", None, [vp]),
    ("fb_hll_clear", C.c_int, [vp]),
    ("fb_hll_add_hashes", C.c_int, [vp, u64p, C.c_uint64]),
    ("fb_hll_add_bank", C.c_int, [vp, vp, C.c_uint32]),
    ("fb_hll_add_canonical_hashes", C.c_int, [vp, u64p, C.c_uint64]),
    ("fb_hll_merge", C.c_int, [vp, vp]),
    ("fb_cms_add_and_estimate", C.c_int, [vp, u64p, C.c_uint64, u32p]),
    ("fb_cms_add_and_estimate_static", C.c_int, [vp, u64p, C.c_uint64, u32p]),
    ("fb_nthash_windows", C.c_int, [vp, u8p, C.c_uint64, C.c_uint32, u64p]),
    ("fb_hll_registers", C.c_int, [vp, u8p]),
    ("fb_hll_estimate", C.c_int, [vp, C.POINTER(C.c_double)]),
    ("fb_hll_estimate_registers", C.c_double, [u8p, C.c_int]),
    ("fb_build_params_default", None, [C.POINTER(BuildParams), C.c_uint32]),
    ("fb_build_solid", C.c_int, [vp, vp, u8p, u32p, C.c_uint64, C.POINTER(BuildParams), C.POINTER(vp)]),
    ("fb_solid_destroy", None, [vp]),
    ("fb_solid_count", C.c_uint64, [vp]),
    ("fb_solid_key_words", C.c_int, [vp]),
    ("fb_solid_download", C.c_int, [vp, u64p, u16p, u8p]),
    ("fb_solid_cms_bytes", C.c_uint64, [vp]),
    ("fb_solid_cms_download", C.c_int, [vp, vp, C.c_uint64]),
    ("fb_solid_hll_registers", C.c_int, [vp, u8p]),
    ("fb_solid_timings", C.c_int, [vp, C.POINTER(C.c_float)]),
    ("fb_solid_kernel_timings", C.c_int, [vp, C.POINTER(C.c_float)]),
    ("fb_diag_copy", C.c_int, [vp, C.c_uint64]),
    ("fb_diag_atomic_rate", C.c_int, [vp, C.c_uint32, C.c_uint64, C.c_int, C.POINTER(C.c_float)]),
    ("fb_solid_stats", C.c_int, [vp, u64p]),
    ("fb_solid_pass2_gate", C.c_int, [vp]),
    ("fb_solid_pass2_layout", C.c_int, [vp]),
    ("fb_compact_solid", C.c_int, [vp, C.POINTER(vp)]),
    ("fb_compact_arrays", C.c_int, [vp, u64p, u16p, u8p, C.c_uint64, C.c_uint32, C.POINTER(vp)]),
    ("fb_graph_destroy", None, [vp]),
    ("fb_graph_unitig_count", C.c_uint64, [vp]),
    ("fb_graph_base_count", C.c_uint64, [vp]),
    ("fb_graph_edge_count", C.c_uint64, [vp]),
    ("fb_graph_unitigs", C.c_int, [vp, u8p, u64p, u32p, C.POINTER(C.c_float), u8p]),
    ("fb_graph_edges", C.c_int, [vp, u32p, u32p, u8p, u16p]),
    ("fb_graph_timings", C.c_int, [vp, C.POINTER(C.c_float)]),
    ("fb_fastx_read", C.c_int, [C.c_char_p, C.POINTER(vp)]),
    ("fb_fastx_last_error", C.c_char_p, []),
    ("fb_fastx_count", C.c_uint64, [vp]),
    ("fb_fastx_view", C.c_int, [vp, C.POINTER(RecordsView)]),
    ("fb_fastx_destroy", None, [vp]),
    ("fb_ingest_params_default", None, [C.POINTER(IngestParams)]),
    ("fb_ingest_records", C.c_int, [vp, C.POINTER(RecordsView), C.POINTER(RecordsView), C.POINTER(IngestParams),
                                    C.POINTER(vp), C.POINTER(IngestStats)]),
    ("fb_ingest_files", C.c_int, [vp, C.c_char_p, C.c_char_p, C.POINTER(IngestParams), C.POINTER(vp),
                                  C.POINTER(IngestStats)]),
    ("fb_graph_view_of", C.c_int, [vp, C.POINTER(GraphView)]),
    ("fb_map_paired_reads", C.c_int, [vp, C.POINTER(GraphView), vp, C.c_uint32, C.POINTER(vp)]),
    ("fb_polish_contigs", C.c_int, [vp, C.POINTER(GraphView), vp, C.c_uint32, C.POINTER(C.c_uint8),
                                    C.POINTER(PolishStats)]),
    ("fb_pair_map_link_count", C.c_uint64, [vp]),
    ("fb_pair_map_links", C.c_int, [vp, C.POINTER(PairLink)]),
    ("fb_pair_map_sizes", C.c_int, [vp, C.POINTER(C.c_uint32)]),
    ("fb_pair_map_dist", C.c_int, [vp, C.POINTER(InsertSizeDist)]),
    ("fb_pair_map_destroy", None, [vp]),
    ("fb_filter_active_reads", C.c_int, [vp, vp, C.POINTER(GraphView), C.POINTER(vp), u8p, u8p]),
    ("fb_upgrade_multik", C.c_int, [vp, vp, C.POINTER(GraphView), C.c_uint32, C.c_uint64, C.POINTER(BuildParams),
                                    C.POINTER(MultiKResult), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
    ("fb_batch_create", C.c_int, [vp, C.POINTER(vp), u64p, C.c_uint32, C.POINTER(BuildParams), C.POINTER(vp)]),
    ("fb_batch_destroy", None, [vp]),
    ("fb_batch_footprint", C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(BuildParams), C.c_uint32]),
    ("fb_solid_out_bytes", None, [C.c_uint64, C.c_uint32, u64p]),
    ("fb_solid_out_fits", C.c_int, [C.c_uint64, C.c_uint32, u64p]),
    ("fb_plan_pass2", C.c_int, [C.c_double, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(Pass2Plan)]),
    ("fb_batch_run", C.c_int, [vp, C.c_uint32, C.c_int32]),
    ("fb_batch_size", C.c_uint32, [vp]),
    ("fb_batch_count", C.c_uint64, [vp, C.c_uint32]),
    ("fb_batch_key_words", C.c_int, [vp]),
    ("fb_batch_download", C.c_int, [vp, C.c_uint32, u64p, u16p, u8p]),
    ("fb_batch_cms_bytes", C.c_uint64, [vp, C.c_uint32]),
    ("fb_batch_cms_download", C.c_int, [vp, C.c_uint32, vp, C.c_uint64]),
    ("fb_batch_hll_registers", C.c_int, [vp, C.c_uint32, u8p]),
    ("fb_batch_timings", C.c_int, [vp, C.POINTER(C.c_float)]),
    ("fb_batch_routes", C.c_int, [vp, u32p]),
    ("fb_batch_autotune", C.c_int, [vp, C.c_uint32, C.c_int32, C.POINTER(TuneResult)]),
    ("fb_autotune_save", C.c_int, [C.c_char_p]),
    ("fb_autotune_load", C.c_int, [C.c_char_p]),
    ("fb_autotune_clear", None, []),
    ("fb_serve_params_default", None, [C.POINTER(ServeParams)]),
    ("fb_server_create", C.c_int, [C.c_int, C.POINTER(ServeParams), C.POINTER(vp)]),
    ("fb_server_destroy", None, [vp]),
    ("fb_server_last_error", C.c_char_p, [vp]),
    ("fb_server_submit", C.c_int, [vp, C.POINTER(ArenaView), C.POINTER(ArenaView), C.c_uint64, u64p]),
    ("fb_server_wait", C.c_int, [vp, C.c_uint64, C.c_double, C.POINTER(ServeResult)]),
    ("fb_server_download", C.c_int, [vp, C.c_uint64, C.c_uint32, u64p, u16p, u8p]),
    ("fb_server_sketches", C.c_int, [vp, C.c_uint64, C.c_uint32, vp, C.c_uint64, u8p]),
    ("fb_server_release", C.c_int, [vp, C.c_uint64]),
    ("fb_server_stats", C.c_int, [vp, C.POINTER(ServeStats)]),
    ("fb_admission_create", C.c_int, [C.c_uint64, C.c_uint32, C.POINTER(vp)]),
    ("fb_admission_destroy", None, [vp]),
    ("fb_admission_submit", C.c_int, [vp, C.c_uint64, C.c_uint64]),
    ("fb_admission_admit", C.c_uint32, [vp, u64p, C.c_uint32]),
    ("fb_admission_retire", C.c_int, [vp, C.c_uint64]),
    ("fb_admission_used", C.c_uint64, [vp]),
    ("fb_admission_queued", C.c_uint32, [vp]),
    ("fb_admission_inflight", C.c_uint32, [vp]),
    ("fb_count_stats_from_lengths", C.c_int, [u32p, u32p, C.c_uint64, C.POINTER(CountStats)]),
    ("fb_compute_dynamic_kmers", C.c_uint32, [C.c_uint64, C.c_int64, u32p]),
    ("fb_count_fastx", C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(CountStats)]),
    ("fb_estimate_genome_files", C.c_int, [vp, C.c_char_p, C.c_char_p, C.c_uint64, C.c_double,
                                           C.POINTER(GenomeEstimateC)]),
    ("fb_auto_subsample_factor", C.c_double, [C.POINTER(GenomeEstimateC), C.c_double, C.c_double, C.c_uint64,
                                              C.POINTER(C.c_int32)]),
    ("fb_estimate_genome", C.c_int, [vp, C.POINTER(ArenaView), C.POINTER(ArenaView), C.POINTER(CountStats),
                                     C.c_uint64, C.c_double, C.POINTER(GenomeEstimateC)]),
    ("fb_pack_ascii", C.c_int, [u8p, u32p, C.c_uint64, u64p, u64p]),
    ("fb_splitmix64", C.c_uint64, [C.c_uint64]),
    ("fb_synth_params_default", None, [C.POINTER(SynthParams)]),
    ("fb_synth_genome", C.c_int, [C.POINTER(SynthParams), u8p]),
    ("fb_synth_n_pairs", C.c_uint64, [C.POINTER(SynthParams)]),
    ("fb_synth_reads", C.c_int, [C.POINTER(SynthParams), u8p, C.c_uint64, C.c_uint64, u8p, u8p]),
]

_LIB = None


def load() -> C.CDLL:
    """Loads the built library once; raises when it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP extension with `python -m fastbact_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib
