"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline.
Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline leg import
this module.  Build with `make -C oracle`.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

u8p = C.POINTER(C.c_uint8)
u16p = C.POINTER(C.c_uint16)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
vp = C.c_void_p


class Arena(C.Structure):
    _fields_ = [("data", u64p), ("offsets", u64p), ("lengths", u32p), ("n_reads", C.c_uint64)]


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


class GenomeEstimate(C.Structure):
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


class BuildParams(C.Structure):
    _fields_ = [("width", C.c_uint64), ("counter_bits", C.c_int), ("t_solid", C.c_int),
                ("min_trans_cov", C.c_int), ("hll_p", C.c_int)]


class SolidSetC(C.Structure):
    _fields_ = [("n", C.c_uint64), ("nw", C.c_int), ("keys", u64p), ("cov", u16p), ("mask", u8p)]


class GraphC(C.Structure):
    _fields_ = [("n_unitigs", C.c_uint64), ("seq", u8p), ("offsets", u64p), ("lengths", u32p),
                ("cov", C.POINTER(C.c_float)), ("circular", u8p), ("n_edges", C.c_uint64), ("e_src", u32p),
                ("e_dst", u32p), ("e_ov", u8p), ("e_support", u16p)]


class GraphView(C.Structure):
    _fields_ = [("k", C.c_uint32), ("n_nodes", C.c_uint64), ("seq", u8p), ("offsets", u64p), ("lengths", u32p),
                ("flags", u8p), ("n_fwd", u32p), ("n_rev", u32p)]


class Records(C.Structure):
    _fields_ = [("seq", u8p), ("seq_off", u64p), ("seq_len", u32p), ("qual", u8p), ("qual_off", u64p),
                ("qual_len", u32p), ("n", C.c_uint64)]


class IngestParams(C.Structure):
    _fields_ = [("subsample_factor", C.c_double), ("min_len", C.c_uint32), ("min_qual", C.c_uint8),
                ("enable_trim", C.c_int), ("enable_merge", C.c_int), ("seed", C.c_uint64)]


class IngestStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("input_read_pairs", "retained_read_pairs", "bases_before_trim",
                                           "bases_after_trim", "merged_pairs_count", "merged_bases",
                                           "unmerged_pairs_count")]


class IngestOut(C.Structure):
    _fields_ = [("seq", u8p * 3), ("len", u32p * 3), ("n", C.c_uint64 * 3), ("stats", IngestStats)]


class PairLink(C.Structure):
    _fields_ = [("u1", C.c_uint32), ("u2", C.c_uint32), ("pair_id", C.c_uint32), ("dist_u1_exit", C.c_uint32),
                ("dist_u2_entry", C.c_uint32), ("r1_len", C.c_uint16), ("r2_len", C.c_uint16)]


class InsertDist(C.Structure):
    _fields_ = [("median", C.c_double), ("mad", C.c_double), ("min_insert", C.c_uint64), ("max_insert", C.c_uint64)]


class PairMap(C.Structure):
    _fields_ = [("n_links", C.c_uint64), ("n_sizes", C.c_uint64), ("links", C.POINTER(PairLink)), ("sizes", u32p),
                ("dist", InsertDist)]


class PolishStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("bases_modified", "fwd_mapped_reads", "rev_mapped_reads",
                                           "fwd_pileup_bases", "rev_pileup_bases", "rejected_anchors")]


class PolishOut(C.Structure):
    _fields_ = [("seq", u8p), ("stats", PolishStats)]


_SIGS = [
    ("orc_splitmix64", C.c_uint64, [C.c_uint64]),
    ("orc_rev_comp_word", C.c_uint64, [C.c_uint64, C.c_uint32]),
    ("orc_ascii_to_2bit", C.c_uint8, [C.c_uint8]),
    ("orc_arena_pack", None, [u8p, u32p, C.c_uint64, u64p, u64p]),
    ("orc_arena_extract_ascii", None, [C.POINTER(Arena), C.c_uint64, u8p]),
    ("orc_kmer_words_for_k", C.c_int, [C.c_uint32]),
    ("orc_kmer_from_ascii", None, [u8p, C.c_uint32, C.c_int, u64p]),
    ("orc_kmer_rev_comp", None, [u64p, C.c_uint32, C.c_int, u64p]),
    ("orc_kmer_roll_next", None, [u64p, u64p, C.c_uint8, C.c_uint32, C.c_int, u64p, u64p]),
    ("orc_kmer_to_u64", C.c_uint64, [u64p, C.c_int]),
    ("orc_kmer_canonical_windows", None, [u8p, C.c_uint32, C.c_uint32, C.c_int, u64p, u8p]),
    ("orc_nthash_fwd", C.c_uint64, [u8p, C.c_uint32]),
    ("orc_nthash_rev", C.c_uint64, [u8p, C.c_uint32]),
    ("orc_nthash_canonical", C.c_uint64, [u8p, C.c_uint32]),
    ("orc_nthash_rolling", None, [u8p, C.c_uint32, C.c_uint32, u64p]),
    ("orc_cms_new", vp, [C.c_uint64, C.c_int]),
    ("orc_cms_free", None, [vp]),
    ("orc_cms_width", C.c_uint64, [vp]),
    ("orc_cms_add", None, [vp, C.c_uint64]),
    ("orc_cms_estimate", C.c_uint32, [vp, C.c_uint64]),
    ("orc_cms_add_and_estimate", C.c_uint32, [vp, C.c_uint64]),
    ("orc_cms_table_bytes", C.c_size_t, [vp]),
    ("orc_cms_table", vp, [vp]),
    ("orc_cms_width_for_genome_size", C.c_uint64, [C.c_uint64]),
    ("orc_hll_add_hash", None, [u8p, C.c_int, C.c_uint64]),
    ("orc_hll_merge", None, [u8p, u8p, C.c_int]),
    ("orc_hll_estimate", C.c_double, [u8p, C.c_int]),
    ("orc_count_stats_from_lengths", C.c_int, [u32p, u32p, C.c_uint64, C.POINTER(CountStats)]),
    ("orc_compute_dynamic_kmers", C.c_uint32, [C.c_uint64, C.c_int64, u32p]),
    ("orc_estimate_genome", C.c_int, [C.POINTER(Arena), C.POINTER(Arena), C.POINTER(CountStats), C.c_uint64,
                                      C.c_double, C.POINTER(GenomeEstimate)]),
    ("orc_estimator_sketch", None, [C.POINTER(Arena), C.POINTER(Arena), C.c_uint64, u8p, u8p, u64p]),
    ("orc_estimate_finish", None, [C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(CountStats), C.c_double,
                                   C.POINTER(GenomeEstimate)]),
    ("orc_f64_to_u64", C.c_uint64, [C.c_double]),
    ("orc_build_params_default", None, [C.POINTER(BuildParams), C.c_int, C.c_uint64]),
    ("orc_build_solid", C.c_int, [C.POINTER(Arena), C.c_int, u8p, u32p, C.c_uint64, C.c_uint32,
                                  C.POINTER(BuildParams), C.POINTER(vp), u8p, C.POINTER(SolidSetC)]),
    ("orc_build_solid_mt", C.c_int, [C.POINTER(Arena), C.c_int, u8p, u32p, C.c_uint64, C.c_uint32,
                                     C.POINTER(BuildParams), C.c_int, C.POINTER(vp), u8p, C.POINTER(SolidSetC)]),
    ("orc_solid_free", None, [C.POINTER(SolidSetC)]),
    ("orc_compact", C.c_int, [u64p, u16p, u8p, C.c_uint64, C.c_uint32, C.POINTER(GraphC)]),
    ("orc_graph_free", None, [C.POINTER(GraphC)]),
    ("orc_filter_active_reads", C.c_int, [C.POINTER(GraphView), C.POINTER(Arena), u8p, u8p]),
    ("orc_trim_adapter", C.c_uint32, [u8p, C.c_uint32]),
    ("orc_trim_quality", C.c_int64, [u8p, C.c_uint32, C.c_uint32, C.c_uint8, C.c_uint32]),
    ("orc_merge_read_pair", C.c_uint32, [u8p, u8p, C.c_uint32, C.c_uint32, u8p, u8p, C.c_uint32, C.c_uint32,
                                         C.c_uint32, C.c_double, u8p, u32p]),
    ("orc_ingest_records", C.c_int, [C.POINTER(Records), C.POINTER(Records), C.POINTER(IngestParams),
                                     C.POINTER(IngestOut)]),
    ("orc_ingest_free", None, [C.POINTER(IngestOut)]),
    ("orc_map_paired_reads", C.c_int, [C.POINTER(GraphView), C.POINTER(Arena), C.POINTER(Arena), C.c_uint32,
                                       C.POINTER(PairMap)]),
    ("orc_pair_map_free", None, [C.POINTER(PairMap)]),
    ("orc_insert_dist_of", None, [u32p, C.c_uint64, C.POINTER(InsertDist)]),
    ("orc_polish_contigs", C.c_int, [C.POINTER(GraphView), C.POINTER(Arena), C.c_int, C.c_uint32,
                                     C.POINTER(PolishOut)]),
    ("orc_polish_free", None, [C.POINTER(PolishOut)]),
]

_LIB = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None and a.size else None


def _bytes(s) -> np.ndarray:
    return np.frombuffer(bytes(s), dtype=np.uint8).copy() if not isinstance(s, np.ndarray) else \
        np.ascontiguousarray(s, dtype=np.uint8)


# ---- primitives ------------------------------------------------------------

def splitmix64(z: int) -> int:
    return int(lib().orc_splitmix64(z & 0xFFFFFFFFFFFFFFFF))


def rev_comp_word(w: int, nb: int) -> int:
    return int(lib().orc_rev_comp_word(w, nb))


def words_for_k(k: int) -> int:
    return int(lib().orc_kmer_words_for_k(k))


def kmer_from_ascii(s: bytes, nw: int | None = None) -> np.ndarray:
    k = len(s)
    nw = nw or words_for_k(k)
    out = np.zeros(nw, np.uint64)
    b = _bytes(s)
    lib().orc_kmer_from_ascii(_p(b, C.c_uint8), k, nw, _p(out, C.c_uint64))
    return out


def kmer_rev_comp(x: np.ndarray, k: int) -> np.ndarray:
    x = np.ascontiguousarray(x, np.uint64)
    out = np.zeros_like(x)
    lib().orc_kmer_rev_comp(_p(x, C.c_uint64), k, x.size, _p(out, C.c_uint64))
    return out


def kmer_roll_next(fwd, rev, base: int, k: int):
    fwd = np.ascontiguousarray(fwd, np.uint64)
    rev = np.ascontiguousarray(rev, np.uint64)
    of, orv = np.zeros_like(fwd), np.zeros_like(rev)
    lib().orc_kmer_roll_next(_p(fwd, C.c_uint64), _p(rev, C.c_uint64), base, k, fwd.size, _p(of, C.c_uint64),
                             _p(orv, C.c_uint64))
    return of, orv


def kmer_to_u64(x) -> int:
    x = np.ascontiguousarray(x, np.uint64)
    return int(lib().orc_kmer_to_u64(_p(x, C.c_uint64), x.size))


def canonical_windows(seq: bytes, k: int, nw: int | None = None):
    nw = nw or words_for_k(k)
    n = max(0, len(seq) - k + 1)
    out = np.zeros((n, nw), np.uint64)
    rc = np.zeros(n, np.uint8)
    b = _bytes(seq)
    lib().orc_kmer_canonical_windows(_p(b, C.c_uint8), len(seq), k, nw, _p(out, C.c_uint64), _p(rc, C.c_uint8))
    return out, rc


def nthash_fwd(s: bytes) -> int:
    b = _bytes(s)
    return int(lib().orc_nthash_fwd(_p(b, C.c_uint8), len(s)))


def nthash_rev(s: bytes) -> int:
    b = _bytes(s)
    return int(lib().orc_nthash_rev(_p(b, C.c_uint8), len(s)))


def nthash_canonical(s: bytes) -> int:
    b = _bytes(s)
    return int(lib().orc_nthash_canonical(_p(b, C.c_uint8), len(s)))


def nthash_rolling(seq: bytes, k: int) -> np.ndarray:
    n = max(0, len(seq) - k + 1)
    out = np.zeros(n, np.uint64)
    b = _bytes(seq)
    lib().orc_nthash_rolling(_p(b, C.c_uint8), len(seq), k, _p(out, C.c_uint64))
    return out


# ---- sketches --------------------------------------------------------------

class Cms:
    """CountMinSketch (sketch.rs:121-254), nibble (4) or u16 (16) counters."""

    def __init__(self, width: int, counter_bits: int = 4, handle=None):
        self.h = handle if handle is not None else lib().orc_cms_new(width, counter_bits)
        self.counter_bits = counter_bits

    @property
    def width(self) -> int:
        return int(lib().orc_cms_width(self.h))

    def add(self, key: int) -> None:
        lib().orc_cms_add(self.h, key & 0xFFFFFFFFFFFFFFFF)

    def estimate(self, key: int) -> int:
        return int(lib().orc_cms_estimate(self.h, key & 0xFFFFFFFFFFFFFFFF))

    def add_and_estimate(self, key: int) -> int:
        return int(lib().orc_cms_add_and_estimate(self.h, key & 0xFFFFFFFFFFFFFFFF))

    def table(self) -> np.ndarray:
        n = int(lib().orc_cms_table_bytes(self.h))
        raw = np.ctypeslib.as_array(C.cast(lib().orc_cms_table(self.h), u8p), shape=(n,)).copy()
        return raw if self.counter_bits == 4 else raw.view(np.uint16)

    def __del__(self):
        try:
            if self.h:
                lib().orc_cms_free(self.h)
                self.h = None
        except Exception:
            pass


def cms_width_for_genome_size(g: int) -> int:
    return int(lib().orc_cms_width_for_genome_size(g))


def hll_merge(regs: np.ndarray, other: np.ndarray, p: int) -> None:
    lib().orc_hll_merge(_p(regs, C.c_uint8), _p(other, C.c_uint8), p)


def hll_add_hashes(regs: np.ndarray, p: int, hashes) -> None:
    for h in np.asarray(hashes, dtype=np.uint64).tolist():
        lib().orc_hll_add_hash(_p(regs, C.c_uint8), p, h)


def hll_estimate(regs: np.ndarray, p: int) -> float:
    r = np.ascontiguousarray(regs, np.uint8)
    return float(lib().orc_hll_estimate(_p(r, C.c_uint8), p))


# ---- arenas ----------------------------------------------------------------

def pack(seqs) -> tuple:
    """BitSeqArena::push_seq for each sequence -> (data, offsets, lengths)."""
    lens = np.array([len(s) for s in seqs], np.uint32)
    return pack_concat(b"".join(bytes(s) for s in seqs), lens)


def pack_concat(ascii_concat, lengths) -> tuple:
    lens = np.ascontiguousarray(lengths, np.uint32)
    buf = _bytes(ascii_concat)
    total = int(lens.sum(dtype=np.uint64))
    data = np.zeros((total + 31) // 32 + 1, np.uint64)
    offs = np.zeros(lens.size, np.uint64)
    lib().orc_arena_pack(_p(buf, C.c_uint8), _p(lens, C.c_uint32), lens.size, _p(data, C.c_uint64),
                         _p(offs, C.c_uint64))
    return data, offs, lens


def arena(data, offsets, lengths) -> Arena:
    a = Arena(_p(data, C.c_uint64), _p(offsets, C.c_uint64), _p(lengths, C.c_uint32), lengths.size)
    a._keep = (data, offsets, lengths)
    return a


# ---- estimator -------------------------------------------------------------

def count_stats(l1, l2) -> CountStats | None:
    a = np.ascontiguousarray(l1, np.uint32)
    b = np.ascontiguousarray(l2, np.uint32)
    st = CountStats()
    rc = lib().orc_count_stats_from_lengths(_p(a, C.c_uint32), _p(b, C.c_uint32), a.size, C.byref(st))
    return None if rc else st


def compute_dynamic_kmers(median_len: int, target_max_k: int | None = None) -> list:
    out = (C.c_uint32 * 8)()
    n = lib().orc_compute_dynamic_kmers(median_len, -1 if target_max_k is None else target_max_k, out)
    return [int(out[i]) for i in range(n)]


def estimate_genome(r1: tuple, r2: tuple, stats: CountStats, max_pairs: int, target_cov: float):
    a1, a2 = arena(*r1), arena(*r2)
    out = GenomeEstimate()
    rc = lib().orc_estimate_genome(C.byref(a1), C.byref(a2), C.byref(stats), max_pairs, float(target_cov),
                                   C.byref(out))
    return None if rc else out


def estimator_sketch(r1: tuple, r2: tuple, n_pairs: int):
    a1, a2 = arena(*r1), arena(*r2)
    ht = np.zeros(1 << 14, np.uint8)
    hs = np.zeros(1 << 14, np.uint8)
    m = C.c_uint64()
    lib().orc_estimator_sketch(C.byref(a1), C.byref(a2), n_pairs, _p(ht, C.c_uint8), _p(hs, C.c_uint8), C.byref(m))
    return ht, hs, int(m.value)


def estimate_finish(u1: int, u2: int, m_sample: int, stats: CountStats, target_cov: float):
    out = GenomeEstimate()
    lib().orc_estimate_finish(u1, u2, m_sample, C.byref(stats), float(target_cov), C.byref(out))
    return out


# ---- solid set -------------------------------------------------------------

def build_solid(arenas: list, k: int, target_solid_cov: int | None = None, width: int = 1 << 22,
                counter_bits: int = 4, hll_p: int = 0, contigs=None, threads: int = 0):
    """build_cdbg_with_contigs_ext passes 1-3 -> dict(keys, cov, mask, cms, hll).

    arenas: list of (data, offsets, lengths); threads 0 = sequential form."""
    p = BuildParams()
    lib().orc_build_params_default(C.byref(p), -1 if target_solid_cov is None else target_solid_cov, width)
    p.counter_bits = counter_bits
    p.hll_p = hll_p
    arr = (Arena * max(1, len(arenas)))(*[arena(*a) for a in arenas])
    keep = [arena(*a) for a in arenas]  # noqa: F841 keep buffers alive
    if contigs:
        cl = np.array([len(c) for c in contigs], np.uint32)
        cb = _bytes(b"".join(contigs))
    else:
        cl, cb = None, None
    cms_h = vp()
    regs = np.zeros(1 << hll_p, np.uint8) if hll_p else None
    ss = SolidSetC()
    if threads:
        rc = lib().orc_build_solid_mt(arr, len(arenas), _p(cb, C.c_uint8), _p(cl, C.c_uint32),
                                      0 if cl is None else cl.size, k, C.byref(p), threads, C.byref(cms_h),
                                      _p(regs, C.c_uint8), C.byref(ss))
    else:
        rc = lib().orc_build_solid(arr, len(arenas), _p(cb, C.c_uint8), _p(cl, C.c_uint32),
                                   0 if cl is None else cl.size, k, C.byref(p), C.byref(cms_h),
                                   _p(regs, C.c_uint8), C.byref(ss))
    if rc:
        raise ValueError(f"orc_build_solid failed ({rc})")
    n, nw = int(ss.n), int(ss.nw)
    keys = np.ctypeslib.as_array(ss.keys, shape=(max(n, 1) * nw,))[: n * nw].copy().reshape(n, nw)
    cov = np.ctypeslib.as_array(ss.cov, shape=(max(n, 1),))[:n].copy()
    mask = np.ctypeslib.as_array(ss.mask, shape=(max(n, 1),))[:n].copy()
    lib().orc_solid_free(C.byref(ss))
    cms = Cms(width, counter_bits, handle=cms_h.value)
    return {"keys": keys, "cov": cov, "mask": mask, "cms": cms, "hll": regs}


def compact(keys, cov, mask, k: int) -> dict:
    """Unitig compaction (build.rs:330-597) of a sorted solid set -> dict with
    seqs (list of bytes), cov (float32), circular (uint8) and the edge arrays
    src, dst, ov, support (each node's edges in insertion order, forward side
    then reverse side)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    cov = np.ascontiguousarray(cov, dtype=np.uint16)
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    n = int(cov.size)
    g = GraphC()
    rc = lib().orc_compact(_p(keys, C.c_uint64), _p(cov, C.c_uint16), _p(mask, C.c_uint8), n, k, C.byref(g))
    if rc:
        raise ValueError(f"orc_compact failed ({rc})")
    nu, ne = int(g.n_unitigs), int(g.n_edges)

    def arr(ptr, count, dt):
        if count == 0:
            return np.zeros(0, dt)
        return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dt, copy=True)

    offs = arr(g.offsets, nu, np.uint64)
    lens = arr(g.lengths, nu, np.uint32)
    total = int(offs[-1] + lens[-1]) if nu else 0
    seq = arr(g.seq, total, np.uint8).tobytes()
    out = {"seqs": [seq[int(o):int(o) + int(ln)] for o, ln in zip(offs, lens)],
           "cov": arr(g.cov, nu, np.float32), "circular": arr(g.circular, nu, np.uint8),
           "src": arr(g.e_src, ne, np.uint32), "dst": arr(g.e_dst, ne, np.uint32),
           "ov": arr(g.e_ov, ne, np.uint8), "support": arr(g.e_support, ne, np.uint16)}
    lib().orc_graph_free(C.byref(g))
    return out


# ---- filter_active_reads (local_multik.rs:259-333) ----------------------------

def edge_counts(graph: dict) -> tuple:
    """edges_fwd.len() and edges_rev.len() per node from the edge arrays
    (src = id << 1 | side)."""
    n = len(graph["seqs"])
    src = np.asarray(graph["src"], np.int64)
    nf = np.bincount(src[(src & 1) == 0] >> 1, minlength=n).astype(np.uint32)
    nr = np.bincount(src[(src & 1) == 1] >> 1, minlength=n).astype(np.uint32)
    return nf, nr


def unpack(data, offsets, lengths, ids=None) -> list:
    """BitSeqArena::extract_ascii for the reads `ids` (all by default)."""
    ids = np.arange(len(lengths)) if ids is None else np.asarray(ids, np.int64)
    out = []
    for i in ids:
        g = int(offsets[i]) + np.arange(int(lengths[i]), dtype=np.uint64)
        codes = (data[(g >> np.uint64(5)).astype(np.int64)] >> ((g & np.uint64(31)) << np.uint64(1))) & np.uint64(3)
        out.append(np.frombuffer(b"ACGT", np.uint8)[codes.astype(np.int64)].tobytes())
    return out


def filter_active_reads(graph: dict, k: int, arenas: list, pack_reads: bool = True) -> dict:
    """filter_active_reads on a graph dict (seqs, circular, src; optional
    deleted) at graph k and a bank of three arenas (merged, pe_forward,
    pe_reverse as (data, offsets, lengths)) -> keep_merged, keep_pairs and the
    filtered bank's three arenas, packed as BitSeqArena::push_seq does."""
    seqs = [bytes(x) for x in graph["seqs"]]
    n = len(seqs)
    lens = np.array([len(x) for x in seqs], np.uint32)
    offs = np.zeros(n, np.uint64)
    if n:
        offs[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
    buf = _bytes(b"".join(seqs))
    flags = (np.asarray(graph["circular"], np.uint8) << 1).astype(np.uint8)
    if "deleted" in graph:
        flags |= np.asarray(graph["deleted"], np.uint8)
    nf, nr = edge_counts(graph)
    gv = GraphView(k, n, _p(buf, C.c_uint8), _p(offs, C.c_uint64), _p(lens, C.c_uint32), _p(flags, C.c_uint8),
                   _p(nf, C.c_uint32), _p(nr, C.c_uint32))
    ars = [tuple(np.ascontiguousarray(x) for x in a) for a in arenas]
    arr = (Arena * 3)(*[arena(*a) for a in ars])
    km = np.zeros(max(1, ars[0][2].size), np.uint8)
    npairs = min(ars[1][2].size, ars[2][2].size)
    kp = np.zeros(max(1, npairs), np.uint8)
    rc = lib().orc_filter_active_reads(C.byref(gv), arr, _p(km, C.c_uint8), _p(kp, C.c_uint8))
    if rc:
        raise ValueError(f"orc_filter_active_reads failed ({rc})")
    km, kp = km[:ars[0][2].size], kp[:npairs]
    if not pack_reads:
        return {"keep_merged": km, "keep_pairs": kp}
    merged = unpack(*ars[0], np.flatnonzero(km))
    r1 = unpack(*ars[1], np.flatnonzero(kp))
    r2 = unpack(*ars[2], np.flatnonzero(kp))
    return {"keep_merged": km, "keep_pairs": kp, "arenas": [pack(merged), pack(r1), pack(r2)],
            "reads": [merged, r1, r2]}


# ---- ingest (trim_merge.rs:17-417) -----------------------------------------

def trim_adapter(seq: bytes) -> int:
    b = _bytes(seq)
    return int(lib().orc_trim_adapter(_p(b, C.c_uint8), b.size))


def trim_quality(qual: bytes, min_qual: int, min_len: int, seq_len: int | None = None):
    q = _bytes(qual)
    n = q.size if seq_len is None else seq_len
    r = int(lib().orc_trim_quality(_p(q, C.c_uint8), n, q.size, min_qual, min_len))
    return None if r < 0 else r


def merge_read_pair(r1: bytes, q1: bytes, r2: bytes, q2: bytes, min_overlap: int = 15, ratio: float = 0.10):
    """-> (merged bytes, overlap, mismatches) or None when unmerged."""
    a, qa, b, qb = (_bytes(x) for x in (r1, q1, r2, q2))
    out = np.zeros(a.size + b.size + 1, np.uint8)
    mm = np.zeros(1, np.uint32)
    d = int(lib().orc_merge_read_pair(_p(a, C.c_uint8), _p(qa, C.c_uint8), a.size, qa.size, _p(b, C.c_uint8),
                                      _p(qb, C.c_uint8), b.size, qb.size, min_overlap, ratio, _p(out, C.c_uint8),
                                      _p(mm, C.c_uint32)))
    if d == 0:
        return None
    return out[:a.size + b.size - d].tobytes(), d, int(mm[0])


def records(seqs, quals=None) -> tuple:
    """Parsed records as (seq bytes, seq offsets, seq lengths, qual bytes, qual offsets, qual lengths)."""
    quals = quals if quals is not None else [b""] * len(seqs)
    sl = np.array([len(x) for x in seqs], np.uint32)
    ql = np.array([len(x) for x in quals], np.uint32)
    so = np.zeros(len(seqs), np.uint64)
    qo = np.zeros(len(seqs), np.uint64)
    if len(seqs):
        so[1:] = np.cumsum(sl[:-1], dtype=np.uint64)
        qo[1:] = np.cumsum(ql[:-1], dtype=np.uint64)
    return (_bytes(b"".join(bytes(x) for x in seqs)), so, sl, _bytes(b"".join(bytes(x) for x in quals)), qo, ql)


def _records_c(rec) -> Records:
    s, so, sl, q, qo, ql = rec
    r = Records(_p(s, C.c_uint8), _p(so, C.c_uint64), _p(sl, C.c_uint32), _p(q, C.c_uint8), _p(qo, C.c_uint64),
                _p(ql, C.c_uint32), sl.size)
    r._keep = rec
    return r


INGEST_ERRORS = {-1: "Mismatched paired-end FASTQ mate counts between R1 and R2",
                 -2: "No valid reads survived trimming and quality filtering",
                 -3: "quality string shorter than its adapter-trimmed read"}


def ingest_records(r1, r2, subsample_factor: float = 1.0, min_len: int = 50, min_qual: int = 20,
                   enable_trim: bool = True, enable_merge: bool = True, seed: int = 0) -> dict:
    """ingest_reads_ext_with_seed's record loop on parsed records (records()):
    -> {"reads": [merged, pe_forward, pe_reverse] as lists of bytes, "arenas":
    the three arenas packed as push_seq does, "stats": IngestStats fields}.
    Raises ValueError with the reference's message on its error paths."""
    a, b = _records_c(r1), _records_c(r2)
    prm = IngestParams(subsample_factor, min_len, min_qual, int(enable_trim), int(enable_merge), seed)
    out = IngestOut()
    rc = lib().orc_ingest_records(C.byref(a), C.byref(b), C.byref(prm), C.byref(out))
    try:
        if rc:
            raise ValueError(INGEST_ERRORS.get(rc, f"orc_ingest_records failed ({rc})"))
        reads = []
        for i in range(3):
            n = int(out.n[i])
            lens = np.ctypeslib.as_array(out.len[i], shape=(n,)).copy() if n else np.zeros(0, np.uint32)
            tot = int(lens.sum(dtype=np.uint64))
            raw = np.ctypeslib.as_array(out.seq[i], shape=(tot,)).tobytes() if tot else b""
            offs = np.concatenate([[0], np.cumsum(lens, dtype=np.uint64)]).astype(np.int64)
            reads.append([raw[offs[j]:offs[j + 1]] for j in range(n)])
        st = {f: int(getattr(out.stats, f)) for f, _ in IngestStats._fields_}
    finally:
        lib().orc_ingest_free(C.byref(out))
    return {"reads": reads, "arenas": [pack(r) for r in reads], "stats": st}


# ---- PE mapper (mapper.rs:89-279) --------------------------------------------

LINK_FIELDS = ("u1", "u2", "pair_id", "dist_u1_exit", "dist_u2_entry", "r1_len", "r2_len")


def _graph_view(graph: dict, k: int):
    seqs = [bytes(x) for x in graph["seqs"]]
    n = len(seqs)
    lens = np.array([len(x) for x in seqs], np.uint32)
    offs = np.zeros(n, np.uint64)
    if n:
        offs[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
    buf = _bytes(b"".join(seqs))
    flags = (np.asarray(graph.get("circular", np.zeros(n)), np.uint8) << 1).astype(np.uint8)
    if "deleted" in graph:
        flags |= np.asarray(graph["deleted"], np.uint8)
    gv = GraphView(k, n, _p(buf, C.c_uint8), _p(offs, C.c_uint64), _p(lens, C.c_uint32), _p(flags, C.c_uint8),
                   None, None)
    gv._keep = (buf, offs, lens, flags)
    return gv


def map_paired_reads(graph: dict, pe_forward, pe_reverse, k: int) -> dict:
    """map_paired_reads over a graph dict and the two mate arenas ((data,
    offsets, lengths)) -> {"links": structured array of the evidence records
    in pair order, "sizes": insert sizes in pair order, "dist": InsertSizeDist}."""
    gv = _graph_view(graph, k)
    a, b = arena(*pe_forward), arena(*pe_reverse)
    m = PairMap()
    rc = lib().orc_map_paired_reads(C.byref(gv), C.byref(a), C.byref(b), k, C.byref(m))
    if rc:
        raise ValueError(f"orc_map_paired_reads failed ({rc})")
    try:
        dt = np.dtype([(f, np.uint32) for f in LINK_FIELDS[:5]] + [("r1_len", np.uint16), ("r2_len", np.uint16)])
        nl = int(m.n_links)
        links = np.frombuffer(C.string_at(m.links, nl * dt.itemsize), dt).copy() if nl else np.zeros(0, dt)
        ns = int(m.n_sizes)
        sizes = np.ctypeslib.as_array(m.sizes, shape=(ns,)).copy() if ns else np.zeros(0, np.uint32)
        dist = {"median": m.dist.median, "mad": m.dist.mad, "min_insert": int(m.dist.min_insert),
                "max_insert": int(m.dist.max_insert)}
    finally:
        lib().orc_pair_map_free(C.byref(m))
    return {"links": links, "sizes": sizes, "dist": dist}


def insert_dist(sizes) -> dict:
    s = np.ascontiguousarray(sizes, np.uint32)
    d = InsertDist()
    lib().orc_insert_dist_of(_p(s, C.c_uint32), s.size, C.byref(d))
    return {"median": d.median, "mad": d.mad, "min_insert": int(d.min_insert), "max_insert": int(d.max_insert)}


# ---- consensus polishing (consensus.rs:64-271) --------------------------------

def polish_contigs(graph: dict, read_arenas: list, min_output_len: int) -> dict:
    """polish_contigs over a graph dict (seqs; optional deleted) and read
    arenas (the reference passes merged reads then pe_forward) -> {"seqs":
    every node's sequence after polishing, "stats": PolishStats fields}."""
    gv = _graph_view(graph, 31)
    ars = [arena(*a) for a in read_arenas]
    arr = (Arena * max(1, len(ars)))(*ars)
    out = PolishOut()
    rc = lib().orc_polish_contigs(C.byref(gv), arr, len(ars), min_output_len, C.byref(out))
    if rc:
        raise ValueError(f"orc_polish_contigs failed ({rc})")
    try:
        lens = [len(bytes(x)) for x in graph["seqs"]]
        raw = C.string_at(out.seq, sum(lens)) if sum(lens) else b""
        offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        seqs = [raw[offs[i]:offs[i + 1]] for i in range(len(lens))]
        st = {f: int(getattr(out.stats, f)) for f, _ in PolishStats._fields_}
    finally:
        lib().orc_polish_free(C.byref(out))
    return {"seqs": seqs, "stats": st}
