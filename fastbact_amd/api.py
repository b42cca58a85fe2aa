"""Host-side mirror of fastBact's k-mer spectrum interface over the C ABI.

Names and argument meaning follow the reference:
  ReadBank / BitSeqArena            src/common/bitseq.rs:145-399
  CountMinSketch                    src/common/sketch.rs:121-254
  HyperLogLog                       src/common/sketch.rs:24-115
  build_solid_kmers                 src/graph/build.rs:37-328 (passes 1-3 of
                                    build_cdbg_with_contigs_ext)
  estimate_genome, count_stats, compute_dynamic_kmers
                                    src/ingest/estimator.rs:104-440
Every call runs on the GPU through libfastbact_gpu.so; errors raise
FastBactError carrying the library's status and message.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import ArenaView, BuildParams, CountStats, GenomeEstimateC, SynthParams


class FastBactError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_lib.ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None and a.size else None


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def splitmix64(z: int) -> int:
    return int(_lib.load().fb_splitmix64(C.c_uint64(z & 0xFFFFFFFFFFFFFFFF)))


# ---------------------------------------------------------------------------
# Arenas


@dataclass
class BitSeqArena:
    """2-bit packed reads (bitseq.rs:145-149): base g at bits 2*(g&31) of data[g>>5]."""

    data: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    offsets: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    lengths: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))

    @classmethod
    def from_ascii(cls, seqs: Sequence[bytes]) -> "BitSeqArena":
        lens = np.array([len(s) for s in seqs], dtype=np.uint32)
        return cls.from_concat(b"".join(seqs), lens)

    @classmethod
    def from_concat(cls, ascii_concat, lengths) -> "BitSeqArena":
        """Packs concatenated ASCII reads (BitSeqArena::push_seq for each)."""
        lens = _u32(lengths)
        buf = np.frombuffer(ascii_concat, dtype=np.uint8) if isinstance(ascii_concat, (bytes, bytearray)) \
            else np.ascontiguousarray(ascii_concat, dtype=np.uint8)
        total = int(lens.sum(dtype=np.uint64))
        if buf.size < total:
            raise ValueError("ascii shorter than the sum of lengths")
        data = np.zeros((total + 31) // 32, dtype=np.uint64)
        offs = np.zeros(lens.size, dtype=np.uint64)
        rc = _lib.load().fb_pack_ascii(_ptr(buf, C.c_uint8), _ptr(lens, C.c_uint32), lens.size,
                                       _ptr(data, C.c_uint64), _ptr(offs, C.c_uint64))
        if rc:
            raise FastBactError(rc, "fb_pack_ascii failed")
        return cls(data, offs, lens)

    def __len__(self) -> int:
        return int(self.lengths.size)

    def total_bases(self) -> int:
        return int(self.lengths.sum(dtype=np.uint64))

    def view(self) -> ArenaView:
        self.data = _u64(self.data)
        self.offsets = _u64(self.offsets)
        self.lengths = _u32(self.lengths)
        return ArenaView(_ptr(self.data, C.c_uint64), self.data.size, _ptr(self.offsets, C.c_uint64),
                         _ptr(self.lengths, C.c_uint32), self.lengths.size)


@dataclass
class ReadBank:
    """{merged_reads, pe_forward, pe_reverse} (bitseq.rs:350-354)."""

    merged_reads: BitSeqArena = field(default_factory=BitSeqArena)
    pe_forward: BitSeqArena = field(default_factory=BitSeqArena)
    pe_reverse: BitSeqArena = field(default_factory=BitSeqArena)

    def arenas(self):
        return [self.merged_reads, self.pe_forward, self.pe_reverse]

    def total_bases(self) -> int:
        return sum(a.total_bases() for a in self.arenas())

    def total_reads(self) -> int:
        return sum(len(a) for a in self.arenas())


# ---------------------------------------------------------------------------
# Device context and objects


class Context:
    """One GPU context: a HIP stream and a device memory pool."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.fb_ctx_create(device, hbm_budget, C.byref(h))
        if rc:
            raise FastBactError(rc, f"fb_ctx_create(device={device}) failed")
        self.h = h
        self.device = device

    def check(self, rc: int, what: str) -> None:
        if rc:
            raise FastBactError(rc, f"{what}: {self.lib.fb_last_error(self.h).decode(errors='replace')}")

    def synchronize(self) -> None:
        self.check(self.lib.fb_ctx_synchronize(self.h), "fb_ctx_synchronize")

    def memory(self) -> tuple:
        """(free, total) bytes of the context's device."""
        f, t = C.c_uint64(), C.c_uint64()
        self.check(self.lib.fb_device_memory(self.device, C.byref(f), C.byref(t)), "fb_device_memory")
        return int(f.value), int(t.value)

    def trim(self) -> None:
        """Returns the context's cached free device buffers (fb_ctx_trim)."""
        self.check(self.lib.fb_ctx_trim(self.h), "fb_ctx_trim")

    def device_bytes(self) -> int:
        return int(self.lib.fb_ctx_device_bytes(self.h))

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.fb_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Handle:
    _destroy = ""

    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            getattr(self.ctx.lib, self._destroy)(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBank(_Handle):
    """A ReadBank resident in HBM."""

    _destroy = "fb_bank_destroy"

    def __init__(self, ctx: Context, bank: ReadBank):
        views = (ArenaView * 3)(*[a.view() for a in bank.arenas()])
        h = C.c_void_p()
        ctx.check(ctx.lib.fb_bank_upload(ctx.h, views, C.byref(h)), "fb_bank_upload")
        super().__init__(ctx, h)

    def total_bases(self) -> int:
        return int(self.ctx.lib.fb_bank_total_bases(self.h))

    def total_reads(self) -> int:
        return int(self.ctx.lib.fb_bank_total_reads(self.h))

    def bases_for_k(self, k: int) -> int:
        return int(self.ctx.lib.fb_bank_bases_for_k(self.h, k))

    @classmethod
    def _adopt(cls, ctx: Context, h) -> "DeviceBank":
        obj = cls.__new__(cls)
        _Handle.__init__(obj, ctx, h)
        return obj

    def download(self) -> ReadBank:
        """The three arenas back on the host (BitSeqArena layout)."""
        out = []
        for a in range(3):
            nr, nb, nw = C.c_uint64(), C.c_uint64(), C.c_uint64()
            self.ctx.check(self.ctx.lib.fb_bank_arena_info(self.h, a, C.byref(nr), C.byref(nb), C.byref(nw)),
                           "fb_bank_arena_info")
            data = np.zeros(max(nw.value, 1), np.uint64)
            offs = np.zeros(max(nr.value, 1), np.uint64)
            lens = np.zeros(max(nr.value, 1), np.uint32)
            self.ctx.check(self.ctx.lib.fb_bank_download_arena(self.h, a, _ptr(data, C.c_uint64),
                                                               _ptr(offs, C.c_uint64), _ptr(lens, C.c_uint32)),
                           "fb_bank_download_arena")
            out.append(BitSeqArena(data[:nw.value], offs[:nr.value], lens[:nr.value]))
        return ReadBank(*out)


class CountMinSketch(_Handle):
    """CountMinSketch (sketch.rs:121-254); counter_bits 4 = reference nibbles, 16 = u16."""

    _destroy = "fb_cms_destroy"

    def __init__(self, ctx: Context, width: int, counter_bits: int = 4):
        h = C.c_void_p()
        ctx.check(ctx.lib.fb_cms_create(ctx.h, width, counter_bits, C.byref(h)), "fb_cms_create")
        super().__init__(ctx, h)
        self.counter_bits = counter_bits

    @staticmethod
    def width_for_genome_size(g: int) -> int:
        return int(_lib.load().fb_cms_width_for_genome_size(g))

    @property
    def width(self) -> int:
        return int(self.ctx.lib.fb_cms_width(self.h))

    def add(self, keys) -> None:
        k = _u64(np.atleast_1d(keys))
        self.ctx.check(self.ctx.lib.fb_cms_add_keys(self.h, _ptr(k, C.c_uint64), k.size), "fb_cms_add_keys")

    def estimate(self, keys) -> np.ndarray:
        k = _u64(np.atleast_1d(keys))
        out = np.zeros(k.size, np.uint32)
        self.ctx.check(self.ctx.lib.fb_cms_query(self.h, _ptr(k, C.c_uint64), k.size, _ptr(out, C.c_uint32)),
                       "fb_cms_query")
        return out

    def add_and_estimate(self, keys) -> np.ndarray:
        """add_and_estimate (sketch.rs:248-253) for each key in order."""
        k = _u64(np.atleast_1d(keys))
        out = np.zeros(k.size, np.uint32)
        self.ctx.check(self.ctx.lib.fb_cms_add_and_estimate(self.h, _ptr(k, C.c_uint64), k.size,
                                                            _ptr(out, C.c_uint32)), "fb_cms_add_and_estimate")
        return out

    def add_bank(self, bank: DeviceBank, k: int) -> None:
        self.ctx.check(self.ctx.lib.fb_cms_add_bank(self.h, bank.h, k), "fb_cms_add_bank")

    def clear(self) -> None:
        self.ctx.check(self.ctx.lib.fb_cms_clear(self.h), "fb_cms_clear")

    def table(self) -> np.ndarray:
        """Counters in the reference layout (uint8 nibble bytes or uint16)."""
        nbytes = int(self.ctx.lib.fb_cms_table_bytes(self.h))
        out = np.zeros(nbytes, np.uint8)
        self.ctx.check(self.ctx.lib.fb_cms_download(self.h, out.ctypes.data, nbytes), "fb_cms_download")
        return out if self.counter_bits == 4 else out.view(np.uint16)


class HyperLogLog(_Handle):
    """HyperLogLog (sketch.rs:24-115) at precision p."""

    _destroy = "fb_hll_destroy"

    def __init__(self, ctx: Context, p: int = 14):
        h = C.c_void_p()
        ctx.check(ctx.lib.fb_hll_create(ctx.h, p, C.byref(h)), "fb_hll_create")
        super().__init__(ctx, h)
        self.p = p

    def add_hash(self, hashes) -> None:
        a = _u64(np.atleast_1d(hashes))
        self.ctx.check(self.ctx.lib.fb_hll_add_hashes(self.h, _ptr(a, C.c_uint64), a.size), "fb_hll_add_hashes")

    def add_canonical_hash(self, hashes) -> None:
        """add_canonical_hash (sketch.rs:66-70): add_hash(splitmix64(h))."""
        a = _u64(np.atleast_1d(hashes))
        self.ctx.check(self.ctx.lib.fb_hll_add_canonical_hashes(self.h, _ptr(a, C.c_uint64), a.size),
                       "fb_hll_add_canonical_hashes")

    def merge(self, other: "HyperLogLog") -> None:
        """HyperLogLog::merge (sketch.rs:73-80)."""
        self.ctx.check(self.ctx.lib.fb_hll_merge(self.h, other.h), "fb_hll_merge")

    def add_bank(self, bank: DeviceBank, k: int) -> None:
        self.ctx.check(self.ctx.lib.fb_hll_add_bank(self.h, bank.h, k), "fb_hll_add_bank")

    def registers(self) -> np.ndarray:
        out = np.zeros(1 << self.p, np.uint8)
        self.ctx.check(self.ctx.lib.fb_hll_registers(self.h, _ptr(out, C.c_uint8)), "fb_hll_registers")
        return out

    def estimate(self) -> float:
        v = C.c_double()
        self.ctx.check(self.ctx.lib.fb_hll_estimate(self.h, C.byref(v)), "fb_hll_estimate")
        return v.value

    def clear(self) -> None:
        self.ctx.check(self.ctx.lib.fb_hll_clear(self.h), "fb_hll_clear")


def nthash_windows(ctx: Context, seq: bytes, k: int) -> np.ndarray:
    """RollingNtHash (kmer.rs:433-477): the canonical ntHash of every window
    of seq (new on the first k bases, then roll)."""
    buf = np.frombuffer(bytes(seq), np.uint8).copy()
    out = np.zeros(max(buf.size - k + 1, 0), np.uint64)
    ctx.check(ctx.lib.fb_nthash_windows(ctx.h, _ptr(buf, C.c_uint8), buf.size, k, _ptr(out, C.c_uint64)),
              "fb_nthash_windows")
    return out


def hll_estimate_registers(regs: np.ndarray, p: int) -> float:
    r = np.ascontiguousarray(regs, dtype=np.uint8)
    return float(_lib.load().fb_hll_estimate_registers(_ptr(r, C.c_uint8), p))



# pass-2 routes (fb_solid_stats[4]): the abandoned LDS routes finish on the global table
ROUTES = ("global", "lds", "lds_region_full", "lds_spill_full", "lds_output_full")


@dataclass
class SolidSet:
    """Solid k-mers {K -> (cov, mask)} of one pass, sorted by K.

    keys: (n, words) uint64, words most significant first (1: u64, 2: u128
    {hi, lo}, 4: [u64;4]) as the reference's KmerInt types store them."""

    k: int
    keys: np.ndarray
    cov: np.ndarray
    mask: np.ndarray
    cms_table: Optional[np.ndarray] = None
    hll_registers: Optional[np.ndarray] = None
    timings_ms: tuple = ()

    def __len__(self) -> int:
        return int(self.cov.size)


class SolidHandle(_Handle):
    """Device-resident result of one pass (CMS, HLL, solid set)."""

    _destroy = "fb_solid_destroy"

    def count(self) -> int:
        return int(self.ctx.lib.fb_solid_count(self.h))

    def key_words(self) -> int:
        return int(self.ctx.lib.fb_solid_key_words(self.h))

    def timings(self) -> tuple:
        t = (C.c_float * 4)()
        self.ctx.check(self.ctx.lib.fb_solid_timings(self.h, t), "fb_solid_timings")
        return tuple(float(x) for x in t)

    def kernel_timings(self) -> dict:
        """Milliseconds of the LDS-binned route's kernels (0 for another route)."""
        t = (C.c_float * 5)()
        self.ctx.check(self.ctx.lib.fb_solid_kernel_timings(self.h, t), "fb_solid_kernel_timings")
        names = ("bin_scatter", "bin_count", "p2_gate_count", "p2_scatter", "p2_bins")
        return {n: float(x) for n, x in zip(names, t)}

    def stats(self) -> dict:
        t = (C.c_uint64 * 11)()
        self.ctx.check(self.ctx.lib.fb_solid_stats(self.h, t), "fb_solid_stats")
        return {"windows": int(t[0]), "records": int(t[1]), "keys": int(t[2]), "slots": int(t[3]),
                "pass2_route": ROUTES[int(t[4])],
                "lds_bins": int(t[5]), "lds_bin_keys": int(t[6]), "est_keys": int(t[7]),
                "pass2_records": int(t[8]), "reads_used": int(t[9]), "pass2_rounds": int(t[10])}

    def pass2_gate(self) -> str:
        """'window': the Count-Min gate applied per window pair; 'key': once
        per kept key and low edge (the ungated LDS route)."""
        r = int(self.ctx.lib.fb_solid_pass2_gate(self.h))
        if r < 0:
            self.ctx.check(r, "fb_solid_pass2_gate")
        return ("window", "key")[r]

    def pass2_layout(self) -> str:
        """'runs': a count walk gave exact record runs; 'chunks': records in
        chunked regions without a count walk; 'chunks_then_runs': a chunked
        region filled up and the pass ran again with exact runs."""
        r = int(self.ctx.lib.fb_solid_pass2_layout(self.h))
        if r < 0:
            self.ctx.check(r, "fb_solid_pass2_layout")
        return ("runs", "chunks", "chunks_then_runs")[r]

    def download(self, k: int, with_cms: bool = True, counter_bits: int = 4, hll_p: int = 0) -> SolidSet:
        n, nw = self.count(), self.key_words()
        keys = np.zeros((n, nw), np.uint64)
        cov = np.zeros(n, np.uint16)
        mask = np.zeros(n, np.uint8)
        self.ctx.check(self.ctx.lib.fb_solid_download(self.h, _ptr(keys, C.c_uint64), _ptr(cov, C.c_uint16),
                                                      _ptr(mask, C.c_uint8)), "fb_solid_download")
        cms = None
        if with_cms:
            nb = int(self.ctx.lib.fb_solid_cms_bytes(self.h))
            buf = np.zeros(nb, np.uint8)
            self.ctx.check(self.ctx.lib.fb_solid_cms_download(self.h, buf.ctypes.data, nb), "fb_solid_cms_download")
            cms = buf if counter_bits == 4 else buf.view(np.uint16)
        regs = None
        if hll_p:
            regs = np.zeros(1 << hll_p, np.uint8)
            self.ctx.check(self.ctx.lib.fb_solid_hll_registers(self.h, _ptr(regs, C.c_uint8)),
                           "fb_solid_hll_registers")
        return SolidSet(k, keys, cov, mask, cms, regs, self.timings())


def _pack_contigs(contigs: Optional[Sequence[bytes]]):
    if not contigs:
        return None, None, 0
    lens = _u32([len(c) for c in contigs])
    buf = np.frombuffer(b"".join(contigs), dtype=np.uint8).copy()
    return buf, lens, len(contigs)


def build_params(k: int, target_solid_cov: Optional[int] = None, genome_size: Optional[int] = None,
                 cms_width: Optional[int] = None, counter_bits: int = 4, hll_p: int = 0, sort: bool = True,
                 split: int = 0, placement: int = 0, bins2: int = 0, chunks1: int = 0, chunks2: int = 0) -> BuildParams:
    p = BuildParams()
    _lib.load().fb_build_params_default(C.byref(p), k)
    p.target_solid_cov = -1 if target_solid_cov is None else int(target_solid_cov)
    p.genome_size = 0 if genome_size is None else int(genome_size)
    p.cms_width = 0 if cms_width is None else int(cms_width)
    p.counter_bits = counter_bits
    p.hll_p = hll_p
    p.sort = 1 if sort else 0
    p.split = split
    p.placement = placement
    p.bins2 = bins2
    p.chunks1 = chunks1
    p.chunks2 = chunks2
    return p


def build_solid_handle(ctx: Context, bank: DeviceBank, k: int, contigs: Optional[Sequence[bytes]] = None,
                       **kw) -> SolidHandle:
    """Runs passes 1-3 on the device and keeps the result resident."""
    p = build_params(k, **kw)
    buf, lens, n = _pack_contigs(contigs)
    h = C.c_void_p()
    ctx.check(ctx.lib.fb_build_solid(ctx.h, bank.h, _ptr(buf, C.c_uint8), _ptr(lens, C.c_uint32), n,
                                     C.byref(p), C.byref(h)), "fb_build_solid")
    return SolidHandle(ctx, h)


def build_solid_kmers(ctx: Context, bank: DeviceBank, k: int, target_solid_cov: Optional[int] = None,
                      genome_size: Optional[int] = None, contigs: Optional[Sequence[bytes]] = None,
                      cms_width: Optional[int] = None, counter_bits: int = 4, hll_p: int = 0,
                      with_cms: bool = True, split: int = 0, placement: int = 0, bins2: int = 0,
                      chunks1: int = 0, chunks2: int = 0) -> SolidSet:
    """build_cdbg_with_contigs_ext passes 1-3 (build.rs:67-328): the solid set,
    the pass's CMS table and (hll_p > 0) the pass's HLL registers."""
    h = build_solid_handle(ctx, bank, k, contigs, target_solid_cov=target_solid_cov, genome_size=genome_size,
                           cms_width=cms_width, counter_bits=counter_bits, hll_p=hll_p, split=split,
                           placement=placement, bins2=bins2, chunks1=chunks1, chunks2=chunks2)
    try:
        return h.download(k, with_cms=with_cms, counter_bits=counter_bits, hll_p=hll_p)
    finally:
        h.close()


# ---------------------------------------------------------------------------
# Unitig compaction


def _graph_result(ctx: Context, h) -> dict:
    lib = ctx.lib
    try:
        nu = int(lib.fb_graph_unitig_count(h))
        nb = int(lib.fb_graph_base_count(h))
        ne = int(lib.fb_graph_edge_count(h))
        seq = np.zeros(max(nb, 1), np.uint8)
        off = np.zeros(max(nu, 1), np.uint64)
        ln = np.zeros(max(nu, 1), np.uint32)
        cov = np.zeros(max(nu, 1), np.float32)
        circ = np.zeros(max(nu, 1), np.uint8)
        ctx.check(lib.fb_graph_unitigs(h, _ptr(seq, C.c_uint8), _ptr(off, C.c_uint64), _ptr(ln, C.c_uint32),
                                       cov.ctypes.data_as(C.POINTER(C.c_float)), _ptr(circ, C.c_uint8)),
                  "fb_graph_unitigs")
        src = np.zeros(max(ne, 1), np.uint32)
        dst = np.zeros(max(ne, 1), np.uint32)
        ov = np.zeros(max(ne, 1), np.uint8)
        sup = np.zeros(max(ne, 1), np.uint16)
        ctx.check(lib.fb_graph_edges(h, _ptr(src, C.c_uint32), _ptr(dst, C.c_uint32), _ptr(ov, C.c_uint8),
                                     _ptr(sup, C.c_uint16)), "fb_graph_edges")
        t = (C.c_float * 3)()
        ctx.check(lib.fb_graph_timings(h, t), "fb_graph_timings")
        raw = seq[:nb].tobytes()
        return {"seqs": [raw[int(o):int(o) + int(n)] for o, n in zip(off[:nu], ln[:nu])], "cov": cov[:nu],
                "circular": circ[:nu], "src": src[:ne], "dst": dst[:ne], "ov": ov[:ne], "support": sup[:ne],
                "timings_ms": tuple(float(x) for x in t)}
    finally:
        lib.fb_graph_destroy(h)


def compact_unitigs(ctx: Context, solid: SolidHandle) -> dict:
    """Unitig compaction (build.rs:330-597) of a sorted build's solid set on
    the device: seqs, cov (float32), circular, and the edges (src, dst, ov,
    support) in each node's insertion order."""
    h = C.c_void_p()
    ctx.check(ctx.lib.fb_compact_solid(solid.h, C.byref(h)), "fb_compact_solid")
    return _graph_result(ctx, h)


def compact_arrays(ctx: Context, keys, cov, mask, k: int) -> dict:
    """The same compaction of a solid set given as host arrays (keys
    strictly ascending, fb_solid_download layout)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    cov = np.ascontiguousarray(cov, dtype=np.uint16)
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    h = C.c_void_p()
    ctx.check(ctx.lib.fb_compact_arrays(ctx.h, _ptr(keys, C.c_uint64), _ptr(cov, C.c_uint16), _ptr(mask, C.c_uint8),
                                        int(cov.size), k, C.byref(h)), "fb_compact_arrays")
    return _graph_result(ctx, h)


# ---------------------------------------------------------------------------
# Multi-k flow: filter_active_reads and upgrade_multik_ext (local_multik.rs)


class _GraphArrays:
    """fb_graph_view over a graph dict (seqs, circular, src; optional cov,
    deleted), keeping the arrays alive."""

    def __init__(self, graph: dict, k: int):
        seqs = [bytes(x) for x in graph["seqs"]]
        n = len(seqs)
        self.lens = np.array([len(x) for x in seqs], np.uint32)
        self.offs = np.zeros(n, np.uint64)
        if n:
            self.offs[1:] = np.cumsum(self.lens[:-1], dtype=np.uint64)
        self.seq = np.frombuffer(b"".join(seqs), np.uint8).copy()
        self.flags = (np.asarray(graph.get("circular", np.zeros(n)), np.uint8) << 1).astype(np.uint8)
        if "deleted" in graph:
            self.flags |= np.asarray(graph["deleted"], np.uint8)
        src = np.asarray(graph.get("src", np.zeros(0)), np.int64)
        self.nf = np.bincount(src[(src & 1) == 0] >> 1, minlength=n).astype(np.uint32)
        self.nr = np.bincount(src[(src & 1) == 1] >> 1, minlength=n).astype(np.uint32)
        self.cov = np.asarray(graph.get("cov", np.zeros(n)), np.float32).copy()
        self.view = _lib.GraphView(k, n, _ptr(self.seq, C.c_uint8), _ptr(self.offs, C.c_uint64),
                                   _ptr(self.lens, C.c_uint32),
                                   self.cov.ctypes.data_as(C.POINTER(C.c_float)) if n else None,
                                   _ptr(self.flags, C.c_uint8), _ptr(self.nf, C.c_uint32), _ptr(self.nr, C.c_uint32))


def filter_active_reads(ctx: Context, bank: DeviceBank, graph: dict, k: int) -> tuple:
    """filter_active_reads (local_multik.rs:259-333) of a device bank against a
    graph dict at graph k -> (DeviceBank of the kept reads, keep_merged,
    keep_pairs)."""
    ga = _GraphArrays(graph, k)
    nm, npair = _arena_reads(ctx, bank)
    km = np.zeros(max(nm, 1), np.uint8)
    kp = np.zeros(max(npair, 1), np.uint8)
    h = C.c_void_p()
    ctx.check(ctx.lib.fb_filter_active_reads(ctx.h, bank.h, C.byref(ga.view), C.byref(h), _ptr(km, C.c_uint8),
                                             _ptr(kp, C.c_uint8)), "fb_filter_active_reads")
    return DeviceBank._adopt(ctx, h), km[:nm], kp[:npair]


def _arena_reads(ctx: Context, bank: DeviceBank) -> tuple:
    n = []
    for a in range(3):
        nr = C.c_uint64()
        ctx.check(ctx.lib.fb_bank_arena_info(bank.h, a, C.byref(nr), None, None), "fb_bank_arena_info")
        n.append(int(nr.value))
    return n[0], min(n[1], n[2])


MULTIK_EARLY_STOP = {0: None, 1: "K2_LE_PREV_K", 2: "NO_BRANCHING_REGIONS"}
GENOME_SIZE_SOURCE = {0: "SKIPPED", 1: "EXPLICIT_PASSED", 2: "FALLBACK_READ_BASES_DIV_60"}


def upgrade_multik(ctx: Context, bank: DeviceBank, graph: dict, k: int, k2: int, genome_size: Optional[int] = None,
                   keep_solid: bool = False, keep_active: bool = False, **params) -> dict:
    """upgrade_multik_ext (local_multik.rs:38-147) without simplify_graph:
    result fields of MultiKUpgradeResult, the new graph dict ("graph", None on
    an early stop), and optionally the k2 solid set ("solid", a SolidHandle)
    and the filtered bank ("active", a DeviceBank).  params: counter_bits,
    cms_width, hll_p and launch-shape fields of build_params."""
    ga = _GraphArrays(graph, k)
    p = build_params(k2, **params) if params else None
    res = _lib.MultiKResult()
    hg, hs, hb = C.c_void_p(), C.c_void_p(), C.c_void_p()
    ctx.check(ctx.lib.fb_upgrade_multik(ctx.h, bank.h, C.byref(ga.view), k2, genome_size or 0,
                                        C.byref(p) if p is not None else None, C.byref(res), C.byref(hg),
                                        C.byref(hs) if keep_solid else None, C.byref(hb) if keep_active else None),
              "fb_upgrade_multik")
    out = {f: getattr(res, f) for f, _ in _lib.MultiKResult._fields_ if f != "reserved"}
    out["executed"] = bool(out["executed"])
    out["early_stop_reason"] = MULTIK_EARLY_STOP[out.pop("early_stop")]
    out["genome_size_source"] = GENOME_SIZE_SOURCE[out["genome_size_source"]]
    out["graph"] = _graph_result(ctx, hg) if hg.value else None
    if keep_solid:
        out["solid"] = SolidHandle(ctx, hs) if hs.value else None
    if keep_active:
        out["active"] = DeviceBank._adopt(ctx, hb) if hb.value else None
    return out


# ---------------------------------------------------------------------------
# PE mapper (mapper.rs:89-279)

LINK_DTYPE = np.dtype([("u1", np.uint32), ("u2", np.uint32), ("pair_id", np.uint32), ("dist_u1_exit", np.uint32),
                       ("dist_u2_entry", np.uint32), ("r1_len", np.uint16), ("r2_len", np.uint16)])


def map_paired_reads(ctx: Context, bank: DeviceBank, graph: dict, k: int) -> dict:
    """map_paired_reads<K> on the device: {"links": evidence records in pair
    order (LINK_DTYPE), "sizes": intra-unitig insert sizes in pair order,
    "dist": InsertSizeDist fields}."""
    ga = _GraphArrays(graph, k)
    h = C.c_void_p()
    ctx.check(ctx.lib.fb_map_paired_reads(ctx.h, C.byref(ga.view), bank.h, k, C.byref(h)), "fb_map_paired_reads")
    try:
        n = int(ctx.lib.fb_pair_map_link_count(h))
        links = np.zeros(max(n, 1), LINK_DTYPE)
        ctx.check(ctx.lib.fb_pair_map_links(h, links.ctypes.data_as(C.POINTER(_lib.PairLink))), "fb_pair_map_links")
        d = _lib.InsertSizeDist()
        ctx.check(ctx.lib.fb_pair_map_dist(h, C.byref(d)), "fb_pair_map_dist")
        sizes = np.zeros(max(int(d.n_sizes), 1), np.uint32)
        ctx.check(ctx.lib.fb_pair_map_sizes(h, _ptr(sizes, C.c_uint32)), "fb_pair_map_sizes")
        return {"links": links[:n], "sizes": sizes[:int(d.n_sizes)],
                "dist": {"median": d.median, "mad": d.mad, "min_insert": int(d.min_insert),
                         "max_insert": int(d.max_insert)}}
    finally:
        ctx.lib.fb_pair_map_destroy(h)


def polish_contigs(ctx: Context, bank: DeviceBank, graph: dict, min_output_len: int) -> tuple:
    """polish_contigs (consensus.rs:64-271) on the device with the bank's
    merged and pe_forward reads -> (every node's sequence after polishing,
    PolishStats dict)."""
    ga = _GraphArrays(graph, 31)
    total = int(ga.lens.sum(dtype=np.uint64))
    out = np.zeros(max(total, 1), np.uint8)
    st = _lib.PolishStats()
    ctx.check(ctx.lib.fb_polish_contigs(ctx.h, C.byref(ga.view), bank.h, min_output_len, _ptr(out, C.c_uint8),
                                        C.byref(st)), "fb_polish_contigs")
    raw = out[:total].tobytes()
    offs = np.concatenate([[0], np.cumsum(ga.lens, dtype=np.uint64)]).astype(np.int64)
    return [raw[offs[i]:offs[i + 1]] for i in range(len(ga.lens))], {f: int(getattr(st, f)) for f, _ in st._fields_}


# ---------------------------------------------------------------------------
# Ingest front end (trim_merge.rs:17-417)


def read_fastx(path) -> list:
    """open_fastx + the record reader: [(seq bytes, qual bytes), ...] (qual
    b"" for FASTA); gzip when the name ends in ".gz"."""
    lib = _lib.load()
    h = C.c_void_p()
    rc = lib.fb_fastx_read(str(path).encode(), C.byref(h))
    if rc:
        raise FastBactError(rc, lib.fb_fastx_last_error().decode(errors="replace"))
    try:
        v = _lib.RecordsView()
        lib.fb_fastx_view(h, C.byref(v))
        n = int(v.n)
        if n == 0:
            return []
        so = np.ctypeslib.as_array(v.seq_offsets, shape=(n,))
        sl = np.ctypeslib.as_array(v.seq_lengths, shape=(n,))
        qo = np.ctypeslib.as_array(v.qual_offsets, shape=(n,))
        ql = np.ctypeslib.as_array(v.qual_lengths, shape=(n,))
        sb = int((so + sl).max())
        qb = int((qo + ql).max())
        seq = C.string_at(v.seq, sb) if sb else b""
        qual = C.string_at(v.qual, qb) if qb else b""
        return [(seq[int(o):int(o) + int(a)], qual[int(p):int(p) + int(b)]) for o, a, p, b in zip(so, sl, qo, ql)]
    finally:
        lib.fb_fastx_destroy(h)


def ingest_params(subsample_factor: float = 1.0, min_len: int = 30, min_qual: int = 20, trim: bool = True,
                  merge: bool = True, seed: int = 0) -> _lib.IngestParams:
    p = _lib.IngestParams()
    _lib.load().fb_ingest_params_default(C.byref(p))
    p.subsample_factor, p.min_len, p.min_qual = float(subsample_factor), int(min_len), int(min_qual)
    p.enable_trim, p.enable_merge, p.seed = int(bool(trim)), int(bool(merge)), int(seed)
    return p


class _Records:
    """fb_records_view over [(seq, qual), ...], keeping the arrays alive."""

    def __init__(self, recs):
        seqs = [bytes(s) for s, _ in recs]
        quals = [bytes(q) for _, q in recs]
        self.sl = np.array([len(x) for x in seqs], np.uint32)
        self.ql = np.array([len(x) for x in quals], np.uint32)
        self.so = np.zeros(len(seqs), np.uint64)
        self.qo = np.zeros(len(seqs), np.uint64)
        if len(seqs):
            self.so[1:] = np.cumsum(self.sl[:-1], dtype=np.uint64)
            self.qo[1:] = np.cumsum(self.ql[:-1], dtype=np.uint64)
        self.s = np.frombuffer(b"".join(seqs) or b"\0", np.uint8).copy()
        self.q = np.frombuffer(b"".join(quals) or b"\0", np.uint8).copy()
        self.view = _lib.RecordsView(_ptr(self.s, C.c_uint8), _ptr(self.so, C.c_uint64), _ptr(self.sl, C.c_uint32),
                                     _ptr(self.q, C.c_uint8), _ptr(self.qo, C.c_uint64), _ptr(self.ql, C.c_uint32),
                                     len(seqs))


class _FixedRecords:
    """fb_records_view over equal-length records given as 2-D uint8 arrays."""

    def __init__(self, seq2d: np.ndarray, qual2d: np.ndarray):
        self.s = np.ascontiguousarray(seq2d, np.uint8)
        self.q = np.ascontiguousarray(qual2d, np.uint8)
        n, L = self.s.shape
        self.so = np.arange(n, dtype=np.uint64) * np.uint64(L)
        self.sl = np.full(n, L, np.uint32)
        self.qo = np.arange(n, dtype=np.uint64) * np.uint64(self.q.shape[1])
        self.ql = np.full(n, self.q.shape[1], np.uint32)
        self.view = _lib.RecordsView(_ptr(self.s, C.c_uint8), _ptr(self.so, C.c_uint64), _ptr(self.sl, C.c_uint32),
                                     _ptr(self.q, C.c_uint8), _ptr(self.qo, C.c_uint64), _ptr(self.ql, C.c_uint32), n)


def ingest_arrays(ctx: Context, r1, q1, r2, q2, **params) -> tuple:
    """ingest_records for equal-length mates given as (pairs, length) uint8
    arrays of bases and qualities -> (DeviceBank, stats dict)."""
    a, b = _FixedRecords(r1, q1), _FixedRecords(r2, q2)
    p = ingest_params(**params)
    st = _lib.IngestStats()
    h = C.c_void_p()
    ctx.check(ctx.lib.fb_ingest_records(ctx.h, C.byref(a.view), C.byref(b.view), C.byref(p), C.byref(h),
                                        C.byref(st)), "fb_ingest_records")
    return DeviceBank._adopt(ctx, h), _ingest_stats(st)


def _ingest_stats(st) -> dict:
    return {f: (float(getattr(st, f)) if f.endswith("_ms") else int(getattr(st, f))) for f, _ in st._fields_}


def ingest_records(ctx: Context, r1, r2, **params) -> tuple:
    """ingest_reads_ext_with_seed's record loop on the device over parsed
    mates [(seq, qual), ...] -> (DeviceBank, stats dict)."""
    a, b = _Records(r1), _Records(r2)
    p = ingest_params(**params)
    st = _lib.IngestStats()
    h = C.c_void_p()
    ctx.check(ctx.lib.fb_ingest_records(ctx.h, C.byref(a.view), C.byref(b.view), C.byref(p), C.byref(h),
                                        C.byref(st)), "fb_ingest_records")
    return DeviceBank._adopt(ctx, h), _ingest_stats(st)


def ingest_files(ctx: Context, r1_path, r2_path, **params) -> tuple:
    """ingest_reads_ext_with_seed: FASTQ/FASTA (.gz) mates to a device bank
    -> (DeviceBank, stats dict with parse_ms and device_ms)."""
    p = ingest_params(**params)
    st = _lib.IngestStats()
    h = C.c_void_p()
    ctx.check(ctx.lib.fb_ingest_files(ctx.h, str(r1_path).encode(), str(r2_path).encode(), C.byref(p), C.byref(h),
                                      C.byref(st)), "fb_ingest_files")
    return DeviceBank._adopt(ctx, h), _ingest_stats(st)


# ---------------------------------------------------------------------------
# Batched form


class Batch(_Handle):
    """Many isolates per launch (include/fastbact_gpu.h, fb_batch_*): one
    launch per kernel covers every isolate and each pass replays a HIP graph."""

    _destroy = "fb_batch_destroy"

    def __init__(self, ctx: Context, banks: Sequence[DeviceBank], genome_sizes: Sequence[int],
                 cms_width: Optional[int] = None, counter_bits: int = 4, hll_p: int = 0, sort: bool = True,
                 split: int = 0, placement: int = 0, bins2: int = 0, chunks1: int = 0, chunks2: int = 0):
        p = build_params(0, cms_width=cms_width, counter_bits=counter_bits, hll_p=hll_p, sort=sort, split=split,
                         placement=placement, bins2=bins2, chunks1=chunks1, chunks2=chunks2)
        arr = (C.c_void_p * len(banks))(*[b.h.value if hasattr(b.h, "value") else b.h for b in banks])
        gs = _u64(genome_sizes)
        h = C.c_void_p()
        ctx.check(ctx.lib.fb_batch_create(ctx.h, arr, _ptr(gs, C.c_uint64), len(banks), C.byref(p), C.byref(h)),
                  "fb_batch_create")
        super().__init__(ctx, h)
        self.banks = list(banks)  # keep the device banks alive
        self.counter_bits, self.hll_p = counter_bits, hll_p

    def __len__(self) -> int:
        return int(self.ctx.lib.fb_batch_size(self.h))

    def run(self, k: int, target_solid_cov: Optional[int] = None) -> None:
        self.ctx.check(self.ctx.lib.fb_batch_run(self.h, k, -1 if target_solid_cov is None else target_solid_cov),
                       "fb_batch_run")

    def timings(self) -> tuple:
        t = (C.c_float * 4)()
        self.ctx.check(self.ctx.lib.fb_batch_timings(self.h, t), "fb_batch_timings")
        return tuple(float(x) for x in t)

    def autotune(self, k: int, target_solid_cov: Optional[int] = None) -> dict:
        """fb_batch_autotune: times launch shapes at k on this batch and keeps the fastest."""
        r = _lib.TuneResult()
        self.ctx.check(self.ctx.lib.fb_batch_autotune(self.h, k, -1 if target_solid_cov is None else target_solid_cov,
                                                      C.byref(r)), "fb_batch_autotune")
        return {name: getattr(r, name) for name, _ in _lib.TuneResult._fields_}

    def routes(self) -> list:
        """Pass-2 route of every isolate in the last run."""
        r = np.zeros(len(self), np.uint32)
        self.ctx.check(self.ctx.lib.fb_batch_routes(self.h, _ptr(r, C.c_uint32)), "fb_batch_routes")
        return [ROUTES[int(x)] for x in r]

    def result(self, i: int, k: int, with_cms: bool = True) -> SolidSet:
        n = int(self.ctx.lib.fb_batch_count(self.h, i))
        nw = int(self.ctx.lib.fb_batch_key_words(self.h))
        keys = np.zeros((n, nw), np.uint64)
        cov = np.zeros(n, np.uint16)
        mask = np.zeros(n, np.uint8)
        self.ctx.check(self.ctx.lib.fb_batch_download(self.h, i, _ptr(keys, C.c_uint64), _ptr(cov, C.c_uint16),
                                                      _ptr(mask, C.c_uint8)), "fb_batch_download")
        cms = regs = None
        if with_cms:
            nb = int(self.ctx.lib.fb_batch_cms_bytes(self.h, i))
            buf = np.zeros(nb, np.uint8)
            self.ctx.check(self.ctx.lib.fb_batch_cms_download(self.h, i, buf.ctypes.data, nb), "fb_batch_cms_download")
            cms = buf if self.counter_bits == 4 else buf.view(np.uint16)
        if self.hll_p:
            regs = np.zeros(1 << self.hll_p, np.uint8)
            self.ctx.check(self.ctx.lib.fb_batch_hll_registers(self.h, i, _ptr(regs, C.c_uint8)),
                           "fb_batch_hll_registers")
        return SolidSet(k, keys, cov, mask, cms, regs, self.timings())


def batch_footprint(genome_size: int, bases: int, reads: int, k: int, **kw) -> int:
    p = build_params(k, **kw)
    return int(_lib.load().fb_batch_footprint(genome_size, bases, reads, C.byref(p), k))


# ---------------------------------------------------------------------------
# Estimator

CONFIDENCE = {0: "HIGH", 1: "LOW_COVERAGE_WARNING", 2: "UNRELIABLE"}


@dataclass
class GenomeEstimate:
    genome_size: int
    exact_read_pairs: int
    total_bases: int
    median_read_len: int
    max_read_len: int
    estimated_cov: float
    subsample_factor: float
    kmers: list
    confidence: str
    u_ge_1: int
    u_ge_2: int
    sample_kmer_depth: float
    pairs_sampled: int
    m_sample: int


def count_stats(len_r1, len_r2) -> CountStats:
    """count_fastx_exact's statistics from the read lengths of the pairs."""
    a, b = _u32(len_r1), _u32(len_r2)
    if a.size != b.size:
        raise FastBactError(_lib.FB_ERR_INVALID, "R1 and R2 hold different numbers of reads")
    st = CountStats()
    rc = _lib.load().fb_count_stats_from_lengths(_ptr(a, C.c_uint32), _ptr(b, C.c_uint32), a.size, C.byref(st))
    if rc:
        raise FastBactError(rc, "Input FASTQ files contain 0 valid read pairs" if rc == _lib.FB_ERR_EMPTY
                            else "fb_count_stats_from_lengths failed")
    return st


def compute_dynamic_kmers(median_len: int, target_max_k: Optional[int] = None) -> list:
    out = (C.c_uint32 * 8)()
    n = _lib.load().fb_compute_dynamic_kmers(median_len, -1 if target_max_k is None else target_max_k, out)
    return [int(out[i]) for i in range(n)]


def estimate_genome(ctx: Context, r1: BitSeqArena, r2: BitSeqArena, max_pairs: int, target_cov: float,
                    stats: Optional[CountStats] = None) -> GenomeEstimate:
    """estimate_genome (estimator.rs:286-440) over R1[i]/R2[i] pairs."""
    if stats is None:
        stats = count_stats(r1.lengths, r2.lengths)
    v1, v2 = r1.view(), r2.view()
    out = GenomeEstimateC()
    ctx.check(ctx.lib.fb_estimate_genome(ctx.h, C.byref(v1), C.byref(v2), C.byref(stats), max_pairs,
                                         float(target_cov), C.byref(out)), "fb_estimate_genome")
    return GenomeEstimate(out.genome_size, out.exact_read_pairs, out.total_bases, out.median_read_len,
                          out.max_read_len, out.estimated_cov, out.subsample_factor,
                          [int(out.kmers[i]) for i in range(out.n_kmers)], CONFIDENCE[out.confidence],
                          out.u_ge_1, out.u_ge_2, out.sample_kmer_depth, out.pairs_sampled, out.m_sample)


SUBSAMPLE_REASONS = ("genome_size", "estimate", "within_trigger", "estimate_not_high")


def auto_subsample_factor(est: GenomeEstimate, depth: float, threshold_ratio: float = 1.5,
                          genome_size: int = 0) -> tuple:
    """The pipeline's auto-subsampling policy (main.rs:66-185): (factor, reason)."""
    e = GenomeEstimateC()
    e.total_bases = est.total_bases
    e.estimated_cov = est.estimated_cov
    e.confidence = {v: k for k, v in CONFIDENCE.items()}[est.confidence]
    why = C.c_int32()
    f = _lib.load().fb_auto_subsample_factor(C.byref(e), float(depth), float(threshold_ratio), int(genome_size),
                                             C.byref(why))
    return float(f), SUBSAMPLE_REASONS[why.value]


def _estimate_result(out) -> GenomeEstimate:
    return GenomeEstimate(out.genome_size, out.exact_read_pairs, out.total_bases, out.median_read_len,
                          out.max_read_len, out.estimated_cov, out.subsample_factor,
                          [int(out.kmers[i]) for i in range(out.n_kmers)], CONFIDENCE[out.confidence],
                          out.u_ge_1, out.u_ge_2, out.sample_kmer_depth, out.pairs_sampled, out.m_sample)


def count_fastx(r1_path, r2_path) -> CountStats:
    """count_fastx_exact (estimator.rs:167-283) on two read files."""
    lib = _lib.load()
    st = CountStats()
    rc = lib.fb_count_fastx(str(r1_path).encode(), str(r2_path).encode(), C.byref(st))
    if rc:
        raise FastBactError(rc, lib.fb_fastx_last_error().decode(errors="replace"))
    return st


def estimate_genome_files(ctx: Context, r1_path, r2_path, max_pairs: int = 150_000,
                          target_cov: float = 100.0) -> GenomeEstimate:
    """estimate_genome (estimator.rs:286-440) on two FASTQ/FASTA (.gz) files."""
    out = GenomeEstimateC()
    ctx.check(ctx.lib.fb_estimate_genome_files(ctx.h, str(r1_path).encode(), str(r2_path).encode(), max_pairs,
                                               float(target_cov), C.byref(out)), "fb_estimate_genome_files")
    return _estimate_result(out)


# ---------------------------------------------------------------------------
# Synthetic isolates


def synth_params(**kw) -> SynthParams:
    p = SynthParams()
    _lib.load().fb_synth_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


@dataclass
class SynthIsolate:
    genome: np.ndarray      # ASCII uint8
    r1: np.ndarray          # (n_pairs, read_len) ASCII
    r2: np.ndarray

    def bank(self) -> ReadBank:
        """Unmerged pairs: R1 -> pe_forward, R2 -> pe_reverse."""
        n, L = self.r1.shape
        lens = np.full(n, L, np.uint32)
        return ReadBank(BitSeqArena(), BitSeqArena.from_concat(self.r1.reshape(-1), lens),
                        BitSeqArena.from_concat(self.r2.reshape(-1), lens))


def synth_isolate(n_pairs: Optional[int] = None, **kw) -> SynthIsolate:
    lib = _lib.load()
    p = synth_params(**kw)
    g = np.zeros(p.genome_len, np.uint8)
    rc = lib.fb_synth_genome(C.byref(p), _ptr(g, C.c_uint8))
    if rc:
        raise FastBactError(rc, "fb_synth_genome failed")
    n = int(lib.fb_synth_n_pairs(C.byref(p))) if n_pairs is None else int(n_pairs)
    r1 = np.zeros((n, p.read_len), np.uint8)
    r2 = np.zeros((n, p.read_len), np.uint8)
    rc = lib.fb_synth_reads(C.byref(p), _ptr(g, C.c_uint8), g.size, n, _ptr(r1, C.c_uint8), _ptr(r2, C.c_uint8))
    if rc:
        raise FastBactError(rc, "fb_synth_reads failed")
    return SynthIsolate(g, r1, r2)
