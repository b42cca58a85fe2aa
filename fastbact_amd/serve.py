"""Serving form over the C ABI (include/fastbact_gpu.h, fb_server_* and
fb_admission_*).

A Server runs a request loop with continuous batching in a native worker
thread: requests (one isolate's read pairs each) are admitted by device
footprint, every iteration runs one batched pass per k-ladder position that
in-flight requests are at, and finished requests leave at once.  Each request
follows the reference flow of one isolate: estimate_genome
(src/ingest/estimator.rs:286-440), then passes 1-3 of
build_cdbg_with_contigs_ext (src/graph/build.rs:67-328) at every k.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import ServeParams, ServeResult, ServeStats
from .api import CONFIDENCE, ROUTES, BitSeqArena, FastBactError, GenomeEstimate, SolidSet, _ptr


def serve_params(ks: Sequence[int] = (21, 55, 77), targets: Sequence[Optional[int]] = (None, 4, 4),
                 **kw) -> ServeParams:
    if len(ks) != len(targets) or not 1 <= len(ks) <= 4:
        raise ValueError("one target per k, 1..4 passes")
    p = ServeParams()
    _lib.load().fb_serve_params_default(C.byref(p))
    p.n_passes = len(ks)
    for j, (k, t) in enumerate(zip(ks, targets)):
        p.k[j] = k
        p.target_solid_cov[j] = -1 if t is None else t
    for name, v in kw.items():
        if not hasattr(p, name):
            raise KeyError(name)
        setattr(p, name, v)
    return p


@dataclass
class ServeOutcome:
    ticket: int
    genome_size: int
    estimate: Optional[GenomeEstimate]
    passes: list          # SolidSet per pass (cms_table / hll_registers with keep_sketches)
    routes: list
    footprint: int
    t_submit: float
    t_admit: float
    t_done: float

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


class Server:
    """fb_server: submit() may be called while earlier requests run."""

    def __init__(self, device: int = 0, ks=(21, 55, 77), targets=(None, 4, 4), **kw):
        self.lib = _lib.load()
        self.params = serve_params(ks, targets, **kw)
        self.ks = list(ks)
        h = C.c_void_p()
        rc = self.lib.fb_server_create(device, C.byref(self.params), C.byref(h))
        if rc:
            raise FastBactError(rc, "fb_server_create failed")
        self.h = h

    def _check(self, rc: int, what: str) -> None:
        if rc:
            msg = self.lib.fb_server_last_error(self.h).decode(errors="replace")
            raise FastBactError(rc, f"{what}: {msg}")

    def submit(self, r1: BitSeqArena, r2: BitSeqArena, genome_size: int = 0) -> int:
        v1, v2 = r1.view(), r2.view()
        t = C.c_uint64()
        self._check(self.lib.fb_server_submit(self.h, C.byref(v1), C.byref(v2), genome_size, C.byref(t)),
                    "fb_server_submit")
        return int(t.value)

    def wait(self, ticket: int, timeout: float = -1.0, download: bool = True, release: bool = True) -> ServeOutcome:
        res = ServeResult()
        rc = self.lib.fb_server_wait(self.h, ticket, float(timeout), C.byref(res))
        if rc == _lib.FB_ERR_TIMEOUT:
            raise TimeoutError(f"request {ticket} not finished after {timeout} s")
        if rc:
            if release:
                self.lib.fb_server_release(self.h, ticket)
            raise FastBactError(rc, f"request {ticket}: {res.error.decode(errors='replace')}")
        est = None
        if res.has_estimate:
            e = res.estimate
            est = GenomeEstimate(e.genome_size, e.exact_read_pairs, e.total_bases, e.median_read_len, e.max_read_len,
                                 e.estimated_cov, e.subsample_factor, [int(e.kmers[i]) for i in range(e.n_kmers)],
                                 CONFIDENCE[e.confidence], e.u_ge_1, e.u_ge_2, e.sample_kmer_depth,
                                 e.pairs_sampled, e.m_sample)
        passes = []
        if download:
            for j in range(res.n_passes):
                k = self.ks[j]
                nw = 1 if k <= 31 else (2 if k <= 63 else 4)
                n = int(res.n_solid[j])
                keys = np.zeros((n, nw), np.uint64)
                cov = np.zeros(n, np.uint16)
                mask = np.zeros(n, np.uint8)
                self._check(self.lib.fb_server_download(self.h, ticket, j, _ptr(keys, C.c_uint64),
                                                        _ptr(cov, C.c_uint16), _ptr(mask, C.c_uint8)),
                            "fb_server_download")
                cms = regs = None
                if self.params.keep_sketches:
                    w = self.params.cms_width or _lib.load().fb_cms_width_for_genome_size(max(res.genome_size, 50000))
                    nb = w * 2 if self.params.counter_bits == 4 else w * 8
                    buf = np.zeros(nb, np.uint8)
                    regs = np.zeros(1 << self.params.hll_p, np.uint8) if self.params.hll_p else None
                    self._check(self.lib.fb_server_sketches(self.h, ticket, j, buf.ctypes.data, nb,
                                                            _ptr(regs, C.c_uint8) if regs is not None else None),
                                "fb_server_sketches")
                    cms = buf if self.params.counter_bits == 4 else buf.view(np.uint16)
                passes.append(SolidSet(k, keys, cov, mask, cms, regs, ()))
        out = ServeOutcome(ticket, int(res.genome_size), est, passes, [ROUTES[int(res.route[j])]
                                                                       for j in range(res.n_passes)],
                           int(res.footprint), res.t_submit, res.t_admit, res.t_done)
        if release:
            self.lib.fb_server_release(self.h, ticket)
        return out

    def stats(self) -> dict:
        st = ServeStats()
        self._check(self.lib.fb_server_stats(self.h, C.byref(st)), "fb_server_stats")
        return {name: getattr(st, name) for name, _ in ServeStats._fields_}

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.fb_server_destroy(self.h)
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


class Admission:
    """fb_admission: the server's admission policy alone (host only)."""

    def __init__(self, budget: int, max_inflight: int = 64):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.fb_admission_create(budget, max_inflight, C.byref(h))
        if rc:
            raise FastBactError(rc, "fb_admission_create failed")
        self.h = h

    def submit(self, rid: int, nbytes: int) -> int:
        return int(self.lib.fb_admission_submit(self.h, rid, nbytes))

    def admit(self) -> list:
        ids = np.zeros(1024, np.uint64)
        n = int(self.lib.fb_admission_admit(self.h, _ptr(ids, C.c_uint64), ids.size))
        return [int(x) for x in ids[:n]]

    def retire(self, rid: int) -> int:
        return int(self.lib.fb_admission_retire(self.h, rid))

    @property
    def used(self) -> int:
        return int(self.lib.fb_admission_used(self.h))

    @property
    def queued(self) -> int:
        return int(self.lib.fb_admission_queued(self.h))

    @property
    def inflight(self) -> int:
        return int(self.lib.fb_admission_inflight(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.fb_admission_destroy(self.h)
            self.h = None
