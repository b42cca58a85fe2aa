"""Host-only replay of the solid-output sizing across a request's k ladder.

A request of the served form (and an isolate of a batch) keeps one set of
solid output buffers for its passes at k 21, 55 and 77, while an output entry
widens from 1 to 2 and 4 key words.  These tests replay, with the library's
own pass-2 plan (fb_plan_pass2) and output guard (fb_solid_out_fits), the
sizing rule the batched form once used -- grow the buffers only when the
entry capacity grows -- and show that it left the key buffer 2-4x short at
k 55 and 77 for requests of the serving workload's sizes, so emission wrote
past it; and that sizing in bytes at each k, as the library now does, always
passes the guard the library checks before every such launch.
No device is used.
"""
import ctypes as C

import numpy as np
import pytest

from fastbact_amd import _lib

LADDER = (21, 55, 77)
READ_LEN = 150


def _nref(k):
    return 1 if k <= 31 else (2 if k <= 63 else 4)


def _plan(est_keys, windows, n_reads, k):
    p = _lib.Pass2Plan()
    assert _lib.load().fb_plan_pass2(float(est_keys), int(windows), int(n_reads), READ_LEN, k, C.byref(p)) == 0
    return p


def _need(cap, nref):
    out = np.zeros(3, np.uint64)
    _lib.load().fb_solid_out_bytes(int(cap), nref, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out


def _fits(cap, nref, have):
    have = np.asarray(have, np.uint64)
    return bool(_lib.load().fb_solid_out_fits(int(cap), nref, have.ctypes.data_as(C.POINTER(C.c_uint64))))


def _request(genome_mbp, cov):
    """(windows per k, reads) of a PE150 request of genome_mbp Mbp at cov x."""
    bases = int(genome_mbp * 1e6 * cov)
    reads = bases // READ_LEN
    return {k: reads * (READ_LEN - k + 1) for k in LADDER}, reads


# request sizes of the serving workload (102 to 933 Mbases: 1.5-7 Mbp at 30-200x)
SHAPES = [(1.5, 68), (3.2, 80), (4.7, 120), (6.9, 135), (5.0, 187)]


def _est_cases(genome_mbp, windows):
    g = genome_mbp * 1e6
    # keys pass 2 inserts: near the genome's distinct k-mers once errors are
    # gated out, and up to every window when the sketch is overloaded
    yield {k: 1.1 * g for k in LADDER}
    yield {k: 1.1 * g + 0.02 * windows[k] for k in LADDER}
    yield {k: float(windows[k]) for k in LADDER}


@pytest.mark.parametrize("shape", SHAPES)
def test_old_growth_by_entries_leaves_the_key_buffer_short(shape):
    windows, reads = _request(*shape)
    short = []
    for est in _est_cases(shape[0], windows):
        have = np.zeros(3, np.uint64)
        cap_used = 0
        for k in LADDER:
            q = _plan(est[k], windows[k], reads, k)
            assert q.ok
            if q.out_cap > cap_used:  # the old rule: regrow only for more entries
                have = _need(q.out_cap, _nref(k))
                cap_used = q.out_cap
            # the launch writes up to cap_used entries of this k's width
            if not _fits(cap_used, _nref(k), have):
                need = _need(cap_used, _nref(k))
                short.append((k, int(need[0]) / int(have[0])))
    ks = {k for k, _ in short}
    assert 55 in ks and 77 in ks, short
    assert max(r for _, r in short) >= 4.0 - 1e-9


@pytest.mark.parametrize("shape", SHAPES)
def test_growth_in_bytes_always_fits(shape):
    windows, reads = _request(*shape)
    for est in _est_cases(shape[0], windows):
        have = np.zeros(3, np.uint64)
        for k in LADDER:
            q = _plan(est[k], windows[k], reads, k)
            have = np.maximum(have, _need(q.out_cap, _nref(k)))  # grown, never shrunk
            assert _fits(q.out_cap, _nref(k), have)


def test_guard_arithmetic():
    assert list(_need(1000, 4)) == [32000, 2000, 1000]
    assert _fits(1000, 4, [32000, 2000, 1000])
    assert not _fits(1000, 4, [31999, 2000, 1000])
    assert not _fits(1000, 4, [32000, 1999, 1000])
    assert not _fits(1000, 4, [32000, 2000, 999])
    assert not _fits(1000, 2, [8000, 2000, 1000])  # k 21 sizing at k 55


def test_plan_matches_footprint_bounds():
    # the plan's output capacity never exceeds what the admission footprint reserves
    windows, reads = _request(4.7, 120)
    for k in LADDER:
        q = _plan(windows[k], windows[k], reads, k)
        bp = _lib.BuildParams()
        _lib.load().fb_build_params_default(C.byref(bp), k)
        bp.counter_bits = 16
        bp.hll_p = 14
        fp = _lib.load().fb_batch_footprint(int(4.7e6), int(4.7e6 * 120), reads, C.byref(bp), k)
        assert q.out_cap * (_nref(k) * 8 + 3) < fp
