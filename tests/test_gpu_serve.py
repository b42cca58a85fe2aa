"""GPU tests of the serving form: a request's result is the same alone and in
a mix of overlapping requests, and equals the plain path and the oracle."""
import time

import numpy as np
import pytest

import fastbact_amd as F
from tests.helpers import isolate, oracle_arenas

pytestmark = pytest.mark.gpu

KS, TARGETS = (21, 55, 77), (None, 4, 4)


@pytest.fixture(scope="module")
def requests_():
    # the first two are the largest: they are still on the device when the rest arrive
    shapes = [(90_000, 14_000, 0.38), (60_000, 9000, 0.45), (25_000, 2500, 0.62), (40_000, 4000, 0.5),
              (15_000, 1200, 0.55)]
    return [isolate(seed=500 + i, genome_len=g, n_pairs=n, gc=gc).bank() for i, (g, n, gc) in enumerate(shapes)]


def _same(a: F.ServeOutcome, b: F.ServeOutcome):
    assert a.genome_size == b.genome_size
    assert a.estimate == b.estimate
    assert len(a.passes) == len(b.passes) == len(KS)
    for pa, pb in zip(a.passes, b.passes):
        assert np.array_equal(pa.keys, pb.keys)
        assert np.array_equal(pa.cov, pb.cov) and np.array_equal(pa.mask, pb.mask)
        assert np.array_equal(pa.cms_table, pb.cms_table)
        assert np.array_equal(pa.hll_registers, pb.hll_registers)


def test_serve_result_same_alone_and_in_mix(requests_):
    kw = dict(cms_width=1 << 17, counter_bits=16, hll_p=12, keep_sketches=1)
    alone = []
    with F.Server(0, KS, TARGETS, **kw) as sv:
        for bank in requests_:
            t = sv.submit(bank.pe_forward, bank.pe_reverse)
            alone.append(sv.wait(t, timeout=60))
    mix = {}
    with F.Server(0, KS, TARGETS, **kw) as sv:
        first = [sv.submit(b.pe_forward, b.pe_reverse) for b in requests_[:2]]
        # the rest arrive while the first two are on the device
        deadline = time.time() + 30
        while sv.stats()["iterations"] == 0 and time.time() < deadline:
            time.sleep(0.001)
        later = [sv.submit(b.pe_forward, b.pe_reverse) for b in requests_[2:]]
        for i, t in enumerate(first + later):
            mix[i] = sv.wait(t, timeout=60)
        st = sv.stats()
    for i in range(len(requests_)):
        _same(alone[i], mix[i])
    assert st["completed"] == len(requests_) and st["failed"] == 0
    # requests shared iterations: some ran together, at different ladder positions
    assert st["max_batch"] >= 2 and st["max_groups"] >= 2, st
    outs = list(mix.values())
    assert any(o.t_admit < p.t_done and o.ticket != p.ticket and o.t_admit > p.t_admit for o in outs for p in outs)


def test_serve_equals_plain_path_and_oracle(ctx, orc, requests_):
    bank = requests_[0]
    with F.Server(0, KS, TARGETS, cms_width=1 << 17, counter_bits=16, hll_p=12, keep_sketches=1) as sv:
        out = sv.wait(sv.submit(bank.pe_forward, bank.pe_reverse), timeout=60)
    est = F.estimate_genome(ctx, bank.pe_forward, bank.pe_reverse, 250000, 100.0)
    assert out.estimate == est and out.genome_size == est.genome_size
    dev = F.DeviceBank(ctx, bank)
    for j, (k, t) in enumerate(zip(KS, TARGETS)):
        plain = F.build_solid_kmers(ctx, dev, k, target_solid_cov=t, genome_size=out.genome_size, cms_width=1 << 17,
                                    counter_bits=16, hll_p=12)
        got = out.passes[j]
        assert np.array_equal(got.keys, plain.keys) and np.array_equal(got.cov, plain.cov)
        assert np.array_equal(got.mask, plain.mask) and np.array_equal(got.cms_table, plain.cms_table)
        assert np.array_equal(got.hll_registers, plain.hll_registers)
    ref = orc.build_solid(oracle_arenas(bank), 55, target_solid_cov=4, width=1 << 17, counter_bits=16, hll_p=12)
    assert np.array_equal(out.passes[1].keys, ref["keys"]) and np.array_equal(out.passes[1].cov, ref["cov"])


def test_serve_sketch_sized_from_the_estimate(ctx, requests_):
    """cms_width 0: each request's sketch is CountMinSketch::for_genome_size
    of its estimated genome (build.rs:74-84)."""
    bank = requests_[3]
    with F.Server(0, (31,), (3,), cms_width=0, counter_bits=4, hll_p=0, keep_sketches=1) as sv:
        out = sv.wait(sv.submit(bank.pe_forward, bank.pe_reverse), timeout=60)
    w = F.CountMinSketch.width_for_genome_size(out.genome_size)
    assert out.passes[0].cms_table.size == w * 2
    plain = F.build_solid_kmers(ctx, F.DeviceBank(ctx, bank), 31, target_solid_cov=3, genome_size=out.genome_size,
                                counter_bits=4)
    assert np.array_equal(out.passes[0].keys, plain.keys) and np.array_equal(out.passes[0].cms_table, plain.cms_table)


def test_serve_budget_admits_one_at_a_time():
    """A budget that holds any one request but no two: the next is admitted
    only after the one before it has finished and left; results are the same
    as with room for all."""
    kw = dict(cms_width=1 << 17, counter_bits=16, hll_p=12)
    # equal-sized requests: any one fits the budget below, no two do
    banks = [isolate(seed=600 + i, genome_len=50_000, n_pairs=6000, gc=0.4 + 0.1 * i).bank() for i in range(3)]
    with F.Server(0, KS, TARGETS, **kw) as sv:
        free = [sv.wait(sv.submit(b.pe_forward, b.pe_reverse), timeout=60) for b in banks]
    fps = [o.footprint for o in free]
    budget = 2 * min(fps) - 1
    assert max(fps) <= budget  # each fits alone
    with F.Server(0, KS, TARGETS, hbm_budget=budget, **kw) as sv:
        ts = [sv.submit(b.pe_forward, b.pe_reverse) for b in banks]
        outs = [sv.wait(t, timeout=60) for t in ts]
        st = sv.stats()
    assert st["max_batch"] == 1, st
    for a, b in zip(outs, outs[1:]):
        assert b.t_admit >= a.t_done
    for a, b in zip(outs, free):
        for pa, pb in zip(a.passes, b.passes):
            assert np.array_equal(pa.keys, pb.keys) and np.array_equal(pa.cov, pb.cov)


def test_serve_errors():
    empty = F.BitSeqArena()
    with F.Server(0, (21,), (None,), cms_width=1 << 16) as sv:
        t = sv.submit(empty, empty)
        with pytest.raises(F.FastBactError) as e:
            sv.wait(t, timeout=60)
        assert e.value.code == F._lib.FB_ERR_EMPTY
        assert sv.stats()["failed"] == 1
    with F.Server(0, (21,), (None,), cms_width=1 << 16, hbm_budget=1 << 20) as sv:
        bank = isolate(seed=9, genome_len=20_000, n_pairs=1000).bank()
        with pytest.raises(F.FastBactError) as e:
            sv.submit(bank.pe_forward, bank.pe_reverse)
        assert e.value.code == F._lib.FB_ERR_OOM


def test_failed_pass_released_at_once(requests_):
    """A pass that fails fails every request of its group.  Here the server's
    context is capped (ctx_budget 8 MiB) below what pass 1 of these requests
    needs, while upload and sketches fit: each request ends with FB_ERR_OOM and
    is released the moment its wait returns (Server.wait releases on error),
    while the worker goes on with the others; the server then closes."""
    kw = dict(cms_width=1 << 17, counter_bits=16, hll_p=12, estimate=0, ctx_budget=8 << 20)
    with F.Server(0, KS, TARGETS, **kw) as sv:
        tickets = [sv.submit(b.pe_forward, b.pe_reverse) for b in requests_[:3]]
        for t in tickets:
            with pytest.raises(F.FastBactError) as e:
                sv.wait(t, timeout=60)
            assert e.value.code == F._lib.FB_ERR_OOM, e.value
            # released: the ticket is gone
            assert sv.lib.fb_server_release(sv.h, t) == F._lib.FB_ERR_INVALID
        st = sv.stats()
    assert st["failed"] == 3 and st["completed"] == 0, st


def test_destroy_wakes_waiters(requests_):
    """fb_server_destroy with a waiter blocked on a queued request: the
    waiter returns (FB_ERR_BUSY when the request was still queued, its
    result when it had been admitted) and destroy returns once it has."""
    import ctypes as C
    import threading

    kw = dict(cms_width=1 << 17, counter_bits=16, hll_p=12, max_inflight=1)
    sv = F.Server(0, KS, TARGETS, **kw)
    ts = [sv.submit(b.pe_forward, b.pe_reverse) for b in requests_[:3]]
    rcs = {}

    def waiter():
        res = F._lib.ServeResult()
        rcs["last"] = sv.lib.fb_server_wait(sv.h, ts[-1], -1.0, C.byref(res))

    th = threading.Thread(target=waiter)
    th.start()
    time.sleep(0.05)
    h, sv.h = sv.h, None
    sv.lib.fb_server_destroy(h)  # returns only after the waiter has
    th.join(timeout=60)
    assert not th.is_alive()
    assert rcs["last"] in (0, F._lib.FB_ERR_BUSY), rcs
