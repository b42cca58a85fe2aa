"""GPU checks at the sizes of the batched and served workloads.

Batched workload: 64 isolates, genome U[1.5, 7] Mbp, GC U[0.30, 0.68],
coverage lognormal (median 80x, sigma 0.5) clipped to 30-200x, PE150 at 0.3%
substitutions, k = 55, CMS 4 x 2^23 u16 + HLL p14 per isolate.  The first
wave the benchmark forms (isolates 0-2, 1.95 Gbases: the footprints of the
first three fill 60% of the device) runs as one batch and every isolate's
CMS, HLL and solid set equal a plain single-isolate build of it.

Served workload: requests of the same distributions, each the k ladder
21 / 55 / 77 (gates None / 4 / 4) after estimate_genome.  Three requests of
that size (60 to 300 Mbases) give the same results alone and in an
overlapping mix, and equal plain builds at the request's estimated genome
size.
"""
import time

import numpy as np
import pytest

import bench
import fastbact_amd as F

pytestmark = pytest.mark.gpu


def _workload_isolate(i):
    g, gc, cov = bench.batch_shapes(64)[i]
    iso = F.synth_isolate(seed=0xFB03 * 1000 + i, genome_len=g, gc=gc, coverage=cov, read_len=150,
                          insert_mean=350.0, insert_sd=35.0, sub_rate=0.003, indel_rate=0.0)
    return g, iso.bank()


def _same(a, b, what):
    assert np.array_equal(a.keys, b.keys), what
    assert np.array_equal(a.cov, b.cov) and np.array_equal(a.mask, b.mask), what
    assert np.array_equal(a.cms_table, b.cms_table), what
    assert np.array_equal(a.hll_registers, b.hll_registers), what


def test_first_wave_of_batched_workload(ctx):
    gs, devs = [], []
    for i in range(3):
        g, bank = _workload_isolate(i)
        gs.append(g)
        devs.append(F.DeviceBank(ctx, bank))
        del bank
    kw = dict(cms_width=bench.BATCH_WIDTH, counter_bits=16, hll_p=bench.BATCH_HLL_P)
    batch = F.Batch(ctx, devs, gs, **kw)
    for rep in range(2):  # the second run replays the captured graphs
        batch.run(bench.BATCH_K, None)
        assert all(r == "lds" for r in batch.routes()), batch.routes()
        for i, d in enumerate(devs):
            got = batch.result(i, bench.BATCH_K)
            plain = F.build_solid_kmers(ctx, d, bench.BATCH_K, target_solid_cov=None, genome_size=gs[i], **kw)
            print(f"rep {rep} isolate {i}: {len(got)} solid 55-mers", flush=True)
            assert len(got) > 0.9 * gs[i]
            _same(got, plain, (rep, i))
    batch.close()
    for d in devs:
        d.close()


KS, TARGETS = (21, 55, 77), (None, 4, 4)


def test_served_workload_alone_and_in_mix(ctx):
    # three request sizes of the served workload: 1.5 Mbp at 40x, 3 Mbp at 80x, 5 Mbp at 60x
    shapes = [(1_500_000, 0.45, 40.0), (3_000_000, 0.60, 80.0), (5_000_000, 0.35, 60.0)]
    banks = []
    for j, (g, gc, cov) in enumerate(shapes):
        iso = F.synth_isolate(seed=0xFB05 * 10 + j, genome_len=g, gc=gc, coverage=cov, read_len=150,
                              insert_mean=350.0, insert_sd=35.0, sub_rate=0.003, indel_rate=0.0)
        banks.append(iso.bank())
        del iso
    kw = dict(cms_width=bench.BATCH_WIDTH, counter_bits=16, hll_p=bench.BATCH_HLL_P, keep_sketches=1)
    alone = []
    with F.Server(0, KS, TARGETS, **kw) as sv:
        for b in banks:
            alone.append(sv.wait(sv.submit(b.pe_forward, b.pe_reverse), timeout=120))
    with F.Server(0, KS, TARGETS, **kw) as sv:
        ts = [sv.submit(banks[2].pe_forward, banks[2].pe_reverse)]
        deadline = time.time() + 60
        while sv.stats()["iterations"] == 0 and time.time() < deadline:
            time.sleep(0.001)
        ts += [sv.submit(b.pe_forward, b.pe_reverse) for b in banks[:2]]
        mix = [sv.wait(t, timeout=120) for t in ts]
        st = sv.stats()
    mix = [mix[1], mix[2], mix[0]]
    assert st["completed"] == 3 and st["failed"] == 0
    for j in range(3):
        assert alone[j].genome_size == mix[j].genome_size and alone[j].estimate == mix[j].estimate
        for pa, pb, k in zip(alone[j].passes, mix[j].passes, KS):
            _same(pa, pb, (j, k))
            print(f"request {j} k={k}: {len(pa)} solid k-mers, latency alone {alone[j].latency:.3f} s", flush=True)
    # the smallest request against plain builds at its estimated genome size
    dev = F.DeviceBank(ctx, banks[0])
    for pa, k, t in zip(alone[0].passes, KS, TARGETS):
        plain = F.build_solid_kmers(ctx, dev, k, target_solid_cov=t, genome_size=alone[0].genome_size,
                                    cms_width=bench.BATCH_WIDTH, counter_bits=16, hll_p=bench.BATCH_HLL_P)
        _same(pa, plain, k)
    dev.close()
