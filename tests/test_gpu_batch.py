"""GPU parity of the batched form: every isolate's result from one batched
launch equals the oracle's on that isolate alone."""
import numpy as np
import pytest

import fastbact_amd as F
from tests.helpers import isolate, oracle_arenas

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def isolates():
    out = []
    for i, (g, n, gc) in enumerate([(20_000, 1500, 0.35), (35_000, 2500, 0.5), (12_000, 900, 0.66),
                                    (50_000, 3000, 0.45), (8_000, 0, 0.5), (25_000, 2000, 0.6)]):
        out.append(isolate(seed=100 + i, genome_len=g, n_pairs=n, gc=gc) if n else None)
    return out


def _bank(iso):
    return iso.bank() if iso is not None else F.ReadBank()


@pytest.mark.parametrize("placement", [1, 2])
@pytest.mark.parametrize("k,target", [(21, None), (55, 4), (77, 3)])
def test_batch_equals_oracle_per_isolate(ctx, orc, isolates, k, target, placement):
    banks = [_bank(x) for x in isolates]
    dev = [F.DeviceBank(ctx, b) for b in banks]
    gs = [50_000] * len(banks)
    width = 1 << 16
    batch = F.Batch(ctx, dev, gs, cms_width=width, counter_bits=16, hll_p=12, placement=placement)
    for rep in range(2):  # second run replays the captured graph
        batch.run(k, target)
        for i, b in enumerate(banks):
            got = batch.result(i, k)
            ref = orc.build_solid(oracle_arenas(b), k, target_solid_cov=target, width=width, counter_bits=16,
                                  hll_p=12)
            assert np.array_equal(got.keys, ref["keys"]), (rep, i)
            assert np.array_equal(got.cov, ref["cov"]) and np.array_equal(got.mask, ref["mask"])
            assert np.array_equal(got.cms_table, ref["cms"].table())
            assert np.array_equal(got.hll_registers, ref["hll"])


@pytest.mark.parametrize("width", [1 << 22, 1 << 24, 1 << 25])
def test_batch_pass1_forms_equal_oracle(ctx, orc, isolates, width):
    """Pass 1 of a batch runs the per-bin LDS buffer kernel when the isolates
    have 1025 to 2048 bins (width 2^24: 2048 bins) and the staged scatter
    otherwise (2^22: 512 bins, 2^25: 4096 bins); both give each isolate the
    oracle's table, registers and solid set, the empty isolate included."""
    banks = [_bank(x) for x in isolates]
    dev = [F.DeviceBank(ctx, b) for b in banks]
    batch = F.Batch(ctx, dev, [50_000] * len(banks), cms_width=width, counter_bits=16, hll_p=12, placement=2)
    for rep in range(2):
        batch.run(31, 3)
        for i, b in enumerate(banks):
            got = batch.result(i, 31)
            ref = orc.build_solid(oracle_arenas(b), 31, target_solid_cov=3, width=width, counter_bits=16, hll_p=12)
            assert np.array_equal(got.keys, ref["keys"]), (rep, i)
            assert np.array_equal(got.cov, ref["cov"]) and np.array_equal(got.mask, ref["mask"])
            assert np.array_equal(got.cms_table, ref["cms"].table())
            assert np.array_equal(got.hll_registers, ref["hll"])


def test_batch_multi_k_reuses_buffers_and_matches_single(ctx, isolates):
    banks = [_bank(x) for x in isolates[:4]]
    dev = [F.DeviceBank(ctx, b) for b in banks]
    batch = F.Batch(ctx, dev, [60_000] * 4, cms_width=1 << 17, counter_bits=4)
    for k, t in ((21, None), (55, 4), (77, 4), (21, None)):
        batch.run(k, t)
        for i in range(4):
            single = F.build_solid_kmers(ctx, dev[i], k, target_solid_cov=t, cms_width=1 << 17, counter_bits=4)
            got = batch.result(i, k)
            assert np.array_equal(got.keys, single.keys) and np.array_equal(got.cov, single.cov)
            assert np.array_equal(got.cms_table, single.cms_table)


def test_batch_small_table_grows_exactly(ctx, orc, isolates):
    # tiny genome sizes force the per-isolate tables to overflow and grow
    banks = [_bank(x) for x in isolates[:3]]
    dev = [F.DeviceBank(ctx, b) for b in banks]
    batch = F.Batch(ctx, dev, [1, 1, 1], cms_width=1 << 15, counter_bits=16)
    batch.run(31, 1)
    for i, b in enumerate(banks):
        ref = orc.build_solid(oracle_arenas(b), 31, target_solid_cov=1, width=1 << 15, counter_bits=16)
        got = batch.result(i, 31)
        assert np.array_equal(got.keys, ref["keys"]) and np.array_equal(got.cov, ref["cov"])


@pytest.mark.parametrize("k,bins2", [(21, 0), (55, 1 << 15), (99, 1 << 16)])
def test_batch_lds_route_in_rounds_equals_oracle(ctx, orc, isolates, k, bins2):
    """Passes 2-3 of every isolate on LDS bins in one launch per kernel,
    including pass-2 rounds; the replayed graphs give the same sets."""
    banks = [_bank(x) for x in isolates]
    dev = [F.DeviceBank(ctx, b) for b in banks]
    width = 1 << 16
    batch = F.Batch(ctx, dev, [50_000] * len(banks), cms_width=width, counter_bits=16, hll_p=12, placement=2,
                    bins2=bins2)
    refs = [orc.build_solid(oracle_arenas(b), k, target_solid_cov=2, width=width, counter_bits=16, hll_p=12)
            for b in banks]
    for rep in range(3):
        batch.run(k, 2)
        routes = batch.routes()
        for i, b in enumerate(banks):
            got = batch.result(i, k)
            assert np.array_equal(got.keys, refs[i]["keys"]), (rep, i)
            assert np.array_equal(got.cov, refs[i]["cov"]) and np.array_equal(got.mask, refs[i]["mask"])
            assert np.array_equal(got.hll_registers, refs[i]["hll"])
            assert routes[i] == ("lds" if b.total_bases() else "global"), (i, routes)


def test_batch_matches_plain_path_on_larger_isolates(ctx):
    """Uneven isolates whose keys need several pass-2 rounds at a small
    sketch: batch results equal fb_build_solid's, isolate by isolate."""
    isos = [isolate(seed=300 + i, genome_len=g, n_pairs=n, gc=gc)
            for i, (g, n, gc) in enumerate([(200_000, 40_000, 0.4), (90_000, 12_000, 0.55), (300_000, 60_000, 0.62)])]
    dev = [F.DeviceBank(ctx, x.bank()) for x in isos]
    batch = F.Batch(ctx, dev, [100_000] * 3, cms_width=1 << 17, counter_bits=16, hll_p=14)
    for k, t in ((31, None), (77, 4)):
        batch.run(k, t)
        for i, d in enumerate(dev):
            single = F.build_solid_kmers(ctx, d, k, target_solid_cov=t, cms_width=1 << 17, counter_bits=16,
                                         hll_p=14)
            got = batch.result(i, k)
            assert np.array_equal(got.keys, single.keys) and np.array_equal(got.cov, single.cov)
            assert np.array_equal(got.mask, single.mask)
            assert np.array_equal(got.cms_table, single.cms_table)
            assert np.array_equal(got.hll_registers, single.hll_registers)
        assert batch.routes() == ["lds"] * 3


def test_autotune_keeps_results_and_carries_to_same_shape(ctx, orc, isolates):
    """The tuner times launch shapes on the batch; every shape computes the
    same sets, the choice is kept, and a new batch of the same shape uses it."""
    F._lib.load().fb_autotune_clear()
    banks = [_bank(x) for x in isolates[:4]]
    dev = [F.DeviceBank(ctx, b) for b in banks]
    width = 1 << 16
    batch = F.Batch(ctx, dev, [50_000] * 4, cms_width=width, counter_bits=16, hll_p=12)
    r = batch.autotune(55, 4)
    # the defaults (all 0) may stay the choice; a choice is never slower than them
    assert r["trials"] >= 10 and r["placement"] in (0, 1, 2) and r["best_ms"] <= r["default_ms"]
    for i, b in enumerate(banks):
        ref = orc.build_solid(oracle_arenas(b), 55, target_solid_cov=4, width=width, counter_bits=16, hll_p=12)
        got = batch.result(i, 55)
        assert np.array_equal(got.keys, ref["keys"]) and np.array_equal(got.cov, ref["cov"])
    again = F.Batch(ctx, dev, [50_000] * 4, cms_width=width, counter_bits=16, hll_p=12)
    again.run(55, 4)
    for i in range(4):
        assert np.array_equal(again.result(i, 55).keys, batch.result(i, 55).keys)
    F._lib.load().fb_autotune_clear()


@pytest.mark.parametrize("chunks1,chunks2,split", [(1, 1, 1), (16, 16, 6), (3, 5, 2)])
def test_launch_shapes_do_not_change_results(ctx, orc, chunks1, chunks2, split):
    iso = isolate(seed=41, genome_len=150_000, n_pairs=30_000)
    bank = iso.bank()
    dev = F.DeviceBank(ctx, bank)
    ref = orc.build_solid(oracle_arenas(bank), 31, target_solid_cov=3, width=1 << 18, counter_bits=16, hll_p=14)
    g = F.build_solid_kmers(ctx, dev, 31, target_solid_cov=3, cms_width=1 << 18, counter_bits=16, hll_p=14,
                            placement=2, chunks1=chunks1, chunks2=chunks2, split=split)
    assert np.array_equal(g.keys, ref["keys"]) and np.array_equal(g.cov, ref["cov"])
    assert np.array_equal(g.mask, ref["mask"]) and np.array_equal(g.cms_table, ref["cms"].table())
    assert np.array_equal(g.hll_registers, ref["hll"])
