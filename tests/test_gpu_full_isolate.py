"""GPU checks on the benchmark's own isolate: 5 Mbp at GC 0.508, PE150 at
100x (1,666,667 pairs, insert N(400,40), 0.3% substitutions + 0.01% indels),
passes at k = 21 / 55 / 77 with gates None / 4 / 4, CMS 4 x 2^24 saturating
u16 counters, HLL p = 16 -- the workload whose throughput bench.py reports.

The oracle needs minutes per pass at this size, so the full isolate is
checked through properties that do not depend on its size, and a 300k-pair
prefix of the same reads against the oracle itself:
- the LDS route (binned pass 1, LDS tables in pass 2; its pass-2 gate
  applied once per kept key and low edge, and per window pair as placement
  3) and the global route (global atomics, global hash table, gate per
  window pair) give the same CMS, HLL and solid set;
- the set is sorted by key and free of duplicates;
- a 4-bit run's table equals min(u16 counter, 15) cell by cell in the
  reference's nibble layout, with the same solid set (gates <= 15);
- a p = 14 run's registers equal the fold of the p = 16 registers;
- on the first 300k pairs every pass equals the oracle's threaded form.
"""
import numpy as np
import pytest

import fastbact_amd as F
from tests.helpers import head_pairs, hll_fold_p16_to_p14, nibble_pack, oracle_arenas

pytestmark = pytest.mark.gpu

PASSES = [(21, None), (55, 4), (77, 4)]
WIDTH = 1 << 24
GENOME = dict(seed=0xFB02, genome_len=5_000_000, gc=0.508, coverage=100.0, read_len=150, insert_mean=400.0,
              insert_sd=40.0, sub_rate=0.003, indel_rate=0.0001)


@pytest.fixture(scope="module")
def full(ctx):
    iso = F.synth_isolate(**GENOME)
    bank = iso.bank()
    assert len(bank.pe_forward) == 1_666_667
    dev = F.DeviceBank(ctx, bank)
    yield bank, dev
    dev.close()


def _build(ctx, dev, k, tgt, **kw):
    kw.setdefault("counter_bits", 16)
    kw.setdefault("hll_p", 16)
    return F.build_solid_kmers(ctx, dev, k, target_solid_cov=tgt, genome_size=GENOME["genome_len"], cms_width=WIDTH,
                               **kw)


def _sorted_unique(keys: np.ndarray) -> bool:
    # rows of MSB-first words: lexicographic order = integer order
    if len(keys) < 2:
        return True
    a, b = keys[:-1], keys[1:]
    gt = np.zeros(len(a), bool)
    eq = np.ones(len(a), bool)
    for w in range(keys.shape[1]):
        gt |= eq & (b[:, w] > a[:, w])
        eq &= b[:, w] == a[:, w]
    return bool(gt.all())


@pytest.mark.parametrize("k,tgt", PASSES)
def test_routes_agree_sorted_unique(ctx, full, k, tgt):
    _, dev = full
    lds = _build(ctx, dev, k, tgt, placement=2)
    glb = _build(ctx, dev, k, tgt, placement=1)
    print(f"k={k}: {len(lds)} solid k-mers", flush=True)
    assert 4_000_000 < len(lds) < 20_000_000
    assert np.array_equal(lds.keys, glb.keys)
    assert np.array_equal(lds.cov, glb.cov) and np.array_equal(lds.mask, glb.mask)
    win = _build(ctx, dev, k, tgt, placement=3, with_cms=False)
    assert np.array_equal(win.keys, lds.keys)
    assert np.array_equal(win.cov, lds.cov) and np.array_equal(win.mask, lds.mask)
    assert np.array_equal(lds.cms_table, glb.cms_table)
    assert np.array_equal(lds.hll_registers, glb.hll_registers)
    assert _sorted_unique(lds.keys)
    assert (lds.mask != 0).all() and (lds.cov >= 1).all()


@pytest.mark.parametrize("k,tgt", [(21, None), (77, 4)])
def test_nibble_table_is_clamped_u16(ctx, full, k, tgt):
    _, dev = full
    u16 = _build(ctx, dev, k, tgt, hll_p=0)
    nib = _build(ctx, dev, k, tgt, counter_bits=4, hll_p=0)
    assert u16.cms_table.max() > 15
    assert np.array_equal(nib.cms_table, nibble_pack(u16.cms_table))
    assert np.array_equal(nib.keys, u16.keys) and np.array_equal(nib.cov, u16.cov)
    assert np.array_equal(nib.mask, u16.mask)


def test_hll_p16_folds_to_p14(ctx, full):
    _, dev = full
    p16 = _build(ctx, dev, 21, None, hll_p=16)
    p14 = _build(ctx, dev, 21, None, hll_p=14)
    assert np.array_equal(hll_fold_p16_to_p14(p16.hll_registers), p14.hll_registers)


def test_prefix_300k_pairs_equals_oracle(ctx, orc, full):
    bank, _ = full
    sub = head_pairs(bank, 300_000)
    dev = F.DeviceBank(ctx, sub)
    try:
        for k, tgt in PASSES:
            g = _build(ctx, dev, k, tgt)
            r = orc.build_solid(oracle_arenas(sub), k, target_solid_cov=tgt, width=WIDTH, counter_bits=16, hll_p=16,
                                threads=16)
            print(f"k={k}: {len(g)} solid k-mers on the prefix", flush=True)
            assert np.array_equal(g.keys, r["keys"])
            assert np.array_equal(g.cov, r["cov"]) and np.array_equal(g.mask, r["mask"])
            assert np.array_equal(g.cms_table, r["cms"].table())
            assert np.array_equal(g.hll_registers, r["hll"])
    finally:
        dev.close()


def test_batch_of_full_isolate_and_prefix_matches_plain(ctx, full):
    """One batch holding the full isolate and its 300k-pair prefix: at
    k = 77 the full isolate's pass 2 takes 4096 bins (the direct record
    scatter) and the prefix's at most 2048 (the buffered one), both in the
    same launch; each isolate's set, table and registers equal its plain
    build's."""
    bank, dev = full
    sub = F.DeviceBank(ctx, head_pairs(bank, 300_000))
    try:
        batch = F.Batch(ctx, [dev, sub], [GENOME["genome_len"]] * 2, cms_width=WIDTH, counter_bits=16, hll_p=16,
                        placement=2)
        for k, tgt in ((77, 4), (55, 4)):
            batch.run(k, tgt)
            assert batch.routes() == ["lds", "lds"]
            for i, d in enumerate((dev, sub)):
                single = _build(ctx, d, k, tgt)
                got = batch.result(i, k)
                assert np.array_equal(got.keys, single.keys), (k, i)
                assert np.array_equal(got.cov, single.cov) and np.array_equal(got.mask, single.mask)
                assert np.array_equal(got.cms_table, single.cms_table)
                assert np.array_equal(got.hll_registers, single.hll_registers)
    finally:
        sub.close()
