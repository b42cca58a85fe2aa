"""The repeat-rich, GC-rich isolate (6.3 Mbp at GC 0.666 with seven copies of a
5 kbp operon and 20 tandem arrays of 100-500 bp units in 2-6 copies, PE150 at
180x auto-subsampled to 110x, k = 77/99/127 with Some(4), CMS 4 x 2^24).
The subsample factor comes from the pipeline's own steps: estimate_genome
on the raw pairs (its coverage from the HLL-based genome size), then the
auto-subsampling policy at --depth 110 (main.rs:66-185), then the ingest
keep-test (trim_merge.rs:324-348).  At a reduced genome size every stage
equals the oracle (estimate, ingest bank, CMS table, HLL registers, solid
set); at full size, where the oracle would take minutes, the two pass-2
routes (LDS bins and the global table) and the two pass-1 placements must
agree bit for bit, and the set must be sorted and duplicate-free."""
import numpy as np
import pytest

import fastbact_amd as F
from tests.helpers import oracle_arenas

pytestmark = pytest.mark.gpu

SHAPE = dict(seed=0xFB04, gc=0.666, operon_copies=7, operon_len=5000, tandem_arrays=20, tandem_unit_min=100,
             tandem_unit_max=500, tandem_copies_min=2, tandem_copies_max=6, coverage=180.0, read_len=150,
             insert_mean=400.0, insert_sd=40.0, sub_rate=0.003)
DEPTH = 110.0


def _quals(n, L, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(63, 74, size=(n, L), dtype=np.uint8)  # Phred 30-40


def _auto_factor(ctx, iso):
    """estimate_genome on the raw pairs (max_pairs 250000, target --depth),
    then the auto-subsampling policy: the factor the pipeline would use."""
    raw = iso.bank()
    est = F.estimate_genome(ctx, raw.pe_forward, raw.pe_reverse, 250_000, DEPTH)
    factor, why = F.auto_subsample_factor(est, DEPTH)
    assert why == "estimate" and est.confidence == "HIGH", (why, est)
    # 180x to about 110x; the repeats lower the reduced isolate's genome
    # estimate (207 kbp for 250 kbp), so its coverage reads higher
    assert 0.45 < factor < 0.70, (factor, est.estimated_cov)
    assert factor == 110.0 / est.estimated_cov
    return factor, est, raw


def test_reduced_isolate_every_stage_equals_oracle(ctx, orc):
    iso = F.synth_isolate(genome_len=250_000, **SHAPE)
    n, L = iso.r1.shape
    KEEP, est, raw = _auto_factor(ctx, iso)
    r1, r2 = raw.pe_forward, raw.pe_reverse
    o = orc.estimate_genome((r1.data, r1.offsets, r1.lengths), (r2.data, r2.offsets, r2.lengths),
                            orc.count_stats(r1.lengths, r2.lengths), 250_000, DEPTH)
    assert est.genome_size == o.genome_size and est.estimated_cov == o.estimated_cov
    q1, q2 = _quals(n, L, 1), _quals(n, L, 2)
    bank, st = F.ingest_arrays(ctx, iso.r1, q1, iso.r2, q2, subsample_factor=KEEP, seed=0xFB04)
    ref = orc.ingest_records(orc.records([bytes(r) for r in iso.r1], [bytes(q) for q in q1]),
                             orc.records([bytes(r) for r in iso.r2], [bytes(q) for q in q2]),
                             subsample_factor=KEEP, seed=0xFB04, min_len=30, min_qual=20)
    assert {k: st[k] for k in ref["stats"]} == ref["stats"]
    assert abs(st["retained_read_pairs"] / n - KEEP) < 0.03
    host = bank.download()
    for a, (d, o, ln) in zip(host.arenas(), ref["arenas"]):
        assert np.array_equal(a.lengths, ln) and np.array_equal(a.offsets, o)
        assert np.array_equal(a.data[:len(d)], d[:len(a.data)])
    for k in (77, 99, 127):
        g = F.build_solid_kmers(ctx, bank, k, target_solid_cov=4, cms_width=1 << 20, counter_bits=16, hll_p=14)
        r = orc.build_solid(oracle_arenas(host), k, target_solid_cov=4, width=1 << 20, counter_bits=16, hll_p=14,
                            threads=8)
        assert np.array_equal(g.keys, r["keys"]), k
        assert np.array_equal(g.cov, r["cov"]) and np.array_equal(g.mask, r["mask"])
        assert np.array_equal(g.cms_table, r["cms"].table())
        assert np.array_equal(g.hll_registers, r["hll"])
        assert len(g) > 150_000


def test_full_size_routes_agree():
    ctx = F.Context()
    iso = F.synth_isolate(genome_len=6_300_000, **SHAPE)
    n, L = iso.r1.shape
    KEEP, est, raw = _auto_factor(ctx, iso)
    print(f"estimate: genome {est.genome_size}, coverage {est.estimated_cov:.1f}x, factor {KEEP:.4f}", flush=True)
    del raw
    q = np.full((n, L), 73, np.uint8)  # Phred 40: nothing is quality-trimmed
    bank, st = F.ingest_arrays(ctx, iso.r1, q, iso.r2, q, subsample_factor=KEEP, seed=0xFB04)
    del iso, q
    assert abs(st["retained_read_pairs"] / n - KEEP) < 0.01
    kw = dict(target_solid_cov=4, cms_width=1 << 24, counter_bits=16, hll_p=16)
    for k in (77, 99, 127):
        lds = F.build_solid_kmers(ctx, bank, k, placement=2, **kw)
        glob = F.build_solid_kmers(ctx, bank, k, placement=1, **kw)
        assert np.array_equal(lds.keys, glob.keys), k
        assert np.array_equal(lds.cov, glob.cov) and np.array_equal(lds.mask, glob.mask)
        assert np.array_equal(lds.cms_table, glob.cms_table)
        assert np.array_equal(lds.hll_registers, glob.hll_registers)
        # sorted and duplicate-free under the reference's word order (most significant first)
        keys = lds.keys
        order = np.lexsort(tuple(keys[:, j] for j in range(keys.shape[1] - 1, -1, -1)))
        assert np.array_equal(order, np.arange(keys.shape[0]))
        assert not (keys[1:] == keys[:-1]).all(axis=1).any()
        # about one solid k-mer per genome position
        assert 0.9 * 6_300_000 < len(lds) < 1.2 * 6_300_000
