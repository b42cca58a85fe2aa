"""Pins the oracle's genome estimator (src/ingest/estimator.rs)."""
import json
import math
from pathlib import Path

import numpy as np
import pytest
from scipy.stats import poisson

from tests import pyref as P

GOLD = Path(__file__).resolve().parent / "golden"
REF = json.loads((GOLD / "reference_vectors.json").read_text())


def test_count_stats_known_answer(orc):
    g = REF["count_fastx_exact"]
    st = orc.count_stats(np.full(g["pairs"], g["r1_len"]), np.full(g["pairs"], g["r2_len"]))
    for f, v in g["expect"].items():
        assert getattr(st, f) == v
    assert orc.count_stats(np.zeros(0), np.zeros(0)) is None  # "0 valid read pairs"


def test_count_stats_median_and_overflow_lengths(orc):
    st = orc.count_stats([10, 20, 30], [40, 50, 60])
    assert st.median_read_len == 30  # (6+1)/2 = 3rd smallest
    assert st.min_read_len == 10 and st.max_read_len == 60
    # every length >= 4096: median from the sorted overflow list
    st = orc.count_stats([5000, 7000], [6000, 8000])
    assert st.median_read_len == 7000


def test_dynamic_kmers_reference_and_derived(orc):
    g = REF["dynamic_kmers"]
    ks = orc.compute_dynamic_kmers(150)
    assert ks[0] == g["median150"]["first"] and ks[-1] == g["median150"]["last"]
    assert all(k % 2 == 1 and 31 <= k <= 127 for k in ks)
    ks = orc.compute_dynamic_kmers(50)
    assert ks[0] == g["median50"]["first"] and all(k % 2 == 1 for k in ks)
    # ladders worked by hand from estimator.rs:104-161
    assert orc.compute_dynamic_kmers(150) == [31, 63, 127]
    assert orc.compute_dynamic_kmers(50) == [21, 45, 67]
    assert orc.compute_dynamic_kmers(75) == [31, 67, 101]
    assert orc.compute_dynamic_kmers(30) == [21, 39]
    assert orc.compute_dynamic_kmers(20) == [21, 27]
    assert orc.compute_dynamic_kmers(10) == [21]
    assert orc.compute_dynamic_kmers(100) == [31, 63, 127]
    assert orc.compute_dynamic_kmers(150, 61) == [31, 61]


@pytest.mark.parametrize("u1,u2,m", [(600_000, 520_000, 9_000_000), (1_200_000, 1_000_000, 40_000_000),
                                     (30_000, 20_000, 60_000), (5_000, 4_000, 1_000_000), (100, 0, 500)])
def test_poisson_tail_against_scipy(orc, u1, u2, m):
    # P(X >= 2) of Poisson(lambda) is the survival function at 1
    st = orc.count_stats(np.full(1000, 150), np.full(1000, 150))
    out = orc.estimate_finish(u1, u2, m, st, 100.0)
    lam = (m - max(u1 - u2, 0)) / u2 if u2 else 0.0
    assert out.sample_kmer_depth == pytest.approx(lam, rel=1e-15)
    p2 = poisson.sf(1, lam) if lam > 0 else 0.0
    if p2 > 0.05:
        assert abs(int(out.genome_size) - max(round(u2 / p2), 10_000)) <= 1
    else:
        assert out.genome_size == max(u2, 10_000)
    assert out.estimated_cov == st.total_bases / out.genome_size


def test_confidence_and_subsample_branches(orc):
    st = orc.count_stats(np.full(100_000, 150), np.full(100_000, 150))  # 30 Mbases
    high = orc.estimate_finish(1_100_000, 1_000_000, 30_000_000, st, 10.0)
    assert high.confidence == 0  # D_k = 30e6/G * 130/150 >= 20
    assert high.subsample_factor == pytest.approx(10.0 / high.estimated_cov)
    low = orc.estimate_finish(2_000_000, 400_000, 8_000_000, st, 0.0)
    assert low.subsample_factor == 1.0  # target_cov 0: no subsampling
    unreliable = orc.estimate_finish(10_000, 900, 100_000, st, 50.0)
    assert unreliable.confidence == 2  # u_ge_2 < 1000
    # solid ratio 0.1 with adequate depth -> LOW_COVERAGE_WARNING
    warn = orc.estimate_finish(10_000_000, 1_000_000, 30_000_000, st, 50.0)
    assert warn.confidence == 1


def test_f64_to_u64_saturates(orc):
    f = orc.lib().orc_f64_to_u64
    assert f(-3.0) == 0 and f(float("nan")) == 0 and f(2.0**70) == 2**64 - 1 and f(12.9) == 12


def _pairs(seed, n, ln=60, g=3000):
    rng = np.random.default_rng(seed)
    genome = "".join(rng.choice(list("ACGT"), size=g))
    out = []
    for _ in range(n):
        a = int(rng.integers(0, g - ln))
        r1 = list(genome[a:a + ln])
        b = int(rng.integers(0, g - ln))
        r2 = list(P.revcomp(genome[b:b + ln]))
        for r in (r1, r2):
            for i in range(ln):
                if rng.random() < 0.02:
                    r[i] = "ACGTN"[int(rng.integers(0, 5))]
        out.append(("".join(r1).encode(), "".join(r2).encode()))
    return out


def test_estimator_sketch_order_dependent_solid_hll(orc):
    # hll_solid takes a k-mer at the occurrence where every CMS row already
    # held an insertion (estimator.rs:324-337): checked against a sequential
    # Python emulation that evaluates c >= 2 at insertion time
    pairs = _pairs(4, 120)
    r1 = orc.pack([a for a, _ in pairs])
    r2 = orc.pack([b for _, b in pairs])
    ht, hs, m = orc.estimator_sketch(r1, r2, len(pairs))
    eht, ehs, em = P.estimator_registers(pairs, len(pairs))
    assert m == em
    assert ht.tolist() == eht
    assert hs.tolist() == ehs
    assert (ht >= hs).all() and hs.sum() > 0


def test_estimate_genome_small_sample_end_to_end(orc):
    pairs = _pairs(5, 200, ln=80, g=5000)
    r1 = orc.pack([a for a, _ in pairs])
    r2 = orc.pack([b for _, b in pairs])
    st = orc.count_stats(r1[2], r2[2])
    out = orc.estimate_genome(r1, r2, st, 250_000, 100.0)
    ht, hs, m = P.estimator_registers(pairs, len(pairs))
    u1 = int(round(P.hll_estimate(ht, 14)))
    u2 = int(round(P.hll_estimate(hs, 14)))
    assert (out.u_ge_1, out.u_ge_2, out.m_sample, out.pairs_sampled) == (u1, u2, m, 200)
    ref = orc.estimate_finish(u1, u2, m, st, 100.0)
    assert out.genome_size == ref.genome_size
    assert out.kmers[0] == 31  # median 80 -> min_k 31
    assert orc.estimate_genome(r1, r2, st, 0, 100.0) is None  # no pairs sampled -> error
    assert not math.isnan(out.estimated_cov)
