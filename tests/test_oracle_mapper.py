"""The oracle's PE mapper (mapper.rs:89-279) pinned on the CPU: insert sizes
of pairs drawn from a long unitig equal the true fragment lengths (closed
form), links between two unitigs carry the distances to their ends, the
terminal-window and uniqueness rules of the index, the InsertSizeDist rule
(median, MAD floor 20, 3.5 MAD bounds, the default under 20 sizes), and a
dictionary restatement (tests/pyref.py) on noisy pairs."""
import numpy as np
import pytest

from tests import pyref as P
from tests.test_oracle_compact import _rand_seq


def _g(seqs):
    return {"seqs": [s.encode() for s in seqs], "circular": np.zeros(len(seqs), np.uint8)}


def _arenas(orc, pairs):
    return orc.pack([a.encode() for a, _ in pairs]), orc.pack([b.encode() for _, b in pairs])


def _run(orc, seqs, pairs, k):
    a, b = _arenas(orc, pairs)
    return orc.map_paired_reads(_g(seqs), a, b, k)


def _tuples(links):
    return [tuple(int(x) for x in r) for r in links]


U = _rand_seq(3_000_000, 6000)
V = _rand_seq(4_000_000, 3000)


@pytest.mark.parametrize("k", [21, 31, 55, 77])
def test_intra_unitig_spans_are_fragment_lengths(orc, k):
    """Pairs from the first 1000 windows of a 6 kb unitig (R1 forward, R2 the
    reverse mate, and the mirrored orientation): span = fragment length."""
    pairs, want = [], []
    for s, ins in ((0, 300), (40, 420), (100, 700), (200, 180), (10, 1000 - k)):
        frag = U[s:s + ins]
        pairs.append((frag[:150], P.revcomp(frag[-150:])))
        pairs.append((P.revcomp(frag[-150:]), frag[:150]))
        want += [ins, ins]
    res = _run(orc, [U], pairs, k)
    assert res["sizes"].tolist() == want
    assert res["links"].size == 0
    assert res["sizes"].tolist() == P.map_paired_reads([U], pairs, k)[1]


def test_spans_outside_the_window_or_limits_are_dropped(orc):
    k = 31
    pairs = [(U[2000:2150], P.revcomp(U[2250:2400])),     # interior: not indexed
             (U[0:150], P.revcomp(U[0:150])),               # span 150 < 100? no: 150 ok, but FF below
             (U[0:150], U[200:350]),                         # same strand: rejected
             (U[0:80], P.revcomp(U[20:100]))]               # span 100 >= max(80, 100)
    res = _run(orc, [U], pairs, k)
    assert res["sizes"].tolist() == [150, 100]
    short = U[:1400]  # node shorter than 1500: no sizes
    assert _run(orc, [short], [(short[0:150], P.revcomp(short[100:400]))], k)["sizes"].size == 0


def test_links_between_two_unitigs(orc):
    k = 31
    # R1 near U's end (forward), R2 from V's start (reverse mate): u1 = U fwd, u2 forward = V fwd
    r1 = U[-400:-250]
    r2 = P.revcomp(V[100:250])
    res = _run(orc, [U, V], [(r1, r2)], k)
    d1 = len(U) - (len(U) - 400)
    d2 = 250
    assert _tuples(res["links"]) == [(0, 2, 0, d1, d2, 150, 150), (3, 1, 0, d2, d1, 150, 150)]
    assert _tuples(res["links"]) == P.map_paired_reads([U, V], [(r1, r2)], k)[0]


def test_repeated_kmers_are_not_indexed(orc):
    k = 31
    rep = _rand_seq(9_000_000, 200)
    a = _rand_seq(9_100_000, 800) + rep
    b = rep + _rand_seq(9_200_000, 800)
    pairs = [(rep[:120], P.revcomp(a[:150])), (a[:150], P.revcomp(b[-150:]))]
    res = _run(orc, [a, b], pairs, k)
    # the first pair's R1 lies in the shared 200 bp: no hit at steps 0..100 until it leaves the repeat
    assert _tuples(res["links"]) == P.map_paired_reads([a, b], pairs, k)[0]
    assert all(r[2] == 1 for r in _tuples(res["links"]))


def test_insert_dist_rule(orc):
    assert orc.insert_dist([300] * 19) == {"median": 450.0, "mad": 50.0, "min_insert": 250, "max_insert": 700}
    s = [300] * 11 + [310] * 9 + [1000]  # sorted index 10 is 300
    d = orc.insert_dist(s)
    assert d["median"] == 300.0 and d["mad"] == 20.0  # MAD 0 raised to 20
    assert (d["min_insert"], d["max_insert"]) == (230, 370)
    rng = np.random.default_rng(3)
    x = rng.integers(100, 1500, size=1001)
    med = float(np.sort(x)[500])
    mad = max(float(np.sort(np.abs(x - med))[500]), 20.0)
    d = orc.insert_dist(x)
    assert d["median"] == med and d["mad"] == mad
    assert d["min_insert"] == int(max(med - 3.5 * mad, 50.0)) and d["max_insert"] == int(med + 3.5 * mad)


@pytest.mark.parametrize("k", [21, 31, 33, 63, 77])
def test_matches_dictionary_restatement(orc, k):
    rng = np.random.default_rng(k)
    seqs = [_rand_seq(5_000_000 + 10_000 * i, int(rng.integers(200, 4000))) for i in range(6)]
    seqs.append(seqs[0][:300])  # shared k-mers: not unique
    genome = "".join(seqs[:6])
    pairs = []
    for _ in range(400):
        ins = int(rng.integers(150, 800))
        s = int(rng.integers(0, len(genome) - ins))
        frag = genome[s:s + ins]
        r1, r2 = frag[:150], P.revcomp(frag[-150:])
        if rng.random() < 0.5:
            r1, r2 = r2, r1
        if rng.random() < 0.3:
            j = int(rng.integers(0, 150))
            r1 = r1[:j] + "ACGT"[(("ACGT".index(r1[j])) + 1) % 4] + r1[j + 1:]
        pairs.append((r1, r2))
    res = _run(orc, seqs, pairs, k)
    links, sizes = P.map_paired_reads(seqs, pairs, k)
    assert _tuples(res["links"]) == links
    assert res["sizes"].tolist() == sizes
    assert len(links) > 50 and len(sizes) > 20
