"""The oracle's unitig compaction (build.rs:330-597) pinned on the CPU:
closed forms (a genome with unique k-mers compacts to itself; a circular one
to one circular unitig), the reference's own 10 kb repeat case, hairpins and
self-reverse-complement cycles, and an independent string-based restatement
(tests/pyref.py) on solid sets built from noisy reads."""
import numpy as np
import pytest

from tests import pyref as P
from tests.test_oracle_solid import _circular_case, _oracle_set, _reads, REF

ALPHA = "ACGT"


def _arrays(solid, k):
    nw = 1 if k <= 31 else (2 if k <= 63 else 4)
    keys = np.array([P.ref_words(x, k) for x, _, _ in solid], np.uint64).reshape(len(solid), nw)
    cov = np.array([c for _, c, _ in solid], np.uint16)
    mask = np.array([m for _, _, m in solid], np.uint8)
    return keys, cov, mask


def _rand_seq(seed, n):
    return "".join(ALPHA[P.splitmix64(seed + i) % 4] for i in range(n))


def _unique_kmers(seq, k, circular=False):
    t = seq + seq[:k - 1] if circular else seq
    kms = [P.canonical(t[i:i + k])[0] for i in range(len(t) - k + 1)]
    return len(set(kms)) == len(kms)


def _same(a, b):
    assert a["seqs"] == b["seqs"]
    assert np.array_equal(a["cov"].view(np.uint32), b["cov"].view(np.uint32))
    assert np.array_equal(a["circular"], b["circular"])
    for f in ("src", "dst", "ov", "support"):
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("k", [15, 21, 31, 33, 55, 63, 65, 77, 127])
def test_linear_genome_compacts_to_itself(orc, k):
    g = _rand_seq(1000 + k, 1500)
    assert _unique_kmers(g, k)
    solid = P.solid_from_sequences([g], k)
    res = orc.compact(*_arrays(solid, k), k)
    assert res["seqs"] == [min(g, P.revcomp(g)).encode()]
    assert res["circular"].tolist() == [0] and res["cov"].tolist() == [10.0]
    assert res["src"].size == 0  # no overlaps, no self edges on a linear unitig
    _same(res, P.compact(solid, k))


@pytest.mark.parametrize("k", [21, 31, 55, 77])
def test_circular_genome_is_one_circular_unitig(orc, k):
    g = _rand_seq(2000 + k, 1200)
    assert _unique_kmers(g, k, circular=True)
    solid = P.solid_from_sequences([g], k, circular=True)
    res = orc.compact(*_arrays(solid, k), k)
    assert len(res["seqs"]) == 1 and res["circular"].tolist() == [1]
    s = res["seqs"][0].decode()
    assert len(s) == len(g) and (s in g + g or s in P.revcomp(g) * 2)
    # the cycle starts at its smallest k-mer, read forward
    smallest = solid[0][0]
    assert P.kmer_int(s[:k]) == smallest
    # the creation edge of a circular unitig: itself, forward, overlap k-1
    assert (int(res["src"][0]), int(res["dst"][0]), int(res["ov"][0])) == (0, 0, k - 1)
    _same(res, P.compact(solid, k))


def test_reference_repeat_case_gives_the_three_expected_unitigs(orc):
    """build.rs:604-687: R F1 R F2 circular, k = 31: the repeat and the two
    flanks, each flank unitig running from the repeat's last k-mer to its first."""
    g = REF["circular_repeat"]
    genome, reads = _circular_case()
    k = g["k"]
    rep = genome[:g["repeat_len"]]
    f1 = genome[g["repeat_len"]:g["repeat_len"] + g["flank_len"]]
    f2 = genome[2 * g["repeat_len"] + g["flank_len"]:]
    bases = sum(len(r) for r in reads)
    width = orc.cms_width_for_genome_size(max(bases // 60, 50_000))
    res = orc.build_solid([orc.pack(reads)], k, target_solid_cov=g["target_solid_cov"], width=width)
    solid = _oracle_set(res)
    graph = orc.compact(res["keys"], res["cov"], res["mask"], k)
    # the repeat extends over flanking bases the two copies share (here one
    # base on one side); the flanks lose those bases
    left = 0
    while f1[-1 - left] == f2[-1 - left]:
        left += 1
    right = 0
    while f1[right] == f2[right]:
        right += 1
    rep_x = f1[len(f1) - left:] + rep + f1[:right]
    f1x, f2x = f1[right:len(f1) - left], f2[right:len(f2) - left]
    expect = {min(s, P.revcomp(s)).encode()
              for s in (rep_x, rep_x[-k:] + f1x + rep_x[:k], rep_x[-k:] + f2x + rep_x[:k])}
    assert len(graph["seqs"]) == 3 and set(graph["seqs"]) == expect
    assert graph["circular"].tolist() == [0, 0, 0]
    _same(graph, P.compact(solid, k))


@pytest.mark.parametrize("k", [11, 21, 31])
def test_hairpin_and_self_reverse_complement_cycle(orc, k):
    """A palindromic (k+1)-mer makes a walk meet its own twin: the reference
    stops right after it.  A circular X + revcomp(X) is its own reverse
    complement with two such junctions."""
    x = _rand_seq(77 + k, 60)
    lin = _rand_seq(5, 40) + x + P.revcomp(x) + _rand_seq(9, 40)
    solid = P.solid_from_sequences([lin], k)
    _same(orc.compact(*_arrays(solid, k), k), P.compact(solid, k))
    circ = P.solid_from_sequences([x + P.revcomp(x)], k, circular=True)
    _same(orc.compact(*_arrays(circ, k), k), P.compact(circ, k))


@pytest.mark.parametrize("k,target", [(7, 2), (11, 2), (21, 3), (31, None), (33, 2), (65, 2)])
def test_compaction_matches_string_restatement_on_noisy_reads(orc, k, target):
    reads = []
    for seed in range(4):
        reads += _reads(200 + seed, n=80, g=500, ln=(k, 120), err=0.02)[0]
    res = orc.build_solid([orc.pack(reads)], k, target_solid_cov=target, width=1 << 14)
    solid = _oracle_set(res)
    graph = orc.compact(res["keys"], res["cov"], res["mask"], k)
    assert len(graph["seqs"]) > 3
    _same(graph, P.compact(solid, k))


def test_empty_solid_set(orc):
    res = orc.compact(np.zeros((0, 1), np.uint64), np.zeros(0, np.uint16), np.zeros(0, np.uint8), 31)
    assert res["seqs"] == [] and res["src"].size == 0
