"""The oracle's filter_active_reads (local_multik.rs:259-333) pinned on the
CPU: closed forms at the edges of the 350-window ends of a long unitig, the
short / circular / branching rules, canonical hashing, pairs kept by either
mate, k' = min(k, 31), and an independent string restatement (tests/pyref.py)
on graphs compacted from noisy reads."""
import numpy as np
import pytest

import fastbact_amd as F
from tests import pyref as P
from tests.helpers import oracle_arenas
from tests.test_oracle_compact import _rand_seq


def _graph(seqs, circular=None, nf=None, nr=None, deleted=None):
    """A graph dict as compaction returns it, edges given as counts per side."""
    n = len(seqs)
    src = []
    for i in range(n):
        src += [i << 1] * (0 if nf is None else nf[i])
        src += [i << 1 | 1] * (0 if nr is None else nr[i])
    g = {"seqs": [s.encode() for s in seqs], "circular": np.array(circular or [0] * n, np.uint8),
         "src": np.array(src, np.uint32)}
    if deleted is not None:
        g["deleted"] = np.array(deleted, np.uint8)
    return g


def _bank(merged=(), pairs=()):
    return F.ReadBank(F.BitSeqArena.from_ascii([m.encode() for m in merged]),
                      F.BitSeqArena.from_ascii([a.encode() for a, _ in pairs]),
                      F.BitSeqArena.from_ascii([b.encode() for _, b in pairs]))


def _run(orc, graph, k, merged=(), pairs=()):
    return orc.filter_active_reads(graph, k, oracle_arenas(_bank(merged, pairs)))


def _pyref(graph, k, merged=(), pairs=()):
    nodes = [(s.decode(), bool(c), bool(graph.get("deleted", [0] * len(graph["seqs"]))[i]), f, r)
             for i, (s, c, f, r) in enumerate(zip(graph["seqs"], graph["circular"], *_counts(graph)))]
    return P.filter_active_reads(nodes, k, list(merged), list(pairs))


def _counts(graph):
    n = len(graph["seqs"])
    src = np.asarray(graph["src"], np.int64)
    return (np.bincount(src[(src & 1) == 0] >> 1, minlength=n).tolist(),
            np.bincount(src[(src & 1) == 1] >> 1, minlength=n).tolist())


U = _rand_seq(90_001, 5000)  # a long unitig with unique 31-mers


@pytest.mark.parametrize("k", [21, 31])
def test_long_unitig_ends_and_interior(orc, k):
    """Windows p < 350 and p >= len-350 are in the set, nothing between:
    a read is active exactly when its first, middle or last window is."""
    L, R = len(U), 150
    g = _graph([U])
    starts = [0, 200, 349 - (R - k), 349 - (R - k) + 1, 349, 350, 351, 1000, 2400,
              L - 350 - R, L - 350 - R + 1, L - 350 - (R - k) // 2 - 1, L - 350 - (R - k) // 2, L - R]
    reads = [U[q:q + R] for q in starts]
    res = _run(orc, g, k, merged=reads)
    expect = []
    for q in starts:
        wins = (q, q + (R - k) // 2, q + R - k)
        expect.append(any(p < 350 or p >= L - 350 for p in wins))
    assert res["keep_merged"].astype(bool).tolist() == expect
    assert expect.count(False) >= 5 and expect.count(True) >= 5
    assert res["keep_merged"].astype(bool).tolist() == _pyref(g, k, reads)[0]


def test_reverse_strand_reads_hit_canonically(orc):
    g = _graph([U])
    reads = [P.revcomp(U[q:q + 150]) for q in (0, 300, 2000, len(U) - 150)]
    assert _run(orc, g, 31, merged=reads)["keep_merged"].tolist() == [1, 1, 0, 1]


@pytest.mark.parametrize("rule", ["short", "circular", "branch_fwd", "branch_rev"])
def test_short_circular_branching_unitigs_keep_every_window(orc, rule):
    seq = U[:1999] if rule == "short" else U
    g = _graph([seq], circular=[1] if rule == "circular" else None,
               nf=[2] if rule == "branch_fwd" else [1], nr=[2] if rule == "branch_rev" else [1])
    reads = [seq[q:q + 150] for q in range(0, len(seq) - 150, 97)]
    assert _run(orc, g, 31, merged=reads)["keep_merged"].all()
    # one edge per side and >= 2000 bp: interior reads go
    plain = _graph([U], nf=[1], nr=[1])
    assert not _run(orc, plain, 31, merged=[U[2000:2150]])["keep_merged"].any()


def test_deleted_nodes_and_short_reads(orc):
    g = _graph([U[:1500], U[3000:4500]], deleted=[1, 0])
    reads = [U[100:250], U[3100:3250], U[3100:3130], U[3100:3131]]
    # node 0 is deleted; a read shorter than k' = 31 is never active
    assert _run(orc, g, 31, merged=reads)["keep_merged"].tolist() == [0, 1, 0, 1]


def test_pairs_kept_by_either_mate(orc):
    g = _graph([U])
    hit, miss = U[10:160], U[2000:2150]
    pairs = [(hit, miss), (miss, hit), (miss, miss), (hit, hit), (miss, U[2000:2020])]
    res = _run(orc, g, 31, pairs=pairs)
    assert res["keep_pairs"].tolist() == [1, 1, 0, 1, 0]
    # the filtered bank holds the kept pairs in order, mates aligned
    assert res["reads"][1] == [hit.encode(), miss.encode(), hit.encode()]
    assert res["reads"][2] == [miss.encode(), hit.encode(), hit.encode()]


def test_pairs_beyond_the_shorter_mate_arena_are_dropped(orc):
    g = _graph([U[:1000]])
    r = U[100:250]
    bank = F.ReadBank(F.BitSeqArena.from_ascii([]), F.BitSeqArena.from_ascii([r.encode()] * 3),
                      F.BitSeqArena.from_ascii([r.encode()] * 2))
    res = orc.filter_active_reads(g, 31, oracle_arenas(bank))
    assert res["keep_pairs"].tolist() == [1, 1] and len(res["reads"][1]) == 2


def test_graph_k_above_31_hashes_31_mers(orc):
    """k' = min(k, 31): at graph k 77 a 40-base read still has 31-mer windows."""
    g = _graph([U[:1000]])
    res = _run(orc, g, 77, merged=[U[100:140], U[100:130]])
    assert res["keep_merged"].tolist() == [1, 0]


def test_filtered_arenas_are_push_seq_packed(orc):
    g = _graph([U])
    reads = [U[q:q + L] for q, L in ((0, 150), (2000, 150), (100, 33), (4800, 199), (10, 31))]
    res = _run(orc, g, 31, merged=reads)
    kept = [r.encode() for r, f in zip(reads, res["keep_merged"]) if f]
    data, offs, lens = res["arenas"][0]
    assert lens.tolist() == [len(r) for r in kept]
    assert offs.tolist() == np.concatenate([[0], np.cumsum(lens)[:-1]]).tolist()
    assert orc.unpack(data, offs, lens) == kept


def _sampled_reads(seed, genome, n, L=150, err=0.002):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = int(rng.integers(0, len(genome) - L + 1))
        s = list(genome[a:a + L])
        for i in np.flatnonzero(rng.random(L) < err):
            s[i] = "ACGT"[int(rng.integers(0, 4))]
        r = "".join(s)
        out.append(P.revcomp(r) if rng.random() < 0.5 else r)
    return out


@pytest.mark.parametrize("k,target", [(15, 3), (21, 3), (31, None), (33, 3), (55, 3), (77, 2)])
def test_filter_matches_string_restatement_on_compacted_graphs(orc, k, target):
    """Graphs compacted from reads of a 12 kb genome with a repeat: long
    linear unitigs (ends only), short ones and branches (every window)."""
    rep = _rand_seq(700_000, 300)
    genome = _rand_seq(500_000, 4000) + rep + _rand_seq(600_000, 4000) + rep + _rand_seq(800_000, 3500)
    reads = _sampled_reads(k, genome, 1600)
    res = orc.build_solid([orc.pack([r.encode() for r in reads])], k, target_solid_cov=target, width=1 << 18)
    graph = orc.compact(res["keys"], res["cov"], res["mask"], k)
    assert max(len(s) for s in graph["seqs"]) >= 2000
    merged, pairs = reads[:800], list(zip(reads[800::2], reads[801::2]))
    out = _run(orc, graph, k, merged=merged, pairs=pairs)
    km, kp = _pyref(graph, k, merged, pairs)
    assert out["keep_merged"].astype(bool).tolist() == km
    assert out["keep_pairs"].astype(bool).tolist() == kp
    assert 0 < sum(km) < len(km) and 0 < sum(kp) < len(kp)
