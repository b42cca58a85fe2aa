"""The oracle's consensus polishing (consensus.rs:64-271) pinned on the CPU:
a contig with planted substitutions is restored to the genome its reads come
from (closed form), forward and reverse anchors vote at the reference's
coordinates (consensus.rs:284-307), short and deleted nodes stay untouched,
repeated seeds are not used, and a string restatement (tests/pyref.py) agrees
on noisy reads."""
import numpy as np
import pytest

from tests import pyref as P
from tests.test_oracle_compact import _rand_seq


def _reads(genome, n, L, seed, err=0.0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        s = int(rng.integers(0, len(genome) - L + 1))
        r = list(genome[s:s + L])
        for i in np.flatnonzero(rng.random(L) < err):
            r[i] = "ACGT"[("ACGT".index(r[i]) + int(rng.integers(1, 4))) % 4]
        r = "".join(r)
        out.append(P.revcomp(r) if rng.random() < 0.5 else r)
    return out


def _run(orc, seqs, reads, min_len=500, deleted=None):
    g = {"seqs": [s.encode() for s in seqs], "circular": np.zeros(len(seqs), np.uint8)}
    if deleted is not None:
        g["deleted"] = np.array(deleted, np.uint8)
    res = orc.polish_contigs(g, [orc.pack([r.encode() for r in reads])], min_len)
    return [s.decode() for s in res["seqs"]], res["stats"]


def test_planted_substitutions_are_restored(orc):
    """Seeds sit every 25 bp and reads try windows every 30 bp, so about one
    read in eight anchors; at 150x that leaves >= 5 votes everywhere."""
    G = _rand_seq(11_000_000, 8000)
    contig = list(G)
    planted = list(range(300, 7700, 250))
    for p in planted:
        contig[p] = "ACGT"[("ACGT".index(contig[p]) + 1) % 4]
    contig = "".join(contig)
    reads = _reads(G, 8000, 150, 1)
    seqs, st = _run(orc, [contig], reads)
    assert seqs[0] == G
    assert st["bases_modified"] == len(planted)
    assert st["fwd_mapped_reads"] > 300 and st["rev_mapped_reads"] > 300
    assert st["fwd_pileup_bases"] + st["rev_pileup_bases"] <= 150 * 8000


def test_reverse_anchor_coordinates(orc):
    """A read that is the reverse complement of contig bases [981, 1131)
    anchors in reverse on the seed at 1100 (its window at r = 0) and votes the
    complement of its base i at c + k - 1 + r - i (consensus.rs:284-307):
    five such reads fix the base at 1000."""
    G = _rand_seq(12_000_000, 3000)
    bad = G[:1000] + ("A" if G[1000] != "A" else "C") + G[1001:]
    reads = [P.revcomp(G[981:1131])] * 5
    seqs, st = _run(orc, [bad], reads)
    assert seqs[0] == G and st["rev_mapped_reads"] == 5 and st["fwd_mapped_reads"] == 0
    assert st["rev_pileup_bases"] == 5 * 150
    # four reads are not enough coverage; 4 of 5 votes are not > 75%
    assert _run(orc, [bad], reads[:4])[0][0] == bad
    mixed = reads[:3] + [P.revcomp(bad[981:1131])] * 2
    assert _run(orc, [bad], mixed)[0][0] == bad


def test_eligibility_and_repeats(orc):
    G = _rand_seq(13_000_000, 3000)
    short = _rand_seq(14_000_000, 400)
    bad = G[:1500] + ("T" if G[1500] != "T" else "G") + G[1501:]
    reads = _reads(G, 3000, 150, 2) + _reads(short, 50, 150, 3)
    seqs, st = _run(orc, [bad, short], reads, min_len=500)
    assert seqs == [G, short]
    # deleted or short nodes are neither seeds nor targets
    seqs, st = _run(orc, [bad], reads, deleted=[1])
    assert seqs == [bad] and st["fwd_mapped_reads"] + st["rev_mapped_reads"] == 0
    # a contig repeated in full has no unique seed: nothing maps
    seqs, st = _run(orc, [bad, bad], reads)
    assert seqs == [bad, bad] and st["fwd_mapped_reads"] + st["rev_mapped_reads"] == 0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_matches_string_restatement(orc, seed):
    rng = np.random.default_rng(seed)
    G = _rand_seq(15_000_000 + seed, 6000)
    rep = G[1000:1200]
    contigs = [G[:2500], G[2400:6000], rep + _rand_seq(16_000_000 + seed, 600), G[100:400]]
    for c in range(2):
        s = list(contigs[c])
        for p in rng.integers(0, len(s), size=15):
            s[p] = "ACGT"[("ACGT".index(s[p]) + 1) % 4]
        contigs[c] = "".join(s)
    reads = _reads(G, 5000, 150, seed, err=0.01) + _reads(contigs[2], 300, 120, seed + 7)
    got, st = _run(orc, contigs, reads, min_len=300)
    want, wst = P.polish(contigs, [len(c) >= 300 for c in contigs], reads)
    assert got == want and st == wst
    assert st["bases_modified"] > 10 and st["fwd_mapped_reads"] > 100 and st["rev_mapped_reads"] > 100
