"""Pins the oracle's passes 1-3 (src/graph/build.rs:67-328): brute force with
exact counts, the reference's 10 kb circular-repeat case, and agreement of the
threaded (reference-shaped) form with the sequential form."""
import json
from pathlib import Path

import numpy as np
import pytest

from tests import pyref as P

GOLD = Path(__file__).resolve().parent / "golden"
REF = json.loads((GOLD / "reference_vectors.json").read_text())


def _reads(seed, n=60, g=400, ln=(0, 90), err=0.03):
    rng = np.random.default_rng(seed)
    genome = "".join(rng.choice(list("ACGT"), size=g))
    out = []
    for _ in range(n):
        L = int(rng.integers(ln[0], ln[1] + 1))
        a = int(rng.integers(0, g - L + 1))
        s = list(genome[a:a + L])
        if rng.random() < 0.5:
            s = list(P.revcomp("".join(s)))
        for i in range(L):
            if rng.random() < err:
                s[i] = "ACGTNacgtn"[int(rng.integers(0, 10))]
        out.append("".join(s).encode())
    return out, genome


def _oracle_set(res):
    nw = res["keys"].shape[1] if res["keys"].size else 1
    out = []
    for row, c, m in zip(res["keys"].tolist(), res["cov"].tolist(), res["mask"].tolist()):
        x = 0
        for w in row:
            x = (x << 64) | w
        out.append((x, c, m))
    del nw
    return out


@pytest.mark.parametrize("k,width,target,bits", [
    (5, 64, None, 4), (11, 256, 2, 16), (21, 1024, 3, 4), (31, 128, 1, 16), (32, 2048, None, 4),
    (33, 512, 2, 4), (63, 4096, 2, 16), (64, 256, 1, 4), (65, 1024, 2, 16), (99, 2048, 1, 4), (127, 512, 1, 16)])
def test_solid_set_brute_force(orc, k, width, target, bits):
    reads, genome = _reads(k, n=80, g=300, ln=(0, 160))
    contigs = [genome[10:10 + k + 30].encode(), genome[100:100 + k].encode()] if k in (21, 33, 65) else []
    arenas = [orc.pack(reads[:30]), orc.pack(reads[30:55]), orc.pack(reads[55:])]
    res = orc.build_solid(arenas, k, target_solid_cov=target, width=width, counter_bits=bits, contigs=contigs)
    expect, cells = P.build_solid(reads, k, width, target, bits, contigs)
    assert _oracle_set(res) == expect
    t = P.cms_table_from_counts(cells, width, bits)
    got = res["cms"].table()
    assert (got.tobytes() == t) if bits == 4 else (got.tolist() == t)


def test_threaded_form_equals_sequential(orc):
    reads, genome = _reads(77, n=3000, g=20000, ln=(60, 250), err=0.01)
    arenas = [orc.pack(reads[:1000]), orc.pack(reads[1000:2000]), orc.pack(reads[2000:])]
    contigs = [genome[:3000].encode(), genome[5000:5100].encode()]
    for k in (21, 55, 77, 127):
        a = orc.build_solid(arenas, k, width=1 << 14, hll_p=12, contigs=contigs)
        b = orc.build_solid(arenas, k, width=1 << 14, hll_p=12, contigs=contigs, threads=6)
        for f in ("keys", "cov", "mask", "hll"):
            assert np.array_equal(a[f], b[f]), (k, f)
        assert np.array_equal(a["cms"].table(), b["cms"].table())


def test_per_pass_hll_is_splitmix_of_canonical_nthash(orc):
    reads, _ = _reads(3, n=40, g=500, ln=(20, 120))
    res = orc.build_solid([orc.pack(reads)], 25, width=1024, hll_p=10)
    regs = [0] * 1024
    for r in reads:
        s = P.norm(r)
        for i in range(len(s) - 25 + 1):
            idx, rho = P.hll_index_rho(P.splitmix64(P.nthash_canonical(s[i:i + 25])), 10)
            regs[idx] = max(regs[idx], rho)
    assert res["hll"].tolist() == regs


def _circular_case():
    g = REF["circular_repeat"]
    alpha = "ACGT"

    def gen(seed, n):
        return "".join(alpha[P.splitmix64(seed + i) % 4] for i in range(n))

    rep = gen(g["seeds"][0], g["repeat_len"])
    f1 = gen(g["seeds"][1], g["flank_len"])
    f2 = gen(g["seeds"][2], g["flank_len"])
    genome = rep + f1 + rep + f2
    G = len(genome)
    n_reads = G * g["depth"] // g["read_len"]
    rr = g["read_seed"]
    reads = []
    for i in range(n_reads):  # build.rs:645-671
        rr = P.splitmix64(rr + i)
        start = rr % G
        read = []
        for j in range(g["read_len"]):
            base = genome[(start + j) % G]
            if P.splitmix64(rr ^ j) % 100 == 0:
                alt = 1 + P.splitmix64(rr + j * 31) % 3
                read.append(alpha[("ACGT".index(base) + alt) % 4])
            else:
                read.append(base)
        reads.append("".join(read).encode())
    return genome, reads


def test_circular_repeat_known_answer(orc):
    """build.rs:604-687: 10 kb circular genome with one 200 bp repeat, 40x,
    1% errors, solid cutoff 5 -> errors filtered and exactly 3 unitigs."""
    g = REF["circular_repeat"]
    genome, reads = _circular_case()
    k = g["k"]
    bases = sum(len(r) for r in reads)
    width = orc.cms_width_for_genome_size(max(bases // 60, 50_000))  # build.rs:74-84 fallback
    res = orc.build_solid([orc.pack(reads)], k, target_solid_cov=g["target_solid_cov"], width=width)
    solid = _oracle_set(res)
    circ = genome + genome[:k - 1]
    genome_kmers = {P.canonical(circ[i:i + k])[0] for i in range(len(genome))}
    keys = {x for x, _, _ in solid}
    assert keys == genome_kmers  # every genome k-mer solid, every error k-mer filtered
    assert P.count_unitigs(solid, k) == g["expect_unitigs"]


def test_planted_edge_mistake_breaks_unitig_count(orc):
    # a transposed edge encoding (edge_u <-> edge_v halves) must not give 3
    g = REF["circular_repeat"]
    genome, reads = _circular_case()
    bases = sum(len(r) for r in reads)
    width = orc.cms_width_for_genome_size(max(bases // 60, 50_000))
    res = orc.build_solid([orc.pack(reads)], g["k"], target_solid_cov=g["target_solid_cov"], width=width)
    solid = [(x, c, ((m & 15) << 4) | (m >> 4)) for x, c, m in _oracle_set(res)]
    assert P.count_unitigs(solid, g["k"]) != g["expect_unitigs"]
