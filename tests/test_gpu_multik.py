"""GPU parity of the multi-k flow (local_multik.rs:38-147,259-333):
filter_active_reads on the device keeps exactly the oracle's reads and packs
them byte-identically; upgrade_multik's k2 solid set equals the oracle's
build over the oracle-filtered bank with the prior contigs, and its graph the
oracle's compaction plus the transferred circular unitigs."""
import numpy as np
import pytest

import fastbact_amd as F
from tests import pyref as P
from tests.helpers import isolate, oracle_arenas, ragged_bank
from tests.test_oracle_compact import _rand_seq
from tests.test_oracle_filter import U, _bank, _graph

pytestmark = pytest.mark.gpu


def _same_bank(dev_bank: F.DeviceBank, arenas):
    got = dev_bank.download()
    for a, (data, offs, lens) in zip(got.arenas(), arenas):
        nw = (int(lens.sum(dtype=np.uint64)) + 31) // 32
        assert np.array_equal(a.lengths, lens)
        assert np.array_equal(a.offsets, offs)
        assert np.array_equal(a.data, data[:nw])


def _filter_both(ctx, orc, graph, k, bank: F.ReadBank):
    dev = F.DeviceBank(ctx, bank)
    out, km, kp = F.filter_active_reads(ctx, dev, graph, k)
    ref = orc.filter_active_reads(graph, k, oracle_arenas(bank))
    assert np.array_equal(km, ref["keep_merged"])
    assert np.array_equal(kp, ref["keep_pairs"])
    _same_bank(out, ref["arenas"])
    assert out.total_bases() == sum(int(x[2].sum()) for x in ref["arenas"])
    return out, ref


def test_filter_crafted_graph_edges_of_the_rules(ctx, orc):
    L, R = len(U), 150
    g = _graph([U, U[:1999].replace("A", "C")], nf=[1, 0], nr=[1, 0])
    merged = [U[q:q + R] for q in range(0, L - R, 37)] + [U[:30], U[:31], ""]
    pairs = [(U[q:q + R], P.revcomp(U[q + 200:q + 200 + R])) for q in range(0, L - 400, 53)]
    _, ref = _filter_both(ctx, orc, g, 31, _bank(merged, pairs))
    assert 0 < ref["keep_merged"].sum() < len(merged) and 0 < ref["keep_pairs"].sum() < len(pairs)


@pytest.mark.parametrize("k", [15, 21, 31, 55, 77])
def test_filter_on_compacted_gpu_graph(ctx, orc, k):
    iso = isolate(seed=70 + k, genome_len=60_000, n_pairs=6000, sub_rate=0.004)
    bank = iso.bank()
    dev = F.DeviceBank(ctx, bank)
    h = F.build_solid_handle(ctx, dev, k, target_solid_cov=3, cms_width=1 << 20, counter_bits=16)
    try:
        graph = F.compact_unitigs(ctx, h)
    finally:
        h.close()
    assert max(len(s) for s in graph["seqs"]) >= 2000
    out, ref = _filter_both(ctx, orc, graph, k, bank)
    assert 0 < ref["keep_pairs"].sum() < len(ref["keep_pairs"])


def test_filter_ragged_bank_and_short_words(ctx, orc):
    """Reads of every length (0 .. 260, shorter than a word and than k') in all
    three arenas, unequal mate counts."""
    bank = ragged_bank(seed=9, n=500)
    bank.pe_reverse = F.BitSeqArena(bank.pe_reverse.data, bank.pe_reverse.offsets[:450],
                                    bank.pe_reverse.lengths[:450])
    seqs = []
    for a in bank.arenas():
        seqs += orc.unpack(a.data, a.offsets, a.lengths, range(0, len(a.lengths), 7))
    g = {"seqs": [s for s in seqs if len(s) >= 40][:40], "circular": np.zeros(40, np.uint8),
         "src": np.zeros(0, np.uint32)}
    g["circular"] = g["circular"][:len(g["seqs"])]
    _filter_both(ctx, orc, g, 21, bank)


def test_filter_empty_graph_and_empty_bank(ctx, orc):
    empty = {"seqs": [], "circular": np.zeros(0, np.uint8), "src": np.zeros(0, np.uint32)}
    bank = _bank([U[:150]], [(U[:150], U[200:350])])
    out, ref = _filter_both(ctx, orc, empty, 31, bank)
    assert out.total_reads() == 0
    out, _ = _filter_both(ctx, orc, _graph([U]), 31, F.ReadBank())
    assert out.total_reads() == 0


# ---- upgrade_multik --------------------------------------------------------

def _reads_of(genome, n, L, seed, circular=False):
    rng = np.random.default_rng(seed)
    G = len(genome)
    ext = genome + genome[:L] if circular else genome
    out = []
    for _ in range(n):
        a = int(rng.integers(0, G if circular else G - L + 1))
        r = ext[a:a + L]
        out.append(P.revcomp(r) if rng.random() < 0.5 else r)
    return out


def _repeat_case():
    """R F1 R F2 (a 45 bp repeat: branching at k 31, resolved at k 55), and a
    2 kb circular plasmid read only in 50 bp pieces (no 55-mers), so the
    upgrade transfers it."""
    rep = _rand_seq(4_000_000, 45)
    chrom = _rand_seq(1_000_000, 3000) + rep + _rand_seq(2_000_000, 3000) + rep + _rand_seq(3_000_000, 2500)
    plasmid = _rand_seq(5_000_000, 2000)
    chrom_reads = _reads_of(chrom, len(chrom) * 40 // 150, 150, 1)
    pl_reads = _reads_of(plasmid, len(plasmid) * 60 // 50, 50, 2, circular=True)
    merged = chrom_reads[:len(chrom_reads) // 3] + pl_reads
    rest = chrom_reads[len(chrom_reads) // 3:]
    pairs = list(zip(rest[0::2], rest[1::2]))
    return _bank(merged, pairs)


def _oracle_upgrade(orc, bank, graph, k, k2, genome_size):
    filt = orc.filter_active_reads(graph, k, oracle_arenas(bank))
    prior = [s for s, c in zip(graph["seqs"], graph["circular"]) if not c]
    est_g = max(genome_size, 50_000) if genome_size else max(bank.total_bases() // 60, 50_000)
    width = orc.cms_width_for_genome_size(est_g)
    res = orc.build_solid(filt["arenas"], k2, target_solid_cov=4, width=width, counter_bits=4, contigs=prior)
    g2 = orc.compact(res["keys"], res["cov"], res["mask"], k2)
    return filt, est_g, width, res, g2


@pytest.mark.parametrize("genome_size", [None, 9000])
def test_upgrade_matches_oracle_flow(ctx, orc, genome_size):
    bank = _repeat_case()
    dev = F.DeviceBank(ctx, bank)
    k, k2 = 31, 55
    h = F.build_solid_handle(ctx, dev, k, target_solid_cov=5)
    try:
        g1 = F.compact_unitigs(ctx, h)
    finally:
        h.close()
    assert any(c for c in g1["circular"])  # the plasmid at k 31
    out = F.upgrade_multik(ctx, dev, g1, k, k2, genome_size=genome_size, keep_solid=True, keep_active=True)
    filt, est_g, width, res, g2 = _oracle_upgrade(orc, bank, g1, k, k2, genome_size)
    assert out["executed"] and out["early_stop_reason"] is None
    assert out["genome_size_used"] == est_g and out["cms_width"] == width
    assert out["genome_size_source"] == ("EXPLICIT_PASSED" if genome_size else "FALLBACK_READ_BASES_DIV_60")
    assert out["cms_allocated_bytes"] == 4 * width // 2
    assert out["kept_merged"] == int(filt["keep_merged"].sum()) and out["kept_pairs"] == int(filt["keep_pairs"].sum())
    _same_bank(out["active"], filt["arenas"])
    s = out["solid"].download(k2, with_cms=False)
    assert np.array_equal(s.keys, res["keys"]) and np.array_equal(s.cov, res["cov"])
    assert np.array_equal(s.mask, res["mask"])
    # the graph: the oracle's compaction, then the plasmid (no 55-mer reads)
    g = out["graph"]
    n2 = len(g2["seqs"])
    assert g["seqs"][:n2] == g2["seqs"]
    assert not any(g2["circular"]) and out["transferred_circular"] == 1
    pl = [s for s, c in zip(g1["seqs"], g1["circular"]) if c]
    assert g["seqs"][n2:] == pl and g["circular"][n2:].tolist() == [1]
    assert (int(g["src"][-1]), int(g["dst"][-1]), int(g["ov"][-1]), int(g["support"][-1])) == (n2 << 1, n2 << 1, k2 - 1, 10)
    assert out["active_unitigs_after"] == n2 + 1
    assert out["active_edges_after"] == len(g2["src"]) + 1
    # the repeat is resolved at k 55: the chromosome is one unitig
    assert max(len(x) for x in g["seqs"]) >= 8500
    out["solid"].close()


def test_upgrade_early_stops(ctx):
    bank = _bank([U[q:q + 150] for q in range(0, len(U) - 150, 20)])
    dev = F.DeviceBank(ctx, bank)
    lin = {"seqs": [U.encode()], "circular": np.zeros(1, np.uint8), "src": np.zeros(0, np.uint32)}
    r = F.upgrade_multik(ctx, dev, lin, 31, 55)
    assert not r["executed"] and r["early_stop_reason"] == "NO_BRANCHING_REGIONS" and r["graph"] is None
    assert r["active_unitigs_before"] == r["active_unitigs_after"] == 1
    br = _graph([U], nf=[2], nr=[0])
    r = F.upgrade_multik(ctx, dev, br, 55, 31)
    assert not r["executed"] and r["early_stop_reason"] == "K2_LE_PREV_K"
    assert r["genome_size_source"] == "SKIPPED" and r["active_edges_before"] == 2
