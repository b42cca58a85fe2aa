"""GPU parity of consensus polishing (consensus.rs:64-271): the device seeds,
anchors, pileups and consensus give the oracle's polished sequences and
PolishStats, on crafted contigs with planted errors (restored to the genome)
and on graphs compacted from GPU builds."""
import numpy as np
import pytest

import fastbact_amd as F
from tests.helpers import isolate
from tests.test_oracle_compact import _rand_seq
from tests.test_oracle_polish import _reads

pytestmark = pytest.mark.gpu


def _both(ctx, orc, graph, bank, min_len):
    dev = F.DeviceBank(ctx, bank)
    seqs, st = F.polish_contigs(ctx, dev, graph, min_len)
    ref = orc.polish_contigs(graph, [(a.data, a.offsets, a.lengths) for a in (bank.merged_reads, bank.pe_forward)],
                             min_len)
    assert seqs == ref["seqs"]
    assert st == ref["stats"]
    return seqs, st


def test_planted_errors(ctx, orc):
    G = _rand_seq(21_000_000, 20_000)
    contigs = [G[:9000], G[8900:20_000], G[3000:3400]]
    planted = []
    for c in (0, 1):
        s = list(contigs[c])
        for p in range(200, len(s) - 200, 333):
            s[p] = "ACGT"[("ACGT".index(s[p]) + 1) % 4]
            planted.append((c, p))
        contigs[c] = "".join(s)
    reads = _reads(G, 25_000, 150, 5, err=0.002)
    merged = [r.encode() for r in reads[:8000]]
    fwd = [r.encode() for r in reads[8000:]]
    bank = F.ReadBank(F.BitSeqArena.from_ascii(merged), F.BitSeqArena.from_ascii(fwd),
                      F.BitSeqArena.from_ascii([b"ACGT"] * len(fwd)))  # pe_reverse is not used
    graph = {"seqs": [c.encode() for c in contigs], "circular": np.zeros(3, np.uint8),
             "deleted": np.array([0, 0, 0], np.uint8)}
    seqs, st = _both(ctx, orc, graph, bank, 500)
    assert seqs[0] == G[:9000].encode() and seqs[1] == G[8900:20_000].encode()
    assert st["bases_modified"] == len(planted) and seqs[2] == contigs[2].encode()


@pytest.mark.parametrize("k", [31, 55])
def test_compacted_graph_of_a_gpu_build(ctx, orc, k):
    iso = isolate(seed=130 + k, genome_len=100_000, n_pairs=15_000, sub_rate=0.004)
    bank = iso.bank()
    dev = F.DeviceBank(ctx, bank)
    h = F.build_solid_handle(ctx, dev, k, target_solid_cov=3, cms_width=1 << 20, counter_bits=16)
    try:
        graph = F.compact_unitigs(ctx, h)
    finally:
        h.close()
    seqs, st = _both(ctx, orc, graph, bank, 200)
    assert st["fwd_mapped_reads"] + st["rev_mapped_reads"] > 500


def test_empty_inputs(ctx, orc):
    _both(ctx, orc, {"seqs": [], "circular": np.zeros(0, np.uint8)}, F.ReadBank(), 500)
    g = {"seqs": [_rand_seq(1, 2000).encode()], "circular": np.zeros(1, np.uint8)}
    _both(ctx, orc, g, F.ReadBank(), 500)
