"""GPU parity of the PE mapper (mapper.rs:89-279): the device index and
pair mapping give the oracle's link evidence and insert sizes record for
record, and the same InsertSizeDist, on crafted graphs and on graphs compacted
from GPU builds of synthetic isolates."""
import numpy as np
import pytest

import fastbact_amd as F
from tests import pyref as P
from tests.helpers import isolate
from tests.test_oracle_compact import _rand_seq

pytestmark = pytest.mark.gpu


def _bank(pairs):
    return F.ReadBank(F.BitSeqArena(), F.BitSeqArena.from_ascii([a.encode() for a, _ in pairs]),
                      F.BitSeqArena.from_ascii([b.encode() for _, b in pairs]))


def _same(g, r):
    assert g["links"].size == r["links"].size
    for f in g["links"].dtype.names:
        assert np.array_equal(g["links"][f], r["links"][f]), f
    assert np.array_equal(g["sizes"], r["sizes"])
    assert g["dist"] == r["dist"]


def _both(ctx, orc, graph, bank, k):
    dev = F.DeviceBank(ctx, bank)
    g = F.map_paired_reads(ctx, dev, graph, k)
    r = orc.map_paired_reads(graph, (bank.pe_forward.data, bank.pe_forward.offsets, bank.pe_forward.lengths),
                             (bank.pe_reverse.data, bank.pe_reverse.offsets, bank.pe_reverse.lengths), k)
    _same(g, r)
    return g


@pytest.mark.parametrize("k", [21, 31, 33, 55, 63, 65, 77, 127])
def test_crafted_graph(ctx, orc, k):
    rng = np.random.default_rng(k)
    seqs = [_rand_seq(7_000_000 + 10_000 * i, int(rng.integers(300, 5000))) for i in range(8)]
    seqs.append(seqs[1][:400])  # shared k-mers
    genome = "".join(seqs[:8])
    pairs = []
    for _ in range(3000):
        ins = int(rng.integers(150, 900))
        s = int(rng.integers(0, len(genome) - ins))
        frag = genome[s:s + ins]
        r1, r2 = frag[:150], P.revcomp(frag[-150:])
        if rng.random() < 0.5:
            r1, r2 = r2, r1
        if rng.random() < 0.2:
            j = int(rng.integers(0, len(r1)))
            r1 = r1[:j] + "ACGT"[("ACGT".index(r1[j]) + 1) % 4] + r1[j + 1:]
        pairs.append((r1, r2[:int(rng.integers(k - 1, 151))]))
    graph = {"seqs": [x.encode() for x in seqs], "circular": np.zeros(len(seqs), np.uint8),
             "deleted": (np.arange(len(seqs)) == 3).astype(np.uint8)}
    g = _both(ctx, orc, graph, _bank(pairs), k)
    assert g["links"].size > 100 and g["sizes"].size > 20


@pytest.mark.parametrize("k", [31, 55])
def test_compacted_graph_of_a_gpu_build(ctx, orc, k):
    iso = isolate(seed=90 + k, genome_len=120_000, n_pairs=12_000, sub_rate=0.003)
    bank = iso.bank()
    dev = F.DeviceBank(ctx, bank)
    h = F.build_solid_handle(ctx, dev, k, target_solid_cov=3, cms_width=1 << 20, counter_bits=16)
    try:
        graph = F.compact_unitigs(ctx, h)
    finally:
        h.close()
    g = _both(ctx, orc, graph, bank, k)
    assert g["sizes"].size > 100
    # the synthetic inserts are N(400, 40)-like: the median lands near them
    assert 250 < g["dist"]["median"] < 600


def test_empty_inputs(ctx, orc):
    empty = {"seqs": [], "circular": np.zeros(0, np.uint8)}
    g = _both(ctx, orc, empty, _bank([("ACGT" * 40, "TTGCA" * 30)]), 31)
    assert g["links"].size == 0 and g["dist"]["median"] == 450.0
    _both(ctx, orc, {"seqs": [_rand_seq(1, 3000).encode()], "circular": np.zeros(1, np.uint8)}, F.ReadBank(), 31)
