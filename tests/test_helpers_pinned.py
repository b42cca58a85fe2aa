"""The test helpers that GPU tests compare against, pinned on the CPU:
the p16 -> p14 HLL fold and the nibble packing against the oracle, and the
exact distinct canonical k-mer count against a string brute force."""
import numpy as np

import fastbact_amd as F
from tests.helpers import distinct_canonical_kmers, hll_fold_p16_to_p14, isolate, nibble_pack, oracle_arenas


def test_hll_fold_matches_p14_registers(orc):
    rng = np.random.default_rng(17)
    hs = rng.integers(0, 2**64 - 1, size=200_000, dtype=np.uint64)
    hs[:3] = 0
    r16 = np.zeros(1 << 16, np.uint8)
    r14 = np.zeros(1 << 14, np.uint8)
    orc.hll_add_hashes(r16, 16, hs)
    orc.hll_add_hashes(r14, 14, hs)
    assert np.array_equal(hll_fold_p16_to_p14(r16), r14)
    # a planted mistake (one p16 register raised) breaks the identity
    j = int(np.nonzero((r16 > 0) & ((np.arange(1 << 16) & 3) == 0))[0][0])
    bad = r16.copy()
    bad[j] += 3
    assert not np.array_equal(hll_fold_p16_to_p14(bad), r14)


def test_nibble_pack_matches_nibble_table(orc):
    iso = isolate(seed=3, genome_len=8_000, n_pairs=600)
    bank = iso.bank()
    a16 = orc.build_solid(oracle_arenas(bank), 21, width=1 << 10, counter_bits=16)
    a4 = orc.build_solid(oracle_arenas(bank), 21, width=1 << 10, counter_bits=4)
    t16 = a16["cms"].table()
    assert t16.max() > 15  # saturation is exercised
    assert np.array_equal(nibble_pack(t16), a4["cms"].table())
    assert np.array_equal(a16["keys"], a4["keys"])


def _brute_distinct(reads, k):
    comp = bytes.maketrans(b"ACGT", b"TGCA")
    seen = set()
    for r in reads:
        s = bytes(r)
        for i in range(len(s) - k + 1):
            w = s[i:i + k]
            rc = w.translate(comp)[::-1]
            seen.add(min(w, rc))  # lexicographic order on ACGT = integer order
    return len(seen)


def test_distinct_canonical_kmers_matches_brute_force():
    rng = np.random.default_rng(5)
    g = rng.choice(np.frombuffer(b"ACGT", np.uint8), size=3000)
    starts = rng.integers(0, 3000 - 60, size=400)
    reads = np.stack([g[s:s + 60] for s in starts])
    flip = rng.random(400) < 0.5
    reads[flip] = np.frombuffer(bytes(reads[flip]).translate(bytes.maketrans(b"ACGT", b"TGCA")),
                                np.uint8).reshape(-1, 60)[:, ::-1]
    mut = rng.random(reads.shape) < 0.01
    reads[mut] = rng.choice(np.frombuffer(b"ACGT", np.uint8), size=int(mut.sum()))
    for k in (5, 21, 31):
        assert distinct_canonical_kmers([reads[:200], reads[200:]], k) == _brute_distinct(reads, k)
