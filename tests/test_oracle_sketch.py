"""Pins the oracle's Count-Min Sketch and HyperLogLog (src/common/sketch.rs)."""
import json
import threading
from pathlib import Path

import numpy as np
import pytest

from tests import pyref as P

GOLD = Path(__file__).resolve().parent / "golden"
REF = json.loads((GOLD / "reference_vectors.json").read_text())
PYV = json.loads((GOLD / "pyref_vectors.json").read_text())


def test_cms_basic_known_answer(orc):
    g = REF["cms_basic"]
    c = orc.Cms(g["width"])
    assert c.estimate(g["key"]) == g["after0"]
    for _ in range(g["adds1"]):
        c.add(g["key"])
    assert c.estimate(g["key"]) == g["after1"]
    for _ in range(g["adds2"]):
        c.add(g["key"])
    assert c.estimate(g["key"]) == g["after2"]


def test_cms_concurrent_known_answer(orc):
    g = REF["cms_concurrent"]
    c = orc.Cms(g["width"])

    def work():
        for _ in range(g["per_thread"]):
            c.add(g["key"])
    ts = [threading.Thread(target=work) for _ in range(g["threads"])]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert c.estimate(g["key"]) == g["expect"]


def test_cms_cells_vectors_and_table_layout(orc):
    # every counter equals min(#adds landing in its cell, cap): the table is a
    # clamped histogram of the cells computed independently in Python
    for key, w, cells in PYV["cms_cells"]:
        assert P.cms_cells(key, w) == cells
    rng = np.random.default_rng(8)
    keys = rng.integers(0, 2**64 - 1, size=3000, dtype=np.uint64).tolist()
    keys += keys[:200] * 20  # pushes some cells past 15
    for bits in (4, 16):
        width = 512
        c = orc.Cms(width, bits)
        hist = {}
        for x in keys:
            c.add(x)
            for cell in P.cms_cells(x, width):
                hist[cell] = hist.get(cell, 0) + 1
        t = P.cms_table_from_counts(hist, width, bits)
        got = c.table()
        if bits == 4:
            assert got.tobytes() == t
        else:
            assert got.tolist() == t
        for x in keys[:300]:
            assert c.estimate(x) == P.cms_estimate(hist, x, width, bits)


def test_cms_width_rounding_and_genome_sizing(orc):
    assert orc.Cms(1000).width == 1024
    assert orc.Cms(1024).width == 1024
    # sketch.rs:165-168: clamp(next_pow2(16 G), 2^21, 2^27)
    assert orc.cms_width_for_genome_size(50_000) == 1 << 21
    assert orc.cms_width_for_genome_size(1_000_000) == 1 << 24
    assert orc.cms_width_for_genome_size(5_000_000) == 1 << 27
    assert orc.cms_width_for_genome_size(200_000) == 1 << 22
    assert orc.cms_width_for_genome_size(10**9) == 1 << 27


def test_hll_index_rho_vectors(orc):
    for h, p, idx, rho in PYV["hll_index_rho"]:
        regs = np.zeros(1 << p, np.uint8)
        orc.hll_add_hashes(regs, p, [h])
        assert regs[idx] == rho and regs.sum() == rho


def test_hll_estimate_matches_python(orc):
    rng = np.random.default_rng(2)
    for p in (10, 14, 16):
        for n in (10, 1000, 200_000):
            regs = np.zeros(1 << p, np.uint8)
            orc.hll_add_hashes(regs, p, rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64))
            assert orc.hll_estimate(regs, p) == pytest.approx(P.hll_estimate(regs.tolist(), p), rel=1e-12)


def test_hll_p16_folds_to_p14(orc):
    # register j of p=16 maps to j>>2 of p=14: the p=14 rank is the p=16 rank
    # plus the leading zeros contributed by the two index bits that moved
    rng = np.random.default_rng(9)
    hs = rng.integers(0, 2**64 - 1, size=300_000, dtype=np.uint64)
    r16 = np.zeros(1 << 16, np.uint8)
    r14 = np.zeros(1 << 14, np.uint8)
    orc.hll_add_hashes(r16, 16, hs)
    orc.hll_add_hashes(r14, 14, hs)
    fold = np.zeros(1 << 14, np.uint8)
    for j in np.nonzero(r16)[0].tolist():
        low = j & 3
        v = int(r16[j])
        contrib = 1 if low >= 2 else (2 if low == 1 else 2 + v)
        fold[j >> 2] = max(fold[j >> 2], contrib)
    assert np.array_equal(fold, r14)


def _err(est, n):
    return abs(est - n) / n


def test_hll_reference_error_gates(orc):
    g = REF["hll_gates"]
    n = g["single_100k"]["n"]
    regs = np.zeros(1 << 14, np.uint8)
    orc.hll_add_hashes(regs, 14, [P.splitmix64(i) for i in range(n)])
    assert _err(orc.hll_estimate(regs, 14), n) < g["single_100k"]["max_err"]
    a = np.zeros(1 << 14, np.uint8)
    b = np.zeros(1 << 14, np.uint8)
    orc.hll_add_hashes(a, 14, [P.splitmix64(i) for i in range(g["merge"]["split"])])
    orc.hll_add_hashes(b, 14, [P.splitmix64(i) for i in range(g["merge"]["split"], g["merge"]["n"])])
    import ctypes as C
    orc.lib().orc_hll_merge(a.ctypes.data_as(C.POINTER(C.c_uint8)), b.ctypes.data_as(C.POINTER(C.c_uint8)), 14)
    assert _err(orc.hll_estimate(a, 14), g["merge"]["n"]) < g["merge"]["max_err"]
    rm = g["raw_min"]
    raw = np.zeros(1 << 14, np.uint8)
    fixed = np.zeros(1 << 14, np.uint8)
    mins = [min(P.splitmix64(rm["fwd_xor"] ^ i), P.splitmix64(rm["rev_xor"] ^ i)) for i in range(rm["n"])]
    orc.hll_add_hashes(raw, 14, mins)
    orc.hll_add_hashes(fixed, 14, [P.splitmix64(x) for x in mins])
    assert (orc.hll_estimate(raw, 14) - rm["n"]) / rm["n"] < rm["raw_err_below"]
    assert _err(orc.hll_estimate(fixed, 14), rm["n"]) <= rm["fixed_max_err"]


@pytest.mark.parametrize("n", REF["hll_gates"]["multi_scale"]["ns"][:3])
def test_hll_multi_scale_gate(orc, n):
    g = REF["hll_gates"]["multi_scale"]
    i = np.arange(n, dtype=np.uint64)
    f = _sm64_np(np.uint64(g["fwd_xor"]) ^ i)
    r = _sm64_np(np.uint64(g["rev_xor"]) ^ i)
    hs = _sm64_np(np.minimum(f, r))
    regs = np.zeros(1 << 14, np.uint8)
    orc.hll_add_hashes(regs, 14, hs)
    assert _err(orc.hll_estimate(regs, 14), n) <= g["max_err"]


def _sm64_np(z):
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def test_numpy_splitmix_matches(orc):
    xs = np.array([x for x, _ in PYV["splitmix64"]], np.uint64)
    assert _sm64_np(xs).tolist() == [y for _, y in PYV["splitmix64"]]


@pytest.mark.parametrize("bits", [4, 16])
def test_add_and_estimate_is_sequential_add_then_estimate(orc, bits):
    """add_and_estimate (sketch.rs:248-253) = add, then estimate: against a
    dictionary of exact per-cell counts updated one key at a time, on a
    narrow sketch (shared cells) with repeated keys, saturating at 15 / 65535."""
    width = 1 << 6
    cap = 15 if bits == 4 else 65535
    rng = np.random.default_rng(bits)
    keys = rng.integers(0, 40, size=3000).tolist() + [7] * 200
    c = orc.Cms(width, bits)
    counts = {}
    for key in keys:
        for cell in P.cms_cells(key, width):
            counts[cell] = counts.get(cell, 0) + 1
        want = min(min(counts[cell], cap) for cell in P.cms_cells(key, width))
        assert c.add_and_estimate(key) == want
    assert any(v > 15 for v in counts.values())
