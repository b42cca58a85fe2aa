"""Pins the ungated pass-2 form on the CPU against the oracle.

The reference (src/graph/build.rs:253-289) pushes a pair (u, v) of adjacent
windows only when both pass the Count-Min gate est >= t; pass 3 keeps a key
when one of its edge counters reaches m >= t (build.rs:302-328).  The
library's LDS route drops that per-window gate: it counts every pair, takes
the mask from the ungated counts, and subtracts from cov the low edges
(0 < count < t) whose neighbour fails the gate, probing the gate once per
kept key and low edge.  This test restates that algorithm with Python
dictionaries over the oracle's own Count-Min table (`Cms.estimate`) and
checks it against the oracle's per-window build, on sketches narrow enough
that every neighbour passes and wide enough that thousands fail; a planted
mistake (the correction left out) must break it.
"""
import numpy as np
import pytest

import fastbact_amd as F
from tests.helpers import oracle_arenas

CODE = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}


def _canon(x: int, k: int):
    r, y = 0, x
    for _ in range(k):
        r = (r << 2) | (3 - (y & 3))
        y >>= 2
    return (r, True) if r < x else (x, False)


def _fold(x: int) -> int:
    f = 0
    while x:
        f ^= x & 0xFFFFFFFFFFFFFFFF
        x >>= 64
    return f


def _neighbour(u: int, k: int, q: int) -> int:
    """Slot q < 4 appends base q to canonical u; slot 4 + b prepends 3 - b."""
    mask = (1 << (2 * k)) - 1
    y = ((u << 2) | q) & mask if q < 4 else (u >> 2) | ((7 - q) << (2 * k - 2))
    return _canon(y, k)[0]


def _reads(seed, k, glen=12_000, n=2500, err=0.01):
    rng = np.random.default_rng(seed)
    g = rng.choice(list(b"ACGT"), size=glen).astype(np.uint8)
    comp = bytes.maketrans(b"ACGT", b"TGCA")
    out = []
    for _ in range(n):
        L = int(rng.integers(k - 2, 140))
        s0 = int(rng.integers(0, glen - L))
        r = g[s0:s0 + L].copy()
        e = rng.random(L) < err
        r[e] = rng.choice(list(b"ACGT"), size=int(e.sum()))
        b = r.tobytes()
        out.append(b[::-1].translate(comp) if rng.random() < 0.5 else b)
    return out


def _per_key_gate(reads, k, t, m, est, correct=True):
    occ = {}
    for s in reads:
        if len(s) < k + 1:
            continue
        codes = [CODE[c] for c in s]
        mask = (1 << (2 * k)) - 1
        fw = 0
        for c in codes[:k]:
            fw = (fw << 2) | c
        u, urc = _canon(fw, k)
        for i in range(len(codes) - k):
            b = codes[i + k]
            fw = ((fw << 2) | b) & mask
            v, vrc = _canon(fw, k)
            p = codes[i]
            for key, e, other in ((u, b + 4 if urc else b, v), (v, (3 - p) if vrc else (3 - p) + 4, u)):
                occ.setdefault(key, [0] * 8)[e] += 1
                assert _neighbour(key, k, e) == other  # the slot names the neighbour
            u, urc = v, vrc
    gate = lambda x: est(_fold(x)) >= t
    out = {}
    for key, a in occ.items():
        msk = sum(1 << e for e in range(8) if a[e] >= m)
        if not msk:
            continue
        cov = sum(a)
        if correct:
            cov -= sum(a[e] for e in range(8) if 0 < a[e] < t and not gate(_neighbour(key, k, e)))
        out[key] = (max(min(cov, 65535) // 2, 1), msk)
    return out


@pytest.mark.parametrize("seed,k,target,width", [(1, 21, None, 1 << 11), (2, 21, None, 1 << 16), (3, 25, 4, 1 << 16),
                                                  (4, 31, 3, 1 << 10), (5, 15, 2, 1 << 14), (6, 17, 1, 1 << 12)])
def test_gate_per_kept_key_restated_equals_oracle(orc, seed, k, target, width):
    reads = _reads(seed, k)
    half = len(reads) // 2
    bank = F.ReadBank(F.BitSeqArena.from_ascii(reads[:half]), F.BitSeqArena.from_ascii(reads[half:]))
    ref = orc.build_solid(oracle_arenas(bank), k, target_solid_cov=target, width=width, counter_bits=16)
    t = 2 if target is None else target
    m = 3 if target is None else target
    got = _per_key_gate(reads, k, t, m, ref["cms"].estimate)
    want = {int(kk[0]): (int(c), int(mm)) for kk, c, mm in zip(ref["keys"], ref["cov"], ref["mask"])}
    assert got == want
    assert len(want) > 5000


def test_gate_correction_is_needed(orc):
    """Planted mistake: without the low-edge correction cov differs wherever a
    kept key has an error neighbour that the wide sketch leaves below the
    gate, so the comparison above can fail."""
    reads = _reads(2, 21)
    half = len(reads) // 2
    bank = F.ReadBank(F.BitSeqArena.from_ascii(reads[:half]), F.BitSeqArena.from_ascii(reads[half:]))
    ref = orc.build_solid(oracle_arenas(bank), 21, width=1 << 16, counter_bits=16)
    wrong = _per_key_gate(reads, 21, 2, 3, ref["cms"].estimate, correct=False)
    want = {int(kk[0]): (int(c), int(mm)) for kk, c, mm in zip(ref["keys"], ref["cov"], ref["mask"])}
    assert set(wrong) == set(want)  # the mask and the kept keys never depend on the gate
    assert sum(wrong[key] != want[key] for key in want) > 100
