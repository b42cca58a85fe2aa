"""Input builders shared by the tests (seeded, deterministic)."""
from __future__ import annotations

import numpy as np

import fastbact_amd as F


def isolate(seed=11, genome_len=40_000, n_pairs=3000, **kw):
    kw.setdefault("sub_rate", 0.005)
    return F.synth_isolate(n_pairs=n_pairs, seed=seed, genome_len=genome_len, **kw)


def ragged_bank(seed=5, n=400, max_len=260, n_frac=0.02):
    """Reads of assorted lengths (0, 1, k-1, k, k+1, ..., max_len) in all three
    arenas, with some non-ACGT bytes (mapped to A, bitseq.rs:67-81)."""
    rng = np.random.default_rng(seed)
    alpha = np.frombuffer(b"ACGTacgtN", dtype=np.uint8)
    genome = rng.choice(np.frombuffer(b"ACGT", np.uint8), size=5000)
    arenas = []
    fixed = [0, 1, 20, 21, 22, 30, 31, 32, 33, 54, 55, 56, 62, 63, 64, 65, 76, 77, 78, 126, 127, 128, 129]
    for a in range(3):
        seqs = []
        for i in range(n):
            L = fixed[i] if i < len(fixed) else int(rng.integers(0, max_len + 1))
            start = int(rng.integers(0, genome.size - max_len - 1))
            s = genome[start:start + L].copy()
            # lowercase / N noise
            m = rng.random(L) < n_frac
            s[m] = rng.choice(alpha, size=int(m.sum()))
            seqs.append(s.tobytes())
        arenas.append(F.BitSeqArena.from_ascii(seqs))
    return F.ReadBank(*arenas)


def oracle_arenas(bank: F.ReadBank):
    return [(a.data, a.offsets, a.lengths) for a in bank.arenas()]
