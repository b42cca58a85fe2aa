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


def head_arena(a: F.BitSeqArena, n: int) -> F.BitSeqArena:
    """The first n reads of an arena (its words up to the last of them)."""
    if n >= len(a):
        return a
    end = int(a.offsets[n - 1] + a.lengths[n - 1]) if n else 0
    return F.BitSeqArena(a.data[: (end + 31) // 32 + 1].copy(), a.offsets[:n].copy(), a.lengths[:n].copy())


def head_pairs(bank: F.ReadBank, n: int) -> F.ReadBank:
    """The first n read pairs of a bank (merged reads dropped)."""
    return F.ReadBank(F.BitSeqArena(), head_arena(bank.pe_forward, n), head_arena(bank.pe_reverse, n))


def distinct_canonical_kmers(reads2d, k: int) -> int:
    """Exact number of distinct canonical k-mers (k <= 31) over equal-length
    ACGT reads given as 2-D byte arrays, by sorting their 2k-bit integers
    (first base most significant; canonical = min(forward, reverse
    complement))."""
    lut = np.zeros(256, np.uint64)
    for ch, v in zip(b"ACGT", range(4)):
        lut[ch] = v
    keys = []
    for arr in reads2d:
        codes = lut[np.asarray(arr)]
        n, L = codes.shape
        w = L - k + 1
        fwd = np.zeros((n, w), np.uint64)
        rev = np.zeros((n, w), np.uint64)
        for j in range(k):
            c = codes[:, j:j + w]
            fwd = (fwd << np.uint64(2)) | c
            rev |= (np.uint64(3) - c) << np.uint64(2 * j)
        keys.append(np.minimum(fwd, rev).ravel())
    return int(np.unique(np.concatenate(keys)).size)


def hll_fold_p16_to_p14(r16: np.ndarray) -> np.ndarray:
    """p=14 registers implied by p=16 registers of the same hashes: register
    j of p=16 lands in j >> 2, and the two index bits that move into the p=14
    rank word give rank 1 (top bit set), 2 (next bit set) or 2 + the p=16
    rank (both clear)."""
    j = np.nonzero(r16)[0]
    low = j & 3
    v = r16[j].astype(np.int64)
    contrib = np.where(low >= 2, 1, np.where(low == 1, 2, 2 + v))
    out = np.zeros(1 << 14, np.int64)
    np.maximum.at(out, j >> 2, contrib)
    return out.astype(np.uint8)


def nibble_pack(u16_table: np.ndarray) -> np.ndarray:
    """The reference's nibble layout (byte idx >> 1, high nibble for odd idx,
    sketch.rs:191-223) of min(u16 counter, 15)."""
    c = np.minimum(u16_table, 15).astype(np.uint8)
    return c[0::2] | (c[1::2] << 4)
