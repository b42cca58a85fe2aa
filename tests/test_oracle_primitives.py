"""Pins the oracle's primitives: 2-bit arena, KmerInt (1/2/4 words), ntHash,
splitmix64 -- against the reference tests' known answers, the independent
big-integer restatement (tests/pyref.py) and brute force on strings."""
import json
import random
from pathlib import Path

import numpy as np
import pytest

from tests import pyref as P

GOLD = Path(__file__).resolve().parent / "golden"
REF = json.loads((GOLD / "reference_vectors.json").read_text())
PYV = json.loads((GOLD / "pyref_vectors.json").read_text())


def test_splitmix64_vectors(orc):
    assert orc.splitmix64(0) == 0  # no gamma add: f(0) = 0
    for x, y in PYV["splitmix64"]:
        assert orc.splitmix64(x) == y


def test_rev_comp_word_reference_vectors(orc):
    for c in REF["rev_comp_word"]["cases"]:
        assert orc.rev_comp_word(c["word"], c["bases"]) == c["expect"]


def test_rev_comp_word_brute_force(orc):
    rnd = random.Random(1)
    for nb in range(0, 33):
        s = "".join(rnd.choice("ACGT") for _ in range(nb))
        # arena order: base i at bits 2i
        w = sum("ACGT".index(c) << (2 * i) for i, c in enumerate(s))
        rc = P.revcomp(s)
        expect = sum("ACGT".index(c) << (2 * i) for i, c in enumerate(rc))
        assert orc.rev_comp_word(w, nb) == expect


def test_reverse_complement_reference(orc):
    for s, rc in REF["reverse_complement_ascii"]["cases"]:
        assert P.revcomp(s) == rc
        x = orc.kmer_from_ascii(s.encode())
        assert np.array_equal(orc.kmer_rev_comp(x, len(s)), orc.kmer_from_ascii(rc.encode()))


def test_arena_pack_extract_lengths_1_to_500(orc):
    # bitseq.rs:473-506: pack/unpack every length 1..500 with the reference's LCG
    rng = 0x123456789ABCDEF
    seqs = []
    for ln in range(1, 501):
        s = []
        for _ in range(ln):
            rng = (rng * 6364136223846793005 + 1442695040888963407) & P.M64
            s.append("ACGT"[rng % 4])
        seqs.append("".join(s).encode())
    data, offs, lens = orc.pack(seqs)
    a = orc.arena(data, offs, lens)
    import ctypes as C
    for i, s in enumerate(seqs):
        buf = np.zeros(len(s), np.uint8)
        orc.lib().orc_arena_extract_ascii(C.byref(a), i, buf.ctypes.data_as(C.POINTER(C.c_uint8)))
        assert buf.tobytes() == s


def test_ascii_to_2bit_non_acgt_is_a(orc):
    for c in range(256):
        ch = chr(c)
        expect = "ACGT".index(ch.upper()) if ch.upper() in "ACGT" and ch.isalpha() else 0
        assert orc.lib().orc_ascii_to_2bit(c) == expect


@pytest.mark.parametrize("name,k", [("u64", 24), ("u128", 56), ("u256", 96)])
def test_kmer_basic_round_trip(orc, name, k):
    # kmer.rs:543-577
    s = REF["kmer_basic"][name]
    assert len(s) == k
    x = orc.kmer_from_ascii(s.encode())
    assert P.ref_words(P.kmer_int(s), k) == [int(v) for v in x]
    rc = orc.kmer_rev_comp(x, k)
    assert [int(v) for v in rc] == P.ref_words(P.kmer_int(P.revcomp(s)), k)


def test_canonical_windows_vs_bigint(orc):
    for case in PYV["canonical"]:
        k, s = case["k"], case["seq"]
        wins, rcs = orc.canonical_windows(s.encode(), k)
        for i, words, rc, fold in case["windows"]:
            assert [int(v) for v in wins[i]] == words
            assert int(rcs[i]) == rc
            assert orc.kmer_to_u64(wins[i]) == fold


def test_rolling_gate_fwd_equals_rc(orc):
    # kmer.rs:580-651: canonical k-mers of a sequence and of its reverse
    # complement agree window by window (k = 21/55/99/127)
    g = REF["rolling_gate"]
    rng, seq = g["seed"], []
    for _ in range(g["length"]):
        rng = (rng * g["mul"] + g["add"]) & P.M64
        seq.append("ACGT"[rng % 4])
    s = "".join(seq)
    for k in g["ks"]:
        a, _ = orc.canonical_windows(s.encode(), k)
        b, _ = orc.canonical_windows(P.revcomp(s).encode(), k)
        assert np.array_equal(a, b[::-1])


@pytest.mark.parametrize("k", [1, 2, 16, 31, 32, 33, 62, 63, 64, 65, 96, 127])
def test_roll_next_equals_fresh(orc, k):
    rnd = random.Random(k)
    s = "".join(rnd.choice("ACGT") for _ in range(k + 40))
    nw = orc.words_for_k(k)
    fwd = orc.kmer_from_ascii(s[:k].encode(), nw)
    rev = orc.kmer_rev_comp(fwd, k)
    for i in range(k, len(s)):
        fwd, rev = orc.kmer_roll_next(fwd, rev, "ACGT".index(s[i]), k)
        w = s[i - k + 1:i + 1]
        assert [int(v) for v in fwd] == P.ref_words(P.kmer_int(w), k)
        assert [int(v) for v in rev] == P.ref_words(P.kmer_int(P.revcomp(w)), k)


def test_nthash_vectors_and_rolling(orc):
    for v in PYV["nthash"]:
        s = v["seq"].encode()
        assert orc.nthash_fwd(s) == v["fwd"]
        assert orc.nthash_rev(s) == v["rev"]
        assert orc.nthash_canonical(s) == v["can"]
    g = REF["nthash_rolling"]  # kmer.rs:653-664: roll == fresh
    s, k = g["seq"], g["k"]
    rolled = orc.nthash_rolling(s.encode(), k)
    for i in range(len(s) - k + 1):
        assert int(rolled[i]) == orc.nthash_canonical(s[i:i + k].encode())
    rnd = random.Random(7)
    for k in (21, 63, 64, 65, 127):
        s = "".join(rnd.choice("ACGT") for _ in range(k + 60))
        rolled = orc.nthash_rolling(s.encode(), k)
        assert [int(x) for x in rolled] == [P.nthash_canonical(s[i:i + k]) for i in range(len(s) - k + 1)]


def test_nthash_canonical_strand_symmetric(orc):
    rnd = random.Random(3)
    for k in (21, 31, 77):
        s = "".join(rnd.choice("ACGT") for _ in range(k))
        assert orc.nthash_canonical(s.encode()) == orc.nthash_canonical(P.revcomp(s).encode())


def test_planted_fold_mistake_is_caught():
    # a dropped limb in the fold must disagree with the big-int fold
    x = P.kmer_int("ACGT" * 20)
    assert P.fold(x) != (x & P.M64)
