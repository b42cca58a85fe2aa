"""The oracle's ingest steps (trim_merge.rs:17-417) pinned on the CPU: the
reference's own tests (a 230 bp fragment merged through two errors, a 450 bp
fragment left unmerged, adapter clipping), closed forms of the quality trim,
and a string restatement (tests/pyref.py) on random pairs with adapters,
low-quality tails, lowercase and N bytes."""
import numpy as np
import pytest

from tests import pyref as P

ALPHA = "ACGT"


def _frag(seed, n):
    return "".join(ALPHA[P.splitmix64(seed + i) % 4] for i in range(n))


def test_reference_merge_gate_230bp_with_two_errors(orc):
    """trim_merge.rs:423-471"""
    frag = _frag(0x5EED_CAFE_1234_5678, 230)
    r1 = list(frag[:150])
    q1 = ["I"] * 150
    r1[130] = "C" if r1[130] == "A" else "A"
    q1[130] = "!"
    r2 = list(P.revcomp(frag[80:230]))
    q2 = ["I"] * 150
    r2[129] = "T" if r2[129] == "A" else "A"
    q2[129] = "!"
    r1, q1, r2, q2 = ("".join(x).encode() for x in (r1, q1, r2, q2))
    seq, d, m = orc.merge_read_pair(r1, q1, r2, q2, 15, 0.10)
    assert (len(seq), d, m) == (230, 70, 2)
    assert seq == frag.encode()


def test_reference_long_fragment_stays_unmerged(orc):
    """trim_merge.rs:473-495"""
    frag = _frag(0xABCD_EF01_2345_6789, 450)
    assert orc.merge_read_pair(frag[:150].encode(), b"I" * 150, P.revcomp(frag[300:450]).encode(), b"I" * 150) is None


def test_reference_adapter_trimming(orc):
    """trim_merge.rs:497-507"""
    seq = b"ACGTACGTACGTACGT" + b"AGATCGGAAGAGCextra"
    assert orc.trim_adapter(seq) == 16


def test_adapter_edges(orc):
    base = _frag(11, 200)
    for ad in ("AGATCGGAAGAGC", "AGATGTGTATAAGAGACAG", "CTGTCTCTTATACACATCT"):
        for at in (95, 100, 101, 150, 190):
            s = (base[:at] + ad + base)[:200]
            want = P.trim_adapter(s)
            assert orc.trim_adapter(s.encode()) == want
            if at >= 100 and at <= 190:
                assert want <= at
    # one mismatch in the first 10 bases is tolerated, two are not
    s1 = base[:150] + "AGTTCGGAAGAGC" + base[:37]
    s2 = base[:150] + "AGTTCGCAAGAGC" + base[:37]
    assert orc.trim_adapter(s1.encode()) <= 150 and orc.trim_adapter(s2.encode()) == P.trim_adapter(s2)
    assert orc.trim_adapter(b"AGATCGGAA") == 9  # shorter than 10: untouched


@pytest.mark.parametrize("min_qual,min_len", [(20, 50), (30, 1), (0, 4), (40, 100)])
def test_quality_trim_closed_forms_and_restatement(orc, min_qual, min_len):
    rng = np.random.default_rng(min_qual * 7 + min_len)
    # all high: kept whole
    assert orc.trim_quality(b"I" * 150, min_qual, min_len) == (150 if min_len <= 150 else None)
    # a tail of Phred 2: cut at the last window that still averages min_qual
    q = b"I" * 120 + b"#" * 30
    want = P.trim_quality(q.decode(), min_qual, min_len)
    assert orc.trim_quality(q, min_qual, min_len) == want
    if min_qual >= 20 and min_len <= 120:
        # window [cut-4, cut) must average >= min_qual: 40*a + 2*(4-a) >= 4*min_qual
        a = next(a for a in range(5) if 40 * a + 2 * (4 - a) >= 4 * min_qual)
        assert want == 120 + (4 - a)
    for _ in range(200):
        n = int(rng.integers(0, 160))
        qs = bytes(int(x) for x in rng.integers(33, 75, size=n))
        if rng.random() < 0.2:
            qs = qs[:max(n - 20, 0)] + bytes(int(x) for x in rng.integers(20, 40, size=min(20, n)))
        assert orc.trim_quality(qs, min_qual, min_len) == P.trim_quality(qs.decode("latin1"), min_qual, min_len)
    assert orc.trim_quality(b"IIII", min_qual, min_len, seq_len=5) is None  # lengths differ


def _mutate(s, rng, rate, alphabet="ACGTacgtN"):
    out = list(s)
    for i in np.flatnonzero(rng.random(len(s)) < rate):
        out[i] = alphabet[int(rng.integers(0, len(alphabet)))]
    return "".join(out)


def test_merge_matches_string_restatement(orc):
    rng = np.random.default_rng(5)
    merged = unmerged = 0
    for t in range(400):
        flen = int(rng.integers(60, 420))
        frag = _frag(1000 * t + 17, flen)
        l1, l2 = int(rng.integers(20, 160)), int(rng.integers(20, 160))
        r1 = _mutate(frag[:l1], rng, 0.02)
        r2 = _mutate(P.revcomp(frag[max(flen - l2, 0):]), rng, 0.02)
        q1 = "".join(chr(int(x)) for x in rng.integers(35, 74, size=int(rng.integers(0, len(r1) + 1))))
        q2 = "".join(chr(int(x)) for x in rng.integers(35, 74, size=len(r2)))
        got = orc.merge_read_pair(r1.encode(), q1.encode(), r2.encode(), q2.encode())
        want = P.merge_pair(r1, q1, r2, q2)
        if want is None:
            assert got is None
            unmerged += 1
        else:
            assert got == (want[0].encode(), want[1], want[2])
            merged += 1
    assert merged > 50 and unmerged > 50


def _pairs(seed, n, L=150, insert=(120, 420), adapter_rate=0.1, lowq_rate=0.1):
    rng = np.random.default_rng(seed)
    genome = _frag(seed * 1_000_003, 20_000)
    r1s, q1s, r2s, q2s = [], [], [], []
    for _ in range(n):
        ins = int(rng.integers(*insert))
        a = int(rng.integers(0, len(genome) - ins))
        frag = genome[a:a + ins]
        r1 = frag[:L]
        r2 = P.revcomp(frag[-L:]) if ins >= L else P.revcomp(frag)
        if ins < L:  # read-through: adapter after the insert
            r1 = (frag + "AGATCGGAAGAGCACACGTCTGAACTCCAGTCA" * 3)[:L]
            r2 = (P.revcomp(frag) + "AGATCGGAAGAGCGTCGTGTAGGGAAAGAGTGT" * 3)[:L]
        r1, r2 = _mutate(r1, rng, 0.005), _mutate(r2, rng, 0.005)
        q1 = [int(x) for x in rng.integers(58, 74, size=len(r1))]
        q2 = [int(x) for x in rng.integers(58, 74, size=len(r2))]
        if rng.random() < lowq_rate:
            t = int(rng.integers(10, 100))
            q1[-t:] = [int(x) for x in rng.integers(35, 45, size=t)]
        if rng.random() < adapter_rate:
            at = int(rng.integers(60, L - 12))
            r2 = (r2[:at] + "CTGTCTCTTATACACATCT" + r2)[:len(r2)]
        r1s.append(r1.encode())
        r2s.append(r2.encode())
        q1s.append(bytes(q1))
        q2s.append(bytes(q2))
    return r1s, q1s, r2s, q2s


def _pyref_ingest(r1s, q1s, r2s, q2s, factor=1.0, min_len=50, min_qual=20, trim=True, merge=True, seed=0):
    thr = int(factor * 2.0 ** 64) if factor < 0.999 else None
    reads = [[], [], []]
    st = dict(input_read_pairs=0, retained_read_pairs=0, bases_before_trim=0, bases_after_trim=0,
              merged_pairs_count=0, merged_bases=0, unmerged_pairs_count=0)
    for i, (r1, q1, r2, q2) in enumerate(zip(r1s, q1s, r2s, q2s)):
        st["input_read_pairs"] += 1
        st["bases_before_trim"] += len(r1) + len(r2)
        if thr is not None and P.splitmix64((i + 1) ^ seed) > thr:
            continue
        s1, s2, a1, a2 = r1.decode(), r2.decode(), q1.decode("latin1"), q2.decode("latin1")
        if trim:
            c1, c2 = P.trim_adapter(s1), P.trim_adapter(s2)
            t1 = P.trim_quality(a1[:c1], min_qual, min_len)
            t2 = P.trim_quality(a2[:c2], min_qual, min_len)
            if t1 is None or t2 is None:
                continue
            s1, a1, s2, a2 = s1[:t1], a1[:t1], s2[:t2], a2[:t2]
        elif len(s1) < min_len or len(s2) < min_len:
            continue
        st["retained_read_pairs"] += 1
        st["bases_after_trim"] += len(s1) + len(s2)
        m = P.merge_pair(s1, a1, s2, a2) if merge else None
        if m is not None:
            st["merged_pairs_count"] += 1
            st["merged_bases"] += len(m[0])
            reads[0].append(m[0].encode())
        else:
            st["unmerged_pairs_count"] += 1
            reads[1].append(s1.encode())
            reads[2].append(s2.encode())
    return reads, st


@pytest.mark.parametrize("factor,trim,merge,seed", [(1.0, True, True, 0), (0.5, True, True, 7),
                                                    (1.0, False, True, 0), (1.0, True, False, 0),
                                                    (0.25, False, False, 99)])
def test_ingest_records_match_restatement(orc, factor, trim, merge, seed):
    r1s, q1s, r2s, q2s = _pairs(3 + seed, 600)
    out = orc.ingest_records(orc.records(r1s, q1s), orc.records(r2s, q2s), subsample_factor=factor,
                             enable_trim=trim, enable_merge=merge, seed=seed)
    reads, st = _pyref_ingest(r1s, q1s, r2s, q2s, factor, trim=trim, merge=merge, seed=seed)
    assert out["reads"] == reads
    assert out["stats"] == st
    assert st["merged_pairs_count"] + st["unmerged_pairs_count"] == st["retained_read_pairs"]
    if merge:
        assert st["merged_pairs_count"] > 20
    if factor < 1:
        assert abs(st["retained_read_pairs"] / 600 - factor) < 0.12 or not trim
    # the arenas hold the reads packed as push_seq packs them
    for (data, offs, lens), rs in zip(out["arenas"], reads):
        assert orc.unpack(data, offs, lens) == [P.norm(r).encode() for r in rs]


def test_ingest_error_paths(orc):
    r1s, q1s, r2s, q2s = _pairs(1, 10)
    with pytest.raises(ValueError, match="Mismatched"):
        orc.ingest_records(orc.records(r1s, q1s), orc.records(r2s[:9], q2s[:9]))
    with pytest.raises(ValueError, match="No valid reads"):
        orc.ingest_records(orc.records(r1s, q1s), orc.records(r2s, q2s), min_len=1000)
    # FASTA records (no qualities) with trimming: the reference's slice panics
    with pytest.raises(ValueError, match="quality"):
        orc.ingest_records(orc.records(r1s), orc.records(r2s))
    # without trimming they merge with 'I' for the missing qualities
    out = orc.ingest_records(orc.records(r1s), orc.records(r2s), enable_trim=False)
    reads, st = _pyref_ingest(r1s, [b""] * 10, r2s, [b""] * 10, trim=False)
    assert out["reads"] == reads and out["stats"] == st
