"""GPU parity of the ingest front end (trim_merge.rs:17-417): the device
record loop gives the oracle's bank byte for byte (merged reads, unmerged
mates, BitSeqArena packing) and the oracle's IngestStats, through records and
through FASTQ / FASTQ.gz / FASTA files; the reference's own merge and adapter
cases hold on the device; error paths match."""
import gzip

import numpy as np
import pytest

import fastbact_amd as F
from tests import pyref as P
from tests.test_oracle_ingest import _frag, _pairs

pytestmark = pytest.mark.gpu

STATS = ("input_read_pairs", "retained_read_pairs", "bases_before_trim", "bases_after_trim", "merged_pairs_count",
         "merged_bases", "unmerged_pairs_count")


def _same_bank(bank: F.DeviceBank, arenas):
    got = bank.download()
    for a, (data, offs, lens) in zip(got.arenas(), arenas):
        nw = (int(lens.sum(dtype=np.uint64)) + 31) // 32
        assert np.array_equal(a.lengths, lens)
        assert np.array_equal(a.offsets, offs)
        assert np.array_equal(a.data, data[:nw])


def _both(ctx, orc, r1s, q1s, r2s, q2s, **kw):
    bank, st = F.ingest_records(ctx, list(zip(r1s, q1s)), list(zip(r2s, q2s)), **kw)
    okw = dict(subsample_factor=kw.get("subsample_factor", 1.0), min_len=kw.get("min_len", 30),
               min_qual=kw.get("min_qual", 20), enable_trim=kw.get("trim", True),
               enable_merge=kw.get("merge", True), seed=kw.get("seed", 0))
    ref = orc.ingest_records(orc.records(r1s, q1s), orc.records(r2s, q2s), **okw)
    _same_bank(bank, ref["arenas"])
    assert {k: st[k] for k in STATS} == ref["stats"]
    return bank, st, ref


@pytest.mark.parametrize("factor,trim,merge,seed,min_len,min_qual", [
    (1.0, True, True, 0, 30, 20), (0.5, True, True, 7, 30, 20), (1.0, False, True, 0, 30, 20),
    (1.0, True, False, 0, 50, 30), (0.25, False, False, 99, 1, 0), (0.9, True, True, 3, 100, 25)])
def test_records_match_oracle(ctx, orc, factor, trim, merge, seed, min_len, min_qual):
    r1s, q1s, r2s, q2s = _pairs(11 + seed, 3000)
    _, st, _ = _both(ctx, orc, r1s, q1s, r2s, q2s, subsample_factor=factor, trim=trim, merge=merge, seed=seed,
                     min_len=min_len, min_qual=min_qual)
    if merge:
        assert st["merged_pairs_count"] > 100
    assert st["unmerged_pairs_count"] > 100


def test_reference_cases_on_device(ctx, orc):
    """trim_merge.rs:423-507 through the device: the 230 bp fragment merges
    through two errors, the 450 bp one does not, an adapter is clipped."""
    frag = _frag(0x5EED_CAFE_1234_5678, 230)
    r1 = list(frag[:150])
    q1 = [ord("I")] * 150
    r1[130] = "C" if r1[130] == "A" else "A"
    q1[130] = ord("!")
    r2 = list(P.revcomp(frag[80:230]))
    q2 = [ord("I")] * 150
    r2[129] = "T" if r2[129] == "A" else "A"
    q2[129] = ord("!")
    long = _frag(0xABCD_EF01_2345_6789, 450)
    ad = b"ACGTACGTACGTACGTAGATCGGAAGAGCextra" + b"C" * 40
    pairs1 = [("".join(r1).encode(), bytes(q1)), (long[:150].encode(), b"I" * 150), (ad, b"I" * len(ad))]
    pairs2 = [("".join(r2).encode(), bytes(q2)), (P.revcomp(long[300:450]).encode(), b"I" * 150),
              (b"G" * 60, b"I" * 60)]
    bank, st = F.ingest_records(ctx, pairs1, pairs2, min_len=10)
    got = bank.download()
    merged = [bytes(x) for x in
              F_unpack(got.merged_reads)]
    assert merged == [frag.encode()]
    assert F_unpack(got.pe_forward) == [long[:150].encode(), b"ACGTACGTACGTACGT"]
    assert st["merged_pairs_count"] == 1 and st["unmerged_pairs_count"] == 2


def F_unpack(arena: F.BitSeqArena) -> list:
    out = []
    for o, n in zip(arena.offsets, arena.lengths):
        g = int(o) + np.arange(int(n), dtype=np.uint64)
        codes = (arena.data[(g >> np.uint64(5)).astype(np.int64)] >> ((g & np.uint64(31)) << np.uint64(1))) & np.uint64(3)
        out.append(np.frombuffer(b"ACGT", np.uint8)[codes.astype(np.int64)].tobytes())
    return out


def _write_fastq(path, seqs, quals, gz=False, crlf=False):
    nl = b"\r\n" if crlf else b"\n"
    body = b"".join(b"@p%d/x%s%s%s+%s%s%s" % (i, nl, s, nl, nl, q, nl) for i, (s, q) in enumerate(zip(seqs, quals)))
    path.write_bytes(gzip.compress(body) if gz else body)


@pytest.mark.parametrize("gz,crlf", [(False, False), (True, False), (True, True)])
def test_files_match_records_and_oracle(ctx, orc, tmp_path, gz, crlf):
    r1s, q1s, r2s, q2s = _pairs(21, 2500)
    ext = ".fq.gz" if gz else ".fq"
    p1, p2 = tmp_path / ("r1" + ext), tmp_path / ("r2" + ext)
    _write_fastq(p1, r1s, q1s, gz, crlf)
    _write_fastq(p2, r2s, q2s, gz, crlf)
    bank, st = F.ingest_files(ctx, p1, p2, seed=5, subsample_factor=0.8)
    ref = orc.ingest_records(orc.records(r1s, q1s), orc.records(r2s, q2s), subsample_factor=0.8, seed=5,
                             min_len=30, min_qual=20)
    _same_bank(bank, ref["arenas"])
    assert {k: st[k] for k in STATS} == ref["stats"]
    assert st["parse_ms"] > 0 and st["device_ms"] > 0


@pytest.mark.parametrize("L,insert,n", [(250, (180, 700), 2000), (1020, (700, 2400), 300), (1100, (800, 2600), 300)])
def test_long_mates_match_oracle(ctx, orc, L, insert, n):
    """Mates up to 1024 bases are merged with every overlap counted at once
    from bit planes; longer ones take the overlap-by-overlap search.  Both
    give the oracle's bank, with overlaps of 15 up to the whole mate."""
    r1s, q1s, r2s, q2s = _pairs(31 + L, n, L=L, insert=insert)
    _, st, _ = _both(ctx, orc, r1s, q1s, r2s, q2s)
    assert st["merged_pairs_count"] > n // 10 and st["unmerged_pairs_count"] > n // 10


def test_scale_100k_pairs(ctx, orc):
    rng = np.random.default_rng(77)
    genome = rng.choice(np.frombuffer(b"ACGT", np.uint8), size=300_000)
    n, L = 100_000, 150
    ins = rng.integers(100, 500, size=n)
    start = rng.integers(0, genome.size - 600, size=n)
    comp = np.zeros(256, np.uint8)
    for a, b in zip(b"ACGT", b"TGCA"):
        comp[a] = b
    r1s, r2s, q1s, q2s = [], [], [], []
    for s, i in zip(start, ins):
        frag = genome[s:s + i]
        r1 = frag[:L]
        r2 = comp[frag[::-1]][:L]
        r1s.append(r1.tobytes())
        r2s.append(r2.tobytes())
        q = rng.integers(50, 74, size=2 * L).astype(np.uint8)
        q1s.append(q[:len(r1)].tobytes())
        q2s.append(q[L:L + len(r2)].tobytes())
    _, st, _ = _both(ctx, orc, r1s, q1s, r2s, q2s, subsample_factor=0.7, seed=1)
    assert st["merged_pairs_count"] > 10_000 and st["unmerged_pairs_count"] > 10_000


def test_errors(ctx, tmp_path):
    r1s, q1s, r2s, q2s = _pairs(1, 10)
    with pytest.raises(F.FastBactError, match="Mismatched"):
        F.ingest_records(ctx, list(zip(r1s, q1s)), list(zip(r2s[:9], q2s[:9])))
    with pytest.raises(F.FastBactError, match="No valid reads"):
        F.ingest_records(ctx, list(zip(r1s, q1s)), list(zip(r2s, q2s)), min_len=1000)
    # FASTA mates: trimming needs qualities; without trimming they merge with 'I'
    fa1, fa2 = tmp_path / "a.fa", tmp_path / "b.fa"
    fa1.write_bytes(b"".join(b">r\n%s\n" % s for s in r1s))
    fa2.write_bytes(b"".join(b">r\n%s\n" % s for s in r2s))
    with pytest.raises(F.FastBactError, match="quality"):
        F.ingest_files(ctx, fa1, fa2)
    bank, st = F.ingest_files(ctx, fa1, fa2, trim=False)
    assert st["retained_read_pairs"] == 10
    with pytest.raises(F.FastBactError, match="Failed to open"):
        F.ingest_files(ctx, tmp_path / "none.fq", fa2)


@pytest.mark.parametrize("max_pairs,gz", [(150_000, True), (7_000, False)])
def test_estimate_genome_from_files_matches_oracle(ctx, orc, tmp_path, max_pairs, gz):
    """estimate_genome on FASTQ files (count_fastx_exact, then the sketch
    loop over the first pairs) equals the oracle's on the same reads."""
    from tests.helpers import isolate
    from tests.test_gpu_parity import _est_equal
    iso = isolate(seed=41, genome_len=200_000, n_pairs=20_000, sub_rate=0.004)
    n, L = iso.r1.shape
    ext = ".fq.gz" if gz else ".fq"
    p1, p2 = tmp_path / ("e1" + ext), tmp_path / ("e2" + ext)
    q = b"I" * L
    _write_fastq(p1, [bytes(r) for r in iso.r1], [q] * n, gz)
    _write_fastq(p2, [bytes(r) for r in iso.r2], [q] * n, gz)
    g = F.estimate_genome_files(ctx, p1, p2, max_pairs, 100.0)
    bank = iso.bank()
    r1, r2 = bank.pe_forward, bank.pe_reverse
    st = orc.count_stats(r1.lengths, r2.lengths)
    o = orc.estimate_genome((r1.data, r1.offsets, r1.lengths), (r2.data, r2.offsets, r2.lengths), st, max_pairs,
                            100.0)
    _est_equal(g, o)
