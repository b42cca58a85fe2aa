"""The host FASTQ/FASTA reader of the ingest front end (open_fastx and the
record reader, trim_merge.rs:233-252): records as written, gzip (one or
several members), CRLF line ends, multi-line FASTA, and the error paths.
Host code only: no GPU call."""
import gzip

import pytest

import fastbact_amd as F

RECS = [(b"ACGTNacgtn", b"IIIII#####"), (b"", b""), (b"GATTACA" * 30, b"F" * 210), (b"T", b"!")]


def _fastq(recs, nl=b"\n", plus=b"+"):
    return b"".join(b"@r%d desc%s%s%s%s%s%s%s" % (i, nl, s, nl, plus, nl, q, nl) for i, (s, q) in enumerate(recs))


def test_fastq_plain_gzip_and_crlf(tmp_path):
    p = tmp_path / "a.fq"
    p.write_bytes(_fastq(RECS))
    assert F.read_fastx(p) == RECS
    g = tmp_path / "a.fq.gz"
    g.write_bytes(gzip.compress(_fastq(RECS)))
    assert F.read_fastx(g) == RECS
    # two gzip members, CRLF, a "+header" line, no final newline
    m = tmp_path / "m.fastq.gz"
    body = _fastq(RECS[:2], nl=b"\r\n", plus=b"+r0")
    m.write_bytes(gzip.compress(body) + gzip.compress(_fastq(RECS[2:])[:-1]))
    assert F.read_fastx(m) == RECS


def test_fasta_multiline(tmp_path):
    p = tmp_path / "a.fa"
    p.write_bytes(b">x one\nACGT\nacgtN\r\n>y\n>z\nGG\nC")
    assert F.read_fastx(p) == [(b"ACGTacgtN", b""), (b"", b""), (b"GGC", b"")]


@pytest.mark.parametrize("body,msg", [
    (b"", "empty"),
    (b"ACGT\n", "unknown format"),
    (b"@r\nACGT\n+\nIII\n", "lengths differ"),
    (b"@r\nACGT\n+\nIIII\nACGT\n", "expected '@'"),
    (b"@r\nACGT\n", "truncated"),
    (b"@r\nACGT\nIIII\nIIII\n", "truncated"),
])
def test_errors(tmp_path, body, msg):
    p = tmp_path / "bad.fq"
    p.write_bytes(body)
    with pytest.raises(F.FastBactError, match=msg):
        F.read_fastx(p)
    with pytest.raises(F.FastBactError, match="Failed to open"):
        F.read_fastx(tmp_path / "missing.fq")


# ---- count_fastx_exact (estimator.rs:167-283) on files ----------------------

def _pairs_files(tmp_path, r1_lines, r2_lines):
    a, b = tmp_path / "x_1.fq", tmp_path / "x_2.fq"
    a.write_text("".join(r1_lines))
    b.write_text("".join(r2_lines))
    return a, b


def test_count_fastx_reference_cases(tmp_path):
    """estimator.rs:463-528: ten 12 bp pairs; R1 longer than R2; desynchronised IDs."""
    a, b = _pairs_files(tmp_path, [f"@pair_{i}/1\nACGTACGTACGT\n+\nIIIIIIIIIIII\n" for i in range(1, 11)],
                        [f"@pair_{i}/2\nTGCATGCATGCA\n+\nIIIIIIIIIIII\n" for i in range(1, 11)])
    st = F.count_fastx(a, b)
    assert (st.exact_read_pairs, st.total_bases_r1, st.total_bases_r2, st.total_bases) == (10, 120, 120, 240)
    assert (st.min_read_len, st.max_read_len, st.median_read_len) == (12, 12, 12)
    assert abs(st.mean_read_len - 12.0) < 1e-6
    a, b = _pairs_files(tmp_path, [f"@pair_{i}/1\nACGT\n+\nIIII\n" for i in range(1, 6)],
                        [f"@pair_{i}/2\nACGT\n+\nIIII\n" for i in range(1, 5)])
    with pytest.raises(F.FastBactError, match="R1 has more reads than R2"):
        F.count_fastx(a, b)
    with pytest.raises(F.FastBactError, match="R2 has more reads than R1"):
        F.count_fastx(b, a)
    a, b = _pairs_files(tmp_path, ["@pair_A/1\nACGT\n+\nIIII\n"], ["@pair_B/2\nACGT\n+\nIIII\n"])
    with pytest.raises(F.FastBactError, match="synchronization error"):
        F.count_fastx(a, b)


def test_count_fastx_id_prefix_rule_and_lengths(tmp_path, orc):
    """extract_read_id_prefix (estimator.rs:89-99): up to the first whitespace,
    less a trailing /1 or /2; lengths as the oracle's count statistics."""
    import numpy as np
    rng = np.random.default_rng(4)
    l1 = rng.integers(0, 300, size=50)
    l2 = rng.integers(0, 5000, size=50)
    ids1 = [f"SRR1.{i} 1 length=1" if i % 3 else f"SRR1.{i}/1" for i in range(50)]
    ids2 = [f"SRR1.{i}\tother" if i % 3 else f"SRR1.{i}/2 x" for i in range(50)]
    a, b = _pairs_files(tmp_path, [f"@{d}\n{'A' * n}\n+\n{'I' * n}\n" for d, n in zip(ids1, l1)],
                        [f"@{d}\n{'C' * n}\n+\n{'I' * n}\n" for d, n in zip(ids2, l2)])
    st = F.count_fastx(a, b)
    ref = orc.count_stats(l1.astype(np.uint32), l2.astype(np.uint32))
    for f in ("exact_read_pairs", "total_bases", "min_read_len", "max_read_len", "median_read_len"):
        assert getattr(st, f) == getattr(ref, f)
    assert st.mean_read_len == ref.mean_read_len
    a, b = _pairs_files(tmp_path, ["@r/1x\nA\n+\nI\n"], ["@r/2x\nA\n+\nI\n"])  # "/1x" is not a mate suffix
    with pytest.raises(F.FastBactError, match="synchronization"):
        F.count_fastx(a, b)
