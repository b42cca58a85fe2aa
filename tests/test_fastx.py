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
