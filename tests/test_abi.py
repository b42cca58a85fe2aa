"""C ABI checks that need no GPU: the library loads and exports every symbol
its headers declare; host-side functions agree with the oracle; the device
entry points refuse to run without a GPU instead of falling back to the CPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import fastbact_amd as F
from fastbact_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: Path):
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    names = set()
    for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(fb_\w+)\s*\(", text, flags=re.M):
        names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = set()
    for h in (ROOT / "include").glob("*.h"):
        declared |= _declared(h)
    assert len(declared) > 40
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert declared <= bound | {"fb_version"}, sorted(declared - bound)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "absent.so")
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load()


def _has_gpu():
    n = C.c_int(0)
    try:
        hip = C.CDLL("libamdhip64.so")
        return hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.mark.skipif(_has_gpu(), reason="a GPU is present")
def test_no_gpu_means_no_context():
    with pytest.raises(F.FastBactError) as e:
        F.Context(0)
    assert e.value.code == _lib.FB_ERR_NODEV


def test_pack_ascii_matches_oracle(orc):
    rng = np.random.default_rng(1)
    seqs = [bytes(rng.choice(list(b"ACGTacgtNRY"), size=int(n))) for n in rng.integers(0, 300, size=500)]
    a = F.BitSeqArena.from_ascii(seqs)
    data, offs, lens = orc.pack(seqs)
    assert np.array_equal(a.offsets, offs) and np.array_equal(a.lengths, lens)
    assert np.array_equal(a.data, data[: a.data.size]) and not data[a.data.size:].any()


def test_count_stats_and_ladder_match_oracle(orc):
    rng = np.random.default_rng(2)
    for _ in range(20):
        n = int(rng.integers(1, 200))
        l1 = rng.integers(0, 5000, size=n).astype(np.uint32)
        l2 = rng.integers(0, 300, size=n).astype(np.uint32)
        a, b = F.count_stats(l1, l2), orc.count_stats(l1, l2)
        for f, _ in _lib.CountStats._fields_:
            assert getattr(a, f) == getattr(b, f), f
    with pytest.raises(F.FastBactError, match="0 valid read pairs"):
        F.count_stats([], [])
    for med in list(range(0, 400, 3)) + [4095, 10000]:
        for tgt in (None, 41, 80, 200):
            assert F.compute_dynamic_kmers(med, tgt) == orc.compute_dynamic_kmers(med, tgt)


def test_hll_estimator_and_width_match_oracle(orc):
    rng = np.random.default_rng(3)
    for p in (4, 10, 14, 16, 18):
        for fill in (0, 5, 1000, 10**6):
            regs = np.zeros(1 << p, np.uint8)
            orc.hll_add_hashes(regs, p, rng.integers(0, 2**64 - 1, size=fill, dtype=np.uint64)) if fill else None
            assert F.hll_estimate_registers(regs, p) == orc.hll_estimate(regs, p)
    for g in (0, 1, 50_000, 131_072, 200_000, 5_000_000, 10**10):
        assert F.CountMinSketch.width_for_genome_size(g) == orc.cms_width_for_genome_size(g)
    for x in (0, 1, 2**63, 2**64 - 1):
        assert F.splitmix64(x) == orc.splitmix64(x)


def test_synthetic_isolate_is_seeded_and_shaped():
    a = F.synth_isolate(n_pairs=500, seed=9, genome_len=20_000, gc=0.66, sub_rate=0.0)
    b = F.synth_isolate(n_pairs=500, seed=9, genome_len=20_000, gc=0.66, sub_rate=0.0)
    c = F.synth_isolate(n_pairs=500, seed=10, genome_len=20_000, gc=0.66, sub_rate=0.0)
    assert np.array_equal(a.r1, b.r1) and np.array_equal(a.genome, b.genome)
    assert not np.array_equal(a.genome, c.genome)
    g = a.genome
    gc = np.isin(g, np.frombuffer(b"GC", np.uint8)).mean()
    assert abs(gc - 0.66) < 0.02
    # error-free reads are substrings of the genome or of its reverse complement
    gs = g.tobytes()
    rc = gs[::-1].translate(bytes.maketrans(b"ACGT", b"TGCA"))
    for r in list(a.r1[:50]) + list(a.r2[:50]):
        s = r.tobytes()
        assert s in gs or s in rc
    p = F.synth_params(genome_len=5_000_000, coverage=100.0, read_len=150)
    assert _lib.load().fb_synth_n_pairs(C.byref(p)) == 1_666_667


def test_synthetic_repeats_present():
    iso = F.synth_isolate(n_pairs=10, seed=4, genome_len=200_000, gc=0.5, operon_copies=7, operon_len=5000,
                          tandem_arrays=5)
    g = iso.genome.tobytes()
    # the 5 kbp operon occurs 7 times: its 60-mers are the most frequent ones
    from collections import Counter
    cnt = Counter(g[i:i + 60] for i in range(0, len(g) - 60))
    assert max(cnt.values()) >= 7
