"""GPU parity of the sketch and hash entry points beyond the passes:
HyperLogLog::add_canonical_hash and merge, CountMinSketch::add_and_estimate
(in-order semantics over a batch) and RollingNtHash over a sequence, each
against the oracle."""
import numpy as np
import pytest

import fastbact_amd as F

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("p", [10, 14, 16])
def test_hll_add_canonical_hash_and_merge(ctx, orc, p):
    rng = np.random.default_rng(p)
    a = rng.integers(0, 2**64 - 1, size=50_000, dtype=np.uint64)
    b = rng.integers(0, 2**64 - 1, size=80_000, dtype=np.uint64)
    ha, hb = F.HyperLogLog(ctx, p), F.HyperLogLog(ctx, p)
    ha.add_canonical_hash(a)
    hb.add_canonical_hash(b)
    ra = np.zeros(1 << p, np.uint8)
    rb = np.zeros(1 << p, np.uint8)
    orc.hll_add_hashes(ra, p, [orc.splitmix64(int(x)) for x in a])
    orc.hll_add_hashes(rb, p, [orc.splitmix64(int(x)) for x in b])
    assert np.array_equal(ha.registers(), ra) and np.array_equal(hb.registers(), rb)
    ha.merge(hb)
    orc.hll_merge(ra, rb, p)
    assert np.array_equal(ha.registers(), ra)
    assert ha.estimate() == orc.hll_estimate(ra, p)
    with pytest.raises(F.FastBactError):
        ha.merge(F.HyperLogLog(ctx, p - 1))


@pytest.mark.parametrize("bits,width", [(4, 1 << 6), (16, 1 << 6), (16, 1 << 16)])
def test_cms_add_and_estimate_in_order(ctx, orc, bits, width):
    rng = np.random.default_rng(width + bits)
    # repeated keys and a narrow sketch: later adds in a batch see earlier ones
    keys = rng.integers(0, 500, size=20_000).astype(np.uint64)
    keys[5000:5300] = 12345
    g = F.CountMinSketch(ctx, width, bits)
    g.add(keys[:1000])
    c = orc.Cms(width, bits)
    for key in keys[:1000].tolist():
        c.add(key)
    got = np.concatenate([g.add_and_estimate(keys[1000:9000]), g.add_and_estimate(keys[9000:])])
    want = np.array([c.add_and_estimate(int(key)) for key in keys[1000:].tolist()], np.uint32)
    assert np.array_equal(got, want)
    assert np.array_equal(g.table(), c.table())
    assert got.max() == (15 if bits == 4 else got.max())


@pytest.mark.parametrize("k", [1, 5, 21, 31, 63, 64, 65, 127, 200])
def test_rolling_nthash_windows(ctx, orc, k):
    rng = np.random.default_rng(k)
    seq = rng.choice(np.frombuffer(b"ACGTacgtN", np.uint8), size=3000).tobytes()
    got = F.nthash_windows(ctx, seq, k)
    assert np.array_equal(got, orc.nthash_rolling(seq, k))
    assert F.nthash_windows(ctx, seq[: k - 1], k).size == 0
