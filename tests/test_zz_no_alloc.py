"""fb_cms_add_and_estimate_static, the entry point that allocates nothing,
called through the C ABI: its estimates and final table equal
fb_cms_add_and_estimate's on the same sketch state, for two calls in a row
(the second smaller, so the static array still holds the first call's
larger contents), and it refuses batches over its limit."""
import ctypes as C

import numpy as np
import pytest

import fastbact_amd as F
from fastbact_amd import _lib

pytestmark = pytest.mark.gpu

LIMIT = 1 << 18


def _static(cms, keys):
    k = np.ascontiguousarray(keys, np.uint64)
    out = np.zeros(k.size, np.uint32)
    rc = cms.ctx.lib.fb_cms_add_and_estimate_static(cms.h, k.ctypes.data_as(C.POINTER(C.c_uint64)), k.size,
                                                    out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return rc, out


@pytest.mark.parametrize("bits,width", [(4, 1 << 10), (16, 1 << 20)])
def test_static_entry_point_equals_allocating_one(bits, width):
    with F.Context(0) as ctx:
        rng = np.random.default_rng(width)
        a = F.CountMinSketch(ctx, width, bits)
        b = F.CountMinSketch(ctx, width, bits)
        warm = rng.integers(0, 5000, size=10_000).astype(np.uint64)
        a.add(warm)
        b.add(warm)
        first = rng.integers(0, 5000, size=LIMIT).astype(np.uint64)  # the largest batch
        second = rng.integers(0, 5000, size=3_000).astype(np.uint64)
        second[:500] = 77
        for keys in (first, second):
            rc, got = _static(a, keys)
            assert rc == 0, ctx.lib.fb_last_error(ctx.h)
            want = b.add_and_estimate(keys)
            assert np.array_equal(got, want)
            assert np.array_equal(a.table(), b.table())
        rc, _ = _static(a, np.zeros(LIMIT + 1, np.uint64))
        assert rc == _lib.FB_ERR_INVALID
        assert np.array_equal(a.table(), b.table())  # a refused call changes nothing
