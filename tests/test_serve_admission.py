"""Serving admission policy and admission footprints (host only, no GPU)."""
import ctypes as C

import pytest

import fastbact_amd as F
from fastbact_amd import _lib


def test_admits_in_arrival_order_within_budget():
    q = F.Admission(budget=100, max_inflight=8)
    for rid, b in [(1, 40), (2, 30), (3, 50), (4, 10)]:
        assert q.submit(rid, b) == _lib.FB_OK
    assert q.admit() == [1, 2]          # 3 (50) does not fit next to 70
    assert q.used == 70 and q.queued == 2 and q.inflight == 2
    assert q.admit() == []              # nothing changed
    assert q.retire(1) == _lib.FB_OK
    assert q.admit() == [3, 4]          # 30 + 50 + 10 = 90
    assert q.used == 90


def test_head_is_never_overtaken():
    q = F.Admission(budget=100, max_inflight=8)
    q.submit(1, 60)
    q.submit(2, 60)                     # waits for memory
    q.submit(3, 10)                     # would fit, but 2 is ahead of it
    assert q.admit() == [1]
    assert q.admit() == []
    q.retire(1)
    assert q.admit() == [2, 3]


def test_max_inflight_and_oversize_refusal():
    q = F.Admission(budget=1000, max_inflight=2)
    assert q.submit(9, 1001) == _lib.FB_ERR_OOM   # can never fit
    for rid in (1, 2, 3):
        q.submit(rid, 1)
    assert q.admit() == [1, 2]
    q.retire(2)
    assert q.admit() == [3]
    assert q.retire(2) == _lib.FB_ERR_INVALID   # not in flight any more
    q.retire(1)
    q.retire(3)
    assert q.used == 0 and q.inflight == 0 and q.queued == 0


def test_footprint_grows_with_reads_and_sketch():
    lib = _lib.load()
    fp = lambda g, bases, reads, k, w: F.batch_footprint(g, bases, reads, k, cms_width=w, counter_bits=16, hll_p=14)
    small = fp(2_000_000, 100_000_000, 666_667, 55, 1 << 23)
    large = fp(2_000_000, 400_000_000, 2_666_667, 55, 1 << 23)
    assert 0 < small < large
    # the sketch alone: 4 rows of exact u32 counts
    assert fp(2_000_000, 1_000, 10, 21, 1 << 24) > 4 * (1 << 24) * 4
    # a wider k needs wider records
    assert fp(5_000_000, 500_000_000, 3_333_334, 127, 1 << 24) > fp(5_000_000, 500_000_000, 3_333_334, 21, 1 << 24)
    assert lib.fb_batch_footprint(1, 1, 1, None, 21) == 0


def test_server_needs_a_gpu_or_refuses_bad_params():
    p = F.serve_params((21, 55), (None, 4))
    p.counter_bits = 7
    h = C.c_void_p()
    assert _lib.load().fb_server_create(0, C.byref(p), C.byref(h)) == _lib.FB_ERR_INVALID
    with pytest.raises(ValueError):
        F.serve_params((21, 55), (None,))


def test_autotune_choices_save_and_load(tmp_path):
    lib = _lib.load()
    lib.fb_autotune_clear()
    good = tmp_path / "tune.txt"
    good.write_text("# kind limbs bits log2w log2_bases log2_n placement split chunks1 chunks2 ms\n"
                    "0 2 16 23 29 4 2 3 8 4 12.5\n0 1 16 24 29 0 2 2 16 8 25.0\n")
    assert lib.fb_autotune_load(str(good).encode()) == _lib.FB_OK
    out = tmp_path / "saved.txt"
    assert lib.fb_autotune_save(str(out).encode()) == _lib.FB_OK
    rows = [l.split() for l in out.read_text().splitlines() if not l.startswith("#")]
    assert sorted(r[:10] for r in rows) == sorted([["0", "1", "16", "24", "29", "0", "2", "2", "16", "8"],
                                                  ["0", "2", "16", "23", "29", "4", "2", "3", "8", "4"]])
    bad = tmp_path / "bad.txt"
    bad.write_text("0 2 16 23 29 4 9 3 8 4 1.0\n")  # placement 9 does not exist
    assert lib.fb_autotune_load(str(bad).encode()) == _lib.FB_ERR_INVALID
    assert lib.fb_autotune_load(str(tmp_path / "absent.txt").encode()) == _lib.FB_ERR_INVALID
    lib.fb_autotune_clear()
    assert lib.fb_autotune_save(str(out).encode()) == _lib.FB_OK
    assert [l for l in out.read_text().splitlines() if not l.startswith("#")] == []
