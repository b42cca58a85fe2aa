"""GPU parity of unitig compaction (build.rs:330-597): the device walks and
the edges equal the oracle's on crafted graphs (linear, circular, repeats,
hairpins, self-reverse-complement cycles) and on solid sets of the GPU build."""
import numpy as np
import pytest

import fastbact_amd as F
from tests import pyref as P
from tests.helpers import isolate, oracle_arenas
from tests.test_oracle_compact import _arrays, _rand_seq
from tests.test_oracle_solid import REF, _circular_case

pytestmark = pytest.mark.gpu


def _same(g, r):
    assert len(g["seqs"]) == len(r["seqs"])
    assert g["seqs"] == r["seqs"]
    assert np.array_equal(g["cov"].view(np.uint32), r["cov"].view(np.uint32))
    assert np.array_equal(g["circular"], r["circular"])
    for f in ("src", "dst", "ov", "support"):
        assert np.array_equal(g[f], r[f]), f


@pytest.mark.parametrize("k", [11, 21, 31, 33, 55, 63, 65, 77, 99, 127])
def test_crafted_graphs(ctx, orc, k):
    cases = []
    g = _rand_seq(3000 + k, 900)
    cases.append(P.solid_from_sequences([g], k))                           # one linear unitig
    cases.append(P.solid_from_sequences([_rand_seq(4000 + k, 700)], k, circular=True))  # one cycle
    rep = _rand_seq(11, 2 * k)
    cases.append(P.solid_from_sequences([_rand_seq(1, 300) + rep + _rand_seq(2, 300) + rep + _rand_seq(3, 300)], k))
    x = _rand_seq(77 + k, 3 * k)
    cases.append(P.solid_from_sequences([_rand_seq(5, 80) + x + P.revcomp(x) + _rand_seq(9, 80)], k))  # hairpin
    cases.append(P.solid_from_sequences([x + P.revcomp(x)], k, circular=True))  # its own reverse complement
    cases.append(P.solid_from_sequences([_rand_seq(21, 400), _rand_seq(22, 500)], k, circular=True))
    for solid in cases:
        keys, cov, mask = _arrays(solid, k)
        cov = (np.arange(len(solid)) % 97 + 3).astype(np.uint16)  # uneven cov: direction shows in the means
        _same(F.compact_arrays(ctx, keys, cov, mask, k), orc.compact(keys, cov, mask, k))


@pytest.mark.parametrize("k,target", [(21, None), (31, 3), (55, 4), (77, 2), (99, 2), (127, 2)])
def test_compaction_of_gpu_builds(ctx, orc, k, target):
    iso = isolate(seed=31 + k, genome_len=40_000, n_pairs=4000, sub_rate=0.01)
    bank = iso.bank()
    dev = F.DeviceBank(ctx, bank)
    h = F.build_solid_handle(ctx, dev, k, target_solid_cov=target, cms_width=1 << 18, counter_bits=16)
    try:
        s = h.download(k, with_cms=False)
        g = F.compact_unitigs(ctx, h)
    finally:
        h.close()
    r = orc.compact(s.keys, s.cov, s.mask, k)
    assert len(r["seqs"]) > 1
    _same(g, r)


def test_reference_repeat_case_on_gpu(ctx, orc):
    """build.rs:604-687 through the GPU build and compaction: three unitigs."""
    gref = REF["circular_repeat"]
    _, reads = _circular_case()
    k = gref["k"]
    bank = F.ReadBank(F.BitSeqArena.from_ascii(reads))
    dev = F.DeviceBank(ctx, bank)
    h = F.build_solid_handle(ctx, dev, k, target_solid_cov=gref["target_solid_cov"])
    try:
        s = h.download(k, with_cms=False)
        g = F.compact_unitigs(ctx, h)
    finally:
        h.close()
    assert len(g["seqs"]) == gref["expect_unitigs"]
    _same(g, orc.compact(s.keys, s.cov, s.mask, k))


def test_1mbp_isolate_compaction(ctx, orc):
    iso = F.synth_isolate(seed=0xFB01, genome_len=1_000_000, gc=0.5, coverage=30.0, insert_mean=350,
                          insert_sd=30, sub_rate=0.005)
    dev = F.DeviceBank(ctx, iso.bank())
    h = F.build_solid_handle(ctx, dev, 31, target_solid_cov=3, cms_width=1 << 22, counter_bits=16)
    try:
        s = h.download(31, with_cms=False)
        g = F.compact_unitigs(ctx, h)
    finally:
        h.close()
    r = orc.compact(s.keys, s.cov, s.mask, 31)
    _same(g, r)
    # the unitigs spell the solid k-mers: total k-mers = solid k-mers + path ends outside the set
    assert sum(len(x) for x in g["seqs"]) >= len(s.cov)


def test_empty_and_unsorted(ctx):
    g = F.compact_arrays(ctx, np.zeros((0, 1), np.uint64), np.zeros(0, np.uint16), np.zeros(0, np.uint8), 31)
    assert g["seqs"] == [] and g["src"].size == 0
    with pytest.raises(F.FastBactError):
        F.compact_arrays(ctx, np.array([[5], [3]], np.uint64), np.ones(2, np.uint16), np.ones(2, np.uint8), 31)
    dev = F.DeviceBank(ctx, isolate(seed=3, genome_len=5000, n_pairs=300).bank())
    h = F.build_solid_handle(ctx, dev, 31, target_solid_cov=2, sort=False)
    try:
        with pytest.raises(F.FastBactError):
            F.compact_unitigs(ctx, h)
    finally:
        h.close()
