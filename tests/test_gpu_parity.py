"""GPU parity: the HIP path against the CPU oracle, bit-exact.

Every test calls through the C ABI (libfastbact_gpu.so) on cuda:0 and
compares with oracle/liboracle.so on the same seeded inputs.
"""
import numpy as np
import pytest

import fastbact_amd as F
from tests.helpers import distinct_canonical_kmers, isolate, oracle_arenas, ragged_bank

pytestmark = pytest.mark.gpu


def _solid_equal(gpu: F.SolidSet, ref: dict):
    assert gpu.keys.shape == ref["keys"].shape, (gpu.keys.shape, ref["keys"].shape)
    assert np.array_equal(gpu.keys, ref["keys"])
    assert np.array_equal(gpu.cov, ref["cov"])
    assert np.array_equal(gpu.mask, ref["mask"])


@pytest.fixture(scope="module")
def small_iso():
    return isolate(seed=21, genome_len=30_000, n_pairs=2500)


@pytest.mark.parametrize("placement", [1, 2, 3])
@pytest.mark.parametrize("k", [21, 31, 32, 33, 55, 63, 64, 65, 77, 99, 127])
def test_build_solid_matches_oracle_all_widths(ctx, orc, small_iso, k, placement):
    bank = small_iso.bank()
    dev = F.DeviceBank(ctx, bank)
    width = 1 << 18
    g = F.build_solid_kmers(ctx, dev, k, target_solid_cov=3, cms_width=width, hll_p=14, placement=placement)
    r = orc.build_solid(oracle_arenas(bank), k, target_solid_cov=3, width=width, hll_p=14)
    _solid_equal(g, r)
    assert np.array_equal(g.cms_table, r["cms"].table())
    assert np.array_equal(g.hll_registers, r["hll"])
    assert len(g) > 1000


@pytest.mark.parametrize("k,target,bits", [(31, None, 4), (31, 5, 16), (55, 4, 4), (77, 4, 16), (127, 2, 4)])
def test_build_solid_thresholds_and_counter_modes(ctx, orc, small_iso, k, target, bits):
    bank = small_iso.bank()
    dev = F.DeviceBank(ctx, bank)
    width = 1 << 16  # small width: collisions inflate estimates, exercised bit-exactly
    g = F.build_solid_kmers(ctx, dev, k, target_solid_cov=target, cms_width=width, counter_bits=bits)
    r = orc.build_solid(oracle_arenas(bank), k, target_solid_cov=target, width=width, counter_bits=bits)
    _solid_equal(g, r)
    assert np.array_equal(g.cms_table, r["cms"].table())


def test_reference_cms_width_from_genome_size(ctx, orc, small_iso):
    bank = small_iso.bank()
    dev = F.DeviceBank(ctx, bank)
    # genome_size None -> total_bases/60 fallback; explicit -> max(g, 50000) (build.rs:74-84)
    for gs in (None, 30_000, 400_000):
        g = F.build_solid_kmers(ctx, dev, 31, genome_size=gs)
        est = max(gs, 50_000) if gs else max(bank.total_bases() // 60, 50_000)
        w = orc.cms_width_for_genome_size(est)
        r = orc.build_solid(oracle_arenas(bank), 31, width=w)
        _solid_equal(g, r)
        assert g.cms_table.size == w * 2


@pytest.mark.parametrize("placement", [1, 2])
@pytest.mark.parametrize("k", [21, 31, 55, 77, 127])
def test_ragged_reads_and_non_acgt(ctx, orc, k, placement):
    bank = ragged_bank(seed=k)
    dev = F.DeviceBank(ctx, bank)
    g = F.build_solid_kmers(ctx, dev, k, target_solid_cov=1, cms_width=1 << 14, hll_p=12, placement=placement)
    r = orc.build_solid(oracle_arenas(bank), k, target_solid_cov=1, width=1 << 14, hll_p=12)
    _solid_equal(g, r)
    assert np.array_equal(g.cms_table, r["cms"].table())
    assert np.array_equal(g.hll_registers, r["hll"])


@pytest.mark.parametrize("k", [31, 55, 99])
def test_prior_contigs_weights(ctx, orc, small_iso, k):
    bank = small_iso.bank()
    dev = F.DeviceBank(ctx, bank)
    gen = small_iso.genome.tobytes()
    contigs = [gen[1000:5000], gen[12000:12000 + k], gen[20000:20000 + k + 1], gen[25000:26500]]
    g = F.build_solid_kmers(ctx, dev, k, target_solid_cov=4, cms_width=1 << 17, contigs=contigs)
    r = orc.build_solid(oracle_arenas(bank), k, target_solid_cov=4, width=1 << 17, contigs=contigs)
    _solid_equal(g, r)
    assert np.array_equal(g.cms_table, r["cms"].table())


@pytest.mark.parametrize("k", [21, 31, 55, 77, 127])
def test_prior_contigs_both_pass2_routes(ctx, orc, small_iso, k):
    """Prior contigs through the LDS pass 2 (their windows in shares of at
    most 63 pairs after the reads', records flagged as contig) and through
    the global table: a contig over many shares, contigs of k and k + 1 bases,
    contigs repeated until their edge counters (+1000 each) and cov (+100 a
    side) saturate, and few bins, which sends keys past the LDS tables into
    the global table.  Every form gives the oracle's solid set."""
    bank = small_iso.bank()
    dev = F.DeviceBank(ctx, bank)
    gen = small_iso.genome.tobytes()
    contigs = ([gen[0:9000], gen[9000:9000 + k], gen[9100:9100 + k + 1], gen[15000:15000 + 3 * k]]
               + [gen[20000:20600]] * 70 + [gen[22000:22050 + k]] * 400)
    r = orc.build_solid(oracle_arenas(bank), k, target_solid_cov=4, width=1 << 18, contigs=contigs)
    for placement, bins2 in ((1, 0), (2, 0), (2, 2), (3, 0), (3, 2)):
        g = F.build_solid_kmers(ctx, dev, k, target_solid_cov=4, cms_width=1 << 18, contigs=contigs,
                                placement=placement, bins2=bins2)
        _solid_equal(g, r)
        h = F.build_solid_handle(ctx, dev, k, target_solid_cov=4, cms_width=1 << 18, contigs=contigs,
                                 placement=placement, bins2=bins2)
        st = h.stats()
        gate = h.pass2_gate()
        h.close()
        assert st["pass2_route"] == ("global" if placement == 1 else "lds"), st["pass2_route"]
        assert gate == ("key" if placement == 2 else "window"), gate
    assert (r["cov"] == 65535 // 2).any()  # saturated cov of the contig repeated 400 times


@pytest.mark.parametrize("k", [21, 55, 77])
def test_contig_and_read_edges_share_counters(ctx, orc, k):
    """Reads and prior-contig records of the same keys and edges meet in one
    bin at once: 2000 error-free reads tile a 700 bp sequence (every key seen
    hundreds of times, edges counted by reads) while 400 copies of the same
    sequence as prior contigs saturate the same edge counters (+1000 each)
    and the interior keys' cov (+100 a side).
    The LDS route (1 and 2 bins, so each bin holds many keys) and the global
    route give the oracle's solid set."""
    rng = np.random.default_rng(k)
    seq = rng.choice(np.frombuffer(b"ACGT", np.uint8), size=700)
    reads = []
    for _ in range(2000):
        s = int(rng.integers(0, 700 - 150 + 1))
        r = seq[s:s + 150].tobytes()
        if rng.random() < 0.5:
            r = r[::-1].translate(bytes.maketrans(b"ACGT", b"TGCA"))
        reads.append(r)
    bank = F.ReadBank(F.BitSeqArena.from_ascii(reads[:1000]), F.BitSeqArena.from_ascii(reads[1000:]))
    dev = F.DeviceBank(ctx, bank)
    contigs = [seq.tobytes()] * 400
    r = orc.build_solid(oracle_arenas(bank), k, target_solid_cov=4, width=1 << 16, contigs=contigs)
    assert len(r["cov"]) == 700 - k + 1 and (r["cov"] == 65535 // 2).sum() >= 600 - k
    for placement, bins2 in ((2, 1), (2, 2), (2, 0), (1, 0)):
        g = F.build_solid_kmers(ctx, dev, k, target_solid_cov=4, cms_width=1 << 16, contigs=contigs,
                                placement=placement, bins2=bins2)
        _solid_equal(g, r)


def test_empty_bank(ctx, orc):
    bank = F.ReadBank()
    dev = F.DeviceBank(ctx, bank)
    g = F.build_solid_kmers(ctx, dev, 31, cms_width=1 << 12, hll_p=10)
    assert len(g) == 0
    assert not g.cms_table.any()
    assert not g.hll_registers.any()


def test_saturation_nibble_and_u16(ctx, orc):
    # one k-mer repeated many times: counters saturate at 15 / 65535, the solid
    # cov at 65535 // 2 (build.rs:185-186,319)
    seq = b"A" * 20000
    bank = F.ReadBank(F.BitSeqArena.from_ascii([seq] * 5), F.BitSeqArena(), F.BitSeqArena())
    dev = F.DeviceBank(ctx, bank)
    # width 2^16 with LDS bins: every update lands in at most 4 of 16 bins, so
    # the bin regions overflow and the direct-atomic spill path runs
    for bits, width, placement in ((4, 1 << 12, 1), (16, 1 << 12, 2), (4, 1 << 16, 2), (16, 1 << 16, 1)):
        g = F.build_solid_kmers(ctx, dev, 31, cms_width=width, counter_bits=bits, placement=placement)
        r = orc.build_solid(oracle_arenas(bank), 31, width=width, counter_bits=bits)
        _solid_equal(g, r)
        assert np.array_equal(g.cms_table, r["cms"].table())
        assert len(g) == 1 and g.cov[0] == 65535 // 2


@pytest.mark.parametrize("k", [21, 55, 127])
def test_pass1_scatter_forms_agree(ctx, orc, small_iso, k):
    """Both pass-1 scatter kernels give the oracle's table and solid set: the
    per-bin LDS buffers (resident grid, 1025 to 2048 bins: width 2^24) and
    the staged counting sort (32 and 512 bins at widths 2^18 and 2^22, a
    chunked grid, 4096 bins at 2^25).  Low-complexity reads pile their cells
    into a few bins in every round, so the buffers' retries and their
    direct-add fallback run too."""
    base = small_iso.bank()
    low = [b"A" * 150] * 300 + [b"AC" * 75] * 300 + [b"AAC" * 50] * 300
    bank = F.ReadBank(F.BitSeqArena.from_ascii(low), base.pe_forward, base.pe_reverse)
    dev = F.DeviceBank(ctx, bank)
    for width, chunks1 in ((1 << 18, 0), (1 << 18, 2), (1 << 22, 0), (1 << 24, 0), (1 << 25, 0)):
        g = F.build_solid_kmers(ctx, dev, k, target_solid_cov=3, cms_width=width, hll_p=14, placement=2,
                                chunks1=chunks1)
        r = orc.build_solid(oracle_arenas(bank), k, target_solid_cov=3, width=width, hll_p=14)
        _solid_equal(g, r)
        assert np.array_equal(g.cms_table, r["cms"].table()), (width, chunks1)
        assert np.array_equal(g.hll_registers, r["hll"])


def test_cms_reference_known_answers(ctx):
    # sketch.rs:304-319: 0 -> 5 -> saturates at 15
    cms = F.CountMinSketch(ctx, 1024, 4)
    assert cms.estimate(42)[0] == 0
    for _ in range(5):
        cms.add(42)
    assert cms.estimate(42)[0] == 5
    cms.add(np.full(20, 42, np.uint64))
    assert cms.estimate(42)[0] == 15
    # sketch.rs:321-342: 4 threads x 3 adds = 12 (here 12 concurrent GPU adds)
    cms2 = F.CountMinSketch(ctx, 4096, 4)
    cms2.add(np.full(12, 12345, np.uint64))
    assert cms2.estimate(12345)[0] == 12


def test_cms_keys_and_table_match_oracle(ctx, orc):
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 2**63, size=50_000, dtype=np.uint64)
    keys = np.concatenate([keys, keys[:5000], keys[:100].repeat(30)])
    for bits in (4, 16):
        cms = F.CountMinSketch(ctx, 1 << 12, bits)
        cms.add(keys)
        o = orc.Cms(1 << 12, bits)
        for x in keys.tolist():
            o.add(x)
        assert np.array_equal(cms.table(), o.table())
        q = keys[:2000]
        assert np.array_equal(cms.estimate(q), np.array([o.estimate(int(x)) for x in q.tolist()], np.uint32))


def test_hll_hashes_and_bank_match_oracle(ctx, orc, small_iso):
    rng = np.random.default_rng(4)
    hs = rng.integers(0, 2**64 - 1, size=100_000, dtype=np.uint64)
    hs[:10] = 0  # w == 0 -> rho = 64 - p + 1
    for p in (10, 14, 16):
        h = F.HyperLogLog(ctx, p)
        h.add_hash(hs)
        regs = np.zeros(1 << p, np.uint8)
        orc.hll_add_hashes(regs, p, hs)
        assert np.array_equal(h.registers(), regs)
        assert h.estimate() == orc.hll_estimate(regs, p)
    bank = small_iso.bank()
    dev = F.DeviceBank(ctx, bank)
    h = F.HyperLogLog(ctx, 14)
    h.add_bank(dev, 21)
    r = orc.build_solid(oracle_arenas(bank), 21, width=1 << 12, hll_p=14)
    assert np.array_equal(h.registers(), r["hll"])


def test_cms_add_bank_matches_pass1(ctx, orc, small_iso):
    bank = small_iso.bank()
    dev = F.DeviceBank(ctx, bank)
    for k in (21, 63, 127):
        cms = F.CountMinSketch(ctx, 1 << 15, 16)
        cms.add_bank(dev, k)
        r = orc.build_solid(oracle_arenas(bank), k, width=1 << 15, counter_bits=16)
        assert np.array_equal(cms.table(), r["cms"].table())


def _est_equal(g: F.GenomeEstimate, o):
    assert g.genome_size == o.genome_size
    assert g.u_ge_1 == o.u_ge_1 and g.u_ge_2 == o.u_ge_2
    assert g.pairs_sampled == o.pairs_sampled and g.m_sample == o.m_sample
    assert g.estimated_cov == o.estimated_cov
    assert g.subsample_factor == o.subsample_factor
    assert g.sample_kmer_depth == o.sample_kmer_depth
    assert g.confidence == F.api.CONFIDENCE[o.confidence]
    assert g.kmers == [o.kmers[i] for i in range(o.n_kmers)]


@pytest.mark.parametrize("genome_len,n_pairs,max_pairs", [
    (50_000, 20_000, 250_000),    # early stop at the first checkpoint (lambda_s >= 15)
    (1_000_000, 12_000, 250_000),  # no early stop: every sampled pair counts
    (200_000, 9_000, 7_000),       # max_pairs cuts the sample
    (300_000, 3_000, 250_000),     # fewer pairs than the first checkpoint
])
def test_estimate_genome_matches_oracle(ctx, orc, genome_len, n_pairs, max_pairs):
    iso = isolate(seed=genome_len % 97, genome_len=genome_len, n_pairs=n_pairs, sub_rate=0.004)
    bank = iso.bank()
    r1, r2 = bank.pe_forward, bank.pe_reverse
    g = F.estimate_genome(ctx, r1, r2, max_pairs, 100.0)
    st = orc.count_stats(r1.lengths, r2.lengths)
    o = orc.estimate_genome((r1.data, r1.offsets, r1.lengths), (r2.data, r2.offsets, r2.lengths), st, max_pairs,
                            100.0)
    _est_equal(g, o)


def test_estimate_genome_ragged_and_empty(ctx, orc):
    bank = ragged_bank(seed=9, n=3000, max_len=300)
    r1, r2 = bank.pe_forward, bank.pe_reverse
    g = F.estimate_genome(ctx, r1, r2, 250_000, 50.0)
    st = orc.count_stats(r1.lengths, r2.lengths)
    o = orc.estimate_genome((r1.data, r1.offsets, r1.lengths), (r2.data, r2.offsets, r2.lengths), st, 250_000, 50.0)
    _est_equal(g, o)
    with pytest.raises(F.FastBactError) as e:
        F.estimate_genome(ctx, F.BitSeqArena(), F.BitSeqArena(), 1000, 50.0)
    assert "0 valid read pairs" in str(e.value)


def test_1mbp_isolate_100k_pairs(ctx, orc):
    """1 Mbp at 50% GC, 100k PE150 pairs (30x), 0.5% errors, k=31, CMS 4 x 2^22
    u16 counters, HLL p=14, solid threshold 3: bit-exact against the oracle."""
    iso = F.synth_isolate(seed=0xFB01, genome_len=1_000_000, gc=0.5, coverage=30.0, insert_mean=350,
                          insert_sd=30, sub_rate=0.005)
    bank = iso.bank()
    dev = F.DeviceBank(ctx, bank)
    r = orc.build_solid(oracle_arenas(bank), 31, target_solid_cov=3, width=1 << 22, counter_bits=16, hll_p=14,
                        threads=8)
    for placement in (1, 2):
        g = F.build_solid_kmers(ctx, dev, 31, target_solid_cov=3, cms_width=1 << 22, counter_bits=16, hll_p=14,
                                placement=placement)
        _solid_equal(g, r)
        assert np.array_equal(g.cms_table, r["cms"].table())
        assert np.array_equal(g.hll_registers, r["hll"])
    assert 0.5e6 < len(g) < 1.1e6
    # the pass's HLL (splitmix64 of the canonical ntHash) estimates the
    # distinct canonical 31-mers within the reference's 3% gate
    # (sketch.rs:262-284) of their exact count
    est = F.hll_estimate_registers(g.hll_registers, 14)
    exact = distinct_canonical_kmers([iso.r1, iso.r2], 31)
    assert abs(est - exact) / exact < 0.03, (est, exact)


@pytest.mark.parametrize("k", [33, 55, 64, 77, 127])
def test_sort_with_shared_top_bits_matches_oracle(ctx, orc, k):
    """Keys over 64 bits are sorted by their first 16 bases first; runs of
    keys that share those are then sorted in place when at most 64 long,
    else the whole set is sorted limb by limb.  Reads that start with the
    same 40 bases give solid k-mers sharing their first 32 bases, in runs
    of up to 20 or of 300 over their top 64 bits (asserted below); over the
    first 16 bases the runs are 46-75 and 2,000-3,700 keys long, so k = 33
    takes the in-place sort here and the other k the limb-by-limb one (runs
    of 47-56 at k = 55 and 77: the test below).  A random isolate has no
    runs over its top 64 bits.  Every order equals the oracle's."""
    rng = np.random.default_rng(k)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    head = np.frombuffer(b"AC" * 20, np.uint8)
    banks = []
    for n_tails in (6, 300):
        tails = [rng.choice(acgt, size=160) for _ in range(n_tails)]
        seqs = [np.concatenate([head, t]).tobytes() for t in tails for _ in range(3)]
        banks.append(F.ReadBank(F.BitSeqArena.from_ascii(seqs)))
    runs = []
    for b in banks + [isolate(seed=k, genome_len=20_000, n_pairs=1500, sub_rate=0.0).bank()]:
        r = orc.build_solid(oracle_arenas(b), k, target_solid_cov=2, width=1 << 18)
        g = F.build_solid_kmers(ctx, F.DeviceBank(ctx, b), k, target_solid_cov=2, cms_width=1 << 18)
        _solid_equal(g, r)
        nref = r["keys"].shape[1]
        top = [sum(int(x) << (64 * (nref - 1 - j)) for j, x in enumerate(row)) >> (2 * k - 64) for row in r["keys"]]
        runs.append(max(np.unique(np.array(top, dtype=object), return_counts=True)[1]))
    # the data's runs over the top 64 bits: short, long, none (an
    # error-free isolate)
    assert 2 <= runs[0] <= 64 and runs[2] == 1, runs
    assert runs[1] > 64 or k == 33, runs


@pytest.mark.parametrize("k,bins2", [(31, 1), (55, 2), (77, 1), (127, 4)])
def test_lds_table_full_spills_to_global_table(ctx, orc, k, bins2):
    """Few pass-2 bins force keys past the LDS table into the global table;
    each key lives in one of them, so the set is still exact."""
    iso = isolate(seed=77, genome_len=120_000, n_pairs=9000)
    bank = iso.bank()
    dev = F.DeviceBank(ctx, bank)
    g = F.build_solid_kmers(ctx, dev, k, target_solid_cov=2, cms_width=1 << 18, placement=2, bins2=bins2)
    r = orc.build_solid(oracle_arenas(bank), k, target_solid_cov=2, width=1 << 18)
    _solid_equal(g, r)
    h = F.build_solid_handle(ctx, dev, k, target_solid_cov=2, cms_width=1 << 18, placement=2, bins2=bins2)
    st = h.stats()
    h.close()
    assert st["keys"] > 6000 * bins2  # more keys than the LDS tables hold


@pytest.mark.parametrize("placement", [1, 3])
def test_pass2_placements_agree_on_stats(ctx, small_iso, placement):
    """The per-window gate gives the same records and keys on both routes."""
    dev = F.DeviceBank(ctx, small_iso.bank())
    h = F.build_solid_handle(ctx, dev, 31, target_solid_cov=3, cms_width=1 << 18, placement=placement)
    st = h.stats()
    h.close()
    h1 = F.build_solid_handle(ctx, dev, 31, target_solid_cov=3, cms_width=1 << 18, placement=1)
    st1 = h1.stats()
    h1.close()
    assert st["records"] == st1["records"] and st["keys"] == st1["keys"] and st["windows"] == st1["windows"]


@pytest.mark.parametrize("k,target,bits,width", [(21, None, 16, 1 << 22), (21, 2, 4, 1 << 20), (31, 4, 16, 1 << 21),
                                                  (55, 4, 4, 1 << 22), (77, None, 16, 1 << 20), (127, 3, 16, 1 << 22),
                                                  (33, 1, 16, 1 << 20), (21, 16, 4, 1 << 18)])
def test_gate_per_kept_key_equals_oracle(ctx, orc, k, target, bits, width):
    """A sketch far wider than the input leaves most error k-mers below the
    gate, so many kept keys have low edges whose neighbour fails it: the
    ungated route (records of every pair, the gate probed per kept key and
    low edge, error k-mers dropped by the bin kernel's record filter) must
    still give the oracle's cov, mask and key set.  Target 16 with nibble
    counters (cap 15 < t) takes the per-window gate (nothing is solid), and
    target 1 has no record filter."""
    iso = isolate(seed=100 + k, genome_len=50_000, n_pairs=4000, sub_rate=0.01)
    bank = iso.bank()
    dev = F.DeviceBank(ctx, bank)
    r = orc.build_solid(oracle_arenas(bank), k, target_solid_cov=target, width=width, counter_bits=bits)
    for placement in (2, 3):
        h = F.build_solid_handle(ctx, dev, k, target_solid_cov=target, cms_width=width, counter_bits=bits,
                                 placement=placement)
        gate, st = h.pass2_gate(), h.stats()
        g = h.download(k, counter_bits=bits)
        h.close()
        _solid_equal(g, r)
        assert gate == ("key" if placement == 2 and (target is None or target <= (15 if bits == 4 else 255))
                        else "window"), (placement, gate, st, len(r["cov"]))
    if target != 16:
        assert len(r["cov"]) > 4000


@pytest.mark.parametrize("k,bins2", [(21, 1 << 15), (55, 1 << 16), (127, 1 << 18)])
def test_pass2_rounds_match_oracle(ctx, orc, small_iso, k, bins2):
    """More pass-2 bins than one launch of the bin kernels takes: the keys are
    split into rounds by the top bits of their bin hash, each round counted,
    scattered and tabled alone.  The set is the same exact set."""
    bank = small_iso.bank()
    dev = F.DeviceBank(ctx, bank)
    g = F.build_solid_kmers(ctx, dev, k, target_solid_cov=2, cms_width=1 << 18, placement=2, bins2=bins2)
    r = orc.build_solid(oracle_arenas(bank), k, target_solid_cov=2, width=1 << 18)
    _solid_equal(g, r)
    h = F.build_solid_handle(ctx, dev, k, target_solid_cov=2, cms_width=1 << 18, placement=2, bins2=bins2)
    st = h.stats()
    h.close()
    h1 = F.build_solid_handle(ctx, dev, k, target_solid_cov=2, cms_width=1 << 18, placement=2)
    st1 = h1.stats()
    h1.close()
    assert st["pass2_route"] == "lds" and st["pass2_rounds"] == bins2 // 16384
    assert st1["pass2_rounds"] == 1
    # every round writes its own records: together they are the one-round records
    assert st["pass2_records"] == st1["pass2_records"]
    # with the per-window gate the tables hold the same keys in any number of
    # rounds (the ungated route's record filter lets through keys by how they
    # share bins, so its table contents vary with the bin count)
    raw = []
    for b2 in (bins2, 0):
        h3 = F.build_solid_handle(ctx, dev, k, target_solid_cov=2, cms_width=1 << 18, placement=3, bins2=b2)
        raw.append(h3.stats()["records"])
        h3.close()
    assert raw[0] == raw[1]


def _repeat_heavy_bank(seed: int, n_random: int = 12_000, n_repeat: int = 12_000):
    """Reads of a random 150 kb genome plus as many copies of one read made
    of a period-8 tandem repeat, whose few k-mers carry half the records."""
    rng = np.random.default_rng(seed)
    genome = rng.choice(list(b"ACGT"), size=150_000).astype(np.uint8).tobytes()
    reads = []
    for _ in range(n_random):
        s0 = int(rng.integers(0, len(genome) - 150))
        reads.append(genome[s0:s0 + 150])
    rep = (b"AACGTTGC" * 19)[:150]
    reads += [rep] * n_repeat
    half = len(reads) // 2
    return F.ReadBank(F.BitSeqArena.from_ascii(reads[:half]), F.BitSeqArena.from_ascii(reads[half:]))


@pytest.mark.parametrize("k", [21, 55])
def test_chunked_regions_fill_up_then_exact_runs(ctx, orc, small_iso, k):
    """The ungated pass 2 without its count walk writes each bin's records
    into per-slice regions sized for an even spread.  A tandem-repeat read
    present 12,000 times sends half of all records to the few bins of its
    k-mers: their regions fill up, the pass runs again with exact runs, and
    the set is the oracle's.  An ordinary isolate stays on the chunked layout."""
    bank = _repeat_heavy_bank(seed=k)
    dev = F.DeviceBank(ctx, bank)
    kw = dict(target_solid_cov=3, cms_width=1 << 20, placement=2, bins2=64)
    g = F.build_solid_kmers(ctx, dev, k, **kw)
    r = orc.build_solid(oracle_arenas(bank), k, target_solid_cov=3, width=1 << 20)
    _solid_equal(g, r)
    h = F.build_solid_handle(ctx, dev, k, **kw)
    assert h.pass2_gate() == "key" and h.pass2_layout() == "chunks_then_runs"
    h.close()
    dev.close()
    dev = F.DeviceBank(ctx, small_iso.bank())
    h = F.build_solid_handle(ctx, dev, k, target_solid_cov=3, cms_width=1 << 18, placement=2)
    assert h.pass2_layout() == "chunks"
    h.close()


@pytest.mark.parametrize("k", [55, 77])
@pytest.mark.parametrize("n_shared", [40, 400])
def test_sort_runs_of_keys_sharing_their_top_bases(ctx, orc, k, n_shared):
    """Keys of more than 32 bases are sorted by their first 16 bases, then
    every run of keys sharing those is ordered in place by the whole key;
    a run longer than 64 sends the set to the word-by-word sort.  Reads that
    open with 16 A's give their first windows the same 16 leading bases
    (runs of 47-56 keys for 40 such reads, over 500 for 400): both forms
    give the oracle's sorted set."""
    rng = np.random.default_rng(k * 1000 + n_shared)
    genome = rng.choice(list(b"ACGT"), size=60_000).astype(np.uint8).tobytes()
    reads = []
    for _ in range(n_shared):
        reads.append(b"A" * 16 + rng.choice(list(b"ACGT"), size=134).astype(np.uint8).tobytes())
    for _ in range(600):
        s0 = int(rng.integers(0, len(genome) - 150))
        reads.append(genome[s0:s0 + 150])
    half = len(reads) // 2
    bank = F.ReadBank(F.BitSeqArena.from_ascii(reads[:half]), F.BitSeqArena.from_ascii(reads[half:]))
    dev = F.DeviceBank(ctx, bank)
    try:
        for placement in (1, 2):
            g = F.build_solid_kmers(ctx, dev, k, target_solid_cov=1, cms_width=1 << 18, placement=placement)
            r = orc.build_solid(oracle_arenas(bank), k, target_solid_cov=1, width=1 << 18)
            _solid_equal(g, r)
    finally:
        dev.close()
