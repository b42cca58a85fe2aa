#!/usr/bin/env python3
"""Benchmark of the k-mer spectrum path on one MI355X.

A step is one full multi-k spectrum of one isolate held in HBM: for each of
k = 21, 55, 77 one pass of Count-Min Sketch + HyperLogLog + solid-k-mer set
(build_cdbg_with_contigs_ext passes 1-3, src/graph/build.rs:67-328), ending
with the sorted solid set.  Workload: a seeded 5 Mbp isolate at GC 0.508,
PE150 at 100x (insert N(400,40), 0.3% substitutions, 0.01% indels), CMS
4 x 2^24 saturating u16 counters, HLL p=16; solid gates None at k=21 (2/3, as
the first graph build) and 4 at k=55/77 (the multi-k upgrades).

Prints one JSON line.  `value` is Gbases/s = sum over passes of the bases of
reads with len >= k, divided by the wall time of the timed steps (inputs
resident in HBM).  The CPU baseline is the oracle's threaded form of the same
passes on a bounded sample of the same reads.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Gbases/s k-mer counted (CMS+HLL+solid set) 1 GPU; p50/p99 isolate latency"
HBM_PEAK_GBS = 8000.0
PASSES = [(21, None), (55, 4), (77, 4)]
WIDTH = 1 << 24
COUNTER_BITS = 16
HLL_P = 16
GENOME = dict(seed=0xFB02, genome_len=5_000_000, gc=0.508, coverage=100.0, read_len=150, insert_mean=400.0,
              insert_sd=40.0, sub_rate=0.003, indel_rate=0.0001)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ref_words(k: int) -> int:
    return 1 if k <= 31 else (2 if k <= 63 else 4)


def kernel_bytes(k: int, st: dict, bases: int, solid: int) -> dict:
    """Algorithmic bytes one launch of each kernel must move for pass k.

    stream = packed bases (L/4) plus 12 bytes of offset and length per read.
    k_bin_scatter: stream + per window the 4 u16 cell offsets it emits (8) and
      the HLL register read-modify-write (2).
    k_bin_count: the offsets read back (8 per window), the u32 counter
      read-modify-write of every cell (8) and the gate bit per cell (1/8).
    k_p2_scatter: stream + per record its rw 64-bit words.
    k_p2_bins: the records read back plus the solid set written
      (8 * ref_words + 3 per solid k-mer)."""
    rw = (2 * k + 8 + 63) // 64
    stream = bases / 4.0 + 12.0 * st["reads_used"]
    cells = 4.0 * WIDTH
    recs = st["pass2_records"] * rw * 8.0
    return {"bin_scatter": stream + st["windows"] * 10.0,
            "bin_count": st["windows"] * 8.0 + cells * (8.0 + 1.0 / 8.0),
            "p2_scatter": stream + recs,
            "p2_bins": recs + solid * (8.0 * ref_words(k) + 3.0)}


KERNEL_NAMES = {"bin_scatter": "k_bin_buffer (pass-1 Count-Min scatter through per-bin LDS buffers + HLL)",
                "bin_count": "k_bin_count (pass-1 LDS counts + gate bits)",
                "p2_scatter": "k_p2_scatter (pass-2 adjacency record scatter)",
                "p2_bins": "k_p2_bins (pass-2 LDS tables + pass 3)"}


def pass_model_bytes(k: int, st: dict, bases: int, solid: int) -> float:
    """Whole-pass algorithmic bytes: per read L/4 + 12; per window the 4-row
    CMS update (8c) and query (4c) and the HLL RMW (2); per adjacency record
    the key and its cov/edge RMW (K_B + 4); per solid k-mer K_B + 3."""
    c = 2 if COUNTER_BITS == 16 else 1
    kb = 8 * ref_words(k)
    stream = bases / 4.0 + 12.0 * st["reads_used"]
    return stream + st["windows"] * (12 * c + 2) + st["records"] * (kb + 4) + solid * (kb + 3)


def make_isolate():
    import fastbact_amd as F
    t = time.time()
    iso = F.synth_isolate(**GENOME)
    bank = iso.bank()
    log(f"synthetic isolate: {iso.r1.shape[0]} pairs, {bank.total_bases()} bases ({time.time() - t:.1f} s)")
    return iso, bank


def run_step(ctx, dev, F, placement=0):
    handles = []
    for k, tgt in PASSES:
        handles.append(F.build_solid_handle(ctx, dev, k, target_solid_cov=tgt, genome_size=GENOME["genome_len"],
                                            cms_width=WIDTH, counter_bits=COUNTER_BITS, hll_p=HLL_P,
                                            placement=placement))
    ctx.synchronize()
    info = [(h.timings(), h.stats(), h.count(), h.kernel_timings()) for h in handles]
    for h in handles:
        h.close()
    return info


def copy_calibration(ctx, mib: int) -> None:
    """One k_copy_calibration launch: reads and writes mib MiB."""
    ctx.check(ctx.lib.fb_diag_copy(ctx.h, mib << 20), "fb_diag_copy")
    log(f"copy calibration: {mib} MiB device-to-device (k_copy_calibration)")


def atomic_sweep(ctx) -> list:
    """Random 4-byte access rate (k_atomic_rate) at table sizes from
    L2-resident (4 MiB) to HBM-resident (1 GiB): atomicAdd without and with
    the returned value, and plain loads -- the device figures the passes'
    atomics and random probes are compared with."""
    import ctypes as C
    out = []
    n_updates = 1 << 28
    for log2n in (20, 22, 24, 26, 28):
        for mode, name in ((0, "atomic"), (1, "atomic_ret"), (2, "load"), (3, "store")):
            ms = C.c_float()
            ctx.check(ctx.lib.fb_diag_atomic_rate(ctx.h, log2n, n_updates, mode, C.byref(ms)), "fb_diag_atomic_rate")
            out.append({"table_mib": (4 << log2n) >> 20, "op": name, "accesses": n_updates,
                        "ms": round(ms.value, 3), "g_per_s": round(n_updates / ms.value / 1e6, 1)})
    log(f"random access sweep: {[(o['table_mib'], o['op'], o['g_per_s']) for o in out]}")
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_full_isolate_k21_55_77.json")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE + WRITE_SIZE, scaled by the copy
    calibration), or None when no summary covers it."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        v = t["kernels"][kernel]["hbm_bytes_per_launch"]
        return float(v), os.path.relpath(TRAFFIC_FILE, ROOT)
    except (OSError, KeyError, ValueError, TypeError):
        return None, None


def host_info() -> dict:
    """Cores of this host: nproc, the CPUs this process may use, the model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "allowed_cpus": allowed, "model": model}


def cpu_baseline(ctx, F, bank, pairs: int, thread_counts) -> dict:
    """The oracle's threaded form (the reference's CPU structures: CAS
    Count-Min, 64 mutex-sharded maps fed by 2048-entry buffers, 4096-read
    pass-2 chunks) on the first `pairs` pairs of the isolate, at each thread
    count; the GPU's passes on the same sample are checked equal to its
    output.  Gbases/s is the rate on that sample: the per-window work does not
    depend on the sample size, while the sample's maps are smaller than the
    whole isolate's (fewer distinct error k-mers), so the whole isolate would
    run a little slower per base, not faster."""
    from oracle import oracle as O
    n = min(pairs, len(bank.pe_forward))
    sub = F.ReadBank(F.BitSeqArena(), _head(bank.pe_forward, n), _head(bank.pe_reverse, n))
    arenas = [(a.data, a.offsets, a.lengths) for a in sub.arenas()]
    bases = sum(int(a.lengths[a.lengths >= k].sum(dtype=np.uint64)) for a in sub.arenas() for k, _ in PASSES)
    runs, ref = [], None
    for threads in thread_counts:
        t = time.time()
        res = [O.build_solid(arenas, k, target_solid_cov=tgt, width=WIDTH, counter_bits=COUNTER_BITS, hll_p=HLL_P,
                             threads=threads) for k, tgt in PASSES]
        dt = time.time() - t
        ref = ref or res
        runs.append({"value": bases / dt / 1e9, "unit": "Gbases/s", "cores": threads, "seconds": round(dt, 2)})
        log(f"CPU baseline: {threads} threads, {dt:.1f} s")
    dev = F.DeviceBank(ctx, sub)
    equal = True
    for (k, tgt), r in zip(PASSES, ref):
        g = F.build_solid_kmers(ctx, dev, k, target_solid_cov=tgt, genome_size=GENOME["genome_len"], cms_width=WIDTH,
                                counter_bits=COUNTER_BITS, hll_p=HLL_P)
        equal = equal and bool(np.array_equal(g.keys, r["keys"]) and np.array_equal(g.cov, r["cov"])
                               and np.array_equal(g.mask, r["mask"]) and np.array_equal(g.cms_table, r["cms"].table())
                               and np.array_equal(g.hll_registers, r["hll"]))
    dev.close()
    out = dict(runs[0])
    out.update({"kind": "port",
                "sample": f"first {n} pairs of the same isolate ({sub.total_bases()} bases, "
                          f"{100.0 * n / len(bank.pe_forward):.0f}% of its pairs), the same 3 passes (k=21/55/77, "
                          f"CMS 4x2^24 u16, HLL p16), oracle threaded form",
                "runs": runs, "host": host_info(), "gpu_equal_to_oracle_on_sample": equal})
    return out


def _head(a, n):
    import fastbact_amd as F
    if n >= len(a):
        return a
    end = int(a.offsets[n - 1] + a.lengths[n - 1]) if n else 0
    return F.BitSeqArena(a.data[: (end + 31) // 32 + 1].copy(), a.offsets[:n].copy(), a.lengths[:n].copy())


# ---------------------------------------------------------------------------
# Batched and served forms on isolates drawn from the 64-isolate workload:
# genome 1.5-7 Mbp uniform, GC 0.30-0.68 uniform, coverage lognormal (median
# 80x, sigma 0.5) clipped to 30-200x, PE150 at 0.3% substitutions.

BATCH_K, BATCH_WIDTH, BATCH_HLL_P = 55, 1 << 23, 14


def batch_shapes(n: int, seed: int = 0xFB03):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        g = int(rng.uniform(1.5e6, 7e6))
        gc = float(rng.uniform(0.30, 0.68))
        cov = float(np.clip(80.0 * np.exp(0.5 * rng.standard_normal()), 30.0, 200.0))
        out.append((g, gc, cov))
    return out


def make_batch_isolates(n: int, threads: int):
    """n seeded isolates of the batched workload as packed read banks."""
    import concurrent.futures as cf
    import fastbact_amd as F
    shapes = batch_shapes(n)

    def gen(i):
        g, gc, cov = shapes[i]
        iso = F.synth_isolate(seed=0xFB03 * 1000 + i, genome_len=g, gc=gc, coverage=cov, read_len=150,
                              insert_mean=350.0, insert_sd=35.0, sub_rate=0.003, indel_rate=0.0)
        return iso.bank()

    t = time.time()
    with cf.ThreadPoolExecutor(max(1, threads)) as ex:  # the generator releases the GIL
        banks = list(ex.map(gen, range(n)))
    log(f"{n} batched-workload isolates generated ({sum(b.total_bases() for b in banks) / 1e9:.1f} Gbases, "
        f"{time.time() - t:.1f} s)")
    return shapes, banks


def waves_by_footprint(F, shapes, devs, k: int, budget: int):
    """Consecutive isolates grouped so that each wave's footprint fits the budget."""
    waves, cur, used = [], [], 0
    for i, d in enumerate(devs):
        fp = F.batch_footprint(shapes[i][0], d.total_bases(), d.total_reads(), k, cms_width=BATCH_WIDTH,
                               counter_bits=16, hll_p=BATCH_HLL_P)
        if cur and used + fp > budget:
            waves.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += fp
    if cur:
        waves.append(cur)
    return waves


def bench_batched(ctx, F, shapes, devs, reps: int) -> dict:
    """The batched form (waves of isolates per launch, launch shapes from the
    autotuner) against the plain per-isolate drop-in on the same isolates."""
    bases = sum(d.bases_for_k(BATCH_K) for d in devs)
    budget = int(0.6 * ctx.memory()[0])
    waves = waves_by_footprint(F, shapes, devs, BATCH_K, budget)
    log(f"batched form: {len(devs)} isolates in {len(waves)} waves (budget {budget / 2**30:.0f} GiB)")
    tune = None
    batch_ms = []
    for wi, w in enumerate(waves):
        b = F.Batch(ctx, [devs[i] for i in w], [shapes[i][0] for i in w], cms_width=BATCH_WIDTH, counter_bits=16,
                    hll_p=BATCH_HLL_P)
        if wi == 0:
            tune = b.autotune(BATCH_K, None)
            log(f"autotune: {tune}")
        else:
            b.run(BATCH_K, None)  # warm-up
        ms = []
        for _ in range(reps):
            t = time.perf_counter()
            b.run(BATCH_K, None)
            ms.append((time.perf_counter() - t) * 1e3)
        batch_ms.append(min(ms))
        b.close()
    plain_ms = []
    for _ in range(reps):
        t = time.perf_counter()
        for d in devs:
            h = F.build_solid_handle(ctx, d, BATCH_K, target_solid_cov=None, genome_size=None, cms_width=BATCH_WIDTH,
                                     counter_bits=16, hll_p=BATCH_HLL_P)
            h.close()
        ctx.synchronize()
        plain_ms.append((time.perf_counter() - t) * 1e3)
    tb, tp = sum(batch_ms), min(plain_ms)
    return {"workload": f"{len(devs)} isolates: genome U[1.5,7] Mbp, GC U[0.30,0.68], coverage lognormal median 80x "
                        f"sigma 0.5 clipped 30-200x, PE150 0.3% subs; k=55; CMS 4x2^23 u16 + HLL p14 per isolate",
            "bases": bases, "waves": len(waves), "batched_ms": round(tb, 1),
            "batched_gbases_s": round(bases / tb / 1e6, 3), "plain_per_isolate_ms": round(tp, 1),
            "plain_per_isolate_gbases_s": round(bases / tp / 1e6, 3), "batched_over_plain": round(tp / tb, 3),
            "autotune": tune}


def bench_small_mix(ctx, F, n: int, threads: int, reps: int = 2, seed: int = 0xFB07) -> dict:
    """Where batching pays: n small isolates (genome U[50, 300] kbp, GC
    U[0.30, 0.68], coverage U[30, 100]x, PE150 0.3% subs; k = 55, each CMS
    sized by the reference's genome-size rule, HLL p14), which one at a time
    cannot fill the device; batched in waves against plain per-isolate calls."""
    import concurrent.futures as cf
    rng = np.random.default_rng(seed)
    shapes = [(int(rng.uniform(5e4, 3e5)), float(rng.uniform(0.30, 0.68)), float(rng.uniform(30.0, 100.0)))
              for _ in range(n)]

    def gen(i):
        g, gc, cov = shapes[i]
        return F.synth_isolate(seed=seed * 1000 + i, genome_len=g, gc=gc, coverage=cov, read_len=150,
                               insert_mean=350.0, insert_sd=35.0, sub_rate=0.003, indel_rate=0.0).bank()

    with cf.ThreadPoolExecutor(max(1, threads)) as ex:
        banks = list(ex.map(gen, range(n)))
    devs = [F.DeviceBank(ctx, b) for b in banks]
    del banks
    ctx.synchronize()
    bases = sum(d.bases_for_k(BATCH_K) for d in devs)
    kw = dict(counter_bits=16, hll_p=BATCH_HLL_P)
    budget = int(0.6 * ctx.memory()[0])
    waves, cur, used = [], [], 0
    for i, d in enumerate(devs):
        fp = F.batch_footprint(shapes[i][0], d.total_bases(), d.total_reads(), BATCH_K, **kw)
        if cur and used + fp > budget:
            waves.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += fp
    waves.append(cur)
    batch_ms = []
    for w in waves:
        b = F.Batch(ctx, [devs[i] for i in w], [shapes[i][0] for i in w], **kw)
        b.run(BATCH_K, None)  # warm-up, captures the graphs
        ms = []
        for _ in range(reps):
            t = time.perf_counter()
            b.run(BATCH_K, None)
            ms.append((time.perf_counter() - t) * 1e3)
        batch_ms.append(min(ms))
        b.close()
    plain_ms = []
    for _ in range(reps):
        t = time.perf_counter()
        for i, d in enumerate(devs):
            F.build_solid_handle(ctx, d, BATCH_K, target_solid_cov=None, genome_size=shapes[i][0], **kw).close()
        ctx.synchronize()
        plain_ms.append((time.perf_counter() - t) * 1e3)
    for d in devs:
        d.close()
    tb, tp = sum(batch_ms), min(plain_ms)
    return {"workload": f"{n} isolates: genome U[50,300] kbp, GC U[0.30,0.68], coverage U[30,100]x, PE150 0.3% subs; "
                        f"k=55; CMS from the genome-size rule (u16) + HLL p14 per isolate",
            "bases": bases, "waves": len(waves), "batched_ms": round(tb, 1),
            "batched_gbases_s": round(bases / tb / 1e6, 3), "plain_per_isolate_ms": round(tp, 1),
            "plain_per_isolate_gbases_s": round(bases / tp / 1e6, 3), "batched_over_plain": round(tp / tb, 3)}


def bench_serving(F, shapes, banks, n_req: int, rate: float, seed: int = 0xFB05) -> dict:
    """Poisson stream of isolate requests (k ladder 21/55/77 plus the genome
    estimate) through the serving loop; latency from submission to the last
    pass's solid set on the host."""
    rng = np.random.default_rng(seed)
    picks = rng.integers(0, len(banks), size=n_req)
    gaps = rng.exponential(1.0 / rate, size=n_req)
    kw = dict(cms_width=BATCH_WIDTH, counter_bits=16, hll_p=BATCH_HLL_P)
    out = {}
    with F.Server(0, (21, 55, 77), (None, 4, 4), **kw) as sv:
        # warm-up request (first use of the kernels and the pool)
        sv.wait(sv.submit(banks[0].pe_forward, banks[0].pe_reverse), download=False)
        t0 = time.perf_counter()
        tickets = []
        for i in range(n_req):
            target = t0 + float(np.sum(gaps[: i + 1]))
            while time.perf_counter() < target:
                time.sleep(min(0.002, max(0.0, target - time.perf_counter())))
            b = banks[int(picks[i])]
            tickets.append(sv.submit(b.pe_forward, b.pe_reverse))
            if (i + 1) % 24 == 0:
                log(f"serving: {i + 1}/{n_req} requests submitted")
        res = [sv.wait(t, download=False) for t in tickets]
        wall = time.perf_counter() - t0
        st = sv.stats()
        lat = sorted(r.latency * 1e3 for r in res)
        log(f"serving: Poisson stream done in {wall:.1f} s")
        bases = sum(3 * banks[int(p)].total_bases() for p in picks)
        out["poisson"] = {"requests": n_req, "rate_per_s": rate, "wall_s": round(wall, 2),
                          "isolates_per_hour": round(n_req / wall * 3600, 1),
                          "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 1),
                                         "p99": round(float(np.percentile(lat, 99)), 1)},
                          "gbases_s": round(bases / wall / 1e9, 3), "max_batch": st["max_batch"],
                          "max_ladder_positions": st["max_groups"],
                          "worker_s": {n: round(st[n], 2) for n in ("start_s", "pass_s", "fetch_s")}}
        st0 = st
        # capacity: the same requests all at once
        t0 = time.perf_counter()
        tickets = [sv.submit(banks[int(p)].pe_forward, banks[int(p)].pe_reverse) for p in picks]
        res = [sv.wait(t, download=False) for t in tickets]
        wall = time.perf_counter() - t0
        st = sv.stats()
        lat = sorted(r.latency * 1e3 for r in res)
        out["saturated"] = {"requests": n_req, "wall_s": round(wall, 2),
                            "isolates_per_hour": round(n_req / wall * 3600, 1),
                            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 1),
                                           "p99": round(float(np.percentile(lat, 99)), 1)},
                            "gbases_s": round(bases / wall / 1e9, 3), "max_batch": st["max_batch"],
                            "worker_s": {n: round(st[n] - st0[n], 2) for n in ("start_s", "pass_s", "fetch_s")}}
    out["workload"] = ("requests drawn from the batched-workload isolates, each the k ladder 21/55/77 (gates None/4/4) "
                       "after estimate_genome; CMS 4x2^23 u16, HLL p14; latency = submit to solid sets on the host")
    return out


def bench_compaction(ctx, dev, F, cpu: bool) -> dict:
    """Unitig compaction (build.rs:330-597) of the first pass's solid set:
    device walks, downloads and device overlap edges, with the oracle's
    single-threaded compaction of the same set timed beside it."""
    k, tgt = PASSES[0]
    h = F.build_solid_handle(ctx, dev, k, target_solid_cov=tgt, genome_size=GENOME["genome_len"], cms_width=WIDTH,
                             counter_bits=COUNTER_BITS, hll_p=HLL_P)
    try:
        F.compact_unitigs(ctx, h)  # warm-up
        n_solid = h.count()
        t = time.perf_counter()
        g = F.compact_unitigs(ctx, h)
        wall = (time.perf_counter() - t) * 1e3
        s = h.download(k, with_cms=False) if cpu else None
    finally:
        h.close()
    out = {"k": k, "solid_kmers": n_solid, "unitigs": len(g["seqs"]), "circular": int(g["circular"].sum()),
           "bases": int(sum(len(x) for x in g["seqs"])), "edges": int(g["src"].size),
           "device_walk_ms": round(g["timings_ms"][0], 3), "download_ms": round(g["timings_ms"][1], 3),
           "device_edges_ms": round(g["timings_ms"][2], 3), "wall_ms": round(wall, 3),
           "solid_kmers_per_s_device": round(n_solid / max(g["timings_ms"][0], 1e-6) * 1e3, 1)}
    if s is not None:
        from oracle import oracle as O
        t = time.perf_counter()
        r = O.compact(s.keys, s.cov, s.mask, k)
        out["cpu_baseline_ms"] = round((time.perf_counter() - t) * 1e3, 1)
        out["cpu_baseline"] = "oracle compaction (single thread) of the same solid set"
        out["equal_to_oracle"] = bool(r["seqs"] == g["seqs"] and np.array_equal(r["cov"], g["cov"])
                                      and np.array_equal(r["src"], g["src"]) and np.array_equal(r["dst"], g["dst"]))
    return out, g


def bench_next_rows(ctx, dev, F, iso, bank, g21, cpu: bool, cpu_pairs: int = 200_000) -> dict:
    """The rows either side of the path on the benchmark isolate, each timed
    on the device with the oracle's single-threaded form beside it:
    ingest (trim / merge / subsample of the isolate's pairs as parsed FASTQ
    records), the multi-k upgrade from the first pass's graph to k 55, the PE
    mapper and consensus polishing on the k 55 graph."""
    out = {}
    n, L = iso.r1.shape
    rng = np.random.default_rng(0xFB06)
    q1 = rng.integers(63, 74, size=(n, L), dtype=np.uint8)  # Phred 30-40
    q2 = rng.integers(63, 74, size=(n, L), dtype=np.uint8)
    F.ingest_arrays(ctx, iso.r1[:1000], q1[:1000], iso.r2[:1000], q2[:1000])[0].close()  # warm-up
    t = time.perf_counter()
    ib, st = F.ingest_arrays(ctx, iso.r1, q1, iso.r2, q2)
    wall = (time.perf_counter() - t) * 1e3
    out["ingest"] = {"pairs": n, "retained_pairs": st["retained_read_pairs"], "merged_pairs": st["merged_pairs_count"],
                     "device_ms": round(st["device_ms"], 2), "wall_ms": round(wall, 2),
                     "pairs_per_s_device": round(n / max(st["device_ms"], 1e-6) * 1e3)}
    ib.close()
    if cpu:
        from oracle import oracle as O
        m = min(cpu_pairs, n)
        recs1 = O.records([bytes(r) for r in iso.r1[:m]], [bytes(q) for q in q1[:m]])
        recs2 = O.records([bytes(r) for r in iso.r2[:m]], [bytes(q) for q in q2[:m]])
        t = time.perf_counter()
        ref = O.ingest_records(recs1, recs2, min_len=30, min_qual=20)
        cpu_ms = (time.perf_counter() - t) * 1e3
        sb, sst = F.ingest_arrays(ctx, iso.r1[:m], q1[:m], iso.r2[:m], q2[:m])
        got = sb.download()
        same = all(np.array_equal(a.data[:len(d)], d[:len(a.data)]) and np.array_equal(a.lengths, ln)
                   for a, (d, _, ln) in zip(got.arenas(), ref["arenas"]))
        sb.close()
        out["ingest"].update({"cpu_baseline_pairs_per_s": round(m / cpu_ms * 1e3), "cpu_sample_pairs": m,
                              "equal_to_oracle_on_sample": bool(same and {k: sst[k] for k in ref["stats"]} == ref["stats"])})
    # multi-k upgrade 21 -> 55 (filtered bank + prior contigs), then map and polish on its graph
    kw = dict(genome_size=GENOME["genome_len"], cms_width=WIDTH, counter_bits=COUNTER_BITS)
    t = time.perf_counter()
    up = F.upgrade_multik(ctx, dev, g21, PASSES[0][0], 55, keep_active=True, **kw)
    up_ms = (time.perf_counter() - t) * 1e3
    g55 = up["graph"]
    out["multik"] = {"k1": PASSES[0][0], "k2": 55, "wall_ms": round(up_ms, 2), "executed": up["executed"],
                     "kept_pairs": up["kept_pairs"], "pairs": n, "prior_contigs": up["prior_contigs"],
                     "unitigs_before": up["active_unitigs_before"], "unitigs_after": up["active_unitigs_after"]}
    up["active"].close()
    t = time.perf_counter()
    fb, km, kp = F.filter_active_reads(ctx, dev, g21, PASSES[0][0])
    out["multik"]["filter_ms"] = round((time.perf_counter() - t) * 1e3, 2)
    fb.close()
    t = time.perf_counter()
    mp = F.map_paired_reads(ctx, dev, g55, 55)
    out["mapper"] = {"k": 55, "wall_ms": round((time.perf_counter() - t) * 1e3, 2), "links": int(mp["links"].size),
                     "insert_sizes": int(mp["sizes"].size), "dist": mp["dist"]}
    t = time.perf_counter()
    pseqs, pst = F.polish_contigs(ctx, dev, g55, 500)
    out["polish"] = {"wall_ms": round((time.perf_counter() - t) * 1e3, 2), **pst}
    if cpu:
        from oracle import oracle as O
        arenas = [(a.data, a.offsets, a.lengths) for a in bank.arenas()]
        t = time.perf_counter()
        fr = O.filter_active_reads(g21, PASSES[0][0], arenas, pack_reads=False)
        out["multik"]["filter_cpu_baseline_ms"] = round((time.perf_counter() - t) * 1e3, 1)
        out["multik"]["filter_equal_to_oracle"] = bool(np.array_equal(fr["keep_pairs"], kp))
        t = time.perf_counter()
        mr = O.map_paired_reads(g55, arenas[1], arenas[2], 55)
        out["mapper"]["cpu_baseline_ms"] = round((time.perf_counter() - t) * 1e3, 1)
        out["mapper"]["equal_to_oracle"] = bool(mr["links"].size == mp["links"].size and
                                                np.array_equal(mr["sizes"], mp["sizes"]) and mr["dist"] == mp["dist"])
        t = time.perf_counter()
        pr = O.polish_contigs(g55, arenas[:2], 500)
        out["polish"]["cpu_baseline_ms"] = round((time.perf_counter() - t) * 1e3, 1)
        out["polish"]["equal_to_oracle"] = bool(pr["seqs"] == pseqs and pr["stats"] == pst)
        out["cpu_baseline"] = "oracle single-threaded forms; ingest on the first pairs (cpu_sample_pairs)"
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-pairs", type=int, default=300_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--placement", type=int, default=0, help="pass-1 sketch placement: 0 auto, 1 global, 2 LDS bins")
    ap.add_argument("--batch-isolates", type=int, default=64, help="isolates of the batched workload (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=96,
                    help="requests of the Poisson serving stream on batched-workload isolates (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="mean request arrivals per second")
    ap.add_argument("--no-next-rows", action="store_true",
                    help="skip compaction and the next rows (a short run of the timed steps alone, for profiling)")
    ap.add_argument("--no-atomic-sweep", action="store_true",
                    help="skip the random access / atomic rate sweep (k_atomic_rate)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the steps, copy a buffer of this many MiB device to device once (a known-size "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the k-mer spectrum path is measured per GPU); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2
    import fastbact_amd as F

    iso, bank = make_isolate()
    ctx = F.Context(0)
    t = time.time()
    dev = F.DeviceBank(ctx, bank)
    ctx.synchronize()
    h2d_s = time.time() - t
    bases_step = sum(dev.bases_for_k(k) for k, _ in PASSES)
    log(f"bank in HBM ({bank.total_bases() / 4e6:.0f} MB packed, upload {h2d_s:.2f} s); {bases_step} bases per step")

    for i in range(args.warmup):
        run_step(ctx, dev, F, args.placement)
        log(f"warmup step {i + 1}/{args.warmup} done")
    ctx.synchronize()
    step_ms, infos = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        infos.append(run_step(ctx, dev, F, args.placement))
        step_ms.append((time.perf_counter() - ts) * 1e3)
        log(f"step {i + 1}/{args.steps}: {step_ms[-1]:.1f} ms")
    ctx.synchronize()
    total_s = time.perf_counter() - t0
    value = bases_step * args.steps / total_s / 1e9

    # per-kernel event times and algorithmic bytes over the timed steps
    per_pass = []
    k_ms = {n: 0.0 for n in KERNEL_NAMES}
    k_bytes = {n: 0.0 for n in KERNEL_NAMES}
    model_bytes = 0.0
    for j, (k, _) in enumerate(PASSES):
        st, solid = infos[-1][j][1], infos[-1][j][2]
        bases = dev.bases_for_k(k)
        kb = kernel_bytes(k, st, bases, solid)
        kt = {n: sum(info[j][3][n] for info in infos) / args.steps for n in KERNEL_NAMES}
        for n in KERNEL_NAMES:
            k_ms[n] += kt[n] * args.steps
            k_bytes[n] += kb[n] * args.steps
        model_bytes += pass_model_bytes(k, st, bases, solid) * args.steps
        per_pass.append({"k": k,
                         "pass1_ms": sum(info[j][0][0] for info in infos) / args.steps,
                         "pass2_ms": sum(info[j][0][1] for info in infos) / args.steps,
                         "pass3_sort_ms": sum(info[j][0][2] for info in infos) / args.steps,
                         "kernel_ms": {n: round(v, 3) for n, v in kt.items()},
                         "windows": st["windows"], "pass2_records": st["pass2_records"],
                         "raw_cov_sum": st["records"], "solid_kmers": solid, "pass2_route": st["pass2_route"],
                         "lds_bins": st["lds_bins"], "est_keys": st["est_keys"]})
    dom = max(KERNEL_NAMES, key=lambda n: k_ms[n])
    launches = args.steps * len(PASSES)
    avg_ms = k_ms[dom] / launches
    achieved = k_bytes[dom] / (k_ms[dom] / 1e3) / 1e9 if k_ms[dom] > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": KERNEL_NAMES[dom],
                "avg_launch_ms": round(avg_ms, 3), "launches": launches,
                "algorithmic_bytes_per_launch": round(k_bytes[dom] / launches)}
    model_gbs = model_bytes / total_s / 1e9
    pass_model = {"achieved": round(model_gbs, 1), "unit": "GB/s", "frac": round(model_gbs / HBM_PEAK_GBS, 4),
                  "bytes_per_step": round(model_bytes / args.steps)}

    if args.copy_calibration_mib:
        copy_calibration(ctx, args.copy_calibration_mib)
    roofline["traffic"], roofline["traffic_source"] = measured_traffic(dom)

    compaction = next_rows = None
    if not args.no_next_rows:
        compaction, g21 = bench_compaction(ctx, dev, F, cpu=not args.no_cpu_baseline)
        log(f"compaction: {compaction}")
        next_rows = bench_next_rows(ctx, dev, F, iso, bank, g21, cpu=not args.no_cpu_baseline)
        log(f"next rows: {next_rows}")
        del g21
    atomics = atomic_sweep(ctx) if not args.no_atomic_sweep else None

    batched = served = None
    if args.batch_isolates or args.serve_requests:
        shapes, banks = make_batch_isolates(max(args.batch_isolates, 1 if args.serve_requests else 0),
                                            min(16, os.cpu_count() or 1))
        if args.batch_isolates:
            devs = [F.DeviceBank(ctx, b) for b in banks]
            ctx.synchronize()
            batched = bench_batched(ctx, F, shapes, devs, reps=2)
            log(f"batched: {batched['batched_gbases_s']} Gbases/s vs plain {batched['plain_per_isolate_gbases_s']}")
            for d in devs:
                d.close()
            ctx.trim()  # hand the waves' buffers back to the device
            batched["small_isolate_mix"] = bench_small_mix(ctx, F, 256, min(16, os.cpu_count() or 1))
            log(f"small-isolate mix: {batched['small_isolate_mix']}")
            ctx.trim()
        if args.serve_requests:
            served = bench_serving(F, shapes, banks, args.serve_requests, args.serve_rate)
            log(f"serving: {served['poisson']}")
        del banks

    cpu = None
    if not args.no_cpu_baseline:
        # the host's share of cores (16 per GPU on these boxes), then the
        # reference's default --threads 8 (cli.rs:26-27)
        share = args.cpu_threads or min(16, host_info()["allowed_cpus"])
        counts = [share] + ([8] if share != 8 else [])
        log(f"CPU baseline: oracle threaded form at {counts} threads, {args.cpu_pairs} pairs")
        cpu = cpu_baseline(ctx, F, bank, args.cpu_pairs, counts)

    sm = sorted(step_ms)
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gbases/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_s / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "plain full isolate: 5 Mbp GC 0.508, PE150 100x (1,666,667 pairs), insert N(400,40), "
                        "0.3% subs + 0.01% indels; k=21/55/77 passes; CMS 4x2^24 u16; HLL p16; "
                        "solid gates None/4/4; sorted solid set",
            "bases_per_step": bases_step,
            # per-isolate latency of the served form (Poisson arrivals, submit to
            # solid sets on the host), when the serving stream ran
            "isolate_latency_ms": served["poisson"]["latency_ms"] if served else None,
            "plain_step_ms": {"p50": round(float(np.percentile(sm, 50)), 3),
                              "p99": round(float(np.percentile(sm, 99)), 3)},
            "h2d_upload_s": round(h2d_s, 3),
            "passes": per_pass,
        },
        "roofline": roofline,
        "pass_model_roofline": pass_model,
        "cpu_baseline": cpu,
        "compaction": compaction,
        "next_rows": next_rows,
        "batched": batched,
        "serving": served,
        "atomic_sweep": atomics,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
