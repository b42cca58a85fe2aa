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


def pass_bytes(k: int, windows: int, records: int, n_reads: int, bases: int) -> tuple[float, float]:
    """Algorithmic HBM bytes of the pass-1 and pass-2 kernels.

    Per read: packed bases L/4 plus 12 bytes of offset and length.  Per
    window, pass 1: the 4-row CMS read-modify-write (2 x 4 x c, c = counter
    bytes) and the HLL register read-modify-write (2).  Pass 2: the 4-row CMS
    query (4 x c) per window and, per adjacency record, the key plus its
    cov/edge u16 read-modify-write (K_B + 4)."""
    c = 2 if COUNTER_BITS == 16 else 1
    stream = bases / 4.0 + 12.0 * n_reads
    p1 = stream + windows * (8 * c + 2)
    p2 = stream + windows * 4 * c + records * (8 * ref_words(k) + 4)
    return p1, p2


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
    info = [(h.timings(), h.stats(), h.count()) for h in handles]
    for h in handles:
        h.close()
    return info


def cpu_baseline(bank, pairs: int, threads: int) -> dict:
    from oracle import oracle as O
    import fastbact_amd as F
    n = min(pairs, len(bank.pe_forward))
    sub = F.ReadBank(F.BitSeqArena(), _head(bank.pe_forward, n), _head(bank.pe_reverse, n))
    arenas = [(a.data, a.offsets, a.lengths) for a in sub.arenas()]
    bases = sum(int(a.lengths[a.lengths >= k].sum(dtype=np.uint64)) for a in sub.arenas() for k, _ in PASSES)
    t = time.time()
    for k, tgt in PASSES:
        O.build_solid(arenas, k, target_solid_cov=tgt, width=WIDTH, counter_bits=COUNTER_BITS, hll_p=HLL_P,
                      threads=threads)
    dt = time.time() - t
    return {"value": bases / dt / 1e9, "unit": "Gbases/s", "cores": threads, "kind": "port",
            "sample": f"first {n} pairs of the same isolate ({sub.total_bases()} bases), the same 3 passes "
                      f"(k=21/55/77, CMS 4x2^24 u16, HLL p16), oracle threaded form, {dt:.1f} s"}


def _head(a, n):
    import fastbact_amd as F
    if n >= len(a):
        return a
    end = int(a.offsets[n - 1] + a.lengths[n - 1]) if n else 0
    return F.BitSeqArena(a.data[: (end + 31) // 32 + 1].copy(), a.offsets[:n].copy(), a.lengths[:n].copy())


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-pairs", type=int, default=300_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--placement", type=int, default=0, help="pass-1 sketch placement: 0 auto, 1 global, 2 LDS bins")
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
    n_reads = dev.total_reads()
    p1_ms = p2_ms = p3_ms = 0.0
    p1_bytes = p2_bytes = 0.0
    per_pass = []
    for (k, _), j in zip(PASSES, range(len(PASSES))):
        t1 = sum(info[j][0][0] for info in infos)
        t2 = sum(info[j][0][1] for info in infos)
        t3 = sum(info[j][0][2] for info in infos)
        st = infos[-1][j][1]
        b1, b2 = pass_bytes(k, st["windows"], st["records"], n_reads, dev.bases_for_k(k))
        p1_ms += t1
        p2_ms += t2
        p3_ms += t3
        p1_bytes += b1 * args.steps
        p2_bytes += b2 * args.steps
        per_pass.append({"k": k, "pass1_ms": t1 / args.steps, "pass2_ms": t2 / args.steps,
                         "pass3_sort_ms": t3 / args.steps, "windows": st["windows"], "records": st["records"],
                         "solid_kmers": infos[-1][j][2], "pass2_route": st["pass2_route"],
                         "lds_bins": st["lds_bins"], "est_keys": st["est_keys"],
                         "pass1_GBps": b1 / (t1 / args.steps) / 1e6, "pass2_GBps": b2 / (t2 / args.steps) / 1e6})
    if p1_ms >= p2_ms:
        kname, kms, kbytes = "k_pass1 (CMS update + HLL)", p1_ms, p1_bytes
    else:
        kname, kms, kbytes = "k_pass2 (CMS gate + solid table)", p2_ms, p2_bytes
    launches = args.steps * len(PASSES)
    achieved = kbytes / launches / (kms / launches / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kname,
                "avg_launch_ms": round(kms / launches, 3),
                "algorithmic_bytes_per_launch": round(kbytes / launches)}

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"CPU baseline: oracle threaded form, {threads} threads, {args.cpu_pairs} pairs")
        cpu = cpu_baseline(bank, args.cpu_pairs, threads)

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
            "isolate_latency_ms": {"p50": round(float(np.percentile(sm, 50)), 3),
                                   "p99": round(float(np.percentile(sm, 99)), 3)},
            "h2d_upload_s": round(h2d_s, 3),
            "passes": per_pass,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
