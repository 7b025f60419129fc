#!/usr/bin/env python3
"""Benchmark of the tessellation hot path on one MI355X.

A step is one full compute_tessellation (contacts, per-ball SAS and volumes)
of the 100,000-ball workload with the balls already resident in HBM: random
sequential addition in a sphere at 0.056 balls/A^3, 2.0 A minimum centre
distance, radii from {1.52, 1.55, 1.70, 1.80} A, probe 1.4, f64.

The JSON line carries the throughput, the roofline of the dominant kernel
(k_clip, which classifies and clips every candidate pair: algorithmic FP64
work = FLOP per candidate pair x pairs per launch, over its HIP-event time on
the library stream) and the CPU baseline (the
oracle's C++ restatement of the reference on this host's cores, on the same
input).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "voronota-ltr_amd"))

METRIC = "balls/s full tessellation f64 @100k: plain vs batched; fwd+adjoint/fwd ratio"
# Algorithmic FP64 work of one candidate pair (clip + solid angles + pyramid
# volumes), counted by the oracle's counting scalar on this workload
# (add/sub/mul/div/sqrt = 1, fma = 2; acos/sin/cos excluded).
FLOP_PER_PAIR = 2914.39
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--seed", type=int, default=100000)
    ap.add_argument("--probe", type=float, default=1.4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the tessellation of one structure or one batch runs on a single MI355X).", file=sys.stderr)
        sys.exit(2)

    import ctypes as C

    import numpy as np

    from voronota_ltr import _hip, _native, workload

    balls, _ = workload.rsa_sphere(args.n, seed=args.seed)
    lib = _native.load_library()
    ctx = _native.context(0)
    dbuf = _hip.DeviceBuffer(balls.nbytes)
    dbuf.upload(np.ascontiguousarray(balls))

    def step(contacts_only=0):
        rc = lib.vltr_compute_tessellation_device(ctx, dbuf.ptr, args.n, args.probe, None, 0, contacts_only)
        _native.check(rc, "vltr_compute_tessellation_device")

    phases = ["grid", "neighbors", "pairs", "clip", "slow", "cells", "total"]

    def timed(contacts_only):
        for _ in range(args.warmup):
            step(contacts_only)
        _hip.synchronize()
        acc = np.zeros(len(phases))
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(contacts_only)
            acc += np.array([_native.last_timings()[k] for k in phases])
        _hip.synchronize()
        wall = time.perf_counter() - t0
        return wall / args.steps * 1e3, dict(zip(phases, (acc / args.steps).round(4).tolist())), _native.last_counts()

    ms_full, ph_full, counts = timed(0)
    ms_co, ph_co, counts_co = timed(1)

    value = args.n / (ms_full / 1e3)
    cut_ms = ph_full["clip"]
    flops = FLOP_PER_PAIR * counts["pairs"]
    achieved = flops / (cut_ms / 1e3) / 1e12 if cut_ms > 0 else 0.0

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle

        threads = min(os.cpu_count() or 1, 16)
        t1 = pyoracle.time_tessellation(balls, args.probe, threads=threads, reps=1)
        reps = max(1, int(args.cpu_seconds / max(t1, 1e-3)))
        tcpu = pyoracle.time_tessellation(balls, args.probe, threads=threads, reps=reps)
        cpu = {"value": args.n / tcpu, "unit": "balls/s", "cores": threads, "kind": "port",
               "sample": f"full compute_tessellation of the same {args.n}-ball input, {reps} run(s) "
                         f"after one warm-up, C++ restatement (-O3, OpenMP over balls and pairs)"}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "balls/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_full, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"plain full compute_tessellation of {args.n} RSA balls in a sphere "
                        f"(0.056 balls/A^3, min distance 2.0 A, radii 1.52-1.80 A, seed {args.seed}), "
                        f"probe {args.probe}, inputs resident in HBM",
            "balls": args.n,
            "probe": args.probe,
        },
        "roofline": {
            "bound": "mfma",
            "kernel": "k_clip",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": None,
            "note": "FP64 vector work (no matrix op on this path); flops = "
                    f"{FLOP_PER_PAIR} per candidate pair x {counts['pairs']} pairs",
        },
        "cpu_baseline": cpu,
        "phases_ms": ph_full,
        "counts": counts,
        "contacts_only": {"balls_per_s": round(args.n / (ms_co / 1e3), 1), "ms_per_step": round(ms_co, 4),
                          "phases_ms": ph_co, "contacts": counts_co["contacts"]},
    }
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
