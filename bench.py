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
# Algorithmic work per workload (FLOPs: add/sub/mul/div/sqrt = 1, fma = 2;
# acos/sin/cos counted separately as T-ops), counted by the oracle's counting
# scalar and committed as data by tests/golden/make_work_counts.py.
WORK_COUNTS = os.path.join(ROOT, "tests", "golden", "data", "work_counts.json")
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes one full-mode k_clip launch moves on the 100,000-ball workload
# (seed 100000, probe 1.4), per launch shape: FETCH_SIZE x 2 + WRITE_SIZE x 1
# KiB per dispatch, factors from a 256 MiB device copy read with the same
# counters (profiles/k_clip_100k_counters.txt). Other shapes and workloads:
# not measured.
K_CLIP_TRAFFIC_BYTES_100K = {"g8w1o4nq": 180.6e6, "g8w1o4q": 192.5e6}


def work_counts(n, seed, probe):
    """The counted work of this workload, or None when it is not one of the
    counted workloads."""
    with open(WORK_COUNTS) as f:
        data = json.load(f)
    for w in data["workloads"].values():
        if (w["balls"], w["seed"], w["probe"]) == (n, seed, probe):
            return w
    return None


def host_cpu():
    """What the CPU leg runs on: logical CPUs of the host, the CPUs this
    process may use, and the CPU model."""
    model = None
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
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "cpus_allowed": allowed, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "cpu_model": model}


def cpu_threads():
    """Threads of the CPU leg: every CPU this process may use, capped by
    OMP_NUM_THREADS when it is set (the share of the host a one-GPU job gets)."""
    info = host_cpu()
    t = info["cpus_allowed"]
    omp = info["omp_num_threads"]
    if omp and omp.isdigit() and int(omp) > 0:
        t = min(t, int(omp))
    return max(t, 1)


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
    ap.add_argument("--batch-structs", type=int, default=4096,
                    help="structures in the batched leg (0 skips it); sizes log-uniform in [200, 20000]")
    ap.add_argument("--dyn-balls", type=int, default=50000,
                    help="balls of the periodic dynamics leg (0 skips it)")
    ap.add_argument("--dyn-steps", type=int, default=20)
    ap.add_argument("--adj-balls", type=int, default=20000,
                    help="balls of the forward + adjoint leg (0 skips it)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default clip launch shape instead of timing every shape first")
    ap.add_argument("--clip-shape", default=None,
                    help="use this clip launch shape (a name from the library's table) instead of the autotuner's")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (reference kernel for the "
                         "HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the tessellation of one structure or one batch runs on a single MI355X).", file=sys.stderr)
        sys.exit(2)

    import ctypes as C

    import numpy as np

    from voronota_ltr import _hip, _native, workload

    if args.copy_mb > 0:
        nbytes = args.copy_mb << 20
        src, dst = _hip.DeviceBuffer(nbytes), _hip.DeviceBuffer(nbytes)
        _hip.copy_d2d(dst, src, nbytes)
        _hip.synchronize()
        src.free()
        dst.free()

    balls, _ = workload.rsa_sphere(args.n, seed=args.seed)
    lib = _native.load_library()
    ctx = _native.context(0)
    dbuf = _hip.DeviceBuffer(balls.nbytes)
    dbuf.upload(np.ascontiguousarray(balls))

    # contacts-only runs with group 1 for x < 0 and 0 otherwise: only contacts
    # between the two halves are kept
    groups = np.ascontiguousarray((balls[:, 0] < 0.0).astype(np.int32))
    gbuf = _hip.DeviceBuffer(groups.nbytes)
    gbuf.upload(groups)

    def step(contacts_only=0):
        gp, ng = (gbuf.ptr, args.n) if contacts_only else (None, 0)
        rc = lib.vltr_compute_tessellation_device(ctx, dbuf.ptr, args.n, args.probe, gp, ng, contacts_only)
        _native.check(rc, "vltr_compute_tessellation_device")

    phases = ["grid", "neighbors", "pairs", "clip", "slow", "cells", "total"]

    def timed(contacts_only):
        """Wall time of K steps (nothing else in the loop), then a separate
        pass of K steps that reads the per-phase HIP-event times after each."""
        for _ in range(args.warmup):
            step(contacts_only)
        _hip.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(contacts_only)
        _hip.synchronize()
        wall = time.perf_counter() - t0
        acc = np.zeros(len(phases))
        for _ in range(args.steps):
            step(contacts_only)
            t = _native.last_timings()
            acc += np.array([t[k] for k in phases])
        return wall / args.steps * 1e3, dict(zip(phases, (acc / args.steps).round(4).tolist())), _native.last_counts()

    tune = None
    if args.clip_shape is not None:
        _native.set_clip_variant(_native.clip_variants().index(args.clip_shape))
    elif not args.no_autotune:
        tune = autotune(lib, ctx, dbuf.ptr, np.array([0, args.n], dtype=np.uint64), args.probe)
    ms_full, ph_full, counts = timed(0)
    ms_co, ph_co, counts_co = timed(1)
    host_api = host_api_leg(args, lib, ctx, balls)

    value = args.n / (ms_full / 1e3)
    shape = _native.clip_variants()[_native.clip_variant()]
    roofline = clip_roofline(args, ph_full["clip"], counts, shape)

    batched = batched_leg(args, lib, ctx) if args.batch_structs > 0 else None
    dynamics = dynamics_leg(args) if args.dyn_balls > 0 else None
    adjoint = adjoint_leg(args, lib, ctx) if args.adj_balls > 0 else None

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle

        threads = cpu_threads()
        t1 = pyoracle.time_tessellation(balls, args.probe, threads=threads, reps=1)
        reps = max(5, int(args.cpu_seconds / max(t1, 1e-3)))
        tcpu = pyoracle.time_tessellation(balls, args.probe, threads=threads, reps=reps)
        t_single = pyoracle.time_tessellation(balls, args.probe, threads=1, reps=1)
        cpu = {"value": args.n / tcpu, "unit": "balls/s", "cores": threads, "kind": "port",
               "sample": f"full compute_tessellation of the same {args.n}-ball input, mean of {reps} runs after one "
                         f"warm-up, C++ restatement of the reference (-O3, OpenMP over balls and pairs as Rayon "
                         f"does), {threads} threads: every CPU this process may use, capped by OMP_NUM_THREADS",
               "host": host_cpu(),
               "single_thread_balls_per_s": round(args.n / t_single, 1),
               "calibration": cpu_calibration(pyoracle, threads)}

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
        "roofline": roofline,
        "cpu_baseline": cpu,
        "host_api": host_api,
        "clip_launch_shape": shape,
        "autotune": tune,
        "phases_ms": ph_full,
        "counts": counts,
        "grid_build": grid_build(ph_full["grid"], counts),
        "contacts_only": {"balls_per_s": round(args.n / (ms_co / 1e3), 1), "ms_per_step": round(ms_co, 4),
                          "phases_ms": ph_co, "contacts": counts_co["contacts"], "pairs": counts_co["pairs"],
                          "groups": "1 for x < 0, else 0 (inter-group contacts only)"},
    }
    if batched:
        out["batched"] = batched
    if dynamics:
        out["periodic_dynamics"] = dynamics
    if adjoint:
        out["adjoint"] = adjoint
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 2)
    print(json.dumps(out), flush=True)


def clip_roofline(args, clip_ms, counts, shape):
    """FP64 roofline of k_clip: the counted FLOPs of contact construction per
    candidate pair (tests/golden/data/work_counts.json) times the pairs the
    launch clips itself (candidate pairs minus those it hands to the
    fallback kernels), over its HIP-event time on the library stream."""
    w = work_counts(args.n, args.seed, args.probe)
    clipped = counts["pairs"] - counts["slow_pairs"]
    if w is None or clip_ms <= 0:
        return {"bound": "fp64-valu", "kernel": "k_clip", "achieved": None, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": None, "traffic": None,
                "note": "work not counted for this workload (tests/golden/make_work_counts.py)"}
    m = w["mode0"]
    flop_per_pair = m["flops_contacts"] / w["candidate_pairs"]
    tops_per_pair = m["tops_contacts"] / w["candidate_pairs"]
    flops = flop_per_pair * clipped
    achieved = flops / (clip_ms / 1e3) / 1e12
    traffic = K_CLIP_TRAFFIC_BYTES_100K.get(shape) if (args.n, args.seed, args.probe) == (100000, 100000, 1.4) else None
    return {"bound": "fp64-valu", "kernel": "k_clip", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic,
            "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)",
            "flops_per_launch": round(flops), "flop_per_pair": round(flop_per_pair, 2),
            "tops_per_launch": round(tops_per_pair * clipped), "tops_per_s": round(tops_per_pair * clipped /
                                                                                 (clip_ms / 1e3), 1),
            "clip_ms": clip_ms, "pairs_clipped": clipped,
            "note": "FP64 vector work, no matrix op on this path. FLOPs counted by the oracle's counting scalar "
                    "(add/sub/mul/div/sqrt = 1, fma = 2); acos/sin/cos are T-ops, reported apart"}


def host_api_leg(args, lib, ctx, balls):
    """End to end through the host entry point (vltr_compute_tessellation:
    host balls in, host contacts and cells out, PCIe copies included)."""
    import ctypes as C

    import numpy as np

    from voronota_ltr import _native

    b = np.ascontiguousarray(balls)
    bp = b.ctypes.data_as(C.c_void_p)

    def once():
        h = C.c_void_p()
        _native.check(lib.vltr_compute_tessellation(ctx, bp, len(b), args.probe, None, None, 0, 0, C.byref(h)),
                      "vltr_compute_tessellation")
        nc = lib.vltr_result_num_contacts(h)
        lib.vltr_result_free(h)
        return nc

    for _ in range(max(1, args.warmup)):
        once()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        once()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    return {"balls_per_s": round(args.n / (ms / 1e3), 1), "ms_per_call": round(ms, 4),
            "note": "vltr_compute_tessellation: host balls in, host contacts and cells out (result object built "
                    "and freed), full mode"}


# The reference's published times (README.md:375-381, Apple M4, 10 cores,
# Rust/Rayon): balls_2zsk 12 ms, balls_3dlb 30 ms, 5.9x over one thread.
PUBLISHED_M4_MS = {"balls_2zsk": 12.0, "balls_3dlb": 30.0}


def cpu_calibration(pyoracle, threads):
    """The oracle on the reference's benchmark inputs, all cores and one core,
    beside the published numbers, to place this host's CPU path."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import load_xyzr

    out = {}
    for name, ms in PUBLISHED_M4_MS.items():
        b = load_xyzr(name + ".xyzr")
        pyoracle.time_tessellation(b, 1.4, threads=threads, reps=1)
        t_all = pyoracle.time_tessellation(b, 1.4, threads=threads, reps=5)
        t_one = pyoracle.time_tessellation(b, 1.4, threads=1, reps=2)
        out[name] = {"balls": len(b), "ms_all_cores": round(t_all * 1e3, 2), "ms_one_core": round(t_one * 1e3, 2),
                     "published_rayon_m4_10_cores_ms": ms}
    return out


# Algorithmic bytes of the cell-list build: each ball's sphere read (32 B),
# its sorted copy (32 B) and permutation entry (4 B) written, one 4-byte
# offset per cell-hash bucket.
GRID_BYTES_PER_BALL = 68.0
GRID_BYTES_PER_BUCKET = 4.0


def grid_build(grid_ms, counts):
    """HBM roofline of the grid phase (prepare, keys, counting sort, scatter)
    from its HIP-event time."""
    nbytes = GRID_BYTES_PER_BALL * counts["spheres"] + GRID_BYTES_PER_BUCKET * counts["buckets"]
    gbs = nbytes / (grid_ms * 1e-3) / 1e9
    return {"ms": round(grid_ms, 4), "bytes": int(nbytes), "achieved_GBs": round(gbs, 1), "peak_GBs": HBM_PEAK_GBS,
            "frac": round(gbs / HBM_PEAK_GBS, 4), "buckets": counts["buckets"]}


def autotune(lib, ctx, dptr, offsets, probe, repeats=3):
    """The clip kernel's launch shapes timed on this workload (an untimed
    round over all shapes, then `repeats` rounds over all of them; the median
    pipeline time per shape; outside every timed region); the fastest stays
    set on the context."""
    import ctypes as C

    import numpy as np

    from voronota_ltr import _native

    nv = lib.vltr_clip_variant_count()
    ms = np.zeros(nv)
    chosen = C.c_int(0)
    _native.check(lib.vltr_context_autotune(ctx, dptr, offsets.ctypes.data_as(C.POINTER(C.c_uint64)),
                                            len(offsets) - 1, probe, 0, repeats, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                            nv, C.byref(chosen)), "vltr_context_autotune")
    names = _native.clip_variants()
    return {"chosen": names[chosen.value], "pipeline_ms": {names[v]: round(float(ms[v]), 4) for v in range(nv)}}


def batched_leg(args, lib, ctx):
    """Many independent structures: one batched pass vs one call per
    structure, both on device-resident balls (structure k: RSA cluster of
    n_k balls, n_k log-uniform in [200, 20000], seed 1000000 + k)."""
    import ctypes as C
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    from voronota_ltr import _hip, _native, workload

    sizes = workload.batch_sizes(args.batch_structs, seed=4096)
    with ThreadPoolExecutor(max_workers=min(os.cpu_count() or 1, 16)) as ex:
        structs = list(ex.map(lambda k: workload.rsa_sphere(int(sizes[k]), seed=1000000 + k)[0], range(len(sizes))))
    off = _native.offsets_of(structs)
    allb = np.ascontiguousarray(np.concatenate(structs))
    del structs
    total = int(off[-1])
    dbuf = _hip.DeviceBuffer(allb.nbytes)
    dbuf.upload(allb)
    offp = off.ctypes.data_as(C.POINTER(C.c_uint64))

    def batched_once():
        _native.check(lib.vltr_compute_tessellation_batched_device(ctx, dbuf.ptr, offp, len(sizes), args.probe, 0),
                      "vltr_compute_tessellation_batched_device")

    tune = (None if (args.no_autotune or args.clip_shape is not None)
            else autotune(lib, ctx, dbuf.ptr, off, args.probe, repeats=2))
    batched_once()
    _hip.synchronize()
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        batched_once()
    _hip.synchronize()
    t_batched = (time.perf_counter() - t0) / reps
    counts = _native.last_counts()
    grid = grid_build(_native.last_timings()["grid"], counts)

    def graph_vs_direct(k_structs, reps):
        """The first k_structs structures as one batch: direct runs against
        replays of the captured HIP graph (same device buffer and layout)."""
        sub = np.ascontiguousarray(off[:k_structs + 1])
        subp = sub.ctypes.data_as(C.POINTER(C.c_uint64))
        times = {}
        for graphs in (False, True):
            _native.set_graphs(graphs)
            for _ in range(2):
                _native.check(lib.vltr_compute_tessellation_batched_device(ctx, dbuf.ptr, subp, k_structs, args.probe,
                                                                           0), "batched device")
            _hip.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                _native.check(lib.vltr_compute_tessellation_batched_device(ctx, dbuf.ptr, subp, k_structs, args.probe,
                                                                           0), "batched device")
            _hip.synchronize()
            times[graphs] = (time.perf_counter() - t0) / reps
        _native.set_graphs(False)
        return {"structures": k_structs, "balls": int(sub[-1]), "direct_ms": round(times[False] * 1e3, 4),
                "graph_replay_ms": round(times[True] * 1e3, 4),
                "direct_over_graph": round(times[False] / times[True], 3)}

    graphs = {"all": graph_vs_direct(len(sizes), reps), "first_16": graph_vs_direct(min(16, len(sizes)), 50)}

    base = dbuf.ptr.value
    for k in range(min(8, len(sizes))):  # warm the per-call path
        lib.vltr_compute_tessellation_device(ctx, C.c_void_p(base + int(off[k]) * 32), int(sizes[k]), args.probe,
                                             None, 0, 0)
    _hip.synchronize()
    t0 = time.perf_counter()
    for k in range(len(sizes)):
        _native.check(lib.vltr_compute_tessellation_device(ctx, C.c_void_p(base + int(off[k]) * 32), int(sizes[k]),
                                                           args.probe, None, 0, 0), "vltr_compute_tessellation_device")
    _hip.synchronize()
    t_plain = time.perf_counter() - t0
    dbuf.free()
    return {"structures": len(sizes), "balls": total, "pairs": counts["pairs"], "contacts": counts["contacts"],
            "batched_s": round(t_batched, 4), "batched_balls_per_s": round(total / t_batched, 1),
            "plain_s": round(t_plain, 4), "plain_balls_per_s": round(total / t_plain, 1),
            "batched_over_plain": round(t_plain / t_batched, 3), "grid_build": grid, "autotune": tune,
            "hip_graph": graphs,
            "note": "full compute_tessellation of every structure, balls resident in HBM; plain = one "
                    "vltr_compute_tessellation_device call per structure, batched = one "
                    "vltr_compute_tessellation_batched_device call"}


def adjoint_leg(args, lib, ctx):
    """Forward vs forward + adjoint on the differentiation workload: N balls by
    RSA in a sphere, type = radius index, seeded symmetric W in [-1, 1]^(4x4);
    E = sum_contacts W[t_a][t_b] area + sum SAS + sum volume, gradient with
    respect to all centres; inputs and gradient resident in HBM."""
    import ctypes as C

    import numpy as np

    from voronota_ltr import _hip, _native, workload

    n = args.adj_balls
    balls, types = workload.rsa_sphere(n, seed=n)
    W = np.ascontiguousarray(workload.type_matrix(n).reshape(16))
    db = _hip.DeviceBuffer(balls.nbytes)
    db.upload(np.ascontiguousarray(balls))
    dt = _hip.DeviceBuffer(4 * n)
    dt.upload(np.ascontiguousarray(types.astype(np.int32)))
    dg = _hip.DeviceBuffer(24 * n)
    e = C.c_double()
    wp = W.ctypes.data_as(C.POINTER(C.c_double))

    def fwd():
        _native.check(lib.vltr_compute_tessellation_device(ctx, db.ptr, n, args.probe, None, 0, 0), "forward")

    def fwd_adj():
        _native.check(lib.vltr_energy_gradient_device(ctx, db.ptr, n, args.probe, None, dt.ptr, wp, 1.0, 1.0, dg.ptr,
                                                      C.byref(e)), "forward + adjoint")

    def timeit(fn, reps=10):
        for _ in range(2):
            fn()
        _hip.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        _hip.synchronize()
        return (time.perf_counter() - t0) / reps

    t_f = timeit(fwd)
    t_fa = timeit(fwd_adj)
    for b in (db, dt, dg):
        b.free()
    return {"balls": n, "fwd_ms": round(t_f * 1e3, 3), "fwd_adjoint_ms": round(t_fa * 1e3, 3),
            "fwd_adjoint_over_fwd": round(t_fa / t_f, 3), "energy": e.value,
            "note": "E = sum_contacts W[t_a][t_b] area + sum SAS + sum volume; gradient of all 3N centre "
                    "coordinates; device-resident calls"}


def dynamics_leg(args):
    """Periodic dynamics through the updateable tessellation: N balls placed
    by RSA in a periodic cube at 0.056 balls/A^3; each step moves 500 seeded
    balls by N(0, 0.05^2) A per axis, wraps them into the box and updates
    (update_with_changed). The oracle's updateable (C++, host cores) runs the
    first steps of the same sequence for comparison."""
    import numpy as np

    from voronota_ltr import _native, workload

    n = args.dyn_balls
    L = (n / 0.056) ** (1.0 / 3.0)
    balls, _ = workload.rsa_box(n, L, seed=50000)
    box = [L, 0, 0, 0, L, 0, 0, 0, L]
    rng = np.random.default_rng(50000)
    moves = []
    cur = balls.copy()
    for _ in range(args.dyn_steps):
        ch = np.sort(rng.choice(n, 500, replace=False))
        cur[ch, :3] = np.mod(cur[ch, :3] + rng.normal(scale=0.05, size=(500, 3)), L)
        moves.append((cur.copy(), ch))
    u = _native.Updateable()
    assert u.init(balls, args.probe, box)
    t0 = time.perf_counter()
    reinit = 0
    affected = 0
    for b, ch in moves:
        assert u.update_with_changed(b, ch)
        reinit += int(u.last_update_was_full_reinit())
        affected += len(u.affected_ids())
    t_gpu = (time.perf_counter() - t0) / len(moves)
    ph = u.last_timings()
    u.close()
    # the same steps with the contact energy's adjoint in every pass
    # (energy of the differentiation workload: W seeded, type = radius index)
    types = np.searchsorted(np.array([1.52, 1.55, 1.70, 1.80]), balls[:, 3]).astype(np.int32)
    u = _native.Updateable()
    u.set_energy(types, workload.type_matrix(20000))
    assert u.init(balls, args.probe, box)
    u.energy_device()
    t0 = time.perf_counter()
    for b, ch in moves:
        assert u.update_with_changed(b, ch)
        e_last, _ = u.energy_device()
    t_adj = (time.perf_counter() - t0) / len(moves)
    u.close()
    out = {"balls": n, "box_edge_A": round(L, 3), "steps": len(moves), "moved_per_step": 500,
           "ms_per_step": round(t_gpu * 1e3, 3), "steps_per_s": round(1.0 / t_gpu, 1),
           "ms_per_step_with_adjoint": round(t_adj * 1e3, 3), "last_energy": e_last,
           "mean_affected": round(affected / len(moves), 1), "full_reinits": reinit,
           "last_step_phases_ms": {k: round(v, 4) for k, v in ph.items()},
           "note": "host API (the moved balls are sent each step; summary not downloaded); with_adjoint: the "
                   "energy of the differentiation workload and its gradient (left on the device) in every step's "
                   "pass"}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle

        threads = cpu_threads()
        o = pyoracle.Updateable(threads=threads)
        o.init(balls, args.probe, box)
        o.update_with_changed(*moves[0])  # warm-up
        k = min(5, len(moves) - 1)
        t0 = time.perf_counter()
        for b, ch in moves[1:1 + k]:
            o.update_with_changed(b, ch)
        out["cpu_oracle_ms_per_step"] = round((time.perf_counter() - t0) / max(k, 1) * 1e3, 2)
        out["cpu_oracle_steps_timed"] = k
        out["cpu_cores"] = threads
    return out


if __name__ == "__main__":
    main()
