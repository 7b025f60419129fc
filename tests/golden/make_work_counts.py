"""Writes tests/golden/data/work_counts.json: the algorithmic work of the
benchmark workloads, counted by the oracle's counting scalar (oracle/,
the C++ restatement of the reference templated on a scalar that counts its
operations): FLOPs with add/sub/mul/div/sqrt = 1 and fma = 2, and
transcendental ops (acos, sin, cos) separately; per mode (0 = full
compute_tessellation, 1 = contacts only) and split into the collision search
and the rest: the contact construction (classification, clip, restriction,
area, solid angles and pyramid volumes: the work of the clip kernels) plus
the cells' few FLOPs per ball.

bench.py reads this file to price the clip kernel's roofline; the test
tests/test_oracle_pins.py::test_work_counts_fixture recounts it.
Run: python tests/golden/make_work_counts.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "data", "work_counts.json")

WORKLOADS = {
    # name: (balls, seed, probe); RSA in a sphere at 0.056 balls/A^3
    "rsa_sphere_100000": (100000, 100000, 1.4),
    "rsa_sphere_20000": (20000, 20000, 1.4),
}


def count(name):
    sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "voronota-ltr_amd")]
    import pyoracle
    from voronota_ltr import workload

    n, seed, probe = WORKLOADS[name]
    balls, _ = workload.rsa_sphere(n, seed=seed)
    off, idx, _, _, pairs = pyoracle.collisions(balls, probe)
    out = {"balls": n, "seed": seed, "probe": probe, "candidate_pairs": int(len(pairs))}
    sfl, stops = pyoracle.count_flops(balls, probe, mode=2)  # the collision search alone
    for mode in (0, 1):
        fl, tops = pyoracle.count_flops(balls, probe, mode=mode)
        out[f"mode{mode}"] = {"flops_total": int(fl), "tops_total": int(tops), "flops_search": int(sfl),
                              "tops_search": int(stops), "flops_contacts": int(fl - sfl),
                              "tops_contacts": int(tops - stops)}
    return out


def main():
    data = {"cost_table": "add/sub/mul/div/sqrt = 1 FLOP, fma = 2 FLOP; acos/sin/cos counted as T-ops, not FLOPs; "
                          "abs/neg/min/max/compare = 0",
            "workloads": {name: count(name) for name in WORKLOADS}}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps(data, indent=1))


if __name__ == "__main__":
    main()
