"""Pseudo-solvent spheres (compute_solvent_spheres) on the GPU against the
oracle: the same spheres in the same order with bit-identical positions and
parents (the samples are the same double arithmetic on the same inputs) and
weights to the parity bar (1e-9 relative, floor 1e-12 of the sample's
sphere volume). Cases: the reference's unit tests (solvent_spheres.rs:352-481),
a protein at every subdivision depth, a volume probe, a dense RSA cluster,
periodic boxes, the depth clamp and the validation errors with the reference's
messages.
"""
import math

import numpy as np
import pytest

from conftest import corners_box, load_xyzr

pytestmark = pytest.mark.gpu


def _same(gpu, ref, what):
    assert len(gpu) == len(ref), f"{what}: {len(gpu)} spheres on the GPU, {len(ref)} in the oracle"
    if len(ref) == 0:
        return
    assert np.array_equal(gpu["x"], ref[:, 0]) and np.array_equal(gpu["y"], ref[:, 1]) \
        and np.array_equal(gpu["z"], ref[:, 2]), f"{what}: positions differ"
    assert np.array_equal(gpu["parent_index"].astype(np.int64), ref[:, 5].astype(np.int64)), f"{what}: parents"
    assert np.array_equal(gpu["radius"], ref[:, 3]), f"{what}: radii"
    vs = 4.0 / 3.0 * math.pi * np.maximum(ref[:, 3], 0.01) ** 3
    err = np.abs(gpu["weight"] - ref[:, 4])
    bad = err > 1e-9 * np.abs(ref[:, 4]) + 1e-12 * vs
    assert not bad.any(), f"{what}: weight {np.nonzero(bad)[0][:5]} gpu {gpu['weight'][bad][:3]} ref {ref[bad, 4][:3]}"


REFERENCE_CASES = [
    ("single", [[0.0, 0.0, 0.0, 1.5]]),
    ("touching", [[0, 0, 0, 1.5], [2.5, 0, 0, 1.5]]),
    ("separated", [[0, 0, 0, 1.5], [5, 0, 0, 1.5]]),
    ("contained", [[0, 0, 0, 3.0], [2.5, 0, 0, 1.0]]),
]


@pytest.mark.parametrize("depth", [0, 2])
def test_reference_cases(oracle, depth):
    from voronota_ltr import _native

    for name, balls in REFERENCE_CASES:
        print(f"\n[gpu] solvent {name} depth {depth}", flush=True)
        g = _native.solvent_spheres(balls, 1.4, 0.0, depth)
        _same(g, oracle.solvent_spheres(balls, 1.4, 0.0, depth), name)
    # solvent_spheres.rs:395-425: 24 spheres, 12 per parent, total weight ~249.1
    g = _native.solvent_spheres(REFERENCE_CASES[2][1], 1.4, 0.0, 0)
    assert len(g) == 24 and np.count_nonzero(g["parent_index"] == 0) == 12
    assert abs(g["weight"].sum() - 249.1) < 5.0
    assert len(_native.solvent_spheres(np.zeros((0, 4)), 1.4)) == 0


@pytest.mark.parametrize("depth", [0, 1, 2, 3, 4])
def test_protein_depths(oracle, depth):
    from voronota_ltr import _native

    balls = load_xyzr("1ctf_pdb.xyzr")
    print(f"\n[gpu] solvent 1ctf depth {depth}", flush=True)
    g = _native.solvent_spheres(balls, 1.4, 0.0, depth)
    _same(g, oracle.solvent_spheres(balls, 1.4, 0.0, depth), f"1ctf depth {depth}")
    assert len(g) > 0


def test_volume_probe_and_dense_cluster(oracle):
    from voronota_ltr import _native, workload

    balls = load_xyzr("balls_2zsk.xyzr")
    print("\n[gpu] solvent 2zsk volume probe 0.5", flush=True)
    _same(_native.solvent_spheres(balls, 1.4, 0.5, 2), oracle.solvent_spheres(balls, 1.4, 0.5, 2), "2zsk vp 0.5")
    dense, _ = workload.rsa_sphere(3000, seed=91)
    print("[gpu] solvent rsa 3000", flush=True)
    _same(_native.solvent_spheres(dense, 1.2, 0.0, 1), oracle.solvent_spheres(dense, 1.2, 0.0, 1), "rsa 3000")


def test_depth_clamp():
    """SubdivisionDepth::from clamps to 4 (subdivided_icosahedron.rs:156-166)."""
    from voronota_ltr import _native

    balls = load_xyzr("1ctf_pdb.xyzr")[:200]
    a = _native.solvent_spheres(balls, 1.4, 0.0, 4)
    b = _native.solvent_spheres(balls, 1.4, 0.0, 9)
    assert np.array_equal(a, b)


def test_periodic(oracle):
    from voronota_ltr import _native, workload

    L = 30.0
    balls, _ = workload.rsa_box(1200, L, seed=17)
    box = corners_box((0, 0, 0), (L, L, L))
    print("\n[gpu] solvent periodic box", flush=True)
    g = _native.solvent_spheres(balls, 1.4, 0.0, 1, box)
    _same(g, oracle.solvent_spheres(balls, 1.4, 0.0, 1, box), "periodic rsa")
    # a box far larger than the cluster: the same spheres as without a box
    small = load_xyzr("1ctf_pdb.xyzr")
    lo = small[:, :3].min(0) - 50.0
    hi = small[:, :3].max(0) + 50.0
    gp = _native.solvent_spheres(small, 1.4, 0.0, 0, corners_box(lo, hi))
    gn = _native.solvent_spheres(small, 1.4, 0.0, 0)
    assert len(gp) == len(gn)
    assert np.array_equal(gp["parent_index"], gn["parent_index"])
    assert np.array_equal(gp["x"], gn["x"])
    assert np.allclose(gp["weight"], gn["weight"], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("args,msg", [
    (([[0, 0, 0, 1.5]], -1.0), "invalid probe radius: -1 (must be non-negative and finite)"),
    (([[0, 0, 0, 1.5]], float("nan")), "invalid probe radius: NaN (must be non-negative and finite)"),
    (([[0, 0, 0, 1.5]], 1.4, -0.25), "invalid volume probe: -0.25 (must be non-negative and finite)"),
    (([[0, 0, 0, -1.0]], 1.4), "invalid ball at index 0: radius must be positive and finite"),
    (([[0, 0, 0, 1.0], [float("nan"), 0, 0, 1.5]], 1.4), "invalid ball at index 1: coordinates must be finite"),
])
def test_validation_errors(args, msg):
    """solvent_spheres.rs:448-481 and its Display impl (solvent_spheres.rs:51-65)."""
    from voronota_ltr import _native

    with pytest.raises(ValueError) as e:
        _native.solvent_spheres(*args)
    assert str(e.value) == msg
