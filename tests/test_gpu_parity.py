"""HIP path vs the oracle on the reference's fixtures and the seeded workloads.

Each case runs compute_tessellation (and compute_contacts_only) through the C
ABI and checks the contact pair sequence and flags exactly and the values to
1e-9 relative (see parity.py for the floors).
"""
import numpy as np
import pytest

from conftest import CENTRAL8, RING17, corners_box, load_xyzr
from parity import compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vltr():
    import voronota_ltr

    return voronota_ltr


def _check_neighbors(vltr, oracle, balls, probe, pbox):
    from voronota_ltr import _native

    counts, ids = _native.neighbor_lists(len(balls))
    off, idx, _, excluded, _ = oracle.collisions(balls, probe, pbox=pbox)
    assert np.array_equal(counts.astype(np.int64), np.diff(off.astype(np.int64))), "neighbor list lengths"
    assert np.array_equal(ids.astype(np.int64), idx.astype(np.int64)), "neighbor list contents/order"


def _run(vltr, oracle, balls, probe, pbox=None, groups=None, name=""):
    print(f"\n[gpu] {name}: n={len(balls)} probe={probe} periodic={pbox is not None} groups={groups is not None}",
          flush=True)
    gpu = vltr.tessellate(balls, probe, periodic_box=pbox, groups=groups)
    _check_neighbors(vltr, oracle, balls, probe, pbox)
    ref = oracle.compute_tessellation(balls, probe, pbox=pbox, groups=groups)
    stats = compare(gpu, ref, balls, probe)
    print(f"[gpu] {name}: {stats}", flush=True)
    gco = vltr.tessellate(balls, probe, periodic_box=pbox, groups=groups, contacts_only=True)
    rco = oracle.compute_contacts_only(balls, probe, pbox=pbox, groups=groups)
    compare(gco, rco, balls, probe, check_cells=False, ordered=pbox is None)
    return gpu, ref


def test_two_balls(vltr, oracle):
    b = np.array([[0.0, 0.0, 0.0, 1.5], [3.0, 0.0, 0.0, 1.5]])
    _run(vltr, oracle, b, 1.4, name="two balls")


def test_ring17(vltr, oracle):
    gpu, _ = _run(vltr, oracle, RING17, 1.0, name="ring17")
    assert len(gpu.contacts) == 44 and len(gpu.cells) == 17


def test_central8(vltr, oracle):
    _run(vltr, oracle, CENTRAL8, 1.4, name="central8")


def test_cs_1x1(vltr, oracle):
    b = load_xyzr("balls_cs_1x1.xyzr")
    gpu, _ = _run(vltr, oracle, b, 2.0, name="cs_1x1")
    assert len(gpu.contacts) == 153
    assert abs(gpu.total_contact_area - 3992.55) < 0.01
    assert abs(gpu.total_sas_area - 21979.64) < 0.01
    assert abs(gpu.total_volume - 46419.87) < 0.01


def test_1ctf(vltr, oracle):
    b = load_xyzr("1ctf_pdb.xyzr")
    gpu, _ = _run(vltr, oracle, b, 1.4, name="1ctf")
    assert len(gpu.contacts) == 3078 and len(gpu.cells) == 492


def test_2zsk(vltr, oracle):
    b = load_xyzr("balls_2zsk.xyzr")
    gpu, _ = _run(vltr, oracle, b, 1.4, name="2zsk")
    assert len(gpu.contacts) == 23855 and len(gpu.cells) == 3545


@pytest.mark.parametrize("probe", [6.0, 7.0])
def test_long_neighbor_lists(vltr, oracle, probe):
    """2zsk with large probes: up to 1122 neighbors per ball, beyond one
    wavefront's candidate list (block-per-ball neighbor lists) and beyond the
    group clip kernels' neighbor slots."""
    _run(vltr, oracle, load_xyzr("balls_2zsk.xyzr"), probe, name=f"2zsk probe {probe}")


def test_3dlb(vltr, oracle):
    b = load_xyzr("balls_3dlb.xyzr")
    _run(vltr, oracle, b, 1.4, name="3dlb")


def test_groups(vltr, oracle):
    b = load_xyzr("balls_2zsk.xyzr")
    groups = (np.arange(len(b)) // 10).astype(np.int32)
    _run(vltr, oracle, b, 1.4, groups=groups, name="2zsk groups/10")
    tb = np.array([[0, 0, 0, 1.5], [1, 0, 0, 1.5], [4, 0, 0, 2.0], [5, 0, 0, 2.0]], float)
    gco = vltr.tessellate(tb, 1.4, groups=[0, 0, 1, 1], contacts_only=True)
    assert len(gco.contacts) == 1 and abs(gco.contacts["area"][0] - 23.43431770037136) < 1e-6


def test_periodic(vltr, oracle):
    b = load_xyzr("balls_cs_1x1.xyzr")
    gpu, _ = _run(vltr, oracle, b, 2.0, pbox=corners_box((0, 0, 0), (200, 250, 300)), name="cs_1x1 periodic")
    assert len(gpu.contacts) == 189
    _run(vltr, oracle, RING17, 1.0, pbox=corners_box((-1.6, -1.6, -0.6), (1.6, 1.6, 3.1)), name="ring17 periodic")


def test_periodic_balls_outside_the_box(vltr, oracle):
    """Balls need not lie in the box: a 3,000-ball cluster of a periodic box
    (RSA at 0.056 balls/A^3) moved 1.5 box lengths away from the box and a
    box smaller than the cluster's span; the images that can touch a ball are
    those near the balls, wherever the box is."""
    from voronota_ltr import workload

    n = 3000
    L = (n / 0.056) ** (1.0 / 3.0)
    b, _ = workload.rsa_box(n, L, seed=3000)
    shifted = b.copy()
    shifted[:, :3] += 1.5 * L
    _run(vltr, oracle, shifted, 1.4, pbox=corners_box((0, 0, 0), (L, L, L)), name="rsa box shifted")
    _run(vltr, oracle, b, 1.4, pbox=corners_box((0, 0, 0), (0.9 * L, 0.95 * L, L)), name="rsa box, smaller box")


def _lattice(n_side, spacing, r):
    g = np.arange(n_side) * spacing
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return np.column_stack([x.ravel(), y.ravel(), z.ravel(), np.full(x.size, r)])


@pytest.mark.parametrize("case", ["lattice", "lattice_jitter", "duplicates", "nested", "chain", "probe0", "plane"])
def test_degenerate_inputs(vltr, oracle, case):
    """Ties and degeneracies: a cubic lattice (equal neighbor keys, radical
    planes through shared points), the same with 1e-9 jitter, duplicate
    spheres (hidden by index), spheres inside others, a straight chain, zero
    probe, and a flat sheet."""
    rng = np.random.default_rng(7)
    probe = 1.4
    if case == "lattice":
        b = _lattice(6, 2.0, 1.5)
    elif case == "lattice_jitter":
        b = _lattice(6, 2.0, 1.5)
        b[:, :3] += rng.normal(scale=1e-9, size=(len(b), 3))
    elif case == "duplicates":
        b = _lattice(4, 3.0, 1.6)
        b = np.concatenate([b, b[::3], b[::5]])
    elif case == "nested":
        b = _lattice(4, 3.2, 1.7)
        inner = b[::2].copy()
        inner[:, 3] = 0.5
        inner[:, :3] += 0.3
        b = np.concatenate([b, inner])
    elif case == "chain":
        b = np.column_stack([np.arange(40) * 1.1, np.zeros(40), np.zeros(40), np.full(40, 1.2)])
    elif case == "probe0":
        b = _lattice(5, 2.5, 1.6)
        b[:, :3] += rng.normal(scale=0.3, size=(len(b), 3))
        probe = 0.0
    else:
        g = _lattice(8, 2.4, 1.5)
        b = g[g[:, 2] == 0.0]
    _run(vltr, oracle, b, probe, name=case)


def test_rsa_2500(vltr, oracle):
    from voronota_ltr import workload

    b, _ = workload.rsa_sphere(2500, seed=2500)
    _run(vltr, oracle, b, 1.4, name="rsa 2500")


def test_rsa_20000(vltr, oracle):
    from voronota_ltr import workload

    b, _ = workload.rsa_sphere(20000, seed=20000)
    _run(vltr, oracle, b, 1.4, name="rsa 20000")


def test_empty_and_single(vltr):
    r = vltr.tessellate(np.zeros((0, 4)), 1.4)
    assert r.num_balls == 0 and len(r.contacts) == 0 and len(r.cells) == 0
    r = vltr.tessellate(np.array([[0.0, 0.0, 0.0, 1.0]]), 0.5)
    assert len(r.contacts) == 0 and len(r.cells) == 1
    assert abs(r.cells["sas_area"][0] - 4 * np.pi * 1.5 ** 2) < 1e-12


def test_neighbor_slots_after_a_dense_call(vltr):
    """A dense call (2zsk at probe 7, up to 1,122 neighbors per ball) grows
    the context's neighbor slots per ball; a later large call halves them
    until its neighbor and incoming lists fit the list budget (2 GiB), instead
    of sizing 200,000 balls' lists for the dense call's 1,536 or 3,072
    slots.""" 
    from voronota_ltr import _native, workload

    print("\n[gpu] 2zsk probe 7, then 200k balls", flush=True)
    vltr.tessellate(load_xyzr("balls_2zsk.xyzr"), 7.0)
    grown = _native.last_counts()["nb_cap"]
    assert grown >= 1536
    b, _ = workload.rsa_sphere(200000, seed=200000)
    r = vltr.tessellate(b, 1.4)
    cap = _native.last_counts()["nb_cap"]
    assert cap < grown and 200000 * cap * 8 <= 2 * 2 ** 30
    assert len(r.cells) == 200000 and len(r.contacts) > 1400000
