"""Adjoint of E = sum_contacts W[t_a][t_b] area + alpha sum SAS + beta sum volume
with respect to the ball centres, on the GPU.

* against the oracle's forward-mode dual numbers (the exact derivative of the
  reference's computation with its contour combinatorics fixed), all three
  coordinates of 80 balls of a 300-ball cluster;
* against central finite differences (h = 1e-6 A) of the GPU's own forward
  results, differenced term by term (each contact and cell separately, so the
  large total does not cancel), on the 20,000-ball differentiation workload,
  64 random balls x 3 axes, skipping perturbations that change the contact
  set; a subset also against differences of the CPU path (the oracle) and
  against the oracle's exact derivative. 1e-6 relative to the difference;
  where the difference quotient is too noisy for that (round-off of the
  terms over 2h, truncation near nearly degenerate vertices), a floor of
  twice its change between h = 1e-6 and 2e-6, with those coordinates
  counted (at most 10%) and checked against the oracle's dual number;
* total-SAS-only and total-volume-only weights;
* the torch.autograd.Function wrapper.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cluster(n, seed):
    from voronota_ltr import workload

    balls, types = workload.rsa_sphere(n, seed=seed)
    return balls, types, workload.type_matrix(seed)


def _terms(res, types, W, alpha=1.0, beta=1.0):
    c = res.contacts
    ids = zip(c["id_a"].astype(np.int64), c["id_b"].astype(np.int64))
    ct = {(a, b): W[types[a], types[b]] * ar for (a, b), ar in zip(ids, c["area"])}
    cl = {int(i): alpha * s + beta * v for i, s, v in zip(res.cells["index"], res.cells["sas_area"],
                                                          res.cells["volume"])}
    return ct, cl


def _fd(balls, types, W, ball, axis, h, alpha=1.0, beta=1.0):
    """Central difference of E from GPU forward runs, term by term; None when
    the contact set changes between the two runs."""
    from voronota_ltr import _native

    bp, bm = balls.copy(), balls.copy()
    bp[ball, axis] += h
    bm[ball, axis] -= h
    cp, lp = _terms(_native.run(bp, 1.4), types, W, alpha, beta)
    cm, lm = _terms(_native.run(bm, 1.4), types, W, alpha, beta)
    if cp.keys() != cm.keys() or lp.keys() != lm.keys():
        return None
    d = sum(cp[k] - cm[k] for k in cp) + sum(lp[k] - lm[k] for k in lp)
    return d / (2 * h)


def test_gradient_matches_oracle_dual(oracle):
    from voronota_ltr import _native

    balls, types, W = _cluster(300, 31)
    print("\n[gpu] adjoint vs dual: 300 balls", flush=True)
    e, g = _native.energy_gradient(balls, 1.4, types, W)
    e_ref = oracle.energy(balls, 1.4, types, W)
    assert e == pytest.approx(e_ref, rel=1e-10)
    worst = 0.0
    for ball in np.random.default_rng(3).choice(len(balls), 80, replace=False):
        for axis in range(3):
            _, de = oracle.energy_dual(balls, 1.4, types, W, int(ball), axis)
            err = abs(g[ball, axis] - de) / max(1.0, abs(de))
            worst = max(worst, err)
            assert err < 1e-7, (ball, axis, g[ball, axis], de)
    print(f"[gpu] adjoint vs dual: worst {worst:.2e}", flush=True)


def _self_image_box():
    """Four balls at seeded random positions in a 5.3 x 5.9 x 6.1 box: every
    ball touches its own images (4 self-contacts), generic geometry."""
    rng = np.random.default_rng(7)
    L = np.array([5.3, 5.9, 6.1])
    balls = np.c_[rng.uniform(0, 1, (4, 3)) * L, rng.choice([1.52, 1.55, 1.7, 1.8], 4)]
    return balls, [L[0], 0, 0, 0, L[1], 0, 0, 0, L[2]]


@pytest.mark.parametrize("case", ["rsa_box", "self_images"])
def test_periodic_gradient_matches_oracle_dual(oracle, case):
    """Periodic box: contacts through images, the deduplicated contact list
    (one W * area term per canonical pair among the contacts through images,
    the first in pair order), cells of canonical balls only; energy and
    gradient against the oracle's dual numbers."""
    from conftest import corners_box
    from voronota_ltr import _native, workload

    if case == "rsa_box":
        L = 30.0
        balls, _ = workload.rsa_box(1500, L, seed=8)
        box = corners_box((0, 0, 0), (L, L, L))
        picks = np.random.default_rng(9).choice(len(balls), 40, replace=False)
    else:
        balls, box = _self_image_box()
        picks = np.arange(len(balls))
    types = np.arange(len(balls), dtype=np.int32) % 4
    W = workload.type_matrix(44)
    print(f"\n[gpu] periodic adjoint {case}", flush=True)
    e, g = _native.energy_gradient(balls, 1.4, types, W, pbox=box)
    assert e == pytest.approx(oracle.energy(balls, 1.4, types, W, pbox=box), rel=1e-10)
    worst = 0.0
    for ball in picks:
        for axis in range(3):
            _, de = oracle.energy_dual(balls, 1.4, types, W, int(ball), axis, pbox=box)
            err = abs(g[ball, axis] - de) / max(1.0, abs(de))
            worst = max(worst, err)
            assert err < 1e-7, (int(ball), axis, g[ball, axis], de)
    print(f"[gpu] periodic adjoint {case}: worst {worst:.2e}", flush=True)


def test_periodic_degenerate_ring17():
    """Ring-17 in the reference's small periodic box (self-contacts through
    several images, exact mirror symmetry, coincident vertices): the energy
    equals the oracle's and the gradient stays finite; its y and x components
    of the top ball vanish by the mirror symmetry. (The oracle's dual follows
    one branch of the reference's arithmetic at these non-differentiable
    points, so it is not a reference here.)"""
    from conftest import RING17, corners_box
    from voronota_ltr import _native, workload

    import pyoracle

    box = corners_box((-1.6, -1.6, -0.6), (1.6, 1.6, 3.1))
    types = np.arange(17, dtype=np.int32) % 4
    W = workload.type_matrix(44)
    e, g = _native.energy_gradient(RING17, 1.0, types, np.zeros((4, 4)), pbox=box)
    assert e == pytest.approx(pyoracle.energy(RING17, 1.0, types, np.zeros((4, 4)), pbox=box), rel=1e-10)
    assert np.all(np.isfinite(g))
    assert abs(g[0, 0]) < 1e-9 and abs(g[0, 1]) < 1e-9
    e, g = _native.energy_gradient(RING17, 1.0, types, W, pbox=box)
    assert e == pytest.approx(pyoracle.energy(RING17, 1.0, types, W, pbox=box), rel=1e-10)
    assert np.all(np.isfinite(g))


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (0.0, 1.0)])
def test_sas_and_volume_gradients(alpha, beta):
    from voronota_ltr import _native

    balls, types, _ = _cluster(400, 32)
    Wz = np.zeros((4, 4))
    e, g = _native.energy_gradient(balls, 1.4, types, Wz, sas_weight=alpha, volume_weight=beta)
    r = _native.run(balls, 1.4)
    assert e == pytest.approx(alpha * r.cells["sas_area"].sum() + beta * r.cells["volume"].sum(), rel=1e-12)
    rng = np.random.default_rng(5)
    checked = 0
    for ball in rng.choice(len(balls), 12, replace=False):
        for axis in range(3):
            fd = _fd(balls, types, Wz, int(ball), axis, 1e-6, alpha, beta)
            if fd is None:
                continue
            checked += 1
            assert abs(g[ball, axis] - fd) <= 1e-6 * abs(fd) + 1e-6, (ball, axis, g[ball, axis], fd)
    assert checked >= 30


def test_differentiation_workload_fd(oracle):
    """20,000 balls (the differentiation workload), 64 random balls x 3 axes:
    |adjoint - fd| <= 1e-6 |fd| + floor, fd the central difference at
    h = 1e-6 (term by term). The floor is 0 unless the relative bar alone
    fails; it is then twice the difference between the estimates at h = 1e-6
    and h = 2e-6, the difference quotient's own noise (round-off of the terms
    over 2h = 2e-6, and truncation near a nearly degenerate contour vertex),
    the coordinates that needed it are counted and reported, and each of them
    is also checked against the oracle's forward-mode dual number (the exact
    derivative of the reference computation) to 1e-7 relative."""
    from voronota_ltr import _native

    balls, types, W = _cluster(20000, 20000)
    print("\n[gpu] 20k adjoint", flush=True)
    e, g = _native.energy_gradient(balls, 1.4, types, W)
    rng = np.random.default_rng(64)
    picks = rng.choice(len(balls), 64, replace=False)
    checked = skipped = floored = 0
    worst = worst_floor = 0.0
    for k, ball in enumerate(picks):
        for axis in range(3):
            fd = _fd(balls, types, W, int(ball), axis, 1e-6)
            if fd is None:
                skipped += 1
                continue
            checked += 1
            dev = abs(g[ball, axis] - fd)
            if dev > 1e-6 * abs(fd):
                fd2 = _fd(balls, types, W, int(ball), axis, 2e-6)
                assert fd2 is not None, (int(ball), axis)
                floor = 2.0 * abs(fd - fd2)
                floored += 1
                worst_floor = max(worst_floor, floor)
                assert dev <= 1e-6 * abs(fd) + floor, (int(ball), axis, g[ball, axis], fd, fd2)
                # where the difference quotient is this noisy, the exact
                # derivative of the reference computation decides
                _, de = oracle.energy_dual(balls, 1.4, types, W, int(ball), axis)
                assert abs(g[ball, axis] - de) <= 1e-7 * abs(de) + 1e-9, (int(ball), axis, g[ball, axis], de)
            worst = max(worst, dev / max(abs(fd), 1e-300))
        if k % 16 == 15:
            print(f"[gpu] 20k adjoint: {k + 1} balls, worst relative deviation {worst:.2e}", flush=True)
    print(f"[gpu] 20k adjoint: checked {checked}, skipped (contact set changes within h) {skipped}, within 1e-6 "
          f"relative {checked - floored}, floor-limited {floored} (largest floor {worst_floor:.2e}), worst relative "
          f"deviation {worst:.2e}", flush=True)
    assert checked >= 150
    assert floored <= 0.1 * checked
    # the oracle's exact derivative (forward-mode dual numbers) at full size
    for ball in picks[:4]:
        _, de = oracle.energy_dual(balls, 1.4, types, W, int(ball), 1)
        assert abs(g[ball, 1] - de) <= 1e-8 * max(abs(de), 1.0), (int(ball), g[ball, 1], de)
    # reference gradient by differencing the CPU path (oracle), term by term
    for ball in picks[:4]:
        bp, bm = balls.copy(), balls.copy()
        bp[ball, 0] += 1e-6
        bm[ball, 0] -= 1e-6
        tp = oracle.energy_terms(bp, 1.4, types, W)
        tm = oracle.energy_terms(bm, 1.4, types, W)
        kp = {tuple(int(x) for x in i): a for i, a in zip(tp.contact_ids, tp.contact_area)}
        km = {tuple(int(x) for x in i): a for i, a in zip(tm.contact_ids, tm.contact_area)}
        if kp.keys() != km.keys():
            continue
        d = sum(kp[k] - km[k] for k in kp)
        d += float(np.sum(tp.cell_sas - tm.cell_sas) + np.sum(tp.cell_volume - tm.cell_volume))
        fd = d / 2e-6
        assert abs(g[ball, 0] - fd) <= 1e-5 * max(abs(fd), 1.0), (int(ball), g[ball, 0], fd)


TORCH_CHECK = r"""
import sys
import numpy as np
import torch
sys.path.insert(0, sys.argv[1])
if not torch.cuda.is_available():
    print("NO_GPU"); sys.exit(0)
from voronota_ltr import _native, workload
from voronota_ltr.autograd import tessellation_energy
balls, types = workload.rsa_sphere(500, seed=33)
W = workload.type_matrix(33)
dev = torch.device("cuda", 0)
centers = torch.tensor(balls[:, :3], dtype=torch.float64, device=dev, requires_grad=True)
radii = torch.tensor(balls[:, 3], dtype=torch.float64, device=dev)
e = tessellation_energy(centers, radii, torch.tensor(types, device=dev), torch.tensor(W), 1.4)
(2.0 * e).backward()
e_ref, g_ref = _native.energy_gradient(balls, 1.4, types, W)
assert abs(float(e.detach()) - e_ref) <= 1e-12 * abs(e_ref), (float(e.detach()), e_ref)
assert np.allclose(centers.grad.cpu().numpy(), 2.0 * g_ref, rtol=1e-12, atol=1e-12)
print("TORCH_OK")
"""


def test_torch_autograd():
    """The torch.autograd.Function wrapper, in a fresh process that imports
    torch before the library (so both use torch's HIP runtime)."""
    import importlib.util
    import os
    import subprocess
    import sys

    # only look for torch here: importing it would load its own HIP runtime
    # into this process next to the library's
    if importlib.util.find_spec("torch") is None:
        pytest.skip("torch not installed")
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "voronota-ltr_amd")
    out = subprocess.run([sys.executable, "-c", TORCH_CHECK, pkg], capture_output=True, text=True, timeout=300)
    print(out.stdout[-2000:], out.stderr[-2000:], flush=True)
    if "NO_GPU" in out.stdout:
        pytest.skip("torch sees no GPU")
    assert out.returncode == 0 and "TORCH_OK" in out.stdout
