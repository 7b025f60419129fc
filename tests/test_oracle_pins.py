"""Pins of the oracle beyond the reference's goldens.

Each check is an independent statement of the mathematics, chosen so that a
plausible restatement error (dropped term, wrong sign, swapped operand) fails:
  * two-ball closed forms (disk area, arc, cap SAS, cap volume);
  * an exact face area by a different algorithm: the disk of the contact
    clipped by every other ball's radical half-plane (Sutherland-Hodgman on a
    square, then circle/polygon intersection by signed circular sectors);
  * SAS by quadrature on a spherical Fibonacci lattice (points of sphere i
    outside every other ball) and cell volume by quadrature (points of ball i
    whose power distance to i is smallest);
  * rigid-motion and scale invariance;
  * forward-mode dual derivative against central differences.
"""
import math
import os

import numpy as np
import pytest

from conftest import CENTRAL8


def _rsa_cluster(n, seed):
    from voronota_ltr import workload

    return workload.rsa_sphere(n, seed=seed)[0]


# ------------------------------------------------------------ closed forms --
@pytest.mark.parametrize("ra,rb,d,probe", [(1.5, 1.5, 2.5, 1.4), (1.0, 2.0, 2.2, 0.5), (0.6, 2.0, 1.6, 0.1),
                                           (1.9, 1.2, 3.0, 1.4), (1.5, 1.5, 0.3, 1.0)])
def test_two_ball_closed_form(oracle, ra, rb, d, probe):
    b = np.array([[0.0, 0.0, 0.0, ra], [d, 0.0, 0.0, rb]])
    r = oracle.compute_tessellation(b, probe)
    Ra, Rb = ra + probe, rb + probe
    ha = (Ra * Ra + d * d - Rb * Rb) / (2 * d)
    hb = d - ha
    ric = math.sqrt(Ra * Ra - ha * ha)
    assert len(r.contact_area) == 1
    assert r.contact_area[0] == pytest.approx(math.pi * ric * ric, rel=1e-12)
    assert r.contact_arc[0] == pytest.approx(2 * math.pi * ric, rel=1e-12)
    for i, (R, h) in enumerate([(Ra, ha), (Rb, hb)]):
        t = R - h
        sas = 2 * math.pi * R * (R + h)
        vol = 4.0 / 3.0 * math.pi * R ** 3 - math.pi * t * t * (3 * R - t) / 3.0
        assert r.cell_sas[i] == pytest.approx(sas, rel=1e-10, abs=1e-12 * 4 * math.pi * R * R)
        assert r.cell_volume[i] == pytest.approx(vol, rel=1e-10)


# ---------------------------------------------------- independent face area --
def _cross2(a, b):
    return a[0] * b[1] - a[1] * b[0]


def _sector(u, v, r):
    return 0.5 * r * r * math.atan2(_cross2(u, v), u[0] * v[0] + u[1] * v[1])


def _tri_disk(p, q, r):
    """Signed area of disk(0, r) intersected with triangle (0, p, q)."""
    d = q - p
    a = d @ d
    if a == 0.0:
        return 0.0
    b = p @ d
    c = p @ p - r * r
    pin, qin = p @ p <= r * r, q @ q <= r * r
    if pin and qin:
        return 0.5 * _cross2(p, q)
    disc = b * b - a * c
    if disc <= 0.0:
        return _sector(p, q, r)
    s = math.sqrt(disc)
    t1, t2 = (-b - s) / a, (-b + s) / a
    if pin:
        m = p + t2 * d
        return 0.5 * _cross2(p, m) + _sector(m, q, r)
    if qin:
        m = p + t1 * d
        return _sector(p, m, r) + 0.5 * _cross2(m, q)
    if 0.0 < t1 < t2 < 1.0:
        m1, m2 = p + t1 * d, p + t2 * d
        return _sector(p, m1, r) + 0.5 * _cross2(m1, m2) + _sector(m2, q, r)
    return _sector(p, q, r)


def _face_area(S, a, b):
    """Disk of the a|b intersection circle, clipped by pow_a <= pow_c for all c."""
    ca, Ra = S[a, :3], S[a, 3]
    cb, Rb = S[b, :3], S[b, 3]
    ab = cb - ca
    d = np.linalg.norm(ab)
    n = ab / d
    h = (Ra * Ra + d * d - Rb * Rb) / (2 * d)
    r2 = Ra * Ra - h * h
    if r2 <= 0:
        return 0.0
    r = math.sqrt(r2)
    o = ca + n * h
    u = np.cross(n, [1.0, 0.0, 0.0] if abs(n[0]) < 0.9 else [0.0, 1.0, 0.0])
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    poly = [np.array(p) for p in ((-2 * r, -2 * r), (2 * r, -2 * r), (2 * r, 2 * r), (-2 * r, 2 * r))]
    for c in range(len(S)):
        if c in (a, b):
            continue
        cc, Rc = S[c, :3], S[c, 3]
        # 2 x.(cc - ca) <= |cc|^2 - |ca|^2 - Rc^2 + Ra^2 with x = o + s u + t v
        w = 2.0 * (cc - ca)
        k = cc @ cc - ca @ ca - Rc * Rc + Ra * Ra - w @ o
        nrm = np.array([w @ u, w @ v])
        out = []
        for i in range(len(poly)):
            p, q = poly[i], poly[(i + 1) % len(poly)]
            fp, fq = nrm @ p - k, nrm @ q - k
            if fp <= 0:
                out.append(p)
            if (fp < 0) != (fq < 0) and fp != fq:
                out.append(p + (q - p) * (fp / (fp - fq)))
        poly = out
        if len(poly) < 3:
            return 0.0
    return abs(sum(_tri_disk(poly[i], poly[(i + 1) % len(poly)], r) for i in range(len(poly))))


@pytest.mark.parametrize("seed", [3, 4])
def test_face_area_independent(oracle, seed):
    balls = _rsa_cluster(60, seed)
    probe = 1.4
    S = balls.copy()
    S[:, 3] += probe
    r = oracle.compute_tessellation(balls, probe)
    got = {tuple(int(x) for x in ids): a for ids, a in zip(r.contact_ids, r.contact_area)}
    n = len(S)
    bad = 0
    for a in range(n):
        for b in range(a + 1, n):
            if np.linalg.norm(S[a, :3] - S[b, :3]) >= S[a, 3] + S[b, 3]:
                continue
            ref = _face_area(S, a, b)
            g = got.get((a, b), 0.0)
            scale = math.pi * max(S[a, 3], S[b, 3]) ** 2
            if abs(g - ref) > 1e-9 * ref + 1e-11 * scale:
                bad += 1
                assert abs(g - ref) <= 1e-6 * scale, f"pair {(a, b)}: oracle {g} vs exact {ref}"
    assert bad <= 2
    assert len(got) > 100


# -------------------------------------------------------------- quadrature --
def _fib_sphere(m):
    k = np.arange(m) + 0.5
    z = 1 - 2 * k / m
    phi = k * math.pi * (3 - math.sqrt(5))
    s = np.sqrt(1 - z * z)
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)


def test_sas_and_volume_quadrature(oracle):
    balls = _rsa_cluster(40, 9)
    probe = 1.4
    S = balls.copy()
    S[:, 3] += probe
    r = oracle.compute_tessellation(balls, probe)
    sas = dict(zip(r.cell_index.astype(int), r.cell_sas))
    vol = dict(zip(r.cell_index.astype(int), r.cell_volume))
    unit = _fib_sphere(40000)
    rng = np.random.default_rng(5)
    inner = rng.normal(size=(60000, 3))
    inner /= np.linalg.norm(inner, axis=1)[:, None]
    inner *= rng.random(60000)[:, None] ** (1.0 / 3.0)
    for i in range(len(S)):
        c, R = S[i, :3], S[i, 3]
        pts = c + R * unit
        d2 = ((pts[:, None, :] - S[None, :, :3]) ** 2).sum(-1)
        other = np.ones(len(S), bool)
        other[i] = False
        outside = np.all(d2[:, other] >= S[other, 3] ** 2, axis=1)
        est = outside.mean() * 4 * math.pi * R * R
        assert abs(sas[i] - est) < 4e-3 * 4 * math.pi * R * R, (i, sas[i], est)
        q = c + R * inner
        pw = ((q[:, None, :] - S[None, :, :3]) ** 2).sum(-1) - S[None, :, 3] ** 2
        mine = np.argmin(pw, axis=1) == i
        vest = mine.mean() * 4.0 / 3.0 * math.pi * R ** 3
        assert abs(vol[i] - vest) < 1.5e-2 * 4.0 / 3.0 * math.pi * R ** 3, (i, vol[i], vest)


# -------------------------------------------------------------- invariance --
def _keyed(r):
    return {tuple(int(x) for x in ids): (a, arc) for ids, a, arc in zip(r.contact_ids, r.contact_area, r.contact_arc)}


def test_rigid_motion_and_scale(oracle):
    balls = _rsa_cluster(300, 11)
    probe = 1.4
    base = oracle.compute_tessellation(balls, probe)
    th = 0.7
    Rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    moved = balls.copy()
    moved[:, :3] = balls[:, :3] @ Rz.T + np.array([3.5, -1.25, 2.0])
    scaled = balls * 2.0
    for other, probe2, fa, fv in [(moved, probe, 1.0, 1.0), (scaled, 2 * probe, 4.0, 8.0)]:
        r = oracle.compute_tessellation(other, probe2)
        kb, ko = _keyed(base), _keyed(r)
        assert kb.keys() == ko.keys()
        for k in kb:
            assert ko[k][0] == pytest.approx(fa * kb[k][0], rel=1e-8, abs=1e-9)
        assert np.allclose(r.cell_sas, fa * base.cell_sas, rtol=1e-8, atol=1e-8)
        assert np.allclose(r.cell_volume, fv * base.cell_volume, rtol=1e-8, atol=1e-8)


# ------------------------------------------------------------ derivatives --
def test_dual_matches_central_differences(oracle):
    # Generic geometry: at exactly axis-aligned pairs (CENTRAL8 has several)
    # any_normal_of_vector (geometry.rs:179-193) switches branch within 1e-10
    # of the point, so the hexagon is not a smooth function of the input there
    # and the computed energy has no derivative.
    balls = _rsa_cluster(40, 21)
    types = np.arange(len(balls), dtype=np.int32) % 4
    W = np.array([[0.3, -0.2, 0.5, 0.1], [-0.2, 0.7, -0.4, 0.2], [0.5, -0.4, -0.6, 0.9], [0.1, 0.2, 0.9, -0.3]])
    h = 1e-6
    for ball in (0, 13, 27):
        for axis in range(3):
            e0, de = oracle.energy_dual(balls, 1.4, types, W, ball, axis)
            bp, bm = balls.copy(), balls.copy()
            bp[ball, axis] += h
            bm[ball, axis] -= h
            fd = (oracle.energy(bp, 1.4, types, W) - oracle.energy(bm, 1.4, types, W)) / (2 * h)
            assert e0 == pytest.approx(oracle.energy(balls, 1.4, types, W), rel=1e-14)
            assert de == pytest.approx(fd, rel=1e-6, abs=1e-6)


def test_periodic_dual_matches_central_differences(oracle):
    """The same in a periodic box (contacts through images, the deduplicated
    contact list, cells of canonical balls only): the dual derivative of the
    reference computation equals central differences of the energy."""
    from conftest import corners_box
    from voronota_ltr import workload

    L = 13.0
    balls, _ = workload.rsa_box(120, L, seed=5)
    box = corners_box((0, 0, 0), (L, L, L))
    types = np.arange(len(balls), dtype=np.int32) % 4
    W = np.array([[0.3, -0.2, 0.5, 0.1], [-0.2, 0.7, -0.4, 0.2], [0.5, -0.4, -0.6, 0.9], [0.1, 0.2, 0.9, -0.3]])
    h = 1e-6
    for ball in (0, 7, 50, 119):
        for axis in range(3):
            e0, de = oracle.energy_dual(balls, 1.4, types, W, ball, axis, pbox=box)
            bp, bm = balls.copy(), balls.copy()
            bp[ball, axis] += h
            bm[ball, axis] -= h
            fd = (oracle.energy(bp, 1.4, types, W, pbox=box) - oracle.energy(bm, 1.4, types, W, pbox=box)) / (2 * h)
            assert e0 == pytest.approx(oracle.energy(balls, 1.4, types, W, pbox=box), rel=1e-14)
            assert de == pytest.approx(fd, rel=1e-6, abs=1e-6), (ball, axis)


def test_counted_flops_are_deterministic(oracle):
    balls = _rsa_cluster(500, 13)
    f1 = oracle.count_flops(balls, 1.4, mode=0)
    f2 = oracle.count_flops(balls, 1.4, mode=0)
    s = oracle.count_flops(balls, 1.4, mode=2)
    assert f1 == f2 and f1[0] > s[0] > 0


def test_work_counts_fixture(oracle):
    """tests/golden/data/work_counts.json (the algorithmic work bench.py
    prices the clip kernel with) is what the counting scalar gives now: the
    20k workload recounted in full, the 100k one in its full-mode total."""
    import json

    from voronota_ltr import workload

    with open(os.path.join(os.path.dirname(__file__), "golden", "data", "work_counts.json")) as f:
        wl = json.load(f)["workloads"]
    w = wl["rsa_sphere_20000"]
    balls, _ = workload.rsa_sphere(w["balls"], seed=w["seed"])
    sfl, stops = oracle.count_flops(balls, w["probe"], mode=2)
    assert (sfl, stops) == (w["mode0"]["flops_search"], w["mode0"]["tops_search"])
    for mode in (0, 1):
        fl, tops = oracle.count_flops(balls, w["probe"], mode=mode)
        assert (fl, tops) == (w[f"mode{mode}"]["flops_total"], w[f"mode{mode}"]["tops_total"])
        assert w[f"mode{mode}"]["flops_contacts"] == fl - sfl
    off, idx, _, _, pairs = oracle.collisions(balls, w["probe"])
    assert len(pairs) == w["candidate_pairs"]
    w = wl["rsa_sphere_100000"]
    balls, _ = workload.rsa_sphere(w["balls"], seed=w["seed"])
    assert oracle.count_flops(balls, w["probe"], mode=0) == (w["mode0"]["flops_total"], w["mode0"]["tops_total"])
