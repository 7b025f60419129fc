"""The oracle's pseudo-solvent spheres (solvent_spheres.rs, subdivided_icosahedron.rs)
against the reference's own unit tests (solvent_spheres.rs:352-481,
subdivided_icosahedron.rs:146-190) and against checks the geometry fixes:

* the icosahedron's base vertices are the normalised cyclic permutations of
  (+-t, +-1, 0) and every subdivision level adds the normalised midpoints of
  the previous level's edges (brute force over nearest-neighbour pairs);
* the sample filter, restated in numpy as a brute force over all balls, for
  clusters in which every pair of overlapping probe-expanded spheres is also
  a contact of the first tessellation (then "closer than every contact
  partner" and "closer than every other ball" agree);
* mirror and icosahedral symmetry of the weights (stage-2 cell volumes).
"""
import math

import numpy as np
import pytest

T = (1.0 + math.sqrt(5.0)) / 2.0


def test_icosahedron_vertex_counts(oracle):
    # subdivided_icosahedron.rs:147-154
    assert [len(oracle.icosahedron(d)) for d in range(5)] == [12, 42, 162, 642, 2562]


def test_icosahedron_unit_sphere(oracle):
    # subdivided_icosahedron.rs:169-178 (and points_on_sphere, 181-190)
    for d in range(5):
        v = oracle.icosahedron(d)
        assert np.all(np.abs(np.linalg.norm(v, axis=1) - 1.0) < 1e-10)
    c = np.array([1.0, 2.0, 3.0])
    p = c + oracle.icosahedron(1) * 5.0
    assert np.all(np.abs(np.linalg.norm(p - c, axis=1) - 5.0) < 1e-10)


def test_icosahedron_base_and_midpoints(oracle):
    base = []
    for s1 in (1.0, -1.0):
        for s2 in (1.0, -1.0):
            for perm in ((s2 * T, s1, 0), (s1, 0, s2 * T), (0, s2 * T, s1)):
                base.append(perm)
    base = np.array(base) / math.sqrt(1.0 + T * T)
    v0 = oracle.icosahedron(0)
    assert sorted(map(tuple, np.round(v0, 12))) == sorted(map(tuple, np.round(base, 12)))
    for d in range(1, 4):
        prev = oracle.icosahedron(d - 1)
        cur = oracle.icosahedron(d)
        # the previous level comes first, unchanged
        assert np.array_equal(cur[: len(prev)], prev)
        # its edges: pairs at the minimal distance (each vertex has 5 or 6)
        dist = np.linalg.norm(prev[:, None, :] - prev[None, :, :], axis=2)
        np.fill_diagonal(dist, np.inf)
        mind = dist.min()
        ii, jj = np.nonzero(np.triu(dist < mind * 1.25, 1))
        assert len(ii) == len(cur) - len(prev)  # one new vertex per edge
        mids = prev[ii] + prev[jj]
        mids /= np.linalg.norm(mids, axis=1)[:, None]
        new = cur[len(prev):]
        assert sorted(map(tuple, np.round(new, 10))) == sorted(map(tuple, np.round(mids, 10)))


def test_single_sphere(oracle):
    # solvent_spheres.rs:352-367
    s = oracle.solvent_spheres([[0.0, 0.0, 0.0, 1.5]], 1.4, 0.0, 0)
    assert len(s) == 12
    assert np.all(s[:, 5] == 0) and np.all(np.abs(s[:, 3] - 1.4) < 1e-10) and np.all(s[:, 4] > 0)
    # icosahedral symmetry: the twelve cells are congruent
    assert np.ptp(s[:, 4]) <= 1e-9 * s[:, 4].max()
    # positions: centre + v (r + probe)
    assert np.allclose(s[:, :3], oracle.icosahedron(0) * 2.9, rtol=0, atol=1e-15)


def test_two_touching(oracle):
    # solvent_spheres.rs:369-387
    s = oracle.solvent_spheres([[0, 0, 0, 1.5], [2.5, 0, 0, 1.5]], 1.4, 0.0, 0)
    assert 0 < len(s) < 24


def test_empty(oracle):
    # solvent_spheres.rs:389-393
    assert len(oracle.solvent_spheres(np.zeros((0, 4)), 1.4)) == 0


def test_two_separated_cpp_values(oracle):
    # solvent_spheres.rs:395-425: 24 spheres, 12 per ball, total weight ~249.1
    s = oracle.solvent_spheres([[0, 0, 0, 1.5], [5, 0, 0, 1.5]], 1.4, 0.0, 0)
    assert len(s) == 24
    assert np.count_nonzero(s[:, 5] == 0) == 12 and np.count_nonzero(s[:, 5] == 1) == 12
    assert abs(s[:, 4].sum() - 249.1) < 5.0
    assert np.all(np.abs(s[:, 3] - 1.4) < 1e-10)
    # mirror symmetry about x = 2.5 (the icosahedron is symmetric under x -> -x)
    w0 = np.sort(s[s[:, 5] == 0, 4])
    w1 = np.sort(s[s[:, 5] == 1, 4])
    assert np.allclose(w0, w1, rtol=1e-9, atol=0)


def test_contained_sphere(oracle):
    # solvent_spheres.rs:427-446
    s = oracle.solvent_spheres([[0, 0, 0, 3.0], [2.5, 0, 0, 1.0]], 1.4, 0.0, 0)
    assert np.count_nonzero(s[:, 5] == 0) > np.count_nonzero(s[:, 5] == 1)


def _brute_samples(oracle, balls, probe, depth):
    """Kept samples by brute force over all other balls (see module doc)."""
    balls = np.asarray(balls, float)
    r = oracle.compute_tessellation(balls, probe)
    v = oracle.icosahedron(depth)
    out = []
    for i, sas in zip(r.cell_index, r.cell_sas):
        if not sas > 0.0:
            continue
        i = int(i)
        c = balls[i, :3]
        pts = c + v * (balls[i, 3] + probe)
        d_atom = np.linalg.norm(pts - c, axis=1) - balls[i, 3]
        keep = np.ones(len(pts), bool)
        for j in range(len(balls)):
            if j != i:
                keep &= np.linalg.norm(pts - balls[j, :3], axis=1) - balls[j, 3] > d_atom
        out += [(p[0], p[1], p[2], i) for p in pts[keep]]
    return np.array(out).reshape(-1, 4), r


@pytest.mark.parametrize("seed", [4, 5, 6])
def test_filter_brute_force(oracle, seed):
    rng = np.random.default_rng(seed)
    probe = 1.4
    balls = []
    while len(balls) < 14:
        c = rng.uniform(-5, 5, 3)
        rad = rng.uniform(1.2, 1.9)
        if all(np.linalg.norm(c - b[:3]) > 0.8 * (rad + b[3]) for b in balls):
            balls.append(np.array([*c, rad]))
    balls = np.array(balls)
    want, r = _brute_samples(oracle, balls, probe, 2)
    contacts = {(int(a), int(b)) for a, b in r.contact_ids}
    n = len(balls)
    for a in range(n):
        for b in range(a + 1, n):
            if np.linalg.norm(balls[a, :3] - balls[b, :3]) < balls[a, 3] + balls[b, 3] + 2 * probe:
                if (a, b) not in contacts:
                    pytest.skip("an overlapping pair without a contact: brute force does not apply")
    got = oracle.solvent_spheres(balls, probe, 0.0, 2)
    assert len(got) == len(want) and len(got) > 0
    assert np.array_equal(got[:, :3], want[:, :3])
    assert np.array_equal(got[:, 5], want[:, 3])


def test_weights_are_stage2_volumes(oracle):
    """weight = volume of the sample's cell in the probe-0 tessellation of
    atoms (radius max(r - probe + volume_probe, 0.01)) and samples (radius
    probe + volume_probe), solvent_spheres.rs:381-425."""
    balls = np.array([[0, 0, 0, 1.6], [2.8, 0.3, 0, 1.5], [1.2, 2.4, 0.5, 1.8]])
    probe, vp = 1.2, 0.3
    s = oracle.solvent_spheres(balls, probe, vp, 1)
    comb = np.vstack([np.c_[balls[:, :3], np.maximum(balls[:, 3] - probe + vp, 0.01)],
                      np.c_[s[:, :3], np.full(len(s), probe + vp)]])
    r = oracle.compute_tessellation(comb, 0.0)
    vol = dict(zip((int(i) for i in r.cell_index), r.cell_volume))
    n = len(balls)
    for k in range(len(s)):
        assert s[k, 4] == vol.get(n + k, 0.0)
