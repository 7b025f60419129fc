"""The oracle's UpdateableTessellation against the reference's tests
(tests/updateable.rs:63-321, src/updateable.rs:754-816)."""
import numpy as np
import pytest

from conftest import RING17, corners_box, load_xyzr

LINE3 = np.array([[0, 0, 0, 1.0], [2, 0, 0, 1.0], [4, 0, 0, 1.0]])


def totals(r):
    return len(r.contact_area), len(r.cell_index), r.contact_area.sum(), r.cell_sas.sum(), r.cell_volume.sum()


def verify_matches_full(oracle, balls, probe, pbox, changed, iters):
    # tests/updateable.rs:63-116
    balls = balls.copy()
    t = oracle.Updateable(with_backup=True)
    assert t.init(balls, probe, pbox)
    for it in range(iters):
        for i in changed:
            balls[i, 0] += 0.05 * (it + 1.0)
        assert t.update_with_changed(balls, changed)
        full = oracle.compute_tessellation(balls, probe, pbox=pbox)
        inc = t.summary()
        ti, tf = totals(inc), totals(full)
        assert ti[0] == tf[0] and ti[1] == tf[1]
        assert abs(ti[2] - tf[2]) < 0.1 and abs(ti[3] - tf[3]) < 0.1
    assert t.restore()
    assert len(t.summary().contact_area) > 0


def test_ring17_no_periodic(oracle):
    verify_matches_full(oracle, RING17, 1.0, None, [0, 8], 3)


def test_three_line(oracle):
    verify_matches_full(oracle, LINE3, 0.5, None, [1], 5)


def test_ring_periodic_port(oracle):
    # tests/updateable.rs:134-192
    balls = RING17.copy()
    box = corners_box((-1.6, -1.6, -0.6), (1.6, 1.6, 3.1))
    t = oracle.Updateable(with_backup=True)
    assert t.init(balls, 1.0, box)
    full = oracle.compute_tessellation(balls, 1.0, pbox=box)
    up = t.summary()
    assert totals(up)[:2] == totals(full)[:2]
    assert abs(totals(up)[2] - totals(full)[2]) < 0.01
    for _ in range(5):
        balls[[0, 1], 0] += 0.1
        assert t.update_with_changed(balls, [0, 1])
        full = oracle.compute_tessellation(balls, 1.0, pbox=box)
        up = t.summary()
        tu, tf = totals(up), totals(full)
        assert tu[:2] == tf[:2]
        assert abs(tu[2] - tf[2]) < 0.1
        assert abs(tu[3] - tf[3]) < 0.01
        assert abs(tu[4] - tf[4]) < 0.01
    before = t.summary()
    assert t.restore()
    after = t.summary()
    assert abs(before.contact_area.sum() - after.contact_area.sum()) > 0.01


def test_exclusion(oracle):
    # tests/updateable.rs:194-220
    t = oracle.Updateable(with_backup=True)
    assert t.init(LINE3, 0.5)
    assert totals(t.summary())[:2] == (2, 3)
    assert t.set_exclusion(1, True)
    assert totals(t.summary())[:2] == (0, 2)
    assert t.restore()
    assert totals(t.summary())[:2] == (2, 3)


def test_auto_detect_and_no_change(oracle):
    # tests/updateable.rs:222-255
    b = np.array([[0, 0, 0, 1.0], [2, 0, 0, 1.0]])
    t = oracle.Updateable()
    assert t.init(b, 0.5)
    n1 = len(t.summary().contact_area)
    b2 = b.copy()
    b2[0, 0] += 0.1
    assert t.update(b2)
    assert len(t.summary().contact_area) == n1
    assert 0 in t.changed_ids()
    t2 = oracle.Updateable()
    assert t2.init(b, 0.5)
    assert not t2.update(b)


def test_full_reinit_threshold(oracle):
    # tests/updateable.rs:270-288
    n = 20
    b = np.array([[2.0 * i, 0, 0, 1.0] for i in range(n)])
    t = oracle.Updateable()
    assert t.init(b, 0.5)
    ch = list(range(n // 2 + 1))
    m = b.copy()
    m[ch, 0] += 0.1
    assert t.update_with_changed(m, ch)
    assert t.last_update_was_full_reinit()


def test_2zsk_incremental(oracle):
    # tests/updateable.rs:290-321
    balls = load_xyzr("balls_2zsk.xyzr")
    t = oracle.Updateable()
    assert t.init(balls, 1.4)
    s = t.summary()
    assert len(s.contact_area) == 23855 and len(s.cell_index) == 3545
    ch = [0, 100, 500]
    balls = balls.copy()
    balls[ch, 0] += 0.1
    assert t.update_with_changed(balls, ch)
    assert not t.last_update_was_full_reinit()
    full = oracle.compute_tessellation(balls, 1.4)
    inc = t.summary()
    assert len(inc.contact_area) == len(full.contact_area)
    assert len(inc.cell_index) == len(full.cell_index)


def test_updateable_unit_cases(oracle):
    # updateable.rs:760-815
    t = oracle.Updateable()
    assert t.init(LINE3, 0.5)
    assert totals(t.summary())[:2] == (2, 3)
    b = np.array([[0, 0, 0, 1.0], [2, 0, 0, 1.0]])
    t = oracle.Updateable(with_backup=True)
    assert t.init(b, 0.5)
    b1 = b.copy()
    b1[0, 0] += 0.1
    assert t.update_with_changed(b1, [0])
    b2 = b1.copy()
    b2[0, 0] += 0.1
    assert t.update_with_changed(b2, [0])
    after = t.summary().contact_area[0]
    assert t.restore()
    assert abs(t.summary().contact_area[0] - after) > 1e-4
    # updateable.rs:215-220: an empty change list is "no update"
    assert not t.update_with_changed(b2, [])


@pytest.mark.slow
def test_incremental_equals_full_on_rsa(oracle):
    """Contacts and cells of the incremental path equal a full recompute to
    1e-12 on a dense cluster moved a few balls at a time."""
    from voronota_ltr import workload

    balls, _ = workload.rsa_sphere(3000, seed=77)
    rng = np.random.default_rng(1)
    t = oracle.Updateable()
    assert t.init(balls, 1.4)
    for _ in range(4):
        ch = np.sort(rng.choice(len(balls), 15, replace=False))
        balls[ch, :3] += rng.normal(scale=0.05, size=(15, 3))
        assert t.update_with_changed(balls, ch)
        inc = t.summary()
        full = oracle.compute_tessellation(balls, 1.4)
        fi = {tuple(int(x) for x in k): v for k, v in zip(full.contact_ids, full.contact_area)}
        ii = {tuple(int(x) for x in k): v for k, v in zip(inc.contact_ids, inc.contact_area)}
        assert fi.keys() == ii.keys()
        assert max(abs(fi[k] - ii[k]) for k in fi) < 1e-12
        assert np.allclose(inc.cell_sas, full.cell_sas, rtol=0, atol=1e-10)
        assert np.allclose(inc.cell_volume, full.cell_volume, rtol=0, atol=1e-10)


def _lattice(n_side, seed):
    g = np.arange(n_side) * 3.0
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    balls = np.column_stack([x.ravel(), y.ravel(), z.ravel(), np.full(x.size, 1.6)])
    balls[:, :3] += np.random.default_rng(seed).normal(scale=0.2, size=(len(balls), 3))
    return balls


def _cross_box_moves(oracle, remove_from_old_box, steps=25):
    """Three balls at a time moved about 0.4 A per axis (many cross a grid
    box of the searcher) on a 6x6x6 lattice; yields (summary, full
    tessellation) after each update."""
    rng = np.random.default_rng(5)
    balls = _lattice(6, 5)
    t = oracle.Updateable(remove_from_old_box=remove_from_old_box)
    assert t.init(balls, 1.4)
    for _ in range(steps):
        ids = rng.choice(len(balls), 3, replace=False)
        balls[ids, :3] += rng.normal(scale=0.4, size=(3, 3))
        assert t.update_with_changed(balls, list(ids))
        yield t.summary(), oracle.compute_tessellation(balls, 1.4)


def test_searcher_keeps_stale_box_entries(oracle):
    """spheres_searcher.rs:144-175 writes a moved sphere before
    remove_sphere_from_grid looks up its box, so a sphere moved to another
    box stays listed in the old one: a search reaching both boxes reports it
    twice, collect_pairs (updateable.rs:516-561) lists the pair twice and
    update_using_affected stores the contact twice, so summary() reports
    duplicate contacts and the cell sums count them twice. The restatement
    keeps this by default; it appears only once a moved ball has changed box."""
    dups = [len(s.contact_ids) - len({tuple(int(v) for v in k) for k in s.contact_ids})
            for s, _ in _cross_box_moves(oracle, False)]
    assert max(dups) > 0


def test_corrected_searcher_equals_full(oracle):
    """With moved spheres taken out of their old boxes, the incremental
    state equals a full recompute after every move (contact set exactly,
    values to 1e-12)."""
    for inc, full in _cross_box_moves(oracle, True):
        ki = [tuple(int(v) for v in k) for k in inc.contact_ids]
        kf = [tuple(int(v) for v in k) for k in full.contact_ids]
        assert sorted(ki) == sorted(kf)
        a = dict(zip(ki, inc.contact_area))
        assert max(abs(a[k] - v) for k, v in zip(kf, full.contact_area)) < 1e-12
        assert list(inc.cell_index) == list(full.cell_index)
        assert np.allclose(inc.cell_volume, full.cell_volume, rtol=0, atol=1e-10)
        assert np.allclose(inc.cell_sas, full.cell_sas, rtol=0, atol=1e-10)
