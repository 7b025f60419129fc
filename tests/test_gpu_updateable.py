"""The device UpdateableTessellation against the oracle's restatement of
updateable.rs / spheres_container.rs, operation by operation: the returned
booleans, changed and affected ids and the full-reinit flag exactly; the
summary's contact set and cell set exactly, their values to the parity bar
(1e-9 relative with the full-sphere / full-disk floors). Scenarios follow the
reference's tests/updateable.rs:63-321 and updateable.rs:754-816, plus seeded
random moves of dense and periodic clusters.
"""
import math

import numpy as np
import pytest

from conftest import RING17, corners_box, load_xyzr

pytestmark = pytest.mark.gpu

LINE3 = np.array([[0, 0, 0, 1.0], [2, 0, 0, 1.0], [4, 0, 0, 1.0]])


def _close(a, b, scale, what):
    tol = 1e-9 * abs(b) + 1e-12 * scale
    assert abs(a - b) <= tol, f"{what}: gpu {a!r} oracle {b!r}"


def _same_summary(gpu, ref, balls, probe):
    n = len(balls)
    R = np.asarray(balls, float)[:, 3] + probe if n else np.zeros(0)
    g = {(int(c["id_a"]), int(c["id_b"])): c for c in gpu.contacts}
    r = {(int(a), int(b)): k for k, (a, b) in enumerate(ref.contact_ids)}
    assert g.keys() == r.keys(), (f"contact sets differ: only on the GPU {sorted(set(g) - set(r))[:6]}, "
                                  f"only in the oracle {sorted(set(r) - set(g))[:6]}")
    for key, k in r.items():
        rm = max(R[key[0]], R[key[1]])
        _close(g[key]["area"], ref.contact_area[k], math.pi * rm * rm, f"area {key}")
        _close(g[key]["arc_length"], ref.contact_arc[k], 2 * math.pi * rm, f"arc {key}")
        assert bool(g[key]["central"]) == bool(ref.contact_central[k]), f"central {key}"
    assert [int(i) for i in gpu.cells["index"]] == [int(i) for i in ref.cell_index]
    for c, sas, vol in zip(gpu.cells, ref.cell_sas, ref.cell_volume):
        Ri = R[int(c["index"])]
        _close(c["sas_area"], sas, 4 * math.pi * Ri * Ri, f"sas {int(c['index'])}")
        _close(c["volume"], vol, 4.0 / 3.0 * math.pi * Ri ** 3, f"volume {int(c['index'])}")


class Pair:
    """Runs every operation on the device updateable and on the oracle's."""

    def __init__(self, oracle, with_backup=False, remove_from_old_box=False):
        from voronota_ltr import _native

        self.g = _native.Updateable(with_backup=with_backup)
        self.o = oracle.Updateable(with_backup=with_backup, remove_from_old_box=remove_from_old_box)
        self.balls = None
        self.probe = None

    def check(self):
        assert [int(i) for i in self.g.changed_ids()] == [int(i) for i in self.o.changed_ids()]
        assert [int(i) for i in self.g.affected_ids()] == [int(i) for i in self.o.affected_ids()]
        assert self.g.last_update_was_full_reinit() == self.o.last_update_was_full_reinit()
        _same_summary(self.g.summary(), self.o.summary(), self.balls, self.probe)

    def op(self, name, *args):
        a = getattr(self.g, name)(*args)
        b = getattr(self.o, name)(*args)
        assert a == b, f"{name} returned {a} on the GPU, {b} in the oracle"
        self.check()
        return a

    def init(self, balls, probe, pbox=None):
        self.balls = np.array(balls, float)
        self.probe = probe
        return self.op("init", self.balls, probe, pbox)

    def update(self, balls, changed=None):
        self.balls = np.array(balls, float)
        if changed is None:
            return self.op("update", self.balls)
        return self.op("update_with_changed", self.balls, list(changed))


def _moves(p, balls, changed, iters):
    balls = balls.copy()
    for it in range(iters):
        for i in changed:
            balls[i, 0] += 0.05 * (it + 1.0)
        print(f"[gpu] updateable step {it}: {len(changed)} changed of {len(balls)}", flush=True)
        assert p.update(balls, changed)
    return balls


def test_ring17_and_restore(oracle):
    p = Pair(oracle, with_backup=True)
    assert p.init(RING17, 1.0)
    _moves(p, RING17, [0, 8], 3)
    assert p.op("restore")
    assert not p.op("restore")


def test_three_line(oracle):
    p = Pair(oracle)
    assert p.init(LINE3, 0.5)
    _moves(p, LINE3, [1], 5)


def test_ring17_periodic(oracle):
    p = Pair(oracle, with_backup=True)
    box = corners_box((-1.6, -1.6, -0.6), (1.6, 1.6, 3.1))
    assert p.init(RING17, 1.0, box)
    balls = RING17.copy()
    for _ in range(5):
        balls[[0, 1], 0] += 0.1
        assert p.update(balls, [0, 1])
    assert p.op("restore")


def test_exclusion_and_restore(oracle):
    p = Pair(oracle, with_backup=True)
    assert p.init(LINE3, 0.5)
    assert p.op("set_exclusion", 1, True)
    assert not p.op("set_exclusion", 1, True)
    assert p.op("restore")
    assert p.op("set_exclusion", 0, True)
    assert p.op("set_exclusion", 0, False)
    assert not p.op("set_exclusion", 7, True)


def test_auto_detect_no_change_and_reinit(oracle):
    p = Pair(oracle)
    b = np.array([[0, 0, 0, 1.0], [2, 0, 0, 1.0]])
    assert p.init(b, 0.5)
    b2 = b.copy()
    b2[0, 0] += 0.1
    assert p.update(b2)
    assert not p.update(b2)
    assert not p.update(b2, [])
    # below the comparison tolerance: not a change
    b3 = b2.copy()
    b3[1, 1] += 1e-12
    assert not p.update(b3, [1])
    # more than half the balls changed: full reinit
    n = 20
    line = np.array([[2.0 * i, 0, 0, 1.0] for i in range(n)])
    q = Pair(oracle)
    assert q.init(line, 0.5)
    m = line.copy()
    m[: n // 2 + 1, 0] += 0.1
    assert q.update(m, list(range(n // 2 + 1)))
    assert q.g.last_update_was_full_reinit()
    # a different ball count: full reinit
    assert q.update(line[:7])
    # few changed balls with many old neighbors: changed + old neighbors past
    # the threshold (24 balls that all collide: 2 + 22 > 12) is a full reinit,
    # and a repeated changed id counts twice; below it the update is incremental
    rng = np.random.default_rng(24)
    blob = np.column_stack([rng.uniform(0, 3, size=(24, 3)), np.full(24, 1.0)])
    r = Pair(oracle)
    assert r.init(blob, 0.5)
    bm = blob.copy()
    bm[[3, 17], 0] += 0.1
    assert r.update(bm, [3, 17])
    assert r.g.last_update_was_full_reinit()
    bm[5, 1] += 0.1
    assert r.update(bm, [5, 5])
    assert r.g.last_update_was_full_reinit()
    # (a repeated id is added twice to its grid box by the reference's
    # searcher, spheres_searcher.rs:164-173, and then listed twice: the
    # oracle adds it once here, as the device path counts each pair once)
    chain = np.array([[2.0 * i, 0.1 * (i % 3), 0, 1.0] for i in range(30)])
    c = Pair(oracle, remove_from_old_box=True)
    assert c.init(chain, 0.5)
    cm = chain.copy()
    cm[[4, 15, 26], 1] += 0.1
    assert c.update(cm, [4, 15, 26])
    assert not c.g.last_update_was_full_reinit()
    cm[[7, 8], 2] += 0.1
    assert c.update(cm, [7, 8, 8])
    assert not c.g.last_update_was_full_reinit()


def test_2zsk_incremental(oracle):
    p = Pair(oracle)
    balls = load_xyzr("balls_2zsk.xyzr")
    assert p.init(balls, 1.4)
    b = balls.copy()
    b[[0, 100, 500], 0] += 0.1
    assert p.update(b, [0, 100, 500])
    assert not p.g.last_update_was_full_reinit()


def test_rsa_random_moves(oracle):
    from voronota_ltr import workload

    balls, _ = workload.rsa_sphere(3000, seed=77)
    rng = np.random.default_rng(1)
    p = Pair(oracle, with_backup=True)
    assert p.init(balls, 1.4)
    for step in range(4):
        ch = np.sort(rng.choice(len(balls), 15, replace=False))
        balls[ch, :3] += rng.normal(scale=0.05, size=(15, 3))
        print(f"[gpu] rsa step {step}", flush=True)
        assert p.update(balls, ch)
    assert p.op("restore")
    assert p.op("set_exclusion", 5, True)


def test_periodic_box_moves(oracle):
    """A small version of the periodic dynamics workload: RSA in a cube,
    balls displaced by N(0, 0.05^2) per axis and wrapped back into the box."""
    from voronota_ltr import workload

    L = 32.0
    balls, _ = workload.rsa_box(1500, L, seed=11)
    box = corners_box((0, 0, 0), (L, L, L))
    rng = np.random.default_rng(4)
    p = Pair(oracle)
    assert p.init(balls, 1.4, box)
    for step in range(3):
        ch = np.sort(rng.choice(len(balls), 40, replace=False))
        balls[ch, :3] = np.mod(balls[ch, :3] + rng.normal(scale=0.05, size=(40, 3)), L)
        print(f"[gpu] periodic step {step}", flush=True)
        assert p.update(balls, ch)


def _same_gradient(g, g2):
    """Gradients to 1e-9 relative plus 1e-9 of the gradient's scale. An
    incremental periodic step takes a cell's terms of each canonical pair
    through an image from that pair's first record (the reference's
    deduplicated per-ball lists, updateable.rs:565-727), a standalone run
    from each ball's own record: the two frames agree to roundoff, which the
    adjoint can amplify to a few 1e-9 where a contour has a near-degenerate
    vertex."""
    tol = 1e-9 * np.abs(g2) + 1e-9 * np.abs(g2).max()
    d = np.abs(g - g2)
    assert np.all(d <= tol), f"gradient differs by {d.max():.3e} (rows {sorted(set(np.argwhere(d > tol)[:, 0]))[:5]})"


def test_energy_gradient_with_updates(oracle):
    """The energy and gradient computed in the updateable's own pass equal a
    standalone forward + adjoint of the same balls (periodic box), after
    init, incremental updates, set_energy after the fact and restore; and the
    oracle's dual numbers on a few coordinates."""
    from voronota_ltr import _native, workload

    L = 32.0
    balls, _ = workload.rsa_box(1500, L, seed=12)
    box = corners_box((0, 0, 0), (L, L, L))
    types = np.arange(len(balls), dtype=np.int32) % 4
    W = workload.type_matrix(3)
    u = _native.Updateable(with_backup=True)
    u.set_energy(types, W)
    assert u.init(balls, 1.4, box)
    rng = np.random.default_rng(6)

    def check(b):
        e, g = u.energy_gradient()
        e2, g2 = _native.energy_gradient(b, 1.4, types, W, pbox=box)
        assert e == pytest.approx(e2, rel=1e-12)
        _same_gradient(g, g2)
        return g

    check(balls)
    for step in range(3):
        ch = np.sort(rng.choice(len(balls), 30, replace=False))
        balls[ch, :3] = np.mod(balls[ch, :3] + rng.normal(scale=0.05, size=(30, 3)), L)
        print(f"[gpu] updateable energy step {step}", flush=True)
        assert u.update_with_changed(balls, ch)
        g = check(balls)
    for ball in rng.choice(len(balls), 4, replace=False):
        _, de = oracle.energy_dual(balls, 1.4, types, W, int(ball), 0, pbox=box)
        assert abs(g[ball, 0] - de) <= 1e-7 * max(1.0, abs(de))
    prev = balls.copy()
    ch = np.arange(10)
    balls[ch, 0] = np.mod(balls[ch, 0] + 0.1, L)
    assert u.update_with_changed(balls, ch)
    assert u.restore()
    check(prev)
    # switched on after the state exists: computed on request
    v = _native.Updateable()
    assert v.init(prev, 1.4, box)
    v.set_energy(types, W)
    e, g = v.energy_gradient()
    e2, g2 = _native.energy_gradient(prev, 1.4, types, W, pbox=box)
    assert e == pytest.approx(e2, rel=1e-12)
    _same_gradient(g, g2)


def test_hiding_and_exclusion_moves(oracle):
    """Degenerate transitions with the re-clipping of affected pairs only: a
    ball moved onto another (a duplicate, hidden by index), inside another
    (hidden), back out, and moves next to a ball the caller excluded; every
    step against the oracle's updateable. The moves cross grid boxes of the
    reference's searcher, which then lists the moved ball twice
    (spheres_searcher.rs:144-175): the oracle runs with that corrected, as the
    device path computes every pair once. The stale statuses the reference
    keeps are matched: a hidden ball's exclusion stays until the ball is
    affected again, and a ball the caller excluded is included again when a
    neighbor moves."""
    g = np.arange(5) * 3.0
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    balls = np.column_stack([x.ravel(), y.ravel(), z.ravel(), np.full(x.size, 1.6)])
    balls[:, :3] += np.random.default_rng(11).normal(scale=0.2, size=(len(balls), 3))
    p = Pair(oracle, with_backup=True, remove_from_old_box=True)
    print("\n[gpu] updateable hiding moves", flush=True)
    assert p.init(balls, 1.4)
    cur = balls.copy()
    home = cur[62].copy()
    cur[62, :3] = cur[63, :3]  # onto another ball: equal spheres, the higher index hidden
    cur[62, 3] = cur[63, 3]
    p.update(cur, [62])
    cur[62, :3] = cur[37, :3] + 0.1  # inside a bigger ball
    cur[62, 3] = 0.4
    p.update(cur, [62])
    cur[62] = home  # back out
    p.update(cur, [62])
    p.op("set_exclusion", 31, True)
    for k in (30, 32, 56):  # neighbors of the excluded ball move
        cur[k, :3] += 0.35
        p.update(cur, [k])
    p.op("set_exclusion", 31, False)
    cur[31, :3] += 0.2
    p.update(cur, [31])
    p.op("restore")


def test_cross_box_moves(oracle):
    """Three balls of a 6x6x6 lattice moved at a time, about 0.4 A per axis,
    so that many cross a grid box: every step against the oracle's updateable
    with moved spheres taken out of their old boxes (the reference keeps them
    listed there as well, spheres_searcher.rs:144-175, and then reports
    duplicate contacts)."""
    g = np.arange(6) * 3.0
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    balls = np.column_stack([x.ravel(), y.ravel(), z.ravel(), np.full(x.size, 1.6)])
    balls[:, :3] += np.random.default_rng(5).normal(scale=0.2, size=(len(balls), 3))
    rng = np.random.default_rng(5)
    p = Pair(oracle, remove_from_old_box=True)
    ref = oracle.Updateable()  # the reference's searcher, stale entries and all
    print("\n[gpu] updateable cross-box moves", flush=True)
    assert p.init(balls, 1.4) and ref.init(balls, 1.4)
    for _ in range(25):
        ids = rng.choice(len(balls), 3, replace=False)
        balls[ids, :3] += rng.normal(scale=0.4, size=(3, 3))
        assert p.update(balls, ids)
        assert ref.update_with_changed(balls, list(ids))
    # kept: the corrected behaviour. The reference reports some contacts
    # twice by now; this library lists each pair once.
    ref_ids = [tuple(int(v) for v in k) for k in ref.summary().contact_ids]
    gpu_ids = [(int(c["id_a"]), int(c["id_b"])) for c in p.g.summary().contacts]
    assert len(ref_ids) > len(set(ref_ids))
    assert len(gpu_ids) == len(set(gpu_ids)) and set(gpu_ids) == set(ref_ids)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_operation_sequences(oracle, seed):
    """Seeded random sequences of 120 operations on a jittered 5x5x5 lattice
    of mixed radii (with backups): moves of 1-4 balls by N(0, 0.35^2) A per
    axis (some crossing grid boxes), jumps of one ball to anywhere in the
    box, caller exclusions and inclusions of random balls (also next to balls
    that then move) and restores; after every operation the device
    updateable and the oracle's (moved spheres taken out of their old grid
    boxes) agree on the returned booleans, changed and affected ids, the
    reinit flag and the summary to the parity bar. Moves that make a ball
    hidden are pinned by test_hidden_by_a_move."""
    rng = np.random.default_rng(seed)
    g = np.arange(5) * 3.0
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    balls = np.column_stack([x.ravel(), y.ravel(), z.ravel(), np.full(x.size, 1.6)])
    balls[:, :3] += rng.normal(scale=0.25, size=(len(balls), 3))
    balls[:, 3] = rng.choice([1.5, 1.6, 1.7, 1.8], size=len(balls))
    n = len(balls)
    p = Pair(oracle, with_backup=True, remove_from_old_box=True)
    print(f"\n[gpu] random updateable sequence, seed {seed}", flush=True)
    assert p.init(balls, 1.4)
    cur = balls.copy()
    for it in range(120):
        op = rng.choice(["move", "move", "move", "jump", "exclude", "include", "restore"])
        if op == "move":
            ids = rng.choice(n, rng.integers(1, 5), replace=False)
            cur[ids, :3] += rng.normal(scale=0.35, size=(len(ids), 3))
            p.update(cur, ids)
        elif op == "jump":
            i = int(rng.integers(n))
            cur[i, :3] = rng.uniform(-1.0, 13.0, size=3)
            p.update(cur, [i])
        elif op == "exclude":
            p.op("set_exclusion", int(rng.integers(n)), True)
        elif op == "include":
            p.op("set_exclusion", int(rng.integers(n)), False)
        else:
            p.op("restore")  # later moves are taken relative to the restored state
        if it % 40 == 39:
            print(f"[gpu] seed {seed}: {it + 1} operations agree", flush=True)


def test_hidden_by_a_move(oracle):
    """A ball moved exactly onto another with a higher index hides that one
    (equal spheres: the higher index is hidden, spheres_searcher.rs:295-303).
    The moved ball's collision list is then empty, so the hidden ball's other
    neighbors are not marked affected (spheres_container.rs:105-182) and the
    reference keeps their stored contacts with the hidden ball
    (updateable.rs:476-514): those in the neighbor's row (pairs (x, h), x < h)
    and those in the hidden ball's own row (pairs (h, y), y > h). This
    library, which lists every ball's neighbors afresh in each step, keeps the
    first kind (the pair stays in x's list and keeps its record) and drops
    the second (h's list is empty). Pinned here: the reference's extra
    contacts are exactly pairs (h, y) of the hidden ball's own row, every
    other contact and every id list agrees, and a full recompute has neither
    kind."""
    g = np.arange(5) * 3.0
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    balls = np.column_stack([x.ravel(), y.ravel(), z.ravel(), np.full(x.size, 1.6)])
    balls[:, :3] += np.random.default_rng(9).normal(scale=0.25, size=(len(balls), 3))
    from voronota_ltr import _native

    o = oracle.Updateable(remove_from_old_box=True)
    u = _native.Updateable()
    assert o.init(balls, 1.4) and u.init(balls, 1.4)
    mover, h = 0, 62  # 62 is interior, 0 a corner far from it
    cur = balls.copy()
    cur[mover] = cur[h]
    print("\n[gpu] ball 0 moved onto ball 62", flush=True)
    assert o.update_with_changed(cur, [mover]) and u.update_with_changed(cur, [mover])
    assert [int(i) for i in u.changed_ids()] == [int(i) for i in o.changed_ids()]
    assert [int(i) for i in u.affected_ids()] == [int(i) for i in o.affected_ids()]
    rs, gs = o.summary(), u.summary()
    ref = {(int(a), int(b)): k for k, (a, b) in enumerate(rs.contact_ids)}
    gpu = {(int(c["id_a"]), int(c["id_b"])): c for c in gs.contacts}
    full = {(int(a), int(b)) for a, b in oracle.compute_tessellation(cur, 1.4).contact_ids}
    extra_ref = set(ref) - set(gpu)
    assert set(gpu) <= set(ref)
    assert extra_ref and all(a == h and b > h for a, b in extra_ref), sorted(extra_ref)
    assert not (extra_ref & full)
    assert any(b == h and a < h for a, b in set(ref) & set(gpu) - full)  # the first kind, kept by both
    R = cur[:, 3] + 1.4
    for key in set(gpu):
        rm = max(R[key[0]], R[key[1]])
        _close(gpu[key]["area"], rs.contact_area[ref[key]], math.pi * rm * rm, f"area {key}")
