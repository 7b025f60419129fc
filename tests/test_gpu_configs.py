"""The HIP path against the oracle on the benchmark workloads at their own
sizes (the inputs bench.py times):

* 100,000 RSA balls in a sphere (seed 100000, probe 1.4): full
  compute_tessellation and compute_contacts_only, the latter also with group
  1 for x < 0 and 0 otherwise (only contacts between the halves kept),
  through the host entry points and the device-resident one;
* 50,000 RSA balls in a periodic cube at 0.056 balls/A^3: 20 steps of
  update_with_changed, 500 seeded balls moved by N(0, 0.05^2) A per axis and
  wrapped into the box, with the contact energy switched on; every step's
  changed/affected ids, reinit flag and summary against the oracle's
  Updateable, the energy against the oracle's and the gradient against a
  standalone forward + adjoint of the same balls;
* 4,096 structures of 200-20,000 balls (log-uniform, 18.2 M balls): every
  structure of one batched call bit-identical to its own single call, and 64
  of them (the 8 largest, the 8 smallest and 48 seeded) against the oracle.

The bar is parity.py's: pair sequences, flags and indices exact, values to
1e-9 relative with the full-sphere / full-disk floors.
"""
import ctypes as C
import math

import numpy as np
import pytest

from parity import FLOOR, REL, compare

pytestmark = pytest.mark.gpu

PROBE = 1.4


def _device_run(lib, ctx, dbuf, n, gbuf=None, contacts_only=0):
    """vltr_compute_tessellation_device on device balls; outputs downloaded."""
    from voronota_ltr import _hip, _native

    gp, ng = (gbuf.ptr, n) if gbuf is not None else (None, 0)
    _native.check(lib.vltr_compute_tessellation_device(ctx, dbuf.ptr, n, PROBE, gp, ng, contacts_only), "device run")
    dc, nc, dl, nl = C.c_void_p(), C.c_size_t(), C.c_void_p(), C.c_size_t()
    _native.check(lib.vltr_context_device_outputs(ctx, C.byref(dc), C.byref(nc), C.byref(dl), C.byref(nl)), "outputs")
    con = np.zeros(nc.value, dtype=_native.CONTACT_DTYPE)
    cel = np.zeros(nl.value, dtype=_native.CELL_DTYPE)
    h = _hip._load()
    if nc.value:
        assert h.hipMemcpy(con.ctypes.data_as(C.c_void_p), dc, con.nbytes, _hip.D2H) == 0
    if nl.value:
        assert h.hipMemcpy(cel.ctypes.data_as(C.c_void_p), dl, cel.nbytes, _hip.D2H) == 0
    return con, cel


def test_plain_100k(oracle):
    """Config: 100,000 balls, full tessellation and contacts only (plain and
    grouped by the sign of x)."""
    import voronota_ltr
    from voronota_ltr import _hip, _native, workload

    n = 100000
    balls, _ = workload.rsa_sphere(n, seed=100000)
    groups = (balls[:, 0] < 0.0).astype(np.int32)
    print(f"\n[gpu] 100k full: n={n}", flush=True)
    gpu = voronota_ltr.tessellate(balls, PROBE)
    counts, ids = _native.neighbor_lists(n)
    off, idx, _, _, _ = oracle.collisions(balls, PROBE)
    assert np.array_equal(counts.astype(np.int64), np.diff(off.astype(np.int64))), "neighbor list lengths"
    assert np.array_equal(ids.astype(np.int64), idx.astype(np.int64)), "neighbor list contents/order"
    ref = oracle.compute_tessellation(balls, PROBE)
    stats = compare(gpu, ref, balls, PROBE)
    print(f"[gpu] 100k full: {stats}", flush=True)
    assert stats["contacts"] > 700000 and stats["cells"] == n

    print("[gpu] 100k contacts only", flush=True)
    gco = voronota_ltr.tessellate(balls, PROBE, contacts_only=True)
    compare(gco, oracle.compute_contacts_only(balls, PROBE), balls, PROBE, check_cells=False)

    print("[gpu] 100k contacts only, groups by the sign of x", flush=True)
    ggr = voronota_ltr.tessellate(balls, PROBE, groups=groups, contacts_only=True)
    rgr = oracle.compute_contacts_only(balls, PROBE, groups=groups)
    st = compare(ggr, rgr, balls, PROBE, check_cells=False)
    print(f"[gpu] 100k grouped: {st}", flush=True)
    ia, ib = ggr.contacts["id_a"].astype(np.int64), ggr.contacts["id_b"].astype(np.int64)
    assert len(ia) > 5000 and np.all(groups[ia] != groups[ib]), "grouped contacts-only kept a same-group contact"

    # the device-resident entry point (the one bench.py times) gives exactly
    # what the host entry points give
    print("[gpu] 100k device-resident entry", flush=True)
    lib = _native.load_library()
    ctx = _native.context(0)
    dbuf = _hip.DeviceBuffer(balls.nbytes)
    gbuf = _hip.DeviceBuffer(groups.nbytes)
    try:
        dbuf.upload(np.ascontiguousarray(balls))
        gbuf.upload(np.ascontiguousarray(groups))
        con, cel = _device_run(lib, ctx, dbuf, n)
        assert np.array_equal(con, gpu.contacts) and np.array_equal(cel, gpu.cells)
        con, cel = _device_run(lib, ctx, dbuf, n, gbuf=gbuf, contacts_only=1)
        assert np.array_equal(con, ggr.contacts) and len(cel) == 0
    finally:
        dbuf.free()
        gbuf.free()


def _sorted_contacts(ids, area, arc, central):
    o = np.lexsort((ids[:, 1], ids[:, 0]))
    return ids[o], area[o], arc[o], central[o]


def _close(gpu, ref, scale, what):
    tol = REL * np.abs(ref) + FLOOR * scale
    bad = np.flatnonzero(np.abs(gpu - ref) > tol)
    assert len(bad) == 0, f"{what}: {len(bad)} differ; first gpu={gpu[bad[0]]!r} ref={ref[bad[0]]!r}"


def _same_summary(gpu, ref, R):
    """Updateable summaries: the same contact set and cell set, values to the
    bar (contacts compared in id order: the reference lists them in per-ball
    append order after incremental updates, this library in pair order)."""
    g_ids = np.stack([gpu.contacts["id_a"], gpu.contacts["id_b"]], axis=1).astype(np.uint64)
    gi, ga, gl, gc = _sorted_contacts(g_ids, gpu.contacts["area"], gpu.contacts["arc_length"],
                                      gpu.contacts["central"].astype(np.uint8))
    ri, ra, rl, rc = _sorted_contacts(ref.contact_ids.astype(np.uint64), ref.contact_area, ref.contact_arc,
                                      ref.contact_central)
    assert gi.shape == ri.shape, f"contact count {len(gi)} != {len(ri)}"
    assert np.array_equal(gi, ri), "contact sets differ"
    assert np.array_equal(gc, rc), "central flags differ"
    rmax = np.maximum(R[ri[:, 0].astype(np.int64)], R[ri[:, 1].astype(np.int64)])
    _close(ga, ra, math.pi * rmax ** 2, "contact area")
    _close(gl, rl, 2 * math.pi * rmax, "arc length")
    assert np.array_equal(gpu.cells["index"].astype(np.uint64), ref.cell_index.astype(np.uint64)), "cell indices"
    Rc = R[ref.cell_index.astype(np.int64)]
    _close(gpu.cells["sas_area"], ref.cell_sas, 4 * math.pi * Rc ** 2, "cell SAS")
    _close(gpu.cells["volume"], ref.cell_volume, 4.0 / 3.0 * math.pi * Rc ** 3, "cell volume")
    return len(gi)


def test_periodic_dynamics_50k(oracle):
    """Config: 50,000 balls in a periodic cube, 20 updateable steps of 500
    moved balls with the contact energy and its gradient in every pass. The
    oracle's Updateable takes moved spheres out of their old grid boxes
    (remove_from_old_box): the reference's searcher keeps a sphere that
    changed box listed in its old box too and then reports duplicate contacts
    (spheres_searcher.rs:144-175), which this library does not reproduce.
    Cells of affected balls follow the reference's incremental rule: they sum
    the deduplicated per-ball contact lists, so of a canonical pair through a
    periodic image both balls take the first record in pair order, where a
    full run gives each ball its own record (the two frames can differ where
    a cut is near-degenerate: at step 7 a buried ball's volume differs by 1e-8
    relative between the rules)."""
    from conftest import corners_box
    from test_gpu_updateable import _same_gradient
    from voronota_ltr import _native, workload

    n = 50000
    L = (n / 0.056) ** (1.0 / 3.0)
    balls, _ = workload.rsa_box(n, L, seed=50000)
    box = corners_box((0, 0, 0), (L, L, L))
    R = balls[:, 3] + PROBE
    types = np.searchsorted(np.array([1.52, 1.55, 1.70, 1.80]), balls[:, 3]).astype(np.int32)
    W = workload.type_matrix(20000)
    rng = np.random.default_rng(50000)
    u = _native.Updateable()
    u.set_energy(types, W)
    o = oracle.Updateable(remove_from_old_box=True)
    print(f"\n[gpu] periodic dynamics: n={n} L={L:.3f}", flush=True)
    assert u.init(balls, PROBE, box) and o.init(balls, PROBE, box)
    _same_summary(u.summary(), o.summary(), R)
    cur = balls.copy()
    reinits = 0
    for step in range(20):
        ch = np.sort(rng.choice(n, 500, replace=False))
        cur[ch, :3] = np.mod(cur[ch, :3] + rng.normal(scale=0.05, size=(500, 3)), L)
        a = u.update_with_changed(cur, ch)
        b = o.update_with_changed(cur, ch)
        assert a and b
        assert np.array_equal(u.changed_ids().astype(np.int64), o.changed_ids().astype(np.int64))
        aff_g, aff_o = u.affected_ids().astype(np.int64), o.affected_ids().astype(np.int64)
        assert np.array_equal(aff_g, aff_o), "affected ids"
        assert u.last_update_was_full_reinit() == o.last_update_was_full_reinit()
        reinits += int(o.last_update_was_full_reinit())
        nc = _same_summary(u.summary(), o.summary(), R)
        e, g = u.energy_gradient()
        e_ref = oracle.energy(cur, PROBE, types, W, pbox=box)
        assert abs(e - e_ref) <= 1e-9 * abs(e_ref), f"energy {e!r} vs oracle {e_ref!r}"
        e2, g2 = _native.energy_gradient(cur, PROBE, types, W, pbox=box)
        assert e == pytest.approx(e2, rel=1e-12)
        _same_gradient(g, g2)  # the updateable's pass vs a standalone forward + adjoint
        print(f"[gpu] dynamics step {step}: affected={len(aff_g)} contacts={nc} energy={e:.6f}", flush=True)
    assert reinits == 0  # 500 moved of 50,000 stays below the reinit threshold


def test_batch_4096(oracle):
    """Config: 4,096 structures in one batched call (device-resident, as
    bench.py runs it): every structure bit-identical to its own
    device-resident single call; 64 structures against the oracle."""
    from concurrent.futures import ThreadPoolExecutor

    from test_gpu_batched import _device_outputs
    from voronota_ltr import _hip, _native, workload

    sizes = workload.batch_sizes(4096, seed=4096)
    with ThreadPoolExecutor(max_workers=8) as ex:
        structs = list(ex.map(lambda k: workload.rsa_sphere(int(sizes[k]), seed=1000000 + k)[0], range(len(sizes))))
    off = _native.offsets_of(structs)
    allb = np.ascontiguousarray(np.concatenate(structs))
    total = int(off[-1])
    print(f"\n[gpu] batch: {len(sizes)} structures, {total} balls", flush=True)
    lib = _native.load_library()
    ctx = _native.context(0)
    dbuf = _hip.DeviceBuffer(allb.nbytes)
    try:
        dbuf.upload(allb)
        _native.check(lib.vltr_compute_tessellation_batched_device(ctx, dbuf.ptr,
                                                                   off.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                                   len(sizes), PROBE, 0), "batched device")
        pairs = int(_native.last_counts()["pairs"])
        con, cel, coff, loff = _device_outputs(lib, ctx, len(sizes))
        print(f"[gpu] batch: {pairs} pairs, {len(con)} contacts, {len(cel)} cells", flush=True)
        assert len(cel) == total
        base = dbuf.ptr.value
        for k in range(len(sizes)):
            sub = C.c_void_p(base + int(off[k]) * 32)
            nk = int(sizes[k])
            _native.check(lib.vltr_compute_tessellation_device(ctx, sub, nk, PROBE, None, 0, 0), "single")
            dc, nc, dl, nl = C.c_void_p(), C.c_size_t(), C.c_void_p(), C.c_size_t()
            _native.check(lib.vltr_context_device_outputs(ctx, C.byref(dc), C.byref(nc), C.byref(dl), C.byref(nl)),
                          "outputs")
            sc = np.zeros(nc.value, dtype=_native.CONTACT_DTYPE)
            sl = np.zeros(nl.value, dtype=_native.CELL_DTYPE)
            h = _hip._load()
            assert h.hipMemcpy(sc.ctypes.data_as(C.c_void_p), dc, sc.nbytes, _hip.D2H) == 0
            assert h.hipMemcpy(sl.ctypes.data_as(C.c_void_p), dl, sl.nbytes, _hip.D2H) == 0
            pc = con[int(coff[k]):int(coff[k + 1])]
            pl = cel[int(loff[k]):int(loff[k + 1])]
            assert np.array_equal(pc, sc), f"structure {k} (n={nk}): contacts differ from its single call"
            assert np.array_equal(pl, sl), f"structure {k} (n={nk}): cells differ from its single call"
            if k % 512 == 0:
                print(f"[gpu] batch: structures 0..{k} equal their single calls", flush=True)
    finally:
        dbuf.free()

    order = np.argsort(sizes, kind="stable")
    rng = np.random.default_rng(64)
    rest = order[8:-8]
    picked = list(order[-8:]) + list(order[:8]) + list(rng.choice(rest, 48, replace=False))
    for k in picked:
        k = int(k)
        part = _native.BatchPart(int(sizes[k]), con[int(coff[k]):int(coff[k + 1])], cel[int(loff[k]):int(loff[k + 1])])
        ref = oracle.compute_tessellation(structs[k], PROBE)
        compare(part, ref, structs[k], PROBE)
    print(f"[gpu] batch: {len(picked)} structures match the oracle (sizes {int(sizes[order[0]])}.."
          f"{int(sizes[order[-1]])})", flush=True)
