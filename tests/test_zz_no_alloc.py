"""The allocation-free entry point (vltr_compute_tessellation_device_noalloc)
against vltr_compute_tessellation_device through the C ABI: identical
contacts and cells, full and contacts-only (with groups), on two calls in a
row, the second with a smaller input than the first; and its error paths
(a run beyond the workspace, too small output capacities)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PROBE = 1.4
VLTR_ERR_INVALID_ARGUMENT = 1  # include/voronota_ltr.h


def _download(ptr, count, dtype):
    from voronota_ltr import _hip

    out = np.zeros(count, dtype=dtype)
    if count:
        assert _hip._load().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes, _hip.D2H) == 0
    return out


def _reference(lib, ctx, dballs, n, dgroups, contacts_only):
    from voronota_ltr import _native

    gp, ng = (dgroups.ptr, n) if dgroups is not None else (None, 0)
    _native.check(lib.vltr_compute_tessellation_device(ctx, dballs.ptr, n, PROBE, gp, ng, int(contacts_only)),
                  "vltr_compute_tessellation_device")
    dc, nc, dl, nl = C.c_void_p(), C.c_size_t(), C.c_void_p(), C.c_size_t()
    _native.check(lib.vltr_context_device_outputs(ctx, C.byref(dc), C.byref(nc), C.byref(dl), C.byref(nl)), "outputs")
    return (_download(dc.value, nc.value, _native.CONTACT_DTYPE), _download(dl.value, nl.value, _native.CELL_DTYPE))


def _noalloc(lib, ctx, dballs, n, dgroups, contacts_only, dcon, ccap, dcel, lcap):
    from voronota_ltr import _native

    gp, ng = (dgroups.ptr, n) if dgroups is not None else (None, 0)
    nc, nl = C.c_size_t(), C.c_size_t()
    rc = lib.vltr_compute_tessellation_device_noalloc(ctx, dballs.ptr, n, PROBE, gp, ng, int(contacts_only),
                                                      dcon.ptr, ccap, dcel.ptr, lcap, C.byref(nc), C.byref(nl))
    return rc, nc.value, nl.value


@pytest.mark.parametrize("contacts_only", [False, True])
def test_noalloc_equals_device_entry(contacts_only):
    from voronota_ltr import _hip, _native, workload

    lib = _native.load_library()
    ctx = _native.context(0)
    big, _ = workload.rsa_sphere(12000, seed=12000)
    small, _ = workload.rsa_sphere(3000, seed=3000)
    ccap, lcap = 200000, 12000
    dcon = _hip.DeviceBuffer(ccap * _native.CONTACT_DTYPE.itemsize)
    dcel = _hip.DeviceBuffer(lcap * _native.CELL_DTYPE.itemsize)
    try:
        for name, balls in (("12000", big), ("3000", small)):
            n = len(balls)
            print(f"\n[gpu] noalloc {name} balls, contacts_only={contacts_only}", flush=True)
            db = _hip.DeviceBuffer(balls.nbytes)
            groups = (balls[:, 0] < 0.0).astype(np.int32)
            dg = _hip.DeviceBuffer(groups.nbytes)
            try:
                db.upload(np.ascontiguousarray(balls))
                dg.upload(groups)
                for g in ((None, dg) if contacts_only else (None,)):
                    ref_c, ref_l = _reference(lib, ctx, db, n, g, contacts_only)
                    rc, nc, nl = _noalloc(lib, ctx, db, n, g, contacts_only, dcon, ccap, dcel, lcap)
                    assert rc == 0, _native.last_error()
                    assert nc == len(ref_c) and nl == len(ref_l)
                    assert np.array_equal(_download(dcon.ptr.value, nc, _native.CONTACT_DTYPE), ref_c)
                    assert np.array_equal(_download(dcel.ptr.value, nl, _native.CELL_DTYPE), ref_l)
                    assert nc > 0 and (contacts_only or nl == n)
            finally:
                db.free()
                dg.free()
    finally:
        dcon.free()
        dcel.free()


def test_noalloc_limits():
    """Beyond the workspace, and too small output capacities: clear errors,
    no outputs written beyond the given capacities."""
    from voronota_ltr import _hip, _native, workload

    lib = _native.load_library()
    ctx = _native.context(0)
    balls, _ = workload.rsa_sphere(2000, seed=2000)
    db = _hip.DeviceBuffer(balls.nbytes)
    dcon = _hip.DeviceBuffer(100 * _native.CONTACT_DTYPE.itemsize)
    dcel = _hip.DeviceBuffer(2000 * _native.CELL_DTYPE.itemsize)
    try:
        db.upload(np.ascontiguousarray(balls))
        print("\n[gpu] noalloc: contacts beyond contacts_cap", flush=True)
        rc, nc, _ = _noalloc(lib, ctx, db, 2000, None, False, dcon, 100, dcel, 2000)
        assert rc == VLTR_ERR_INVALID_ARGUMENT and nc > 100
        assert "contacts_cap" in _native.last_error()
        ref_c, _ = _reference(lib, ctx, db, 2000, None, False)
        assert np.array_equal(_download(dcon.ptr.value, 100, _native.CONTACT_DTYPE), ref_c[:100])
        print("[gpu] noalloc: cells_cap below the ball count", flush=True)
        rc, _, _ = _noalloc(lib, ctx, db, 2000, None, False, dcon, 100, dcel, 1999)
        assert rc == VLTR_ERR_INVALID_ARGUMENT and "cells_cap" in _native.last_error()
    finally:
        db.free()
        dcon.free()
        dcel.free()
    big, _ = workload.rsa_sphere(60000, seed=60000)
    db = _hip.DeviceBuffer(big.nbytes)
    dcon = _hip.DeviceBuffer(_native.CONTACT_DTYPE.itemsize)
    dcel = _hip.DeviceBuffer(60000 * _native.CELL_DTYPE.itemsize)
    try:
        db.upload(np.ascontiguousarray(big))
        print("[gpu] noalloc: zero contacts_cap", flush=True)
        rc, _, _ = _noalloc(lib, ctx, db, 60000, None, False, dcon, 0, dcel, 60000)
        assert rc == VLTR_ERR_INVALID_ARGUMENT and "contacts_cap" in _native.last_error()
        print("[gpu] noalloc: 60000 balls, beyond the workspace", flush=True)
        rc, _, _ = _noalloc(lib, ctx, db, 60000, None, False, dcon, 1, dcel, 60000)
        assert rc == VLTR_ERR_INVALID_ARGUMENT and "MiB workspace" in _native.last_error()
    finally:
        db.free()
        dcon.free()
        dcel.free()
