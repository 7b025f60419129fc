"""Direct runs that enqueue the back half without waiting for the pair total
(a call repeating the last plain call's shape): their outputs are the bytes a
synchronized run of the same balls gives, also when the pair total outgrows
the capacity taken from the last run (the run is then redone, counted as a
miss)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PROBE = 1.4


def _run(lib, ctx, dbuf, n, contacts_only=0):
    from voronota_ltr import _hip, _native

    _native.check(lib.vltr_compute_tessellation_device(ctx, dbuf.ptr, n, PROBE, None, 0, contacts_only), "device run")
    dc, nc, dl, nl = C.c_void_p(), C.c_size_t(), C.c_void_p(), C.c_size_t()
    _native.check(lib.vltr_context_device_outputs(ctx, C.byref(dc), C.byref(nc), C.byref(dl), C.byref(nl)), "outputs")
    con = np.zeros(nc.value, dtype=_native.CONTACT_DTYPE)
    cel = np.zeros(nl.value, dtype=_native.CELL_DTYPE)
    h = _hip._load()
    if nc.value:
        assert h.hipMemcpy(con.ctypes.data_as(C.c_void_p), dc, con.nbytes, _hip.D2H) == 0
    if nl.value:
        assert h.hipMemcpy(cel.ctypes.data_as(C.c_void_p), dl, cel.nbytes, _hip.D2H) == 0
    return con.tobytes(), cel.tobytes(), int(nc.value)


@pytest.mark.parametrize("contacts_only", [0, 1])
def test_speculative_runs_match_synchronized(contacts_only):
    from voronota_ltr import _hip, _native, workload

    n = 6000
    dense, _ = workload.rsa_sphere(n, seed=6000)
    sparse = dense.copy()
    sparse[:, :3] *= 1.6  # the same balls spread out: far fewer pairs
    lib = _native.load_library()
    ctx = _native.context(0)
    db = _hip.DeviceBuffer(dense.nbytes)
    other = _hip.DeviceBuffer(dense.nbytes)
    other.upload(np.ascontiguousarray(dense))

    def synchronized(balls):
        # a call of another shape first, so that this one waits for the total
        _run(lib, ctx, other, n - 1, contacts_only)
        db.upload(np.ascontiguousarray(balls))
        runs0, _ = _native.speculative_runs()
        out = _run(lib, ctx, db, n, contacts_only)
        assert _native.speculative_runs()[0] == runs0
        return out

    print(f"\n[gpu] speculative runs, {n} balls, contacts_only={contacts_only}", flush=True)
    ref_dense = synchronized(dense)
    ref_sparse = synchronized(sparse)
    assert ref_dense[2] > ref_sparse[2] > 0

    # dense again: repeats the shape, capacity from the last (sparse) run is
    # too small: redone
    db.upload(np.ascontiguousarray(dense))
    runs0, miss0 = _native.speculative_runs()
    assert _run(lib, ctx, db, n, contacts_only) == ref_dense
    runs1, miss1 = _native.speculative_runs()
    assert (runs1 - runs0, miss1 - miss0) == (1, 1)
    # and again: capacity from the dense run, no miss
    assert _run(lib, ctx, db, n, contacts_only) == ref_dense
    runs2, miss2 = _native.speculative_runs()
    assert (runs2 - runs1, miss2 - miss1) == (1, 0)
    # sparse balls under the dense capacity: no miss
    db.upload(np.ascontiguousarray(sparse))
    assert _run(lib, ctx, db, n, contacts_only) == ref_sparse
    runs3, miss3 = _native.speculative_runs()
    assert (runs3 - runs2, miss3 - miss2) == (1, 0)
