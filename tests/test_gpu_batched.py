"""Batched runs (many structures in one pipeline pass) through the C ABI.

Every structure of a batch must give exactly what a call on that structure
alone gives: the same contact pair sequence, flags and cell indices, and
bit-identical values (the same arithmetic runs on the same inputs), and so
the same parity with the oracle as single calls. The batch mixes the
reference's fixtures, seeded RSA clusters of different sizes, one ball and an
empty structure.
"""
import numpy as np
import pytest

from conftest import CENTRAL8, RING17, load_xyzr
from parity import compare

pytestmark = pytest.mark.gpu


def _structures():
    from voronota_ltr import workload

    return [
        ("cs_1x1", load_xyzr("balls_cs_1x1.xyzr"), None),
        ("ring17", RING17, None),
        ("empty", np.zeros((0, 4)), None),
        ("central8", CENTRAL8, None),
        ("one ball", np.array([[1.0, 2.0, 3.0, 1.5]]), None),
        ("1ctf", load_xyzr("1ctf_pdb.xyzr"), None),
        ("rsa 2500", workload.rsa_sphere(2500, seed=2500)[0], None),
        ("rsa 700", workload.rsa_sphere(700, seed=7)[0], None),
        ("2zsk", load_xyzr("balls_2zsk.xyzr"), None),
    ]


def _same(part, single, check_cells=True):
    assert np.array_equal(part.contacts, single.contacts)
    if check_cells:
        assert np.array_equal(part.cells, single.cells)


@pytest.mark.parametrize("contacts_only", [False, True])
def test_batched_equals_single_calls(oracle, contacts_only):
    from voronota_ltr import _native

    probe = 1.4
    structs = _structures()
    print(f"\n[gpu] batched: {len(structs)} structures, contacts_only={contacts_only}", flush=True)
    parts = _native.run_batched([s for _, s, _ in structs], probe, contacts_only=contacts_only)
    assert len(parts) == len(structs)
    for (name, balls, _), part in zip(structs, parts):
        print(f"[gpu] batched part {name}: n={len(balls)}", flush=True)
        single = _native.run(balls, probe, contacts_only=contacts_only)
        _same(part, single, check_cells=not contacts_only)
        if len(balls) == 0:
            assert len(part.contacts) == 0 and len(part.cells) == 0
            continue
        if contacts_only:
            ref = oracle.compute_contacts_only(balls, probe)
            compare(part, ref, balls, probe, check_cells=False)
        else:
            ref = oracle.compute_tessellation(balls, probe)
            compare(part, ref, balls, probe)


def test_batched_device_offsets():
    """Device-resident batched entry: the per-structure offsets partition the
    outputs as the host entry's parts do."""
    import ctypes as C

    from voronota_ltr import _hip, _native, workload

    structs = [workload.rsa_sphere(n, seed=n)[0] for n in (300, 1200, 50, 2000)]
    parts = _native.run_batched(structs, 1.4)
    lib = _native.load_library()
    ctx = _native.context(0)
    allb = np.ascontiguousarray(np.concatenate(structs))
    off = _native.offsets_of(structs)
    buf = _hip.DeviceBuffer(allb.nbytes)
    buf.upload(allb)
    _native.check(lib.vltr_compute_tessellation_batched_device(ctx, buf.ptr, off.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                               len(structs), 1.4, 0), "batched device")
    coff = np.zeros(len(structs) + 1, np.uint64)
    loff = np.zeros(len(structs) + 1, np.uint64)
    _native.check(lib.vltr_context_batch_offsets(ctx, len(structs), coff.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                 loff.ctypes.data_as(C.POINTER(C.c_uint64))), "batch offsets")
    assert list(np.diff(coff)) == [len(p.contacts) for p in parts]
    assert list(np.diff(loff)) == [len(p.cells) for p in parts]


def test_clip_variants_give_identical_records(oracle):
    """Every launch shape of the clip kernel gives the same records (the
    autotuner only picks among equals); the autotuner keeps one of them set
    on the context."""
    from voronota_ltr import _native, workload

    names = _native.clip_variants()
    assert len(names) >= 2 and names[0] == "g8w1o4nq"
    structs = [s for _, s, _ in _structures()] + [workload.rsa_sphere(20000, seed=20000)[0]]
    before = _native.clip_variant()
    try:
        ref_parts = None
        for v, name in enumerate(names):
            print(f"\n[gpu] clip variant {name}", flush=True)
            _native.set_clip_variant(v)
            parts = _native.run_batched(structs, 1.4)
            plain = _native.run(structs[-1], 1.4, pbox=[100, 0, 0, 0, 100, 0, 0, 0, 100])
            if ref_parts is None:
                ref_parts, ref_plain = parts, plain
                for part, balls in zip(parts[-2:], structs[-2:]):
                    compare(part, oracle.compute_tessellation(balls, 1.4), balls, 1.4)
                continue
            for a, b in zip(parts, ref_parts):
                _same(a, b)
            _same(plain, ref_plain)
        print("[gpu] autotune", flush=True)
        chosen, ms = _native.autotune(structs, 1.4, repeats=1)
        assert chosen in names and set(ms) == set(names) and all(t > 0 for t in ms.values())
        assert _native.clip_variant() == names.index(chosen)
        with pytest.raises(_native.VltrError, match="clip variant out of range"):
            _native.set_clip_variant(len(names))
    finally:
        _native.set_clip_variant(before)


def _device_outputs(lib, ctx, n_structs):
    import ctypes as C

    from voronota_ltr import _hip, _native

    dc, nc, dl, nl = C.c_void_p(), C.c_size_t(), C.c_void_p(), C.c_size_t()
    _native.check(lib.vltr_context_device_outputs(ctx, C.byref(dc), C.byref(nc), C.byref(dl), C.byref(nl)), "outputs")
    con = np.zeros(nc.value, dtype=_native.CONTACT_DTYPE)
    cel = np.zeros(nl.value, dtype=_native.CELL_DTYPE)
    h = _hip._load()
    if nc.value:
        assert h.hipMemcpy(con.ctypes.data_as(C.c_void_p), dc, con.nbytes, _hip.D2H) == 0
    if nl.value:
        assert h.hipMemcpy(cel.ctypes.data_as(C.c_void_p), dl, cel.nbytes, _hip.D2H) == 0
    coff = np.zeros(n_structs + 1, np.uint64)
    loff = np.zeros(n_structs + 1, np.uint64)
    _native.check(lib.vltr_context_batch_offsets(ctx, n_structs, coff.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                 loff.ctypes.data_as(C.POINTER(C.c_uint64))), "batch offsets")
    return con, cel, coff, loff


def test_graph_replays():
    """The batched device run captured in a HIP graph: replays give exactly
    the direct run's outputs, follow new ball values at the same address, and
    a replay that finds more pairs than the capture holds falls back to a
    direct run (then captures again)."""
    import ctypes as C

    from voronota_ltr import _hip, _native, workload

    lib = _native.load_library()
    ctx = _native.context(0)
    sizes = (300, 1200, 50, 2000, 5000)

    def batch(seed, scale=1.0):
        structs = [workload.rsa_sphere(n, seed=seed + n)[0] for n in sizes]
        for s in structs:
            s[:, :3] *= scale
        return np.ascontiguousarray(np.concatenate(structs))

    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    offp = off.ctypes.data_as(C.POINTER(C.c_uint64))
    first = batch(1)
    buf = _hip.DeviceBuffer(first.nbytes)

    def run(balls, graphs, contacts_only=0):
        buf.upload(balls)
        _native.set_graphs(graphs)
        _native.check(lib.vltr_compute_tessellation_batched_device(ctx, buf.ptr, offp, len(sizes), 1.4,
                                                                   contacts_only), "batched")
        return _device_outputs(lib, ctx, len(sizes))

    def same(a, b):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)

    try:
        for label, balls in (("first", first), ("second", batch(2)), ("denser", batch(3, scale=0.8))):
            print(f"\n[gpu] graph replays: {label}", flush=True)
            ref = run(balls, False)
            assert len(ref[0]) > 0
            r0 = _native.graph_replays()
            for k in range(3):
                same(run(balls, True), ref)
            # a layout seen for the first time runs directly; it is captured
            # when it comes again
            assert _native.graph_replays() - r0 >= 2
        # captured on a sparse batch, replayed on a denser one at the same
        # address: about twice the pairs, past the capture's capacity
        sparse, dense = batch(4), batch(3, scale=0.8)
        ref_sparse, ref_dense = run(sparse, False), run(dense, False)
        same(run(sparse, True), ref_sparse)
        r0 = _native.graph_replays()
        same(run(dense, True), ref_dense)  # overflow: direct run, no replay
        assert _native.graph_replays() == r0
        same(run(dense, True), ref_dense)  # captured again, replayed
        assert _native.graph_replays() == r0 + 1
        # contacts only: no cells, its own capture
        ref_co = run(first, False, contacts_only=1)
        assert len(ref_co[0]) > 0 and len(ref_co[1]) == 0
        for _ in range(3):
            same(run(first, True, contacts_only=1), ref_co)
        assert _native.graph_replays() == r0 + 3  # new layout: direct, then captured and replayed twice
    finally:
        _native.set_graphs(False)
        buf.free()
