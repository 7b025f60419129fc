"""ctypes binding of the C ABI in include/voronota_ltr.h (libvltr_hip.so).

The HIP library is the only compute path: if it is missing or no GPU is
visible, every call raises. ctypes releases the GIL around each foreign call,
as the reference's PyO3 module does around compute_tessellation
(src/python.rs:253).
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvltr_hip.so")

VLTR_OK = 0
NULL_ID = np.iinfo(np.uint64).max


class VltrBall(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("r", C.c_double)]


class VltrPeriodicBox(C.Structure):
    _fields_ = [("shift_a", C.c_double * 3), ("shift_b", C.c_double * 3), ("shift_c", C.c_double * 3)]


# numpy views of the ABI structs
CONTACT_DTYPE = np.dtype([("id_a", "<u8"), ("id_b", "<u8"), ("area", "<f8"), ("arc_length", "<f8"),
                          ("central", "u1"), ("reserved", "u1", (7,))])
CELL_DTYPE = np.dtype([("index", "<u8"), ("sas_area", "<f8"), ("volume", "<f8")])
VERTEX_DTYPE = np.dtype([("ball_indices", "<u8", (4,)), ("x", "<f8"), ("y", "<f8"), ("z", "<f8")])
EDGE_DTYPE = np.dtype([("ball_indices", "<u8", (3,)), ("length", "<f8")])
SOLVENT_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("radius", "<f8"), ("weight", "<f8"),
                          ("parent_index", "<u8")])
assert CONTACT_DTYPE.itemsize == 40 and CELL_DTYPE.itemsize == 24

_vp = C.c_void_p
_sz = C.c_size_t
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)

# every symbol include/voronota_ltr.h declares, with (restype, argtypes)
SIGNATURES = {
    "vltr_context_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "vltr_context_destroy": (None, [_vp]),
    "vltr_last_error": (C.c_char_p, []),
    "vltr_context_last_timings": (C.c_int, [_vp, _f64p, C.c_int]),
    "vltr_context_last_counts": (C.c_int, [_vp, _u64p, C.c_int]),
    "vltr_periodic_box_from_corners": (None, [_f64p, _f64p, C.POINTER(VltrPeriodicBox)]),
    "vltr_periodic_box_from_vectors": (None, [_f64p, _f64p, _f64p, C.POINTER(VltrPeriodicBox)]),
    "vltr_compute_tessellation": (C.c_int, [_vp, _vp, _sz, C.c_double, C.POINTER(VltrPeriodicBox), _i32p, _sz,
                                            C.c_int, C.POINTER(_vp)]),
    "vltr_compute_contacts_only": (C.c_int, [_vp, _vp, _sz, C.c_double, C.POINTER(VltrPeriodicBox), _i32p, _sz,
                                             C.POINTER(_vp)]),
    "vltr_compute_tessellation_device": (C.c_int, [_vp, _vp, _sz, C.c_double, _vp, _sz, C.c_int]),
    "vltr_context_device_outputs": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_sz), C.POINTER(_vp), C.POINTER(_sz)]),
    "vltr_compute_tessellation_device_noalloc": (C.c_int, [_vp, _vp, _sz, C.c_double, _vp, _sz, C.c_int, _vp, _sz,
                                                           _vp, _sz, C.POINTER(_sz), C.POINTER(_sz)]),
    "vltr_context_neighbor_lists": (C.c_int64, [_vp, _sz, _i32p, _i32p, _sz]),
    "vltr_result_num_balls": (_sz, [_vp]),
    "vltr_result_num_contacts": (_sz, [_vp]),
    "vltr_result_contacts": (_vp, [_vp]),
    "vltr_result_num_cells": (_sz, [_vp]),
    "vltr_result_cells": (_vp, [_vp]),
    "vltr_result_has_cell_vertices": (C.c_int, [_vp]),
    "vltr_result_num_cell_vertices": (_sz, [_vp]),
    "vltr_result_cell_vertices": (_vp, [_vp]),
    "vltr_result_num_cell_edges": (_sz, [_vp]),
    "vltr_result_cell_edges": (_vp, [_vp]),
    "vltr_result_total_sas_area": (C.c_double, [_vp]),
    "vltr_result_total_volume": (C.c_double, [_vp]),
    "vltr_result_total_contact_area": (C.c_double, [_vp]),
    "vltr_result_sas_areas": (None, [_vp, _f64p]),
    "vltr_result_volumes": (None, [_vp, _f64p]),
    "vltr_result_free": (None, [_vp]),
    "vltr_compute_tessellation_batched": (C.c_int, [_vp, _vp, _u64p, _sz, C.c_double, C.c_int, C.POINTER(_vp)]),
    "vltr_batch_num_structs": (_sz, [_vp]),
    "vltr_batch_contacts": (_vp, [_vp, _sz, C.POINTER(_sz)]),
    "vltr_batch_cells": (_vp, [_vp, _sz, C.POINTER(_sz)]),
    "vltr_batch_free": (None, [_vp]),
    "vltr_compute_tessellation_batched_device": (C.c_int, [_vp, _vp, _u64p, _sz, C.c_double, C.c_int]),
    "vltr_context_set_graphs": (C.c_int, [_vp, C.c_int]),
    "vltr_context_graph_replays": (C.c_uint64, [_vp]),
    "vltr_context_speculative_runs": (C.c_uint64, [_vp, C.POINTER(C.c_uint64)]),
    "vltr_clip_variant_count": (C.c_int, []),
    "vltr_clip_variant_name": (C.c_char_p, [C.c_int]),
    "vltr_context_set_clip_variant": (C.c_int, [_vp, C.c_int]),
    "vltr_context_clip_variant": (C.c_int, [_vp]),
    "vltr_context_autotune": (C.c_int, [_vp, _vp, _u64p, _sz, C.c_double, C.c_int, C.c_int, _f64p, C.c_int,
                                        C.POINTER(C.c_int)]),
    "vltr_context_batch_offsets": (C.c_int, [_vp, _sz, _u64p, _u64p]),
    "vltr_energy_gradient": (C.c_int, [_vp, _vp, _sz, C.c_double, C.POINTER(VltrPeriodicBox), _i32p, _f64p,
                                       C.c_double, C.c_double, _f64p, _f64p]),
    "vltr_energy_gradient_device": (C.c_int, [_vp, _vp, _sz, C.c_double, C.POINTER(VltrPeriodicBox), _vp, _f64p,
                                              C.c_double, C.c_double, _vp, _f64p]),
    "vltr_radii_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "vltr_radii_load_text": (C.c_int, [_vp, C.c_char_p]),
    "vltr_radii_get": (C.c_double, [_vp, C.c_char_p, C.c_char_p]),
    "vltr_radii_free": (None, [_vp]),
    "vltr_parse_file": (C.c_int, [C.c_char_p, _vp, _vp, C.POINTER(_vp)]),
    "vltr_parse_text": (C.c_int, [C.c_char_p, _sz, C.c_int, _vp, _vp, C.POINTER(_vp)]),
    "vltr_structure_format": (C.c_int, [_vp]),
    "vltr_structure_num_balls": (_sz, [_vp]),
    "vltr_structure_balls": (_vp, [_vp]),
    "vltr_structure_num_records": (_sz, [_vp]),
    "vltr_record_string": (C.c_char_p, [_vp, _sz, C.c_int]),
    "vltr_record_numbers": (C.c_int, [_vp, _sz, _i32p, _i32p, _f64p]),
    "vltr_structure_free": (None, [_vp]),
    "vltr_grouping_chain": (C.c_int, [_vp, _i32p]),
    "vltr_grouping_residue": (C.c_int, [_vp, _i32p]),
    "vltr_grouping_selections": (C.c_int, [_vp, C.POINTER(C.c_char_p), _sz, _i32p]),
    "vltr_selection_match": (C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_uint8)]),
    "vltr_compute_tessellation_from_file": (C.c_int, [_vp, C.c_char_p, C.c_double, C.POINTER(VltrPeriodicBox),
                                                      C.c_int, C.POINTER(C.c_char_p), _sz, C.POINTER(_vp)]),
    "vltr_result_to_json": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_sz)]),
    "vltr_pymol_graphics": (C.c_int, [_vp, _vp, _sz, C.c_double, _i32p, _sz, C.c_double, C.c_char_p,
                                      C.POINTER(_vp), C.POINTER(_sz)]),
    "vltr_string_free": (None, [_vp]),
    "vltr_compute_solvent_spheres": (C.c_int, [_vp, _vp, _sz, C.c_double, C.c_double, C.c_int,
                                               C.POINTER(VltrPeriodicBox), C.POINTER(_vp)]),
    "vltr_solvent_count": (_sz, [_vp]),
    "vltr_solvent_spheres": (_vp, [_vp]),
    "vltr_solvent_free": (None, [_vp]),
    "vltr_updateable_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(_vp)]),
    "vltr_updateable_destroy": (None, [_vp]),
    "vltr_updateable_init": (C.c_int, [_vp, _vp, _sz, C.c_double, C.POINTER(VltrPeriodicBox), C.POINTER(C.c_int)]),
    "vltr_updateable_update": (C.c_int, [_vp, _vp, _sz, C.POINTER(C.c_int)]),
    "vltr_updateable_update_with_changed": (C.c_int, [_vp, _vp, _sz, _u64p, _sz, C.POINTER(C.c_int)]),
    "vltr_updateable_set_exclusion": (C.c_int, [_vp, _sz, C.c_int, C.POINTER(C.c_int)]),
    "vltr_updateable_restore": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "vltr_updateable_summary": (C.c_int, [_vp, C.POINTER(_vp)]),
    "vltr_updateable_changed_ids": (_sz, [_vp, _u64p, _sz]),
    "vltr_updateable_affected_ids": (_sz, [_vp, _u64p, _sz]),
    "vltr_updateable_last_update_was_full_reinit": (C.c_int, [_vp]),
    "vltr_updateable_backup_enabled": (C.c_int, [_vp]),
    "vltr_updateable_last_timings": (C.c_int, [_vp, _f64p, C.c_int]),
    "vltr_updateable_set_energy": (C.c_int, [_vp, _i32p, _sz, _f64p, C.c_double, C.c_double]),
    "vltr_updateable_energy_gradient": (C.c_int, [_vp, _f64p, _f64p]),
    "vltr_updateable_energy_gradient_device": (C.c_int, [_vp, _f64p, C.POINTER(_vp)]),
}

_lock = threading.Lock()
_lib = None
_contexts: dict[int, int] = {}


class VltrError(RuntimeError):
    pass


def load_library():
    """Loads libvltr_hip.so and binds its signatures (no GPU needed)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise VltrError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
            lib = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def last_error() -> str:
    msg = load_library().vltr_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str):
    if rc != VLTR_OK:
        raise VltrError(f"{what} failed (status {rc}): {last_error()}")


def context(device: int = 0) -> int:
    """Per-device library context, created on first use."""
    lib = load_library()
    with _lock:
        h = _contexts.get(device)
        if h is None:
            out = C.c_void_p()
            check(lib.vltr_context_create(device, C.byref(out)), "vltr_context_create")
            h = out.value
            _contexts[device] = h
    return h


def last_timings(device: int = 0):
    lib = load_library()
    buf = (C.c_double * 7)()
    k = lib.vltr_context_last_timings(context(device), buf, 7)
    names = ["grid", "neighbors", "pairs", "clip", "slow", "cells", "total"]
    return {names[i]: buf[i] for i in range(k)}


def last_counts(device: int = 0):
    lib = load_library()
    buf = (C.c_uint64 * 8)()
    k = lib.vltr_context_last_counts(context(device), buf, 8)
    names = ["balls", "spheres", "pairs", "contacts", "cells", "slow_pairs", "nb_cap", "buckets"]
    return {names[i]: int(buf[i]) for i in range(k)}


def neighbor_lists(num_balls: int, device: int = 0):
    """(counts, ids) of the last pipeline run on the device's context."""
    lib = load_library()
    ctx = context(device)
    counts = np.zeros(num_balls, dtype=np.int32)
    total = lib.vltr_context_neighbor_lists(ctx, num_balls, counts.ctypes.data_as(_i32p), None, 0)
    if total < 0:
        raise VltrError(last_error())
    ids = np.zeros(max(total, 1), dtype=np.int32)
    lib.vltr_context_neighbor_lists(ctx, num_balls, counts.ctypes.data_as(_i32p), ids.ctypes.data_as(_i32p), total)
    return counts, ids[:total]


def make_pbox(pbox) -> VltrPeriodicBox | None:
    """Accepts None, a VltrPeriodicBox, or 9 numbers (shift_a | shift_b | shift_c)."""
    if pbox is None:
        return None
    if isinstance(pbox, VltrPeriodicBox):
        return pbox
    v = np.asarray(pbox, dtype=np.float64).reshape(9)
    b = VltrPeriodicBox()
    for k in range(3):
        b.shift_a[k] = v[k]
        b.shift_b[k] = v[3 + k]
        b.shift_c[k] = v[6 + k]
    return b


class Result:
    """Host copy of a vltr_result: numpy structured arrays."""

    def __init__(self, handle):
        lib = load_library()
        self.num_balls = lib.vltr_result_num_balls(handle)
        nc = lib.vltr_result_num_contacts(handle)
        ncell = lib.vltr_result_num_cells(handle)
        self.contacts = _copy(lib.vltr_result_contacts(handle), nc, CONTACT_DTYPE)
        self.cells = _copy(lib.vltr_result_cells(handle), ncell, CELL_DTYPE)
        self.total_sas_area = lib.vltr_result_total_sas_area(handle)
        self.total_volume = lib.vltr_result_total_volume(handle)
        self.total_contact_area = lib.vltr_result_total_contact_area(handle)
        self.cell_vertices = None
        self.cell_edges = None
        if lib.vltr_result_has_cell_vertices(handle):
            self.cell_vertices = _copy(lib.vltr_result_cell_vertices(handle),
                                       lib.vltr_result_num_cell_vertices(handle), VERTEX_DTYPE)
            self.cell_edges = _copy(lib.vltr_result_cell_edges(handle), lib.vltr_result_num_cell_edges(handle),
                                    EDGE_DTYPE)

    def sas_areas(self):
        """Per-ball SAS, NaN where the reference returns None (types.rs:206-213)."""
        out = np.full(self.num_balls, np.nan)
        out[self.cells["index"].astype(np.int64)] = self.cells["sas_area"]
        return out

    def volumes(self):
        out = np.full(self.num_balls, np.nan)
        out[self.cells["index"].astype(np.int64)] = self.cells["volume"]
        return out


def _copy(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    buf = (C.c_char * (n * dtype.itemsize)).from_address(ptr)
    return np.frombuffer(bytes(buf), dtype=dtype).copy()


def _balls_array(balls) -> np.ndarray:
    b = np.ascontiguousarray(np.asarray(balls, dtype=np.float64))
    if b.size == 0:
        return np.zeros((0, 4), dtype=np.float64)
    return b.reshape(-1, 4)


def run(balls, probe, pbox=None, groups=None, with_cell_vertices=False, contacts_only=False, device=0) -> Result:
    lib = load_library()
    ctx = context(device)
    b = _balls_array(balls)
    box = make_pbox(pbox)
    g = None
    if groups is not None:
        g = np.ascontiguousarray(np.asarray(groups, dtype=np.int32))
    handle = C.c_void_p()
    bp = b.ctypes.data_as(_vp) if len(b) else None
    gp = g.ctypes.data_as(_i32p) if g is not None and len(g) else None
    ng = len(g) if g is not None else 0
    boxp = C.byref(box) if box is not None else None
    if contacts_only:
        rc = lib.vltr_compute_contacts_only(ctx, bp, len(b), float(probe), boxp, gp, ng, C.byref(handle))
        check(rc, "vltr_compute_contacts_only")
    else:
        rc = lib.vltr_compute_tessellation(ctx, bp, len(b), float(probe), boxp, gp, ng, int(bool(with_cell_vertices)),
                                           C.byref(handle))
        check(rc, "vltr_compute_tessellation")
    try:
        return Result(handle)
    finally:
        lib.vltr_result_free(handle)


class BatchPart:
    """One structure of a batched run: contacts and cells with local indices."""

    def __init__(self, num_balls, contacts, cells):
        self.num_balls = num_balls
        self.contacts = contacts
        self.cells = cells

    def sas_areas(self):
        out = np.full(self.num_balls, np.nan)
        out[self.cells["index"].astype(np.int64)] = self.cells["sas_area"]
        return out

    def volumes(self):
        out = np.full(self.num_balls, np.nan)
        out[self.cells["index"].astype(np.int64)] = self.cells["volume"]
        return out


def offsets_of(structures):
    sizes = [len(s) for s in structures]
    off = np.zeros(len(sizes) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(sizes)
    return off


def set_graphs(enable, device=0):
    """HIP-graph capture and replay of device-resident batched runs on the
    device's context."""
    lib = load_library()
    check(lib.vltr_context_set_graphs(context(device), int(bool(enable))), "vltr_context_set_graphs")


def graph_replays(device=0):
    return int(load_library().vltr_context_graph_replays(context(device)))


def speculative_runs(device=0):
    """(runs that skipped the synchronisation on the pair total, of which
    redone) on the device's context."""
    misses = C.c_uint64(0)
    runs = load_library().vltr_context_speculative_runs(context(device), C.byref(misses))
    return int(runs), int(misses.value)


def clip_variants():
    """Names of the clip kernel's launch shapes (index = shape)."""
    lib = load_library()
    return [lib.vltr_clip_variant_name(v).decode() for v in range(lib.vltr_clip_variant_count())]


def set_clip_variant(variant, device=0):
    lib = load_library()
    check(lib.vltr_context_set_clip_variant(context(device), int(variant)), "vltr_context_set_clip_variant")


def clip_variant(device=0):
    return int(load_library().vltr_context_clip_variant(context(device)))


def autotune(structures, probe, contacts_only=False, repeats=2, device=0, d_balls=None):
    """Times every clip launch shape on the batch (a list of structures, or
    one structure) and keeps the fastest on the device's context; returns
    (name, {name: median pipeline ms}). d_balls: optional device pointer that
    already holds the concatenated balls."""
    from . import _hip

    lib = load_library()
    ctx = context(device)
    if len(structures) and np.ndim(structures[0]) == 1:
        structures = [structures]
    off = offsets_of(structures)
    buf = None
    if d_balls is None:
        allb = np.ascontiguousarray(np.concatenate([_balls_array(s) for s in structures]))
        buf = _hip.DeviceBuffer(allb.nbytes, device)
        buf.upload(allb)
        d_balls = buf.ptr
    try:
        nv = lib.vltr_clip_variant_count()
        ms = np.zeros(nv)
        chosen = C.c_int(0)
        check(lib.vltr_context_autotune(ctx, d_balls, off.ctypes.data_as(_u64p), len(structures), float(probe),
                                        int(bool(contacts_only)), int(repeats), ms.ctypes.data_as(_f64p), nv,
                                        C.byref(chosen)), "vltr_context_autotune")
    finally:
        if buf is not None:
            buf.free()
    names = clip_variants()
    return names[chosen.value], {names[v]: float(ms[v]) for v in range(nv)}


def run_batched(structures, probe, contacts_only=False, device=0):
    """Tessellation of many independent structures in one pipeline pass; one
    BatchPart per structure (equal to running each structure alone)."""
    lib = load_library()
    ctx = context(device)
    arrays = [_balls_array(s) for s in structures]
    if not arrays:
        return []
    off = offsets_of(arrays)
    allb = np.ascontiguousarray(np.concatenate(arrays)) if int(off[-1]) else np.zeros((0, 4))
    handle = C.c_void_p()
    rc = lib.vltr_compute_tessellation_batched(ctx, allb.ctypes.data_as(_vp) if len(allb) else None,
                                               off.ctypes.data_as(_u64p), len(arrays), float(probe),
                                               int(bool(contacts_only)), C.byref(handle))
    check(rc, "vltr_compute_tessellation_batched")
    try:
        parts = []
        for k, arr in enumerate(arrays):
            nc, ncell = C.c_size_t(), C.c_size_t()
            cp = lib.vltr_batch_contacts(handle, k, C.byref(nc))
            clp = lib.vltr_batch_cells(handle, k, C.byref(ncell))
            parts.append(BatchPart(len(arr), _copy(cp, nc.value, CONTACT_DTYPE), _copy(clp, ncell.value, CELL_DTYPE)))
        return parts
    finally:
        lib.vltr_batch_free(handle)


class Updateable:
    """UpdateableTessellation (updateable.rs:154-742) on the GPU; the methods
    return the reference's booleans."""

    def __init__(self, with_backup=False, device=0):
        lib = load_library()
        h = C.c_void_p()
        check(lib.vltr_updateable_create(device, int(bool(with_backup)), C.byref(h)), "vltr_updateable_create")
        self._h = h
        self._lib = lib

    def close(self):
        if self._h:
            self._lib.vltr_updateable_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        ok = C.c_int(0)
        check(getattr(self._lib, name)(self._h, *args, C.byref(ok)), name)
        return bool(ok.value)

    def init(self, balls, probe, pbox=None):
        b = _balls_array(balls)
        box = make_pbox(pbox)
        return self._call("vltr_updateable_init", b.ctypes.data_as(_vp) if len(b) else None, len(b), float(probe),
                          C.byref(box) if box is not None else None)

    def update(self, balls):
        b = _balls_array(balls)
        return self._call("vltr_updateable_update", b.ctypes.data_as(_vp) if len(b) else None, len(b))

    def update_with_changed(self, balls, changed):
        b = _balls_array(balls)
        ch = np.ascontiguousarray(np.asarray(changed, dtype=np.uint64).reshape(-1))
        return self._call("vltr_updateable_update_with_changed", b.ctypes.data_as(_vp) if len(b) else None, len(b),
                          ch.ctypes.data_as(_u64p) if len(ch) else None, len(ch))

    def set_exclusion(self, ball_id, excluded):
        return self._call("vltr_updateable_set_exclusion", int(ball_id), int(bool(excluded)))

    def restore(self):
        return self._call("vltr_updateable_restore")

    def set_energy(self, types, W, sas_weight=1.0, volume_weight=1.0):
        """E = sum_contacts W[t_a][t_b] area + sas_weight sum SAS +
        volume_weight sum volume and its gradient, computed with every later
        state change."""
        t = np.ascontiguousarray(np.asarray(types, dtype=np.int32).reshape(-1))
        w = np.ascontiguousarray(np.asarray(W, dtype=np.float64).reshape(16))
        check(self._lib.vltr_updateable_set_energy(self._h, t.ctypes.data_as(_i32p) if len(t) else None, len(t),
                                                   w.ctypes.data_as(_f64p), float(sas_weight), float(volume_weight)),
              "vltr_updateable_set_energy")
        self._n_types = len(t)

    def energy_gradient(self):
        """(E, dE/dcentres (n, 3)) of the current state."""
        e = C.c_double()
        g = np.zeros((getattr(self, "_n_types", 0), 3))
        check(self._lib.vltr_updateable_energy_gradient(self._h, C.byref(e), g.ctypes.data_as(_f64p)),
              "vltr_updateable_energy_gradient")
        return e.value, g

    def energy_device(self):
        """(E, device pointer to dE/dcentres, 3 n doubles) of the current state."""
        e, ptr = C.c_double(), C.c_void_p()
        check(self._lib.vltr_updateable_energy_gradient_device(self._h, C.byref(e), C.byref(ptr)),
              "vltr_updateable_energy_gradient_device")
        return e.value, ptr.value

    def summary(self) -> Result:
        h = C.c_void_p()
        check(self._lib.vltr_updateable_summary(self._h, C.byref(h)), "vltr_updateable_summary")
        try:
            return Result(h)
        finally:
            self._lib.vltr_result_free(h)

    def _ids(self, name):
        fn = getattr(self._lib, name)
        k = fn(self._h, None, 0)
        out = np.zeros(max(k, 1), dtype=np.uint64)
        fn(self._h, out.ctypes.data_as(_u64p), k)
        return out[:k]

    def changed_ids(self):
        return self._ids("vltr_updateable_changed_ids")

    def affected_ids(self):
        return self._ids("vltr_updateable_affected_ids")

    def last_update_was_full_reinit(self):
        return bool(self._lib.vltr_updateable_last_update_was_full_reinit(self._h))

    def backup_enabled(self):
        return bool(self._lib.vltr_updateable_backup_enabled(self._h))

    def last_timings(self):
        buf = (C.c_double * 7)()
        k = self._lib.vltr_updateable_last_timings(self._h, buf, 7)
        names = ["grid", "neighbors", "pairs", "clip", "slow", "cells", "total"]
        return {names[i]: buf[i] for i in range(k)}


def energy_gradient(balls, probe, types, W, sas_weight=1.0, volume_weight=1.0, pbox=None, device=0):
    """(E, dE/dcentres (n, 3)) for E = sum_contacts W[t_a][t_b] area +
    sas_weight * sum SAS + volume_weight * sum volume (periodic box: over the
    deduplicated contact list)."""
    lib = load_library()
    ctx = context(device)
    b = _balls_array(balls)
    t = np.ascontiguousarray(np.asarray(types, dtype=np.int32).reshape(-1))
    w = np.ascontiguousarray(np.asarray(W, dtype=np.float64).reshape(16))
    if len(t) != len(b):
        raise ValueError("one type per ball")
    e = C.c_double()
    g = np.zeros((len(b), 3), dtype=np.float64)
    box = make_pbox(pbox)
    rc = lib.vltr_energy_gradient(ctx, b.ctypes.data_as(_vp) if len(b) else None, len(b), float(probe),
                                  C.byref(box) if box is not None else None,
                                  t.ctypes.data_as(_i32p) if len(t) else None, w.ctypes.data_as(_f64p),
                                  float(sas_weight), float(volume_weight), C.byref(e),
                                  g.ctypes.data_as(_f64p) if len(b) else None)
    check(rc, "vltr_energy_gradient")
    return e.value, g


def solvent_spheres(balls, probe, volume_probe=0.0, depth=2, pbox=None, device=0):
    """compute_solvent_spheres (solvent_spheres.rs:131-168) as a structured
    array (x, y, z, radius, weight, parent_index)."""
    lib = load_library()
    ctx = context(device)
    b = _balls_array(balls)
    box = make_pbox(pbox)
    h = C.c_void_p()
    rc = lib.vltr_compute_solvent_spheres(ctx, b.ctypes.data_as(_vp) if len(b) else None, len(b), float(probe),
                                          float(volume_probe), int(depth), C.byref(box) if box is not None else None,
                                          C.byref(h))
    if rc == 1:
        raise ValueError(last_error())
    check(rc, "vltr_compute_solvent_spheres")
    try:
        return _copy(lib.vltr_solvent_spheres(h), lib.vltr_solvent_count(h), SOLVENT_DTYPE)
    finally:
        lib.vltr_solvent_free(h)


# ------------------------------------------------------------ file front end
VLTR_ERR_IO = 5
FORMATS = {"auto": -1, "pdb": 0, "mmcif": 1, "xyzr": 2}
RECORD_FIELDS = ["record_name", "name", "alt_loc", "res_name", "chain_id", "i_code", "element"]


class ParseOptions(C.Structure):
    """vltr_parse_options (src/input/mod.rs:69-77)."""
    _fields_ = [("exclude_heteroatoms", C.c_int), ("include_hydrogens", C.c_int), ("as_assembly", C.c_int)]


class Radii:
    """RadiiLookup (src/input/radii.rs): default rules, or none with
    default=False; load_text raises ValueError with the reference's message."""

    def __init__(self, default=True):
        lib = load_library()
        self._h = C.c_void_p()
        check(lib.vltr_radii_create(int(bool(default)), C.byref(self._h)), "vltr_radii_create")

    def load_text(self, text):
        if load_library().vltr_radii_load_text(self._h, text.encode()) != VLTR_OK:
            raise ValueError(last_error())

    def get(self, res_name, atom_name):
        return load_library().vltr_radii_get(self._h, res_name.encode(), atom_name.encode())

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.vltr_radii_free(self._h)
            self._h = None


class Structure:
    """A parsed file: balls (n, 4) and atom records (dicts; none for XYZR)."""

    def __init__(self, handle):
        lib = load_library()
        self._h = handle
        self.format = {v: k for k, v in FORMATS.items()}[lib.vltr_structure_format(handle)]
        n = lib.vltr_structure_num_balls(handle)
        self.balls = _copy(lib.vltr_structure_balls(handle), n, np.dtype((np.float64, 4)))
        self.balls = np.asarray(self.balls, dtype=np.float64).reshape(n, 4)
        self.records = []
        serial, res_seq = C.c_int32(), C.c_int32()
        xyzb = (C.c_double * 4)()
        for i in range(lib.vltr_structure_num_records(handle)):
            rec = {f: lib.vltr_record_string(handle, i, k).decode() for k, f in enumerate(RECORD_FIELDS)}
            lib.vltr_record_numbers(handle, i, C.byref(serial), C.byref(res_seq), xyzb)
            rec.update(serial=serial.value, res_seq=res_seq.value, x=xyzb[0], y=xyzb[1], z=xyzb[2],
                       b_factor=xyzb[3])
            self.records.append(rec)

    def _groups(self, fn, *args):
        out = np.zeros(len(self.records), np.int32)
        rc = fn(self._h, *args, out.ctypes.data_as(_i32p))
        if rc != VLTR_OK:
            raise ValueError(last_error())
        return out

    def chain_grouping(self):
        return self._groups(load_library().vltr_grouping_chain)

    def residue_grouping(self):
        return self._groups(load_library().vltr_grouping_residue)

    def selection_grouping(self, selections):
        arr = (C.c_char_p * len(selections))(*[s.encode() for s in selections])
        return self._groups(load_library().vltr_grouping_selections, arr, len(selections))

    def select(self, selection):
        """Boolean mask of the records a selection matches."""
        out = np.zeros(len(self.records), np.uint8)
        rc = load_library().vltr_selection_match(self._h, selection.encode(), out.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc != VLTR_OK:
            raise ValueError(last_error())
        return out.astype(bool)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.vltr_structure_free(self._h)
            self._h = None


def _options(exclude_heteroatoms, include_hydrogens, as_assembly):
    return ParseOptions(int(bool(exclude_heteroatoms)), int(bool(include_hydrogens)), int(bool(as_assembly)))


def parse_file(path, exclude_heteroatoms=False, include_hydrogens=False, as_assembly=False, radii=None):
    """parse_file_with_records (src/input/mod.rs:300-334); OSError when the
    file cannot be read or its format not detected."""
    lib = load_library()
    h = C.c_void_p()
    opts = _options(exclude_heteroatoms, include_hydrogens, as_assembly)
    rc = lib.vltr_parse_file(str(path).encode(), C.byref(opts), radii._h if radii is not None else None, C.byref(h))
    if rc == VLTR_ERR_IO:
        raise OSError(last_error())
    check(rc, "vltr_parse_file")
    return Structure(h)


def parse_text(text, fmt="auto", exclude_heteroatoms=False, include_hydrogens=False, as_assembly=False,
               radii=None):
    lib = load_library()
    h = C.c_void_p()
    data = text.encode() if isinstance(text, str) else bytes(text)
    opts = _options(exclude_heteroatoms, include_hydrogens, as_assembly)
    check(lib.vltr_parse_text(data, len(data), FORMATS[fmt], C.byref(opts),
                              radii._h if radii is not None else None, C.byref(h)), "vltr_parse_text")
    return Structure(h)


def _take_string(ptr, n):
    try:
        return C.string_at(ptr, n).decode()
    finally:
        load_library().vltr_string_free(ptr)


def from_file(path, probe, pbox=None, with_cell_vertices=False, selections=None, device=0, want_json=False):
    """compute_tessellation_from_file (src/input/mod.rs:518-552): Result (and
    the CLI's JSON text with want_json); OSError on every failure, as the
    PyO3 binding raises (src/python.rs:310-321)."""
    lib = load_library()
    ctx = context(device)
    box = make_pbox(pbox)
    arr, ns = None, 0
    if selections is not None:
        arr = (C.c_char_p * max(len(selections), 1))(*[s.encode() for s in selections])
        ns = len(selections)
    h = C.c_void_p()
    rc = lib.vltr_compute_tessellation_from_file(ctx, str(path).encode(), float(probe),
                                                 C.byref(box) if box is not None else None,
                                                 int(bool(with_cell_vertices)), arr, ns, C.byref(h))
    if rc in (VLTR_ERR_IO, 1):
        raise OSError(last_error())
    check(rc, "vltr_compute_tessellation_from_file")
    try:
        res = Result(h)
        if want_json:
            txt, n = C.c_void_p(), C.c_size_t()
            check(lib.vltr_result_to_json(h, C.byref(txt), C.byref(n)), "vltr_result_to_json")
            res.json = _take_string(txt, n.value)
        return res
    finally:
        lib.vltr_result_free(h)


def run_json(balls, probe, pbox=None, groups=None, with_cell_vertices=False, device=0):
    """The CLI's JSON document for a tessellation (src/main.rs:22-34)."""
    lib = load_library()
    ctx = context(device)
    b = _balls_array(balls)
    box = make_pbox(pbox)
    g = None if groups is None else np.ascontiguousarray(np.asarray(groups, np.int32))
    h = C.c_void_p()
    check(lib.vltr_compute_tessellation(ctx, b.ctypes.data_as(_vp) if len(b) else None, len(b), float(probe),
                                        C.byref(box) if box is not None else None,
                                        g.ctypes.data_as(_i32p) if g is not None and len(g) else None,
                                        len(g) if g is not None else 0, int(bool(with_cell_vertices)), C.byref(h)),
          "vltr_compute_tessellation")
    try:
        txt, n = C.c_void_p(), C.c_size_t()
        check(lib.vltr_result_to_json(h, C.byref(txt), C.byref(n)), "vltr_result_to_json")
        return _take_string(txt, n.value)
    finally:
        lib.vltr_result_free(h)


def pymol_graphics(balls, probe, groups=None, angle_step=0.5, object_name="contacts", device=0):
    """GraphicsWriter::from_balls + write_pymol (src/graphics.rs:45-178)."""
    lib = load_library()
    ctx = context(device)
    b = _balls_array(balls)
    g = None if groups is None else np.ascontiguousarray(np.asarray(groups, np.int32))
    txt, n = C.c_void_p(), C.c_size_t()
    check(lib.vltr_pymol_graphics(ctx, b.ctypes.data_as(_vp) if len(b) else None, len(b), float(probe),
                                  g.ctypes.data_as(_i32p) if g is not None and len(g) else None,
                                  len(g) if g is not None else 0, float(angle_step), object_name.encode(),
                                  C.byref(txt), C.byref(n)), "vltr_pymol_graphics")
    return _take_string(txt, n.value)
