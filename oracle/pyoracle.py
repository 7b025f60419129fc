"""ctypes loader for the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker. The product path never does.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VLTR_ORACLE_LIB: another build of the same library (the sanitizer build,
# oracle/_san/, loaded by tests/test_sanitizers.py)
LIB_PATH = os.environ.get("VLTR_ORACLE_LIB") or os.path.join(HERE, "liboracle_vltr.so")

_u64p = C.POINTER(C.c_uint64)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p

_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    sz = C.c_size_t
    L.oracle_compute_tessellation.argtypes = [_f64p, sz, C.c_double, _f64p, _i32p, sz, C.c_int, C.c_int, C.POINTER(_vp)]
    L.oracle_compute_contacts_only.argtypes = [_f64p, sz, C.c_double, _f64p, _i32p, sz, C.c_int, C.POINTER(_vp)]
    L.oracle_collisions.argtypes = [_f64p, sz, C.c_double, _f64p, _i32p, sz, C.c_int, C.POINTER(_vp)]
    for nm in ["num_balls", "num_contacts", "num_cells", "num_vertices", "num_edges", "num_neighbors", "num_pairs"]:
        f = getattr(L, "oracle_result_" + nm)
        f.argtypes = [_vp]
        f.restype = sz
    L.oracle_result_has_network.argtypes = [_vp]
    L.oracle_result_contacts.argtypes = [_vp, _u64p, _f64p, _f64p, _u8p]
    L.oracle_result_cells.argtypes = [_vp, _u64p, _f64p, _f64p]
    L.oracle_result_vertices.argtypes = [_vp, _u64p, _f64p]
    L.oracle_result_edges.argtypes = [_vp, _u64p, _f64p]
    L.oracle_result_neighbors.argtypes = [_vp, _u64p, _u64p, _f64p, _u8p]
    L.oracle_result_pairs.argtypes = [_vp, _u64p]
    L.oracle_result_free.argtypes = [_vp]
    L.oracle_pair_records.argtypes = [_f64p, sz, C.c_double, _f64p, _i32p, sz, C.c_int, C.c_int, _f64p,
                                      C.POINTER(C.c_uint32)]
    L.oracle_count_flops.argtypes = [_f64p, sz, C.c_double, _f64p, _i32p, sz, C.c_int, _u64p, _u64p]
    L.oracle_energy.argtypes = [_f64p, sz, C.c_double, _f64p, _i32p, _f64p, C.c_int, _f64p]
    L.oracle_energy_dual.argtypes = [_f64p, sz, C.c_double, _f64p, _i32p, _f64p, sz, C.c_int, _f64p, _f64p]
    L.oracle_energy_terms.argtypes = [_f64p, sz, C.c_double, _i32p, _f64p, C.c_int, C.POINTER(_vp)]
    L.oracle_updateable_new.argtypes = [C.c_int, C.c_int]
    L.oracle_updateable_new.restype = _vp
    L.oracle_updateable_free.argtypes = [_vp]
    L.oracle_updateable_set_remove_from_old_box.argtypes = [_vp, C.c_int]
    L.oracle_updateable_init.argtypes = [_vp, _f64p, sz, C.c_double, _f64p]
    L.oracle_updateable_update.argtypes = [_vp, _f64p, sz]
    L.oracle_updateable_update_with_changed.argtypes = [_vp, _f64p, sz, _u64p, sz]
    L.oracle_updateable_set_exclusion.argtypes = [_vp, sz, C.c_int]
    L.oracle_updateable_restore.argtypes = [_vp]
    L.oracle_updateable_last_update_was_full_reinit.argtypes = [_vp]
    L.oracle_updateable_num_changed.argtypes = [_vp]
    L.oracle_updateable_num_changed.restype = sz
    L.oracle_updateable_num_affected.argtypes = [_vp]
    L.oracle_updateable_num_affected.restype = sz
    L.oracle_updateable_changed_ids.argtypes = [_vp, _u64p]
    L.oracle_updateable_affected_ids.argtypes = [_vp, _u64p]
    L.oracle_updateable_summary.argtypes = [_vp, C.POINTER(_vp)]
    L.oracle_time_tessellation.argtypes = [_f64p, sz, C.c_double, _f64p, _i32p, sz, C.c_int, C.c_int, C.c_int]
    L.oracle_time_tessellation.restype = C.c_double
    L.oracle_max_threads.restype = C.c_int
    _lib = L
    return L


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _balls(balls):
    b = np.ascontiguousarray(np.asarray(balls, dtype=np.float64).reshape(-1, 4))
    return b


def _pbox(pbox):
    if pbox is None:
        return None
    return np.ascontiguousarray(np.asarray(pbox, dtype=np.float64).reshape(9))


def _groups(groups):
    if groups is None:
        return None, 0
    g = np.ascontiguousarray(np.asarray(groups, dtype=np.int32))
    return g, len(g)


@dataclass
class Result:
    num_balls: int = 0
    contact_ids: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.uint64))
    contact_area: np.ndarray = field(default_factory=lambda: np.zeros(0))
    contact_arc: np.ndarray = field(default_factory=lambda: np.zeros(0))
    contact_central: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint8))
    cell_index: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    cell_sas: np.ndarray = field(default_factory=lambda: np.zeros(0))
    cell_volume: np.ndarray = field(default_factory=lambda: np.zeros(0))
    vertex_ids: np.ndarray | None = None
    vertex_xyz: np.ndarray | None = None
    edge_ids: np.ndarray | None = None
    edge_length: np.ndarray | None = None


def _read(handle) -> Result:
    L = _load()
    r = Result()
    r.num_balls = L.oracle_result_num_balls(handle)
    nc = L.oracle_result_num_contacts(handle)
    r.contact_ids = np.zeros((nc, 2), np.uint64)
    r.contact_area = np.zeros(nc)
    r.contact_arc = np.zeros(nc)
    r.contact_central = np.zeros(nc, np.uint8)
    if nc:
        L.oracle_result_contacts(handle, _ptr(r.contact_ids, _u64p), _ptr(r.contact_area, _f64p),
                                 _ptr(r.contact_arc, _f64p), _ptr(r.contact_central, _u8p))
    ncell = L.oracle_result_num_cells(handle)
    r.cell_index = np.zeros(ncell, np.uint64)
    r.cell_sas = np.zeros(ncell)
    r.cell_volume = np.zeros(ncell)
    if ncell:
        L.oracle_result_cells(handle, _ptr(r.cell_index, _u64p), _ptr(r.cell_sas, _f64p), _ptr(r.cell_volume, _f64p))
    if L.oracle_result_has_network(handle):
        nv = L.oracle_result_num_vertices(handle)
        ne = L.oracle_result_num_edges(handle)
        r.vertex_ids = np.zeros((nv, 4), np.uint64)
        r.vertex_xyz = np.zeros((nv, 3))
        r.edge_ids = np.zeros((ne, 3), np.uint64)
        r.edge_length = np.zeros(ne)
        if nv:
            L.oracle_result_vertices(handle, _ptr(r.vertex_ids, _u64p), _ptr(r.vertex_xyz, _f64p))
        if ne:
            L.oracle_result_edges(handle, _ptr(r.edge_ids, _u64p), _ptr(r.edge_length, _f64p))
    return r


def compute_tessellation(balls, probe, pbox=None, groups=None, with_vertices=False, threads=0) -> Result:
    L = _load()
    b = _balls(balls)
    pb = _pbox(pbox)
    g, ng = _groups(groups)
    h = C.c_void_p()
    L.oracle_compute_tessellation(_ptr(b, _f64p), len(b), float(probe), _ptr(pb, _f64p), _ptr(g, _i32p), ng,
                                  int(with_vertices), int(threads), C.byref(h))
    try:
        return _read(h)
    finally:
        L.oracle_result_free(h)


def compute_contacts_only(balls, probe, pbox=None, groups=None, threads=0) -> Result:
    L = _load()
    b = _balls(balls)
    pb = _pbox(pbox)
    g, ng = _groups(groups)
    h = C.c_void_p()
    L.oracle_compute_contacts_only(_ptr(b, _f64p), len(b), float(probe), _ptr(pb, _f64p), _ptr(g, _i32p), ng,
                                   int(threads), C.byref(h))
    try:
        return _read(h)
    finally:
        L.oracle_result_free(h)


def collisions(balls, probe, pbox=None, groups=None, threads=0):
    """Per-ball sorted neighbor lists as CSR: (offsets, index, value, excluded, pairs)."""
    L = _load()
    b = _balls(balls)
    pb = _pbox(pbox)
    g, ng = _groups(groups)
    h = C.c_void_p()
    L.oracle_collisions(_ptr(b, _f64p), len(b), float(probe), _ptr(pb, _f64p), _ptr(g, _i32p), ng, int(threads),
                        C.byref(h))
    try:
        n = len(b)
        m = L.oracle_result_num_neighbors(h)
        offsets = np.zeros(n + 1, np.uint64)
        index = np.zeros(m, np.uint64)
        value = np.zeros(m)
        excluded = np.zeros(n, np.uint8)
        L.oracle_result_neighbors(h, _ptr(offsets, _u64p), _ptr(index, _u64p), _ptr(value, _f64p),
                                  _ptr(excluded, _u8p))
        npairs = L.oracle_result_num_pairs(h)
        pairs = np.zeros((npairs, 2), np.uint64)
        if npairs:
            L.oracle_result_pairs(h, _ptr(pairs, _u64p))
        return offsets, index, value, excluded, pairs
    finally:
        L.oracle_result_free(h)


def pair_records(balls, probe, n_pairs, pbox=None, groups=None, mode=0, threads=0):
    """(P x 6 values, P flags) for the collision step's pair list, in pair order."""
    L = _load()
    b = _balls(balls)
    pb = _pbox(pbox)
    g, ng = _groups(groups)
    out = np.zeros((n_pairs, 6))
    flags = np.zeros(n_pairs, np.uint32)
    L.oracle_pair_records(_ptr(b, _f64p), len(b), float(probe), _ptr(pb, _f64p), _ptr(g, _i32p), ng, int(mode),
                          int(threads), _ptr(out, _f64p), flags.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out, flags


def count_flops(balls, probe, pbox=None, groups=None, mode=0):
    L = _load()
    b = _balls(balls)
    pb = _pbox(pbox)
    g, ng = _groups(groups)
    fl = C.c_uint64()
    tp = C.c_uint64()
    L.oracle_count_flops(_ptr(b, _f64p), len(b), float(probe), _ptr(pb, _f64p), _ptr(g, _i32p), ng, int(mode),
                         C.byref(fl), C.byref(tp))
    return fl.value, tp.value


def energy(balls, probe, types, W, threads=0, pbox=None) -> float:
    L = _load()
    b = _balls(balls)
    pb = _pbox(pbox)
    t = np.ascontiguousarray(np.asarray(types, np.int32))
    w = np.ascontiguousarray(np.asarray(W, np.float64).reshape(16))
    e = C.c_double()
    L.oracle_energy(_ptr(b, _f64p), len(b), float(probe), _ptr(pb, _f64p), _ptr(t, _i32p), _ptr(w, _f64p),
                    int(threads), C.byref(e))
    return e.value


def energy_dual(balls, probe, types, W, ball_id, axis, pbox=None):
    L = _load()
    b = _balls(balls)
    pb = _pbox(pbox)
    t = np.ascontiguousarray(np.asarray(types, np.int32))
    w = np.ascontiguousarray(np.asarray(W, np.float64).reshape(16))
    e = C.c_double()
    de = C.c_double()
    L.oracle_energy_dual(_ptr(b, _f64p), len(b), float(probe), _ptr(pb, _f64p), _ptr(t, _i32p), _ptr(w, _f64p),
                         int(ball_id), int(axis), C.byref(e), C.byref(de))
    return e.value, de.value


def energy_terms(balls, probe, types, W, threads=0) -> Result:
    """Result whose contact_area holds W[t_a][t_b]*area per contact."""
    L = _load()
    b = _balls(balls)
    t = np.ascontiguousarray(np.asarray(types, np.int32))
    w = np.ascontiguousarray(np.asarray(W, np.float64).reshape(16))
    h = C.c_void_p()
    L.oracle_energy_terms(_ptr(b, _f64p), len(b), float(probe), _ptr(t, _i32p), _ptr(w, _f64p), int(threads),
                          C.byref(h))
    try:
        return _read(h)
    finally:
        L.oracle_result_free(h)


def time_tessellation(balls, probe, pbox=None, groups=None, contacts_only=False, threads=0, reps=1) -> float:
    L = _load()
    b = _balls(balls)
    pb = _pbox(pbox)
    g, ng = _groups(groups)
    return L.oracle_time_tessellation(_ptr(b, _f64p), len(b), float(probe), _ptr(pb, _f64p), _ptr(g, _i32p), ng,
                                      int(contacts_only), int(threads), int(reps))


def max_threads() -> int:
    return _load().oracle_max_threads()


class Updateable:
    """UpdateableTessellation restated on the CPU (updateable.rs:154-742).

    remove_from_old_box=True corrects spheres_searcher.rs:144-175, where a
    moved sphere stays listed in its old grid box (the reference removes it
    from the box of its new position), so that later searches report it twice
    and the updateable state gains duplicate contacts. The default keeps the
    reference's behaviour.
    """

    def __init__(self, with_backup=False, threads=0, remove_from_old_box=False):
        self._L = _load()
        self._h = self._L.oracle_updateable_new(int(with_backup), int(threads))
        self._L.oracle_updateable_set_remove_from_old_box(self._h, int(bool(remove_from_old_box)))
        self._n = 0

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.oracle_updateable_free(self._h)
            self._h = None

    def init(self, balls, probe, pbox=None) -> bool:
        b = _balls(balls)
        pb = _pbox(pbox)
        self._n = len(b)
        return bool(self._L.oracle_updateable_init(self._h, _ptr(b, _f64p), len(b), float(probe), _ptr(pb, _f64p)))

    def update(self, balls) -> bool:
        b = _balls(balls)
        return bool(self._L.oracle_updateable_update(self._h, _ptr(b, _f64p), len(b)))

    def update_with_changed(self, balls, ids) -> bool:
        b = _balls(balls)
        i = np.ascontiguousarray(np.asarray(ids, np.uint64))
        return bool(self._L.oracle_updateable_update_with_changed(self._h, _ptr(b, _f64p), len(b), _ptr(i, _u64p),
                                                                   len(i)))

    def set_exclusion(self, idx, excluded) -> bool:
        return bool(self._L.oracle_updateable_set_exclusion(self._h, int(idx), int(bool(excluded))))

    def restore(self) -> bool:
        return bool(self._L.oracle_updateable_restore(self._h))

    def last_update_was_full_reinit(self) -> bool:
        return bool(self._L.oracle_updateable_last_update_was_full_reinit(self._h))

    def changed_ids(self):
        n = self._L.oracle_updateable_num_changed(self._h)
        out = np.zeros(n, np.uint64)
        if n:
            self._L.oracle_updateable_changed_ids(self._h, _ptr(out, _u64p))
        return out

    def affected_ids(self):
        n = self._L.oracle_updateable_num_affected(self._h)
        out = np.zeros(n, np.uint64)
        if n:
            self._L.oracle_updateable_affected_ids(self._h, _ptr(out, _u64p))
        return out

    def summary(self) -> Result:
        h = C.c_void_p()
        self._L.oracle_updateable_summary(self._h, C.byref(h))
        try:
            return _read(h)
        finally:
            self._L.oracle_result_free(h)


def solvent_spheres(balls, probe, volume_probe=0.0, depth=2, pbox=None, threads=0):
    """solvent_spheres.rs:131-168: (m, 6) array of x, y, z, radius, weight, parent."""
    L = _load()
    f = L.oracle_solvent_spheres
    f.restype = C.c_int64
    f.argtypes = [_vp, C.c_size_t, C.c_double, C.c_double, C.c_int, _vp, C.c_int, _vp, C.c_int64]
    b = np.ascontiguousarray(np.asarray(balls, np.float64).reshape(-1, 4))
    pb = None if pbox is None else np.ascontiguousarray(np.asarray(pbox, np.float64).reshape(9))
    args = (b.ctypes.data if len(b) else None, len(b), float(probe), float(volume_probe), int(depth),
            pb.ctypes.data if pb is not None else None, int(threads))
    m = f(*args, None, 0)
    out = np.zeros((max(m, 1), 6))
    f(*args, out.ctypes.data, m)
    return out[:m]


def icosahedron(depth):
    L = _load()
    f = L.oracle_icosahedron
    f.restype = C.c_int64
    f.argtypes = [C.c_int, _vp, C.c_int64]
    m = f(int(depth), None, 0)
    out = np.zeros((m, 3))
    f(int(depth), out.ctypes.data, m)
    return out


def contact_polygons(balls, probe, groups=None):
    """Per contact with area > 0 (pair order): dict a, b, ic_center, ic_r,
    axis, bary, contour [(x, y, z, angle, left, right), ...]."""
    L = _load()
    f = L.oracle_contact_polygons
    f.restype = C.c_int64
    f.argtypes = [_vp, C.c_size_t, C.c_double, _vp, C.c_size_t, _vp, C.c_int64]
    b = np.ascontiguousarray(np.asarray(balls, np.float64).reshape(-1, 4))
    g = None if groups is None else np.ascontiguousarray(np.asarray(groups, np.int32))
    args = (b.ctypes.data if len(b) else None, len(b), float(probe), g.ctypes.data if g is not None else None,
            len(g) if g is not None else 0)
    m = f(*args, None, 0)
    out = np.zeros(max(m, 1))
    f(*args, out.ctypes.data, m)
    polys, i = [], 0
    while i < m:
        h = out[i:i + 14]
        k = int(h[12])
        pts = out[i + 14:i + 14 + 6 * k].reshape(k, 6)
        polys.append({"a": int(h[0]), "b": int(h[1]), "ic_center": tuple(h[2:5]), "ic_r": float(h[5]),
                      "axis": tuple(h[6:9]), "bary": tuple(h[9:12]),
                      "contour": [(p[0], p[1], p[2], p[3], int(p[4]), int(p[5])) for p in pts]})
        i += 14 + 6 * k
    return polys
