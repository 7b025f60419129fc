"""voronota_ltr - radical tessellation of balls on MI355X.

Drop-in for the reference's Python module (src/python.rs:238-401): same
function names, arguments, dict layout and error types. The computation runs
in the HIP library libvltr_hip.so; there is no CPU fallback.
"""
from __future__ import annotations

import numpy as np

from . import _native
from ._native import Result, VltrError, last_timings
from ._native import Updateable as _Updateable

__all__ = [
    "compute_tessellation",
    "compute_contacts_only",
    "tessellate",
    "Result",
    "VltrError",
    "last_timings",
]


def _ball_from_sequence(seq, type_name):
    # python.rs:28-42
    if len(seq) != 4:
        raise ValueError(f"ball {type_name} must have 4 elements (x, y, z, r)")
    return [float(seq[0]), float(seq[1]), float(seq[2]), float(seq[3])]


def _parse_single_ball(obj):
    # python.rs:45-67: dict, tuple or list
    if isinstance(obj, dict):
        vals = []
        for k in ("x", "y", "z", "r"):
            if k not in obj:
                raise KeyError(f"missing '{k}'")
            vals.append(float(obj[k]))
        return vals
    if isinstance(obj, tuple):
        return _ball_from_sequence(obj, "tuple")
    if isinstance(obj, list):
        return _ball_from_sequence(obj, "list")
    raise TypeError("ball must be tuple (x,y,z,r), list [x,y,z,r], or dict {x,y,z,r}")


def parse_balls(obj) -> np.ndarray:
    """python.rs:69-89: numpy (N, 4) array, or a list of tuples/lists/dicts."""
    if isinstance(obj, np.ndarray):
        if obj.ndim != 2 or obj.shape[1] != 4:
            raise ValueError("numpy array must have shape (N, 4) for (x, y, z, r)")
        return np.ascontiguousarray(obj, dtype=np.float64)
    if not isinstance(obj, list):
        raise TypeError("balls must be a list or a numpy array")
    rows = [_parse_single_ball(item) for item in obj]
    if not rows:
        return np.zeros((0, 4), dtype=np.float64)
    return np.asarray(rows, dtype=np.float64)


def parse_periodic_box(obj):
    """python.rs:92-120: {"corners": [lo, hi]} or {"vectors": [a, b, c]} -> 9 floats."""
    if not isinstance(obj, dict):
        raise TypeError("periodic_box must be a dict")
    if "corners" in obj:
        c = [tuple(float(v) for v in p) for p in obj["corners"]]
        if len(c) != 2:
            raise ValueError("corners must have exactly 2 points")
        lo, hi = c
        return [hi[0] - lo[0], 0.0, 0.0, 0.0, hi[1] - lo[1], 0.0, 0.0, 0.0, hi[2] - lo[2]]
    if "vectors" in obj:
        v = [tuple(float(x) for x in p) for p in obj["vectors"]]
        if len(v) != 3:
            raise ValueError("vectors must have exactly 3 direction vectors")
        return [*v[0], *v[1], *v[2]]
    raise ValueError("periodic_box must have 'corners' or 'vectors' key")


def tessellate(balls, probe, periodic_box=None, groups=None, with_cell_vertices=False, contacts_only=False,
               device=0) -> Result:
    """Array-level entry point: balls as (N, 4) floats, periodic_box as None or
    9 floats (shift_a | shift_b | shift_c); returns numpy structured arrays."""
    return _native.run(balls, probe, periodic_box, groups, with_cell_vertices, contacts_only, device)


def _contacts_to_list(c):
    return [
        {"id_a": int(r["id_a"]), "id_b": int(r["id_b"]), "area": float(r["area"]),
         "arc_length": float(r["arc_length"]), "central": bool(r["central"])}
        for r in c
    ]


def _cells_to_list(c):
    return [{"index": int(r["index"]), "sas_area": float(r["sas_area"]), "volume": float(r["volume"])} for r in c]


def _ids_to_list(ids):
    return [None if int(i) == int(_native.NULL_ID) else int(i) for i in ids]


def compute_tessellation(balls, probe, periodic_box=None, groups=None, with_cell_vertices=False):
    """python.rs:238-264: dict with num_balls, contacts, cells, total_sas_area,
    total_volume, total_contact_area (+ cell_vertices, cell_edges)."""
    b = parse_balls(balls)
    pb = parse_periodic_box(periodic_box) if periodic_box is not None else None
    g = None if groups is None else [int(x) for x in groups]
    return _result_dict(_native.run(b, float(probe), pb, g, with_cell_vertices, False), with_cell_vertices)


def _result_dict(r, with_cell_vertices):
    """result_to_dict (python.rs:189-210)."""
    out = {
        "num_balls": int(r.num_balls),
        "contacts": _contacts_to_list(r.contacts),
        "cells": _cells_to_list(r.cells),
        "total_sas_area": float(r.total_sas_area),
        "total_volume": float(r.total_volume),
        "total_contact_area": float(r.total_contact_area),
    }
    if r.cell_vertices is not None:
        out["cell_vertices"] = [
            {"ball_indices": _ids_to_list(v["ball_indices"]), "x": float(v["x"]), "y": float(v["y"]),
             "z": float(v["z"]), "is_on_sas": int(v["ball_indices"][3]) == int(_native.NULL_ID)}
            for v in r.cell_vertices
        ]
        out["cell_edges"] = [
            {"ball_indices": _ids_to_list(e["ball_indices"]), "length": float(e["length"]),
             "is_on_sas": int(e["ball_indices"][2]) == int(_native.NULL_ID)}
            for e in r.cell_edges
        ]
    return out


def compute_contacts_only(balls, probe, periodic_box=None, groups=None):
    """tessellation.rs:125-166 exposed to Python: list of contact dicts."""
    b = parse_balls(balls)
    pb = parse_periodic_box(periodic_box) if periodic_box is not None else None
    g = None if groups is None else [int(x) for x in groups]
    r = _native.run(b, float(probe), pb, g, False, True)
    return _contacts_to_list(r.contacts)


def compute_tessellation_batched(structures, probe, contacts_only=False):
    """compute_tessellation (or compute_contacts_only) of many independent
    structures in one pass on the GPU. Each element of `structures` is any
    ball collection parse_balls accepts; returns one dict per structure in
    the layout of compute_tessellation / compute_contacts_only."""
    parts = _native.run_batched([parse_balls(s) for s in structures], probe, contacts_only=contacts_only)
    out = []
    for part in parts:
        if contacts_only:
            out.append(_contacts_to_list(part.contacts))
        else:
            out.append({"contacts": _contacts_to_list(part.contacts), "cells": _cells_to_list(part.cells)})
    return out


class UpdateableTessellation:
    """updateable.rs:154-742 from Python: balls in any form parse_balls
    accepts, periodic_box as in compute_tessellation; summary() returns the
    dict layout of compute_tessellation."""

    def __init__(self, with_backup=False):
        self._u = _Updateable(with_backup=with_backup)

    def init(self, balls, probe, periodic_box=None):
        pb = parse_periodic_box(periodic_box) if periodic_box is not None else None
        return self._u.init(parse_balls(balls), probe, pb)

    def update(self, balls):
        return self._u.update(parse_balls(balls))

    def update_with_changed(self, balls, changed_ids):
        return self._u.update_with_changed(parse_balls(balls), [int(i) for i in changed_ids])

    def set_exclusion(self, ball_id, excluded):
        return self._u.set_exclusion(ball_id, excluded)

    def restore(self):
        return self._u.restore()

    def summary(self):
        r = self._u.summary()
        return {"num_balls": int(r.num_balls), "contacts": _contacts_to_list(r.contacts),
                "cells": _cells_to_list(r.cells), "total_sas_area": float(r.total_sas_area),
                "total_volume": float(r.total_volume), "total_contact_area": float(r.total_contact_area)}

    def changed_ids(self):
        return self._u.changed_ids().tolist()

    def affected_ids(self):
        return self._u.affected_ids().tolist()

    def last_update_was_full_reinit(self):
        return self._u.last_update_was_full_reinit()

    def backup_enabled(self):
        return self._u.backup_enabled()


def compute_solvent_spheres(balls, probe, volume_probe=None, subdivision_depth=None, periodic_box=None):
    """solvent_spheres.rs:131-168: list of dicts (x, y, z, radius, weight,
    parent_index); volume_probe defaults to 0.0 and subdivision_depth to 2
    (162 samples per exposed atom)."""
    depth = 2 if subdivision_depth is None else int(subdivision_depth)
    if depth < 0:
        raise OverflowError("can't convert negative int to unsigned")
    pb = parse_periodic_box(periodic_box) if periodic_box is not None else None
    arr = _native.solvent_spheres(parse_balls(balls), probe, 0.0 if volume_probe is None else volume_probe,
                                  min(depth, 4), pb)
    return [{"x": float(r["x"]), "y": float(r["y"]), "z": float(r["z"]), "radius": float(r["radius"]),
             "weight": float(r["weight"]), "parent_index": int(r["parent_index"])} for r in arr]


def compute_tessellation_from_file(path, probe, periodic_box=None, with_cell_vertices=False, group_selections=None):
    """src/python.rs:292-324: parse a PDB / mmCIF / XYZR file (default
    options and radii), optionally group atoms by selections, tessellate; the
    dict of compute_tessellation. Every failure raises OSError (IOError)."""
    pb = parse_periodic_box(periodic_box) if periodic_box is not None else None
    res = _native.from_file(path, probe, pb, with_cell_vertices,
                            None if group_selections is None else list(group_selections))
    return _result_dict(res, with_cell_vertices)
