"""Comparison of a GPU result with the oracle (test helper).

Bar (pairs and indices bit-exact, values to 1e-9 relative):
  * contact pair sequence and central flags: exact, in the same order;
  * cell index sequence: exact;
  * areas, arc lengths, SAS, volumes: |gpu - ref| <= 1e-9 * |ref| + floor,
    floor = 1e-12 x the full-sphere / full-disk value of that ball or contact
    (SAS 4 pi R^2, volume 4/3 pi R^3, area and arc pi r^2 / 2 pi r of the
    intersection circle bounded by the larger ball), so that values obtained
    as a small difference of large solid angles are compared at the precision
    their inputs carry.
"""
from __future__ import annotations

import math

import numpy as np

REL = 1e-9
FLOOR = 1e-12


def _close(gpu, ref, scale, what):
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    tol = REL * np.abs(ref) + FLOOR * scale
    bad = np.abs(gpu - ref) > tol
    if np.any(bad):
        i = int(np.flatnonzero(bad)[0])
        rel = abs(gpu[i] - ref[i]) / max(abs(ref[i]), 1e-300)
        raise AssertionError(f"{what}: {int(bad.sum())} of {len(ref)} values differ; first at {i}: "
                             f"gpu={gpu[i]!r} ref={ref[i]!r} rel={rel:.3e}")
    return int(np.sum(np.abs(gpu - ref) > REL * np.abs(ref)))  # floor-limited count


def compare(gpu, ref, balls, probe, check_cells=True, ordered=True):
    """gpu: voronota_ltr.Result; ref: pyoracle.Result. Returns a stats dict."""
    balls = np.asarray(balls, dtype=np.float64).reshape(-1, 4)
    R = balls[:, 3] + probe
    g_ids = np.stack([gpu.contacts["id_a"], gpu.contacts["id_b"]], axis=1).astype(np.uint64)
    r_ids = ref.contact_ids.astype(np.uint64)
    g_area = gpu.contacts["area"]
    g_arc = gpu.contacts["arc_length"]
    g_central = gpu.contacts["central"].astype(np.uint8)
    r_area, r_arc, r_central = ref.contact_area, ref.contact_arc, ref.contact_central
    if not ordered:
        go = np.lexsort((g_ids[:, 1], g_ids[:, 0]))
        ro = np.lexsort((r_ids[:, 1], r_ids[:, 0]))
        g_ids, g_area, g_arc, g_central = g_ids[go], g_area[go], g_arc[go], g_central[go]
        r_ids, r_area, r_arc, r_central = r_ids[ro], r_area[ro], r_arc[ro], r_central[ro]
    assert len(g_ids) == len(r_ids), f"contact count {len(g_ids)} != {len(r_ids)}"
    if len(g_ids):
        mism = np.flatnonzero(np.any(g_ids != r_ids, axis=1))
        assert len(mism) == 0, f"contact pair sequence differs at {int(mism[0])}: {g_ids[mism[0]]} vs {r_ids[mism[0]]}"
        cm = np.flatnonzero(g_central != r_central)
        assert len(cm) == 0, f"central flag differs at {int(cm[0])} ({r_ids[cm[0]]})"
    n = len(balls)
    ia = (r_ids[:, 0] % max(n, 1)).astype(np.int64) if len(r_ids) else np.zeros(0, np.int64)
    ib = (r_ids[:, 1] % max(n, 1)).astype(np.int64) if len(r_ids) else np.zeros(0, np.int64)
    rmax = np.maximum(R[ia], R[ib]) if len(r_ids) else np.zeros(0)
    stats = {"contacts": len(r_ids)}
    stats["area_floor"] = _close(g_area, r_area, math.pi * rmax ** 2, "contact area")
    stats["arc_floor"] = _close(g_arc, r_arc, 2 * math.pi * rmax, "contact arc length")
    if check_cells:
        assert len(gpu.cells) == len(ref.cell_index), f"cell count {len(gpu.cells)} != {len(ref.cell_index)}"
        assert np.array_equal(gpu.cells["index"].astype(np.uint64), ref.cell_index.astype(np.uint64)), "cell indices"
        Rc = R[ref.cell_index.astype(np.int64)]
        stats["cells"] = len(ref.cell_index)
        stats["sas_floor"] = _close(gpu.cells["sas_area"], ref.cell_sas, 4 * math.pi * Rc ** 2, "cell SAS")
        stats["volume_floor"] = _close(gpu.cells["volume"], ref.cell_volume, 4.0 / 3.0 * math.pi * Rc ** 3,
                                       "cell volume")
    return stats
