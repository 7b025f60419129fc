"""Tessellation network output (with_cell_vertices, contact.rs:120-177,
tessellation.rs:334-401) through the C ABI against the oracle: vertex and edge
id tuples exact and in the same order, coordinates and lengths to the parity
bar; the reference's counts for ring-17 (68 vertices) and cs_1x1 (65 / 215)."""
import numpy as np
import pytest

from conftest import RING17, corners_box, load_xyzr

pytestmark = pytest.mark.gpu


def _check(oracle, balls, probe, pbox=None, name=""):
    import voronota_ltr

    print(f"\n[gpu] network {name}: n={len(balls)}", flush=True)
    g = voronota_ltr.tessellate(balls, probe, periodic_box=pbox, with_cell_vertices=True)
    r = oracle.compute_tessellation(balls, probe, pbox=pbox, with_vertices=True)
    assert g.cell_vertices is not None and g.cell_edges is not None
    gv = g.cell_vertices["ball_indices"].astype(np.uint64)
    ge = g.cell_edges["ball_indices"].astype(np.uint64)
    assert gv.shape == r.vertex_ids.shape and np.array_equal(gv, r.vertex_ids.astype(np.uint64))
    assert ge.shape == r.edge_ids.shape and np.array_equal(ge, r.edge_ids.astype(np.uint64))
    gxyz = np.stack([g.cell_vertices["x"], g.cell_vertices["y"], g.cell_vertices["z"]], axis=1)
    scale = np.abs(r.vertex_xyz).max() if len(r.vertex_xyz) else 1.0
    assert np.allclose(gxyz, r.vertex_xyz, rtol=1e-9, atol=1e-12 * scale)
    assert np.allclose(g.cell_edges["length"], r.edge_length, rtol=1e-9, atol=1e-12)
    return len(gv), len(ge)


def test_ring17(oracle):
    nv, _ = _check(oracle, RING17, 1.0, name="ring17")
    assert nv == 68


def test_cs_1x1(oracle):
    nv, ne = _check(oracle, load_xyzr("balls_cs_1x1.xyzr"), 2.0, name="cs_1x1")
    assert (nv, ne) == (65, 215)


def test_two_balls(oracle):
    _check(oracle, np.array([[0, 0, 0, 1.5], [2.5, 0, 0, 1.5]], float), 0.5, name="two balls")


def test_periodic_ring17(oracle):
    _check(oracle, RING17, 1.0, corners_box((-1.6, -1.6, -0.6), (1.6, 1.6, 3.1)), name="ring17 periodic")


def test_1ctf_and_rsa(oracle):
    from voronota_ltr import workload

    _check(oracle, load_xyzr("1ctf_pdb.xyzr"), 1.4, name="1ctf")
    _check(oracle, workload.rsa_sphere(2500, seed=2500)[0], 1.4, name="rsa 2500")


def test_periodic_rsa_box(oracle):
    """A periodic box of 1,500 balls: the network comes from every contact,
    images included, before the contact dedup (tessellation.rs:474-478)."""
    from voronota_ltr import workload

    n = 1500
    L = (n / 0.056) ** (1.0 / 3.0)
    b, _ = workload.rsa_box(n, L, seed=1500)
    _check(oracle, b, 1.4, corners_box((0, 0, 0), (L, L, L)), name="rsa box periodic")
