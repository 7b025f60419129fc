"""The oracle against every golden value the reference's tests hold.

Sources: tests/cli.rs:9-70, src/tessellation.rs:667-1511, tests/contacts_only.rs,
tests/centrality_golden.rs (+ data/central_golden.tsv), tests/pdb_mmcif.rs:26-42,155-214
(+ data/1ctf_pdb_cells_reference.tsv). Tolerances are the reference's own.
"""
import math
import os

import numpy as np
import pytest

from conftest import CENTRAL8, DATA, RING17, corners_box, load_xyzr

NULL = np.iinfo(np.uint64).max


def contact(r, a, b):
    for k in range(len(r.contact_ids)):
        if tuple(int(x) for x in r.contact_ids[k]) == (a, b):
            return r.contact_area[k], r.contact_arc[k]
    raise AssertionError(f"contact ({a}, {b}) missing")


def cell(r, i):
    k = int(np.flatnonzero(r.cell_index == i)[0])
    return r.cell_sas[k], r.cell_volume[k]


def test_cs_1x1(oracle):
    # tests/cli.rs:9-22, tessellation.rs:887-1016
    b = load_xyzr("balls_cs_1x1.xyzr")
    r = oracle.compute_tessellation(b, 2.0)
    assert len(r.contact_area) == 153 and len(r.cell_index) == 100
    assert abs(r.contact_area.sum() - 3992.55) < 0.01
    assert abs(r.cell_sas.sum() - 21979.64) < 0.01
    assert abs(r.cell_volume.sum() - 46419.87) < 0.01
    area, arc = contact(r, 0, 1)
    assert abs(area - 42.9555) < 0.01 and abs(arc - 23.2335) < 0.01


@pytest.mark.parametrize("form", ["corners", "vectors"])
def test_cs_1x1_periodic(oracle, form):
    # tests/cli.rs:25-61
    b = load_xyzr("balls_cs_1x1.xyzr")
    box = corners_box((0, 0, 0), (200, 250, 300)) if form == "corners" else [200, 0, 0, 0, 250, 0, 0, 0, 300]
    r = oracle.compute_tessellation(b, 2.0, pbox=box)
    assert len(r.contact_area) == 189 and len(r.cell_index) == 100
    assert abs(r.contact_area.sum() - 4812.14) < 0.01
    assert abs(r.cell_sas.sum() - 20023.06) < 0.01
    assert abs(r.cell_volume.sum() - 45173.20) < 0.01


def test_2zsk_counts(oracle):
    # tests/cli.rs:64-70
    r = oracle.compute_tessellation(load_xyzr("balls_2zsk.xyzr"), 1.4)
    assert len(r.contact_area) == 23855 and len(r.cell_index) == 3545


def test_ring17(oracle):
    # tessellation.rs:742-797
    r = oracle.compute_tessellation(RING17, 1.0)
    assert len(r.contact_area) == 44 and len(r.cell_index) == 17
    s0, v0 = cell(r, 0)
    s1, v1 = cell(r, 1)
    assert abs(s0 - 34.8168) < 0.01 and abs(v0 - 29.2302) < 0.01
    assert abs(s1 - 3.29195) < 0.01 and abs(v1 - 2.48022) < 0.01
    a01, arc01 = contact(r, 0, 1)
    assert abs(a01 - 0.747721) < 0.001 and abs(arc01 - 0.726907) < 0.001
    a12, _ = contact(r, 1, 2)
    assert abs(a12 - 5.0216) < 0.01


def test_ring17_periodic(oracle):
    # tessellation.rs:1021-1072 (C++ gives 64 contacts)
    r = oracle.compute_tessellation(RING17, 1.0, pbox=corners_box((-1.6, -1.6, -0.6), (1.6, 1.6, 3.1)))
    assert len(r.contact_area) == 64 and len(r.cell_index) == 17
    s0, v0 = cell(r, 0)
    assert abs(s0 - 3.56578) < 0.1 and abs(v0 - 20.2781) < 0.1
    a00, _ = contact(r, 0, 0)
    assert abs(a00 - 3.36258) < 0.01


def test_cell_vertices_goldens(oracle):
    # tessellation.rs:1350-1387 and 1392-1511
    r = oracle.compute_tessellation(RING17, 1.0, with_vertices=True)
    on_sas = int((r.vertex_ids[:, 3] == NULL).sum())
    assert len(r.vertex_ids) == 68 and on_sas == 32
    r = oracle.compute_tessellation(load_xyzr("balls_cs_1x1.xyzr"), 2.0, with_vertices=True)
    assert len(r.vertex_ids) == 65 and len(r.edge_ids) == 215
    assert (r.vertex_ids[:, 3] == NULL).sum() > len(r.vertex_ids) / 2
    # full circle: one SAS edge, no vertex (tessellation.rs:1327-1346)
    r = oracle.compute_tessellation(np.array([[0, 0, 0, 1.5], [2.5, 0, 0, 1.5]], float), 0.5, with_vertices=True)
    assert len(r.vertex_ids) == 0 and len(r.edge_ids) == 1 and r.edge_ids[0, 2] == NULL


def test_centrality_golden(oracle):
    # tests/centrality_golden.rs
    r = oracle.compute_tessellation(CENTRAL8, 1.4)
    got = {tuple(sorted(int(x) for x in ids)): bool(c) for ids, c in zip(r.contact_ids, r.contact_central)}
    want = {}
    with open(os.path.join(DATA, "central_golden.tsv")) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            a, b, c = line.split()
            want[tuple(sorted((int(a), int(b))))] = c == "1"
    assert got == want


def test_contacts_only_two_body(oracle):
    # tests/contacts_only.rs:163-181
    b = np.array([[0, 0, 0, 1.5], [1, 0, 0, 1.5], [4, 0, 0, 2.0], [5, 0, 0, 2.0]], float)
    r = oracle.compute_contacts_only(b, 1.4, groups=[0, 0, 1, 1])
    assert [tuple(int(x) for x in r.contact_ids[0])] == [(1, 2)]
    assert abs(r.contact_area[0] - 23.43431770037136) < 1e-6


def test_contacts_only_matches_full(oracle):
    # tests/contacts_only.rs:4-87: same pairs, area/arc within 1e-10, same central
    for balls in (np.array([[0, 0, 0, 1.5], [2.5, 0, 0, 1.5], [1.25, 2.0, 0, 1.5]], float), CENTRAL8):
        f = oracle.compute_tessellation(balls, 1.4)
        c = oracle.compute_contacts_only(balls, 1.4)
        assert np.array_equal(f.contact_ids, c.contact_ids)
        assert np.all(np.abs(f.contact_area - c.contact_area) < 1e-10)
        assert np.all(np.abs(f.contact_arc - c.contact_arc) < 1e-10)
        assert np.array_equal(f.contact_central, c.contact_central)
    assert not np.all(oracle.compute_contacts_only(CENTRAL8, 1.4).contact_central)


def test_contacts_only_groups_and_periodic(oracle):
    # tests/contacts_only.rs:89-143
    b = np.array([[0, 0, 0, 1.5], [1, 0, 0, 1.5], [3, 0, 0, 1.5]], float)
    g = [0, 0, 1]
    r = oracle.compute_contacts_only(b, 1.4, groups=g)
    assert len(r.contact_ids) > 0
    assert all(g[int(x)] != g[int(y)] for x, y in r.contact_ids)
    b = np.array([[0, 0, 0, 2.0], [9, 0, 0, 2.0]], float)
    box = corners_box((-5, -5, -5), (5, 5, 5))
    f = oracle.compute_tessellation(b, 1.4, pbox=box)
    c = oracle.compute_contacts_only(b, 1.4, pbox=box)
    assert len(f.contact_ids) == len(c.contact_ids)
    assert np.all(c.contact_ids < 2)
    fo = np.lexsort((f.contact_ids[:, 1], f.contact_ids[:, 0]))
    co = np.lexsort((c.contact_ids[:, 1], c.contact_ids[:, 0]))
    assert np.all(np.abs(f.contact_area[fo] - c.contact_area[co]) < 1e-10)
    assert len(oracle.compute_contacts_only(np.zeros((0, 4)), 1.4).contact_ids) == 0
    assert len(oracle.compute_contacts_only(np.array([[0, 0, 0, 1.0], [100, 0, 0, 1.0]]), 1.4).contact_ids) == 0


def test_1ctf_per_atom(oracle):
    # tests/pdb_mmcif.rs:26-42 and 155-214 (balls converted by golden/make_1ctf_xyzr.py)
    b = load_xyzr("1ctf_pdb.xyzr")
    assert len(b) == 492
    r = oracle.compute_tessellation(b, 1.4)
    assert len(r.contact_area) == 3078 and len(r.cell_index) == 492
    assert abs(r.cell_sas.sum() - 4097.64) < 0.01
    assert abs(r.contact_area.sum() - 10663.91) < 0.01
    ref = np.genfromtxt(os.path.join(DATA, "1ctf_pdb_cells_reference.tsv"), skip_header=1, usecols=(5, 6, 7))
    assert np.array_equal(r.cell_index.astype(np.int64), ref[:, 0].astype(np.int64))
    assert np.max(np.abs(r.cell_sas - ref[:, 1])) < 0.01
    assert np.max(np.abs(r.cell_volume - ref[:, 2])) < 0.01


def test_reference_unit_cases(oracle):
    # tessellation.rs:667-884
    r = oracle.compute_tessellation(np.array([[0, 0, 0, 1.0], [2, 0, 0, 1.0]], float), 0.5)
    assert [tuple(int(x) for x in r.contact_ids[0])] == [(0, 1)] and len(r.cell_index) == 2
    r = oracle.compute_tessellation(np.array([[0, 0, 0, 1.0]], float), 0.5)
    assert len(r.contact_ids) == 0 and len(r.cell_index) == 1
    assert abs(r.cell_sas[0] - 4 * math.pi * 1.5 ** 2) < 1e-12
    r = oracle.compute_tessellation(np.array([[0, 0, 0, 1], [2, 0, 0, 1], [1, 1.7, 0, 1]], float), 0.5)
    assert len(r.contact_ids) == 3 and len(r.cell_index) == 3
    r = oracle.compute_tessellation(np.array([[0, 0, 0, 1], [10, 0, 0, 1]], float), 0.5)
    assert len(r.contact_ids) == 0 and len(r.cell_index) == 2
    assert oracle.compute_tessellation(np.zeros((0, 4)), 0.5).num_balls == 0
    for balls, probe in [([[0, 0, 0, 1], [0.5, 0, 0, 1], [1, 0, 0, 1]], 1.0),
                         ([[0, 0, 0, 1], [0, 0, 1, 1], [0, 1, 0, 1], [0, 1, 1, 1]], 2.0),
                         ([[x, y, z, 1] for x in (0, 1) for y in (0, 1) for z in (0, 1)], 2.0)]:
        r = oracle.compute_tessellation(np.array(balls, float), probe)
        assert len(r.cell_index) == len(balls) and len(r.contact_ids) > 0
    r = oracle.compute_tessellation(np.array([[0, 0, 0, 1], [0, 0, 0, 0.5], [1, 0, 0, 1]], float), 0.5)
    assert len(r.cell_index) <= 3


def test_grouping(oracle):
    # tessellation.rs:1212-1283
    b = np.array([[0, 0, 0, 1], [2, 0, 0, 1], [4, 0, 0, 1]], float)
    assert len(oracle.compute_tessellation(b, 0.5).contact_ids) == 2
    assert len(oracle.compute_tessellation(b, 0.5, groups=[0, 0, 0]).contact_ids) == 0
    r = oracle.compute_tessellation(b, 0.5, groups=[0, 0, 1])
    assert [tuple(int(x) for x in c) for c in r.contact_ids] == [(1, 2)]
    b4 = np.array([[0, 0, 0, 1], [2, 0, 0, 1], [4, 0, 0, 1], [6, 0, 0, 1]], float)
    r = oracle.compute_tessellation(b4, 0.5, groups=[0, 0, 1, 1])
    assert [tuple(int(x) for x in c) for c in r.contact_ids] == [(1, 2)] and len(r.cell_index) == 2
    b3 = np.array([[0, 0, 2, 1], [0, 1, 0, 0.5], [0.382683, 0.92388, 0, 0.5]], float)
    box = corners_box((-2, -2, -1), (2, 2, 4))
    assert len(oracle.compute_tessellation(b3, 1.0, pbox=box, groups=[0, 0, 0]).contact_ids) < \
        len(oracle.compute_tessellation(b3, 1.0, pbox=box).contact_ids)
