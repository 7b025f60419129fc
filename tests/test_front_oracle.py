"""The oracle's restatement of the file front end (oracle/front.py) against the
reference's own tests: the unit tests of src/input/{pdb,mmcif,radii,selection,
mod}.rs and src/graphics.rs, and the golden counts and totals of
tests/pdb_mmcif.rs (parsing + radii + groupings, tessellated by the oracle)."""
import math
import os

import numpy as np
import pytest

from conftest import DATA
from selection_kat import ERRORS, GLOB, GROUPING, MATCH

import front  # noqa: E402  (oracle/ is on sys.path via conftest)

SAMPLE_PDB = """\
ATOM      1  N   MET A   1       0.000   0.000   0.000  1.00  0.00           N
ATOM      2  CA  MET A   1       1.458   0.000   0.000  1.00  0.00           C
ATOM      3  C   MET A   1       2.009   1.420   0.000  1.00  0.00           C
HETATM   99  ZN  ZN  A 100       5.000   5.000   5.000  1.00  0.00          ZN
END
"""

SAMPLE_CIF = """
data_test
loop_
_atom_site.group_PDB
_atom_site.id
_atom_site.auth_atom_id
_atom_site.auth_comp_id
_atom_site.auth_asym_id
_atom_site.auth_seq_id
_atom_site.Cartn_x
_atom_site.Cartn_y
_atom_site.Cartn_z
_atom_site.type_symbol
_atom_site.label_alt_id
_atom_site.pdbx_PDB_model_num
ATOM 1 N MET A 1 0.000 0.000 0.000 N . 1
ATOM 2 CA MET A 1 1.458 0.000 0.000 C . 1
HETATM 99 ZN ZN A 100 5.000 5.000 5.000 ZN . 1
"""

QUOTED_CIF = SAMPLE_CIF.split("ATOM 1")[0] + "ATOM 1 \"C'\" ALA A 1 1.0 2.0 3.0 C . 1\n"


def rec(t):
    r = front.Record(record_name=t[4] if len(t) > 4 else "ATOM", name=t[3], res_name=t[1], chain_id=t[0],
                     res_seq=t[2])
    return r


def read(name):
    with open(os.path.join(DATA, name)) as f:
        return f.read()


# ------------------------------------------------------ unit tests (pdb.rs)
def test_pdb_unit():
    r = front.parse_pdb(SAMPLE_PDB)
    assert len(r) == 4 and [x.name for x in r] == ["N", "CA", "C", "ZN"] and r[0].res_name == "MET"
    r = front.parse_pdb(SAMPLE_PDB, exclude_heteroatoms=True)
    assert len(r) == 3 and r[2].name == "C"
    assert [front.normalize_atom_name(x) for x in ("1HG2", "CA", "2HD1")] == ["HG21", "CA", "HD12"]
    line = "ATOM      1  CA  ALA A   1       1.000   2.000   3.000"
    assert (front.col(line, 1, 6), front.col(line, 13, 16), front.col(line, 18, 20)) == ("ATOM", "CA", "ALA")
    assert front.parse_pdb_line(
        "ATOM      1  N   MET A   1       0.000   0.000   0.000  1.00 49.05           N").b_factor == 49.05
    assert front.parse_pdb_line(line).b_factor == 0.0


def test_mmcif_unit():
    r = front.parse_mmcif(SAMPLE_CIF)
    assert len(r) == 3 and r[0].name == "N" and r[0].res_name == "MET" and r[2].name == "ZN"
    r = front.parse_mmcif(SAMPLE_CIF, exclude_heteroatoms=True)
    assert len(r) == 2 and r[1].name == "CA"
    r = front.parse_mmcif(QUOTED_CIF)
    assert len(r) == 1 and r[0].name == "C'"


def test_xyzr_and_detection():
    assert front.parse_xyzr("1.0 2.0 3.0 4.0\n5.0 6.0 7.0 8.0\n") == [[1, 2, 3, 4], [5, 6, 7, 8]]
    assert front.parse_xyzr("a b c 1.0 2.0 3.0 4.0\n0 0 0 1\n") == [[1, 2, 3, 4], [0, 0, 0, 1]]
    for name, fmt in [("file.pdb", "pdb"), ("file.pdb1", "pdb"), ("file.ent", "pdb"), ("file.cif", "mmcif"),
                      ("file.mmcif", "mmcif"), ("file.xyzr", "xyzr")]:
        assert front.detect(name, "") == fmt


# ---------------------------------------------------- unit tests (radii.rs)
def test_radii_unit():
    R = front.Radii()
    for res, atom, r in [("ALA", "CA", 1.90), ("GLY", "C", 1.75), ("ALA", "N", 1.70), ("ALA", "O", 1.49),
                         ("ALA", "CB", 1.92), ("ARG", "NH1", 1.62), ("ARG", "NH2", 1.67), ("ARG", "CD", 1.88),
                         ("UNK", "CX", 1.80), ("ZN", "ZN", 0.74), ("CA", "CA", 1.00), ("FE", "FE", 0.61),
                         ("UNK", "XYZ", 1.8)]:
        assert abs(R.get(res, atom) - r) < 0.001, (res, atom)
    C = front.Radii(default=False)
    C.load_text("* C* 2.0\nALA CB 1.5")
    assert C.get("ALA", "CB") == 1.5 and C.get("GLY", "CA") == 2.0
    assert len(front.DEFAULT_RULES) == 132


# ------------------------------------------------ unit tests (selection.rs)
@pytest.mark.parametrize("pattern,text,expected", GLOB)
def test_glob(pattern, text, expected):
    assert front.glob_match(pattern, text) == expected


@pytest.mark.parametrize("sel,matches,rejects", MATCH)
def test_selection_matches(sel, matches, rejects):
    s = front.Selection(sel)
    assert all(s.matches(rec(t)) for t in matches)
    assert not any(s.matches(rec(t)) for t in rejects)


@pytest.mark.parametrize("sel", ERRORS)
def test_selection_errors(sel):
    with pytest.raises(front.SelectionError):
        front.Selection(sel)


@pytest.mark.parametrize("sels,records,expected", GROUPING)
def test_custom_grouping(sels, records, expected):
    g = front.custom_grouping([rec(t) for t in records], front.parse_selections(sels))
    # the reference's assertions: first match wins, unmatched atoms unique
    assert g == expected


def test_selection_count_rules():
    assert len(front.parse_selections(["chain A B", "chain L", "resname HOH"])) == 3
    for bad in (["chain A"], []):
        with pytest.raises(front.SelectionError):
            front.parse_selections(bad)


def test_groupings_unit():
    # mod.rs:624-655
    recs = [rec(("A", "ALA", 1, "CA")), rec(("A", "ALA", 2, "CA")), rec(("B", "ALA", 1, "CA")),
            rec(("A", "ALA", 3, "CA")), rec(("B", "ALA", 2, "CA"))]
    assert front.chain_grouping(recs) == [0, 0, 1, 0, 1]
    recs = [rec(("A", "ALA", 1, "CA")), rec(("A", "ALA", 1, "CA")), rec(("A", "ALA", 2, "CA")),
            rec(("B", "ALA", 1, "CA")), rec(("A", "ALA", 1, "CA"))]
    recs[4].i_code = "A"
    assert front.residue_grouping(recs) == [0, 0, 1, 2, 3]


# -------------------------------------------------- golden counts (fixtures)
def test_fixture_counts():
    pdb, cif = read("assembly_1ctf.pdb1"), read("assembly_1ctf.cif")
    assert len(front.parse_pdb(pdb)) == 492
    assert len(front.parse_pdb(pdb, exclude_heteroatoms=True)) == 487
    assert len(front.parse_pdb(pdb, as_assembly=True)) == len(front.parse_mmcif(cif)) == 984
    assert len(front.parse_mmcif(cif, exclude_heteroatoms=True)) == 974


def test_selection_counts_match_mdtraj():
    recs = front.parse_mmcif(read("assembly_1ctf.cif"), exclude_heteroatoms=True)
    for sel, n in [("protein", 974), ("backbone", 544), ("sidechain", 430), ("resname ALA", 150),
                   ("resname ALA GLY", 198), ("hydrophobic", 490), ("aromatic", 22), ("acidic", 226),
                   ("basic", 204), ("polar", 54), ("charged", 430)]:
        s = front.Selection(sel)
        assert sum(s.matches(r) for r in recs) == n, sel


def test_fixture_tessellation_goldens(oracle):
    """tests/pdb_mmcif.rs:26-303 through the oracle's tessellation."""
    pdb, cif = read("assembly_1ctf.pdb1"), read("assembly_1ctf.cif")
    R = front.Radii()
    r = oracle.compute_tessellation(front.balls_of(front.parse_pdb(pdb), R), 1.4)
    assert len(r.contact_ids) == 3078 and len(r.cell_index) == 492
    assert abs(r.cell_sas.sum() - 4097.64) < 0.01 and abs(r.contact_area.sum() - 10663.91) < 0.01
    r = oracle.compute_tessellation(front.balls_of(front.parse_mmcif(cif), R), 1.4)
    assert len(r.cell_index) == 984 and len(r.contact_ids) == 6354
    C = front.Radii(default=False)
    C.load_text(read("custom_radii.txt"))
    recs = front.parse_pdb(pdb, as_assembly=True)
    r = oracle.compute_tessellation(front.balls_of(recs, C), 1.4)
    assert len(recs) == 984 and len(r.contact_ids) == 6354
    assert abs(r.cell_sas.sum() - 7094.42) < 0.01 and abs(r.contact_area.sum() - 21423.54) < 0.01
    recs = front.parse_mmcif(cif, exclude_heteroatoms=True)
    r = oracle.compute_tessellation(front.balls_of(recs, R), 1.4, groups=front.chain_grouping(recs))
    assert len(recs) == 974 and len(r.contact_ids) == 218 and abs(r.contact_area.sum() - 513.032) < 0.001
    recs = front.parse_pdb(pdb)
    r = oracle.compute_tessellation(front.balls_of(recs, R), 1.4, groups=front.residue_grouping(recs))
    assert len(r.contact_ids) == 2013 and abs(r.contact_area.sum() - 4486.49) < 0.01
    recs = front.parse_mmcif(cif)
    g = front.custom_grouping(recs, front.parse_selections(["chain A", "chain A-2"]))
    r = oracle.compute_tessellation(front.balls_of(recs, R), 1.4, groups=g)
    assert len(r.contact_ids) < 6354


def test_per_atom_reference_cells(oracle):
    """tests/pdb_mmcif.rs:155-214: per-atom SAS and volume of the PDB fixture
    against the upstream C++ TSV at 0.01."""
    recs = front.parse_pdb(read("assembly_1ctf.pdb1"))
    r = oracle.compute_tessellation(front.balls_of(recs, front.Radii()), 1.4)
    rows = [l.split("\t") for l in read("1ctf_pdb_cells_reference.tsv").splitlines()[1:]]
    ref = [(int(p[5]), float(p[6]), float(p[7])) for p in rows if len(p) >= 8]
    assert len(ref) == len(r.cell_index)
    for (i, sas, vol), ci, cs, cv in zip(ref, r.cell_index, r.cell_sas, r.cell_volume):
        assert int(ci) == i and abs(cs - sas) < 0.01 and abs(cv - vol) < 0.01


# --------------------------------------------------- graphics.rs unit tests
def test_graphics_unit(oracle):
    poly = {"a": 0, "ic_center": (0.0, 0.0, 0.0), "ic_r": 1.0, "axis": (0.0, 0.0, 1.0), "contour": []}
    pts = front.polygon(poly, math.pi / 6)
    assert len(pts) == 12 and all(abs(math.sqrt(front._dot(p, p)) - 1.0) < 1e-9 for p in pts)
    s = front.pymol_script([[0.0, 0.0, 0.0, 1.0]], [], 0.5, "test")
    assert "from pymol.cgo import *" in s and "SPHERE" in s and "test_balls" in s
    line = [[0, 0, 0, 1.0], [0.5, 0, 0, 1.0], [1, 0, 0, 1.0]]
    s = front.pymol_script(line, oracle.contact_polygons(line, 1.0))
    # tests/cli.rs:144-152
    assert s.count("SPHERE,") == 3 and s.count("TRIANGLE_FAN") == 2 and s.count("LINE_LOOP") == 2
    from conftest import RING17
    s = front.pymol_script(RING17, oracle.contact_polygons(RING17, 1.0))
    assert s.count("SPHERE,") == 17 and s.count("TRIANGLE_FAN") > 0


def test_rotation_and_normal_pins():
    """rotate() is a rotation (norm kept, axis fixed, angle recovered) and
    any_normal() is a unit vector orthogonal to its argument."""
    rng = np.random.default_rng(2)
    for _ in range(50):
        axis = tuple(rng.normal(size=3))
        p = tuple(rng.normal(size=3))
        ang = float(rng.uniform(0.1, 3.0))
        q = front.rotate(axis, ang, p)
        assert abs(front._dot(q, q) - front._dot(p, p)) < 1e-12
        u = front._unit(axis)
        assert abs(front._dot(q, u) - front._dot(p, u)) < 1e-12
        pp = front._sub(p, front._mul(u, front._dot(p, u)))
        qq = front._sub(q, front._mul(u, front._dot(q, u)))
        assert abs(front.directed_angle((0, 0, 0), pp, qq, u) - ang) < 1e-9
        n = front.any_normal(axis)
        assert abs(front._dot(n, n) - 1) < 1e-12 and abs(front._dot(n, axis)) < 1e-12
