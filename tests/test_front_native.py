"""The library's file front end (host C++, no GPU needed) against the oracle's
restatement: every record field and ball of the fixtures under every parse
option, radii lookups, groupings, selections (the reference's known answers
and the MDTraj counts), error messages and format detection; the CLI's
argument handling."""
import os
import subprocess

import numpy as np
import pytest

from conftest import DATA, PKG
from selection_kat import ERRORS, GROUPING, MATCH, as_mmcif

import front

FIELDS = ["record_name", "name", "alt_loc", "res_name", "chain_id", "i_code", "element", "serial", "res_seq", "x",
          "y", "z", "b_factor"]
OPTIONS = [dict(), dict(exclude_heteroatoms=True), dict(include_hydrogens=True), dict(as_assembly=True),
           dict(exclude_heteroatoms=True, include_hydrogens=True, as_assembly=True)]


@pytest.fixture(scope="module")
def native():
    from voronota_ltr import _native

    _native.load_library()
    return _native


def read(name):
    with open(os.path.join(DATA, name)) as f:
        return f.read()


def same_records(st, recs):
    assert len(st.records) == len(recs)
    for a, b in zip(st.records, recs):
        assert [a[f] for f in FIELDS] == [getattr(b, f) for f in FIELDS]


@pytest.mark.parametrize("opts", OPTIONS)
@pytest.mark.parametrize("name,fmt", [("assembly_1ctf.pdb1", "pdb"), ("assembly_1ctf.cif", "mmcif")])
def test_fixture_records_and_balls(native, name, fmt, opts):
    st = native.parse_file(os.path.join(DATA, name), **opts)
    text = read(name)
    recs = front.parse_pdb(text, **opts) if fmt == "pdb" else front.parse_mmcif(text, **opts)
    assert st.format == fmt
    same_records(st, recs)
    assert np.array_equal(st.balls, np.array(front.balls_of(recs, front.Radii())).reshape(-1, 4))


def test_custom_radii(native):
    text = read("custom_radii.txt")
    R = native.Radii(default=False)
    R.load_text(text)
    O = front.Radii(default=False)
    O.load_text(text)
    st = native.parse_file(os.path.join(DATA, "assembly_1ctf.pdb1"), as_assembly=True, radii=R)
    recs = front.parse_pdb(read("assembly_1ctf.pdb1"), as_assembly=True)
    assert np.array_equal(st.balls, np.array(front.balls_of(recs, O)))
    D = native.Radii()
    for res, atom, _ in front.DEFAULT_RULES:
        for a in (atom.rstrip("*"), atom.rstrip("*") + "X", "*", ""):
            assert D.get(res, a) == front.Radii().get(res, a), (res, a)
    for bad, msg in [("ALA CB", "Invalid radii rule: ALA CB"), ("ALA CB x1", "Invalid radius value: x1"),
                     ("ALA CB -1.5", "Negative radius: -1.5")]:
        with pytest.raises(ValueError) as e:
            native.Radii(default=False).load_text(bad)
        assert str(e.value) == msg


def test_groupings(native):
    for name, fmt in [("assembly_1ctf.pdb1", "pdb"), ("assembly_1ctf.cif", "mmcif")]:
        for opts in OPTIONS:
            st = native.parse_file(os.path.join(DATA, name), **opts)
            text = read(name)
            recs = front.parse_pdb(text, **opts) if fmt == "pdb" else front.parse_mmcif(text, **opts)
            assert list(st.chain_grouping()) == front.chain_grouping(recs)
            assert list(st.residue_grouping()) == front.residue_grouping(recs)
            for sels in (["chain A", "chain A-2"], ["protein", "hetatm"], ["backbone", "resname ALA GLY", "resid 60:80"]):
                assert list(st.selection_grouping(sels)) == front.custom_grouping(recs, front.parse_selections(sels))


def test_mdtraj_counts(native):
    st = native.parse_file(os.path.join(DATA, "assembly_1ctf.cif"), exclude_heteroatoms=True)
    for sel, n in [("protein", 974), ("backbone", 544), ("sidechain", 430), ("resname ALA", 150),
                   ("resname ALA GLY", 198), ("hydrophobic", 490), ("aromatic", 22), ("acidic", 226),
                   ("basic", 204), ("polar", 54), ("charged", 430)]:
        assert st.select(sel).sum() == n, sel


@pytest.mark.parametrize("sel,matches,rejects", MATCH)
def test_selection_known_answers(native, sel, matches, rejects):
    """Records built as mmCIF and parsed; the parser drops HOH residues, so
    those cases are the oracle's alone."""
    recs = [t for t in matches + rejects if t[1] != "HOH"]
    want = [t in matches for t in recs]
    st = native.parse_text(as_mmcif(recs))
    assert len(st.records) == len(recs)
    assert list(st.select(sel)) == want


@pytest.mark.parametrize("sels,records,expected", GROUPING)
def test_selection_grouping_known_answers(native, sels, records, expected):
    st = native.parse_text(as_mmcif(records))
    assert list(st.selection_grouping(sels)) == expected


@pytest.mark.parametrize("sel", ERRORS + ["", "chain A )", "resid 5 to", "resid 5 :", "name", "chain A # B",
                                          "resid 99999999999", "(protein", "name é", "resid x"])
def test_selection_error_messages(native, sel):
    st = native.parse_text(as_mmcif([("A", "ALA", 1, "CA")]))
    with pytest.raises(front.SelectionError) as e:
        front.Selection(sel)
    with pytest.raises(ValueError) as g:
        st.select(sel)
    assert str(g.value) == str(e.value)


def test_selection_list_errors(native):
    st = native.parse_text(as_mmcif([("A", "ALA", 1, "CA")]))
    for sels in (["protein"], [], ["protein", "chain"], ["protein", "resid 1", "bogus x"]):
        with pytest.raises(front.SelectionError) as e:
            front.parse_selections(sels)
        with pytest.raises(ValueError) as g:
            st.selection_grouping(sels)
        assert str(g.value) == str(e.value)


def test_text_detection_and_xyzr(native):
    st = native.parse_text("0 0 0 1\n0.5 0 0 1\n1 0 0 1\n")
    assert st.format == "xyzr" and st.balls.tolist() == [[0, 0, 0, 1], [0.5, 0, 0, 1], [1, 0, 0, 1]]
    st = native.parse_text("a b c 1.0 2.0 3.0 4.0\n0 0 0 1\nbad line\n", "xyzr")
    assert st.balls.tolist() == [[1, 2, 3, 4], [0, 0, 0, 1]]
    pdb = read("assembly_1ctf.pdb1")
    assert native.parse_text(pdb).format == "pdb"
    assert native.parse_text(read("assembly_1ctf.cif")).format == "mmcif"
    # the PDB fixture's first line is HEADER: content detection reads it as PDB
    same_records(native.parse_text(pdb), front.parse_pdb(pdb))


def test_file_errors(native, tmp_path):
    with pytest.raises(OSError) as e:
        native.parse_file(str(tmp_path / "missing.pdb"))
    assert str(e.value) == "No such file or directory (os error 2)"
    p = tmp_path / "noext"
    p.write_text("hello world\n")
    with pytest.raises(OSError) as e:
        native.parse_file(str(p))
    assert str(e.value) == "Unable to detect input format"
    p = tmp_path / "balls"
    p.write_text("1 2 3 4\n")
    assert native.parse_file(str(p)).format == "xyzr"


def test_unit_samples(native):
    """pdb.rs / mmcif.rs unit samples through the library."""
    from test_front_oracle import QUOTED_CIF, SAMPLE_CIF, SAMPLE_PDB

    for text, fmt in [(SAMPLE_PDB, "pdb"), (SAMPLE_CIF, "mmcif"), (QUOTED_CIF, "mmcif")]:
        for opts in OPTIONS:
            st = native.parse_text(text, fmt, **opts)
            recs = front.parse_pdb(text, **opts) if fmt == "pdb" else front.parse_mmcif(text, **opts)
            same_records(st, recs)


CLI = os.path.join(PKG, "voronota_ltr", "voronota-ltr")


def test_cli_help_and_argument_errors():
    """tests/cli.rs:99-110 and clap's usage errors (exit status 2)."""
    out = subprocess.run([CLI, "--help"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0
    for s in ("XYZR", "--probe", "--output"):
        assert s in out.stdout
    for args in (["--probe"], ["--periodic-box-corners", "0", "0", "0", "1", "1"], ["-s", "protein"],
                 ["-s", "a", "b", "--inter-chain-only"], ["--bogus"],
                 ["--periodic-box-corners", "0", "0", "0", "1", "1", "1", "--periodic-box-directions", "1", "0",
                  "0", "0", "1", "0", "0", "0", "1"]):
        assert subprocess.run([CLI] + args, capture_output=True, timeout=60).returncode == 2, args
