"""File front end on the GPU: compute_tessellation_from_file and the CLI
against the reference's golden counts and totals (tests/pdb_mmcif.rs,
tests/cli.rs), the CLI's JSON document against the library's own results,
and the PyMOL script against the oracle's restatement of graphics.rs."""
import json
import math
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import DATA, PKG, RING17, corners_box, load_xyzr

import front

pytestmark = pytest.mark.gpu

CLI = os.path.join(PKG, "voronota_ltr", "voronota-ltr")
PDB = os.path.join(DATA, "assembly_1ctf.pdb1")
CIF = os.path.join(DATA, "assembly_1ctf.cif")


def totals(res):
    return float(res.contacts["area"].sum()), float(res.cells["sas_area"].sum()), float(res.cells["volume"].sum())


def test_from_file_goldens():
    """tests/pdb_mmcif.rs:356-445"""
    from voronota_ltr import _native

    print("\n[gpu] from_file 1ctf pdb", flush=True)
    r = _native.from_file(PDB, 1.4)
    assert len(r.contacts) == 3078 and len(r.cells) == 492
    assert abs(r.cells["sas_area"].sum() - 4097.64) < 0.01 and abs(r.contacts["area"].sum() - 10663.91) < 0.01
    r = _native.from_file(PDB, 1.4, with_cell_vertices=True)
    assert r.cell_vertices is not None and len(r.cell_vertices) > 0 and len(r.cell_edges) > 0
    r = _native.from_file(CIF, 1.4)
    assert len(r.cells) == 984 and len(r.contacts) == 6354
    r = _native.from_file(CIF, 1.4, selections=["chain A", "chain A-2"])
    assert 0 < len(r.contacts) < 6354
    with pytest.raises(OSError, match="at least two selections required for grouping"):
        _native.from_file(PDB, 1.4, selections=["protein"])
    with pytest.raises(OSError, match="at least two selections required for grouping"):
        _native.from_file("nonexistent.pdb", 1.4, selections=[])
    with pytest.raises(OSError, match="selections require PDB or mmCIF input"):
        _native.from_file(os.path.join(DATA, "balls_cs_1x1.xyzr"), 1.4, selections=["protein", "hetatm"])
    with pytest.raises(OSError, match=r"I/O error: No such file or directory \(os error 2\)"):
        _native.from_file("nonexistent.pdb", 1.4)
    with pytest.raises(OSError, match="selection error: Unknown keyword: bogus at position 1"):
        _native.from_file(CIF, 1.4, selections=["protein", "bogus"])


def test_from_file_python_api(oracle):
    """The dict of python.rs:292-324 equals compute_tessellation on the oracle
    parser's balls within the parity bar."""
    import voronota_ltr

    d = voronota_ltr.compute_tessellation_from_file(CIF, 1.4, group_selections=["chain A", "chain A-2"])
    recs = front.parse_mmcif(open(CIF).read())
    g = front.custom_grouping(recs, front.parse_selections(["chain A", "chain A-2"]))
    ref = oracle.compute_tessellation(front.balls_of(recs, front.Radii()), 1.4, groups=g)
    assert [(c["id_a"], c["id_b"]) for c in d["contacts"]] == [tuple(int(x) for x in i) for i in ref.contact_ids]
    assert d["total_contact_area"] == pytest.approx(float(ref.contact_area.sum()), rel=1e-9)
    with pytest.raises(OSError):
        voronota_ltr.compute_tessellation_from_file(PDB, 1.4, group_selections=["protein"])


def test_parsed_goldens():
    """tests/pdb_mmcif.rs:217-303: custom radii, inter-chain, inter-residue."""
    from voronota_ltr import _native

    R = _native.Radii(default=False)
    R.load_text(open(os.path.join(DATA, "custom_radii.txt")).read())
    st = _native.parse_file(PDB, as_assembly=True, radii=R)
    r = _native.run(st.balls, 1.4)
    assert len(st.balls) == 984 and len(r.contacts) == 6354
    assert abs(r.cells["sas_area"].sum() - 7094.42) < 0.01 and abs(r.contacts["area"].sum() - 21423.54) < 0.01
    st = _native.parse_file(CIF, exclude_heteroatoms=True)
    r = _native.run(st.balls, 1.4, groups=st.chain_grouping())
    assert len(st.balls) == 974 and len(r.contacts) == 218 and abs(r.contacts["area"].sum() - 513.032) < 0.001
    st = _native.parse_file(PDB)
    r = _native.run(st.balls, 1.4, groups=st.residue_grouping())
    assert len(r.contacts) == 2013 and abs(r.contacts["area"].sum() - 4486.49) < 0.01


def run_cli(args, stdin=None):
    print(f"[gpu] cli {' '.join(args)}", flush=True)
    out = subprocess.run([CLI] + args, input=stdin, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout), out.stdout


def cli_totals(d):
    return (sum(c["area"] for c in d["contacts"]), sum(c["sas_area"] for c in d["cells"]),
            sum(c["volume"] for c in d["cells"]))


def test_cli_goldens():
    """tests/cli.rs:9-97"""
    cs = os.path.join(DATA, "balls_cs_1x1.xyzr")
    d, _ = run_cli([cs, "--probe", "2.0", "-q"])
    assert len(d["contacts"]) == 153 and len(d["cells"]) == 100
    for got, want in zip(cli_totals(d), (3992.55, 21979.64, 46419.87)):
        assert abs(got - want) < 0.01
    for box in (["--periodic-box-corners", "0", "0", "0", "200", "250", "300"],
                ["--periodic-box-directions", "200", "0", "0", "0", "250", "0", "0", "0", "300"]):
        d, _ = run_cli([cs, "--probe", "2.0"] + box + ["-q"])
        assert len(d["contacts"]) == 189 and len(d["cells"]) == 100
        for got, want in zip(cli_totals(d), (4812.14, 20023.06, 45173.20)):
            assert abs(got - want) < 0.01
    d, _ = run_cli([os.path.join(DATA, "balls_2zsk.xyzr"), "--probe", "1.4", "-q"])
    assert len(d["contacts"]) == 23855 and len(d["cells"]) == 3545
    d, _ = run_cli(["--probe", "1.0", "-q"], stdin="0 0 0 1\n0.5 0 0 1\n1 0 0 1\n")
    assert len(d["cells"]) == 3
    d, _ = run_cli([PDB, "--inter-residue-only", "-q"])
    assert len(d["contacts"]) == 2013
    d, _ = run_cli([CIF, "--exclude-heteroatoms", "--inter-chain-only", "-q"])
    assert len(d["contacts"]) == 218
    d, _ = run_cli([PDB, "--radii-file", os.path.join(DATA, "custom_radii.txt"), "-q"])
    assert d["num_balls"] == 492


NUM = re.compile(r"(?<![\w.])-?\d[\d.]*(?:e-?\d+)?")


def test_cli_json_document():
    """The document of main.rs:22-34: the result's fields, per-ball SAS and
    volume (null without a cell), totals; every number is the shortest text
    that reads back to the library's value."""
    from voronota_ltr import _native

    balls = load_xyzr("balls_cs_1x1.xyzr")
    d, text = run_cli([os.path.join(DATA, "balls_cs_1x1.xyzr"), "--probe", "2.0", "-q"])
    assert list(d.keys()) == ["num_balls", "contacts", "cells", "cell_vertices", "cell_edges", "sas_areas",
                              "volumes", "total_sas_area", "total_volume", "total_contact_area"]
    r = _native.run(balls, 2.0)
    assert d["num_balls"] == len(balls) and d["cell_vertices"] is None and d["cell_edges"] is None
    assert [c["area"] for c in d["contacts"]] == list(r.contacts["area"])
    assert [c["central"] for c in d["contacts"]] == [bool(x) for x in r.contacts["central"]]
    assert [c["volume"] for c in d["cells"]] == list(r.cells["volume"])
    sas = r.sas_areas()
    assert [x is None for x in d["sas_areas"]] == list(np.isnan(sas))
    assert d["total_sas_area"] == sum(float(x) for x in r.cells["sas_area"])
    assert text.startswith('{\n  "num_balls": 100,\n  "contacts": [\n    {\n      "id_a": ')
    for tok in NUM.findall(text):
        if "." in tok or "e" in tok:
            v = float(tok)
            digits = lambda s: s.lstrip("-").split("e")[0].replace(".", "").strip("0")
            assert digits(tok) == digits(repr(v)), tok
    # with the network: vertices and edges as in compute_tessellation
    j = json.loads(_native.run_json(RING17, 1.0, with_cell_vertices=True))
    rv = _native.run(RING17, 1.0, with_cell_vertices=True)
    assert len(j["cell_vertices"]) == len(rv.cell_vertices) == 68
    assert [v["ball_indices"][3] is None for v in j["cell_vertices"]].count(True) == 32


def _script_numbers(s):
    return [float(x) for x in re.findall(r"-?\d+\.\d{6}", s)]


@pytest.mark.parametrize("case", ["line", "ring17", "1ctf", "1ctf_groups"])
def test_pymol_graphics_vs_oracle(oracle, case):
    from voronota_ltr import _native

    groups = None
    if case == "line":
        balls, probe = np.array([[0, 0, 0, 1.0], [0.5, 0, 0, 1.0], [1, 0, 0, 1.0]]), 1.0
    elif case == "ring17":
        balls, probe = RING17, 1.0
    else:
        st = _native.parse_file(PDB)
        balls, probe = st.balls[:200], 1.4
        if case == "1ctf_groups":
            groups = st.residue_grouping()[:200]
    print(f"\n[gpu] pymol {case}", flush=True)
    got = _native.pymol_graphics(balls, probe, groups)
    want = front.pymol_script(balls, oracle.contact_polygons(balls, probe, groups))
    # same structure line by line, numbers to 1e-6 (printed with 6 decimals)
    gl, wl = got.splitlines(), want.splitlines()
    assert len(gl) == len(wl)
    assert [re.sub(r"-?\d+\.\d{6}", "#", x) for x in gl] == [re.sub(r"-?\d+\.\d{6}", "#", x) for x in wl]
    a, b = np.array(_script_numbers(got)), np.array(_script_numbers(want))
    assert np.max(np.abs(a - b)) <= 1.01e-6
    if case == "line":
        assert got.count("SPHERE,") == 3 and got.count("TRIANGLE_FAN") == 2 and got.count("LINE_LOOP") == 2


def test_cli_pymol(tmp_path):
    """tests/cli.rs:112-166"""
    p = tmp_path / "g.py"
    run_cli(["--probe", "1.0", "-q", "--pymol", str(p)], stdin="0 0 0 1\n0.5 0 0 1\n1 0 0 1\n")
    s = p.read_text()
    assert "from pymol.cgo import *" in s
    assert s.count("SPHERE,") == 3 and s.count("TRIANGLE_FAN") == 2 and s.count("LINE_LOOP") == 2
    ring = "".join(" ".join(str(v) for v in b) + "\n" for b in RING17)
    run_cli(["--probe", "1.0", "-q", "--pymol", str(p)], stdin=ring)
    s = p.read_text()
    assert s.count("SPHERE,") == 17 and s.count("TRIANGLE_FAN") > 0
