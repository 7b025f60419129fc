"""Writes tests/golden/data/1ctf_pdb.xyzr from the reference's 1ctf PDB fixture.

Run in the build container only (it reads /root/reference); the output is a
committed data fixture so nothing at test time needs the reference tree.

Restates just enough of the reference's input layer to turn the PDB into balls:
  * PDB parsing, src/input/pdb.rs:68-171 (first model only, ATOM/HETATM,
    altLoc in {"", A, 1, .}, no hydrogens, no HOH, atom-name digit rotation);
  * radius lookup, src/input/radii.rs:81-110 (exact, then longest prefix
    pattern, specific residue before "*", then residue wildcard, default 1.8)
    with the default rule table of src/input/radii.rs:113-270.
The per-atom expected SAS/volume live in 1ctf_pdb_cells_reference.tsv, which
the reference's own test compares at 0.01 (tests/pdb_mmcif.rs:155-214).
"""
import os
import sys

SRC = "/root/reference/tests/data/assembly_1ctf.pdb1"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "1ctf_pdb.xyzr")

# (residue, atom pattern, radius): the default rule table of radii.rs.
RULES = [
    ("*", "C*", 1.80), ("*", "N*", 1.60), ("*", "O*", 1.50), ("*", "P*", 1.90), ("*", "S*", 1.90),
    ("*", "H*", 1.30), ("*", "C", 1.75), ("*", "CA", 1.90), ("*", "N", 1.70), ("*", "O", 1.49),
    ("ALA", "CB", 1.92),
    ("ARG", "CB", 1.91), ("ARG", "CD*", 1.88), ("ARG", "CG*", 1.92), ("ARG", "CZ*", 1.80), ("ARG", "NE*", 1.62),
    ("ARG", "NH1", 1.62), ("ARG", "NH2", 1.67),
    ("ASN", "CB", 1.91), ("ASN", "CG*", 1.81), ("ASN", "ND2", 1.62), ("ASN", "OD1", 1.52),
    ("ASP", "CB", 1.91), ("ASP", "CG*", 1.76), ("ASP", "OD1", 1.49), ("ASP", "OD2", 1.49),
    ("CYS", "CB", 1.91), ("CYS", "S*", 1.88),
    ("GLN", "CB", 1.91), ("GLN", "CD*", 1.81), ("GLN", "CG*", 1.80), ("GLN", "NE2", 1.62), ("GLN", "OE1", 1.52),
    ("GLU", "CB", 1.91), ("GLU", "CD*", 1.76), ("GLU", "CG*", 1.88), ("GLU", "OE1", 1.49), ("GLU", "OE2", 1.49),
    ("HIS", "CB", 1.91), ("HIS", "CD*", 1.74), ("HIS", "CE*", 1.74), ("HIS", "CG*", 1.80), ("HIS", "ND1", 1.60),
    ("HIS", "NE2", 1.60),
    ("ILE", "CB", 2.01), ("ILE", "CD1", 1.92), ("ILE", "CG1", 1.92), ("ILE", "CG2", 1.92),
    ("LEU", "CB", 1.91), ("LEU", "CD1", 1.92), ("LEU", "CD2", 1.92), ("LEU", "CG*", 2.01),
    ("LYS", "CB", 1.91), ("LYS", "CD*", 1.92), ("LYS", "CE*", 1.88), ("LYS", "CG*", 1.92), ("LYS", "NZ*", 1.67),
    ("MET", "CB", 1.91), ("MET", "CE*", 1.80), ("MET", "CG*", 1.92), ("MET", "S*", 1.94),
    ("PHE", "CB", 1.91), ("PHE", "CD*", 1.82), ("PHE", "CE*", 1.82), ("PHE", "CG*", 1.74), ("PHE", "CZ*", 1.82),
    ("PRO", "CB", 1.91), ("PRO", "CD*", 1.92), ("PRO", "CG*", 1.92),
    ("SER", "CB", 1.91), ("SER", "OG*", 1.54),
    ("THR", "CB", 2.01), ("THR", "CG2", 1.92), ("THR", "OG*", 1.54),
    ("TRP", "CB", 1.91), ("TRP", "CD*", 1.82), ("TRP", "CE*", 1.82), ("TRP", "CE2", 1.74), ("TRP", "CG*", 1.74),
    ("TRP", "CH*", 1.82), ("TRP", "CZ*", 1.82), ("TRP", "NE1", 1.66),
    ("TYR", "CB", 1.91), ("TYR", "CD*", 1.82), ("TYR", "CE*", 1.82), ("TYR", "CG*", 1.74), ("TYR", "CZ*", 1.80),
    ("TYR", "OH*", 1.54),
    ("VAL", "CB", 2.01), ("VAL", "CG1", 1.92), ("VAL", "CG2", 1.92),
    ("*", "F*", 1.33), ("*", "CL*", 1.81), ("*", "BR*", 1.96), ("*", "I*", 2.20),
]
TABLE = {(r, a): v for r, a, v in RULES}


def radius(res, atom):
    if atom in ("", "*"):
        return 1.8
    for r in (res, "*"):
        if (r, atom) in TABLE:
            return TABLE[(r, atom)]
        for k in range(len(atom), 0, -1):
            if (r, atom[:k] + "*") in TABLE:
                return TABLE[(r, atom[:k] + "*")]
    for r in (res, "*"):
        if (r, "*") in TABLE:
            return TABLE[(r, "*")]
    return 1.8


def col(line, a, b):
    return line[a - 1:b].strip() if len(line) >= a else ""


def norm_name(name):
    for i, ch in enumerate(name):
        if not ch.isdigit():
            return name if i == 0 else name[i:] + name[:i]
    return name


def main():
    out = []
    for line in open(SRC):
        rec = col(line, 1, 6)
        if rec == "ENDMDL" or rec == "END":
            break
        if rec not in ("ATOM", "HETATM"):
            continue
        try:
            x, y, z = float(col(line, 31, 38)), float(col(line, 39, 46)), float(col(line, 47, 54))
            int(col(line, 7, 11))
            rnum = int(col(line, 23, 26))
        except ValueError:
            continue
        raw = col(line, 13, 16)
        if not raw:
            continue
        name = norm_name(raw)
        alt = col(line, 17, 17)
        res = col(line, 18, 20)
        chain = col(line, 22, 22)
        elem = col(line, 77, 78)
        if alt not in ("", "A", "1", "."):
            continue
        if name.startswith("H") or elem in ("H", "D"):
            continue
        if res == "HOH":
            continue
        out.append(f"{chain} {rnum} {res} {name} {x!r} {y!r} {z!r} {radius(res, name)!r}\n")
    with open(OUT, "w") as f:
        f.writelines(out)
    print(f"wrote {len(out)} balls to {OUT}")


if __name__ == "__main__":
    sys.exit(main())
