"""Known answers of the selection language: the cases of the reference's unit
tests (src/input/selection.rs:872-1378) as data. A record is (chain, resname,
resid, name) or (chain, resname, resid, name, "HETATM")."""

GLOB = [  # (pattern, text, expected)
    ("ALA", "ALA", True), ("ALA", "ALAB", False), ("ALA", "ALA ", False), ("ALA", "ala", False),
    ("C*", "C", True), ("C*", "CA", True), ("C*", "CG1", True), ("C*", "NC", False),
    ("?A", "CA", True), ("?A", "NA", True), ("?A", "A", False), ("?A", "CAA", False),
    ("[CNO]", "C", True), ("[CNO]", "N", True), ("[CNO]", "O", True), ("[CNO]", "S", False),
    ("C[AG]*", "CA", True), ("C[AG]*", "CG", True), ("C[AG]*", "CG1", True), ("C[AG]*", "CB", False),
    ("*A", "A", True), ("*A", "CA", True), ("*A", "CGA", True), ("*A", "AB", False),
    ("C*A", "CA", True), ("C*A", "CGA", True), ("C*A", "CG", False),
]

CA = ("A", "ALA", 1, "CA")
MATCH = [  # (selection, matches, rejects)
    ("chain A", [CA], [("B", "ALA", 1, "CA")]),
    ("chain A B C", [CA, ("B", "ALA", 1, "CA"), ("C", "ALA", 1, "CA")], [("D", "ALA", 1, "CA")]),
    ("resname ALA GLY", [CA, ("A", "GLY", 1, "CA")], [("A", "VAL", 1, "CA")]),
    ("resid 10 to 20", [("A", "ALA", 10, "CA"), ("A", "ALA", 15, "CA"), ("A", "ALA", 20, "CA")],
     [("A", "ALA", 9, "CA"), ("A", "ALA", 21, "CA")]),
    ("resid 10:20", [("A", "ALA", 15, "CA")], []),
    ("name C*", [CA, ("A", "ALA", 1, "CB"), ("A", "ALA", 1, "C")], [("A", "ALA", 1, "N")]),
    ("protein", [CA, ("A", "GLY", 1, "N")], [("A", "HOH", 1, "O")]),
    ("backbone", [CA, ("A", "ALA", 1, "C"), ("A", "ALA", 1, "N"), ("A", "ALA", 1, "O")], [("A", "ALA", 1, "CB")]),
    ("sidechain", [("A", "ALA", 1, "CB"), ("A", "ALA", 1, "CG")], [CA, ("A", "HOH", 1, "O")]),
    ("hetatm", [("A", "HOH", 1, "O", "HETATM")], [CA]),
    ("all", [CA, ("A", "HOH", 1, "O", "HETATM")], []),
    ("none", [], [CA]),
    ("hydrophobic", [CA, ("A", "VAL", 1, "CA"), ("A", "PHE", 1, "CA")], [("A", "ASP", 1, "CA"), ("A", "SER", 1, "CA")]),
    ("aromatic", [("A", "PHE", 1, "CA"), ("A", "TYR", 1, "CA"), ("A", "TRP", 1, "CA"), ("A", "HIS", 1, "CA")], [CA]),
    ("acidic", [("A", "ASP", 1, "CA"), ("A", "GLU", 1, "CA")], [("A", "LYS", 1, "CA"), CA]),
    ("basic", [("A", "ARG", 1, "CA"), ("A", "LYS", 1, "CA"), ("A", "HIS", 1, "CA")], [("A", "ASP", 1, "CA"), CA]),
    ("polar", [("A", "SER", 1, "CA"), ("A", "THR", 1, "CA"), ("A", "ASN", 1, "CA"), ("A", "GLN", 1, "CA")], [CA]),
    ("charged", [("A", "ASP", 1, "CA"), ("A", "GLU", 1, "CA"), ("A", "ARG", 1, "CA"), ("A", "LYS", 1, "CA")],
     [CA, ("A", "SER", 1, "CA")]),
    ("chain A and resname ALA", [CA], [("A", "GLY", 1, "CA"), ("B", "ALA", 1, "CA")]),
    ("resname ALA or resname GLY", [CA, ("A", "GLY", 1, "CA")], [("A", "VAL", 1, "CA")]),
    ("not protein", [("A", "ZN", 1, "ZN", "HETATM")], [CA]),
    ("not resname HOH and chain A", [CA], [("A", "HOH", 1, "O"), ("B", "ALA", 1, "CA")]),
    ("resname HOH or chain A and resname ALA", [("A", "HOH", 1, "O"), CA, ("B", "HOH", 1, "O")],
     [("A", "GLY", 1, "CA")]),
    ("protein and (backbone or name CB)", [CA, ("A", "ALA", 1, "CB")], [("A", "ALA", 1, "CG"), ("A", "HOH", 1, "O")]),
    ("(chain A or chain B) and (resname ALA or resname GLY)", [CA, ("B", "GLY", 1, "CA")],
     [("C", "ALA", 1, "CA"), ("A", "VAL", 1, "CA")]),
    ("not not protein", [CA], [("A", "HOH", 1, "O")]),
    ("not (protein and hetatm)", [CA], []),
    ("everything", [CA], []),
    ("nothing", [], [CA]),
    ("segid A", [CA], []),
    ("resn ALA", [CA], []),
    ("resi 1", [CA], []),
    ("resname ALA ASP GLU", [CA, ("A", "ASP", 1, "CA"), ("A", "GLU", 1, "CA")], [("A", "VAL", 1, "CA")]),
    ("resid 100 101 102", [("A", "ALA", 100, "CA"), ("A", "ALA", 101, "CA")], [("A", "ALA", 99, "CA")]),
    ("resid 1 to 10", [CA, ("A", "ALA", 5, "CA"), ("A", "ALA", 10, "CA")], [("A", "ALA", 11, "CA")]),
    ("resid 1:10", [("A", "ALA", 5, "CA")], []),
    ("resid -5 to 5", [("A", "ALA", -3, "CA"), ("A", "ALA", 0, "CA")], [("A", "ALA", 6, "CA")]),
    ("name 'CA'", [CA], []),
    ('name "CA"', [CA], []),
    ('name "O5\'"', [("A", "DA", 1, "O5'")], []),
]

ERRORS = ["unknown_keyword foo", "chain", "(chain A and resname ALA", "or", "and", "chain A and", "not unknown",
          "protein or unknown"]

GROUPING = [  # (selections, records, expected groups)
    (["chain A", "chain L"], [CA, ("A", "ALA", 2, "CA"), ("L", "LIG", 1, "C1")], [0, 0, 1]),
    (["chain A", "chain B"], [CA, ("X", "UNK", 1, "X1"), ("Y", "UNK", 1, "Y1")], [0, 2, 3]),
    (["chain A", "resname ALA"], [CA, ("A", "GLY", 2, "CA"), ("B", "ALA", 1, "CA")], [0, 0, 1]),
]


def as_mmcif(records):
    """The records as an mmCIF _atom_site loop (names quoted)."""
    lines = ["data_kat", "loop_"] + ["_atom_site." + f for f in (
        "group_PDB", "id", "auth_atom_id", "auth_comp_id", "auth_asym_id", "auth_seq_id", "Cartn_x", "Cartn_y",
        "Cartn_z", "type_symbol", "pdbx_PDB_model_num")]
    for k, r in enumerate(records):
        group = r[4] if len(r) > 4 else "ATOM"
        lines.append(f'{group} {k + 1} "{r[3]}" {r[1]} {r[0]} {r[2]} 0.0 0.0 {k}.0 C 1')
    return "\n".join(lines) + "\n"
