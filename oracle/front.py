"""CPU restatement of the reference's file front end, used only by tests/ as
the checker of the library's parsers (test infrastructure, never shipped).

Follows src/input/pdb.rs (PDB records), src/input/mmcif.rs (_atom_site
loops), src/input/mod.rs (XYZR, format detection, groupings,
compute_tessellation_from_file's checks), src/input/radii.rs (lookup rules and
the default table, kept as data) and
src/input/selection.rs (VMD-like selection language). Pinned by the
reference's golden counts and unit tests (tests/test_front_oracle.py).
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass


@dataclass
class Record:
    record_name: str = ""
    serial: int = 0
    name: str = ""
    alt_loc: str = ""
    res_name: str = ""
    chain_id: str = ""
    res_seq: int = 0
    i_code: str = ""
    x: float = 0.0
    y: float = 0.0
    z: float = 0.0
    element: str = ""
    b_factor: float = 0.0


_F64 = re.compile(r"^[+-]?(inf|infinity|nan|(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)$", re.IGNORECASE)
_I32 = re.compile(r"^[+-]?\d+$")


def parse_f64(s):
    """str::parse::<f64> (None on failure)."""
    if not _F64.match(s):
        return None
    return float(s)


def parse_i32(s):
    if not _I32.match(s):
        return None
    v = int(s)
    return v if -2**31 <= v < 2**31 else None


def fix_undefined(s):
    return "" if s in (".", "?") else s


def normalize_atom_name(name):
    """pdb.rs:46-56: leading digits move to the end."""
    k = 0
    while k < len(name) and name[k].isdigit() and name[k].isascii():
        k += 1
    if k == 0 or k == len(name):
        return name
    return name[k:] + name[:k]


def is_hydrogen(r):
    return r.name.startswith("H") or r.element in ("H", "D")


def acceptable(r, exclude_heteroatoms=False, include_hydrogens=False):
    """pdb.rs:112-139 / mmcif.rs:262-285"""
    if r.record_name == "HETATM" and exclude_heteroatoms:
        return False
    if r.alt_loc not in ("", "A", "1", "."):
        return False
    if not include_hydrogens and is_hydrogen(r):
        return False
    return r.res_name != "HOH"


def lines_of(text):
    """BufRead::lines: '\\n' separated, a trailing '\\r' dropped."""
    out = text.split("\n")
    if out and out[-1] == "":
        out.pop()
    return [l[:-1] if l.endswith("\r") else l for l in out]


# ------------------------------------------------------------------ PDB --
def col(line, a, b):
    """1-based inclusive columns, trimmed (pdb.rs:13-25)."""
    if a - 1 >= len(line):
        return ""
    return line[a - 1:min(b, len(line))].strip()


def parse_pdb_line(line):
    rec = col(line, 1, 6)
    if rec not in ("ATOM", "HETATM"):
        return None
    x, y, z = parse_f64(col(line, 31, 38)), parse_f64(col(line, 39, 46)), parse_f64(col(line, 47, 54))
    if x is None or y is None or z is None:
        return None
    serial, res_seq = parse_i32(col(line, 7, 11)), parse_i32(col(line, 23, 26))
    if serial is None or res_seq is None:
        return None
    raw = col(line, 13, 16)
    if not raw:
        return None
    b = parse_f64(col(line, 61, 66))
    return Record(rec, serial, normalize_atom_name(raw), fix_undefined(col(line, 17, 17)),
                  fix_undefined(col(line, 18, 20)), fix_undefined(col(line, 22, 22)), res_seq,
                  fix_undefined(col(line, 27, 27)), x, y, z, fix_undefined(col(line, 77, 78)),
                  b if b is not None else 0.0)


def parse_pdb(text, exclude_heteroatoms=False, include_hydrogens=False, as_assembly=False):
    """pdb.rs:142-171"""
    out, model = [], 1
    for line in lines_of(text):
        rec = col(line, 1, 6)
        if rec in ("ATOM", "HETATM"):
            r = parse_pdb_line(line)
            if r is not None and acceptable(r, exclude_heteroatoms, include_hydrogens):
                if as_assembly and model > 1:
                    r.chain_id += str(model)
                r.alt_loc = ""
                out.append(r)
        elif rec == "ENDMDL":
            model += 1
            if not as_assembly:
                break
        elif rec == "END":
            break
    return out


# ---------------------------------------------------------------- mmCIF --
_CIF_FIELDS = ["pdbx_PDB_model_num", "group_PDB", "id", "auth_atom_id", "label_atom_id", "label_alt_id",
               "auth_comp_id", "label_comp_id", "auth_asym_id", "label_asym_id", "auth_seq_id", "label_seq_id",
               "pdbx_PDB_ins_code", "Cartn_x", "Cartn_y", "Cartn_z", "type_symbol", "B_iso_or_equiv"]


def cif_tokens(text):
    """mmcif.rs:62-161"""
    ws = " \t\n\r\f"
    for line in lines_of(text):
        i = 0
        while i < len(line):
            while i < len(line) and line[i] in ws:
                i += 1
            if i >= len(line) or line[i] == "#":
                break
            c = line[i]
            if c in "'\"":
                j = line.find(c, i + 1)
                if j < 0:
                    yield line[i + 1:]
                    i = len(line)
                else:
                    yield line[i + 1:j]
                    i = j + 1
                continue
            j = i
            while j < len(line) and line[j] not in ws:
                j += 1
            yield line[i:j]
            i = j


def parse_mmcif(text, exclude_heteroatoms=False, include_hydrogens=False, as_assembly=False):
    """mmcif.rs:292-389"""
    toks = cif_tokens(text)
    out = []

    def val(row, fmap, key, fallback=None):
        if key in fmap and fmap[key] < len(row):
            return row[fmap[key]]
        if fallback is not None and fallback in fmap and fmap[fallback] < len(row):
            return row[fmap[fallback]]
        return ""

    def read_rows(fmap, ncols, first):
        first_model = None
        row = [first]
        while True:
            while len(row) < ncols:
                t = next(toks, None)
                if t is None:
                    return
                row.append(t)
            if row[0].startswith("_") or row[0].startswith("data_") or row[0] == "loop_":
                return
            model = val(row, fmap, "pdbx_PDB_model_num")
            if first_model is None:
                first_model = model
            if as_assembly or model == first_model:
                r = row_record(row, fmap)
                if r is not None and acceptable(r, exclude_heteroatoms, include_hydrogens):
                    if as_assembly and model != "1" and model != "":
                        r.chain_id += model
                    r.alt_loc = ""
                    out.append(r)
            t = next(toks, None)
            if t is None:
                return
            row = [t]

    def row_record(row, fmap):
        x, y, z = (parse_f64(val(row, fmap, k)) for k in ("Cartn_x", "Cartn_y", "Cartn_z"))
        if x is None or y is None or z is None:
            return None
        serial = parse_i32(val(row, fmap, "id"))
        res_seq = parse_i32(val(row, fmap, "auth_seq_id", "label_seq_id"))
        if serial is None or res_seq is None:
            return None
        raw = val(row, fmap, "auth_atom_id", "label_atom_id")
        rec = val(row, fmap, "group_PDB")
        if not raw or not rec:
            return None
        b = parse_f64(val(row, fmap, "B_iso_or_equiv"))
        return Record(rec, serial, normalize_atom_name(raw), fix_undefined(val(row, fmap, "label_alt_id")),
                      fix_undefined(val(row, fmap, "auth_comp_id", "label_comp_id")),
                      fix_undefined(val(row, fmap, "auth_asym_id", "label_asym_id")), res_seq,
                      fix_undefined(val(row, fmap, "pdbx_PDB_ins_code")), x, y, z,
                      fix_undefined(val(row, fmap, "type_symbol")), b if b is not None else 0.0)

    for tok in toks:
        if tok != "loop_":
            continue
        fmap, ncols = {}, 0
        for t in toks:
            if not t.startswith("_atom_site."):
                if ncols:
                    read_rows(fmap, ncols, t)
                break
            if t[11:] in _CIF_FIELDS:
                fmap[t[11:]] = ncols
            ncols += 1
    return out


# ----------------------------------------------------------------- XYZR --
def parse_xyzr(text):
    """mod.rs:113-143"""
    out = []
    for line in lines_of(text):
        p = line.split()
        if len(p) < 4:
            continue
        v = [parse_f64(s) for s in p[-4:]]
        if None not in v:
            out.append(v)
    return out


def detect(path, text):
    """mod.rs:80-111: 'pdb' / 'mmcif' / 'xyzr' / None."""
    name = path.rsplit("/", 1)[-1]
    if "." in name[1:]:
        ext = name.rsplit(".", 1)[-1].lower()
        if ext in ("pdb", "ent", "pdb1"):
            return "pdb"
        if ext in ("cif", "mmcif"):
            return "mmcif"
        if ext == "xyzr":
            return "xyzr"
    first = lines_of(text)[0].strip() if text else ""
    if first.startswith(("data_", "_", "loop_")):
        return "mmcif"
    if first.startswith(("ATOM", "HETATM", "HEADER", "REMARK")):
        return "pdb"
    p = first.split()
    if len(p) >= 4 and all(parse_f64(s) is not None for s in p[-4:]):
        return "xyzr"
    return None


# ---------------------------------------------------------------- radii --
# The default rule table of radii.rs:113-270 (data).
DEFAULT_RULES = [
    ("*", "C*", 1.80), ("*", "N*", 1.60), ("*", "O*", 1.50), ("*", "P*", 1.90), ("*", "S*", 1.90),
    ("*", "H*", 1.30), ("*", "C", 1.75), ("*", "CA", 1.90), ("*", "N", 1.70), ("*", "O", 1.49),
    ("ALA", "CB", 1.92), ("ARG", "CB", 1.91), ("ARG", "CD*", 1.88), ("ARG", "CG*", 1.92), ("ARG", "CZ*", 1.80),
    ("ARG", "NE*", 1.62), ("ARG", "NH1", 1.62), ("ARG", "NH2", 1.67), ("ASN", "CB", 1.91), ("ASN", "CG*", 1.81),
    ("ASN", "ND2", 1.62), ("ASN", "OD1", 1.52), ("ASP", "CB", 1.91), ("ASP", "CG*", 1.76), ("ASP", "OD1", 1.49),
    ("ASP", "OD2", 1.49), ("CYS", "CB", 1.91), ("CYS", "S*", 1.88), ("GLN", "CB", 1.91), ("GLN", "CD*", 1.81),
    ("GLN", "CG*", 1.80), ("GLN", "NE2", 1.62), ("GLN", "OE1", 1.52), ("GLU", "CB", 1.91), ("GLU", "CD*", 1.76),
    ("GLU", "CG*", 1.88), ("GLU", "OE1", 1.49), ("GLU", "OE2", 1.49), ("HIS", "CB", 1.91), ("HIS", "CD*", 1.74),
    ("HIS", "CE*", 1.74), ("HIS", "CG*", 1.80), ("HIS", "ND1", 1.60), ("HIS", "NE2", 1.60), ("ILE", "CB", 2.01),
    ("ILE", "CD1", 1.92), ("ILE", "CG1", 1.92), ("ILE", "CG2", 1.92), ("LEU", "CB", 1.91), ("LEU", "CD1", 1.92),
    ("LEU", "CD2", 1.92), ("LEU", "CG*", 2.01), ("LYS", "CB", 1.91), ("LYS", "CD*", 1.92), ("LYS", "CE*", 1.88),
    ("LYS", "CG*", 1.92), ("LYS", "NZ*", 1.67), ("MET", "CB", 1.91), ("MET", "CE*", 1.80), ("MET", "CG*", 1.92),
    ("MET", "S*", 1.94), ("PHE", "CB", 1.91), ("PHE", "CD*", 1.82), ("PHE", "CE*", 1.82), ("PHE", "CG*", 1.74),
    ("PHE", "CZ*", 1.82), ("PRO", "CB", 1.91), ("PRO", "CD*", 1.92), ("PRO", "CG*", 1.92), ("SER", "CB", 1.91),
    ("SER", "OG*", 1.54), ("THR", "CB", 2.01), ("THR", "CG2", 1.92), ("THR", "OG*", 1.54), ("TRP", "CB", 1.91),
    ("TRP", "CD*", 1.82), ("TRP", "CE*", 1.82), ("TRP", "CE2", 1.74), ("TRP", "CG*", 1.74), ("TRP", "CH*", 1.82),
    ("TRP", "CZ*", 1.82), ("TRP", "NE1", 1.66), ("TYR", "CB", 1.91), ("TYR", "CD*", 1.82), ("TYR", "CE*", 1.82),
    ("TYR", "CG*", 1.74), ("TYR", "CZ*", 1.80), ("TYR", "OH*", 1.54), ("VAL", "CB", 2.01), ("VAL", "CG1", 1.92),
    ("VAL", "CG2", 1.92), ("*", "F*", 1.33), ("*", "CL*", 1.81), ("*", "BR*", 1.96), ("*", "I*", 2.20),
    ("AL", "AL", 0.60), ("AS", "AS", 0.58), ("AU", "AU", 1.37), ("BA", "BA", 1.35), ("BE", "BE", 0.45),
    ("BI", "BI", 1.03), ("CA", "CA", 1.00), ("CD", "CD", 0.95), ("CO", "CO", 0.65), ("CR", "CR", 0.73),
    ("CS", "CS", 1.67), ("CU", "CU", 0.73), ("FE", "FE", 0.61), ("HE*", "FE", 0.61), ("GA", "GA", 0.62),
    ("GE", "GE", 0.73), ("HG", "HG", 1.02), ("K", "K", 1.38), ("LI", "LI", 0.76), ("MG", "MG", 0.72),
    ("MN", "MN", 0.83), ("MO", "MO", 0.69), ("NA", "NA", 1.02), ("NI", "NI", 0.69), ("PB", "PB", 1.19),
    ("PD", "PD", 0.86), ("PT", "PT", 0.80), ("RB", "RB", 1.52), ("SB", "SB", 0.76), ("SC", "SC", 0.75),
    ("SN", "SN", 0.69), ("SR", "SR", 1.18), ("TC", "TC", 0.65), ("TI", "TI", 0.86), ("V", "V", 0.79),
    ("ZN", "ZN", 0.74), ("ZR", "ZR", 0.72),
]


class Radii:
    """radii.rs:15-111 over a rule dict {(res, atom): r}; default=True starts
    from DEFAULT_RULES (RadiiLookup::new), False from nothing (empty())."""

    def __init__(self, default=True):
        self.rules = {(r, a): v for r, a, v in DEFAULT_RULES} if default else {}

    def load_text(self, text):
        for raw in lines_of(text):
            line = raw.strip()
            if not line or line.startswith("#"):
                continue
            p = line.split()
            if len(p) < 3:
                raise ValueError(f"Invalid radii rule: {line}")
            v = parse_f64(p[2])
            if v is None:
                raise ValueError(f"Invalid radius value: {p[2]}")
            if v < 0:
                raise ValueError(f"Negative radius: {v}")
            self.rules[(p[0], p[1])] = v

    def get(self, res, atom):
        if atom in ("", "*"):
            return 1.8
        for r in (res, "*"):
            if (r, atom) in self.rules:
                return self.rules[(r, atom)]
            for k in range(len(atom), 0, -1):
                if (r, atom[:k] + "*") in self.rules:
                    return self.rules[(r, atom[:k] + "*")]
        for r in (res, "*"):
            if (r, "*") in self.rules:
                return self.rules[(r, "*")]
        return 1.8


def balls_of(records, radii):
    return [[r.x, r.y, r.z, radii.get(r.res_name, r.name)] for r in records]


def chain_grouping(records):
    ids = {}
    return [ids.setdefault(r.chain_id, len(ids)) for r in records]


def residue_grouping(records):
    ids = {}
    return [ids.setdefault((r.chain_id, r.res_seq, r.i_code), len(ids)) for r in records]


# ------------------------------------------------------------ selection --
class SelectionError(ValueError):
    def __init__(self, message, position):
        super().__init__(f"{message} at position {position}")
        self.message, self.position = message, position


def glob_match(p, t):
    """selection.rs:56-155"""
    star = None
    pi = ti = 0
    while ti < len(t):
        if pi < len(p):
            c = p[pi]
            if c == "?":
                pi += 1
                ti += 1
                continue
            if c == "*":
                star = (pi, ti)
                pi += 1
                continue
            if c == "[":
                cls = _char_class(p, pi, t[ti])
                if cls is not None and cls[0]:
                    pi = cls[1] + 1
                    ti += 1
                    continue
            elif c == t[ti]:
                pi += 1
                ti += 1
                continue
        if star is None:
            return False
        pi = star[0] + 1
        star = (star[0], star[1] + 1)
        ti = star[1]
        if ti > len(t):
            return False
    while pi < len(p) and p[pi] == "*":
        pi += 1
    return pi == len(p)


def _char_class(p, start, c):
    i, m = start + 1, False
    while i < len(p) and p[i] != "]":
        if i + 2 < len(p) and p[i + 1] == "-" and p[i + 2] != "]":
            if p[i] <= c <= p[i + 2]:
                m = True
            i += 3
        else:
            if p[i] == c:
                m = True
            i += 1
    return (m, i) if i < len(p) and p[i] == "]" else None


PROTEIN = {"ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS", "ILE", "LEU", "LYS", "MET", "PHE", "PRO",
           "SER", "THR", "TRP", "TYR", "VAL", "HIE", "HID", "HIP", "CYX", "ASH", "GLH", "LYN"}
DNA = {"DA", "DT", "DG", "DC", "DU"}
RNA = {"A", "U", "G", "C", "RA", "RU", "RG", "RC"}
BACKBONE = {"C", "CA", "N", "O"}
HYDROPHOBIC = {"ALA", "VAL", "ILE", "LEU", "MET", "PHE", "TRP", "PRO", "GLY"}
AROMATIC = {"PHE", "TYR", "TRP", "HIS", "HIE", "HID", "HIP"}
ACIDIC = {"ASP", "GLU", "ASH", "GLH"}
BASIC = {"ARG", "LYS", "HIS", "HIE", "HID", "HIP", "LYN"}
POLAR = {"SER", "THR", "ASN", "GLN", "CYS", "CYX", "TYR"}

_KEYWORDS = {"chain": "chain", "segid": "chain", "resname": "resname", "resn": "resname", "resid": "resid",
             "resi": "resid", "resseq": "resid", "resnum": "resid", "name": "name", "atomname": "name",
             "protein": "protein", "backbone": "backbone", "sidechain": "sidechain", "nucleic": "nucleic",
             "nucleicacid": "nucleic", "hetatm": "hetatm", "hetero": "hetatm", "hydrophobic": "hydrophobic",
             "aromatic": "aromatic", "acidic": "acidic", "basic": "basic", "polar": "polar", "charged": "charged",
             "all": "all", "everything": "all", "none": "none", "nothing": "none", "and": "and", "&&": "and",
             "or": "or", "||": "or", "not": "not", "!": "not", "to": "to"}
_FLAGS = {"protein", "backbone", "sidechain", "nucleic", "hetatm", "hydrophobic", "aromatic", "acidic", "basic",
          "polar", "charged", "all", "none"}


def _is_ident_start(c):
    return (c.isascii() and c.isalnum()) or c in "_*?['"


def _is_ident_char(c, in_bracket):
    return in_bracket or (c.isascii() and c.isalnum()) or c in "_*?[]-'"


def tokenize(s):
    """selection.rs:271-403: list of (kind, value, byte position)."""
    toks, i = [], 0
    pos = [len(s[:k].encode()) for k in range(len(s) + 1)]
    while i < len(s):
        c = s[i]
        if c.isspace():
            i += 1
            continue
        if c in "():":
            toks.append(({"(": "lparen", ")": "rparen", ":": "colon"}[c], None, pos[i]))
            i += 1
            continue
        if c in "\"'":
            j = s.find(c, i + 1)
            j = len(s) if j < 0 else j
            toks.append(("ident", s[i + 1:j], pos[i + 1]))
            i = j + 1
            continue
        if (c.isascii() and c.isdigit()) or (c == "-" and i + 1 < len(s) and s[i + 1].isascii() and s[i + 1].isdigit()):
            j = i + 1
            while j < len(s) and s[j].isascii() and s[j].isdigit():
                j += 1
            v = parse_i32(s[i:j])
            if v is None:
                raise SelectionError(f"Invalid number: {s[i:j]}", pos[i])
            toks.append(("number", v, pos[i]))
            i = j
            continue
        if _is_ident_start(c):
            j, br = i, False
            while j < len(s):
                d = s[j]
                if d == "[":
                    br = True
                elif d == "]":
                    br = False
                if not _is_ident_char(d, br):
                    break
                j += 1
            ident = s[i:j]
            kw = _KEYWORDS.get(ident.lower())
            toks.append((kw, None, pos[i]) if kw else ("ident", ident, pos[i]))
            i = j
            continue
        raise SelectionError(f"Unexpected character: {c}", pos[i])
    return toks


class Selection:
    """Selection::parse + matches (selection.rs:465-738). The expression is a
    nested tuple."""

    def __init__(self, text):
        self.toks = tokenize(text)
        self.i = 0
        if not self.toks:
            raise SelectionError("Empty selection", 0)
        self.expr = self._or()
        if self.i < len(self.toks):
            raise SelectionError("Unexpected token after expression", self._pos())

    def _pos(self):
        return self.toks[self.i][2] if self.i < len(self.toks) else 0

    def _peek(self):
        return self.toks[self.i][0] if self.i < len(self.toks) else None

    def _or(self):
        e = self._and()
        while self._peek() == "or":
            self.i += 1
            e = ("or", e, self._and())
        return e

    def _and(self):
        e = self._not()
        while self._peek() == "and":
            self.i += 1
            e = ("and", e, self._not())
        return e

    def _not(self):
        if self._peek() == "not":
            self.i += 1
            return ("not", self._not())
        return self._primary()

    def _primary(self):
        pos, k = self._pos(), self._peek()
        if k == "lparen":
            self.i += 1
            e = self._or()
            if self._peek() != "rparen":
                raise SelectionError("Missing closing parenthesis", self._pos())
            self.i += 1
            return e
        if k in ("chain", "resname", "name"):
            self.i += 1
            pats = []
            while self._peek() == "ident":
                pats.append(self.toks[self.i][1])
                self.i += 1
            if not pats:
                raise SelectionError(f"{k} requires at least one argument", pos)
            return (k, pats)
        if k == "resid":
            self.i += 1
            return ("resid", self._ranges(pos))
        if k in _FLAGS:
            self.i += 1
            return (k,)
        if k in ("and", "or"):
            raise SelectionError("Unexpected boolean operator", pos)
        if k == "ident":
            raise SelectionError(f"Unknown keyword: {self.toks[self.i][1]}", pos)
        raise SelectionError("Expected selection expression", pos)

    def _ranges(self, pos):
        out = []
        while self.i < len(self.toks):
            k, v, _ = self.toks[self.i]
            if k == "number":
                self.i += 1
                if self._peek() in ("to", "colon"):
                    sep = self._peek()
                    self.i += 1
                    if self._peek() == "number":
                        out.append((v, self.toks[self.i][1]))
                        self.i += 1
                    else:
                        raise SelectionError("Expected number after 'to'" if sep == "to" else
                                             "Expected number after ':'", self._pos())
                else:
                    out.append((v, v))
            elif k == "ident":
                parts = v.split(":")
                if len(parts) == 2 and parse_i32(parts[0]) is not None and parse_i32(parts[1]) is not None:
                    out.append((parse_i32(parts[0]), parse_i32(parts[1])))
                    self.i += 1
                elif parse_i32(v) is not None:
                    out.append((parse_i32(v), parse_i32(v)))
                    self.i += 1
                else:
                    break
            else:
                break
        if not out:
            raise SelectionError("resid requires at least one argument", pos)
        return out

    def matches(self, r):
        return _match(self.expr, r)


def _match(e, r):
    k = e[0]
    res, atom = r.res_name, r.name
    if k == "chain":
        return any(glob_match(p, r.chain_id) for p in e[1])
    if k == "resname":
        return any(glob_match(p, res) for p in e[1])
    if k == "name":
        return any(glob_match(p, atom) for p in e[1])
    if k == "resid":
        return any(lo <= r.res_seq <= hi for lo, hi in e[1])
    if k == "protein":
        return res in PROTEIN
    if k == "backbone":
        return res in PROTEIN and atom in BACKBONE
    if k == "sidechain":
        return res in PROTEIN and atom not in BACKBONE
    if k == "nucleic":
        return res in DNA or res in RNA
    if k == "hetatm":
        return r.record_name == "HETATM"
    if k == "hydrophobic":
        return res in HYDROPHOBIC
    if k == "aromatic":
        return res in AROMATIC
    if k == "acidic":
        return res in ACIDIC
    if k == "basic":
        return res in BASIC
    if k == "polar":
        return res in POLAR
    if k == "charged":
        return res in ACIDIC or res in BASIC
    if k == "all":
        return True
    if k == "none":
        return False
    if k == "and":
        return _match(e[1], r) and _match(e[2], r)
    if k == "or":
        return _match(e[1], r) or _match(e[2], r)
    if k == "not":
        return not _match(e[1], r)
    raise AssertionError(k)


def parse_selections(texts):
    """selection.rs:742-766"""
    if len(texts) < 2:
        raise SelectionError("At least two selections required", 0)
    out = []
    for k, t in enumerate(texts):
        try:
            out.append(Selection(t.strip()))
        except SelectionError as e:
            raise SelectionError(e.message, k) from None
    return out


def custom_grouping(records, selections):
    """selection.rs:773-814"""
    g = []
    for r in records:
        g.append(next((k for k, s in enumerate(selections) if s.matches(r)), -1))
    nxt = len(selections)
    for i, x in enumerate(g):
        if x == -1:
            g[i] = nxt
            nxt += 1
    return g


# ------------------------------------------------------------- graphics --
# graphics.rs restated over plain float tuples, in geometry.rs's operation
# order (dot = (x x' + y y') + z z', normalize = division by the norm).
TAU = 6.283185307179586


def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def _mul(a, s):
    return (a[0] * s, a[1] * s, a[2] * s)


def _dot(a, b):
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _unit(a):
    n = math.sqrt(_dot(a, a))
    return (a[0] / n, a[1] / n, a[2] / n)


def _feq(a, b):
    return abs(a - b) <= 1e-10


def any_normal(a):
    """geometry.rs:179-193"""
    b = list(a)
    if not _feq(b[0], 0.0) and (not _feq(b[1], 0.0) or not _feq(b[2], 0.0)):
        b[0] = -b[0]
        return _unit(_cross(a, tuple(b)))
    if not _feq(b[1], 0.0) and (not _feq(b[0], 0.0) or not _feq(b[2], 0.0)):
        b[1] = -b[1]
        return _unit(_cross(a, tuple(b)))
    if not _feq(b[0], 0.0):
        return (0.0, 1.0, 0.0)
    return (1.0, 0.0, 0.0)


def rotate(axis, angle, p):
    """geometry.rs:196-203 (unit quaternion (cos t/2, u sin t/2) applied as
    (t w + q x t) + p with t = 2 q x p)."""
    if _dot(axis, axis) <= 0.0:
        return p
    u = _unit(axis)
    h = angle / 2.0
    qv = _mul(u, math.sin(h))
    t = _mul(_cross(qv, p), 2.0)
    cr = _cross(qv, t)
    return _add(_add(_mul(t, math.cos(h)), cr), p)


def directed_angle(o, a, b, c):
    """geometry.rs:157-176"""
    v1 = _unit(_sub(a, o))
    v2 = _unit(_sub(b, o))
    ang = math.acos(max(-1.0, min(1.0, _dot(v1, v2))))
    return ang if _dot(_sub(c, o), _cross(v1, v2)) >= 0.0 else TAU - ang


def polygon(poly, angle_step):
    """ContactDescriptor::to_graphics outer points (graphics.rs:203-349)."""
    c, r, axis = poly["ic_center"], poly["ic_r"], poly["axis"]
    if not poly["contour"]:
        first = _mul(any_normal(axis), r)
        n = math.ceil(TAU / angle_step)
        step = TAU / n
        return [_add(c, rotate(axis, i * step, first)) for i in range(n)]
    pts = []
    cont = poly["contour"]
    for i, cp in enumerate(cont):
        nx = cont[(i + 1) % len(cont)]
        p = cp[:3]
        pts.append(p)
        if cp[5] == poly["a"] and cp[3] > 0.0 and not cp[3] <= angle_step:
            ang = min(directed_angle(c, p, nx[:3], _add(c, axis)), cp[3])
            n = math.floor(ang / angle_step)
            if n > 1:
                off = _sub(p, c)
                step = ang / n
                pts += [_add(c, _mul(_unit(rotate(axis, i2 * step, off)), r)) for i2 in range(1, n)]
    return pts


def pymol_script(balls, polys, angle_step=0.5, name="contacts"):
    """GraphicsWriter::write_pymol (graphics.rs:87-178)."""
    faces = []
    for poly in polys:
        outer = polygon(poly, angle_step)
        if len(outer) >= 3:
            faces.append((outer, poly["ic_center"] if not poly["contour"] else poly["bary"], poly["axis"]))
    f6 = "%.6f"
    o = ["from pymol.cgo import *", "from pymol import cmd", "", "cgo_graphics_list_balls = [", "    COLOR, 0, 1, 1,"]
    for b in balls:
        o.append("    SPHERE, " + ", ".join(f6 % v for v in b) + ",")
    o += ["]", f"cmd.load_cgo(cgo_graphics_list_balls, '{name}_balls')", "", "cgo_graphics_list_faces = [",
          "    COLOR, 1, 1, 0,"]

    def vx(head, p):
        return f"    {head}, " + ", ".join(f6 % v for v in p) + ","

    for outer, bary, normal in faces:
        o += ["    BEGIN, TRIANGLE_FAN,", vx("NORMAL", normal), vx("VERTEX", bary)]
        o += [vx("VERTEX", p) for p in outer] + [vx("VERTEX", outer[0]), "    END,"]
    o += ["]", f"cmd.load_cgo(cgo_graphics_list_faces, '{name}_faces')", "", "cgo_graphics_list_wireframe = [",
          "    COLOR, 1, 0, 0,"]
    for outer, _, _ in faces:
        o += ["    BEGIN, LINE_LOOP,"] + [vx("VERTEX", p) for p in outer] + ["    END,"]
    o += ["]", f"cmd.load_cgo(cgo_graphics_list_wireframe, '{name}_wireframe')", "",
          "cmd.set('two_sided_lighting', 1)", "cmd.set('cgo_line_width', 1)"]
    return "\n".join(o) + "\n"
