// Molecular file front end (replaces src/input/{mod,pdb,mmcif,radii,
// selection}.rs): PDB / mmCIF / XYZR parsing with the reference's filters,
// radii lookup, VMD-like selections, groupings and the from-file
// tessellation. Host code; the tessellation itself runs through the device
// pipeline (vltr::host_tessellate).
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "vltr_host.h"
#include "voronota_ltr.h"
#include "voronota_ltr_input.h"

namespace {

using vltr::host_fail;

// ------------------------------------------------------------ text utils --

// char::is_whitespace on ASCII (trim, split_whitespace)
bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }
// u8::is_ascii_whitespace (the mmCIF tokenizer): no vertical tab
bool is_ascii_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f'; }

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && is_space(s[a])) ++a;
  while (b > a && is_space(s[b - 1])) --b;
  return s.substr(a, b - a);
}

bool starts_with(const std::string& s, const char* p) { return s.rfind(p, 0) == 0; }

std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && is_space(s[i])) ++i;
    const size_t j = i;
    while (i < s.size() && !is_space(s[i])) ++i;
    if (i > j) out.push_back(s.substr(j, i - j));
  }
  return out;
}

// UTF-8 well-formedness (Rust reads text line by line and stops at the first
// line that is not valid UTF-8).
bool valid_utf8(const std::string& s) {
  size_t i = 0;
  while (i < s.size()) {
    const unsigned char c = static_cast<unsigned char>(s[i]);
    int len = 0;
    uint32_t cp = 0;
    if (c < 0x80) {
      ++i;
      continue;
    } else if ((c & 0xe0) == 0xc0) {
      len = 2;
      cp = c & 0x1f;
    } else if ((c & 0xf0) == 0xe0) {
      len = 3;
      cp = c & 0x0f;
    } else if ((c & 0xf8) == 0xf0) {
      len = 4;
      cp = c & 0x07;
    } else {
      return false;
    }
    if (i + len > s.size()) return false;
    for (int k = 1; k < len; ++k) {
      const unsigned char d = static_cast<unsigned char>(s[i + k]);
      if ((d & 0xc0) != 0x80) return false;
      cp = (cp << 6) | (d & 0x3f);
    }
    if ((len == 2 && cp < 0x80) || (len == 3 && cp < 0x800) || (len == 4 && (cp < 0x10000 || cp > 0x10ffff)) ||
        (cp >= 0xd800 && cp <= 0xdfff))
      return false;
    i += len;
  }
  return true;
}

// Lines as BufRead::lines yields them: split at '\n', a trailing '\r'
// removed, stopping at the first line that is not UTF-8.
std::vector<std::string> text_lines(const std::string& text) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < text.size()) {
    size_t j = text.find('\n', i);
    const bool nl = j != std::string::npos;
    if (!nl) j = text.size();
    std::string line = text.substr(i, j - i);
    if (nl && !line.empty() && line.back() == '\r') line.pop_back();
    if (!valid_utf8(line)) break;
    out.push_back(std::move(line));
    i = nl ? j + 1 : j;
  }
  return out;
}

// str::parse::<f64>: an optional sign, decimal digits with an optional point
// and exponent, or inf / infinity / nan (any case); nothing else.
bool parse_f64(const std::string& s, double* out) {
  if (s.empty()) return false;
  size_t i = 0;
  if (s[i] == '+' || s[i] == '-') ++i;
  std::string rest = s.substr(i);
  std::string low;
  for (char c : rest) low.push_back(static_cast<char>(std::tolower(static_cast<unsigned char>(c))));
  if (low == "inf" || low == "infinity" || low == "nan") {
    const double v = low == "nan" ? NAN : INFINITY;
    *out = s[0] == '-' ? -v : v;
    return true;
  }
  size_t k = i, digits = 0;
  while (k < s.size() && std::isdigit(static_cast<unsigned char>(s[k]))) ++k, ++digits;
  if (k < s.size() && s[k] == '.') {
    ++k;
    while (k < s.size() && std::isdigit(static_cast<unsigned char>(s[k]))) ++k, ++digits;
  }
  if (digits == 0) return false;
  if (k < s.size() && (s[k] == 'e' || s[k] == 'E')) {
    ++k;
    if (k < s.size() && (s[k] == '+' || s[k] == '-')) ++k;
    size_t ed = 0;
    while (k < s.size() && std::isdigit(static_cast<unsigned char>(s[k]))) ++k, ++ed;
    if (ed == 0) return false;
  }
  if (k != s.size()) return false;
  *out = std::strtod(s.c_str(), nullptr);
  return true;
}

// str::parse::<i32>: optional sign, digits, no overflow.
bool parse_i32(const std::string& s, int32_t* out) {
  if (s.empty()) return false;
  size_t i = 0;
  bool neg = false;
  if (s[0] == '+' || s[0] == '-') {
    neg = s[0] == '-';
    i = 1;
  }
  if (i == s.size()) return false;
  int64_t v = 0;
  for (; i < s.size(); ++i) {
    if (!std::isdigit(static_cast<unsigned char>(s[i]))) return false;
    v = v * 10 + (s[i] - '0');
    if (v > 2147483648ll) return false;
  }
  if (neg) v = -v;
  if (v > 2147483647ll || v < -2147483648ll) return false;
  *out = static_cast<int32_t>(v);
  return true;
}

// Rust's Display of f64 (shortest round-trip digits, positional notation).
std::string display_f64(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
  char buf[64];
  int prec = 1;
  for (; prec < 17; ++prec) {
    std::snprintf(buf, sizeof buf, "%.*e", prec - 1, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  std::snprintf(buf, sizeof buf, "%.*e", prec - 1, v);
  std::string str(buf);
  const bool neg = str[0] == '-';
  if (neg) str.erase(0, 1);
  const size_t epos = str.find('e');
  const int ex = std::atoi(str.c_str() + epos + 1);
  std::string digits = str.substr(0, epos);
  digits.erase(std::remove(digits.begin(), digits.end(), '.'), digits.end());
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  std::string o;
  const int point = ex + 1;
  if (point <= 0) {
    o = "0." + std::string(static_cast<size_t>(-point), '0') + digits;
  } else if (point >= static_cast<int>(digits.size())) {
    o = digits + std::string(static_cast<size_t>(point) - digits.size(), '0');
  } else {
    o = digits.substr(0, static_cast<size_t>(point)) + "." + digits.substr(static_cast<size_t>(point));
  }
  return (neg ? "-" : "") + o;
}

// io::Error Display of an OS error.
std::string os_error(int err) { return std::string(std::strerror(err)) + " (os error " + std::to_string(err) + ")"; }

// ----------------------------------------------------------------- radii --

constexpr double kDefaultRadius = 1.8;

struct RadiiRule {
  const char* res;
  const char* atom;
  double r;
};

// The default rule table of voronota-lt (src/input/radii.rs:113-270), kept
// as data: residue, atom name or prefix pattern ("C*"), radius.
const RadiiRule kDefaultRules[] = {
    {"*", "C*", 1.80},    {"*", "N*", 1.60},    {"*", "O*", 1.50},    {"*", "P*", 1.90},    {"*", "S*", 1.90},
    {"*", "H*", 1.30},    {"*", "C", 1.75},     {"*", "CA", 1.90},    {"*", "N", 1.70},     {"*", "O", 1.49},
    {"ALA", "CB", 1.92},  {"ARG", "CB", 1.91},  {"ARG", "CD*", 1.88}, {"ARG", "CG*", 1.92}, {"ARG", "CZ*", 1.80},
    {"ARG", "NE*", 1.62}, {"ARG", "NH1", 1.62}, {"ARG", "NH2", 1.67}, {"ASN", "CB", 1.91},  {"ASN", "CG*", 1.81},
    {"ASN", "ND2", 1.62}, {"ASN", "OD1", 1.52}, {"ASP", "CB", 1.91},  {"ASP", "CG*", 1.76}, {"ASP", "OD1", 1.49},
    {"ASP", "OD2", 1.49}, {"CYS", "CB", 1.91},  {"CYS", "S*", 1.88},  {"GLN", "CB", 1.91},  {"GLN", "CD*", 1.81},
    {"GLN", "CG*", 1.80}, {"GLN", "NE2", 1.62}, {"GLN", "OE1", 1.52}, {"GLU", "CB", 1.91},  {"GLU", "CD*", 1.76},
    {"GLU", "CG*", 1.88}, {"GLU", "OE1", 1.49}, {"GLU", "OE2", 1.49}, {"HIS", "CB", 1.91},  {"HIS", "CD*", 1.74},
    {"HIS", "CE*", 1.74}, {"HIS", "CG*", 1.80}, {"HIS", "ND1", 1.60}, {"HIS", "NE2", 1.60}, {"ILE", "CB", 2.01},
    {"ILE", "CD1", 1.92}, {"ILE", "CG1", 1.92}, {"ILE", "CG2", 1.92}, {"LEU", "CB", 1.91},  {"LEU", "CD1", 1.92},
    {"LEU", "CD2", 1.92}, {"LEU", "CG*", 2.01}, {"LYS", "CB", 1.91},  {"LYS", "CD*", 1.92}, {"LYS", "CE*", 1.88},
    {"LYS", "CG*", 1.92}, {"LYS", "NZ*", 1.67}, {"MET", "CB", 1.91},  {"MET", "CE*", 1.80}, {"MET", "CG*", 1.92},
    {"MET", "S*", 1.94},  {"PHE", "CB", 1.91},  {"PHE", "CD*", 1.82}, {"PHE", "CE*", 1.82}, {"PHE", "CG*", 1.74},
    {"PHE", "CZ*", 1.82}, {"PRO", "CB", 1.91},  {"PRO", "CD*", 1.92}, {"PRO", "CG*", 1.92}, {"SER", "CB", 1.91},
    {"SER", "OG*", 1.54}, {"THR", "CB", 2.01},  {"THR", "CG2", 1.92}, {"THR", "OG*", 1.54}, {"TRP", "CB", 1.91},
    {"TRP", "CD*", 1.82}, {"TRP", "CE*", 1.82}, {"TRP", "CE2", 1.74}, {"TRP", "CG*", 1.74}, {"TRP", "CH*", 1.82},
    {"TRP", "CZ*", 1.82}, {"TRP", "NE1", 1.66}, {"TYR", "CB", 1.91},  {"TYR", "CD*", 1.82}, {"TYR", "CE*", 1.82},
    {"TYR", "CG*", 1.74}, {"TYR", "CZ*", 1.80}, {"TYR", "OH*", 1.54}, {"VAL", "CB", 2.01},  {"VAL", "CG1", 1.92},
    {"VAL", "CG2", 1.92}, {"*", "F*", 1.33},    {"*", "CL*", 1.81},   {"*", "BR*", 1.96},   {"*", "I*", 2.20},
    {"AL", "AL", 0.60},   {"AS", "AS", 0.58},   {"AU", "AU", 1.37},   {"BA", "BA", 1.35},   {"BE", "BE", 0.45},
    {"BI", "BI", 1.03},   {"CA", "CA", 1.00},   {"CD", "CD", 0.95},   {"CO", "CO", 0.65},   {"CR", "CR", 0.73},
    {"CS", "CS", 1.67},   {"CU", "CU", 0.73},   {"FE", "FE", 0.61},   {"HE*", "FE", 0.61},  {"GA", "GA", 0.62},
    {"GE", "GE", 0.73},   {"HG", "HG", 1.02},   {"K", "K", 1.38},     {"LI", "LI", 0.76},   {"MG", "MG", 0.72},
    {"MN", "MN", 0.83},   {"MO", "MO", 0.69},   {"NA", "NA", 1.02},   {"NI", "NI", 0.69},   {"PB", "PB", 1.19},
    {"PD", "PD", 0.86},   {"PT", "PT", 0.80},   {"RB", "RB", 1.52},   {"SB", "SB", 0.76},   {"SC", "SC", 0.75},
    {"SN", "SN", 0.69},   {"SR", "SR", 1.18},   {"TC", "TC", 0.65},   {"TI", "TI", 0.86},   {"V", "V", 0.79},
    {"ZN", "ZN", 0.74},   {"ZR", "ZR", 0.72},
};

}  // namespace

struct vltr_radii {
  std::map<std::pair<std::string, std::string>, double> rules;

  const double* find(const std::string& res, const std::string& atom) const {
    auto it = rules.find({res, atom});
    return it == rules.end() ? nullptr : &it->second;
  }
  // radii.rs:81-110: exact, then the longest "prefix*" pattern, for the
  // residue and then for "*"; then the residue's and "*"'s "*" rule; 1.8.
  double get(const std::string& res, const std::string& atom) const {
    if (atom.empty() || atom == "*") return kDefaultRadius;
    for (const std::string& r : {res, std::string("*")}) {
      if (const double* v = find(r, atom)) return *v;
      for (size_t k = atom.size(); k >= 1; --k)
        if (const double* v = find(r, atom.substr(0, k) + "*")) return *v;
    }
    for (const std::string& r : {res, std::string("*")})
      if (const double* v = find(r, "*")) return *v;
    return kDefaultRadius;
  }
};

namespace {

void add_default_rules(vltr_radii* r) {
  for (const RadiiRule& d : kDefaultRules) r->rules[{d.res, d.atom}] = d.r;
}

// radii.rs:50-76
int load_radii_text(vltr_radii* r, const std::string& text) {
  for (const std::string& raw : text_lines(text)) {
    const std::string line = trim(raw);
    if (line.empty() || line[0] == '#') continue;
    const std::vector<std::string> parts = split_ws(line);
    if (parts.size() < 3) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "Invalid radii rule: " + line);
    double v = 0.0;
    if (!parse_f64(parts[2], &v)) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "Invalid radius value: " + parts[2]);
    if (v < 0.0) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "Negative radius: " + display_f64(v));
    r->rules[{parts[0], parts[1]}] = v;
  }
  return VLTR_OK;
}

// ---------------------------------------------------------- atom records --

struct AtomRecord {
  std::string record_name;
  int32_t serial = 0;
  std::string name, alt_loc, res_name, chain_id;
  int32_t res_seq = 0;
  std::string i_code;
  double x = 0, y = 0, z = 0;
  std::string element;
  double b_factor = 0.0;
};

std::string fix_undefined(const std::string& s) { return (s == "." || s == "?") ? std::string() : s; }

// Leading digits moved to the end: "1HG2" -> "HG21" (pdb.rs:46-56).
std::string normalize_atom_name(const std::string& name) {
  size_t k = 0;
  while (k < name.size() && std::isdigit(static_cast<unsigned char>(name[k]))) ++k;
  if (k == 0 || k == name.size()) return name;
  return name.substr(k) + name.substr(0, k);
}

bool is_hydrogen(const AtomRecord& r) { return starts_with(r.name, "H") || r.element == "H" || r.element == "D"; }

// pdb.rs:112-139 (the same filter in mmcif.rs:262-285)
bool acceptable(const AtomRecord& r, const vltr_parse_options& o) {
  if (r.record_name == "HETATM" && o.exclude_heteroatoms) return false;
  if (!r.alt_loc.empty() && r.alt_loc != "A" && r.alt_loc != "1" && r.alt_loc != ".") return false;
  if (!o.include_hydrogens && is_hydrogen(r)) return false;
  if (r.res_name == "HOH") return false;
  return true;
}

// ------------------------------------------------------------------- PDB --

// Fixed columns start..end (1-based, inclusive), trimmed (pdb.rs:13-25).
std::string column(const std::string& line, size_t start, size_t end) {
  const size_t a = start - 1;
  if (a >= line.size()) return "";
  const size_t b = std::min(end, line.size());
  // a cut inside a multi-byte character gives nothing, as str::get does
  auto boundary = [&](size_t i) {
    return i == 0 || i >= line.size() || (static_cast<unsigned char>(line[i]) & 0xc0) != 0x80;
  };
  if (!boundary(a) || !boundary(b)) return "";
  return trim(line.substr(a, b - a));
}

// pdb.rs:68-102
bool parse_pdb_atom(const std::string& line, AtomRecord* r) {
  const std::string rec = column(line, 1, 6);
  if (rec != "ATOM" && rec != "HETATM") return false;
  if (!parse_f64(column(line, 31, 38), &r->x) || !parse_f64(column(line, 39, 46), &r->y) ||
      !parse_f64(column(line, 47, 54), &r->z))
    return false;
  if (!parse_i32(column(line, 7, 11), &r->serial) || !parse_i32(column(line, 23, 26), &r->res_seq)) return false;
  const std::string raw = column(line, 13, 16);
  if (raw.empty()) return false;
  r->record_name = rec;
  r->name = normalize_atom_name(raw);
  r->alt_loc = fix_undefined(column(line, 17, 17));
  r->res_name = fix_undefined(column(line, 18, 20));
  r->chain_id = fix_undefined(column(line, 22, 22));
  r->i_code = fix_undefined(column(line, 27, 27));
  r->element = fix_undefined(column(line, 77, 78));
  if (!parse_f64(column(line, 61, 66), &r->b_factor)) r->b_factor = 0.0;
  return true;
}

// pdb.rs:142-171: first model only unless as_assembly (then chain ids of
// later models get the model number appended); stops at END.
std::vector<AtomRecord> parse_pdb(const std::vector<std::string>& lines, const vltr_parse_options& o) {
  std::vector<AtomRecord> out;
  int model = 1;
  for (const std::string& line : lines) {
    const std::string rec = column(line, 1, 6);
    if (rec == "ATOM" || rec == "HETATM") {
      AtomRecord r;
      if (parse_pdb_atom(line, &r) && acceptable(r, o)) {
        if (o.as_assembly && model > 1) r.chain_id += std::to_string(model);
        r.alt_loc.clear();
        out.push_back(std::move(r));
      }
    } else if (rec == "ENDMDL") {
      ++model;
      if (!o.as_assembly) break;
    } else if (rec == "END") {
      break;
    }
  }
  return out;
}

// ----------------------------------------------------------------- mmCIF --

enum CifField {
  kModelNum, kGroupPdb, kId, kAuthAtomId, kLabelAtomId, kLabelAltId, kAuthCompId, kLabelCompId, kAuthAsymId,
  kLabelAsymId, kAuthSeqId, kLabelSeqId, kInsCode, kCartnX, kCartnY, kCartnZ, kTypeSymbol, kBIso, kNumFields
};

int cif_field(const std::string& name) {
  static const std::unordered_map<std::string, int> m = {
      {"pdbx_PDB_model_num", kModelNum}, {"group_PDB", kGroupPdb},         {"id", kId},
      {"auth_atom_id", kAuthAtomId},     {"label_atom_id", kLabelAtomId},   {"label_alt_id", kLabelAltId},
      {"auth_comp_id", kAuthCompId},     {"label_comp_id", kLabelCompId},   {"auth_asym_id", kAuthAsymId},
      {"label_asym_id", kLabelAsymId},   {"auth_seq_id", kAuthSeqId},       {"label_seq_id", kLabelSeqId},
      {"pdbx_PDB_ins_code", kInsCode},   {"Cartn_x", kCartnX},              {"Cartn_y", kCartnY},
      {"Cartn_z", kCartnZ},              {"type_symbol", kTypeSymbol},      {"B_iso_or_equiv", kBIso}};
  auto it = m.find(name);
  return it == m.end() ? -1 : it->second;
}

// mmcif.rs:62-161: whitespace-separated tokens; '#' at a token start skips
// the rest of the line; a token opening with ' or " runs to the next same
// quote on the line.
struct CifTokens {
  const std::vector<std::string>& lines;
  size_t li = 0, pos = 0;

  bool next(std::string* tok) {
    for (;;) {
      if (li >= lines.size()) return false;
      const std::string& b = lines[li];
      while (pos < b.size() && is_ascii_ws(b[pos])) ++pos;
      if (pos >= b.size() || b[pos] == '#') {
        ++li;
        pos = 0;
        continue;
      }
      const char c = b[pos];
      if (c == '\'' || c == '"') {
        const size_t start = ++pos;
        while (pos < b.size() && b[pos] != c) ++pos;
        *tok = b.substr(start, pos - start);
        if (pos < b.size()) ++pos;
        return true;
      }
      const size_t start = pos;
      while (pos < b.size() && !is_ascii_ws(b[pos])) ++pos;
      *tok = b.substr(start, pos - start);
      return true;
    }
  }
};

// mmcif.rs:185-203: the field's value, else the fallback's, else "".
const std::string& cif_value(const std::vector<std::string>& row, const int* fmap, int primary, int fallback) {
  static const std::string empty;
  if (fmap[primary] >= 0 && static_cast<size_t>(fmap[primary]) < row.size()) return row[fmap[primary]];
  if (fallback >= 0 && fmap[fallback] >= 0 && static_cast<size_t>(fmap[fallback]) < row.size())
    return row[fmap[fallback]];
  return empty;
}

// mmcif.rs:206-258
bool parse_cif_row(const std::vector<std::string>& row, const int* fmap, AtomRecord* r) {
  if (!parse_f64(cif_value(row, fmap, kCartnX, -1), &r->x) || !parse_f64(cif_value(row, fmap, kCartnY, -1), &r->y) ||
      !parse_f64(cif_value(row, fmap, kCartnZ, -1), &r->z))
    return false;
  if (!parse_i32(cif_value(row, fmap, kId, -1), &r->serial)) return false;
  if (!parse_i32(cif_value(row, fmap, kAuthSeqId, kLabelSeqId), &r->res_seq)) return false;
  const std::string& raw = cif_value(row, fmap, kAuthAtomId, kLabelAtomId);
  if (raw.empty()) return false;
  const std::string& rec = cif_value(row, fmap, kGroupPdb, -1);
  if (rec.empty()) return false;
  r->record_name = rec;
  r->name = normalize_atom_name(raw);
  r->alt_loc = fix_undefined(cif_value(row, fmap, kLabelAltId, -1));
  r->res_name = fix_undefined(cif_value(row, fmap, kAuthCompId, kLabelCompId));
  r->chain_id = fix_undefined(cif_value(row, fmap, kAuthAsymId, kLabelAsymId));
  r->i_code = fix_undefined(cif_value(row, fmap, kInsCode, -1));
  r->element = fix_undefined(cif_value(row, fmap, kTypeSymbol, -1));
  if (!parse_f64(cif_value(row, fmap, kBIso, -1), &r->b_factor)) r->b_factor = 0.0;
  return true;
}

// mmcif.rs:330-389: rows of num_cols tokens until a row starts a new
// category, block or loop (that row is dropped); only the first model's rows
// unless as_assembly.
void read_atom_site(CifTokens& tk, const int* fmap, size_t ncols, const std::string& first, const vltr_parse_options& o,
                    std::vector<AtomRecord>* out) {
  bool have_first_model = false;
  std::string first_model;
  std::vector<std::string> row;
  row.push_back(first);
  for (;;) {
    std::string t;
    while (row.size() < ncols) {
      if (!tk.next(&t)) return;
      row.push_back(t);
    }
    if (starts_with(row[0], "_") || starts_with(row[0], "data_") || row[0] == "loop_") return;
    const std::string model = cif_value(row, fmap, kModelNum, -1);
    if (!have_first_model) {
      first_model = model;
      have_first_model = true;
    }
    if (o.as_assembly || model == first_model) {
      AtomRecord r;
      if (parse_cif_row(row, fmap, &r) && acceptable(r, o)) {
        if (o.as_assembly && model != "1" && !model.empty()) r.chain_id += model;
        r.alt_loc.clear();
        out->push_back(std::move(r));
      }
    }
    row.clear();
    if (!tk.next(&t)) return;
    row.push_back(t);
  }
}

// mmcif.rs:292-327
std::vector<AtomRecord> parse_mmcif(const std::vector<std::string>& lines, const vltr_parse_options& o) {
  std::vector<AtomRecord> out;
  CifTokens tk{lines};
  std::string tok;
  while (tk.next(&tok)) {
    if (tok != "loop_") continue;
    int fmap[kNumFields];
    std::fill(fmap, fmap + kNumFields, -1);
    size_t ncols = 0;
    std::string t;
    while (tk.next(&t)) {
      if (!starts_with(t, "_atom_site.")) {
        if (ncols > 0) read_atom_site(tk, fmap, ncols, t, o, &out);
        break;
      }
      const int f = cif_field(t.substr(11));
      if (f >= 0) fmap[f] = static_cast<int>(ncols);
      ++ncols;
    }
  }
  return out;
}

// ------------------------------------------------------------------ XYZR --

// mod.rs:113-143: the last four numbers of each line are x y z r.
std::vector<vltr_ball> parse_xyzr(const std::vector<std::string>& lines) {
  std::vector<vltr_ball> out;
  for (const std::string& line : lines) {
    const std::vector<std::string> p = split_ws(line);
    if (p.size() < 4) continue;
    double v[4];
    bool ok = true;
    for (int k = 0; k < 4 && ok; ++k) ok = parse_f64(p[p.size() - 4 + k], &v[k]);
    if (ok) out.push_back(vltr_ball{v[0], v[1], v[2], v[3]});
  }
  return out;
}

// mod.rs:84-111 (content detection from the first line)
int detect_from_first_line(const std::string& first, bool xyzr_fallback) {
  const std::string t = trim(first);
  if (starts_with(t, "data_") || starts_with(t, "_") || starts_with(t, "loop_")) return VLTR_FORMAT_MMCIF;
  if (starts_with(t, "ATOM") || starts_with(t, "HETATM") || starts_with(t, "HEADER") || starts_with(t, "REMARK"))
    return VLTR_FORMAT_PDB;
  if (xyzr_fallback) return VLTR_FORMAT_XYZR;
  const std::vector<std::string> p = split_ws(t);
  if (p.size() >= 4) {
    double v;
    bool ok = true;
    for (size_t k = p.size() - 4; k < p.size() && ok; ++k) ok = parse_f64(p[k], &v);
    if (ok) return VLTR_FORMAT_XYZR;
  }
  return VLTR_FORMAT_AUTO;
}

// mod.rs:80-90: extension (case-insensitive) of the file name.
int detect_from_extension(const std::string& path) {
  const size_t slash = path.find_last_of('/');
  const std::string name = slash == std::string::npos ? path : path.substr(slash + 1);
  const size_t dot = name.find_last_of('.');
  if (dot == std::string::npos || dot == 0) return VLTR_FORMAT_AUTO;
  std::string ext = name.substr(dot + 1);
  for (char& c : ext) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  if (ext == "pdb" || ext == "ent" || ext == "pdb1") return VLTR_FORMAT_PDB;
  if (ext == "cif" || ext == "mmcif") return VLTR_FORMAT_MMCIF;
  if (ext == "xyzr") return VLTR_FORMAT_XYZR;
  return VLTR_FORMAT_AUTO;
}

}  // namespace

struct vltr_structure {
  int format = VLTR_FORMAT_XYZR;
  std::vector<vltr_ball> balls;
  std::vector<AtomRecord> records;
};

namespace {

vltr_structure* build_structure(const std::vector<std::string>& lines, int format, const vltr_parse_options& o,
                                const vltr_radii& radii) {
  auto* s = new vltr_structure();
  s->format = format;
  if (format == VLTR_FORMAT_XYZR) {
    s->balls = parse_xyzr(lines);
    return s;
  }
  s->records = format == VLTR_FORMAT_PDB ? parse_pdb(lines, o) : parse_mmcif(lines, o);
  s->balls.reserve(s->records.size());
  for (const AtomRecord& r : s->records) s->balls.push_back(vltr_ball{r.x, r.y, r.z, radii.get(r.res_name, r.name)});
  return s;
}

const vltr_radii& default_radii() {
  static const vltr_radii* r = [] {
    auto* x = new vltr_radii();
    add_default_rules(x);
    return x;
  }();
  return *r;
}

bool read_file(const std::string& path, std::string* text, int* err) {
  errno = 0;
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    *err = errno;
    return false;
  }
  char buf[1 << 16];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) text->append(buf, k);
  const bool bad = std::ferror(f);
  *err = bad ? errno : 0;
  std::fclose(f);
  return !bad;
}

// parse_file_with_records (mod.rs:300-334). On failure *io_msg holds the
// io::Error text.
int parse_path(const std::string& path, const vltr_parse_options& o, const vltr_radii& radii, vltr_structure** out,
               std::string* io_msg) {
  std::string text;
  int err = 0;
  if (!read_file(path, &text, &err)) {
    *io_msg = os_error(err);
    return VLTR_ERR_IO;
  }
  int fmt = detect_from_extension(path);
  std::vector<std::string> lines = text_lines(text);
  if (fmt == VLTR_FORMAT_AUTO) {
    fmt = detect_from_first_line(lines.empty() ? std::string() : lines[0], false);
    if (fmt == VLTR_FORMAT_AUTO) {
      *io_msg = "Unable to detect input format";
      return VLTR_ERR_IO;
    }
  }
  *out = build_structure(lines, fmt, o, radii);
  return VLTR_OK;
}

// ------------------------------------------------------------- selection --

// Glob match with '*', '?' and '[abc]' / '[a-z]' classes, backtracking to
// the last '*' (selection.rs:56-155).
bool char_class(const std::string& p, size_t start, char c, bool* matched, size_t* end) {
  if (start >= p.size() || p[start] != '[') return false;
  size_t i = start + 1;
  bool m = false;
  while (i < p.size() && p[i] != ']') {
    if (i + 2 < p.size() && p[i + 1] == '-' && p[i + 2] != ']') {
      if (c >= p[i] && c <= p[i + 2]) m = true;
      i += 3;
    } else {
      if (p[i] == c) m = true;
      ++i;
    }
  }
  if (i < p.size() && p[i] == ']') {
    *matched = m;
    *end = i;
    return true;
  }
  return false;
}

bool glob_match(const std::string& p, const std::string& t) {
  bool have_star = false;
  size_t sp = 0, st = 0, pi = 0, ti = 0;
  while (ti < t.size()) {
    if (pi < p.size()) {
      const char c = p[pi];
      if (c == '?') {
        ++pi;
        ++ti;
        continue;
      }
      if (c == '*') {
        have_star = true;
        sp = pi;
        st = ti;
        ++pi;
        continue;
      }
      bool handled = false;
      if (c == '[') {
        bool m = false;
        size_t e = 0;
        if (char_class(p, pi, t[ti], &m, &e) && m) {
          pi = e + 1;
          ++ti;
          handled = true;
        }
      } else if (c == t[ti]) {
        ++pi;
        ++ti;
        handled = true;
      }
      if (handled) continue;
    }
    if (!have_star) return false;
    pi = sp + 1;
    st = st + 1;
    ti = st;
    if (ti > t.size()) return false;
  }
  while (pi < p.size() && p[pi] == '*') ++pi;
  return pi == p.size();
}

const char* const kProtein[] = {"ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS", "ILE",
                                "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL",
                                "HIE", "HID", "HIP", "CYX", "ASH", "GLH", "LYN"};
const char* const kDna[] = {"DA", "DT", "DG", "DC", "DU"};
const char* const kRna[] = {"A", "U", "G", "C", "RA", "RU", "RG", "RC"};
const char* const kBackbone[] = {"C", "CA", "N", "O"};
const char* const kHydrophobic[] = {"ALA", "VAL", "ILE", "LEU", "MET", "PHE", "TRP", "PRO", "GLY"};
const char* const kAromatic[] = {"PHE", "TYR", "TRP", "HIS", "HIE", "HID", "HIP"};
const char* const kAcidic[] = {"ASP", "GLU", "ASH", "GLH"};
const char* const kBasic[] = {"ARG", "LYS", "HIS", "HIE", "HID", "HIP", "LYN"};
const char* const kPolar[] = {"SER", "THR", "ASN", "GLN", "CYS", "CYX", "TYR"};

template <size_t N>
bool in_set(const std::string& s, const char* const (&set)[N]) {
  for (const char* x : set)
    if (s == x) return true;
  return false;
}

enum TokKind {
  tChain, tResname, tResid, tName, tProtein, tBackbone, tSidechain, tNucleic, tHetatm, tHydrophobic, tAromatic,
  tAcidic, tBasic, tPolar, tCharged, tAll, tNone, tAnd, tOr, tNot, tLParen, tRParen, tTo, tIdent, tNumber, tColon
};

struct Tok {
  TokKind k;
  std::string s;
  int32_t num = 0;
  size_t pos = 0;
};

struct SelError {
  std::string message;
  size_t position = 0;
  std::string text() const { return message + " at position " + std::to_string(position); }
};

// selection.rs:232-257
TokKind keyword(const std::string& ident) {
  std::string l = ident;
  for (char& c : l) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  static const std::unordered_map<std::string, TokKind> m = {
      {"chain", tChain},         {"segid", tChain},     {"resname", tResname},   {"resn", tResname},
      {"resid", tResid},         {"resi", tResid},      {"resseq", tResid},      {"resnum", tResid},
      {"name", tName},           {"atomname", tName},   {"protein", tProtein},   {"backbone", tBackbone},
      {"sidechain", tSidechain}, {"nucleic", tNucleic}, {"nucleicacid", tNucleic}, {"hetatm", tHetatm},
      {"hetero", tHetatm},       {"hydrophobic", tHydrophobic}, {"aromatic", tAromatic}, {"acidic", tAcidic},
      {"basic", tBasic},         {"polar", tPolar},     {"charged", tCharged},   {"all", tAll},
      {"everything", tAll},      {"none", tNone},       {"nothing", tNone},      {"and", tAnd},
      {"&&", tAnd},              {"or", tOr},           {"||", tOr},             {"not", tNot},
      {"!", tNot},               {"to", tTo}};
  auto it = m.find(l);
  return it == m.end() ? tIdent : it->second;
}

bool ident_start(char c) {
  return std::isalnum(static_cast<unsigned char>(c)) || c == '_' || c == '*' || c == '?' || c == '[' || c == '\'';
}
bool ident_char(char c, bool in_bracket) {
  return in_bracket || std::isalnum(static_cast<unsigned char>(c)) || c == '_' || c == '*' || c == '?' || c == '[' ||
         c == ']' || c == '-' || c == '\'';
}

// One UTF-8 character starting at i (for error messages).
std::string utf8_char(const std::string& s, size_t i) {
  const unsigned char c = static_cast<unsigned char>(s[i]);
  size_t n = c < 0x80 ? 1 : (c & 0xe0) == 0xc0 ? 2 : (c & 0xf0) == 0xe0 ? 3 : 4;
  return s.substr(i, std::min(n, s.size() - i));
}

// selection.rs:271-403
bool tokenize(const std::string& in, std::vector<Tok>* out, SelError* err) {
  size_t i = 0;
  while (i < in.size()) {
    const char c = in[i];
    if (is_space(c)) {
      ++i;
      continue;
    }
    if (c == '(' || c == ')' || c == ':') {
      out->push_back(Tok{c == '(' ? tLParen : c == ')' ? tRParen : tColon, "", 0, i});
      ++i;
      continue;
    }
    if (c == '"' || c == '\'') {
      const size_t start = i + 1;
      size_t j = start;
      while (j < in.size() && in[j] != c) ++j;
      out->push_back(Tok{tIdent, in.substr(start, j - start), 0, start});
      i = j < in.size() ? j + 1 : j;
      continue;
    }
    const bool digit = std::isdigit(static_cast<unsigned char>(c));
    if (digit || (c == '-' && i + 1 < in.size() && std::isdigit(static_cast<unsigned char>(in[i + 1])))) {
      size_t j = i;
      std::string num;
      if (in[j] == '-') num.push_back(in[j++]);
      while (j < in.size() && std::isdigit(static_cast<unsigned char>(in[j]))) num.push_back(in[j++]);
      int32_t v = 0;
      if (!parse_i32(num, &v)) {
        *err = SelError{"Invalid number: " + num, i};
        return false;
      }
      out->push_back(Tok{tNumber, num, v, i});
      i = j;
      continue;
    }
    if (ident_start(c)) {
      size_t j = i;
      bool in_bracket = false;
      std::string id;
      while (j < in.size()) {
        const char d = in[j];
        if (d == '[') in_bracket = true;
        else if (d == ']') in_bracket = false;
        if (!ident_char(d, in_bracket)) break;
        id.push_back(d);
        ++j;
      }
      const TokKind k = keyword(id);
      out->push_back(Tok{k, k == tIdent ? id : std::string(), 0, i});
      i = j;
      continue;
    }
    *err = SelError{"Unexpected character: " + utf8_char(in, i), i};
    return false;
  }
  return true;
}

struct Expr {
  TokKind k;  // tChain / tResname / tName (patterns), tResid (ranges), the flags, tAnd / tOr / tNot
  std::vector<std::string> patterns;
  std::vector<std::pair<int32_t, int32_t>> ranges;
  std::unique_ptr<Expr> a, b;

  bool matches(const AtomRecord& r) const {
    const std::string& res = r.res_name;
    const std::string& atom = r.name;
    switch (k) {
      case tChain:
        for (const auto& p : patterns)
          if (glob_match(p, r.chain_id)) return true;
        return false;
      case tResname:
        for (const auto& p : patterns)
          if (glob_match(p, res)) return true;
        return false;
      case tName:
        for (const auto& p : patterns)
          if (glob_match(p, atom)) return true;
        return false;
      case tResid:
        for (const auto& g : ranges)
          if (r.res_seq >= g.first && r.res_seq <= g.second) return true;
        return false;
      case tProtein: return in_set(res, kProtein);
      case tBackbone: return in_set(res, kProtein) && in_set(atom, kBackbone);
      case tSidechain: return in_set(res, kProtein) && !in_set(atom, kBackbone);
      case tNucleic: return in_set(res, kDna) || in_set(res, kRna);
      case tHetatm: return r.record_name == "HETATM";
      case tHydrophobic: return in_set(res, kHydrophobic);
      case tAromatic: return in_set(res, kAromatic);
      case tAcidic: return in_set(res, kAcidic);
      case tBasic: return in_set(res, kBasic);
      case tPolar: return in_set(res, kPolar);
      case tCharged: return in_set(res, kAcidic) || in_set(res, kBasic);
      case tAll: return true;
      case tNone: return false;
      case tAnd: return a->matches(r) && b->matches(r);
      case tOr: return a->matches(r) || b->matches(r);
      case tNot: return !a->matches(r);
      default: return false;
    }
  }
};

// Recursive descent: or > and > not > primary (selection.rs:465-699).
struct SelParser {
  const std::vector<Tok>& t;
  size_t i = 0;
  SelError err;

  const Tok* peek() const { return i < t.size() ? &t[i] : nullptr; }
  size_t cur_pos() const { return i < t.size() ? t[i].pos : 0; }
  bool is(TokKind k) const { return i < t.size() && t[i].k == k; }
  std::unique_ptr<Expr> fail_at(const std::string& m, size_t pos) {
    err = SelError{m, pos};
    return nullptr;
  }

  std::unique_ptr<Expr> parse() {
    if (t.empty()) return fail_at("Empty selection", 0);
    auto e = parse_or();
    if (!e) return nullptr;
    if (i < t.size()) return fail_at("Unexpected token after expression", cur_pos());
    return e;
  }
  std::unique_ptr<Expr> binary(TokKind op, std::unique_ptr<Expr> l, std::unique_ptr<Expr> r) {
    auto e = std::make_unique<Expr>();
    e->k = op;
    e->a = std::move(l);
    e->b = std::move(r);
    return e;
  }
  std::unique_ptr<Expr> parse_or() {
    auto l = parse_and();
    if (!l) return nullptr;
    while (is(tOr)) {
      ++i;
      auto r = parse_and();
      if (!r) return nullptr;
      l = binary(tOr, std::move(l), std::move(r));
    }
    return l;
  }
  std::unique_ptr<Expr> parse_and() {
    auto l = parse_not();
    if (!l) return nullptr;
    while (is(tAnd)) {
      ++i;
      auto r = parse_not();
      if (!r) return nullptr;
      l = binary(tAnd, std::move(l), std::move(r));
    }
    return l;
  }
  std::unique_ptr<Expr> parse_not() {
    if (is(tNot)) {
      ++i;
      auto inner = parse_not();
      if (!inner) return nullptr;
      auto e = std::make_unique<Expr>();
      e->k = tNot;
      e->a = std::move(inner);
      return e;
    }
    return parse_primary();
  }
  std::unique_ptr<Expr> leaf(TokKind k) {
    ++i;
    auto e = std::make_unique<Expr>();
    e->k = k;
    return e;
  }
  std::unique_ptr<Expr> parse_primary() {
    const size_t pos = cur_pos();
    const Tok* tk = peek();
    if (!tk) return fail_at("Expected selection expression", pos);
    switch (tk->k) {
      case tLParen: {
        ++i;
        auto inner = parse_or();
        if (!inner) return nullptr;
        if (!is(tRParen)) return fail_at("Missing closing parenthesis", cur_pos());
        ++i;
        return inner;
      }
      case tChain: return patterns("chain", tChain);
      case tResname: return patterns("resname", tResname);
      case tName: return patterns("name", tName);
      case tResid: return resid();
      case tProtein: case tBackbone: case tSidechain: case tNucleic: case tHetatm: case tHydrophobic:
      case tAromatic: case tAcidic: case tBasic: case tPolar: case tCharged: case tAll: case tNone:
        return leaf(tk->k);
      case tAnd: case tOr: return fail_at("Unexpected boolean operator", pos);
      case tIdent: return fail_at("Unknown keyword: " + tk->s, pos);
      default: return fail_at("Expected selection expression", pos);
    }
  }
  std::unique_ptr<Expr> patterns(const char* name, TokKind k) {
    const size_t pos = cur_pos();
    ++i;
    auto e = std::make_unique<Expr>();
    e->k = k;
    while (is(tIdent)) e->patterns.push_back(t[i++].s);
    if (e->patterns.empty()) return fail_at(std::string(name) + " requires at least one argument", pos);
    return e;
  }
  static bool colon_range(const std::string& s, std::pair<int32_t, int32_t>* g) {
    const size_t c = s.find(':');
    if (c == std::string::npos || s.find(':', c + 1) != std::string::npos) return false;
    return parse_i32(s.substr(0, c), &g->first) && parse_i32(s.substr(c + 1), &g->second);
  }
  std::unique_ptr<Expr> resid() {
    const size_t pos = cur_pos();
    ++i;
    auto e = std::make_unique<Expr>();
    e->k = tResid;
    while (const Tok* tk = peek()) {
      if (tk->k == tNumber) {
        const int32_t start = tk->num;
        ++i;
        if (is(tTo) || is(tColon)) {
          const bool to = is(tTo);
          ++i;
          if (is(tNumber)) {
            e->ranges.push_back({start, t[i].num});
            ++i;
          } else {
            return fail_at(to ? "Expected number after 'to'" : "Expected number after ':'", cur_pos());
          }
        } else {
          e->ranges.push_back({start, start});
        }
      } else if (tk->k == tIdent) {
        std::pair<int32_t, int32_t> g;
        int32_t v = 0;
        if (colon_range(tk->s, &g)) {
          e->ranges.push_back(g);
          ++i;
        } else if (parse_i32(tk->s, &v)) {
          e->ranges.push_back({v, v});
          ++i;
        } else {
          break;
        }
      } else {
        break;
      }
    }
    if (e->ranges.empty()) return fail_at("resid requires at least one argument", pos);
    return e;
  }
};

std::unique_ptr<Expr> parse_selection(const std::string& text, SelError* err) {
  std::vector<Tok> toks;
  if (!tokenize(text, &toks, err)) return nullptr;
  SelParser p{toks};
  auto e = p.parse();
  if (!e) *err = p.err;
  return e;
}

// parse_selections (selection.rs:742-766): at least two; an error in
// selection k reports position k.
bool parse_selections(const std::vector<std::string>& in, std::vector<std::unique_ptr<Expr>>* out, SelError* err) {
  if (in.size() < 2) {
    *err = SelError{"At least two selections required", 0};
    return false;
  }
  for (size_t k = 0; k < in.size(); ++k) {
    SelError e;
    auto x = parse_selection(trim(in[k]), &e);
    if (!x) {
      e.position = k;
      *err = e;
      return false;
    }
    out->push_back(std::move(x));
  }
  return true;
}

// build_custom_grouping (selection.rs:773-814): first matching selection's
// index; unmatched atoms get unique ids after the selections.
std::vector<int32_t> custom_grouping(const std::vector<AtomRecord>& recs, const std::vector<std::unique_ptr<Expr>>& sels) {
  std::vector<int32_t> g(recs.size(), -1);
  for (size_t i = 0; i < recs.size(); ++i)
    for (size_t k = 0; k < sels.size(); ++k)
      if (sels[k]->matches(recs[i])) {
        g[i] = static_cast<int32_t>(k);
        break;
      }
  int32_t next = static_cast<int32_t>(sels.size());
  for (auto& x : g)
    if (x == -1) x = next++;
  return g;
}

std::vector<std::string> string_list(const char* const* s, size_t n) {
  std::vector<std::string> v;
  for (size_t k = 0; k < n; ++k) v.push_back(s[k] ? s[k] : "");
  return v;
}

}  // namespace

extern "C" {

int vltr_radii_create(int with_default_rules, vltr_radii** out) {
  if (!out) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  *out = new vltr_radii();
  if (with_default_rules) add_default_rules(*out);
  return VLTR_OK;
}
int vltr_radii_load_text(vltr_radii* r, const char* text) {
  if (!r || !text) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  return load_radii_text(r, text);
}
double vltr_radii_get(const vltr_radii* r, const char* res_name, const char* atom_name) {
  return (r ? *r : default_radii()).get(res_name ? res_name : "", atom_name ? atom_name : "");
}
void vltr_radii_free(vltr_radii* r) { delete r; }

int vltr_parse_file(const char* path, const vltr_parse_options* options, const vltr_radii* radii,
                    vltr_structure** out) {
  if (!path || !out) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  const vltr_parse_options o = options ? *options : vltr_parse_options{0, 0, 0};
  std::string msg;
  const int rc = parse_path(path, o, radii ? *radii : default_radii(), out, &msg);
  if (rc != VLTR_OK) return host_fail(rc, msg);
  return VLTR_OK;
}

int vltr_parse_text(const char* text, size_t len, int format, const vltr_parse_options* options,
                    const vltr_radii* radii, vltr_structure** out) {
  if ((!text && len) || !out) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  if (format < VLTR_FORMAT_AUTO || format > VLTR_FORMAT_XYZR)
    return host_fail(VLTR_ERR_INVALID_ARGUMENT, "unknown input format");
  const vltr_parse_options o = options ? *options : vltr_parse_options{0, 0, 0};
  const std::vector<std::string> lines = text_lines(std::string(text ? text : "", len));
  if (format == VLTR_FORMAT_AUTO) format = detect_from_first_line(lines.empty() ? std::string() : lines[0], true);
  *out = build_structure(lines, format, o, radii ? *radii : default_radii());
  return VLTR_OK;
}

int vltr_structure_format(const vltr_structure* s) { return s ? s->format : VLTR_FORMAT_AUTO; }
size_t vltr_structure_num_balls(const vltr_structure* s) { return s ? s->balls.size() : 0; }
const vltr_ball* vltr_structure_balls(const vltr_structure* s) {
  return s && !s->balls.empty() ? s->balls.data() : nullptr;
}
size_t vltr_structure_num_records(const vltr_structure* s) { return s ? s->records.size() : 0; }
const char* vltr_record_string(const vltr_structure* s, size_t i, int field) {
  if (!s || i >= s->records.size()) return nullptr;
  const AtomRecord& r = s->records[i];
  switch (field) {
    case VLTR_REC_RECORD_NAME: return r.record_name.c_str();
    case VLTR_REC_NAME: return r.name.c_str();
    case VLTR_REC_ALT_LOC: return r.alt_loc.c_str();
    case VLTR_REC_RES_NAME: return r.res_name.c_str();
    case VLTR_REC_CHAIN_ID: return r.chain_id.c_str();
    case VLTR_REC_I_CODE: return r.i_code.c_str();
    case VLTR_REC_ELEMENT: return r.element.c_str();
    default: return nullptr;
  }
}
int vltr_record_numbers(const vltr_structure* s, size_t i, int32_t* serial, int32_t* res_seq, double* xyzb) {
  if (!s || i >= s->records.size()) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "record index out of range");
  const AtomRecord& r = s->records[i];
  if (serial) *serial = r.serial;
  if (res_seq) *res_seq = r.res_seq;
  if (xyzb) {
    xyzb[0] = r.x;
    xyzb[1] = r.y;
    xyzb[2] = r.z;
    xyzb[3] = r.b_factor;
  }
  return VLTR_OK;
}
void vltr_structure_free(vltr_structure* s) { delete s; }

int vltr_grouping_chain(const vltr_structure* s, int32_t* out) {
  if (!s || (!out && !s->records.empty())) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  std::unordered_map<std::string, int32_t> ids;
  for (size_t i = 0; i < s->records.size(); ++i) {
    auto it = ids.emplace(s->records[i].chain_id, static_cast<int32_t>(ids.size())).first;
    out[i] = it->second;
  }
  return VLTR_OK;
}

int vltr_grouping_residue(const vltr_structure* s, int32_t* out) {
  if (!s || (!out && !s->records.empty())) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  std::map<std::tuple<std::string, int32_t, std::string>, int32_t> ids;
  for (size_t i = 0; i < s->records.size(); ++i) {
    const AtomRecord& r = s->records[i];
    auto it = ids.emplace(std::make_tuple(r.chain_id, r.res_seq, r.i_code), static_cast<int32_t>(ids.size())).first;
    out[i] = it->second;
  }
  return VLTR_OK;
}

int vltr_grouping_selections(const vltr_structure* s, const char* const* selections, size_t n_selections,
                             int32_t* out) {
  if (!s || (n_selections && !selections) || (!out && !s->records.empty()))
    return host_fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  std::vector<std::unique_ptr<Expr>> sels;
  SelError err;
  if (!parse_selections(string_list(selections, n_selections), &sels, &err))
    return host_fail(VLTR_ERR_INVALID_ARGUMENT, err.text());
  const std::vector<int32_t> g = custom_grouping(s->records, sels);
  std::copy(g.begin(), g.end(), out);
  return VLTR_OK;
}

int vltr_selection_match(const vltr_structure* s, const char* selection, uint8_t* out) {
  if (!s || !selection || (!out && !s->records.empty())) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  SelError err;
  auto e = parse_selection(selection, &err);
  if (!e) return host_fail(VLTR_ERR_INVALID_ARGUMENT, err.text());
  for (size_t i = 0; i < s->records.size(); ++i) out[i] = e->matches(s->records[i]) ? 1 : 0;
  return VLTR_OK;
}

int vltr_compute_tessellation_from_file(vltr_context* ctx, const char* path, double probe,
                                        const vltr_periodic_box* periodic_box, int with_cell_vertices,
                                        const char* const* selections, size_t n_selections, vltr_result** out) {
  if (!ctx || !path || !out || (n_selections && !selections))
    return host_fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  // mod.rs:526-531: the selection count is checked before the file is read
  if (selections && n_selections < 2)
    return host_fail(VLTR_ERR_INVALID_ARGUMENT, "at least two selections required for grouping");
  vltr_structure* st = nullptr;
  std::string msg;
  const int rc = parse_path(path, vltr_parse_options{0, 0, 0}, default_radii(), &st, &msg);
  if (rc != VLTR_OK) return host_fail(rc, "I/O error: " + msg);
  std::unique_ptr<vltr_structure> hold(st);
  std::vector<int32_t> groups;
  if (selections) {
    if (st->records.empty())
      return host_fail(VLTR_ERR_INVALID_ARGUMENT, "selections require PDB or mmCIF input (not XYZR)");
    std::vector<std::unique_ptr<Expr>> sels;
    SelError err;
    if (!parse_selections(string_list(selections, n_selections), &sels, &err))
      return host_fail(VLTR_ERR_INVALID_ARGUMENT, "selection error: " + err.text());
    groups = custom_grouping(st->records, sels);
  }
  return vltr::host_tessellate(ctx, st->balls.data(), st->balls.size(), probe, periodic_box,
                               groups.empty() ? nullptr : groups.data(), groups.size(), with_cell_vertices, out);
}

void vltr_string_free(char* text) { std::free(text); }

}  // extern "C"
