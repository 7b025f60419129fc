// voronota-ltr command line tool on the MI355X library (replaces src/main.rs):
// reads PDB / mmCIF / XYZR from a file or stdin, optionally groups atoms
// (chains, residues or selections), runs the tessellation on the GPU and
// writes the JSON document (and optionally a PyMOL CGO script).
#include <cctype>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

#include "voronota_ltr.h"
#include "voronota_ltr_input.h"

namespace {

const char* kAbout = "Compute radical Voronoi tessellation of atomic balls";

const char* kHelp =
    "Constructs a radical Voronoi tessellation of atomic balls constrained inside a solvent-accessible surface\n"
    "defined by a rolling probe. Computes inter-atom contact areas, solvent accessible surface areas, and\n"
    "volumes.\n\n"
    "Supports PDB, mmCIF, and XYZR input formats (auto-detected from extension or content).\n\n"
    "Usage: voronota-ltr [OPTIONS] [INPUT]\n\n"
    "Arguments:\n"
    "  [INPUT]  Input file (PDB, mmCIF, or XYZR format). Reads from stdin if not specified\n\n"
    "Options:\n"
    "      --probe <PROBE>\n          Rolling probe radius [default: 1.4]\n"
    "  -o, --output <OUTPUT>\n          Output JSON file. Writes to stdout if not specified\n"
    "      --include-hydrogens\n          Include hydrogen atoms\n"
    "      --exclude-heteroatoms\n          Exclude heteroatoms (HETATM records)\n"
    "      --radii-file <RADII_FILE>\n          Custom radii file (format: residue atom radius per line)\n"
    "      --periodic-box-corners <X1> <Y1> <Z1> <X2> <Y2> <Z2>\n          Periodic box corners: x1 y1 z1 x2 y2 z2\n"
    "      --periodic-box-directions <X1> <Y1> <Z1> <X2> <Y2> <Z2> <X3> <Y3> <Z3>\n"
    "          Periodic box as three direction vectors: x1 y1 z1 x2 y2 z2 x3 y3 z3\n"
    "      --inter-chain-only\n          Only compute inter-chain contacts (exclude intra-chain)\n"
    "      --inter-residue-only\n          Only compute inter-residue contacts (exclude intra-residue)\n"
    "  -s, --selection <EXPR> <EXPR>...\n"
    "          Group atoms for inter-group contacts only (VMD-like syntax) Example: -s \"protein\" \"resname LIG\"\n"
    "  -v, --verbose...\n          Increase verbosity (-v: debug, -vv: trace)\n"
    "  -q, --quiet\n          Reduce verbosity to warnings only\n"
    "      --processors <PROCESSORS>\n          Maximum number of threads to use (default: all available)\n"
    "      --measure-running-time\n          Measure and output running time to stderr\n"
    "      --pymol <PATH>\n          Output `PyMOL` CGO graphics script for visualizing contacts\n"
    "      --device <DEVICE>\n          GPU to run on [default: 0]\n"
    "  -h, --help\n          Print help (see a summary with '-h')\n"
    "  -V, --version\n          Print version\n";

struct Args {
  double probe = 1.4;
  std::string input, output, radii_file, pymol;
  bool include_h = false, exclude_het = false, inter_chain = false, inter_residue = false;
  bool quiet = false, measure = false;
  int verbose = 0, device = 0;
  std::vector<double> corners, directions;
  std::vector<std::string> selections;
};

bool g_quiet = false;

void info(const std::string& m) {
  if (!g_quiet) std::fprintf(stderr, "[INFO  voronota_ltr] %s\n", m.c_str());
}

[[noreturn]] void usage_error(const std::string& m) {
  std::fprintf(stderr, "error: %s\n\nUsage: voronota-ltr [OPTIONS] [INPUT]\n\nFor more information, try '--help'.\n",
               m.c_str());
  std::exit(2);
}

// Errors returned from main (io::Error): printed in Debug form, exit 1.
[[noreturn]] void fail(const std::string& m) {
  std::fprintf(stderr, "Error: %s\n", m.c_str());
  std::exit(1);
}

double number(const std::string& flag, const std::string& v) {
  char* end = nullptr;
  const double d = std::strtod(v.c_str(), &end);
  if (v.empty() || *end) usage_error("invalid value '" + v + "' for '" + flag + "': invalid float literal");
  return d;
}

bool is_flag(const std::string& a) {
  return a.size() > 1 && a[0] == '-' && !(a.size() > 1 && (std::isdigit(static_cast<unsigned char>(a[1])) || a[1] == '.'));
}

Args parse_args(int argc, char** argv) {
  Args a;
  std::vector<std::string> v(argv + 1, argv + argc);
  size_t i = 0;
  auto value = [&](const std::string& flag) -> std::string {
    if (i + 1 >= v.size()) usage_error("a value is required for '" + flag + "' but none was supplied");
    return v[++i];
  };
  auto values = [&](const std::string& flag, size_t n) {
    std::vector<double> out;
    for (size_t k = 0; k < n; ++k) {
      if (i + 1 >= v.size() || is_flag(v[i + 1]))
        usage_error(std::to_string(n) + " values required for '" + flag + "' but " + std::to_string(k) +
                    " were provided");
      out.push_back(number(flag, v[++i]));
    }
    return out;
  };
  for (; i < v.size(); ++i) {
    std::string s = v[i];
    std::string inline_value;
    const size_t eq = s.find('=');
    if (s.rfind("--", 0) == 0 && eq != std::string::npos) {
      inline_value = s.substr(eq + 1);
      s = s.substr(0, eq);
      v.insert(v.begin() + static_cast<long>(i) + 1, inline_value);
    }
    if (s == "-h" || s == "--help") {
      std::printf("%s\n\n%s", kAbout, kHelp);
      std::exit(0);
    } else if (s == "-V" || s == "--version") {
      std::printf("voronota_ltr 0.6.0\n");
      std::exit(0);
    } else if (s == "--probe") {
      a.probe = number(s, value(s));
    } else if (s == "-o" || s == "--output") {
      a.output = value(s);
    } else if (s == "--include-hydrogens") {
      a.include_h = true;
    } else if (s == "--exclude-heteroatoms") {
      a.exclude_het = true;
    } else if (s == "--radii-file") {
      a.radii_file = value(s);
    } else if (s == "--periodic-box-corners") {
      a.corners = values(s, 6);
    } else if (s == "--periodic-box-directions") {
      a.directions = values(s, 9);
    } else if (s == "--inter-chain-only") {
      a.inter_chain = true;
    } else if (s == "--inter-residue-only") {
      a.inter_residue = true;
    } else if (s == "-s" || s == "--selection") {
      // num_args = 2..: every following value up to the next flag
      while (i + 1 < v.size() && !is_flag(v[i + 1])) a.selections.push_back(v[++i]);
      if (a.selections.size() < 2)
        usage_error("2 values required by '--selection <EXPR> <EXPR>...'; only " +
                    std::to_string(a.selections.size()) + " was provided");
    } else if (s == "-q" || s == "--quiet") {
      a.quiet = true;
    } else if (s == "--processors") {
      (void)number(s, value(s));
    } else if (s == "--measure-running-time") {
      a.measure = true;
    } else if (s == "--pymol") {
      a.pymol = value(s);
    } else if (s == "--device") {
      a.device = static_cast<int>(number(s, value(s)));
    } else if (s.size() > 1 && s[0] == '-' && s[1] != '-' && s.find_first_not_of("vq", 1) == std::string::npos) {
      for (size_t k = 1; k < s.size(); ++k) {
        if (s[k] == 'v') ++a.verbose;
        if (s[k] == 'q') a.quiet = true;
      }
    } else if (is_flag(s)) {
      usage_error("unexpected argument '" + s + "' found");
    } else if (a.input.empty()) {
      a.input = s;
    } else {
      usage_error("unexpected argument '" + s + "' found");
    }
  }
  if (!a.corners.empty() && !a.directions.empty())
    usage_error("the argument '--periodic-box-corners <X1> <Y1> <Z1> <X2> <Y2> <Z2>' cannot be used with "
                "'--periodic-box-directions <X1> <Y1> <Z1> <X2> <Y2> <Z2> <X3> <Y3> <Z3>'");
  if (!a.selections.empty() && (a.inter_chain || a.inter_residue))
    usage_error(std::string("the argument '--selection <EXPR> <EXPR>...' cannot be used with '") +
                (a.inter_chain ? "--inter-chain-only" : "--inter-residue-only") + "'");
  return a;
}

std::string last_error() { return vltr_last_error(); }

}  // namespace

int main(int argc, char** argv) {
  const Args a = parse_args(argc, argv);
  g_quiet = a.quiet;

  vltr_radii* radii = nullptr;
  vltr_radii_create(1, &radii);
  if (!a.radii_file.empty()) {
    std::FILE* f = std::fopen(a.radii_file.c_str(), "rb");
    if (!f) fail(std::string(std::strerror(errno)) + " (os error " + std::to_string(errno) + ")");
    std::string text;
    char buf[4096];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, k);
    std::fclose(f);
    if (vltr_radii_load_text(radii, text.c_str()) != VLTR_OK) fail(last_error());
    info("Loaded custom radii from " + a.radii_file);
  }
  const vltr_parse_options opts{a.exclude_het ? 1 : 0, a.include_h ? 1 : 0, 0};

  vltr_structure* st = nullptr;
  std::vector<int32_t> groups;
  if (!a.input.empty()) {
    if (vltr_parse_file(a.input.c_str(), &opts, radii, &st) != VLTR_OK) fail(last_error());
    const size_t nrec = vltr_structure_num_records(st);
    if (!a.selections.empty() || a.inter_chain || a.inter_residue) {
      if (nrec == 0) {
        const char* which = !a.selections.empty() ? "--selection" : a.inter_chain ? "--inter-chain-only"
                                                                                    : "--inter-residue-only";
        fail(std::string(which) + " requires PDB or mmCIF input (not XYZR)");
      }
      groups.resize(nrec);
      int rc;
      if (!a.selections.empty()) {
        std::vector<const char*> sp;
        for (const auto& s : a.selections) sp.push_back(s.c_str());
        rc = vltr_grouping_selections(st, sp.data(), sp.size(), groups.data());
        if (rc == VLTR_OK) info("Parsed " + std::to_string(sp.size()) + " selection groups");
      } else if (a.inter_chain) {
        rc = vltr_grouping_chain(st, groups.data());
      } else {
        rc = vltr_grouping_residue(st, groups.data());
      }
      if (rc != VLTR_OK) fail(last_error());
    }
  } else {
    if (!a.selections.empty() || a.inter_chain || a.inter_residue)
      fail("--selection, --inter-chain-only, and --inter-residue-only require file input");
    std::string text((std::istreambuf_iterator<char>(std::cin)), std::istreambuf_iterator<char>());
    if (vltr_parse_text(text.data(), text.size(), VLTR_FORMAT_AUTO, &opts, radii, &st) != VLTR_OK)
      fail(last_error());
  }
  vltr_radii_free(radii);
  const size_t n = vltr_structure_num_balls(st);
  const vltr_ball* balls = vltr_structure_balls(st);
  info("Read " + std::to_string(n) + " balls");

  vltr_periodic_box box;
  const vltr_periodic_box* pbox = nullptr;
  if (!a.directions.empty()) {
    const double* d = a.directions.data();
    vltr_periodic_box_from_vectors(d, d + 3, d + 6, &box);
    pbox = &box;
  } else if (!a.corners.empty()) {
    vltr_periodic_box_from_corners(a.corners.data(), a.corners.data() + 3, &box);
    pbox = &box;
  }

  vltr_context* ctx = nullptr;
  if (vltr_context_create(a.device, &ctx) != VLTR_OK) fail(last_error());
  const auto t0 = std::chrono::steady_clock::now();
  vltr_result* r = nullptr;
  if (vltr_compute_tessellation(ctx, balls, n, a.probe, pbox, groups.empty() ? nullptr : groups.data(), groups.size(),
                                0, &r) != VLTR_OK)
    fail(last_error());
  const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
  info("Computed " + std::to_string(vltr_result_num_contacts(r)) + " contacts, " +
       std::to_string(vltr_result_num_cells(r)) + " cells");
  if (a.measure) info("Tessellation time: " + std::to_string(ms) + " ms");

  if (!a.pymol.empty()) {
    char* g = nullptr;
    size_t glen = 0;
    if (vltr_pymol_graphics(ctx, balls, n, a.probe, groups.empty() ? nullptr : groups.data(), groups.size(), 0.5,
                            "contacts", &g, &glen) != VLTR_OK)
      fail(last_error());
    std::FILE* f = std::fopen(a.pymol.c_str(), "wb");
    if (!f) fail(std::string(std::strerror(errno)) + " (os error " + std::to_string(errno) + ")");
    std::fwrite(g, 1, glen, f);
    std::fclose(f);
    vltr_string_free(g);
    info("Wrote PyMOL graphics to " + a.pymol);
  }

  char* json = nullptr;
  size_t jlen = 0;
  if (vltr_result_to_json(r, &json, &jlen) != VLTR_OK) fail(last_error());
  if (!a.output.empty()) {
    std::FILE* f = std::fopen(a.output.c_str(), "wb");
    if (!f) fail(std::string(std::strerror(errno)) + " (os error " + std::to_string(errno) + ")");
    std::fwrite(json, 1, jlen, f);
    std::fclose(f);
  } else {
    std::fwrite(json, 1, jlen, stdout);
    std::fputc('\n', stdout);
  }
  vltr_string_free(json);
  vltr_result_free(r);
  vltr_structure_free(st);
  vltr_context_destroy(ctx);
  return 0;
}
