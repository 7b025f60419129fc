// Output writers of the front end: the CLI's JSON document (src/main.rs:22-34
// through serde_json's pretty printer) and the PyMOL CGO script of the
// contact polygons (src/graphics.rs). The polygons come from the restricted
// contours the device clip kernels keep for the network output; everything
// here is host formatting.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vltr_geom.h"
#include "vltr_host.h"
#include "voronota_ltr.h"
#include "voronota_ltr_input.h"

namespace {

using vltr::host_fail;
using vltr::Sph;
using vltr::V3;

char* to_c_string(const std::string& s, size_t* len) {
  char* p = static_cast<char*>(std::malloc(s.size() + 1));
  if (!p) return nullptr;
  std::memcpy(p, s.data(), s.size());
  p[s.size()] = '\0';
  if (len) *len = s.size();
  return p;
}

// serde_json's f64 text (ryu): the shortest digits that read back to the
// value; positional for decimal exponents -5 < k <= 16, else d.ddde±x;
// non-finite values are written as null.
std::string json_f64(double v) {
  if (!std::isfinite(v)) return "null";
  char buf[64];
  int prec = 1;
  for (; prec < 17; ++prec) {
    std::snprintf(buf, sizeof buf, "%.*e", prec - 1, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  std::snprintf(buf, sizeof buf, "%.*e", prec - 1, v);
  std::string s(buf);
  std::string sign;
  if (s[0] == '-') {
    sign = "-";
    s.erase(0, 1);
  }
  const size_t e = s.find('e');
  const int ex = std::atoi(s.c_str() + e + 1);
  std::string d = s.substr(0, e);
  d.erase(std::remove(d.begin(), d.end(), '.'), d.end());
  while (d.size() > 1 && d.back() == '0') d.pop_back();
  const int len = static_cast<int>(d.size());
  const int k = ex - (len - 1);  // value = d x 10^k
  const int kk = len + k;        // position of the decimal point
  std::string o;
  if (k >= 0 && kk <= 16) {
    o = d + std::string(static_cast<size_t>(k), '0') + ".0";
  } else if (kk > 0 && kk <= 16) {
    o = d.substr(0, static_cast<size_t>(kk)) + "." + d.substr(static_cast<size_t>(kk));
  } else if (kk > -5 && kk <= 0) {
    o = "0." + std::string(static_cast<size_t>(-kk), '0') + d;
  } else if (len == 1) {
    o = d + "e" + std::to_string(kk - 1);
  } else {
    o = d.substr(0, 1) + "." + d.substr(1) + "e" + std::to_string(kk - 1);
  }
  return sign + o;
}

std::string json_id(uint64_t id) { return id == VLTR_NULL_ID ? "null" : std::to_string(id); }

// serde_json PrettyFormatter: two-space indent, one element per line, "[]"
// for an empty array.
struct Pretty {
  std::string out;
  int depth = 0;
  std::vector<bool> first;

  void indent() { out.append(static_cast<size_t>(2 * depth), ' '); }
  void open(char c) {
    out.push_back(c);
    ++depth;
    first.push_back(true);
  }
  void close(char c) {
    --depth;
    const bool empty = first.back();
    first.pop_back();
    if (!empty) {
      out.push_back('\n');
      indent();
    }
    out.push_back(c);
  }
  void item() {
    if (!first.back()) out.push_back(',');
    first.back() = false;
    out.push_back('\n');
    indent();
  }
  void key(const char* k) {
    item();
    out += "\"";
    out += k;
    out += "\": ";
  }
};

// Iterator::sum over f64: sequential, starting from -0.0.
template <class T, class F>
double fsum(const std::vector<T>& v, F f) {
  double s = -0.0;
  for (const T& x : v) s += f(x);
  return s;
}

}  // namespace

extern "C" {

int vltr_result_to_json(const vltr_result* r, char** text, size_t* len) {
  if (!r || !text) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  Pretty p;
  p.open('{');
  p.key("num_balls");
  p.out += std::to_string(r->num_balls);
  p.key("contacts");
  p.open('[');
  for (const vltr_contact& c : r->contacts) {
    p.item();
    p.open('{');
    p.key("id_a");
    p.out += std::to_string(c.id_a);
    p.key("id_b");
    p.out += std::to_string(c.id_b);
    p.key("area");
    p.out += json_f64(c.area);
    p.key("arc_length");
    p.out += json_f64(c.arc_length);
    p.key("central");
    p.out += c.central ? "true" : "false";
    p.close('}');
  }
  p.close(']');
  p.key("cells");
  p.open('[');
  for (const vltr_cell& c : r->cells) {
    p.item();
    p.open('{');
    p.key("index");
    p.out += std::to_string(c.index);
    p.key("sas_area");
    p.out += json_f64(c.sas_area);
    p.key("volume");
    p.out += json_f64(c.volume);
    p.close('}');
  }
  p.close(']');
  p.key("cell_vertices");
  if (!r->has_network) {
    p.out += "null";
  } else {
    p.open('[');
    for (const vltr_cell_vertex& v : r->vertices) {
      p.item();
      p.open('{');
      p.key("ball_indices");
      p.open('[');
      for (uint64_t id : v.ball_indices) {
        p.item();
        p.out += json_id(id);
      }
      p.close(']');
      p.key("x");
      p.out += json_f64(v.x);
      p.key("y");
      p.out += json_f64(v.y);
      p.key("z");
      p.out += json_f64(v.z);
      p.close('}');
    }
    p.close(']');
  }
  p.key("cell_edges");
  if (!r->has_network) {
    p.out += "null";
  } else {
    p.open('[');
    for (const vltr_cell_edge& e : r->edges) {
      p.item();
      p.open('{');
      p.key("ball_indices");
      p.open('[');
      for (uint64_t id : e.ball_indices) {
        p.item();
        p.out += json_id(id);
      }
      p.close(']');
      p.key("length");
      p.out += json_f64(e.length);
      p.close('}');
    }
    p.close(']');
  }
  // Results::sas_areas / volumes (types.rs:203-225): null without a cell
  std::vector<const vltr_cell*> by_ball(r->num_balls, nullptr);
  for (const vltr_cell& c : r->cells)
    if (c.index < r->num_balls) by_ball[c.index] = &c;
  p.key("sas_areas");
  p.open('[');
  for (const vltr_cell* c : by_ball) {
    p.item();
    p.out += c ? json_f64(c->sas_area) : "null";
  }
  p.close(']');
  p.key("volumes");
  p.open('[');
  for (const vltr_cell* c : by_ball) {
    p.item();
    p.out += c ? json_f64(c->volume) : "null";
  }
  p.close(']');
  p.key("total_sas_area");
  p.out += json_f64(fsum(r->cells, [](const vltr_cell& c) { return c.sas_area; }));
  p.key("total_volume");
  p.out += json_f64(fsum(r->cells, [](const vltr_cell& c) { return c.volume; }));
  p.key("total_contact_area");
  p.out += json_f64(fsum(r->contacts, [](const vltr_contact& c) { return c.area; }));
  p.close('}');
  *text = to_c_string(p.out, len);
  if (!*text) return host_fail(VLTR_ERR_OUT_OF_MEMORY, "json text");
  return VLTR_OK;
}

}  // extern "C"

namespace {

// geometry.rs:196-203 for any angle: the unit quaternion (cos t/2, u sin t/2).
V3 rotate_by(V3 axis, double angle, V3 p) {
  const double h = angle / 2.0;
  return vltr::rotate(axis, std::sin(h), std::cos(h), p);
}

struct Face {
  std::vector<V3> outer;
  V3 bary;
  V3 normal;
};

// graphics.rs:262-283
std::vector<V3> full_circle(V3 c, double r, V3 axis, double step) {
  const V3 first = vltr::any_normal(axis) * r;
  const size_t n = static_cast<size_t>(std::ceil(vltr::kTau / step));
  const double actual = vltr::kTau / static_cast<double>(n);
  std::vector<V3> out;
  for (size_t i = 0; i < n; ++i) out.push_back(c + rotate_by(axis, static_cast<double>(i) * actual, first));
  return out;
}

// graphics.rs:317-349
void arc_points(V3 c, double r, V3 axis, V3 start, V3 end, double total, double step, std::vector<V3>* out) {
  if (total <= step) return;
  const double ang = std::fmin(vltr::directed_angle(c, start, end, c + axis), total);
  const size_t n = static_cast<size_t>(std::floor(ang / step));
  if (n <= 1) return;
  const V3 off = start - c;
  const double actual = ang / static_cast<double>(n);
  for (size_t i = 1; i < n; ++i) out->push_back(c + vltr::unit(rotate_by(axis, static_cast<double>(i) * actual, off)) * r);
}

void fmt_xyz(std::string* o, const char* head, V3 p) {
  char buf[160];
  std::snprintf(buf, sizeof buf, "    %s, %.6f, %.6f, %.6f,\n", head, p.x, p.y, p.z);
  *o += buf;
}

}  // namespace

extern "C" {

int vltr_pymol_graphics(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe, const int32_t* groups,
                        size_t n_groups, double angle_step, const char* object_name, char** text, size_t* len) {
  if (!ctx || !text || (n && !balls) || !object_name) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  if (!(angle_step > 0.0)) return host_fail(VLTR_ERR_INVALID_ARGUMENT, "angle_step must be positive");
  *text = nullptr;
  vltr_result* r = nullptr;
  std::vector<vltr::ContactNet> nets;
  int rc = vltr::host_tessellate(ctx, balls, n, probe, nullptr, groups, n_groups, 1, &r, &nets);
  if (rc != VLTR_OK) return rc;
  // graphics.rs:45-84, 203-235: one face per contact in contact order
  std::vector<Face> faces;
  for (size_t k = 0; k < r->contacts.size() && k < nets.size(); ++k) {
    const vltr_contact& c = r->contacts[k];
    const vltr::ContactNet& nt = nets[k];
    const vltr_ball& A = balls[c.id_a];
    const vltr_ball& B = balls[c.id_b];
    const Sph sa{V3{A.x, A.y, A.z}, A.r + probe}, sb{V3{B.x, B.y, B.z}, B.r + probe};
    const Sph ic = vltr::intersection_circle(sa, sb);
    const V3 axis = vltr::unit(sb.c - sa.c);
    Face f;
    f.normal = axis;
    if (nt.m == 0) {
      f.outer = full_circle(ic.c, ic.r, axis, angle_step);
      f.bary = ic.c;
    } else {
      V3 sum{0.0, 0.0, 0.0};
      for (int t = 0; t < nt.m; ++t) {
        const V3 p{nt.x[t], nt.y[t], nt.z[t]};
        sum = sum + p;
        const int u = t + 1 < nt.m ? t + 1 : 0;
        f.outer.push_back(p);
        if (static_cast<uint64_t>(nt.r[t]) == c.id_a && nt.ang[t] > 0.0)
          arc_points(ic.c, ic.r, axis, p, V3{nt.x[u], nt.y[u], nt.z[u]}, nt.ang[t], angle_step, &f.outer);
      }
      f.bary = sum / static_cast<double>(nt.m);
    }
    if (f.outer.size() >= 3) faces.push_back(std::move(f));
  }
  vltr_result_free(r);
  // graphics.rs:87-178
  const std::string obj(object_name);
  std::string o = "from pymol.cgo import *\nfrom pymol import cmd\n\n";
  o += "cgo_graphics_list_balls = [\n    COLOR, 0, 1, 1,\n";
  char buf[200];
  for (size_t i = 0; i < n; ++i) {
    std::snprintf(buf, sizeof buf, "    SPHERE, %.6f, %.6f, %.6f, %.6f,\n", balls[i].x, balls[i].y, balls[i].z,
                  balls[i].r);
    o += buf;
  }
  o += "]\ncmd.load_cgo(cgo_graphics_list_balls, '" + obj + "_balls')\n\n";
  o += "cgo_graphics_list_faces = [\n    COLOR, 1, 1, 0,\n";
  for (const Face& f : faces) {
    o += "    BEGIN, TRIANGLE_FAN,\n";
    fmt_xyz(&o, "NORMAL", f.normal);
    fmt_xyz(&o, "VERTEX", f.bary);
    for (const V3& p : f.outer) fmt_xyz(&o, "VERTEX", p);
    fmt_xyz(&o, "VERTEX", f.outer[0]);
    o += "    END,\n";
  }
  o += "]\ncmd.load_cgo(cgo_graphics_list_faces, '" + obj + "_faces')\n\n";
  o += "cgo_graphics_list_wireframe = [\n    COLOR, 1, 0, 0,\n";
  for (const Face& f : faces) {
    o += "    BEGIN, LINE_LOOP,\n";
    for (const V3& p : f.outer) fmt_xyz(&o, "VERTEX", p);
    o += "    END,\n";
  }
  o += "]\ncmd.load_cgo(cgo_graphics_list_wireframe, '" + obj + "_wireframe')\n\n";
  o += "cmd.set('two_sided_lighting', 1)\ncmd.set('cgo_line_width', 1)\n";
  *text = to_c_string(o, len);
  if (!*text) return host_fail(VLTR_ERR_OUT_OF_MEMORY, "graphics text");
  return VLTR_OK;
}

}  // extern "C"
