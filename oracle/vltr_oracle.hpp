// CPU oracle: a statement-by-statement C++ restatement of voronota-ltr's
// tessellation hot path, templated on the scalar type (see scalars.hpp).
//
// TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load the library built from this file, and only
// as the checker. The product library (voronota-ltr_amd/) never links it.
//
// The reference is a Rust crate (mlund/voronota-ltr v0.6.0) that cannot be
// built in this environment (no rustc/cargo, dependencies not vendored). Its
// arithmetic is restated here with the same operation order:
//   * Rust never contracts a*b+c; this file is compiled with -ffp-contract=off
//     and calls fma() exactly where the reference calls mul_add.
//   * nalgebra 0.34.2 small-vector ops: dot = (x*x' + y*y') + z*z',
//     norm = sqrt(norm_squared), normalize = component-wise division by norm,
//     cross = (ay*bz - az*by, az*bx - ax*bz, ax*by - ay*bx),
//     UnitQuaternion(axis, t) = (cos t/2, axis*sin t/2),
//     q*v = (t*w + q_v x t) + v with t = 2 (q_v x v).
//     nalgebra is not in this environment; these orders are parity-unpinned at
//     the ulp level (they only move hexagon vertices by ulps).
//   * acos/sin/cos come from the host libm, as in the reference.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <optional>
#include <utility>
#include <map>
#include <array>
#include <vector>

#include "scalars.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace vltr_oracle {

// ----------------------------------------------------------- constants ----
// geometry.rs:17
constexpr double EPSILON = 1e-10;
// std::f64::consts, correctly rounded (contact.rs:5, geometry.rs:5).
constexpr double PI = 0x1.921fb54442d18p+1;
constexpr double TAU = 0x1.921fb54442d18p+2;
constexpr double FRAC_PI_3 = 0x1.0c152382d7366p+0;
// contact.rs:448
constexpr double HEXAGON_SCALE = 1.19;
// types.rs:98
constexpr size_t NULL_ID = std::numeric_limits<size_t>::max();

// ------------------------------------------------------------- vectors ----
template <class R>
struct V3 {
  R x, y, z;
};
template <class R> inline V3<R> operator-(const V3<R>& a, const V3<R>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator+(const V3<R>& a, const V3<R>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator*(const V3<R>& a, const R& s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator/(const V3<R>& a, const R& s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline V3<R> operator-(const V3<R>& a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline R dot(const V3<R>& a, const V3<R>& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> inline R norm_squared(const V3<R>& a) { return dot(a, a); }
template <class R> inline R norm(const V3<R>& a) { return vsqrt(norm_squared(a)); }
template <class R> inline V3<R> normalize(const V3<R>& a) { return a / norm(a); }
template <class R> inline V3<R> cross(const V3<R>& a, const V3<R>& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ------------------------------------------------------------- types.rs ----
template <class R>
struct Sphere {
  V3<R> center;
  R r;
};

struct BallD {
  double x, y, z, r;
};

template <class R>
struct Ball {
  R x, y, z, r;
};

// types.rs:58-63 Sphere::from_ball
template <class R>
inline Sphere<R> sphere_from_ball(const Ball<R>& b, R probe) {
  return Sphere<R>{V3<R>{b.x, b.y, b.z}, b.r + probe};
}

// types.rs:148-184 ValuedId ordering: by value, NaN/ties fall back to index.
template <class R>
struct ValuedId {
  R value;
  size_t index;
};
template <class R>
inline bool valued_id_less(const ValuedId<R>& a, const ValuedId<R>& b) {
  const double av = value_of(a.value), bv = value_of(b.value);
  if (av < bv) return true;
  if (av > bv) return false;
  return a.index < b.index;
}

// types.rs:292-376 PeriodicBox
template <class R>
struct PeriodicBox {
  V3<R> shift_a, shift_b, shift_c;
};
// types.rs:322-345: mul_add chain, per component c*wc + (b*wb + (a*wa + s)).
template <class R>
inline Sphere<R> shift_sphere(const PeriodicBox<R>& pb, const Sphere<R>& s, double wa, double wb, double wc) {
  Sphere<R> o;
  o.center.x = vfma(pb.shift_c.x, R(wc), vfma(pb.shift_b.x, R(wb), vfma(pb.shift_a.x, R(wa), s.center.x)));
  o.center.y = vfma(pb.shift_c.y, R(wc), vfma(pb.shift_b.y, R(wb), vfma(pb.shift_a.y, R(wa), s.center.y)));
  o.center.z = vfma(pb.shift_c.z, R(wc), vfma(pb.shift_b.z, R(wb), vfma(pb.shift_a.z, R(wa), s.center.z)));
  o.r = s.r;
  return o;
}
// types.rs:349-359 NEIGHBOR_SHIFTS (26 entries, origin excluded).
inline void neighbor_shift(int g, int* sx, int* sy, int* sz) {
  int k = g < 13 ? g : g + 1;  // skip (0,0,0) which is k == 13
  *sx = k / 9 - 1;
  *sy = (k / 3) % 3 - 1;
  *sz = k % 3 - 1;
}
// types.rs:362-375 populate_periodic_spheres: originals, then per shift g all n.
template <class R>
inline std::vector<Sphere<R>> populate_periodic_spheres(const PeriodicBox<R>& pb, const std::vector<Sphere<R>>& s) {
  std::vector<Sphere<R>> out;
  out.reserve(s.size() * 27);
  out.insert(out.end(), s.begin(), s.end());
  for (int g = 0; g < 26; ++g) {
    int sx, sy, sz;
    neighbor_shift(g, &sx, &sy, &sz);
    for (const auto& sp : s) out.push_back(shift_sphere(pb, sp, double(sx), double(sy), double(sz)));
  }
  return out;
}

// ---------------------------------------------------------- geometry.rs ----
namespace fcmp {
// geometry.rs:22-54
template <class R> inline bool eq(const R& a, const R& b) { return vabs(a - b) <= R(EPSILON); }
template <class R> inline bool lt(const R& a, const R& b) { return a + R(EPSILON) < b; }
template <class R> inline bool gt(const R& a, const R& b) { return a - R(EPSILON) > b; }
template <class R> inline bool le(const R& a, const R& b) { return a < b + R(EPSILON); }
template <class R> inline bool ge(const R& a, const R& b) { return a + R(EPSILON) > b; }
}  // namespace fcmp

// geometry.rs:59-61
template <class R>
inline bool point_equals(const V3<R>& a, const V3<R>& b) {
  return fcmp::eq(a.x, b.x) && fcmp::eq(a.y, b.y) && fcmp::eq(a.z, b.z);
}
// geometry.rs:66-69
template <class R>
inline bool sphere_intersects_sphere(const Sphere<R>& a, const Sphere<R>& b) {
  const R sum_r = a.r + b.r;
  return fcmp::lt(norm_squared(b.center - a.center), sum_r * sum_r);
}
// geometry.rs:73-75
template <class R>
inline bool sphere_equals_sphere(const Sphere<R>& a, const Sphere<R>& b) {
  return fcmp::eq(a.r, b.r) && point_equals(a.center, b.center);
}
// geometry.rs:79-82
template <class R>
inline bool sphere_contains_sphere(const Sphere<R>& a, const Sphere<R>& b) {
  const R diff_r = a.r - b.r;
  return fcmp::ge(a.r, b.r) && fcmp::le(norm_squared(b.center - a.center), diff_r * diff_r);
}
// geometry.rs:97-103
template <class R>
inline R signed_distance_to_plane_unit(const V3<R>& pp, const V3<R>& n, const V3<R>& x) {
  return dot(n, x - pp);
}
// geometry.rs:86-92 (normalizes the normal first)
template <class R>
inline R signed_distance_to_plane(const V3<R>& pp, const V3<R>& n, const V3<R>& x) {
  return signed_distance_to_plane_unit(pp, normalize(n), x);
}
// geometry.rs:118-131
template <class R>
inline int halfspace_of_point_unit(const V3<R>& pp, const V3<R>& n, const V3<R>& x) {
  const R sd = signed_distance_to_plane_unit(pp, n, x);
  if (fcmp::gt(sd, R(0.0))) return 1;
  if (fcmp::lt(sd, R(0.0))) return -1;
  return 0;
}
// geometry.rs:134-148
template <class R>
inline V3<R> intersection_of_plane_and_segment(const V3<R>& pp, const V3<R>& n, const V3<R>& a, const V3<R>& b) {
  const R da = signed_distance_to_plane(pp, n, a);
  const R db = signed_distance_to_plane(pp, n, b);
  if (vabs(da - db) < R(EPSILON)) return a;
  const R t = da / (da - db);
  return a + (b - a) * t;
}
// geometry.rs:152-154
template <class R>
inline R triangle_area(const V3<R>& a, const V3<R>& b, const V3<R>& c) {
  return norm(cross(b - a, c - a)) / R(2.0);
}
// Rust f64::clamp(-1, 1): NaN passes through.
template <class R>
inline R clamp_unit(const R& x) {
  R y = x;
  if (y < R(-1.0)) y = R(-1.0);
  if (y > R(1.0)) y = R(1.0);
  return y;
}
// geometry.rs:157-161
template <class R>
inline R min_angle(const V3<R>& o, const V3<R>& a, const V3<R>& b) {
  const V3<R> v1 = normalize(a - o);
  const V3<R> v2 = normalize(b - o);
  return vacos(clamp_unit(dot(v1, v2)));
}
// geometry.rs:165-176
template <class R>
inline R directed_angle(const V3<R>& o, const V3<R>& a, const V3<R>& b, const V3<R>& c) {
  const R angle = min_angle(o, a, b);
  const V3<R> v1 = normalize(a - o);
  const V3<R> v2 = normalize(b - o);
  const V3<R> n = cross(v1, v2);
  if (dot(c - o, n) >= R(0.0)) return angle;
  return R(TAU) - angle;
}
// geometry.rs:179-193
template <class R>
inline V3<R> any_normal_of_vector(const V3<R>& a) {
  V3<R> b = a;
  const R z = R(0.0);
  if (!fcmp::eq(b.x, z) && (!fcmp::eq(b.y, z) || !fcmp::eq(b.z, z))) {
    b.x = -b.x;
    return normalize(cross(a, b));
  } else if (!fcmp::eq(b.y, z) && (!fcmp::eq(b.x, z) || !fcmp::eq(b.z, z))) {
    b.y = -b.y;
    return normalize(cross(a, b));
  } else if (!fcmp::eq(b.x, z)) {
    return V3<R>{R(0.0), R(1.0), R(0.0)};
  }
  return V3<R>{R(1.0), R(0.0), R(0.0)};
}
// geometry.rs:196-203 (nalgebra UnitQuaternion::from_axis_angle and q*v).
template <class R>
inline V3<R> rotate_point_around_axis(const V3<R>& axis, const R& angle, const V3<R>& p) {
  if (norm_squared(axis) <= R(0.0)) return p;
  const V3<R> u = normalize(axis);
  const R half = angle / R(2.0);
  const R s = vsin(half);
  const R w = vcos(half);
  const V3<R> qv = u * s;
  const V3<R> t = cross(qv, p) * R(2.0);
  const V3<R> c = cross(qv, t);
  return (t * w + c) + p;
}
// geometry.rs:208-215
template <class R>
inline R distance_to_intersection_circle_center(const Sphere<R>& a, const Sphere<R>& b) {
  const R cm = norm(b.center - a.center);
  if (cm < R(EPSILON)) return R(0.0);
  const R cos_g = vfma(b.r, -b.r, vfma(a.r, a.r, cm * cm)) / (R(2.0) * a.r * cm);
  return a.r * cos_g;
}
// geometry.rs:218-226
template <class R>
inline V3<R> center_of_intersection_circle(const Sphere<R>& a, const Sphere<R>& b) {
  const V3<R> cv = b.center - a.center;
  const R cm = norm(cv);
  if (cm < R(EPSILON)) return a.center;
  const R cos_g = vfma(b.r, -b.r, vfma(a.r, a.r, cm * cm)) / (R(2.0) * a.r * cm);
  return a.center + cv * (a.r * cos_g / cm);
}
// geometry.rs:229-239
template <class R>
inline Sphere<R> intersection_circle_of_two_spheres(const Sphere<R>& a, const Sphere<R>& b) {
  const V3<R> cv = b.center - a.center;
  const R cm = norm(cv);
  if (cm < R(EPSILON)) return Sphere<R>{a.center, R(0.0)};
  const R cos_g = vfma(b.r, -b.r, vfma(a.r, a.r, cm * cm)) / (R(2.0) * a.r * cm);
  const R sin_g = vsqrt(vmax(R(1.0) - cos_g * cos_g, R(0.0)));
  const V3<R> center = a.center + cv * (a.r * cos_g / cm);
  return Sphere<R>{center, a.r * sin_g};
}
// geometry.rs:243-255
template <class R>
inline bool project_point_inside_line(const V3<R>& o, const V3<R>& a, const V3<R>& b, V3<R>* out) {
  const V3<R> v = normalize(b - a);
  const R l = dot(v, o - a);
  if (l > R(0.0) && l * l <= norm_squared(b - a)) {
    *out = a + v * l;
    return true;
  }
  return false;
}
// geometry.rs:258-280
template <class R>
inline bool intersect_segment_with_circle(const Sphere<R>& circle, const V3<R>& p_in, const V3<R>& p_out, V3<R>* out) {
  const R dist = norm(p_in - p_out);
  if (dist <= R(0.0)) return false;
  const V3<R> v = (p_in - p_out) / dist;
  const V3<R> u = circle.center - p_out;
  const V3<R> s = p_out + v * dot(v, u);
  const R ll = vfma(circle.r, circle.r, -norm_squared(circle.center - s));
  if (ll >= R(0.0)) {
    *out = s - v * vsqrt(ll);
    return true;
  }
  return false;
}
// geometry.rs:283-293
template <class R>
inline R min_dihedral_angle(const V3<R>& o, const V3<R>& a, const V3<R>& b1, const V3<R>& b2) {
  const V3<R> oa = normalize(a - o);
  const V3<R> d1 = b1 - (o + oa * dot(oa, b1 - o));
  const V3<R> d2 = b2 - (o + oa * dot(oa, b2 - o));
  return vacos(clamp_unit(dot(normalize(d1), normalize(d2))));
}

// --------------------------------------------------- spheres_searcher.rs ----
// Rust `f64 as i32` saturates and maps NaN to 0.
inline int32_t saturating_i32(double v) {
  if (std::isnan(v)) return 0;
  if (v >= 2147483647.0) return std::numeric_limits<int32_t>::max();
  if (v <= -2147483648.0) return std::numeric_limits<int32_t>::min();
  return static_cast<int32_t>(v);
}

struct GridPoint {
  int32_t x = 0, y = 0, z = 0;
};

// spheres_searcher.rs:19-27
template <class R>
inline GridPoint grid_point_from_sphere(const Sphere<R>& s, double box) {
  return GridPoint{saturating_i32(std::floor(value_of(s.center.x) / box)),
                   saturating_i32(std::floor(value_of(s.center.y) / box)),
                   saturating_i32(std::floor(value_of(s.center.z) / box))};
}
// spheres_searcher.rs:29-36
template <class R>
inline GridPoint grid_point_with_offset(const Sphere<R>& s, double box, const GridPoint& off) {
  GridPoint g = grid_point_from_sphere(s, box);
  g.x -= off.x;
  g.y -= off.y;
  g.z -= off.z;
  return g;
}
// spheres_searcher.rs:38-51
inline bool grid_index(const GridPoint& p, const GridPoint& size, size_t* out) {
  if (p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x < size.x && p.y < size.y && p.z < size.z) {
    *out = static_cast<size_t>(p.z * size.x * size.y + p.y * size.x + p.x);
    return true;
  }
  return false;
}

struct GridParameters {
  GridPoint offset;
  GridPoint size{1, 1, 1};
  double box_size = 1.0;
};

// spheres_searcher.rs:62-101
template <class R>
inline GridParameters make_grid_parameters(const std::vector<Sphere<R>>& spheres) {
  GridParameters p;
  if (spheres.empty()) return p;
  for (const auto& s : spheres) p.box_size = std::fmax(p.box_size, std::fma(value_of(s.r), 2.0, 0.25));
  const int padding = 1;
  for (size_t i = 0; i < spheres.size(); ++i) {
    const GridPoint gp = grid_point_from_sphere(spheres[i], p.box_size);
    if (i == 0) {
      p.offset = gp;
      p.size = gp;
    } else {
      p.offset.x = std::min(p.offset.x, gp.x - padding);
      p.offset.y = std::min(p.offset.y, gp.y - padding);
      p.offset.z = std::min(p.offset.z, gp.z - padding);
      p.size.x = std::max(p.size.x, gp.x + padding);
      p.size.y = std::max(p.size.y, gp.y + padding);
      p.size.z = std::max(p.size.z, gp.z + padding);
    }
  }
  p.size.x = p.size.x - p.offset.x + 1;
  p.size.y = p.size.y - p.offset.y + 1;
  p.size.z = p.size.z - p.offset.z + 1;
  return p;
}

template <class R>
struct CollisionResult {
  std::vector<ValuedId<R>> colliding_ids;
  bool excluded = false;
};

// spheres_searcher.rs:116-327
template <class R>
class SpheresSearcher {
 public:
  SpheresSearcher() = default;
  explicit SpheresSearcher(std::vector<Sphere<R>> spheres) : spheres_(std::move(spheres)) {
    params_ = make_grid_parameters(spheres_);
    init_boxes();
  }
  const std::vector<Sphere<R>>& spheres() const { return spheres_; }

  // spheres_searcher.rs:144-175. The reference writes the new spheres
  // before remove_sphere_from_grid looks up their boxes, so the removal
  // searches the NEW box: a sphere that moved to another box stays listed in
  // its old one as well, and a search that reaches both boxes reports it
  // twice; a changed id given twice is also added twice. remove_from_old_box
  // = true removes the entries while the old spheres are still in place and
  // adds each sphere once (the evident intent), for the checks of callers
  // that list every colliding pair once.
  void update(const std::vector<Sphere<R>>& spheres, const std::vector<size_t>& changed_in,
              bool remove_from_old_box = false) {
    std::vector<size_t> changed = changed_in;
    if (remove_from_old_box) {
      std::sort(changed.begin(), changed.end());
      changed.erase(std::unique(changed.begin(), changed.end()), changed.end());
      for (size_t id : changed)
        if (id < spheres_.size()) remove_sphere_from_grid(id);
    }
    for (size_t id : changed)
      if (id < spheres_.size() && id < spheres.size()) spheres_[id] = spheres[id];
    const GridParameters np = make_grid_parameters(spheres_);
    if (std::fabs(np.box_size - params_.box_size) > 0.01 || np.size.x != params_.size.x ||
        np.size.y != params_.size.y || np.size.z != params_.size.z) {
      params_ = np;
      init_boxes();
    } else {
      if (!remove_from_old_box)
        for (size_t id : changed)
          if (id < spheres_.size()) remove_sphere_from_grid(id);
      for (size_t id : changed)
        if (id < spheres_.size()) add_sphere_to_grid(id);
    }
  }

  // spheres_searcher.rs:253-326
  CollisionResult<R> find_colliding_ids(size_t central_id, bool discard_hidden) const {
    CollisionResult<R> res;
    if (central_id >= spheres_.size()) return res;
    const Sphere<R>& cs = spheres_[central_id];
    const GridPoint gp = grid_point_with_offset(cs, params_.box_size, params_.offset);
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          const GridPoint nb{gp.x + dx, gp.y + dy, gp.z + dz};
          size_t index;
          if (!grid_index(nb, params_.size, &index)) continue;
          const int32_t box_id = map_of_boxes_[index];
          if (box_id < 0) continue;
          for (size_t id : boxes_[static_cast<size_t>(box_id)]) {
            if (id == central_id) continue;
            const Sphere<R>& cand = spheres_[id];
            if (!sphere_intersects_sphere(cs, cand)) continue;
            if (discard_hidden && sphere_contains_sphere(cand, cs) &&
                (!sphere_equals_sphere(cand, cs) || central_id > id)) {
              res.colliding_ids.clear();
              res.excluded = true;
              return res;
            }
            if (discard_hidden && sphere_contains_sphere(cs, cand)) continue;
            res.colliding_ids.push_back(ValuedId<R>{distance_to_intersection_circle_center(cs, cand), id});
          }
        }
    std::sort(res.colliding_ids.begin(), res.colliding_ids.end(), valued_id_less<R>);
    return res;
  }

 private:
  // spheres_searcher.rs:224-248
  void init_boxes() {
    const size_t total = static_cast<size_t>(params_.size.x) * static_cast<size_t>(params_.size.y) *
                         static_cast<size_t>(params_.size.z);
    map_of_boxes_.assign(total, -1);
    boxes_.clear();
    for (size_t i = 0; i < spheres_.size(); ++i) add_sphere_to_grid(i);
  }
  // spheres_searcher.rs:200-222
  void add_sphere_to_grid(size_t id) {
    const GridPoint gp = grid_point_with_offset(spheres_[id], params_.box_size, params_.offset);
    size_t index;
    if (!grid_index(gp, params_.size, &index)) return;
    const int32_t box_id = map_of_boxes_[index];
    if (box_id < 0) {
      map_of_boxes_[index] = static_cast<int32_t>(boxes_.size());
      boxes_.push_back(std::vector<size_t>{id});
    } else {
      boxes_[static_cast<size_t>(box_id)].push_back(id);
    }
  }
  // spheres_searcher.rs:178-194 (swap_remove)
  void remove_sphere_from_grid(size_t id) {
    const GridPoint gp = grid_point_with_offset(spheres_[id], params_.box_size, params_.offset);
    size_t index;
    if (!grid_index(gp, params_.size, &index)) return;
    const int32_t box_id = map_of_boxes_[index];
    if (box_id < 0) return;
    auto& v = boxes_[static_cast<size_t>(box_id)];
    for (size_t k = 0; k < v.size(); ++k)
      if (v[k] == id) {
        v[k] = v.back();
        v.pop_back();
        break;
      }
  }

  std::vector<Sphere<R>> spheres_;
  GridParameters params_;
  std::vector<int32_t> map_of_boxes_;
  std::vector<std::vector<size_t>> boxes_;
};

// ------------------------------------------------------------ contact.rs ----
enum ContourState : int { kInside = 0, kOutside = 1, kArcStart = 2, kArcContinue = 3 };

// contact.rs:36-59
template <class R>
struct ContourPoint {
  V3<R> p;
  R angle;
  size_t left_id, right_id;
  int state;
  ContourPoint() = default;
  ContourPoint(const V3<R>& pp, size_t l, size_t r) : p(pp), angle(R(0.0)), left_id(l), right_id(r), state(kInside) {}
};
template <class R>
using Contour = std::vector<ContourPoint<R>>;

template <class R>
struct ClippedContact {
  Contour<R> contour;
  Sphere<R> ic;
  V3<R> axis;
  V3<R> barycenter;
  R sum_of_arc_angles;
  R area;
};

// contact.rs:455-469
template <class R>
inline void init_contour(Contour<R>& contour, size_t a_id, const Sphere<R>& base, const V3<R>& axis) {
  contour.clear();
  const V3<R> first_point = (any_normal_of_vector(axis) * base.r) * R(HEXAGON_SCALE);
  const R angle_step = R(FRAC_PI_3);
  contour.emplace_back(base.center + first_point, a_id, a_id);
  R rot = angle_step;
  while (rot < R(TAU)) {
    const V3<R> rotated = rotate_point_around_axis(axis, rot, first_point);
    contour.emplace_back(base.center + rotated, a_id, a_id);
    rot = rot + angle_step;
  }
}

// contact.rs:472-481
template <class R>
inline bool test_contour_cuttable(const V3<R>& a_center, const V3<R>& closest_cut_point, const Contour<R>& contour) {
  const R threshold = norm_squared(closest_cut_point - a_center);
  for (const auto& cp : contour)
    if (norm_squared(cp.p - a_center) >= threshold) return true;
  return false;
}

// contact.rs:506-520
template <class R>
inline size_t mark_contour(Contour<R>& contour, const V3<R>& pc, const V3<R>& pn, size_t c_id) {
  size_t count = 0;
  for (auto& cp : contour)
    if (halfspace_of_point_unit(pc, pn, cp.p) >= 0) {
      cp.left_id = c_id;
      cp.right_id = c_id;
      ++count;
    }
  return count;
}

// contact.rs:524-627
template <class R>
inline void cut_contour(Contour<R>& contour, const V3<R>& pc, const V3<R>& pn, size_t c_id) {
  if (contour.size() < 3) return;
  auto is_out = [&](size_t i) { return contour[i].left_id == c_id && contour[i].right_id == c_id; };
  long first = -1, last = -1;
  for (size_t i = 0; i < contour.size(); ++i)
    if (is_out(i)) {
      if (first < 0) first = static_cast<long>(i);
      last = static_cast<long>(i);
    }
  if (first < 0) return;
  size_t start = static_cast<size_t>(first), end = static_cast<size_t>(last);
  if (start == 0 && end == contour.size() - 1) {
    end = 0;
    while (end + 1 < contour.size() && is_out(end + 1)) ++end;
    start = contour.size() - 1;
    while (start > 0 && is_out(start - 1)) --start;
  }
  if (start == end) {
    contour.insert(contour.begin() + static_cast<long>(start), contour[start]);
    end = start + 1;
  } else if (start < end) {
    if (start + 1 < end)
      contour.erase(contour.begin() + static_cast<long>(start + 1), contour.begin() + static_cast<long>(end));
    end = start + 1;
  } else {
    if (start + 1 < contour.size()) contour.erase(contour.begin() + static_cast<long>(start + 1), contour.end());
    if (end > 0) contour.erase(contour.begin(), contour.begin() + static_cast<long>(end));
    start = contour.size() - 1;
    end = 0;
  }
  const size_t n = contour.size();
  const size_t i_left = start > 0 ? start - 1 : n - 1;
  const V3<R> left_x = intersection_of_plane_and_segment(pc, pn, contour[start].p, contour[i_left].p);
  const ContourPoint<R> new_start(left_x, contour[i_left].right_id, contour[start].left_id);
  contour[start] = new_start;
  const size_t i_right = end + 1 < n ? end + 1 : 0;
  const V3<R> right_x = intersection_of_plane_and_segment(pc, pn, contour[end].p, contour[i_right].p);
  const ContourPoint<R> new_end(right_x, contour[end].right_id, contour[i_right].left_id);
  contour[end] = new_end;
  if (!fcmp::gt(norm_squared(contour[end].p - contour[start].p), R(0.0))) {
    contour[start].right_id = contour[end].right_id;
    contour.erase(contour.begin() + static_cast<long>(end));
  }
}

// contact.rs:484-502
template <class R>
inline bool mark_and_cut_contour(Contour<R>& contour, const V3<R>& pc, const V3<R>& pn, size_t c_id) {
  const size_t outsiders = mark_contour(contour, pc, pn, c_id);
  if (outsiders == 0) return false;
  if (outsiders >= contour.size()) {
    contour.clear();
    return true;
  }
  cut_contour(contour, pc, pn, c_id);
  return true;
}

// contact.rs:631-769
template <class R>
inline void restrict_contour_to_circle(Contour<R>& contour, const Sphere<R>& ic, const V3<R>& axis, size_t a_id,
                                       R* sum_angles) {
  *sum_angles = R(0.0);
  size_t outsiders = 0;
  for (auto& cp : contour) {
    if (norm_squared(cp.p - ic.center) <= ic.r * ic.r) {
      cp.state = kInside;
    } else {
      cp.state = kOutside;
      ++outsiders;
    }
  }
  if (outsiders == 0) return;
  size_t insertions = 0;
  size_t i = 0;
  while (i < contour.size()) {
    const size_t next_i = (i + 1) % contour.size();
    const int s1 = contour[i].state;
    const int s2 = contour[next_i].state;
    if (s1 == kOutside || s2 == kOutside) {
      if (s1 == kOutside && s2 == kOutside) {
        V3<R> mp, ip1, ip2;
        if (project_point_inside_line(ic.center, contour[i].p, contour[next_i].p, &mp) &&
            norm_squared(mp - ic.center) <= ic.r * ic.r) {
          const bool ok1 = intersect_segment_with_circle(ic, mp, contour[i].p, &ip1);
          const bool ok2 = intersect_segment_with_circle(ic, mp, contour[next_i].p, &ip2);
          if (ok1 && ok2) {
            const size_t right_id = contour[i].right_id;
            const size_t left_id = contour[next_i].left_id;
            const size_t pos = i + 1;
            contour.insert(contour.begin() + static_cast<long>(pos), ContourPoint<R>(ip1, a_id, right_id));
            contour.insert(contour.begin() + static_cast<long>(pos + 1), ContourPoint<R>(ip2, left_id, a_id));
            insertions += 2;
            i += 2;
          }
        }
      } else if (s1 == kOutside) {
        V3<R> ip;
        if (intersect_segment_with_circle(ic, contour[next_i].p, contour[i].p, &ip)) {
          const size_t right_id = contour[i].right_id;
          contour.insert(contour.begin() + static_cast<long>(i + 1), ContourPoint<R>(ip, a_id, right_id));
          insertions += 1;
          i += 1;
        } else {
          contour[next_i].left_id = a_id;
          contour[next_i].right_id = contour[i].right_id;
        }
      } else {
        V3<R> ip;
        if (intersect_segment_with_circle(ic, contour[i].p, contour[next_i].p, &ip)) {
          const size_t left_id = contour[next_i].left_id;
          contour.insert(contour.begin() + static_cast<long>(i + 1), ContourPoint<R>(ip, left_id, a_id));
          insertions += 1;
          i += 1;
        } else {
          contour[i].left_id = contour[next_i].left_id;
          contour[i].right_id = a_id;
        }
      }
    }
    i += 1;
  }
  if (insertions == 0) {
    contour.clear();
    return;
  }
  contour.erase(std::remove_if(contour.begin(), contour.end(), [](const ContourPoint<R>& cp) { return cp.state == kOutside; }),
                contour.end());
  if (contour.size() < 2) {
    contour.clear();
    return;
  }
  const V3<R> ctrl = ic.center + axis;
  for (size_t k = 0; k < contour.size(); ++k) {
    const size_t nk = (k + 1) % contour.size();
    if (contour[k].right_id == a_id && contour[nk].left_id == a_id) {
      const R angle = directed_angle(ic.center, contour[k].p, contour[nk].p, ctrl);
      contour[k].angle = angle;
      *sum_angles = *sum_angles + angle;
      contour[k].state = kArcStart;
    }
  }
  for (size_t k = 0; k < contour.size(); ++k) {
    const size_t pk = k > 0 ? k - 1 : contour.size() - 1;
    if (contour[pk].state == kArcStart) contour[k].state = kArcContinue;
  }
  if (fcmp::ge(*sum_angles, R(TAU)) || (contour.size() > 2 && fcmp::eq(*sum_angles, R(TAU)))) {
    *sum_angles = R(TAU);
    contour.clear();
  }
}

// contact.rs:773-793
template <class R>
inline R calculate_contour_area(const Contour<R>& contour, const Sphere<R>& ic, V3<R>* barycenter) {
  V3<R> sum{R(0.0), R(0.0), R(0.0)};
  for (const auto& cp : contour) sum = sum + cp.p;
  *barycenter = sum / R(static_cast<double>(contour.size()));
  R area = R(0.0);
  for (size_t i = 0; i < contour.size(); ++i) {
    const size_t ni = (i + 1) % contour.size();
    area = area + triangle_area(*barycenter, contour[i].p, contour[ni].p);
    if (contour[i].angle > R(0.0))
      area = area + ic.r * ic.r * (contour[i].angle - vsin(contour[i].angle)) * R(0.5);
  }
  return area;
}

// contact.rs:196-341
template <class R>
inline std::optional<ClippedContact<R>> clip_contact(const std::vector<Sphere<R>>& spheres, size_t a_id, size_t b_id,
                                                     const std::vector<ValuedId<R>>& neighbors) {
  if (a_id >= spheres.size() || b_id >= spheres.size()) return std::nullopt;
  const Sphere<R>& a = spheres[a_id];
  const Sphere<R>& b = spheres[b_id];
  if (!sphere_intersects_sphere(a, b) || sphere_contains_sphere(a, b) || sphere_contains_sphere(b, a))
    return std::nullopt;
  const Sphere<R> ic = intersection_circle_of_two_spheres(a, b);
  const V3<R> axis = normalize(b.center - a.center);
  if (ic.r <= R(0.0)) return std::nullopt;

  Contour<R> contour;
  bool discarded = false;
  bool initialized = false;
  for (const auto& nb : neighbors) {
    if (discarded) break;
    const size_t c_id = nb.index;
    if (c_id == b_id) continue;
    const Sphere<R>& c = spheres[c_id];
    if (!sphere_intersects_sphere(ic, c) || !sphere_intersects_sphere(b, c)) continue;
    if (sphere_contains_sphere(c, a) || sphere_contains_sphere(c, b)) {
      discarded = true;
      continue;
    }
    const V3<R> ac_center = center_of_intersection_circle(a, c);
    const V3<R> ac_normal = normalize(c.center - a.center);
    const R cos_val = dot(normalize(ic.center - a.center), normalize(ac_center - a.center));
    if (vabs(cos_val) >= R(1.0)) {
      if (halfspace_of_point_unit(ac_center, ac_normal, ic.center) > 0) discarded = true;
      continue;
    }
    const R l = vabs(signed_distance_to_plane_unit(ac_center, ac_normal, ic.center));
    const R xl = l / vsqrt(vfma(cos_val, -cos_val, R(1.0)));
    if (xl >= ic.r) {
      if (halfspace_of_point_unit(ac_center, ac_normal, ic.center) >= 0) discarded = true;
      continue;
    }
    if (!initialized) {
      init_contour(contour, a_id, ic, axis);
      initialized = true;
    } else if (!test_contour_cuttable(a.center, ac_center, contour)) {
      continue;
    }
    mark_and_cut_contour(contour, ac_center, ac_normal, c_id);
    if (contour.empty()) discarded = true;
  }
  if (discarded) return std::nullopt;

  ClippedContact<R> out;
  out.barycenter = V3<R>{R(0.0), R(0.0), R(0.0)};
  out.sum_of_arc_angles = R(0.0);
  out.area = R(0.0);
  if (!initialized) {
    out.barycenter = ic.center;
    out.sum_of_arc_angles = R(TAU);
    out.area = ic.r * ic.r * R(PI);
  } else if (!contour.empty()) {
    restrict_contour_to_circle(contour, ic, axis, a_id, &out.sum_of_arc_angles);
    if (!contour.empty()) out.area = calculate_contour_area(contour, ic, &out.barycenter);
  }
  if (out.area <= R(0.0)) return std::nullopt;
  out.contour = std::move(contour);
  out.ic = ic;
  out.axis = axis;
  return out;
}

// contact.rs:411-444
template <class R>
inline bool contour_is_central(const V3<R>& center, const Contour<R>& contour, const V3<R>& barycenter) {
  const double FE = 1e-10;
  for (const auto& cp : contour)
    if ((cp.angle - R(FE)) > R(PI)) return true;
  const size_t n = contour.size();
  for (size_t i = 0; i < n; ++i) {
    const V3<R> pi = contour[i].p;
    const V3<R> pj = contour[(i + 1) % n].p;
    const V3<R> edge = pj - pi;
    const R len = norm(edge);
    if (len < R(FE)) continue;
    const V3<R> u = edge / len;
    const V3<R> foot = pi + u * dot(u, barycenter - pi);
    const V3<R> n_ijb = barycenter - foot;
    if (dot(n_ijb, center - pi) < R(0.0)) return false;
  }
  return true;
}

// contact.rs:797-848
template <class R>
inline R calculate_solid_angle(const Sphere<R>& a, const Sphere<R>& b, const Sphere<R>& ic, const Contour<R>& contour) {
  R turn = R(0.0);
  if (contour.empty()) {
    turn = R(TAU) * norm(ic.center - a.center) / a.r;
  } else {
    const size_t n = contour.size();
    for (size_t i = 0; i < n; ++i) {
      const size_t pi_ = i > 0 ? i - 1 : n - 1;
      const size_t ni = (i + 1) % n;
      const ContourPoint<R>& pr0 = contour[pi_];
      const ContourPoint<R>& pr1 = contour[i];
      const ContourPoint<R>& pr2 = contour[ni];
      if (pr0.angle > R(0.0)) {
        V3<R> d = cross(b.center - a.center, pr1.p - ic.center);
        const R t = dot(d, pr0.p - pr1.p);
        const bool flip = (pr0.angle < R(PI) && t < R(0.0)) || (pr0.angle > R(PI) && t > R(0.0));
        if (flip) d = -d;
        turn = turn + (R(PI) - min_dihedral_angle(a.center, pr1.p, pr1.p + d, pr2.p));
      } else if (pr1.angle > R(0.0)) {
        V3<R> d = cross(b.center - a.center, pr1.p - ic.center);
        const R t = dot(d, pr2.p - pr1.p);
        const bool flip = (pr1.angle < R(PI) && t < R(0.0)) || (pr1.angle > R(PI) && t > R(0.0));
        if (flip) d = -d;
        turn = turn + (R(PI) - min_dihedral_angle(a.center, pr1.p, pr0.p, pr1.p + d));
        turn = turn + pr1.angle * (norm(ic.center - a.center) / a.r);
      } else {
        turn = turn + (R(PI) - min_dihedral_angle(a.center, pr1.p, pr0.p, pr2.p));
      }
    }
  }
  R solid = R(TAU) - turn;
  const V3<R> ic_to_a = ic.center - a.center;
  const V3<R> ic_to_b = ic.center - b.center;
  if (dot(ic_to_a, ic_to_b) > R(0.0) && norm_squared(a.center - ic.center) < norm_squared(b.center - ic.center))
    solid = -solid;
  return solid;
}

// types.rs:380-415 ContactDescriptorSummary
template <class R>
struct Summary {
  R area, arc_length, solid_angle_a, solid_angle_b, pyramid_volume_a, pyramid_volume_b, distance;
  size_t id_a, id_b;
  bool central;
  // types.rs:408-414
  void ensure_ids_ordered() {
    if (id_a > id_b) {
      std::swap(id_a, id_b);
      std::swap(solid_angle_a, solid_angle_b);
      std::swap(pyramid_volume_a, pyramid_volume_b);
    }
  }
};

struct TessVertex {
  size_t ids[4];
  uint64_t pos[3];
};
struct TessEdge {
  size_t ids[3];
  double length;
};

// types.rs:461-489 (4-element sorting network)
inline void sort_ids4(size_t* v) {
  if (v[0] > v[1]) std::swap(v[0], v[1]);
  if (v[2] > v[3]) std::swap(v[2], v[3]);
  if (v[0] > v[2]) std::swap(v[0], v[2]);
  if (v[1] > v[3]) std::swap(v[1], v[3]);
  if (v[1] > v[2]) std::swap(v[1], v[2]);
}
// types.rs:541-552 (3-element sorting network)
inline void sort_ids3(size_t* v) {
  if (v[0] > v[1]) std::swap(v[0], v[1]);
  if (v[1] > v[2]) std::swap(v[1], v[2]);
  if (v[0] > v[1]) std::swap(v[0], v[1]);
}
inline uint64_t dbits(double x) {
  uint64_t u;
  std::memcpy(&u, &x, 8);
  return u;
}

// contact.rs:66-114, 347-382: the full descriptor plus its summary.
template <class R>
struct Descriptor {
  ClippedContact<R> cc;
  R solid_angle_a, solid_angle_b, pyramid_volume_a, pyramid_volume_b, distance;
  size_t id_a, id_b;

  Summary<R> to_summary() const {
    Summary<R> s;
    s.area = cc.area;
    s.arc_length = cc.sum_of_arc_angles * cc.ic.r;
    s.solid_angle_a = solid_angle_a;
    s.solid_angle_b = solid_angle_b;
    s.pyramid_volume_a = pyramid_volume_a;
    s.pyramid_volume_b = pyramid_volume_b;
    s.distance = distance;
    s.id_a = id_a;
    s.id_b = id_b;
    s.central = contour_is_central(cc.ic.center, cc.contour, cc.barycenter);
    return s;
  }

  // contact.rs:120-177
  void to_tessellation_elements(std::vector<TessEdge>* edges, std::vector<TessVertex>* verts) const {
    if (cc.area <= R(0.0)) return;
    if (cc.contour.empty()) {
      TessEdge e{{id_a, id_b, NULL_ID}, value_of(cc.sum_of_arc_angles * cc.ic.r)};
      sort_ids3(e.ids);
      edges->push_back(e);
      return;
    }
    const size_t n = cc.contour.size();
    for (size_t i = 0; i < n; ++i) {
      const auto& cp = cc.contour[i];
      const auto& nx = cc.contour[(i + 1) % n];
      const R length = cp.right_id == id_a ? cp.angle * cc.ic.r : norm(cp.p - nx.p);
      const size_t edge_third = cp.right_id == id_a ? NULL_ID : cp.right_id;
      TessEdge e{{id_a, id_b, edge_third}, value_of(length)};
      sort_ids3(e.ids);
      edges->push_back(e);
      const size_t v3 = cp.left_id == id_a ? NULL_ID : cp.left_id;
      const size_t v4 = cp.right_id == id_a ? NULL_ID : cp.right_id;
      TessVertex v{{id_a, id_b, v3, v4}, {dbits(value_of(cp.p.x)), dbits(value_of(cp.p.y)), dbits(value_of(cp.p.z))}};
      sort_ids4(v.ids);
      verts->push_back(v);
    }
  }
};

// contact.rs:347-382
template <class R>
inline std::optional<Descriptor<R>> construct_contact_descriptor(const std::vector<Sphere<R>>& spheres, size_t a_id,
                                                                 size_t b_id, const std::vector<ValuedId<R>>& neighbors) {
  auto cc = clip_contact(spheres, a_id, b_id, neighbors);
  if (!cc) return std::nullopt;
  const Sphere<R>& a = spheres[a_id];
  const Sphere<R>& b = spheres[b_id];
  Descriptor<R> d;
  d.solid_angle_a = calculate_solid_angle(a, b, cc->ic, cc->contour);
  d.solid_angle_b = calculate_solid_angle(b, a, cc->ic, cc->contour);
  const R sign_a = d.solid_angle_a < R(0.0) ? R(-1.0) : R(1.0);
  const R sign_b = d.solid_angle_b < R(0.0) ? R(-1.0) : R(1.0);
  d.pyramid_volume_a = norm(cc->ic.center - a.center) * cc->area / R(3.0) * sign_a;
  d.pyramid_volume_b = norm(cc->ic.center - b.center) * cc->area / R(3.0) * sign_b;
  d.distance = norm(b.center - a.center);
  d.id_a = a_id;
  d.id_b = b_id;
  d.cc = std::move(*cc);
  return d;
}

// contact.rs:387-404 (area, arc_length, central)
template <class R>
struct AreaOnly {
  R area, arc_length;
  bool central;
};
template <class R>
inline std::optional<AreaOnly<R>> construct_contact_area(const std::vector<Sphere<R>>& spheres, size_t a_id, size_t b_id,
                                                         const std::vector<ValuedId<R>>& neighbors) {
  auto cc = clip_contact(spheres, a_id, b_id, neighbors);
  if (!cc) return std::nullopt;
  AreaOnly<R> o;
  o.central = contour_is_central(cc->ic.center, cc->contour, cc->barycenter);
  o.area = cc->area;
  o.arc_length = cc->sum_of_arc_angles * cc->ic.r;
  return o;
}

// -------------------------------------------------------------- cells ----
enum CellStage : int { kInit = 0, kContactsAdded = 1, kSasComputed = 2 };

// types.rs:431-454, 587-676
template <class R>
struct CellSummary {
  size_t id = 0;
  R area = R(0.0), arc_length = R(0.0);
  R sa_pos = R(0.0), sa_neg = R(0.0), pv_pos = R(0.0), pv_neg = R(0.0);
  R sas_area = R(0.0), sas_inside_volume = R(0.0);
  size_t count = 0;
  int stage = kInit;

  // types.rs:589-610
  void add(const Summary<R>& cds) {
    if (cds.area > R(0.0) && (cds.id_a == id || cds.id_b == id)) {
      count += 1;
      area = area + cds.area;
      arc_length = arc_length + cds.arc_length;
      const R sa = cds.id_a == id ? cds.solid_angle_a : cds.solid_angle_b;
      const R pv = cds.id_a == id ? cds.pyramid_volume_a : cds.pyramid_volume_b;
      sa_pos = sa_pos + vmax(sa, R(0.0));
      sa_neg = sa_neg - vmin(sa, R(0.0));
      pv_pos = pv_pos + vmax(pv, R(0.0));
      pv_neg = pv_neg - vmin(pv, R(0.0));
      stage = kContactsAdded;
    }
  }
  // types.rs:624-663
  void compute_sas(const R& r) {
    if (stage != kContactsAdded) return;
    sas_area = R(0.0);
    sas_inside_volume = R(0.0);
    const R diff = vabs(sa_pos - sa_neg);
    if (arc_length > R(0.0) && diff > R(1e-10)) {
      if (sa_pos > sa_neg) {
        const R angle_diff = sa_pos - sa_neg;
        sas_area = vfma(R(4.0), R(PI), -vmax(angle_diff, R(0.0))) * r * r;
      } else {
        const R angle_diff = sa_neg - sa_pos;
        sas_area = vmax(angle_diff, R(0.0)) * r * r;
      }
      sas_inside_volume = (sas_area * r / R(3.0)) + pv_pos - pv_neg;
      const R full = R(4.0) / R(3.0) * R(PI) * r * r * r;
      if (sas_inside_volume > full) {
        sas_area = R(0.0);
        sas_inside_volume = pv_pos - pv_neg;
      }
    } else {
      sas_inside_volume = pv_pos - pv_neg;
    }
    stage = kSasComputed;
  }
  // types.rs:666-675
  void compute_sas_detached(size_t new_id, const R& r) {
    if (stage == kInit) {
      id = new_id;
      sas_area = R(4.0) * R(PI) * r * r;
      sas_inside_volume = sas_area * r / R(3.0);
      stage = kSasComputed;
    }
  }
};

// --------------------------------------------------------- outputs ----
template <class R>
struct Contact {
  size_t id_a, id_b;
  R area, arc_length;
  bool central;
};
template <class R>
struct Cell {
  size_t index;
  R sas_area, volume;
};
struct CellVertexOut {
  size_t ids[4];  // NULL_ID = None
  double x, y, z;
};
struct CellEdgeOut {
  size_t ids[3];
  double length;
};
template <class R>
struct TessellationResult {
  size_t num_balls = 0;
  std::vector<Contact<R>> contacts;
  std::vector<Cell<R>> cells;
  bool has_network = false;
  std::vector<CellVertexOut> vertices;
  std::vector<CellEdgeOut> edges;
};

// ------------------------------------------------------- tessellation.rs ----
// tessellation.rs:323-328
inline bool same_group(const int32_t* groups, size_t n_groups, size_t a, size_t b) {
  if (groups != nullptr && a < n_groups && b < n_groups) return groups[a] == groups[b];
  return false;
}

// tessellation.rs:298-319
template <class R>
inline std::vector<std::pair<size_t, size_t>> collect_collision_pairs(const std::vector<std::vector<ValuedId<R>>>& all,
                                                                     const int32_t* groups, size_t n_groups,
                                                                     size_t periodic_n /*0 = none*/) {
  std::vector<std::pair<size_t, size_t>> pairs;
  for (size_t a = 0; a < all.size(); ++a)
    for (const auto& nb : all[a]) {
      const size_t b = nb.index;
      const size_t bc = periodic_n ? b % periodic_n : b;
      if (same_group(groups, n_groups, a, bc)) continue;
      if ((periodic_n && b >= periodic_n) || a < b) pairs.emplace_back(a, b);
    }
  return pairs;
}

// Parallel-for helper (OpenMP when enabled; sequential otherwise).
template <class F>
inline void parallel_for(size_t n, int threads, F&& f) {
#ifdef _OPENMP
  if (threads != 1) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : omp_get_max_threads())
    for (long i = 0; i < static_cast<long>(n); ++i) f(static_cast<size_t>(i));
    return;
  }
#endif
  (void)threads;
  for (size_t i = 0; i < n; ++i) f(i);
}

// tessellation.rs:51-111 CollisionContext
template <class R>
struct CollisionContext {
  SpheresSearcher<R> searcher;
  std::vector<std::vector<ValuedId<R>>> all_collisions;
  std::vector<std::pair<size_t, size_t>> pairs;
  std::vector<Sphere<R>> input_spheres;
  size_t n = 0;
  bool periodic = false;
};

template <class R>
inline CollisionContext<R> make_collision_context(const std::vector<Ball<R>>& balls, R probe, const PeriodicBox<R>* pbox,
                                                  const int32_t* groups, size_t n_groups, int threads) {
  CollisionContext<R> ctx;
  for (const auto& b : balls) ctx.input_spheres.push_back(sphere_from_ball(b, probe));
  ctx.n = ctx.input_spheres.size();
  ctx.periodic = pbox != nullptr;
  if (pbox)
    ctx.searcher = SpheresSearcher<R>(populate_periodic_spheres(*pbox, ctx.input_spheres));
  else
    ctx.searcher = SpheresSearcher<R>(ctx.input_spheres);
  ctx.all_collisions.assign(ctx.n, {});
  parallel_for(ctx.n, threads, [&](size_t id) { ctx.all_collisions[id] = ctx.searcher.find_colliding_ids(id, true).colliding_ids; });
  ctx.pairs = collect_collision_pairs(ctx.all_collisions, groups, n_groups, pbox ? ctx.n : 0);
  return ctx;
}

// tessellation.rs:607-658
template <class R>
inline std::vector<Cell<R>> compute_cells(const std::vector<Summary<R>>& summaries, const std::vector<Sphere<R>>& spheres,
                                          const std::vector<std::vector<ValuedId<R>>>& all_collisions, size_t periodic_n) {
  const size_t n = periodic_n ? periodic_n : spheres.size();
  std::vector<CellSummary<R>> cs(n);
  for (size_t i = 0; i < n; ++i) cs[i].id = i;
  for (const auto& s : summaries) {
    if (!(s.area > R(0.0))) continue;
    if (periodic_n) {
      if (s.id_a < n) cs[s.id_a].add(s);
      if (s.id_b < n && s.id_b != s.id_a) cs[s.id_b].add(s);
    } else {
      cs[s.id_a].add(s);
      cs[s.id_b].add(s);
    }
  }
  for (size_t i = 0; i < n; ++i) {
    if (cs[i].stage == kContactsAdded)
      cs[i].compute_sas(spheres[i].r);
    else if (cs[i].stage == kInit && all_collisions[i].empty())
      cs[i].compute_sas_detached(i, spheres[i].r);
  }
  std::vector<Cell<R>> out;
  for (const auto& c : cs)
    if (c.stage == kSasComputed) out.push_back(Cell<R>{c.id, c.sas_area, c.sas_inside_volume});
  return out;
}

// tessellation.rs:556-601
template <class R>
inline std::vector<size_t> compute_canonical_ids(const std::vector<Summary<R>>& s, size_t n) {
  std::vector<std::vector<size_t>> boundary(n);
  for (size_t i = 0; i < s.size(); ++i)
    if (s[i].id_a >= n || s[i].id_b >= n) {
      boundary[s[i].id_a % n].push_back(i);
      boundary[s[i].id_b % n].push_back(i);
    }
  std::vector<size_t> canon(s.size());
  for (size_t i = 0; i < s.size(); ++i) canon[i] = i;
  for (size_t i = 0; i < s.size(); ++i) {
    if (!(s[i].id_a >= n || s[i].id_b >= n)) continue;
    const size_t sa = s[i].id_a % n, sb = s[i].id_b % n;
    const auto& cand = boundary[sa].size() <= boundary[sb].size() ? boundary[sa] : boundary[sb];
    for (size_t ci : cand) {
      const size_t ca = s[ci].id_a % n, cb = s[ci].id_b % n;
      if ((ca == sa && cb == sb) || (ca == sb && cb == sa)) {
        canon[i] = ci;
        break;
      }
    }
  }
  return canon;
}

// tessellation.rs:527-553 (kept summaries in order, ids not canonicalized)
template <class R>
inline std::vector<Summary<R>> deduplicate_periodic_contacts(const std::vector<Summary<R>>& s, size_t n) {
  const auto canon = compute_canonical_ids(s, n);
  std::vector<Summary<R>> out;
  for (size_t i = 0; i < s.size(); ++i)
    if (canon[i] == i) out.push_back(s[i]);
  return out;
}

// tessellation.rs:334-401
inline void assemble_network(std::vector<TessEdge> edges, std::vector<TessVertex> verts, size_t periodic_n,
                             std::vector<CellVertexOut>* vout, std::vector<CellEdgeOut>* eout) {
  auto canon = [&](size_t id) { return (periodic_n && id != NULL_ID) ? id % periodic_n : id; };
  for (auto& e : edges) {
    for (auto& id : e.ids) id = canon(id);
    sort_ids3(e.ids);
  }
  for (auto& v : verts) {
    for (auto& id : v.ids) id = canon(id);
    sort_ids4(v.ids);
  }
  // The reference uses sort_unstable then keeps the first element of each run
  // of equal ids; which duplicate survives is unspecified there. A stable sort
  // keeps the first in contact order.
  std::stable_sort(edges.begin(), edges.end(), [](const TessEdge& a, const TessEdge& b) {
    return std::lexicographical_compare(a.ids, a.ids + 3, b.ids, b.ids + 3);
  });
  std::stable_sort(verts.begin(), verts.end(), [](const TessVertex& a, const TessVertex& b) {
    return std::lexicographical_compare(a.ids, a.ids + 4, b.ids, b.ids + 4);
  });
  for (size_t i = 0; i < edges.size(); ++i)
    if (i == 0 || !std::equal(edges[i].ids, edges[i].ids + 3, edges[i - 1].ids))
      eout->push_back(CellEdgeOut{{edges[i].ids[0], edges[i].ids[1], edges[i].ids[2]}, edges[i].length});
  for (size_t i = 0; i < verts.size(); ++i)
    if (i == 0 || !std::equal(verts[i].ids, verts[i].ids + 4, verts[i - 1].ids)) {
      double p[3];
      for (int k = 0; k < 3; ++k) std::memcpy(&p[k], &verts[i].pos[k], 8);
      vout->push_back(CellVertexOut{{verts[i].ids[0], verts[i].ids[1], verts[i].ids[2], verts[i].ids[3]}, p[0], p[1], p[2]});
    }
}

// Per-pair record kept by the full drivers.
template <class R>
struct PairResult {
  bool valid = false;
  Summary<R> summary;
  std::vector<TessEdge> edges;
  std::vector<TessVertex> verts;
};

// tessellation.rs:37-48, 199-292 (standard) and 406-520 (periodic)
template <class R>
inline TessellationResult<R> compute_tessellation(const std::vector<Ball<R>>& balls, R probe, const PeriodicBox<R>* pbox,
                                                  const int32_t* groups, size_t n_groups, bool with_vertices, int threads,
                                                  std::vector<Summary<R>>* all_summaries_out = nullptr) {
  TessellationResult<R> res;
  if (balls.empty()) return res;
  CollisionContext<R> ctx = make_collision_context(balls, probe, pbox, groups, n_groups, threads);
  const size_t n = ctx.n;
  std::vector<PairResult<R>> pr(ctx.pairs.size());
  parallel_for(ctx.pairs.size(), threads, [&](size_t k) {
    const size_t a = ctx.pairs[k].first, b = ctx.pairs[k].second;
    auto cd = construct_contact_descriptor(ctx.searcher.spheres(), a, b, ctx.all_collisions[a]);
    if (!cd) return;
    Summary<R> s = cd->to_summary();
    if (pbox) {
      s.id_a = a;
      s.id_b = b;
    }
    s.ensure_ids_ordered();
    pr[k].valid = true;
    pr[k].summary = s;
    if (with_vertices) cd->to_tessellation_elements(&pr[k].edges, &pr[k].verts);
  });
  std::vector<Summary<R>> valid;
  std::vector<TessEdge> all_edges;
  std::vector<TessVertex> all_verts;
  for (auto& p : pr)
    if (p.valid && p.summary.area > R(0.0)) {
      valid.push_back(p.summary);
      if (with_vertices) {
        all_edges.insert(all_edges.end(), p.edges.begin(), p.edges.end());
        all_verts.insert(all_verts.end(), p.verts.begin(), p.verts.end());
      }
    }
  res.num_balls = n;
  res.cells = compute_cells(valid, ctx.input_spheres, ctx.all_collisions, pbox ? n : 0);
  if (with_vertices) {
    res.has_network = true;
    assemble_network(std::move(all_edges), std::move(all_verts), pbox ? n : 0, &res.vertices, &res.edges);
  }
  if (pbox) {
    const auto kept = deduplicate_periodic_contacts(valid, n);
    for (const auto& s : kept) res.contacts.push_back(Contact<R>{s.id_a % n, s.id_b % n, s.area, s.arc_length, s.central});
  } else {
    for (const auto& s : valid) res.contacts.push_back(Contact<R>{s.id_a, s.id_b, s.area, s.arc_length, s.central});
  }
  if (all_summaries_out) *all_summaries_out = std::move(valid);
  return res;
}

// tessellation.rs:125-188. Periodic output: first occurrence per canonical
// (min,max) pair in pair order; the reference returns them in HashMap order.
template <class R>
inline std::vector<Contact<R>> compute_contacts_only(const std::vector<Ball<R>>& balls, R probe, const PeriodicBox<R>* pbox,
                                                     const int32_t* groups, size_t n_groups, int threads) {
  std::vector<Contact<R>> out;
  if (balls.empty()) return out;
  CollisionContext<R> ctx = make_collision_context(balls, probe, pbox, groups, n_groups, threads);
  std::vector<std::optional<AreaOnly<R>>> pr(ctx.pairs.size());
  parallel_for(ctx.pairs.size(), threads, [&](size_t k) {
    const size_t a = ctx.pairs[k].first, b = ctx.pairs[k].second;
    pr[k] = construct_contact_area(ctx.searcher.spheres(), a, b, ctx.all_collisions[a]);
  });
  for (size_t k = 0; k < pr.size(); ++k)
    if (pr[k]) out.push_back(Contact<R>{ctx.pairs[k].first, ctx.pairs[k].second, pr[k]->area, pr[k]->arc_length, pr[k]->central});
  if (!pbox) return out;
  const size_t n = ctx.n;
  std::vector<Contact<R>> dedup;
  std::vector<std::pair<size_t, size_t>> seen;
  for (const auto& c : out) {
    size_t ca = c.id_a % n, cb = c.id_b % n;
    if (ca > cb) std::swap(ca, cb);
    seen.emplace_back(ca, cb);
  }
  std::vector<size_t> order(out.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return seen[x] < seen[y]; });
  std::vector<char> keep(out.size(), 0);
  for (size_t i = 0; i < order.size(); ++i)
    if (i == 0 || seen[order[i]] != seen[order[i - 1]]) keep[order[i]] = 1;
  for (size_t i = 0; i < out.size(); ++i)
    if (keep[i]) {
      Contact<R> c = out[i];
      c.id_a = seen[i].first;
      c.id_b = seen[i].second;
      dedup.push_back(c);
    }
  return dedup;
}

// ---------------------------------------------- spheres_container.rs ----
template <class R>
struct UpdateResult {
  std::vector<size_t> changed_ids, affected_ids;
  bool was_full_reinit = false;
};

// spheres_container.rs:28-405
template <class R>
class SpheresContainer {
 public:
  // spheres_container.rs:55-64
  void init(std::vector<Sphere<R>> spheres, const PeriodicBox<R>* pbox, int threads) {
    spheres_ = std::move(spheres);
    has_pbox_ = pbox != nullptr;
    if (pbox) pbox_ = *pbox;
    populate_spheres();
    exclusion_.assign(populated_.size(), false);
    searcher_ = SpheresSearcher<R>(populated_);
    has_searcher_ = true;
    detect_all_collisions(threads);
  }

  // spheres_container.rs:71-182
  std::optional<UpdateResult<R>> update(const std::vector<Sphere<R>>& ns, const std::vector<size_t>* changed_ids, int threads) {
    if (ns.size() != spheres_.size()) {
      init(ns, has_pbox_ ? &pbox_ : nullptr, threads);
      UpdateResult<R> r;
      r.was_full_reinit = true;
      return r;
    }
    const size_t threshold = size_threshold();
    std::vector<size_t> changed;
    if (changed_ids) {
      for (size_t id : *changed_ids)
        if (id < spheres_.size() && !sphere_equals_sphere(ns[id], spheres_[id])) changed.push_back(id);
    } else {
      for (size_t i = 0; i < ns.size() && changed.size() < threshold + 1; ++i)
        if (!sphere_equals_sphere(ns[i], spheres_[i])) changed.push_back(i);
    }
    if (changed.empty()) return std::nullopt;
    auto full_reinit = [&]() {
      init(ns, has_pbox_ ? &pbox_ : nullptr, threads);
      UpdateResult<R> r;
      r.was_full_reinit = true;
      return r;
    };
    if (changed.size() > threshold) return full_reinit();
    std::vector<size_t> affected = changed;
    std::sort(affected.begin(), affected.end());
    const size_t n = spheres_.size();
    for (size_t sid : changed)
      for (const auto& nb : colliding_[sid]) {
        const size_t cid = nb.index % n;
        auto it = std::lower_bound(affected.begin(), affected.end(), cid);
        if (it == affected.end() || *it != cid) {
          if (affected.size() < threshold)
            affected.insert(it, cid);
          else
            return full_reinit();
        }
      }
    std::vector<size_t> changed_populated;
    for (size_t sid : changed) {
      spheres_[sid] = ns[sid];
      update_sphere_periodic_instances(sid, &changed_populated);
    }
    if (has_searcher_) searcher_.update(populated_, changed_populated, remove_from_old_box_);
    update_collisions_for_spheres(affected, threads);
    std::vector<size_t> more;
    for (size_t sid : changed)
      for (const auto& nb : colliding_[sid]) {
        const size_t cid = nb.index % n;
        if (!std::binary_search(affected.begin(), affected.end(), cid)) {
          auto it = std::lower_bound(more.begin(), more.end(), cid);
          if (it == more.end() || *it != cid) more.insert(it, cid);
        }
      }
    if (!more.empty()) {
      update_collisions_for_spheres(more, threads);
      affected.insert(affected.end(), more.begin(), more.end());
      std::sort(affected.begin(), affected.end());
      affected.erase(std::unique(affected.begin(), affected.end()), affected.end());
    }
    UpdateResult<R> r;
    r.changed_ids = changed;
    r.affected_ids = affected;
    r.was_full_reinit = false;
    return r;
  }

  // spheres_container.rs:186-208
  std::optional<std::vector<size_t>> set_exclusion(size_t id, bool excluded) {
    if (id >= spheres_.size() || id >= exclusion_.size()) return std::nullopt;
    if (exclusion_[id] == excluded) return std::nullopt;
    exclusion_[id] = excluded;
    set_exclusion_status_periodic_instances(id);
    std::vector<size_t> affected{id};
    const size_t n = spheres_.size();
    for (const auto& nb : colliding_[id]) {
      const size_t cid = nb.index % n;
      auto it = std::lower_bound(affected.begin(), affected.end(), cid);
      if (it == affected.end() || *it != cid) affected.insert(it, cid);
    }
    return affected;
  }

  // spheres_container.rs:246-286
  void restore_from(const SpheresContainer& backup, const std::vector<size_t>& affected) {
    if (affected.empty() || spheres_.size() != backup.spheres_.size() || affected.size() > size_threshold()) {
      *this = backup;
      return;
    }
    const size_t n = spheres_.size();
    for (size_t id : affected)
      if (id >= n) {
        *this = backup;
        return;
      }
    for (size_t id : affected) {
      spheres_[id] = backup.spheres_[id];
      colliding_[id] = backup.colliding_[id];
      exclusion_[id] = backup.exclusion_[id];
      if (has_pbox_)
        for (size_t m = 1; m < 27; ++m) {
          const size_t sid = m * n + id;
          if (sid < populated_.size()) {
            populated_[sid] = backup.populated_[sid];
            exclusion_[sid] = backup.exclusion_[sid];
          }
        }
      populated_[id] = backup.populated_[id];
    }
    if (backup.has_searcher_) {
      searcher_ = backup.searcher_;
      has_searcher_ = true;
    }
  }

  const std::vector<Sphere<R>>& spheres() const { return spheres_; }
  const std::vector<Sphere<R>>& populated_spheres() const { return populated_; }
  const std::vector<std::vector<ValuedId<R>>>& all_colliding_ids() const { return colliding_; }
  const std::vector<ValuedId<R>>& colliding_ids(size_t id) const { return colliding_[id]; }
  const std::vector<bool>& exclusion_statuses() const { return exclusion_; }
  bool is_excluded(size_t id) const { return id < exclusion_.size() && exclusion_[id]; }
  void set_remove_from_old_box(bool on) { remove_from_old_box_ = on; }
  bool periodic() const { return has_pbox_; }

 private:
  size_t size_threshold() const { return std::max<size_t>(spheres_.size() / 2, 10); }
  void populate_spheres() {
    if (has_pbox_)
      populated_ = populate_periodic_spheres(pbox_, spheres_);
    else
      populated_ = spheres_;
  }
  // spheres_container.rs:303-326
  void update_sphere_periodic_instances(size_t id, std::vector<size_t>* changed) {
    if (id >= spheres_.size()) return;
    const size_t n = spheres_.size();
    populated_[id] = spheres_[id];
    changed->push_back(id);
    if (has_pbox_)
      for (int g = 0; g < 26; ++g) {
        int sx, sy, sz;
        neighbor_shift(g, &sx, &sy, &sz);
        const size_t sid = static_cast<size_t>(g + 1) * n + id;
        populated_[sid] = shift_sphere(pbox_, spheres_[id], double(sx), double(sy), double(sz));
        changed->push_back(sid);
      }
  }
  // spheres_container.rs:329-341
  void set_exclusion_status_periodic_instances(size_t id) {
    if (!has_pbox_) return;
    const size_t n = spheres_.size();
    if (exclusion_.size() != n * 27) return;
    const bool st = exclusion_[id];
    for (size_t m = 1; m < 27; ++m) exclusion_[m * n + id] = st;
  }
  // spheres_container.rs:344-370
  void detect_all_collisions(int threads) {
    const size_t n = spheres_.size();
    colliding_.assign(n, {});
    if (!has_searcher_) return;
    std::vector<char> excl(n, 0);
    parallel_for(n, threads, [&](size_t id) {
      auto r = searcher_.find_colliding_ids(id, true);
      colliding_[id] = std::move(r.colliding_ids);
      excl[id] = r.excluded ? 1 : 0;
    });
    for (size_t i = 0; i < n; ++i) exclusion_[i] = excl[i] != 0;
    if (has_pbox_)
      for (size_t i = 0; i < n; ++i) set_exclusion_status_periodic_instances(i);
  }
  // spheres_container.rs:373-394
  void update_collisions_for_spheres(const std::vector<size_t>& ids, int threads) {
    if (!has_searcher_) return;
    std::vector<CollisionResult<R>> rs(ids.size());
    parallel_for(ids.size(), threads, [&](size_t k) { rs[k] = searcher_.find_colliding_ids(ids[k], true); });
    for (size_t k = 0; k < ids.size(); ++k) {
      colliding_[ids[k]] = std::move(rs[k].colliding_ids);
      exclusion_[ids[k]] = rs[k].excluded;
    }
    if (has_pbox_)
      for (size_t id : ids) set_exclusion_status_periodic_instances(id);
  }

  std::vector<Sphere<R>> spheres_;
  bool has_pbox_ = false;
  PeriodicBox<R> pbox_{};
  std::vector<Sphere<R>> populated_;
  std::vector<bool> exclusion_;
  // SpheresSearcher::update's remove_from_old_box for this container's updates
  bool remove_from_old_box_ = false;
  std::vector<std::vector<ValuedId<R>>> colliding_;
  SpheresSearcher<R> searcher_;
  bool has_searcher_ = false;
};

// -------------------------------------------------------- updateable.rs ----
template <class R>
struct UpdState {
  SpheresContainer<R> container;
  size_t num_balls = 0;
  std::vector<Cell<R>> cells;
  std::vector<std::vector<Summary<R>>> contacts_by_sphere;
  std::vector<CellSummary<R>> cell_summaries;
  std::vector<size_t> changed_ids, affected_ids;
  bool was_full_reinit = true;
  bool result_empty() const { return cells.empty() || contacts_by_sphere.empty(); }
};

// updateable.rs:154-742
template <class R>
class UpdateableTessellation {
 public:
  explicit UpdateableTessellation(bool backup_enabled, int threads = 1) : backup_enabled_(backup_enabled), threads_(threads) {}

  // updateable.rs:187-202
  bool init(const std::vector<Ball<R>>& balls, R probe, const PeriodicBox<R>* pbox) {
    probe_ = probe;
    prepare_for_update();
    std::vector<Sphere<R>> sp;
    for (const auto& b : balls) sp.push_back(sphere_from_ball(b, probe));
    st_.container.init(sp, pbox, threads_);
    st_.was_full_reinit = true;
    st_.changed_ids.clear();
    st_.affected_ids.clear();
    compute_full_tessellation();
    return !st_.result_empty();
  }
  // Moved spheres leave their old grid boxes (SpheresSearcher::update); the
  // reference's behaviour is the default, off.
  void set_remove_from_old_box(bool on) { st_.container.set_remove_from_old_box(on); }
  // updateable.rs:208-219
  bool update(const std::vector<Ball<R>>& balls) { return update_internal(balls, nullptr); }
  bool update_with_changed(const std::vector<Ball<R>>& balls, const std::vector<size_t>& changed) {
    if (changed.empty()) return false;
    return update_internal(balls, &changed);
  }
  // updateable.rs:225-272
  bool set_exclusion(size_t id, bool excluded) {
    if (st_.result_empty() || id >= st_.contacts_by_sphere.size() || id >= st_.container.exclusion_statuses().size())
      return false;
    if (st_.container.is_excluded(id) == excluded) return false;
    prepare_for_update();
    std::vector<size_t> affected;
    if (excluded) {
      affected.push_back(id);
      for (const auto& c : st_.contacts_by_sphere[id]) {
        const size_t nb = c.id_a == id ? c.id_b : c.id_a;
        auto it = std::lower_bound(affected.begin(), affected.end(), nb);
        if (it == affected.end() || *it != nb) affected.insert(it, nb);
      }
      st_.container.set_exclusion(id, true);
    } else {
      auto a = st_.container.set_exclusion(id, false);
      if (!a) return false;
      affected = *a;
    }
    st_.changed_ids = {id};
    st_.affected_ids = affected;
    st_.was_full_reinit = false;
    update_using_affected();
    return true;
  }
  // updateable.rs:277-322
  bool restore() {
    if (!backup_enabled_ || !backup_) return false;
    UpdState<R> backup = std::move(*backup_);
    backup_.reset();
    const std::vector<size_t> affected = st_.affected_ids;
    st_.container.restore_from(backup.container, affected);
    const size_t n = st_.contacts_by_sphere.size();
    bool needs_rebuild;
    if (st_.was_full_reinit || backup.contacts_by_sphere.size() != n || backup.cell_summaries.size() != n) {
      st_.num_balls = backup.num_balls;
      st_.cells = backup.cells;
      st_.contacts_by_sphere = backup.contacts_by_sphere;
      st_.cell_summaries = backup.cell_summaries;
      needs_rebuild = false;
    } else {
      for (size_t id : affected)
        if (id < n) {
          st_.contacts_by_sphere[id] = backup.contacts_by_sphere[id];
          st_.cell_summaries[id] = backup.cell_summaries[id];
        }
      needs_rebuild = true;
    }
    st_.changed_ids = backup.changed_ids;
    st_.affected_ids = backup.affected_ids;
    st_.was_full_reinit = backup.was_full_reinit;
    if (needs_rebuild) rebuild_cells_from_summaries();
    return true;
  }
  // updateable.rs:333-376
  TessellationResult<R> summary() const {
    TessellationResult<R> r;
    const size_t n = st_.contacts_by_sphere.size();
    r.num_balls = n;
    for (size_t i = 0; i < n; ++i)
      for (const auto& c : st_.contacts_by_sphere[i])
        if (c.id_a == i) r.contacts.push_back(Contact<R>{c.id_a, c.id_b, c.area, c.arc_length, c.central});
    for (size_t i = 0; i < n; ++i) {
      const auto& cs = st_.cell_summaries[i];
      if (cs.stage == kSasComputed) r.cells.push_back(Cell<R>{cs.id, cs.sas_area, cs.sas_inside_volume});
    }
    return r;
  }
  const std::vector<size_t>& changed_ids() const { return st_.changed_ids; }
  const std::vector<size_t>& affected_ids() const { return st_.affected_ids; }
  bool last_update_was_full_reinit() const { return st_.was_full_reinit; }
  bool backup_enabled() const { return backup_enabled_; }
  const std::vector<std::vector<Summary<R>>>& contacts_by_sphere() const { return st_.contacts_by_sphere; }
  const std::vector<Cell<R>>& cells() const { return st_.cells; }
  size_t num_balls() const { return st_.num_balls; }

 private:
  // updateable.rs:404-420
  void prepare_for_update() {
    if (backup_enabled_ && !st_.result_empty()) backup_ = st_;
    st_.changed_ids.clear();
    st_.affected_ids.clear();
    st_.was_full_reinit = false;
  }
  // updateable.rs:422-447
  bool update_internal(const std::vector<Ball<R>>& balls, const std::vector<size_t>* changed) {
    prepare_for_update();
    std::vector<Sphere<R>> ns;
    for (const auto& b : balls) ns.push_back(sphere_from_ball(b, probe_));
    auto ur = st_.container.update(ns, changed, threads_);
    if (!ur) return false;
    if (ur->was_full_reinit) {
      st_.was_full_reinit = true;
      compute_full_tessellation();
      return true;
    }
    st_.changed_ids = ur->changed_ids;
    st_.affected_ids = ur->affected_ids;
    st_.was_full_reinit = false;
    update_using_affected();
    return true;
  }
  // updateable.rs:449-474
  void compute_full_tessellation() {
    const size_t n = st_.container.spheres().size();
    const bool periodic = st_.container.periodic();
    const auto pairs = collect_pairs(nullptr);
    std::vector<Summary<R>> all, dedup;
    summaries_with_dedup(pairs, periodic, n, &all, &dedup);
    st_.num_balls = n;
    st_.contacts_by_sphere.assign(n, {});
    for (const auto& c : dedup)
      if (c.area > R(0.0)) {
        st_.contacts_by_sphere[c.id_a].push_back(c);
        if (c.id_b != c.id_a) st_.contacts_by_sphere[c.id_b].push_back(c);
      }
    cells_from_all_summaries(all, periodic, n);
  }
  // updateable.rs:476-514
  void update_using_affected() {
    const size_t n = st_.container.spheres().size();
    const bool periodic = st_.container.periodic();
    std::vector<char> inv(n, 0);
    for (size_t id : st_.affected_ids)
      if (id < n) inv[id] = 1;
    const auto pairs = collect_pairs(&inv);
    std::vector<Summary<R>> all, fresh;
    summaries_with_dedup(pairs, periodic, n, &all, &fresh);
    for (size_t sid : st_.affected_ids)
      if (sid < n) {
        auto& v = st_.contacts_by_sphere[sid];
        v.erase(std::remove_if(v.begin(), v.end(), [&](const Summary<R>& c) { return inv[c.id_a % n] && inv[c.id_b % n]; }),
                v.end());
      }
    for (const auto& c : fresh)
      if (c.area > R(0.0)) {
        st_.contacts_by_sphere[c.id_a].push_back(c);
        if (c.id_b != c.id_a) st_.contacts_by_sphere[c.id_b].push_back(c);
      }
    recompute_cells_for_affected();
  }
  // updateable.rs:516-561
  std::vector<std::pair<size_t, size_t>> collect_pairs(const std::vector<char>* inv) const {
    const size_t n = st_.container.spheres().size();
    const auto& all = st_.container.all_colliding_ids();
    const auto& excl = st_.container.exclusion_statuses();
    const bool periodic = st_.container.periodic();
    std::vector<std::pair<size_t, size_t>> pairs;
    for (size_t a = 0; a < all.size() && a < n; ++a) {
      const bool a_inv = !inv || (a < inv->size() && (*inv)[a]);
      if (!a_inv || (a < excl.size() && excl[a])) continue;
      for (const auto& nb : all[a]) {
        const size_t b = nb.index;
        const size_t bc = b % n;
        const bool b_inv = !inv || (bc < inv->size() && (*inv)[bc]);
        if (!b_inv || (bc < excl.size() && excl[bc])) continue;
        const bool include = (periodic && b >= n) ? true : a < b;
        if (include) pairs.emplace_back(a, b);
      }
    }
    return pairs;
  }
  // updateable.rs:565-630
  void summaries_with_dedup(const std::vector<std::pair<size_t, size_t>>& pairs, bool periodic, size_t n,
                            std::vector<Summary<R>>* all, std::vector<Summary<R>>* dedup) const {
    const auto& sp = st_.container.populated_spheres();
    const auto& coll = st_.container.all_colliding_ids();
    std::vector<std::optional<Summary<R>>> tmp(pairs.size());
    parallel_for(pairs.size(), threads_, [&](size_t k) {
      const size_t a = pairs[k].first, b = pairs[k].second;
      auto cd = construct_contact_descriptor(sp, a, b, coll[a]);
      if (!cd) return;
      Summary<R> s = cd->to_summary();
      s.id_a = a;
      s.id_b = b;
      s.ensure_ids_ordered();
      if (s.area > R(0.0)) tmp[k] = s;
    });
    all->clear();
    for (auto& t : tmp)
      if (t) all->push_back(*t);
    if (periodic) {
      dedup->clear();
      for (auto s : deduplicate_periodic_contacts(*all, n)) {
        s.id_a %= n;
        s.id_b %= n;
        s.ensure_ids_ordered();
        dedup->push_back(s);
      }
    } else {
      *dedup = *all;
    }
  }
  // updateable.rs:633-689
  void cells_from_all_summaries(const std::vector<Summary<R>>& all, bool periodic, size_t n) {
    const auto& sp = st_.container.spheres();
    st_.cell_summaries.assign(n, CellSummary<R>{});
    for (size_t i = 0; i < n; ++i) st_.cell_summaries[i].id = i;
    if (periodic) {
      for (const auto& c : all)
        if (c.area > R(0.0)) {
          if (c.id_a < n) st_.cell_summaries[c.id_a].add(c);
          if (c.id_b < n && c.id_b != c.id_a) st_.cell_summaries[c.id_b].add(c);
        }
    } else {
      for (size_t i = 0; i < n; ++i)
        for (const auto& c : st_.contacts_by_sphere[i]) st_.cell_summaries[i].add(c);
    }
    for (size_t i = 0; i < n && i < sp.size(); ++i) {
      auto& cs = st_.cell_summaries[i];
      if (cs.stage == kContactsAdded)
        cs.compute_sas(sp[i].r);
      else if (cs.stage == kInit && !st_.container.is_excluded(i) && st_.container.colliding_ids(i).empty())
        cs.compute_sas_detached(i, sp[i].r);
    }
    rebuild_cells_from_summaries();
  }
  // updateable.rs:691-727
  void recompute_cells_for_affected() {
    const auto& sp = st_.container.spheres();
    for (size_t sid : st_.affected_ids) {
      if (sid >= sp.size()) continue;
      CellSummary<R> cs;
      cs.id = sid;
      for (const auto& c : st_.contacts_by_sphere[sid]) cs.add(c);
      if (cs.stage == kContactsAdded) {
        cs.compute_sas(sp[sid].r);
      } else if (cs.stage == kInit && !st_.container.is_excluded(sid)) {
        if (st_.contacts_by_sphere[sid].empty()) cs.compute_sas_detached(sid, sp[sid].r);
      }
      st_.cell_summaries[sid] = cs;
    }
    rebuild_cells_from_summaries();
  }
  // updateable.rs:729-741
  void rebuild_cells_from_summaries() {
    st_.cells.clear();
    for (const auto& cs : st_.cell_summaries)
      if (cs.stage == kSasComputed) st_.cells.push_back(Cell<R>{cs.id, cs.sas_area, cs.sas_inside_volume});
  }

  UpdState<R> st_;
  std::optional<UpdState<R>> backup_;
  bool backup_enabled_;
  int threads_;
  R probe_ = R(0.0);
};


// ---------------------------------------------- subdivided_icosahedron.rs ----
// subdivided_icosahedron.rs:51-131: unit vertices in order of first appearance
// (the 12 base vertices, then each new edge midpoint as the faces are walked).
inline std::vector<V3<double>> subdivided_icosahedron(int depth) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  const double base[12][3] = {{t, 1.0, 0.0},  {-t, 1.0, 0.0},  {t, -1.0, 0.0}, {-t, -1.0, 0.0},
                              {1.0, 0.0, t},  {1.0, 0.0, -t},  {-1.0, 0.0, t}, {-1.0, 0.0, -t},
                              {0.0, t, 1.0},  {0.0, -t, 1.0},  {0.0, t, -1.0}, {0.0, -t, -1.0}};
  std::vector<V3<double>> v;
  for (const auto& b : base) v.push_back(normalize(V3<double>{b[0], b[1], b[2]}));
  std::vector<std::array<size_t, 3>> tri = {{0, 8, 4},  {1, 10, 7}, {2, 9, 11}, {7, 3, 1},  {0, 5, 10},
                                            {3, 9, 6},  {3, 11, 9}, {8, 6, 4},  {2, 4, 9},  {3, 7, 11},
                                            {4, 2, 0},  {9, 4, 6},  {2, 11, 5}, {0, 10, 8}, {5, 0, 2},
                                            {10, 5, 7}, {1, 6, 8},  {1, 8, 10}, {6, 1, 3},  {11, 7, 5}};
  for (int d = 0; d < depth; ++d) {
    std::vector<std::array<size_t, 3>> next;
    std::map<std::pair<size_t, size_t>, size_t> mids;
    for (const auto& f : tri) {
      const std::pair<size_t, size_t> edges[3] = {{f[1], f[2]}, {f[0], f[2]}, {f[0], f[1]}};
      size_t mid[3];
      for (int e = 0; e < 3; ++e) {
        const size_t a = edges[e].first, b = edges[e].second;
        const auto key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
        auto it = mids.find(key);
        if (it == mids.end()) {
          v.push_back(normalize((v[a] + v[b]) * 0.5));
          it = mids.emplace(key, v.size() - 1).first;
        }
        mid[e] = it->second;
      }
      next.push_back({f[0], mid[1], mid[2]});
      next.push_back({f[1], mid[0], mid[2]});
      next.push_back({f[2], mid[0], mid[1]});
      next.push_back({mid[0], mid[1], mid[2]});
    }
    tri.swap(next);
  }
  return v;
}

// ------------------------------------------------------ solvent_spheres.rs ----
struct SolventSphere {
  double x, y, z, radius, weight;
  size_t parent;
};

// solvent_spheres.rs:131-425 (validation is the caller's). Stage 1: exposed
// atoms (cells with sas_area > 0) are sampled on the icosahedron at radius
// r + probe; a sample is kept when its distance to the atom's surface is
// smaller than to the surface of every contact partner (canonical balls).
// Stage 2: tessellation of the atoms (radius max(r - probe + volume_probe,
// 0.01)) and samples (radius probe + volume_probe) with probe 0; a sample's
// weight is its cell volume (0 without a cell).
inline std::vector<SolventSphere> compute_solvent_spheres(const std::vector<Ball<double>>& balls, double probe,
                                                          double volume_probe, int depth,
                                                          const PeriodicBox<double>* pbox, int threads) {
  std::vector<SolventSphere> out;
  if (balls.empty()) return out;
  const auto s1 = compute_tessellation<double>(balls, probe, pbox, nullptr, 0, false, threads);
  const size_t n = balls.size();
  std::vector<char> exposed(n, 0);
  for (const auto& c : s1.cells)
    if (c.sas_area > 0.0) exposed[c.index] = 1;
  std::vector<std::vector<size_t>> nbrs(n);
  for (const auto& c : s1.contacts) {
    if (exposed[c.id_a]) nbrs[c.id_a].push_back(c.id_b);
    if (exposed[c.id_b]) nbrs[c.id_b].push_back(c.id_a);
  }
  const auto sih = subdivided_icosahedron(depth);
  for (const auto& cell : s1.cells) {
    if (!(cell.sas_area > 0.0)) continue;
    const Ball<double>& at = balls[cell.index];
    const V3<double> center{at.x, at.y, at.z};
    const double radius = at.r + probe;
    for (const auto& v : sih) {
      const V3<double> p = center + v * radius;
      const double d_atom = norm(p - center) - at.r;
      bool ok = true;
      for (size_t nb : nbrs[cell.index]) {
        const Ball<double>& o = balls[nb];
        if (norm(p - V3<double>{o.x, o.y, o.z}) - o.r <= d_atom) {
          ok = false;
          break;
        }
      }
      if (ok) out.push_back(SolventSphere{p.x, p.y, p.z, probe, 0.0, cell.index});
    }
  }
  if (out.empty()) return out;
  std::vector<Ball<double>> comb;
  for (const auto& b : balls) comb.push_back(Ball<double>{b.x, b.y, b.z, std::max(b.r - probe + volume_probe, 0.01)});
  for (const auto& sp : out) comb.push_back(Ball<double>{sp.x, sp.y, sp.z, probe + volume_probe});
  const auto s2 = compute_tessellation<double>(comb, 0.0, pbox, nullptr, 0, false, threads);
  for (const auto& c : s2.cells)
    if (c.index >= n) out[c.index - n].weight = c.volume;
  return out;
}

}  // namespace vltr_oracle
