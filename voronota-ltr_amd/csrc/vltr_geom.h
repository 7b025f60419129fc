// FP64 geometry for the tessellation kernels (device and host).
//
// Operation order follows the reference exactly so results are bit-identical
// to it wherever IEEE-754 fixes the result (+ - * / sqrt fma): the file is
// compiled with -ffp-contract=off and fma() appears only where the reference
// calls mul_add. acos/sin come from the device math library (ocml); the
// reference uses the host libm, so those two differ by at most an ulp or two.
//
// Reference: src/geometry.rs (primitives), src/types.rs (Sphere, shifts).
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VLTR_HD __host__ __device__ __forceinline__
#else
// Host-only build of the same code (CPU tests of the kernel logic).
#define VLTR_HD inline
using std::acos;
using std::fabs;
using std::floor;
using std::fma;
using std::fmax;
using std::fmin;
using std::sin;
using std::sqrt;
#endif

namespace vltr {

// geometry.rs:17 and std::f64::consts (correctly rounded).
constexpr double kEps = 1e-10;
constexpr double kPi = 0x1.921fb54442d18p+1;
constexpr double kTau = 0x1.921fb54442d18p+2;
constexpr double kFracPi3 = 0x1.0c152382d7366p+0;
// contact.rs:448
constexpr double kHexagonScale = 1.19;

struct V3 {
  double x, y, z;
};

VLTR_HD V3 v3(double x, double y, double z) { return V3{x, y, z}; }
VLTR_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
VLTR_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
VLTR_HD V3 operator*(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
VLTR_HD V3 operator/(V3 a, double s) { return V3{a.x / s, a.y / s, a.z / s}; }
VLTR_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
// nalgebra 3-vector dot: (x*x' + y*y') + z*z'
VLTR_HD double dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
VLTR_HD double norm2(V3 a) { return dot(a, a); }
VLTR_HD double norm(V3 a) { return sqrt(norm2(a)); }
VLTR_HD V3 unit(V3 a) { return a / norm(a); }
VLTR_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// Ball expanded by the probe: centre + radius R.
struct Sph {
  V3 c;
  double r;
};

// Fuzzy comparisons, geometry.rs:22-54.
VLTR_HD bool f_eq(double a, double b) { return fabs(a - b) <= kEps; }
VLTR_HD bool f_lt(double a, double b) { return a + kEps < b; }
VLTR_HD bool f_gt(double a, double b) { return a - kEps > b; }
VLTR_HD bool f_le(double a, double b) { return a < b + kEps; }
VLTR_HD bool f_ge(double a, double b) { return a + kEps > b; }

// geometry.rs:66-82
VLTR_HD bool intersects(const Sph& a, const Sph& b) {
  const double s = a.r + b.r;
  return f_lt(norm2(b.c - a.c), s * s);
}
VLTR_HD bool equals(const Sph& a, const Sph& b) {
  return f_eq(a.r, b.r) && f_eq(a.c.x, b.c.x) && f_eq(a.c.y, b.c.y) && f_eq(a.c.z, b.c.z);
}
VLTR_HD bool contains(const Sph& a, const Sph& b) {
  const double d = a.r - b.r;
  return f_ge(a.r, b.r) && f_le(norm2(b.c - a.c), d * d);
}

// cos of the angle at a's centre towards the a|b intersection circle, with the
// distance |b - a| (geometry.rs:208-239 share this expression).
VLTR_HD double circle_cos(const Sph& a, const Sph& b, double cm) {
  return fma(b.r, -b.r, fma(a.r, a.r, cm * cm)) / (2.0 * a.r * cm);
}
// geometry.rs:208-215: sort key of a neighbor.
VLTR_HD double circle_center_distance(const Sph& a, const Sph& b) {
  const double cm = norm(b.c - a.c);
  if (cm < kEps) return 0.0;
  return a.r * circle_cos(a, b, cm);
}
// geometry.rs:218-226: a point of the radical plane of a and b.
VLTR_HD V3 circle_center(const Sph& a, const Sph& b) {
  const V3 cv = b.c - a.c;
  const double cm = norm(cv);
  if (cm < kEps) return a.c;
  return a.c + cv * (a.r * circle_cos(a, b, cm) / cm);
}
// geometry.rs:229-239: intersection circle as (centre, radius).
VLTR_HD Sph intersection_circle(const Sph& a, const Sph& b) {
  const V3 cv = b.c - a.c;
  const double cm = norm(cv);
  if (cm < kEps) return Sph{a.c, 0.0};
  const double cg = circle_cos(a, b, cm);
  const double sg = sqrt(fmax(1.0 - cg * cg, 0.0));
  return Sph{a.c + cv * (a.r * cg / cm), a.r * sg};
}

// geometry.rs:118-131 (unit normal)
VLTR_HD int halfspace(V3 pp, V3 n, V3 x) {
  const double sd = dot(n, x - pp);
  if (f_gt(sd, 0.0)) return 1;
  if (f_lt(sd, 0.0)) return -1;
  return 0;
}
// geometry.rs:134-148; `nn` is unit(n) (the reference re-normalizes here).
// plane_segment_d takes the two signed distances da = dot(nn, a - pp) and
// db = dot(nn, b - pp) already computed.
VLTR_HD V3 plane_segment_d(V3 a, V3 b, double da, double db) {
  if (fabs(da - db) < kEps) return a;
  const double t = da / (da - db);
  return a + (b - a) * t;
}
VLTR_HD V3 plane_segment(V3 pp, V3 nn, V3 a, V3 b) { return plane_segment_d(a, b, dot(nn, a - pp), dot(nn, b - pp)); }
// geometry.rs:152-154
VLTR_HD double triangle_area(V3 a, V3 b, V3 c) { return norm(cross(b - a, c - a)) / 2.0; }
// Rust f64::clamp(-1, 1): NaN passes through.
VLTR_HD double clamp1(double x) {
  if (x < -1.0) x = -1.0;
  if (x > 1.0) x = 1.0;
  return x;
}
// geometry.rs:157-176
VLTR_HD double directed_angle(V3 o, V3 a, V3 b, V3 c) {
  const V3 v1 = unit(a - o);
  const V3 v2 = unit(b - o);
  const double ang = acos(clamp1(dot(v1, v2)));
  const V3 n = cross(v1, v2);
  return dot(c - o, n) >= 0.0 ? ang : kTau - ang;
}
// geometry.rs:179-193
VLTR_HD V3 any_normal(V3 a) {
  const bool nx = !f_eq(a.x, 0.0), ny = !f_eq(a.y, 0.0), nz = !f_eq(a.z, 0.0);
  const bool flip_x = nx && (ny || nz);
  const bool flip_y = !flip_x && ny && (nx || nz);
  if (flip_x || flip_y) {
    // one normalization for both cases: a with its x (or else y) negated
    const V3 b{flip_x ? -a.x : a.x, flip_x ? a.y : -a.y, a.z};
    return unit(cross(a, b));
  }
  return nx ? V3{0.0, 1.0, 0.0} : V3{1.0, 0.0, 0.0};
}
// geometry.rs:196-203 with the unit quaternion (cos t/2, u sin t/2) given by
// its sine/cosine (the hexagon angles are fixed, so they are precomputed).
// rotate_unit takes u = unit(axis) already computed.
VLTR_HD V3 rotate_unit(V3 axis, V3 u, double s_half, double c_half, V3 p) {
  if (norm2(axis) <= 0.0) return p;
  const V3 qv = u * s_half;
  const V3 t = cross(qv, p) * 2.0;
  const V3 cr = cross(qv, t);
  return (t * c_half + cr) + p;
}
VLTR_HD V3 rotate(V3 axis, double s_half, double c_half, V3 p) { return rotate_unit(axis, unit(axis), s_half, c_half, p); }
// geometry.rs:243-255
VLTR_HD bool project_inside_segment(V3 o, V3 a, V3 b, V3* out) {
  const V3 v = unit(b - a);
  const double l = dot(v, o - a);
  if (l > 0.0 && l * l <= norm2(b - a)) {
    *out = a + v * l;
    return true;
  }
  return false;
}
// geometry.rs:258-280
VLTR_HD bool segment_circle(const Sph& circle, V3 p_in, V3 p_out, V3* out) {
  const double dist = norm(p_in - p_out);
  if (dist <= 0.0) return false;
  const V3 v = (p_in - p_out) / dist;
  const V3 u = circle.c - p_out;
  const V3 s = p_out + v * dot(v, u);
  const double ll = fma(circle.r, circle.r, -norm2(circle.c - s));
  if (ll >= 0.0) {
    *out = s - v * sqrt(ll);
    return true;
  }
  return false;
}
// geometry.rs:283-293
VLTR_HD double min_dihedral(V3 o, V3 a, V3 b1, V3 b2) {
  const V3 oa = unit(a - o);
  const V3 d1 = b1 - (o + oa * dot(oa, b1 - o));
  const V3 d2 = b2 - (o + oa * dot(oa, b2 - o));
  return acos(clamp1(dot(unit(d1), unit(d2))));
}

VLTR_HD int ctz64(uint64_t x) { return __builtin_ctzll(x); }

// Workgroups are dealt to the 8 XCDs of an MI355X round-robin (block b runs
// on XCD b % 8). xcd_remap turns the hardware block index into a logical one
// so that each XCD walks one contiguous range of logical blocks (neighboring
// work shares that XCD's L2). Bijective on [0, g) for any g.
VLTR_HD int64_t xcd_remap(int64_t b, int64_t g) {
  const int64_t per = g / 8, rem = g % 8;
  const int64_t x = b % 8, i = b / 8;
  return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

// types.rs:322-345: periodic image, per component c*wc + (b*wb + (a*wa + s)).
struct PBox {
  V3 a, b, c;
};
VLTR_HD V3 shift_point(const PBox& pb, V3 s, double wa, double wb, double wc) {
  return V3{fma(pb.c.x, wc, fma(pb.b.x, wb, fma(pb.a.x, wa, s.x))),
            fma(pb.c.y, wc, fma(pb.b.y, wb, fma(pb.a.y, wa, s.y))),
            fma(pb.c.z, wc, fma(pb.b.z, wb, fma(pb.a.z, wa, s.z)))};
}
// types.rs:349-359: shift g (0..25) of the 26 non-zero lattice offsets in
// lexicographic (x, y, z) order over {-1,0,1}^3.
VLTR_HD void lattice_shift(int g, int* sx, int* sy, int* sz) {
  const int k = g < 13 ? g : g + 1;
  *sx = k / 9 - 1;
  *sy = (k / 3) % 3 - 1;
  *sz = k % 3 - 1;
}

}  // namespace vltr
