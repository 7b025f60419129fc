// Contact clipping for one candidate pair (replaces src/contact.rs:196-848).
//
// Host+device code: the kernels in k_contact.hip instantiate it with
// fixed-capacity register/scratch arrays or a global workspace, and the CPU
// tests compile the same header for the host to check it pair by pair.
#pragma once

#include "vltr_geom.h"

namespace vltr {

// Contour capacity of the fast cut kernel (points while cutting / after the
// restriction to the circle). Contacts that would exceed it are flagged and
// re-run by the slow kernel, whose contour lives in a global workspace.
constexpr int kMaxC = 16;
constexpr int kMaxR = 2 * kMaxC;

// Per-pair record written by the cut kernels (types.rs:380-415 summary).
// flags: bit0 = contact exists (area > 0), bit1 = central, bit2 = overflow
// (fast kernel ran out of contour capacity; slow kernel must redo it).
// One record is one 64-byte memory line: records are written and read at
// scattered pair indices, and a 56-byte record straddled two lines at six of
// its eight offsets.
struct alignas(64) PairRec {
  double area;
  double arc_length;
  double solid_a, solid_b;
  double pyr_a, pyr_b;
  uint32_t flags;
  uint32_t pad;
  uint64_t spare;
};
static_assert(sizeof(PairRec) == 64, "a record is one 64-byte line");
constexpr uint32_t kRecValid = 1u;
constexpr uint32_t kRecCentral = 2u;
constexpr uint32_t kRecOverflow = 4u;

// Sphere i of an array of (x, y, z, R) quadruples.
VLTR_HD Sph load_sph(const double* sph4, int64_t i) {
  const double* q = sph4 + 4 * i;
  return Sph{V3{q[0], q[1], q[2]}, q[3]};
}

// Contour storage while cutting: points plus (left, right) ids.
template <int CAP>
struct LocalCut {
  static constexpr int cap = CAP;
  double x[CAP], y[CAP], z[CAP];
  int32_t l[CAP], r[CAP];
  VLTR_HD void init(int) {}
  VLTR_HD V3 p(int k) const { return V3{x[k], y[k], z[k]}; }
  VLTR_HD void set(int k, V3 v, int32_t lid, int32_t rid) {
    x[k] = v.x;
    y[k] = v.y;
    z[k] = v.z;
    l[k] = lid;
    r[k] = rid;
  }
  VLTR_HD void copy(int dst, int src) {
    x[dst] = x[src];
    y[dst] = y[src];
    z[dst] = z[src];
    l[dst] = l[src];
    r[dst] = r[src];
  }
};

// Restricted contour: points, ids and arc angles.
template <int CAP>
struct LocalArc {
  static constexpr int cap = CAP;
  double x[CAP], y[CAP], z[CAP], ang[CAP];
  int32_t l[CAP], r[CAP];
  VLTR_HD void init(int) {}
  VLTR_HD V3 p(int k) const { return V3{x[k], y[k], z[k]}; }
  VLTR_HD void set(int k, V3 v, int32_t lid, int32_t rid) {
    x[k] = v.x;
    y[k] = v.y;
    z[k] = v.z;
    l[k] = lid;
    r[k] = rid;
    ang[k] = 0.0;
  }
};

// Global-workspace variants for the slow kernel (one region per lane).
struct GlobalCut {
  int cap;
  double *x, *y, *z;
  int32_t *l, *r;
  VLTR_HD V3 p(int k) const { return V3{x[k], y[k], z[k]}; }
  VLTR_HD void set(int k, V3 v, int32_t lid, int32_t rid) {
    x[k] = v.x;
    y[k] = v.y;
    z[k] = v.z;
    l[k] = lid;
    r[k] = rid;
  }
  VLTR_HD void copy(int dst, int src) {
    x[dst] = x[src];
    y[dst] = y[src];
    z[dst] = z[src];
    l[dst] = l[src];
    r[dst] = r[src];
  }
};
struct GlobalArc {
  int cap;
  double *x, *y, *z, *ang;
  int32_t *l, *r;
  VLTR_HD V3 p(int k) const { return V3{x[k], y[k], z[k]}; }
  VLTR_HD void set(int k, V3 v, int32_t lid, int32_t rid) {
    x[k] = v.x;
    y[k] = v.y;
    z[k] = v.z;
    l[k] = lid;
    r[k] = rid;
    ang[k] = 0.0;
  }
};

template <class S>
VLTR_HD int cap_of(const S& s) {
  return s.cap;
}

enum ClipStatus : int { kNone = 0, kContact = 1, kOverflow = 2 };

struct ClipOut {
  Sph ic;
  V3 axis;
  V3 bary;
  double sum_arc;
  double area;
  int m;  // restricted contour length (0: uncut full circle)
};

// contact.rs:472-481
template <class S>
VLTR_HD bool cuttable(const S& ct, int n, V3 a_c, V3 cut_pt) {
  const double thr = norm2(cut_pt - a_c);
  for (int k = 0; k < n; ++k)
    if (norm2(ct.p(k) - a_c) >= thr) return true;
  return false;
}

// contact.rs:484-627 (mark_and_cut_contour, mark_contour, cut_contour).
// `nn` is unit(pn), used by the plane-segment intersections.
template <class S>
VLTR_HD int mark_and_cut(S& ct, int& n, V3 pc, V3 pn, V3 nn, int32_t cid) {
  int outs = 0;
  for (int k = 0; k < n; ++k)
    if (halfspace(pc, pn, ct.p(k)) >= 0) {
      ct.l[k] = cid;
      ct.r[k] = cid;
      ++outs;
    }
  if (outs == 0) return kContact;
  if (outs >= n) {
    n = 0;
    return kContact;
  }
  if (n < 3) return kContact;
  int first = -1, last = -1;
  for (int k = 0; k < n; ++k)
    if (ct.l[k] == cid && ct.r[k] == cid) {
      if (first < 0) first = k;
      last = k;
    }
  if (first < 0) return kContact;
  int start = first, end = last;
  if (start == 0 && end == n - 1) {
    end = 0;
    while (end + 1 < n && ct.l[end + 1] == cid && ct.r[end + 1] == cid) ++end;
    start = n - 1;
    while (start > 0 && ct.l[start - 1] == cid && ct.r[start - 1] == cid) --start;
  }
  if (start == end) {
    if (n + 1 > cap_of(ct)) return kOverflow;
    for (int k = n; k > start; --k) ct.copy(k, k - 1);
    ++n;
    end = start + 1;
  } else if (start < end) {
    if (start + 1 < end) {
      const int d = end - (start + 1);
      for (int k = start + 1; k + d < n; ++k) ct.copy(k, k + d);
      n -= d;
    }
    end = start + 1;
  } else {
    if (start + 1 < n) n = start + 1;
    if (end > 0) {
      for (int k = 0; k + end < n; ++k) ct.copy(k, k + end);
      n -= end;
    }
    start = n - 1;
    end = 0;
  }
  const int il = start > 0 ? start - 1 : n - 1;
  const V3 lx = plane_segment(pc, nn, ct.p(start), ct.p(il));
  ct.set(start, lx, ct.r[il], ct.l[start]);
  const int ir = end + 1 < n ? end + 1 : 0;
  const V3 rx = plane_segment(pc, nn, ct.p(end), ct.p(ir));
  ct.set(end, rx, ct.r[end], ct.l[ir]);
  if (!f_gt(norm2(ct.p(end) - ct.p(start)), 0.0)) {
    ct.r[start] = ct.r[end];
    for (int k = end; k + 1 < n; ++k) ct.copy(k, k + 1);
    --n;
  }
  return kContact;
}

template <class A>
VLTR_HD bool arc_push(A& out, int& m, V3 v, int32_t lid, int32_t rid) {
  if (m >= cap_of(out)) return false;
  out.set(m, v, lid, rid);
  ++m;
  return true;
}

// contact.rs:631-769 as one streaming pass. Point i of the cut contour is
// emitted after segment i has been processed (its ids are final then), its
// insertions follow it; point 0's ids may still change in the last segment,
// so its output slot is patched at the end.
template <class S, class A>
VLTR_HD int restrict_to_circle(S& ct, int n, A& out, int& m, const Sph& ic, V3 axis, int32_t a_id,
                                                  double* sum) {
  *sum = 0.0;
  m = 0;
  const double rr = ic.r * ic.r;
  int outs = 0;
  for (int k = 0; k < n; ++k) outs += (norm2(ct.p(k) - ic.c) <= rr) ? 0 : 1;
  if (outs == 0) {
    for (int k = 0; k < n; ++k)
      if (!arc_push(out, m, ct.p(k), ct.l[k], ct.r[k])) return kOverflow;
    return kContact;
  }
  int insertions = 0;
  const bool in0 = norm2(ct.p(0) - ic.c) <= rr;
  if (in0 && !arc_push(out, m, ct.p(0), ct.l[0], ct.r[0])) return kOverflow;
  for (int i = 0; i < n; ++i) {
    const int ni = (i + 1) % n;
    const V3 pi = ct.p(i), pn = ct.p(ni);
    const bool o1 = !(norm2(pi - ic.c) <= rr);
    const bool o2 = !(norm2(pn - ic.c) <= rr);
    V3 ins1 = pi, ins2 = pi;
    int32_t i1l = 0, i1r = 0, i2l = 0, i2r = 0;
    int nins = 0;
    if (o1 || o2) {
      if (o1 && o2) {
        V3 mp;
        if (project_inside_segment(ic.c, pi, pn, &mp) && norm2(mp - ic.c) <= rr) {
          V3 ip1, ip2;
          const bool ok1 = segment_circle(ic, mp, pi, &ip1);
          const bool ok2 = segment_circle(ic, mp, pn, &ip2);
          if (ok1 && ok2) {
            ins1 = ip1;
            i1l = a_id;
            i1r = ct.r[i];
            ins2 = ip2;
            i2l = ct.l[ni];
            i2r = a_id;
            nins = 2;
          }
        }
      } else if (o1) {
        V3 ip;
        if (segment_circle(ic, pn, pi, &ip)) {
          ins1 = ip;
          i1l = a_id;
          i1r = ct.r[i];
          nins = 1;
        } else {
          ct.l[ni] = a_id;
          ct.r[ni] = ct.r[i];
        }
      } else {
        V3 ip;
        if (segment_circle(ic, pi, pn, &ip)) {
          ins1 = ip;
          i1l = ct.l[ni];
          i1r = a_id;
          nins = 1;
        } else {
          ct.l[i] = ct.l[ni];
          ct.r[i] = a_id;
        }
      }
    }
    if (i > 0 && !o1 && !arc_push(out, m, pi, ct.l[i], ct.r[i])) return kOverflow;
    if (nins >= 1 && !arc_push(out, m, ins1, i1l, i1r)) return kOverflow;
    if (nins == 2 && !arc_push(out, m, ins2, i2l, i2r)) return kOverflow;
    insertions += nins;
  }
  if (in0) {
    out.l[0] = ct.l[0];
    out.r[0] = ct.r[0];
  }
  if (insertions == 0 || m < 2) {
    m = 0;
    return kContact;
  }
  const V3 ctrl = ic.c + axis;
  double s = 0.0;
  for (int k = 0; k < m; ++k) {
    const int nk = (k + 1) % m;
    if (out.r[k] == a_id && out.l[nk] == a_id) {
      const double ang = directed_angle(ic.c, out.p(k), out.p(nk), ctrl);
      out.ang[k] = ang;
      s = s + ang;
    }
  }
  *sum = s;
  if (f_ge(s, kTau) || (m > 2 && f_eq(s, kTau))) {
    *sum = kTau;
    m = 0;
  }
  return kContact;
}

// contact.rs:773-793
template <class A>
VLTR_HD double contour_area(const A& ct, int m, const Sph& ic, V3* bary) {
  V3 sum{0.0, 0.0, 0.0};
  for (int k = 0; k < m; ++k) sum = sum + ct.p(k);
  *bary = sum / static_cast<double>(m);
  double area = 0.0;
  for (int k = 0; k < m; ++k) {
    const int nk = (k + 1) % m;
    area = area + triangle_area(*bary, ct.p(k), ct.p(nk));
    const double ang = ct.ang[k];
    if (ang > 0.0) area = area + ic.r * ic.r * (ang - sin(ang)) * 0.5;
  }
  return area;
}

// Geometry of one candidate pair (contact.rs:202-218).
struct PairGeom {
  Sph a, b, ic;
  V3 axis;       // unit a -> b
  V3 ic_dir;     // unit a -> intersection-circle centre
  V3 axis_unit;  // unit(axis), the hexagon's rotation axis (rotate normalizes again)
};

// contact.rs:202-218: false when the pair cannot have a contact.
VLTR_HD bool pair_setup(const Sph& a, const Sph& b, PairGeom* g) {
  if (!intersects(a, b) || contains(a, b) || contains(b, a)) return false;
  g->a = a;
  g->b = b;
  g->ic = intersection_circle(a, b);
  g->axis = unit(b.c - a.c);
  g->axis_unit = unit(g->axis);
  if (g->ic.r <= 0.0) return false;
  g->ic_dir = unit(g->ic.c - a.c);
  return true;
}

// Radical plane of a and neighbor c: a point, the unit normal a -> c, and the
// unit direction a -> that point (contact.rs:246-250). Depends on (a, c) only.
struct Plane {
  V3 pc, pn, pdir;
};
VLTR_HD Plane plane_of(const Sph& a, const Sph& c) {
  Plane p;
  p.pc = circle_center(a, c);
  p.pn = unit(c.c - a.c);
  p.pdir = unit(p.pc - a.c);
  return p;
}

// plane_of(a, c), pair_setup(a, c, g) and contains(c, a) from one evaluation
// of each quantity: the circle's centre is the plane point, its axis the
// plane normal and its direction from a the plane direction, and the squared
// centre distance serves the intersection and containment tests (the same
// expressions, so the same values). Returns pair_setup's result; g's fields
// are set in any case. *pc_dist2 = norm2(pl->pc - a.c), the plane point's
// squared distance from a's centre, which the unit direction also needs.
VLTR_HD bool plane_pair_setup(const Sph& a, const Sph& c, Plane* pl, PairGeom* g, bool* c_contains_a,
                              double* pc_dist2) {
  const V3 cv = c.c - a.c;
  const double cm2 = norm2(cv);
  const double cm = sqrt(cm2);
  const bool apart = !(cm < kEps);
  const double cg = circle_cos(a, c, cm);
  pl->pc = apart ? a.c + cv * (a.r * cg / cm) : a.c;
  pl->pn = cv / cm;
  const V3 pv = pl->pc - a.c;
  *pc_dist2 = norm2(pv);
  pl->pdir = pv / sqrt(*pc_dist2);
  g->a = a;
  g->b = c;
  g->ic = Sph{pl->pc, apart ? a.r * sqrt(fmax(1.0 - cg * cg, 0.0)) : 0.0};
  g->axis = pl->pn;
  g->axis_unit = unit(pl->pn);
  g->ic_dir = pl->pdir;
  const double s = a.r + c.r, dac = a.r - c.r, dca = c.r - a.r;
  const bool meets = f_lt(cm2, s * s);
  const bool a_has_c = f_ge(a.r, c.r) && f_le(cm2, dac * dac);
  *c_contains_a = f_ge(c.r, a.r) && f_le(cm2, dca * dca);
  return meets && !a_has_c && !*c_contains_a && g->ic.r > 0.0;
}

// What neighbor c does to the contact of a pair (contact.rs:224-276). The
// answer depends on (a, b, c) only, never on the contour built so far: the
// pair has no contact as soon as one neighbor kills it, and otherwise its
// contour is cut by the kCut neighbors in sorted order.
enum NbClass : int { kSkip = 0, kCut = 1, kKill = 2 };

// l / sqrt(t) >= r, the quotient and root rounded in double precision as the
// reference computes them (contact.rs:261-262). On the device a
// single-precision estimate decides first when it is far from r: l and t
// rounded to single (2^-24 each), a hardware reciprocal square root (1 ulp,
// 2^-23) and one product (2^-24) put it within 5e-7 of the exact quotient
// (relative), so outside a 1e-5 margin around r its answer is the exact one.
// Near r, and for l or t outside [1e-30, 1e15], the exact quotient decides.
// `need` false skips the test (returns false): lanes whose answer is not
// used never take the exact path.
VLTR_HD bool quotient_reaches(double l, double t, double r, bool need = true) {
#if defined(__HIP_DEVICE_COMPILE__)
  const bool fast = t >= 1e-30 && l >= 1e-30 && l <= 1e15;
  const double xa = static_cast<double>(static_cast<float>(l) * __builtin_amdgcn_rsqf(static_cast<float>(t)));
  const bool above = xa > r * (1.0 + 1e-5), below = xa < r * (1.0 - 1e-5);
  bool reach = fast & above;
  if (need & !(fast & (above | below))) reach = l / sqrt(t) >= r;
  return reach;
#else
  return need && l / sqrt(t) >= r;
#endif
}

// What neighbor c does to the contact of a pair (contact.rs:224-276), in
// the reference's order of tests: no intersection with the circle or with b
// skips, c containing a or b kills, and otherwise the circle centre's side
// of the plane decides when the plane is parallel or misses the circle;
// the rest cut. Every predicate is computed and the answer selected, so a
// wavefront classifying one neighbor per lane runs no branches (only the
// rare exact quotient). intersects(g.b, c) and contains(c, g.b) share the
// squared distance of the two centres (norm2(c - b) and norm2(b - c) are
// equal bit for bit), and the side of the circle centre is the sign of the
// distance l is taken from.
VLTR_HD int classify(const PairGeom& g, const Sph& c, const Plane& pl, bool c_contains_a) {
  const bool meets_ic = intersects(g.ic, c);
  const double dbc = norm2(c.c - g.b.c);
  const double sbc = g.b.r + c.r;
  const bool meets_b = f_lt(dbc, sbc * sbc);
  const double dr = c.r - g.b.r;
  const bool kill = c_contains_a || (f_ge(c.r, g.b.r) && f_le(dbc, dr * dr));
  const double cosv = dot(g.ic_dir, pl.pdir);
  const double sd = dot(pl.pn, g.ic.c - pl.pc);
  const bool pos = f_gt(sd, 0.0);       // halfspace(pl.pc, pl.pn, g.ic.c) > 0
  const bool nonneg = !f_lt(sd, 0.0);   // halfspace(...) >= 0
  const bool par = fabs(cosv) >= 1.0;
  const bool open = meets_ic & meets_b & !kill & !par;
  const bool reach = quotient_reaches(fabs(sd), fma(cosv, -cosv, 1.0), g.ic.r, open);
  int cl = kCut;
  cl = reach ? (nonneg ? kKill : kSkip) : cl;
  cl = par ? (pos ? kKill : kSkip) : cl;
  cl = kill ? kKill : cl;
  return (meets_ic & meets_b) ? cl : kSkip;
}

// One cutting neighbor (contact.rs:278-297): the first one lays the hexagon
// of circumradius 1.19 r around the circle (contact.rs:455-469), later ones
// only cut when some contour point lies beyond their plane point. Returns
// kContact to go on, kNone when the contour vanished, kOverflow.
template <class S>
VLTR_HD int apply_cut(S& ct, int& n, bool& initialized, const PairGeom& g, V3 pc, V3 pn, int32_t a_id, int32_t cid,
                      const double* hex_s, const double* hex_c) {
  if (!initialized) {
    const V3 first = (any_normal(g.axis) * g.ic.r) * kHexagonScale;
    ct.set(0, g.ic.c + first, a_id, a_id);
    for (int h = 0; h < 5; ++h) ct.set(h + 1, g.ic.c + rotate_unit(g.axis, g.axis_unit, hex_s[h], hex_c[h], first), a_id, a_id);
    n = 6;
    initialized = true;
  } else if (!cuttable(ct, n, g.a.c, pc)) {
    return kContact;
  }
  if (mark_and_cut(ct, n, pc, pn, unit(pn), cid) == kOverflow) return kOverflow;
  return n == 0 ? kNone : kContact;
}

// contact.rs:303-340: uncut full circle, or restriction to the circle + area.
template <class S, class A>
VLTR_HD int finish_clip(S& ct, int n, bool initialized, const PairGeom& g, int32_t a_id, A& out, ClipOut* co) {
  co->ic = g.ic;
  co->axis = g.axis;
  co->m = 0;
  co->sum_arc = 0.0;
  co->area = 0.0;
  co->bary = V3{0.0, 0.0, 0.0};
  if (!initialized) {
    co->bary = g.ic.c;
    co->sum_arc = kTau;
    co->area = g.ic.r * g.ic.r * kPi;
  } else {
    int m = 0;
    double s = 0.0;
    if (restrict_to_circle(ct, n, out, m, g.ic, g.axis, a_id, &s) == kOverflow) return kOverflow;
    co->sum_arc = s;
    co->m = m;
    if (m > 0) co->area = contour_area(out, m, g.ic, &co->bary);
  }
  return co->area <= 0.0 ? kNone : kContact;
}

// contact.rs:196-341 (clip_contact). `nbl` is a's sorted neighbor list.
template <class S, class A>
VLTR_HD int clip(const double* sph4, int32_t a_id, int32_t b_id, const int32_t* nbl, int nbn, const double* hex_s,
                 const double* hex_c, S& ct, A& out, ClipOut* co) {
  PairGeom g;
  if (!pair_setup(load_sph(sph4, a_id), load_sph(sph4, b_id), &g)) return kNone;
  int n = 0;
  bool initialized = false;
  for (int k = 0; k < nbn; ++k) {
    const int32_t cid = nbl[k];
    if (cid == b_id) continue;
    const Sph c = load_sph(sph4, cid);
    if (!intersects(g.ic, c) || !intersects(g.b, c)) continue;
    const Plane pl = plane_of(g.a, c);
    const int cls = classify(g, c, pl, contains(c, g.a));
    if (cls == kKill) return kNone;
    if (cls == kSkip) continue;
    const int st = apply_cut(ct, n, initialized, g, pl.pc, pl.pn, a_id, cid, hex_s, hex_c);
    if (st != kContact) return st;
  }
  return finish_clip(ct, n, initialized, g, a_id, out, co);
}

// Clip driven by a precomputed set of cutting neighbors: bit k of cut_mask
// (128 bits, slot order) marks slot k of a's list as a kCut neighbor; the
// pair is known to have no kKill neighbor.
template <class S, class A>
VLTR_HD int clip_masked(const double* sph4, int32_t a_id, int32_t b_id, const int32_t* nbl, uint64_t mlo, uint64_t mhi,
                        const double* hex_s, const double* hex_c, S& ct, A& out, ClipOut* co) {
  PairGeom g;
  if (!pair_setup(load_sph(sph4, a_id), load_sph(sph4, b_id), &g)) return kNone;
  int n = 0;
  bool initialized = false;
  for (int half = 0; half < 2; ++half) {
    uint64_t m = half == 0 ? mlo : mhi;
    while (m) {
      const int k = half * 64 + ctz64(m);
      m &= m - 1;
      const int32_t cid = nbl[k];
      const Sph c = load_sph(sph4, cid);
      const V3 pc = circle_center(g.a, c);
      const V3 pn = unit(c.c - g.a.c);
      const int st = apply_cut(ct, n, initialized, g, pc, pn, a_id, cid, hex_s, hex_c);
      if (st != kContact) return st;
    }
  }
  return finish_clip(ct, n, initialized, g, a_id, out, co);
}

// contact.rs:411-444
template <class A>
VLTR_HD bool is_central(V3 center, const A& ct, int m, V3 bary) {
  const double fe = 1e-10;
  for (int k = 0; k < m; ++k)
    if ((ct.ang[k] - fe) > kPi) return true;
  for (int i = 0; i < m; ++i) {
    const V3 pi = ct.p(i);
    const V3 pj = ct.p((i + 1) % m);
    const V3 edge = pj - pi;
    const double len = norm(edge);
    if (len < fe) continue;
    const V3 u = edge / len;
    const V3 foot = pi + u * dot(u, bary - pi);
    const V3 nb = bary - foot;
    if (dot(nb, center - pi) < 0.0) return false;
  }
  return true;
}

// contact.rs:797-848
template <class A>
VLTR_HD double solid_angle(const Sph& a, const Sph& b, const Sph& ic, const A& ct, int m) {
  double turn = 0.0;
  if (m == 0) {
    turn = kTau * norm(ic.c - a.c) / a.r;
  } else {
    for (int i = 0; i < m; ++i) {
      const int pi = i > 0 ? i - 1 : m - 1;
      const int ni = (i + 1) % m;
      const V3 p0 = ct.p(pi), p1 = ct.p(i), p2 = ct.p(ni);
      const double a0 = ct.ang[pi], a1 = ct.ang[i];
      if (a0 > 0.0) {
        V3 d = cross(b.c - a.c, p1 - ic.c);
        const double t = dot(d, p0 - p1);
        if ((a0 < kPi && t < 0.0) || (a0 > kPi && t > 0.0)) d = -d;
        turn = turn + (kPi - min_dihedral(a.c, p1, p1 + d, p2));
      } else if (a1 > 0.0) {
        V3 d = cross(b.c - a.c, p1 - ic.c);
        const double t = dot(d, p2 - p1);
        if ((a1 < kPi && t < 0.0) || (a1 > kPi && t > 0.0)) d = -d;
        turn = turn + (kPi - min_dihedral(a.c, p1, p0, p1 + d));
        turn = turn + a1 * (norm(ic.c - a.c) / a.r);
      } else {
        turn = turn + (kPi - min_dihedral(a.c, p1, p0, p2));
      }
    }
  }
  double solid = kTau - turn;
  const V3 ic_a = ic.c - a.c;
  const V3 ic_b = ic.c - b.c;
  if (dot(ic_a, ic_b) > 0.0 && norm2(a.c - ic.c) < norm2(b.c - ic.c)) solid = -solid;
  return solid;
}

// Record of a pair without contact.
VLTR_HD PairRec empty_record() {
  PairRec rec;
  rec.area = rec.arc_length = rec.solid_a = rec.solid_b = rec.pyr_a = rec.pyr_b = 0.0;
  rec.flags = 0u;
  rec.pad = 0u;
  rec.spare = 0u;
  return rec;
}

// contact.rs:347-404: record of a clipped pair (descriptor summary, or area,
// arc and centrality only when mode == 1).
template <class A>
VLTR_HD PairRec record_of(int st, const ClipOut& co, const Sph& a, const Sph& b, const A& out, int mode) {
  PairRec rec = empty_record();
  if (st != kContact) {
    if (st == kOverflow) rec.flags = kRecOverflow;
    return rec;
  }
  rec.area = co.area;
  rec.arc_length = co.sum_arc * co.ic.r;
  if (is_central(co.ic.c, out, co.m, co.bary)) rec.flags |= kRecCentral;
  rec.flags |= kRecValid;
  if (mode == 0) {
    rec.solid_a = solid_angle(a, b, co.ic, out, co.m);
    rec.solid_b = solid_angle(b, a, co.ic, out, co.m);
    const double sa = rec.solid_a < 0.0 ? -1.0 : 1.0;
    const double sb = rec.solid_b < 0.0 ? -1.0 : 1.0;
    rec.pyr_a = norm(co.ic.c - a.c) * co.area / 3.0 * sa;
    rec.pyr_b = norm(co.ic.c - b.c) * co.area / 3.0 * sb;
  }
  return rec;
}

template <class S, class A>
VLTR_HD PairRec make_record(const double* sph4, int32_t a_id, int32_t b_id, const int32_t* nbl, int nbn, int mode,
                            const double* hex_s, const double* hex_c, S& ct, A& out, int* status) {
  ClipOut co;
  const int st = clip(sph4, a_id, b_id, nbl, nbn, hex_s, hex_c, ct, out, &co);
  *status = st;
  return record_of(st, co, load_sph(sph4, a_id), load_sph(sph4, b_id), out, mode);
}

template <class S, class A>
VLTR_HD PairRec make_record_masked(const double* sph4, int32_t a_id, int32_t b_id, const int32_t* nbl, uint64_t mlo,
                                   uint64_t mhi, int mode, const double* hex_s, const double* hex_c, S& ct, A& out,
                                   int* status) {
  ClipOut co;
  const int st = clip_masked(sph4, a_id, b_id, nbl, mlo, mhi, hex_s, hex_c, ct, out, &co);
  *status = st;
  return record_of(st, co, load_sph(sph4, a_id), load_sph(sph4, b_id), out, mode);
}

// Record of an uncut pair (no kCut and no kKill neighbor): the full circle.
VLTR_HD PairRec record_uncut(const PairGeom& g, int mode) {
  ClipOut co;
  co.ic = g.ic;
  co.axis = g.axis;
  co.m = 0;
  co.bary = g.ic.c;
  co.sum_arc = kTau;
  co.area = g.ic.r * g.ic.r * kPi;
  struct NoContour {
    double ang[1];
    VLTR_HD V3 p(int) const { return V3{0.0, 0.0, 0.0}; }
  } none{{0.0}};
  return record_of(co.area <= 0.0 ? kNone : kContact, co, g.a, g.b, none, mode);
}

}  // namespace vltr
