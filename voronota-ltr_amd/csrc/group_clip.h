// Contact clipping of one candidate pair by a group of 16 cooperating lanes
// (same results as clip_masked + record_of in contour.h, which restate
// src/contact.rs:196-848).
//
// Lane j holds point j of the contour while it is cut (at most 16 points);
// cuts become lane permutations and the two plane/segment intersections of a
// cut are computed side by side. The contour restricted to the circle (at
// most 32 points) is compacted into a per-group shared buffer, and every
// per-point term (arc angles, triangle areas, dihedral turns) is computed by
// one lane each; sums are then taken in the reference's sequential order, so
// the record is bit-identical to the one-lane code. Contours that would
// outgrow the group, and the degenerate cuts that leave fewer than three
// points, return kOverflow: the caller redoes them with the one-lane code.
//
// G is the lane group of G::kSize lanes (8 or 16): lane(), ballot(bool) ->
// kSize-bit mask, shfl(v, src) for double and int32, sync() (shared-buffer
// writes visible to the group). Contours hold at most kSize points while cut
// and 2 kSize after the restriction.
// All control flow that reaches a ballot, shfl or sync is uniform over the
// group. The kernels implement G with wavefront intrinsics; the CPU tests
// run the same code on 16 lock-stepped host fibers.
#pragma once

#include "contour.h"

namespace vltr {

constexpr int kGroup = 16;
constexpr int kGroupR = 2 * kGroup;

// Restricted contour of one group (CAP = twice the group size): points,
// (left, right) ids, arc angles.
template <int CAP>
struct GroupLdsT {
  static constexpr int kCap = CAP;
  double x[CAP], y[CAP], z[CAP], ang[CAP];
  int32_t l[CAP], r[CAP];
};
using GroupLds = GroupLdsT<kGroupR>;

struct GPt {
  V3 p;
  int32_t l, r;
};

// Where a clip also records the cut topology of its contour (for the adjoint):
// the store record reserved for the pair (written when the pair has a
// contact, its index then put in the pair's slot), and the spheres (to tell
// the two candidate circle points apart).
constexpr int kTopoCap = kGroupR;
// Cut topology of one contact: m points (m = 0: uncut full circle, m < 0: no
// topology, so no gradient), ids l[k], r[k] of the pieces before and after
// point k (a's own id = the circle arc), bit k of roots = which of the two
// candidate circle points (sign of (x - a) . ((b - a) x (c - a))).
struct ContactTopo {
  int32_t m;
  uint32_t roots;
  int32_t l[kTopoCap];
  int32_t r[kTopoCap];
};
constexpr int32_t kTopoNone = -1;
constexpr int32_t kTopoTooLong = -2;

// Topology slots of the contacts: an index into the run's store, or one of
// these values, which need no store record. (Pairs without a contact are not
// read.)
constexpr uint32_t kTopoSlotUncut = 0xffffffffu;    // m = 0: the full circle
constexpr uint32_t kTopoSlotTooLong = 0xfffffffeu;  // more than kTopoCap points: no gradient
constexpr uint32_t kTopoSlotFull = 0xfffffffdu;     // the store region was full (the host reruns)
VLTR_HD bool topo_slot_stored(uint32_t s) { return s < kTopoSlotFull; }

// Restricted contour of one contact as computed (network output,
// contact.rs:120-177): m points (0: uncut full circle, kTopoTooLong: more than
// kTopoCap), their (left, right) ids, coordinates and arc angles.
struct ContactNet {
  int32_t m;
  int32_t pad;
  double r_ic;  // radius of the intersection circle
  int32_t l[kTopoCap], r[kTopoCap];
  double x[kTopoCap], y[kTopoCap], z[kTopoCap], ang[kTopoCap];
};

struct TopoOut {
  ContactTopo* topo;  // the store, or null
  uint32_t* slot;     // the pair's slot
  uint32_t record;    // the store record reserved for the pair, or kTopoSlotFull
  const double* sph4;
  ContactNet* net;    // the pair's contour, or null
};

// Radical plane of a and a cutting neighbor (plane_of), plus unit(pn) and the
// neighbor's id.
struct CutPlane {
  V3 pc, pn, nn;
  int32_t cid;
};

VLTR_HD int popc32(uint32_t x) { return __builtin_popcount(x); }
VLTR_HD int ctz32(uint32_t x) { return __builtin_ctz(x); }
VLTR_HD int clz32(uint32_t x) { return __builtin_clz(x); }

template <class G>
VLTR_HD V3 gshfl3(G& g, V3 v, int src) {
  return V3{g.shfl(v.x, src), g.shfl(v.y, src), g.shfl(v.z, src)};
}
template <class G>
VLTR_HD GPt gshfl(G& g, const GPt& v, int src) {
  return GPt{gshfl3(g, v.p, src), g.shfl(v.l, src), g.shfl(v.r, src)};
}

// One cut of the contour by a neighbor's radical plane (contact.rs:484-627;
// one-lane form: mark_and_cut).
// The point's offset from the plane point serves both the side test (along
// pn: halfspace(pc, pn, p) >= 0) and the plane-segment intersections (along
// the unit normal nn): each point's distance along nn is computed once, by
// its lane, and travels with the point (plane_segment_d).
template <class G>
VLTR_HD int group_cut(G& g, GPt& c, int& n, const CutPlane& cp) {
  constexpr int GS = G::kSize;
  const int j = g.lane();
  const V3 v = c.p - cp.pc;
  const bool o = j < n && !f_lt(dot(cp.pn, v), 0.0);
  double dn = dot(cp.nn, v);
  if (o) {
    c.l = cp.cid;
    c.r = cp.cid;
  }
  const uint32_t M = g.ballot(o);
  const int outs = popc32(M);
  if (outs == 0) return kContact;
  if (outs >= n) {
    n = 0;
    return kNone;
  }
  if (n < 3) return kContact;
  // The marked run is [start, end]; when it wraps around point 0 it ends
  // before the first unmarked point and starts after the last one. The run
  // is replaced by two points, start and start + 1 (one point out: one more
  // point, the others shift up; several: the points after the run shift
  // down), or, for a wrapped run, the unmarked points move to the front and
  // the two new points are the last and the first. Every lane computes the
  // group's indices, so the cases are selects rather than branches.
  const uint32_t full = (1u << n) - 1u;
  const int first = ctz32(M), last = 31 - clz32(M);
  const bool wraps = first == 0 && last == n - 1;
  int start = wraps ? 32 - clz32(~M & full) : first;
  const int stop = wraps ? ctz32(~M) - 1 : last;
  const bool run = !wraps;  // start <= stop
  const int d = stop - (start + 1);  // -1 when a single point is out
  const int nn = run ? n - d : start + 1 - stop;
  const int src = run ? (j > start ? j + d : j) : j + stop;
  if (nn > GS || nn < 3) return kOverflow;
  start = run ? start : nn - 1;
  const int end = run ? start + 1 : 0;
  c = gshfl(g, c, src & (GS - 1));
  dn = g.shfl(dn, src & (GS - 1));
  n = nn;
  const int il = start > 0 ? start - 1 : n - 1;
  const int ir = end + 1 < n ? end + 1 : 0;
  // with n >= 3, il != end and ir != start: both intersections use the
  // points as they were before either is written
  const int os = j == start ? il : ir;
  const GPt other = gshfl(g, c, os);
  const V3 x = plane_segment_d(c.p, other.p, dn, g.shfl(dn, os));
  const bool is_s = j == start, is_e = j == end;
  const int32_t nl = is_s ? other.r : (is_e ? c.r : c.l);
  const int32_t nr = is_s ? c.l : (is_e ? other.l : c.r);
  c.l = nl;
  c.r = nr;
  if (is_s || is_e) c.p = x;
  const V3 pe = gshfl3(g, c.p, end);
  if (g.ballot(j == start && !f_gt(norm2(pe - c.p), 0.0)) != 0u) {
    const int32_t re = g.shfl(c.r, end);
    if (j == start) c.r = re;
    c = gshfl(g, c, (j >= end ? j + 1 : j) & (GS - 1));
    --n;
  }
  return kContact;
}

template <class S>
VLTR_HD void lds_put(S* s, int k, V3 p, int32_t l, int32_t r) {
  s->x[k] = p.x;
  s->y[k] = p.y;
  s->z[k] = p.z;
  s->l[k] = l;
  s->r[k] = r;
}
template <class S>
VLTR_HD V3 lds_p(const S* s, int k) {
  return V3{s->x[k], s->y[k], s->z[k]};
}

// Restriction to the circle (contact.rs:631-769; one-lane form:
// restrict_to_circle). Lane j handles segment (j, j+1). On return the
// restricted contour is s[0..m) with its arc angles; m = 0 means no contour.
template <class G, class S>
VLTR_HD int group_restrict(G& g, S* s, const GPt& c, int n, const Sph& ic, V3 axis, int32_t a_id, int* m_out,
                           double* sum) {
  constexpr int GS = G::kSize;
  const int j = g.lane();
  const double rr = ic.r * ic.r;
  const bool valid = j < n;
  const bool in = valid && norm2(c.p - ic.c) <= rr;
  const uint32_t Min = g.ballot(in);
  const int n_in = popc32(Min);
  *sum = 0.0;
  *m_out = 0;
  if (n_in == n) {
    if (valid) {
      lds_put(s, j, c.p, c.l, c.r);
      s->ang[j] = 0.0;
    }
    g.sync();
    *m_out = n;
    return kContact;
  }
  const int ni = j + 1 < n ? j + 1 : 0;
  const int pj = j > 0 ? j - 1 : n - 1;
  const GPt nx = gshfl(g, c, ni & (GS - 1));
  const bool o1 = !in;
  const bool o2 = !((Min >> (ni & 31)) & 1u);
  V3 ins1 = c.p, ins2 = c.p;
  int32_t i1l = 0, i1r = 0, i2l = 0, i2r = 0;
  int nins = 0;
  bool mod_next = false, mod_self = false;
  if (valid && (o1 || o2)) {
    // o1 && o2: both crossings from the foot of the perpendicular (when it
    // lies inside the circle); o1 only: from the next point (inside) towards
    // this one; o2 only: from this point (inside) towards the next one.
    bool go = true;
    V3 from = o1 ? nx.p : c.p;
    V3 to = o1 ? c.p : nx.p;
    if (o1 && o2) {
      V3 mp;
      go = project_inside_segment(ic.c, c.p, nx.p, &mp) && norm2(mp - ic.c) <= rr;
      from = mp;
    }
    V3 ip1 = c.p, ip2 = c.p;
    bool ok1 = false, ok2 = false;
    if (go) {
#pragma unroll 1
      for (int h = 0; h < 2; ++h) {
        if (h == 1 && !(o1 && o2 && ok1)) break;
        V3 ip;
        const bool ok = segment_circle(ic, from, h == 0 ? to : nx.p, &ip);
        if (h == 0) {
          ok1 = ok;
          ip1 = ip;
        } else {
          ok2 = ok;
          ip2 = ip;
        }
      }
    }
    if (o1 && o2) {
      if (go && ok1 && ok2) {
        ins1 = ip1;
        i1l = a_id;
        i1r = c.r;
        ins2 = ip2;
        i2l = nx.l;
        i2r = a_id;
        nins = 2;
      }
    } else if (ok1) {
      ins1 = ip1;
      i1l = o1 ? a_id : nx.l;
      i1r = o1 ? c.r : a_id;
      nins = 1;
    } else if (o1) {
      mod_next = true;  // point j+1 gets ids (a, r_j)
    } else {
      mod_self = true;  // point j gets ids (l_{j+1}, a)
    }
  }
  // Final ids of an inside point: segment j-1 runs before segment j, except
  // that segment n-1 runs after segment 0 and so wins for point 0.
  const bool prev_mod = (g.ballot(mod_next) >> (pj & 31)) & 1u;
  const int32_t prev_r = g.shfl(c.r, pj & (GS - 1));
  int32_t fl = c.l, fr = c.r;
  if (j == 0) {
    if (mod_self) {
      fl = nx.l;
      fr = a_id;
    }
    if (prev_mod) {
      fl = a_id;
      fr = prev_r;
    }
  } else {
    if (prev_mod) {
      fl = a_id;
      fr = prev_r;
    }
    if (mod_self) {
      fl = nx.l;
      fr = a_id;
    }
  }
  // Output order: point j (if inside), then segment j's crossings.
  const int e = valid ? (in ? 1 : 0) + nins : 0;
  const uint32_t B0 = g.ballot((e & 1) != 0), B1 = g.ballot((e & 2) != 0);
  const uint32_t below = (1u << j) - 1u;
  const int off = popc32(B0 & below) + 2 * popc32(B1 & below);
  const int m = popc32(B0) + 2 * popc32(B1);
  if (m > S::kCap) return kOverflow;
  if (m - n_in == 0 || m < 2) return kContact;
  if (valid) {
    int o = off;
    if (in) lds_put(s, o++, c.p, fl, fr);
    if (nins >= 1) lds_put(s, o++, ins1, i1l, i1r);
    if (nins == 2) lds_put(s, o, ins2, i2l, i2r);
  }
  g.sync();
  const V3 ctrl = ic.c + axis;
#pragma unroll 1
  for (int h = 0; h < 2; ++h) {
    const int t = j + h * GS;
    if (t < m) {
      const int nt = t + 1 < m ? t + 1 : 0;
      double ang = 0.0;
      if (s->r[t] == a_id && s->l[nt] == a_id) ang = directed_angle(ic.c, lds_p(s, t), lds_p(s, nt), ctrl);
      s->ang[t] = ang;
    }
  }
  g.sync();
  // non-arc entries hold +0.0, which leaves the running sum unchanged
  double sm = 0.0;
  for (int k = 0; k < m; ++k) sm = sm + s->ang[k];
  *sum = sm;
  if (f_ge(sm, kTau) || (m > 2 && f_eq(sm, kTau))) {
    *sum = kTau;
    return kContact;
  }
  *m_out = m;
  return kContact;
}

// Sequential sum of per-entry terms held by the lanes (entry t in lane t % 16,
// slot t / 16), in entry order: acc = (acc + t1[k]) + t2[k].
template <class G>
VLTR_HD double group_sum2(G& g, double t1a, double t1b, double t2a, double t2b, int m) {
  constexpr int GS = G::kSize;
  double acc = 0.0;
  for (int k = 0; k < m; ++k) {
    const int src = k & (GS - 1);
    const double v1 = k < GS ? g.shfl(t1a, src) : g.shfl(t1b, src);
    const double v2 = k < GS ? g.shfl(t2a, src) : g.shfl(t2b, src);
    acc = acc + v1;
    acc = acc + v2;
  }
  return acc;
}

// Per-vertex turn of the Gauss-Bonnet sum (contact.rs:797-837) of the sphere
// centred at oc against the other sphere's centre qc: first and second addend
// (second is +0.0 when absent). ratio = norm(icc - oc) / (the sphere's
// radius), the same for every vertex. The three cases differ only in the
// dihedral's two outer points, so min_dihedral is evaluated once.
template <class S>
VLTR_HD void turn_terms(V3 oc, V3 qc, V3 icc, double ratio, const S* s, int i, int m, double* t1, double* t2) {
  const int pi = i > 0 ? i - 1 : m - 1;
  const int ni = i + 1 < m ? i + 1 : 0;
  const V3 p0 = lds_p(s, pi), p1 = lds_p(s, i), p2 = lds_p(s, ni);
  const double a0 = s->ang[pi], a1 = s->ang[i];
  *t2 = 0.0;
  V3 b1 = p0, b2 = p2;
  if (a0 > 0.0 || a1 > 0.0) {
    const bool before = a0 > 0.0;
    const double an = before ? a0 : a1;
    V3 d = cross(qc - oc, p1 - icc);
    const double t = dot(d, (before ? p0 : p2) - p1);
    if ((an < kPi && t < 0.0) || (an > kPi && t > 0.0)) d = -d;
    if (before) {
      b1 = p1 + d;
    } else {
      b2 = p1 + d;
      *t2 = a1 * ratio;
    }
  }
  *t1 = kPi - min_dihedral(oc, p1, b1, b2);
}

// Solid angle of the sphere centred at oc from its turn sum, with its sign
// (contact.rs:836-847).
VLTR_HD double signed_solid(double turn, V3 oc, V3 qc, V3 icc) {
  const double solid = kTau - turn;
  return dot(icc - oc, icc - qc) > 0.0 && norm2(oc - icc) < norm2(qc - icc) ? -solid : solid;
}

template <class G, class S>
VLTR_HD double group_solid_angle(G& g, const Sph& a, const Sph& b, const Sph& ic, const S* s, int m) {
  constexpr int GS = G::kSize;
  const int j = g.lane();
  const double ratio = norm(ic.c - a.c) / a.r;
  // entries j and j + GS of the lane (registers, not an indexed array)
  double t1a = 0.0, t2a = 0.0, t1b = 0.0, t2b = 0.0;
#pragma unroll 1
  for (int h = 0; h < 2; ++h) {
    const int t = j + h * GS;
    double u1 = 0.0, u2 = 0.0;
    if (t < m) turn_terms(a.c, b.c, ic.c, ratio, s, t, m, &u1, &u2);
    if (h == 0) {
      t1a = u1;
      t2a = u2;
    } else {
      t1b = u1;
      t2b = u2;
    }
  }
  const double turn = group_sum2(g, t1a, t1b, t2a, t2b, m);
  return signed_solid(turn, a.c, b.c, ic.c);
}

// Area and centrality terms of restricted-contour entry t (contact.rs:773-793,
// 411-444).
template <class S>
VLTR_HD void area_terms(const S* s, int t, int m, const Sph& ic, V3 bary, double* tri, double* seg, bool* reflex,
                        bool* fail) {
  const int nt = t + 1 < m ? t + 1 : 0;
  const V3 p = lds_p(s, t), q = lds_p(s, nt);
  *tri = triangle_area(bary, p, q);
  const double ang = s->ang[t];
  *seg = ang > 0.0 ? ic.r * ic.r * (ang - sin(ang)) * 0.5 : 0.0;
  const double fe = 1e-10;
  *reflex = (ang - fe) > kPi;
  const V3 edge = q - p;
  const double len = norm(edge);
  *fail = false;
  if (!(len < fe)) {
    const V3 u = edge / len;
    const V3 foot = p + u * dot(u, bary - p);
    const V3 nb = bary - foot;
    *fail = dot(nb, ic.c - p) < 0.0;
  }
}

VLTR_HD bool circle_root(V3 x, V3 a, V3 b, V3 c) { return dot(x - a, cross(b - a, c - a)) > 0.0; }

// Writes the topology of the restricted contour s[0..m).
template <class G, class S>
VLTR_HD void write_topology(G& g, const S* s, const PairGeom& pg, int32_t a_id, int m, const TopoOut& tout) {
  constexpr int GS = G::kSize;
  const int j = g.lane();
  const uint32_t slot = tout.record;
  if (slot == kTopoSlotFull) {
    if (j == 0) *tout.slot = kTopoSlotFull;
    return;
  }
  ContactTopo* topo = tout.topo + slot;
  bool root[2] = {false, false};
  for (int h = 0; h < 2; ++h) {
    const int t = j + h * GS;
    if (t >= m) continue;
    const int32_t l = s->l[t], r = s->r[t];
    topo->l[t] = l;
    topo->r[t] = r;
    if (l == a_id || r == a_id) {
      const int32_t c = l == a_id ? r : l;
      root[h] = circle_root(lds_p(s, t), pg.a.c, pg.b.c, load_sph(tout.sph4, c).c);
    }
  }
  const uint32_t r0 = g.ballot(root[0]), r1 = g.ballot(root[1]);
  if (j == 0) {
    topo->m = m;
    topo->roots = r0 | (r1 << GS);
    *tout.slot = slot;
  }
}

// The whole clip of a pair known to have no killing neighbor and a non-empty
// cutting set: bit k of the cutting mask = slot k of a's sorted neighbor list,
// the mask given as nw 64-bit words (words(w) = bits 64 w .. 64 w + 63).
// f.plane(k) returns slot k's CutPlane and f.thr(k) the squared distance from
// a's centre to its plane point. Every lane returns the same record.
template <class G, class S, class Words, class Fetch>
VLTR_HD PairRec group_clip_words(G& g, S* s, const PairGeom& pg, int32_t a_id, int nw, Words&& words, Fetch&& f,
                                 const double* hex_s, const double* hex_c, int mode, int* status,
                                 const TopoOut* tout = nullptr, int* defer_m = nullptr) {
  constexpr int GS = G::kSize;
  const int j = g.lane();
  GPt c{pg.ic.c, a_id, a_id};
  int n = 0;
  double d2 = 0.0;  // squared distance of this lane's point from a's centre
  *status = kNone;
  bool any = false;
  for (int w = 0; w < nw; ++w) any = any || words(w) != 0ull;
  if (any) {
    // contact.rs:455-469: the first cutting plane lays the hexagon of
    // circumradius 1.19 r, which that plane then cuts
    const V3 first = (any_normal(pg.axis) * pg.ic.r) * kHexagonScale;
    if (j == 0) {
      c.p = pg.ic.c + first;
    } else if (j < 6) {
      c.p = pg.ic.c + rotate_unit(pg.axis, pg.axis_unit, hex_s[j - 1], hex_c[j - 1], first);
    }
    n = 6;
  }
  bool fresh = true;
  for (int w = 0; w < nw; ++w) {
    uint64_t mm = words(w);
    while (mm) {
      const int k = w * 64 + ctz64(mm);
      mm &= mm - 1;
      // contact.rs:472-481: a later plane cuts only when some point is at
      // least as far from a's centre as the plane point
      if (!fresh && g.ballot(j < n && d2 >= f.thr(k)) == 0u) continue;
      fresh = false;
      const CutPlane cp = f.plane(k);
      const int st = group_cut(g, c, n, cp);
      if (st != kContact) {
        *status = st;
        return empty_record();
      }
      d2 = norm2(c.p - pg.a.c);
    }
  }
  int m = 0;
  double sum_arc = 0.0;
  if (group_restrict(g, s, c, n, pg.ic, pg.axis, a_id, &m, &sum_arc) == kOverflow) {
    *status = kOverflow;
    return empty_record();
  }
  if (m == 0) return empty_record();
  // barycenter, area and centrality
  V3 sum{0.0, 0.0, 0.0};
  for (int k = 0; k < m; ++k) sum = sum + lds_p(s, k);
  const V3 bary = sum / static_cast<double>(m);
  double tri_a = 0.0, seg_a = 0.0, tri_b = 0.0, seg_b = 0.0;
  bool reflex = false, fail = false;
#pragma unroll 1
  for (int h = 0; h < 2; ++h) {
    const int t = j + h * GS;
    if (t < m) {
      double tr, sg;
      bool rf, fl;
      area_terms(s, t, m, pg.ic, bary, &tr, &sg, &rf, &fl);
      reflex = reflex || rf;
      fail = fail || fl;
      if (h == 0) {
        tri_a = tr;
        seg_a = sg;
      } else {
        tri_b = tr;
        seg_b = sg;
      }
    }
  }
  const double area = group_sum2(g, tri_a, tri_b, seg_a, seg_b, m);
  const bool central = g.ballot(reflex) != 0u || g.ballot(fail) == 0u;
  if (area <= 0.0) return empty_record();
  *status = kContact;
  if (tout != nullptr && tout->topo != nullptr) write_topology(g, s, pg, a_id, m, *tout);
  if (tout != nullptr && tout->net != nullptr) {
    for (int h = 0; h < 2; ++h) {
      const int t = j + h * GS;
      if (t >= m) continue;
      ContactNet* nt = tout->net;
      nt->l[t] = s->l[t];
      nt->r[t] = s->r[t];
      nt->x[t] = s->x[t];
      nt->y[t] = s->y[t];
      nt->z[t] = s->z[t];
      nt->ang[t] = s->ang[t];
    }
    if (j == 0) {
      tout->net->m = m;
      tout->net->r_ic = pg.ic.r;
    }
  }
  PairRec rec = empty_record();
  rec.area = area;
  rec.arc_length = sum_arc * pg.ic.r;
  rec.flags = kRecValid | (central ? kRecCentral : 0u);
  if (defer_m != nullptr) {
    *defer_m = m;  // the caller computes the solid angles from s[0..m)
  } else if (mode == 0) {
    rec.solid_a = group_solid_angle(g, pg.a, pg.b, pg.ic, s, m);
    rec.solid_b = group_solid_angle(g, pg.b, pg.a, pg.ic, s, m);
    const double sa = rec.solid_a < 0.0 ? -1.0 : 1.0;
    const double sb = rec.solid_b < 0.0 ? -1.0 : 1.0;
    rec.pyr_a = norm(pg.ic.c - pg.a.c) * area / 3.0 * sa;
    rec.pyr_b = norm(pg.ic.c - pg.b.c) * area / 3.0 * sb;
  }
  return rec;
}

// Cutting mask of at most 128 slots as two words.
struct MaskPair {
  uint64_t lo, hi;
  VLTR_HD uint64_t operator()(int w) const { return w == 0 ? lo : hi; }
};

template <class G, class S, class Fetch>
VLTR_HD PairRec group_clip(G& g, S* s, const PairGeom& pg, int32_t a_id, uint64_t mlo, uint64_t mhi, Fetch&& f,
                           const double* hex_s, const double* hex_c, int mode, int* status,
                           const TopoOut* tout = nullptr, int* defer_m = nullptr) {
  return group_clip_words(g, s, pg, a_id, mhi != 0ull ? 2 : 1, MaskPair{mlo, mhi}, f, hex_s, hex_c, mode, status,
                          tout, defer_m);
}

}  // namespace vltr
