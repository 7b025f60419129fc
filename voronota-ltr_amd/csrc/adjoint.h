// Adjoint of the tessellation energy
//   E = sum_contacts W[t_a][t_b] * area + sum_cells (sas_area + volume)
// with respect to the ball centres (radii and probe fixed), holding the
// contour combinatorics of the forward pass fixed.
//
// The forward pass stores, per contact, only the cut topology of its contour
// restricted to the circle (contact.rs:631-769): for each contour point the
// ids of the two boundary pieces meeting there. A point between two radical
// planes (a|l), (a|r) is the triple point of those planes and the contact
// plane (a|b); a point between a plane (a|c) and the circle arc lies on
// sphere a, and one bit per point says which of the two candidates it is.
// The backward pass recomputes the points from the spheres (3x3 solves),
// re-evaluates area (contact.rs:773-793), solid angles (contact.rs:797-848)
// and pyramid volumes (contact.rs:372-380) as functions of them, and runs
// their vector-Jacobian products by hand; the points' own derivatives follow
// from implicit differentiation of their defining equations.
//
// With E = sum_contacts W[t_a][t_b] area + alpha sum_cells sas_area + beta
// sum_cells volume, the cell seeds are (types.rs:624-663): dE/d(solid angle)
// = -R^2 (alpha + beta R / 3) in the standard branch of compute_sas and 0 in
// the others; dE/d(pyramid volume) = beta always.
#pragma once

#include "group_clip.h"

namespace vltr {

// ---------------------------------------------------------------- VJPs ----
struct V3acc {
  double x = 0.0, y = 0.0, z = 0.0;
  VLTR_HD void add(V3 v) {
    x += v.x;
    y += v.y;
    z += v.z;
  }
  VLTR_HD V3 v() const { return V3{x, y, z}; }
};

// At the measure-zero points where a derivative below is unbounded (a zero
// vector normalised, a singular 3x3 system: tangent circle points, vertices
// where more than three planes meet) the contribution is taken as zero.

// u = v / |v|: returns v_bar for u_bar.
VLTR_HD V3 vjp_unit(V3 v, V3 ubar) {
  const double n = norm(v);
  if (!(n > 0.0)) return V3{0.0, 0.0, 0.0};
  const V3 u = v / n;
  return (ubar - u * dot(u, ubar)) / n;
}
// s = |v|
VLTR_HD V3 vjp_norm(V3 v, double sbar) {
  const double n = norm(v);
  if (!(n > 0.0)) return V3{0.0, 0.0, 0.0};
  return v * (sbar / n);
}
// acos(clamp(c)): d/dc (0 where the clamp is active)
VLTR_HD double dacos(double c) {
  if (!(c > -1.0 && c < 1.0)) return 0.0;
  return -1.0 / sqrt(1.0 - c * c);
}

// triangle_area(p, q, s) = |(q - p) x (s - p)| / 2
VLTR_HD void vjp_triangle(V3 p, V3 q, V3 s, double tbar, V3* pbar, V3* qbar, V3* sbar) {
  const V3 A = q - p, B = s - p;
  const V3 n = cross(A, B);
  const double nn = norm(n);
  if (!(nn > 0.0)) {
    *pbar = *qbar = *sbar = V3{0.0, 0.0, 0.0};
    return;
  }
  const V3 nbar = n * (tbar * 0.5 / nn);
  const V3 Abar = cross(B, nbar);
  const V3 Bbar = cross(nbar, A);
  *qbar = Abar;
  *sbar = Bbar;
  *pbar = -(Abar + Bbar);
}

// directed_angle(o, p, q, ctrl) (geometry.rs:157-176): value and the VJP
// (o, p, q) for theta_bar; the branch side is fixed.
VLTR_HD double directed_angle_fwd(V3 o, V3 p, V3 q, V3 ctrl, double* sgn) {
  const V3 v1 = unit(p - o), v2 = unit(q - o);
  const double ang = acos(clamp1(dot(v1, v2)));
  const V3 n = cross(v1, v2);
  const bool first = dot(ctrl - o, n) >= 0.0;
  *sgn = first ? 1.0 : -1.0;
  return first ? ang : kTau - ang;
}
VLTR_HD void vjp_directed_angle(V3 o, V3 p, V3 q, double sgn, double thbar, V3* obar, V3* pbar, V3* qbar) {
  const V3 a = p - o, b = q - o;
  const V3 v1 = unit(a), v2 = unit(b);
  const double cbar = thbar * sgn * dacos(dot(v1, v2));
  if (!(cbar != 0.0 && std::isfinite(cbar))) {
    *obar = *pbar = *qbar = V3{0.0, 0.0, 0.0};
    return;
  }
  const V3 abar = vjp_unit(a, v2 * cbar);
  const V3 bbar = vjp_unit(b, v1 * cbar);
  *pbar = abar;
  *qbar = bbar;
  *obar = -(abar + bbar);
}

// min_dihedral(o, p, b1, b2) (geometry.rs:283-293) and its VJP.
VLTR_HD void vjp_min_dihedral(V3 o, V3 p, V3 b1, V3 b2, double mbar, V3* obar, V3* pbar, V3* b1bar, V3* b2bar) {
  const V3 oa = unit(p - o);
  const V3 w1 = b1 - o, w2 = b2 - o;
  const V3 d1 = w1 - oa * dot(oa, w1);
  const V3 d2 = w2 - oa * dot(oa, w2);
  const V3 u1 = unit(d1), u2 = unit(d2);
  const double cbar = mbar * dacos(dot(u1, u2));
  if (!(cbar != 0.0 && std::isfinite(cbar))) {
    *obar = *pbar = *b1bar = *b2bar = V3{0.0, 0.0, 0.0};
    return;
  }
  const V3 d1bar = vjp_unit(d1, u2 * cbar);
  const V3 d2bar = vjp_unit(d2, u1 * cbar);
  // d = w - oa (oa . w): w_bar = d_bar - oa (oa . d_bar); oa_bar = -((oa . w) d_bar + (oa . d_bar) w)
  const V3 w1bar = d1bar - oa * dot(oa, d1bar);
  const V3 w2bar = d2bar - oa * dot(oa, d2bar);
  const V3 oabar = -(d1bar * dot(oa, w1) + w1 * dot(oa, d1bar) + d2bar * dot(oa, w2) + w2 * dot(oa, d2bar));
  const V3 pobar = vjp_unit(p - o, oabar);
  *b1bar = w1bar;
  *b2bar = w2bar;
  *pbar = pobar;
  *obar = -(w1bar + w2bar + pobar);
}

// ------------------------------------------------------- contour points ----
// Radical plane of sphere a and sphere c in coordinates relative to a:
// e . y = k with e = c - a, k = (|c - a|^2 + Ra^2 - Rc^2) / 2.
struct RelPlane {
  V3 e;
  double k;
};
VLTR_HD RelPlane rel_plane(const Sph& a, const Sph& c) {
  const V3 e = c.c - a.c;
  return RelPlane{e, 0.5 * (dot(e, e) + a.r * a.r - c.r * c.r)};
}

// Point of the contour from its topology: y relative to a. circle: the point
// of sphere a on planes p1, p2 with sign(y . (e1 x e2)) = root; else the
// triple point of p1, p2, p3.
VLTR_HD V3 circle_point(const Sph& a, const RelPlane& p1, const RelPlane& p2, bool root) {
  const double g11 = dot(p1.e, p1.e), g12 = dot(p1.e, p2.e), g22 = dot(p2.e, p2.e);
  const double det = g11 * g22 - g12 * g12;
  const double al = (p1.k * g22 - p2.k * g12) / det;
  const double be = (p2.k * g11 - p1.k * g12) / det;
  const V3 y0 = p1.e * al + p2.e * be;
  const V3 t = cross(p1.e, p2.e);
  const double tt = dot(t, t);
  const double rem = a.r * a.r - dot(y0, y0);
  const double ga = sqrt(rem > 0.0 ? rem / tt : 0.0);
  return y0 + t * (root ? ga : -ga);
}
VLTR_HD V3 triple_point(const RelPlane& p1, const RelPlane& p2, const RelPlane& p3) {
  const V3 c23 = cross(p2.e, p3.e), c31 = cross(p3.e, p1.e), c12 = cross(p1.e, p2.e);
  const double det = dot(p1.e, c23);
  return (c23 * p1.k + c31 * p2.k + c12 * p3.k) / det;
}
// lambda = M^-T ybar for M with rows r1, r2, r3.
VLTR_HD V3 solve_transposed(V3 r1, V3 r2, V3 r3, V3 ybar) {
  const V3 c23 = cross(r2, r3), c31 = cross(r3, r1), c12 = cross(r1, r2);
  const double det = dot(r1, c23);
  const V3 lam = V3{dot(c23, ybar), dot(c31, ybar), dot(c12, ybar)} / det;
  if (!(std::isfinite(lam.x) && std::isfinite(lam.y) && std::isfinite(lam.z))) return V3{0.0, 0.0, 0.0};
  return lam;
}

// ------------------------------------------------------ group backward ----
// Shared memory of one 16-lane group in the backward pass.
struct AdjLds {
  double x[kTopoCap], y[kTopoCap], z[kTopoCap];  // contour points
  double th[kTopoCap], ths[kTopoCap];            // arc angle (0: plane edge), its branch sign
  double px[kTopoCap], py[kTopoCap], pz[kTopoCap];  // entry t's adjoint of point t-1
  double nx[kTopoCap], ny[kTopoCap], nz[kTopoCap];  // entry t's adjoint of point t+1
  int32_t l[kTopoCap], r[kTopoCap];
};

template <class G>
VLTR_HD double gsum(G& g, double v) {
  for (int o = 1; o < G::kSize; o <<= 1) v += g.shfl(v, g.lane() ^ o);
  return v;
}
template <class G>
VLTR_HD V3 gsum3(G& g, V3 v) {
  return V3{gsum(g, v.x), gsum(g, v.y), gsum(g, v.z)};
}

// Gradient of one contact's share of the energy,
//   w * area + sum_{i = a, b} (g_i * solid_i + beta * pyramid_volume_i),
// added to acc (acc.add(sphere id, V3)). g_a, g_b: cell seeds; sgn_a, sgn_b:
// signs of the forward solid angles (they fix the pyramid volume signs).
// Adds v to sphere id unless it is not finite (a non-differentiable point).
template <class Acc>
VLTR_HD void add_finite(Acc& acc, int32_t id, V3 v) {
  if (std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z)) acc.add(id, v);
}

// topo: the contact's cut topology with m points (m = 0: the uncut circle,
// topo not read).
template <class G, class Acc>
VLTR_HD void group_contact_adjoint(G& g, AdjLds* s, const double* sph4, int32_t a_id, int32_t b_id,
                                   const ContactTopo& topo, int m, double w, double g_a, double g_b, double sgn_a,
                                   double sgn_b, Acc& acc, double beta = 1.0) {
  constexpr int GS = G::kSize;
  static_assert(2 * GS <= kTopoCap, "two contour points per lane");
  const int j = g.lane();
  const double own = j == 0 ? 1.0 : 0.0;  // uniform terms are added by lane 0 only
  const Sph A = load_sph(sph4, a_id), B = load_sph(sph4, b_id);
  const V3 e = B.c - A.c;
  const double d = norm(e);
  const V3 u = e / d;
  const double h = (d * d + A.r * A.r - B.r * B.r) / (2.0 * d);
  const V3 ic = A.c + u * h;
  const double r2 = A.r * A.r - h * h;
  const double r = sqrt(r2 > 0.0 ? r2 : 0.0);
  const V3 cen[2] = {A.c, B.c};
  const double Rr[2] = {A.r, B.r}, gs[2] = {g_a, g_b}, sg[2] = {sgn_a, sgn_b};
  double dist[2], turnbar[2];
  for (int i = 0; i < 2; ++i) {
    dist[i] = norm(ic - cen[i]);
    const bool flipped = dot(ic - cen[i], ic - cen[1 - i]) > 0.0 && norm2(cen[i] - ic) < norm2(cen[1 - i] - ic);
    // solid = f (2 pi - turn), f = -1 when flipped
    turnbar[i] = gs[i] * (flipped ? 1.0 : -1.0);
  }
  V3 cbar[2] = {V3{0.0, 0.0, 0.0}, V3{0.0, 0.0, 0.0}};
  V3 icbar{0.0, 0.0, 0.0};
  double rbar = 0.0;
  double distbar[2] = {0.0, 0.0};
  double area;
  if (m == 0) {
    // uncut full circle (contact.rs:303-308, 803-805)
    area = kPi * r * r;
    double Abar = w;
    for (int i = 0; i < 2; ++i) Abar += beta * dist[i] * sg[i] / 3.0;
    rbar += own * Abar * 2.0 * kPi * r;
    for (int i = 0; i < 2; ++i) distbar[i] += own * (beta * area * sg[i] / 3.0 + turnbar[i] * kTau / Rr[i]);
  } else {
    const RelPlane pb = rel_plane(A, B);
    for (int hh = 0; hh < 2; ++hh) {
      const int t = j + hh * GS;
      if (t >= m) continue;
      const int32_t l = topo.l[t], rr = topo.r[t];
      V3 y;
      if (l == a_id || rr == a_id) {
        const int32_t c = l == a_id ? rr : l;
        y = circle_point(A, pb, rel_plane(A, load_sph(sph4, c)), ((topo.roots >> t) & 1u) != 0u);
      } else {
        y = triple_point(pb, rel_plane(A, load_sph(sph4, l)), rel_plane(A, load_sph(sph4, rr)));
      }
      const V3 x = A.c + y;
      s->x[t] = x.x;
      s->y[t] = x.y;
      s->z[t] = x.z;
      s->l[t] = l;
      s->r[t] = rr;
    }
    g.sync();
    auto P = [&](int t) { return V3{s->x[t], s->y[t], s->z[t]}; };
    V3 bs{0.0, 0.0, 0.0};
    for (int hh = 0; hh < 2; ++hh) {
      const int t = j + hh * GS;
      if (t >= m) continue;
      const int nt = t + 1 < m ? t + 1 : 0;
      double th = 0.0, sgn = 1.0;
      if (s->r[t] == a_id && s->l[nt] == a_id) th = directed_angle_fwd(ic, P(t), P(nt), ic + u, &sgn);
      s->th[t] = th;
      s->ths[t] = sgn;
      bs = bs + P(t);
    }
    const V3 bary = gsum3(g, bs) / static_cast<double>(m);
    g.sync();
    double al = 0.0;
    for (int hh = 0; hh < 2; ++hh) {
      const int t = j + hh * GS;
      if (t >= m) continue;
      const int nt = t + 1 < m ? t + 1 : 0;
      al += triangle_area(bary, P(t), P(nt));
      const double th = s->th[t];
      if (th > 0.0) al += 0.5 * r * r * (th - sin(th));
    }
    area = gsum(g, al);
    double Abar = w;
    for (int i = 0; i < 2; ++i) {
      Abar += beta * dist[i] * sg[i] / 3.0;
      distbar[i] += own * beta * area * sg[i] / 3.0;
    }
    V3 self[2] = {V3{0.0, 0.0, 0.0}, V3{0.0, 0.0, 0.0}};
    V3 baryb{0.0, 0.0, 0.0};
    for (int hh = 0; hh < 2; ++hh) {
      const int t = j + hh * GS;
      if (t >= m) continue;
      const int pt = t > 0 ? t - 1 : m - 1;
      const int nt = t + 1 < m ? t + 1 : 0;
      const V3 p0 = P(pt), p1 = P(t), p2 = P(nt);
      const double a0 = s->th[pt], a1 = s->th[t];
      V3 x0b{0.0, 0.0, 0.0}, x1b{0.0, 0.0, 0.0}, x2b{0.0, 0.0, 0.0};
      V3 tb0, tb1, tb2;
      vjp_triangle(bary, p1, p2, Abar, &tb0, &tb1, &tb2);
      baryb = baryb + tb0;
      x1b = x1b + tb1;
      x2b = x2b + tb2;
      double thb = 0.0;
      if (a1 > 0.0) {
        thb += Abar * 0.5 * r * r * (1.0 - cos(a1));
        rbar += Abar * r * (a1 - sin(a1));
      }
      for (int i = 0; i < 2; ++i) {
        const V3 o = cen[i];
        V3 ob, p1b, b1b, b2b;
        if (a0 > 0.0 || a1 > 0.0) {
          const bool before = a0 > 0.0;
          const double an = before ? a0 : a1;
          const V3 ee = cen[1 - i] - o;
          const V3 q = p1 - ic;
          const V3 dd = cross(ee, q);
          const double tt = dot(dd, (before ? p0 : p2) - p1);
          const double sigma = ((an < kPi && tt < 0.0) || (an > kPi && tt > 0.0)) ? -1.0 : 1.0;
          const V3 dv = dd * sigma;
          const V3 B1 = before ? p1 + dv : p0;
          const V3 B2 = before ? p2 : p1 + dv;
          vjp_min_dihedral(o, p1, B1, B2, -turnbar[i], &ob, &p1b, &b1b, &b2b);
          V3 dvb;
          if (before) {
            x1b = x1b + b1b;
            dvb = b1b;
            x2b = x2b + b2b;
          } else {
            x0b = x0b + b1b;
            x1b = x1b + b2b;
            dvb = b2b;
            thb += turnbar[i] * dist[i] / Rr[i];
            distbar[i] += turnbar[i] * a1 / Rr[i];
          }
          const V3 eb = cross(q, dvb) * sigma;
          const V3 qb = cross(dvb, ee) * sigma;
          cbar[1 - i] = cbar[1 - i] + eb;
          cbar[i] = cbar[i] - eb;
          x1b = x1b + qb;
          icbar = icbar - qb;
        } else {
          vjp_min_dihedral(o, p1, p0, p2, -turnbar[i], &ob, &p1b, &b1b, &b2b);
          x0b = x0b + b1b;
          x2b = x2b + b2b;
        }
        cbar[i] = cbar[i] + ob;
        x1b = x1b + p1b;
      }
      if (a1 > 0.0) {
        V3 ob, pb2, qb2;
        vjp_directed_angle(ic, p1, p2, s->ths[t], thb, &ob, &pb2, &qb2);
        icbar = icbar + ob;
        x1b = x1b + pb2;
        x2b = x2b + qb2;
      }
      self[hh] = x1b;
      s->px[t] = x0b.x;
      s->py[t] = x0b.y;
      s->pz[t] = x0b.z;
      s->nx[t] = x2b.x;
      s->ny[t] = x2b.y;
      s->nz[t] = x2b.z;
    }
    const V3 bary_share = gsum3(g, baryb) / static_cast<double>(m);
    g.sync();
    // points -> spheres (implicit differentiation of the defining equations)
    for (int hh = 0; hh < 2; ++hh) {
      const int t = j + hh * GS;
      if (t >= m) continue;
      const int pt = t > 0 ? t - 1 : m - 1;
      const int nt = t + 1 < m ? t + 1 : 0;
      const V3 xb = self[hh] + V3{s->nx[pt], s->ny[pt], s->nz[pt]} + V3{s->px[nt], s->py[nt], s->pz[nt]} + bary_share;
      const V3 y = P(t) - A.c;
      cbar[0] = cbar[0] + xb;  // x = a + y
      const int32_t l = s->l[t], rr = s->r[t];
      if (l == a_id || rr == a_id) {
        const int32_t c = l == a_id ? rr : l;
        const V3 ec = load_sph(sph4, c).c - A.c;
        const V3 lam = solve_transposed(e, ec, y * 2.0, xb);
        const V3 gb = (e - y) * lam.x, gc = (ec - y) * lam.y;
        cbar[1] = cbar[1] + gb;
        cbar[0] = cbar[0] - gb - gc;
        add_finite(acc, c, gc);
      } else {
        const V3 el = load_sph(sph4, l).c - A.c, er = load_sph(sph4, rr).c - A.c;
        const V3 lam = solve_transposed(e, el, er, xb);
        const V3 gb = (e - y) * lam.x, gl = (el - y) * lam.y, gr = (er - y) * lam.z;
        cbar[1] = cbar[1] + gb;
        cbar[0] = cbar[0] - gb - gl - gr;
        add_finite(acc, l, gl);
        add_finite(acc, rr, gr);
      }
    }
  }
  // dist_i = |ic - c_i|
  for (int i = 0; i < 2; ++i) {
    if (distbar[i] != 0.0) {
      const V3 v = vjp_norm(ic - cen[i], distbar[i]);
      icbar = icbar + v;
      cbar[i] = cbar[i] - v;
    }
  }
  // intersection circle: ic = a + u h, r = sqrt(Ra^2 - h^2), h = (d^2 + Ra^2 - Rb^2) / (2 d)
  double hbar = dot(icbar, u);
  if (r > 0.0) hbar -= rbar * h / r;
  const V3 ubar = icbar * h;
  const double dbar = hbar * (0.5 - (A.r * A.r - B.r * B.r) / (2.0 * d * d));
  const V3 ebar = vjp_unit(e, ubar) + u * dbar;
  cbar[0] = cbar[0] + icbar - ebar;
  cbar[1] = cbar[1] + ebar;
  const V3 ga_tot = gsum3(g, cbar[0]);
  const V3 gb_tot = gsum3(g, cbar[1]);
  if (j == 0) {
    add_finite(acc, a_id, ga_tot);
    add_finite(acc, b_id, gb_tot);
  }
}

}  // namespace vltr
