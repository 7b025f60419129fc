// Host build of the kernels' contact-clipping code (voronota-ltr_amd/csrc/
// contour.h), so CPU tests can check it pair by pair against the oracle.
// Test infrastructure: compiled by __graft_entry__.build() into
// tests/native/libclip_host.so; the product library does not use it.
#include <ucontext.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <functional>

#include "adjoint.h"
#include "contour.h"
#include "group_clip.h"
#include "scan_total.h"

namespace {
void hexagon(double* s, double* c) {
  double rot = vltr::kFracPi3;
  for (int k = 0; k < 5 && rot < vltr::kTau; ++k) {
    s[k] = std::sin(rot / 2.0);
    c[k] = std::cos(rot / 2.0);
    rot += vltr::kFracPi3;
  }
}
// The 16-lane group of group_clip.h on the host: one fiber per lane, run in
// lock step. Every collective (ballot, shfl, sync) stores the lane's value
// and yields; the scheduler resumes lanes 0..15 in turn, so when a lane
// resumes, all 16 have contributed. Values alternate between two slot banks,
// so a lane can write the next collective before the others read this one.
struct FiberGroup {
  static constexpr int W = vltr::kGroup;  // at most W lanes; `w` run
  int w = W;
  ucontext_t sched;
  ucontext_t fib[W];
  std::vector<char> stack[W];
  uint64_t slot[2][W];
  int round[W];
  bool done[W];
  FiberGroup() {
    for (int k = 0; k < W; ++k) stack[k].resize(1 << 18);
  }
};

template <int GS>
struct HostLaneT {
  static constexpr int kSize = GS;
  FiberGroup* f;
  int ln;
  int lane() const { return ln; }
  void yield() { swapcontext(&f->fib[ln], &f->sched); }
  uint64_t exchange(uint64_t v, int src) {
    const int b = f->round[ln]++ & 1;
    f->slot[b][ln] = v;
    yield();
    return f->slot[b][src & (GS - 1)];
  }
  uint32_t ballot(bool p) {
    const int b = f->round[ln]++ & 1;
    f->slot[b][ln] = p ? 1u : 0u;
    yield();
    uint32_t m = 0;
    for (int k = 0; k < GS; ++k) m |= static_cast<uint32_t>(f->slot[b][k] & 1u) << k;
    return m;
  }
  double shfl(double v, int src) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    u = exchange(u, src);
    double r;
    std::memcpy(&r, &u, 8);
    return r;
  }
  int32_t shfl(int32_t v, int src) { return static_cast<int32_t>(exchange(static_cast<uint32_t>(v), src)); }
  void sync() {
    f->round[ln]++;
    yield();
  }
};
using HostLane = HostLaneT<vltr::kGroup>;

struct GroupTask {
  const double* sph4;
  const int32_t* nbl;
  vltr::PairGeom g;
  int32_t a_id;
  uint64_t mlo, mhi;
  int mode;
  const double *hs, *hc;
  vltr::GroupLds lds;
  vltr::PairRec rec[FiberGroup::W];
  int st[FiberGroup::W];
  FiberGroup* f;
};
GroupTask* g_task = nullptr;

void lane_main(int ln) {
  GroupTask& t = *g_task;
  HostLane lane{t.f, ln};
  const vltr::Sph A = t.g.a;
  struct {
    GroupTask& t;
    vltr::Sph A;
    double thr(int k) const { return vltr::norm2(vltr::plane_of(A, vltr::load_sph(t.sph4, t.nbl[k])).pc - A.c); }
    vltr::CutPlane plane(int k) const {
      const vltr::Plane pl = vltr::plane_of(A, vltr::load_sph(t.sph4, t.nbl[k]));
      return vltr::CutPlane{pl.pc, pl.pn, vltr::unit(pl.pn), t.nbl[k]};
    }
  } fetch{t, A};
  t.rec[ln] = vltr::group_clip(lane, &t.lds, t.g, t.a_id, t.mlo, t.mhi, fetch, t.hs, t.hc, t.mode, &t.st[ln]);
  t.f->done[ln] = true;
}

// Runs group_clip on 16 fibers; returns lane 0's record, checks all agree.
bool run_group(GroupTask& t, vltr::PairRec* out, int* st) {
  static FiberGroup fg;
  t.f = &fg;
  g_task = &t;
  std::memset(&t.lds, 0xff, sizeof(t.lds));
  for (int k = 0; k < FiberGroup::W; ++k) {
    fg.round[k] = 0;
    fg.done[k] = false;
    getcontext(&fg.fib[k]);
    fg.fib[k].uc_stack.ss_sp = fg.stack[k].data();
    fg.fib[k].uc_stack.ss_size = fg.stack[k].size();
    fg.fib[k].uc_link = &fg.sched;
    makecontext(&fg.fib[k], reinterpret_cast<void (*)()>(lane_main), 1, k);
  }
  for (;;) {
    int running = 0;
    for (int k = 0; k < FiberGroup::W; ++k) {
      if (fg.done[k]) continue;
      swapcontext(&fg.sched, &fg.fib[k]);
      ++running;
    }
    if (running == 0) break;
  }
  bool agree = true;
  for (int k = 1; k < FiberGroup::W; ++k) {
    if (t.st[k] != t.st[0] || std::memcmp(&t.rec[k], &t.rec[0], sizeof(vltr::PairRec)) != 0) agree = false;
    if (fg.round[k] != fg.round[0]) agree = false;
  }
  *out = t.rec[0];
  *st = t.st[0];
  return agree;
}
// Runs fn(lane) on GS lock-stepped fibers.
template <int GS>
struct FiberRun {
  static std::function<void(HostLaneT<GS>&)>* fn;
  static FiberGroup* fg;
  static void lane_main(int ln) {
    HostLaneT<GS> lane{fg, ln};
    (*fn)(lane);
    fg->done[ln] = true;
  }
};
template <int GS>
std::function<void(HostLaneT<GS>&)>* FiberRun<GS>::fn = nullptr;
template <int GS>
FiberGroup* FiberRun<GS>::fg = nullptr;

template <int GS>
void run_fibers_n(std::function<void(HostLaneT<GS>&)> fn) {
  static FiberGroup fg;
  FiberRun<GS>::fn = &fn;
  FiberRun<GS>::fg = &fg;
  for (int k = 0; k < GS; ++k) {
    fg.round[k] = 0;
    fg.done[k] = false;
    getcontext(&fg.fib[k]);
    fg.fib[k].uc_stack.ss_sp = fg.stack[k].data();
    fg.fib[k].uc_stack.ss_size = fg.stack[k].size();
    fg.fib[k].uc_link = &fg.sched;
    makecontext(&fg.fib[k], reinterpret_cast<void (*)()>(&FiberRun<GS>::lane_main), 1, k);
  }
  for (;;) {
    int running = 0;
    for (int k = 0; k < GS; ++k) {
      if (fg.done[k]) continue;
      swapcontext(&fg.sched, &fg.fib[k]);
      ++running;
    }
    if (running == 0) break;
  }
}
void run_fibers(std::function<void(HostLane&)> fn) { run_fibers_n<vltr::kGroup>(fn); }

// One masked pair through group_clip on GS fibers; false when the lanes
// disagree (a bug).
template <int GS>
bool group_record(const double* sph4, const int32_t* nbl, const vltr::PairGeom& g, int32_t a, uint64_t m,
                  const double* hs, const double* hc, int mode, vltr::PairRec* out, int* status) {
  static vltr::GroupLdsT<2 * GS> lds;
  vltr::PairRec rec[GS];
  int st[GS];
  const vltr::Sph A = g.a;
  run_fibers_n<GS>([&](HostLaneT<GS>& lane) {
    struct {
      const double* sph4;
      const int32_t* nbl;
      vltr::Sph A;
      double thr(int k) const { return vltr::norm2(vltr::plane_of(A, vltr::load_sph(sph4, nbl[k])).pc - A.c); }
      vltr::CutPlane plane(int k) const {
        const vltr::Plane pl = vltr::plane_of(A, vltr::load_sph(sph4, nbl[k]));
        return vltr::CutPlane{pl.pc, pl.pn, vltr::unit(pl.pn), nbl[k]};
      }
    } fetch{sph4, nbl, A};
    rec[lane.ln] = vltr::group_clip(lane, &lds, g, a, m, 0ull, fetch, hs, hc, mode, &st[lane.ln]);
  });
  bool agree = true;
  for (int k = 1; k < GS; ++k)
    if (st[k] != st[0] || std::memcmp(&rec[k], &rec[0], sizeof(vltr::PairRec)) != 0) agree = false;
  *out = rec[0];
  *status = st[0];
  return agree;
}

struct HostAcc {
  double* grad;
  int64_t n;
  void add(int64_t id, vltr::V3 v) {
    const int64_t i = id % n;
    grad[3 * i] += v.x;
    grad[3 * i + 1] += v.y;
    grad[3 * i + 2] += v.z;
  }
};
}  // namespace

extern "C" {

// Gradient of E = sum_contacts W[t_a][t_b] area + sum_cells (sas + volume)
// with respect to the centres, by the kernels' adjoint code (adjoint.h) run on
// 16 host fibers: forward records and cut topologies from the group clip,
// cell seeds from the per-ball sums, then the backward of every contact.
// Non-periodic; lists up to 64 neighbors. Returns the number of contacts
// without topology (their gradient is missing), or -1.
int64_t adjoint_host(const double* sph4, int64_t n, const int64_t* nb_off, const int32_t* nb_idx, const int32_t* pairs,
                     int64_t P, const int32_t* types, const double* W16, double* energy, double* grad) {
  double hs[5], hc[5];
  hexagon(hs, hc);
  std::vector<vltr::PairRec> recs(P);
  std::vector<vltr::ContactTopo> topo(P);
  int64_t missing = 0;
  GroupTask* task = new GroupTask;
  for (int64_t t = 0; t < P; ++t) {
    const int32_t a = pairs[2 * t], b = pairs[2 * t + 1];
    const int32_t* nbl = nb_idx + nb_off[a];
    const int nbn = static_cast<int>(nb_off[a + 1] - nb_off[a]);
    recs[t] = vltr::empty_record();
    topo[t].m = vltr::kTopoNone;
    if (nbn > 64) return -1;
    vltr::PairGeom g;
    const vltr::Sph A = vltr::load_sph(sph4, a);
    if (!vltr::pair_setup(A, vltr::load_sph(sph4, b), &g)) continue;
    bool kill = false;
    uint64_t m = 0;
    for (int k = 0; k < nbn; ++k) {
      if (nbl[k] == b) continue;
      const vltr::Sph C = vltr::load_sph(sph4, nbl[k]);
      const int cl = vltr::classify(g, C, vltr::plane_of(A, C), vltr::contains(C, A));
      if (cl == vltr::kKill) kill = true;
      if (cl == vltr::kCut) m |= 1ull << k;
    }
    if (kill) continue;
    if (m == 0) {
      recs[t] = vltr::record_uncut(g, 0);
      topo[t].m = 0;
      continue;
    }
    vltr::PairRec rec[FiberGroup::W];
    int st[FiberGroup::W];
    uint32_t slot = 0u;  // a store of one record: topo[t]
    vltr::TopoOut tout{&topo[t], &slot, 0u, sph4, nullptr};
    run_fibers([&](HostLane& lane) {
      struct {
        const double* sph4;
        const int32_t* nbl;
        vltr::Sph A;
        double thr(int k) const { return vltr::norm2(vltr::plane_of(A, vltr::load_sph(sph4, nbl[k])).pc - A.c); }
        vltr::CutPlane plane(int k) const {
          const vltr::Plane pl = vltr::plane_of(A, vltr::load_sph(sph4, nbl[k]));
          return vltr::CutPlane{pl.pc, pl.pn, vltr::unit(pl.pn), nbl[k]};
        }
      } fetch{sph4, nbl, A};
      rec[lane.ln] = vltr::group_clip(lane, &task->lds, g, a, m, 0ull, fetch, hs, hc, 0, &st[lane.ln], &tout);
    });
    recs[t] = rec[0];
    if (st[0] == vltr::kOverflow) {
      ++missing;
      topo[t].m = vltr::kTopoTooLong;
    }
  }
  delete task;
  // cells (types.rs:587-663) and their seeds
  std::vector<double> sa_pos(n, 0.0), sa_neg(n, 0.0), pv_pos(n, 0.0), pv_neg(n, 0.0), arc(n, 0.0);
  std::vector<int> added(n, 0);
  double e = 0.0;
  for (int64_t t = 0; t < P; ++t) {
    const vltr::PairRec& r = recs[t];
    if (!(r.flags & vltr::kRecValid) || !(r.area > 0.0)) continue;
    const int32_t a = pairs[2 * t], b = pairs[2 * t + 1];
    e += W16[4 * types[a] + types[b]] * r.area;
    for (int side = 0; side < 2; ++side) {
      const int32_t i = side == 0 ? a : b;
      const double sa = side == 0 ? r.solid_a : r.solid_b;
      const double pv = side == 0 ? r.pyr_a : r.pyr_b;
      added[i] = 1;
      arc[i] += r.arc_length;
      sa_pos[i] += std::fmax(sa, 0.0);
      sa_neg[i] -= std::fmin(sa, 0.0);
      pv_pos[i] += std::fmax(pv, 0.0);
      pv_neg[i] -= std::fmin(pv, 0.0);
    }
  }
  std::vector<double> seed(n, 0.0);
  for (int64_t i = 0; i < n; ++i) {
    if (!added[i]) continue;
    const double R = sph4[4 * i + 3];
    double sas = 0.0, vol;
    bool standard = false;
    if (arc[i] > 0.0 && std::fabs(sa_pos[i] - sa_neg[i]) > 1e-10) {
      standard = true;
      if (sa_pos[i] > sa_neg[i])
        sas = std::fma(4.0, vltr::kPi, -std::fmax(sa_pos[i] - sa_neg[i], 0.0)) * R * R;
      else
        sas = std::fmax(sa_neg[i] - sa_pos[i], 0.0) * R * R;
      vol = (sas * R / 3.0) + pv_pos[i] - pv_neg[i];
      if (vol > 4.0 / 3.0 * vltr::kPi * R * R * R) {
        standard = false;
        sas = 0.0;
        vol = pv_pos[i] - pv_neg[i];
      }
    } else {
      vol = pv_pos[i] - pv_neg[i];
    }
    e += sas + vol;
    seed[i] = standard ? -R * R * (1.0 + R / 3.0) : 0.0;
  }
  *energy = e;
  for (int64_t k = 0; k < 3 * n; ++k) grad[k] = 0.0;
  HostAcc acc{grad, n};
  static vltr::AdjLds lds;
  for (int64_t t = 0; t < P; ++t) {
    const vltr::PairRec& r = recs[t];
    if (!(r.flags & vltr::kRecValid) || !(r.area > 0.0) || topo[t].m < 0) continue;
    const int32_t a = pairs[2 * t], b = pairs[2 * t + 1];
    const double w = W16[4 * types[a] + types[b]];
    const double sga = r.solid_a < 0.0 ? -1.0 : 1.0, sgb = r.solid_b < 0.0 ? -1.0 : 1.0;
    run_fibers([&](HostLane& lane) {
      vltr::group_contact_adjoint(lane, &lds, sph4, a, b, topo[t], topo[t].m, w, seed[a], seed[b], sga, sgb, acc);
    });
  }
  return missing;
}

int64_t xcd_remap_host(int64_t b, int64_t g) { return vltr::xcd_remap(b, g); }

// The kernels' two-stage path on the host: classify every neighbor of every
// pair (k_prefilter), then cut with the resulting mask (k_cut_alive). Same
// outputs as clip_host_records; flags bit3 marks pairs decided by the
// prefilter alone (killed or uncut).
int clip_host_records_staged(const double* sph4, const int64_t* nb_off, const int32_t* nb_idx, const int32_t* pairs,
                             int64_t P, int mode, double* out, uint32_t* flags) {
  double hs[5], hc[5];
  hexagon(hs, hc);
  for (int64_t t = 0; t < P; ++t) {
    const int32_t a = pairs[2 * t], b = pairs[2 * t + 1];
    const int32_t* nbl = nb_idx + nb_off[a];
    const int nbn = static_cast<int>(nb_off[a + 1] - nb_off[a]);
    vltr::PairRec r = vltr::empty_record();
    uint32_t extra = 8u;
    vltr::PairGeom g;
    const vltr::Sph A = vltr::load_sph(sph4, a);
    if (nbn <= 128 && vltr::pair_setup(A, vltr::load_sph(sph4, b), &g)) {
      bool kill = false;
      uint64_t m[2] = {0, 0};
      for (int k = 0; k < nbn; ++k) {
        if (nbl[k] == b) continue;
        const vltr::Sph C = vltr::load_sph(sph4, nbl[k]);
        const int cl = vltr::classify(g, C, vltr::plane_of(A, C), vltr::contains(C, A));
        if (cl == vltr::kKill) kill = true;
        if (cl == vltr::kCut) m[k / 64] |= 1ull << (k % 64);
      }
      if (!kill && (m[0] | m[1]) == 0) {
        r = vltr::record_uncut(g, mode);
      } else if (!kill) {
        int st = 0;
        vltr::LocalCut<vltr::kMaxC> ct;
        vltr::LocalArc<vltr::kMaxR> arc;
        r = vltr::make_record_masked(sph4, a, b, nbl, m[0], m[1], mode, hs, hc, ct, arc, &st);
        extra = 0u;
        if (st == 2) {
          std::vector<double> ws(5 * 1024 + 6 * 2048);
          vltr::GlobalCut gc{1024, ws.data(), ws.data() + 1024, ws.data() + 2048,
                             reinterpret_cast<int32_t*>(ws.data() + 3072),
                             reinterpret_cast<int32_t*>(ws.data() + 4096)};
          double* w2 = ws.data() + 5 * 1024;
          vltr::GlobalArc ga{2048, w2, w2 + 2048, w2 + 4096, w2 + 6144, reinterpret_cast<int32_t*>(w2 + 8192),
                             reinterpret_cast<int32_t*>(w2 + 10240)};
          r = vltr::make_record_masked(sph4, a, b, nbl, m[0], m[1], mode, hs, hc, gc, ga, &st);
          extra = 4u;
        }
      }
    } else if (nbn > 128) {
      // lists longer than the mask take the full scan, as on the device
      int st = 0;
      vltr::LocalCut<vltr::kMaxC> ct;
      vltr::LocalArc<vltr::kMaxR> arc;
      r = vltr::make_record(sph4, a, b, nbl, nbn, mode, hs, hc, ct, arc, &st);
      extra = 0u;
    }
    out[6 * t + 0] = r.area;
    out[6 * t + 1] = r.arc_length;
    out[6 * t + 2] = r.solid_a;
    out[6 * t + 3] = r.solid_b;
    out[6 * t + 4] = r.pyr_a;
    out[6 * t + 5] = r.pyr_b;
    flags[t] = (r.flags & 3u) | extra;
  }
  return 0;
}

// The fused kernel's path on the host: classification as in the staged path,
// then the 16-lane group clip (group_clip.h) for pairs with cutting
// neighbors; lists longer than 64 and group overflows take the one-lane code
// with the large contour store, as on the device. flags: bit0 valid, bit1
// central, bit3 decided by classification, bit4 clipped by the group, bit5
// redone by the one-lane code, bit6 the 16 lanes disagreed (a bug).
int clip_host_records_group(const double* sph4, const int64_t* nb_off, const int32_t* nb_idx, const int32_t* pairs,
                            int64_t P, int mode, double* out, uint32_t* flags, int group_size) {
  double hs[5], hc[5];
  hexagon(hs, hc);
  std::vector<double> ws(5 * 1024 + 6 * 2048);
  double* w2 = ws.data() + 5 * 1024;
  vltr::GlobalCut gc{1024, ws.data(), ws.data() + 1024, ws.data() + 2048,
                     reinterpret_cast<int32_t*>(ws.data() + 3072), reinterpret_cast<int32_t*>(ws.data() + 4096)};
  vltr::GlobalArc ga{2048, w2, w2 + 2048, w2 + 4096, w2 + 6144, reinterpret_cast<int32_t*>(w2 + 8192),
                     reinterpret_cast<int32_t*>(w2 + 10240)};
  for (int64_t t = 0; t < P; ++t) {
    const int32_t a = pairs[2 * t], b = pairs[2 * t + 1];
    const int32_t* nbl = nb_idx + nb_off[a];
    const int nbn = static_cast<int>(nb_off[a + 1] - nb_off[a]);
    vltr::PairRec r = vltr::empty_record();
    uint32_t extra = 8u;
    vltr::PairGeom g;
    const vltr::Sph A = vltr::load_sph(sph4, a);
    int st = 0;
    if (nbn > 64) {
      r = vltr::make_record(sph4, a, b, nbl, nbn, mode, hs, hc, gc, ga, &st);
      extra = 32u;
    } else if (vltr::pair_setup(A, vltr::load_sph(sph4, b), &g)) {
      bool kill = false;
      uint64_t m = 0;
      for (int k = 0; k < nbn; ++k) {
        if (nbl[k] == b) continue;
        const vltr::Sph C = vltr::load_sph(sph4, nbl[k]);
        const int cl = vltr::classify(g, C, vltr::plane_of(A, C), vltr::contains(C, A));
        if (cl == vltr::kKill) kill = true;
        if (cl == vltr::kCut) m |= 1ull << k;
      }
      if (!kill && m == 0) {
        r = vltr::record_uncut(g, mode);
      } else if (!kill) {
        const bool agree = group_size == 8 ? group_record<8>(sph4, nbl, g, a, m, hs, hc, mode, &r, &st)
                                           : group_record<16>(sph4, nbl, g, a, m, hs, hc, mode, &r, &st);
        extra = 16u | (agree ? 0u : 64u);
        if (st == vltr::kOverflow) {
          r = vltr::make_record(sph4, a, b, nbl, nbn, mode, hs, hc, gc, ga, &st);
          extra |= 32u;
        }
      }
    }
    out[6 * t + 0] = r.area;
    out[6 * t + 1] = r.arc_length;
    out[6 * t + 2] = r.solid_a;
    out[6 * t + 3] = r.solid_b;
    out[6 * t + 4] = r.pyr_a;
    out[6 * t + 5] = r.pyr_b;
    flags[t] = (r.flags & 3u) | extra;
  }
  return 0;
}

// sph4: n_pop (x, y, z, R); neighbor lists CSR (nb_off[n+1], nb_idx); pairs
// (a, b) x P. out: P x 6 doubles (area, arc, solid_a, solid_b, pyr_a, pyr_b),
// flags: P (bit0 valid, bit1 central, bit2 needed the large contour store).
int clip_host_records(const double* sph4, const int64_t* nb_off, const int32_t* nb_idx, const int32_t* pairs,
                      int64_t P, int mode, double* out, uint32_t* flags) {
  double hs[5], hc[5];
  hexagon(hs, hc);
  for (int64_t t = 0; t < P; ++t) {
    const int32_t a = pairs[2 * t], b = pairs[2 * t + 1];
    const int32_t* nbl = nb_idx + nb_off[a];
    const int nbn = static_cast<int>(nb_off[a + 1] - nb_off[a]);
    int st = 0;
    vltr::LocalCut<vltr::kMaxC> ct;
    vltr::LocalArc<vltr::kMaxR> arc;
    vltr::PairRec r = vltr::make_record(sph4, a, b, nbl, nbn, mode, hs, hc, ct, arc, &st);
    uint32_t extra = 0;
    if (st == 2) {
      std::vector<double> ws(5 * 1024 + 6 * 2048);
      vltr::GlobalCut gc{1024, ws.data(), ws.data() + 1024, ws.data() + 2048,
                         reinterpret_cast<int32_t*>(ws.data() + 3072), reinterpret_cast<int32_t*>(ws.data() + 4096)};
      double* w2 = ws.data() + 5 * 1024;
      vltr::GlobalArc ga{2048, w2, w2 + 2048, w2 + 4096, w2 + 6144, reinterpret_cast<int32_t*>(w2 + 8192),
                         reinterpret_cast<int32_t*>(w2 + 10240)};
      r = vltr::make_record(sph4, a, b, nbl, nbn, mode, hs, hc, gc, ga, &st);
      extra = 4u;
    }
    out[6 * t + 0] = r.area;
    out[6 * t + 1] = r.arc_length;
    out[6 * t + 2] = r.solid_a;
    out[6 * t + 3] = r.solid_b;
    out[6 * t + 4] = r.pyr_a;
    out[6 * t + 5] = r.pyr_b;
    flags[t] = (r.flags & 3u) | extra;
  }
  return 0;
}

// The pair total as the device scan forms it (launch_scan_u32 in k_grid.hip):
// 32-bit tile sums of 2048 counts; the last tile's block adds the tile sums
// before it in 64 bits, then its own into a ScanTotal; checked_pair_total on
// the (32-bit total, beyond-int32) report words. Returns that result (-1: the
// run is refused) and the 32-bit total the offsets would hold.
int64_t scan_pair_guard(const uint32_t* counts, int64_t n, uint32_t* total32) {
  const int64_t tile = 2048;
  std::vector<uint32_t> tiles;
  for (int64_t b = 0; b < n; b += tile) {
    uint32_t s = 0;
    for (int64_t i = b; i < n && i < b + tile; ++i) s += counts[i];
    tiles.push_back(s);
  }
  unsigned long long before = 0;
  for (size_t k = 0; k + 1 < tiles.size(); ++k) before += tiles[k];
  vltr::ScanTotal carry;
  carry.lo = static_cast<uint32_t>(before);
  carry.full = before;
  if (!tiles.empty()) carry.add(tiles.back());
  *total32 = carry.lo;
  return vltr::checked_pair_total(carry.lo, carry.beyond_int32());
}
}
