// Host build of the kernels' contact-clipping code (voronota-ltr_amd/csrc/
// contour.h), so CPU tests can check it pair by pair against the oracle.
// Test infrastructure: compiled by __graft_entry__.build() into
// tests/native/libclip_host.so; the product library does not use it.
#include <cmath>
#include <cstdint>
#include <vector>

#include "contour.h"

namespace {
void hexagon(double* s, double* c) {
  double rot = vltr::kFracPi3;
  for (int k = 0; k < 5 && rot < vltr::kTau; ++k) {
    s[k] = std::sin(rot / 2.0);
    c[k] = std::cos(rot / 2.0);
    rot += vltr::kFracPi3;
  }
}
}  // namespace

extern "C" {
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
}
