// Contact construction (replaces src/contact.rs): per candidate pair (a, b),
// the intersection circle of the two probe-expanded balls is cut by the
// radical planes of a's neighbors in their sorted order, restricted to the
// circle, and reduced to area, arc length, centrality, the two solid angles
// and the two pyramid volumes.
//
// One lane per pair. The contour is a fixed-capacity array; a pair whose
// contour would outgrow it is flagged and redone by k_cut_slow, whose contour
// lives in a global workspace sized for the longest neighbor list. Both run
// the same templated code, so they produce identical results.
#include "contour.h"
#include "vltr_internal.h"

namespace vltr {

// sin/cos of half the five accumulated hexagon rotation angles
// (contact.rs:455-469), computed once on the host with its libm.
__constant__ double c_hex_s[5];
__constant__ double c_hex_c[5];

namespace {

double h_hex_s[5], h_hex_c[5];

__global__ __launch_bounds__(256) void k_cut(Pipeline p) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < p.n_pairs; t += stride) {
    const int32_t a = p.pair_a[t], b = p.pair_b[t];
    const int32_t* nbl = p.nb + static_cast<int64_t>(a) * p.nb_cap;
    LocalCut<kMaxC> ct;
    LocalArc<kMaxR> out;
    int st;
    const PairRec rec = make_record(reinterpret_cast<const double*>(p.sph), a, b, nbl, p.nb_cnt[a], p.mode, c_hex_s, c_hex_c, ct, out, &st);
    p.rec[t] = rec;
    if (st == kOverflow) {
      const uint32_t k = atomicAdd(p.overflow_cnt, 1u);
      p.overflow_list[k] = static_cast<int32_t>(t);
      atomicOr(p.status, kStatContourOverflow);
    }
  }
}

constexpr int kWave = 64;

// Prefilter: one wavefront per ball a, one lane per neighbor slot (two slots
// per lane for lists of up to 128). Each lane keeps its neighbor's sphere and
// radical plane with a in registers and classifies it for every pair (a, b)
// of a (contact.rs:224-276). A pair with a killing neighbor gets an empty
// record; a pair with no cutting neighbor is the uncut full circle and is
// finished here; the rest go to the alive list with their cutting mask.
__global__ __launch_bounds__(256) void k_prefilter(Pipeline p) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t q = xcd_remap(blockIdx.x, gridDim.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (q >= p.n) return;
  const int64_t a = p.periodic ? q : static_cast<int64_t>(p.sorted_id[q]);
  const int64_t base = p.pair_cnt[a];
  const int64_t npairs = static_cast<int64_t>(p.pair_cnt[a + 1]) - base;
  if (npairs <= 0) return;
  const int cnt = p.nb_cnt[a];
  const int32_t* nbl = p.nb + a * static_cast<int64_t>(p.nb_cap);
  const double* sph4 = reinterpret_cast<const double*>(p.sph);
  if (cnt > 2 * kWave) {
    for (int64_t j = lane; j < npairs; j += kWave) {
      p.alive_all[base + j] = AliveRec{static_cast<int32_t>(base + j), kAliveFullScan, 0ull, 0ull};
      p.alive_flag[base + j] = 1u;
    }
    return;
  }
  const Sph A = load_sph(sph4, a);
  int32_t c0 = -1, c1 = -1;
  Sph C0{}, C1{};
  Plane P0{}, P1{};
  bool k0 = false, k1 = false;
  if (lane < cnt) {
    c0 = nbl[lane];
    C0 = load_sph(sph4, c0);
    P0 = plane_of(A, C0);
    k0 = contains(C0, A);
  }
  if (lane + kWave < cnt) {
    c1 = nbl[lane + kWave];
    C1 = load_sph(sph4, c1);
    P1 = plane_of(A, C1);
    k1 = contains(C1, A);
  }
  int64_t j = 0;
  for (int k = 0; k < cnt; ++k) {
    const int32_t b = nbl[k];
    if (!pair_qualifies(p, a, b)) continue;
    const int64_t pidx = base + j;
    ++j;
    PairGeom g;
    if (!pair_setup(A, load_sph(sph4, b), &g)) {
      if (lane == 0) {
        p.rec[pidx] = empty_record();
        p.alive_flag[pidx] = 0u;
      }
      continue;
    }
    const int cl0 = (c0 >= 0 && c0 != b) ? classify(g, C0, P0, k0) : kSkip;
    const int cl1 = (c1 >= 0 && c1 != b) ? classify(g, C1, P1, k1) : kSkip;
    const bool kill = __ballot(cl0 == kKill || cl1 == kKill) != 0ull;
    const uint64_t mlo = __ballot(cl0 == kCut);
    const uint64_t mhi = __ballot(cl1 == kCut);
    if (lane != 0) continue;
    const bool live = !kill && (mlo | mhi) != 0ull;
    if (kill) {
      p.rec[pidx] = empty_record();
    } else if (!live) {
      p.rec[pidx] = record_uncut(g, p.mode);
    } else {
      p.alive_all[pidx] = AliveRec{static_cast<int32_t>(pidx), 0u, mlo, mhi};
    }
    p.alive_flag[pidx] = live ? 1u : 0u;
  }
}

// Alive pairs in pair order (alive_flag holds the exclusive scan).
__global__ void k_alive_compact(Pipeline p) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= p.n_pairs) return;
  const uint32_t k = p.alive_flag[i];
  if (p.alive_flag[i + 1] != k) p.alive[k] = p.alive_all[i];
}

// Contour cutting of the alive pairs, one lane per pair.
__global__ __launch_bounds__(256) void k_cut_alive(Pipeline p) {
  const int64_t n_alive = p.alive_flag[p.n_pairs];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const double* sph4 = reinterpret_cast<const double*>(p.sph);
  for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < n_alive; t += stride) {
    const AliveRec ar = p.alive[t];
    const int32_t a = p.pair_a[ar.pair], b = p.pair_b[ar.pair];
    const int32_t* nbl = p.nb + static_cast<int64_t>(a) * p.nb_cap;
    LocalCut<kMaxC> ct;
    LocalArc<kMaxR> out;
    int st;
    const PairRec rec = (ar.flags & kAliveFullScan)
                            ? make_record(sph4, a, b, nbl, p.nb_cnt[a], p.mode, c_hex_s, c_hex_c, ct, out, &st)
                            : make_record_masked(sph4, a, b, nbl, ar.mlo, ar.mhi, p.mode, c_hex_s, c_hex_c, ct, out, &st);
    p.rec[ar.pair] = rec;
    if (st == kOverflow) {
      const uint32_t k = atomicAdd(p.overflow_cnt, 1u);
      p.overflow_list[k] = ar.pair;
      atomicOr(p.status, kStatContourOverflow);
    }
  }
}

// Slow path: same code with the contour in a global workspace region per
// lane, sized for the longest possible contour (6 + nb_cap points while
// cutting, twice that after the restriction).
__global__ __launch_bounds__(64) void k_cut_slow(Pipeline p, int64_t first, int64_t count) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= count) return;
  const int64_t pi = p.overflow_list[first + t];
  const int cc = 6 + p.nb_cap;
  const int cr = 2 * cc;
  double* ws = p.slow_ws + t * static_cast<int64_t>(5 * cc + 6 * cr);
  GlobalCut ct{cc, ws, ws + cc, ws + 2 * cc, reinterpret_cast<int32_t*>(ws + 3 * cc),
               reinterpret_cast<int32_t*>(ws + 4 * cc)};
  double* w2 = ws + 5 * cc;
  GlobalArc out{cr, w2, w2 + cr, w2 + 2 * cr, w2 + 3 * cr, reinterpret_cast<int32_t*>(w2 + 4 * cr),
                reinterpret_cast<int32_t*>(w2 + 5 * cr)};
  const int32_t a = p.pair_a[pi], b = p.pair_b[pi];
  const int32_t* nbl = p.nb + static_cast<int64_t>(a) * p.nb_cap;
  int st;
  PairRec rec = make_record(reinterpret_cast<const double*>(p.sph), a, b, nbl, p.nb_cnt[a], p.mode, c_hex_s, c_hex_c, ct, out, &st);
  p.rec[pi] = rec;
}

}  // namespace

void set_hexagon_constants(const double* s_half5, const double* c_half5) {
  for (int i = 0; i < 5; ++i) {
    h_hex_s[i] = s_half5[i];
    h_hex_c[i] = c_half5[i];
  }
}

hipError_t upload_hexagon_constants(hipStream_t s) {
  hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_hex_s), h_hex_s, sizeof(h_hex_s), 0, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_hex_c), h_hex_c, sizeof(h_hex_c), 0, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

hipError_t launch_cut(const Pipeline& p, hipStream_t s) {
  if (p.n_pairs <= 0) return hipSuccess;
  int64_t blocks = (p.n_pairs + 255) / 256;
  if (blocks > 65535 * 8) blocks = 65535 * 8;
  hipLaunchKernelGGL(k_cut, dim3(blocks), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_prefilter(const Pipeline& p, uint32_t* scan_tmp, hipStream_t s) {
  if (p.n <= 0 || p.n_pairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_prefilter, dim3((p.n + 3) / 4), dim3(256), 0, s, p);
  hipError_t e = launch_scan_u32(p.alive_flag, p.n_pairs, scan_tmp, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_alive_compact, dim3((p.n_pairs + 255) / 256), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_cut_alive(const Pipeline& p, hipStream_t s) {
  if (p.n_pairs <= 0) return hipSuccess;
  int64_t blocks = (p.n_pairs + 255) / 256;
  if (blocks > 65535 * 8) blocks = 65535 * 8;
  hipLaunchKernelGGL(k_cut_alive, dim3(blocks), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_cut_slow(const Pipeline& p, int64_t n_overflow, hipStream_t s) {
  for (int64_t first = 0; first < n_overflow; first += p.slow_ws_pairs) {
    const int64_t count = (n_overflow - first) < p.slow_ws_pairs ? (n_overflow - first) : p.slow_ws_pairs;
    hipLaunchKernelGGL(k_cut_slow, dim3((count + 63) / 64), dim3(64), 0, s, p, first, count);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace vltr
