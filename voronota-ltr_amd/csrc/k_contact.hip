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
