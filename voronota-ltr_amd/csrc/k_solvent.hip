// Pseudo-solvent spheres (replaces src/solvent_spheres.rs:131-425): sampling
// of the exposed atoms after the first tessellation.
//
// One wavefront per ball a with an exposed cell (stage 2, sas_area > 0). Its
// contact partners are the other ends of its valid pair records: its own
// pairs (second ends, canonical), the pairs it received (first ends) and, in
// a periodic box, the pairs whose second end is an image of a (first ends,
// gathered into a small CSR by k_solv_img_*), which together are the
// canonical contact set the reference's neighbor graph is built from
// (solvent_spheres.rs:205-235 over tessellation.rs:172-188).
// Lanes take the icosahedron samples center + v (r + probe); a sample is kept
// when |p - center| - r is below |p - c| - r_c for every partner c (canonical
// positions, as solvent_spheres.rs:269-288). Kept samples are written in
// (ball, sample) order after a scan of the per-ball counts.
#include "vltr_internal.h"

namespace vltr {
namespace {

constexpr int kWave = 64;
constexpr int kSolvBlock = 256;
constexpr int kSolvWaves = kSolvBlock / kWave;

__device__ __forceinline__ bool exposed(const Pipeline& p, int64_t a) {
  const CellRec c = p.cell[a];
  return c.stage == 2 && c.sas > 0.0;
}

// Periodic box: per canonical ball, the first ends of valid pairs whose second
// end is one of its images (count, then fill after a scan of the counts).
__global__ void k_solv_img_count(Pipeline p, uint32_t* icnt) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= p.n_pairs || p.pair_b[t] < p.n || !has_contact(p, t)) return;
  atomicAdd(&icnt[canon_id(p, p.pair_b[t])], 1u);
}

__global__ void k_solv_img_fill(Pipeline p, const uint32_t* ioff, uint32_t* cur, int32_t* img) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= p.n_pairs || p.pair_b[t] < p.n || !has_contact(p, t)) return;
  const int64_t b = canon_id(p, p.pair_b[t]);
  img[ioff[b] + atomicAdd(&cur[b], 1u)] = p.pair_a[t];
}

__global__ void k_solv_zero(uint32_t* v, int64_t n) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n) v[i] = 0u;
}

__device__ __forceinline__ bool closer(const Pipeline& p, int64_t c, V3 pt, double d_atom) {
  const double4 o = p.balls_in[c];
  return norm(pt - V3{o.x, o.y, o.z}) - o.w <= d_atom;
}

// Keeps sample p of ball A if no partner's surface is as close.
__device__ bool sample_ok(const Pipeline& p, const uint32_t* ioff, const int32_t* img, int64_t a, V3 pt,
                          double d_atom) {
  const int64_t own0 = p.pair_cnt[a], own1 = p.pair_cnt[a + 1];
  for (int64_t t = own0; t < own1; ++t) {
    if (has_contact(p, t) && closer(p, canon_id(p, p.pair_b[t]), pt, d_atom)) return false;
  }
  const int nin = min(static_cast<int>(p.in_cnt[a]), p.nb_cap);
  const int32_t* inc = p.incoming + a * static_cast<int64_t>(p.nb_cap);
  for (int k = 0; k < nin; ++k) {
    const int32_t t = inc[k];
    if (has_contact(p, t) && closer(p, p.pair_a[t], pt, d_atom)) return false;
  }
  if (ioff != nullptr)
    for (uint32_t k = ioff[a]; k < ioff[a + 1]; ++k)
      if (closer(p, img[k], pt, d_atom)) return false;
  return true;
}

// emit == 0: cnt[a] = kept samples of ball a. emit != 0: writes them from
// pos[a] on (solvent balls of radius sr into comb[n + k], parents).
__global__ __launch_bounds__(kSolvBlock) void k_solvent(Pipeline p, const uint32_t* ioff, const int32_t* img,
                                                         const double* sih, int K, double probe, double sr,
                                                         uint32_t* cnt, int emit, double4* comb, uint32_t* parent) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t a = static_cast<int64_t>(blockIdx.x) * kSolvWaves + threadIdx.x / kWave;
  if (a >= p.n) return;
  if (!exposed(p, a)) {
    if (!emit && lane == 0) cnt[a] = 0;
    return;
  }
  const double4 A4 = p.balls_in[a];
  const V3 center{A4.x, A4.y, A4.z};
  const double radius = A4.w + probe;
  uint32_t pos = emit ? cnt[a] : 0u;
  uint32_t kept = 0;
  for (int k0 = 0; k0 < K; k0 += kWave) {
    const int k = k0 + lane;
    bool ok = false;
    V3 pt{0.0, 0.0, 0.0};
    if (k < K) {
      const V3 v{sih[3 * k], sih[3 * k + 1], sih[3 * k + 2]};
      // subdivided_icosahedron.rs:128-130: center + v * radius
      pt = center + v * radius;
      ok = sample_ok(p, ioff, img, a, pt, norm(pt - center) - A4.w);
    }
    const unsigned long long m = __ballot(ok);
    if (emit && ok) {
      const uint32_t q = pos + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
      comb[p.n + q] = make_double4(pt.x, pt.y, pt.z, sr);
      parent[q] = static_cast<uint32_t>(a);
    }
    pos += static_cast<uint32_t>(__popcll(m));
    kept += static_cast<uint32_t>(__popcll(m));
  }
  if (!emit && lane == 0) cnt[a] = kept;
}

// solvent_spheres.rs:394-401: the atoms enter the second tessellation with
// radius max(r - probe + volume_probe, 0.01).
__global__ void k_solvent_atoms(const double4* balls, int64_t n, double probe, double volume_probe, double4* comb) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const double4 b = balls[i];
  comb[i] = make_double4(b.x, b.y, b.z, fmax(b.w - probe + volume_probe, 0.01));
}

inline int64_t blocks_for(int64_t n, int64_t per) { return (n + per - 1) / per; }

}  // namespace

hipError_t launch_solvent_images(const Pipeline& p, uint32_t* ioff, uint32_t* cur, int32_t* img, uint32_t* scan_tmp,
                                 hipStream_t s) {
  if (p.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_solv_zero, dim3(blocks_for(p.n + 1, 256)), dim3(256), 0, s, ioff, p.n + 1);
  hipLaunchKernelGGL(k_solv_zero, dim3(blocks_for(p.n, 256)), dim3(256), 0, s, cur, p.n);
  if (p.n_pairs > 0)
    hipLaunchKernelGGL(k_solv_img_count, dim3(blocks_for(p.n_pairs, 256)), dim3(256), 0, s, p, ioff);
  hipError_t e = launch_scan_u32(ioff, p.n, scan_tmp, s);
  if (e != hipSuccess) return e;
  if (p.n_pairs > 0)
    hipLaunchKernelGGL(k_solv_img_fill, dim3(blocks_for(p.n_pairs, 256)), dim3(256), 0, s, p, ioff, cur, img);
  return hipGetLastError();
}

hipError_t launch_solvent(const Pipeline& p, const uint32_t* ioff, const int32_t* img, const double* sih, int K,
                          double probe, double sr, uint32_t* cnt, int emit, double4* comb, uint32_t* parent,
                          hipStream_t s) {
  if (p.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_solvent, dim3(blocks_for(p.n, kSolvWaves)), dim3(kSolvBlock), 0, s, p, ioff, img, sih, K, probe,
                     sr, cnt, emit, comb, parent);
  return hipGetLastError();
}

hipError_t launch_solvent_atoms(const double4* balls, int64_t n, double probe, double volume_probe, double4* comb,
                                hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_solvent_atoms, dim3(blocks_for(n, 256)), dim3(256), 0, s, balls, n, probe, volume_probe, comb);
  return hipGetLastError();
}

}  // namespace vltr
