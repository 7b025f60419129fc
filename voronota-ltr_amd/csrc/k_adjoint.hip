// Backward pass of the tessellation energy (adjoint.h): one 16-lane group per
// contact, four per wavefront, over all candidate pairs of the last full
// forward run (records, cut topologies and cell seeds written by k_clip,
// k_clip_wide, k_cut_slow and k_cells). Gradients of the centres are added
// with FP64 atomics (image ids fold onto their ball). The same kernel sums
// the energy: contact terms by the groups, cell terms by all lanes, one
// partial per block.
#include "adjoint.h"
#include "vltr_internal.h"
#include "wave_group.h"

namespace vltr {
namespace {

constexpr int kAdjBlock = 256;
constexpr int kAdjGroups = kAdjBlock / kGroup;
constexpr int kAdjBlocks = 2048;

struct W16 {
  double w[16];
};

struct DevAcc {
  double* grad;
  int64_t n;
  __device__ void add(int64_t id, V3 v) {
    const int64_t i = id % n;
    atomicAdd(grad + 3 * i, v.x);
    atomicAdd(grad + 3 * i + 1, v.y);
    atomicAdd(grad + 3 * i + 2, v.z);
  }
};

__global__ void k_zero(double* a, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    a[i] = 0.0;
}

__global__ __launch_bounds__(kAdjBlock) void k_adjoint(Pipeline p, const int32_t* types, W16 W, double* grad,
                                                       double* part) {
  __shared__ AdjLds lds[kAdjGroups];
  __shared__ double red[kAdjBlock / 64];
  const int gl = threadIdx.x / kGroup;
  const int ln = threadIdx.x & (kGroup - 1);
  WaveGroup G{(gl % (64 / kGroup)) * kGroup, ln};
  const double* sph4 = reinterpret_cast<const double*>(p.sph);
  DevAcc acc{grad, p.n};
  double e = 0.0;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * kAdjGroups;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kAdjGroups + gl; t < p.n_pairs; t += groups) {
    const PairRec rec = p.rec[t];
    if (!(rec.flags & kRecValid) || !(rec.area > 0.0)) continue;
    const ContactTopo& topo = p.topo[t];
    if (topo.m < 0) continue;
    const int32_t a = p.pair_a[t], b = p.pair_b[t];
    const int64_t ca = a % p.n, cb = b % p.n;
    const double w = W.w[4 * types[ca] + types[cb]];
    if (ln == 0) e += w * rec.area;
    group_contact_adjoint(G, &lds[gl], sph4, a, b, topo, w, p.seed[ca], p.seed[cb], rec.solid_a < 0.0 ? -1.0 : 1.0,
                          rec.solid_b < 0.0 ? -1.0 : 1.0, acc, p.volume_weight);
  }
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < p.n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const CellRec c = p.cell[i];
    if (c.stage == 2) e += p.sas_weight * c.sas + p.volume_weight * c.volume;
  }
  for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < kAdjBlock / 64; ++k) s += red[k];
    part[blockIdx.x] = s;
  }
}

}  // namespace

int64_t adjoint_blocks() { return kAdjBlocks; }

hipError_t launch_adjoint(const Pipeline& p, const int32_t* types, const double* w16, double* grad, double* part,
                          hipStream_t s) {
  W16 W;
  for (int k = 0; k < 16; ++k) W.w[k] = w16[k];
  hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, s, grad, 3 * p.n);
  hipLaunchKernelGGL(k_adjoint, dim3(kAdjBlocks), dim3(kAdjBlock), 0, s, p, types, W, grad, part);
  return hipGetLastError();
}

}  // namespace vltr
