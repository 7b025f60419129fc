// Backward pass of the tessellation energy (adjoint.h): one 16-lane group per
// contact, four per wavefront, over the contacts of the last full forward run
// (records, cut topologies and cell seeds written by k_clip, k_clip_wide,
// k_cut_slow and k_cells), listed first by a flag / scan / emit pass so that
// no group idles on a candidate pair without a contact. Gradients of the centres are added
// with FP64 atomics (image ids fold onto their ball). The same kernel sums
// the energy: contact terms by the groups, cell terms by all lanes, one
// partial per block.
//
// Periodic box: a pair whose second end is an image adds nothing to that
// ball's cell (tessellation.rs:624-631), so its second-side cell terms are
// zero; and among the contacts through an image the list keeps only the
// first of each canonical pair (tessellation.rs:556-601; contacts between two
// canonical balls are always kept), so the W * area term of such a record
// counts only if it is that first one, found through a hash table of
// canonical pairs filled by k_dedup_insert (smallest pair index per key).
#include "adjoint.h"
#include "dedup.h"
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
  double inv_n;
  __device__ void add(int64_t id, V3 v) {
    const int64_t i = canon_mod(id, n, inv_n);
    atomicAdd(grad + 3 * i, v.x);
    atomicAdd(grad + 3 * i + 1, v.y);
    atomicAdd(grad + 3 * i + 2, v.z);
  }
};

__global__ void k_dedup_clear(DedupTable T) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i <= T.mask;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    T.keys[i] = 0ull;
    T.first[i] = 0xffffffffu;
  }
}

// all == 0: contacts through an image only; all != 0: every contact.
__global__ void k_dedup_insert(Pipeline p, DedupTable T, int all) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= p.n_pairs) return;
  // graph replays: pairs past the device-side total hold no record of this run
  if (p.n_pairs_dev != nullptr && t >= static_cast<int64_t>(*p.n_pairs_dev)) return;
  if ((all == 0 && p.pair_b[t] < p.n) || !has_contact(p, t)) return;
  const unsigned long long key = pair_key(p, t);
  uint64_t slot = mix64(key) & T.mask;
  for (;;) {
    const unsigned long long prev = atomicCAS(T.keys + slot, 0ull, key);
    if (prev == 0ull || prev == key) {
      atomicMin(T.first + slot, static_cast<uint32_t>(t));
      return;
    }
    slot = (slot + 1) & T.mask;
  }
}

__global__ void k_zero_u32(uint32_t* a, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    a[i] = 0u;
}

__global__ void k_zero(double* a, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    a[i] = 0.0;
}

// The topology of contact t: its store record and point count (m = 0: the
// uncut circle, no record read; m < 0: none, no gradient).
__device__ __forceinline__ const ContactTopo* contact_topo(const Pipeline& p, int64_t t, int* m) {
  const uint32_t s = p.topo_slot[t];
  const bool stored = topo_slot_stored(s) && static_cast<int64_t>(s) < kTopoRegions * p.topo_region;
  const ContactTopo* tp = p.topo + (stored ? s : 0u);
  *m = stored ? tp->m : (s == kTopoSlotUncut ? 0 : kTopoNone);
  return tp;
}

// The contacts with a topology (valid record, positive area, m >= 0), listed
// in pair order so that every group of the adjoint takes a real contact.
__global__ void k_adj_flag(Pipeline p, uint32_t* flag) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= p.n_pairs) return;
  bool with = has_contact(p, i);
  if (with) {
    int m = 0;
    contact_topo(p, i, &m);
    with = m >= 0;
  }
  flag[i] = with ? 1u : 0u;
}
__global__ void k_adj_list(Pipeline p, const uint32_t* pos, int32_t* list) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= p.n_pairs || pos[i + 1] == pos[i]) return;
  list[pos[i]] = static_cast<int32_t>(i);
}

// Occupancy target: the vector-Jacobian products want about 256 registers per
// lane; two wavefronts per SIMD (with some spills) measure slightly faster
// than one (20k balls: forward + adjoint 1.39 vs 1.41 ms), three slower (1.62).
#ifndef VLTR_ADJ_WAVES_PER_EU
#define VLTR_ADJ_WAVES_PER_EU 2
#endif
__global__ __launch_bounds__(kAdjBlock) __attribute__((amdgpu_waves_per_eu(VLTR_ADJ_WAVES_PER_EU))) void k_adjoint(Pipeline p, const int32_t* types, W16 W, double* grad,
                                                       double* part, DedupTable T, const int32_t* list,
                                                       const uint32_t* count) {
  __shared__ AdjLds lds[kAdjGroups];
  __shared__ double red[kAdjBlock / 64];
  // W indexed by the contact's types: from shared memory (indexing the
  // by-value argument would copy it to scratch memory)
  __shared__ double ws[16];
  if (threadIdx.x < 16) ws[threadIdx.x] = W.w[threadIdx.x];
  __syncthreads();
  const int gl = threadIdx.x / kGroup;
  const int ln = threadIdx.x & (kGroup - 1);
  WaveGroup G{(gl % (64 / kGroup)) * kGroup, ln};
  const double* sph4 = reinterpret_cast<const double*>(p.sph);
  DevAcc acc{grad, p.n, p.inv_n};
  double e = 0.0;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * kAdjGroups;
  const int64_t n_list = *count;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kAdjGroups + gl; k < n_list; k += groups) {
    const int64_t t = list[k];
    const PairRec rec = p.rec[t];
    int m = 0;
    const ContactTopo& topo = *contact_topo(p, t, &m);  // listed by k_adj_flag: m >= 0
    const int32_t a = p.pair_a[t], b = p.pair_b[t];
    const int64_t ca = canon_id(p, a), cb = canon_id(p, b);
    const bool image = p.periodic && b >= p.n;
    const bool first = image && dedup_first(p, T, t);
    const double w = !image || first ? ws[4 * types[ca] + types[cb]] : 0.0;
    // cell shares: each ball's own records (an image end has no share), or,
    // with dedup cells, the first record of a canonical pair through an
    // image for both ends and the others for neither
    bool a_cell = true, b_cell = b < p.n;
    if (p.dd_keys != nullptr && image) {
      a_cell = first;
      b_cell = first && cb != ca;
    }
    if (ln == 0) e += w * rec.area;
    group_contact_adjoint(G, &lds[gl], sph4, a, b, topo, m, w, a_cell ? p.seed[ca] : 0.0, b_cell ? p.seed[cb] : 0.0,
                          a_cell ? (rec.solid_a < 0.0 ? -1.0 : 1.0) : 0.0,
                          b_cell ? (rec.solid_b < 0.0 ? -1.0 : 1.0) : 0.0, acc, p.volume_weight);
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

hipError_t launch_dedup_table(const Pipeline& p, unsigned long long* keys, uint32_t* first, int all, hipStream_t s) {
  DedupTable T{keys, first, dedup_slots(p.n_pairs) - 1};
  hipLaunchKernelGGL(k_dedup_clear, dim3(1024), dim3(256), 0, s, T);
  if (p.n_pairs > 0) hipLaunchKernelGGL(k_dedup_insert, dim3((p.n_pairs + 255) / 256), dim3(256), 0, s, p, T, all);
  return hipGetLastError();
}

uint64_t dedup_slots(int64_t n_pairs) { return dedup_slots_mask(n_pairs) + 1; }

hipError_t launch_adjoint(const Pipeline& p, const int32_t* types, const double* w16, double* grad, double* part,
                          unsigned long long* dedup_keys, uint32_t* dedup_first, const AdjointWork& work,
                          hipStream_t s) {
  W16 W;
  for (int k = 0; k < 16; ++k) W.w[k] = w16[k];
  DedupTable T{dedup_keys, dedup_first, dedup_slots(p.n_pairs) - 1};
  if (p.dd_keys != nullptr) T = DedupTable{p.dd_keys, p.dd_first, dedup_slots(p.n_pairs) - 1};
  if (p.n_pairs > 0) {
    hipLaunchKernelGGL(k_adj_flag, dim3((p.n_pairs + 255) / 256), dim3(256), 0, s, p, work.flags);
    const hipError_t e = launch_scan_u32(work.flags, p.n_pairs, work.scan_tmp, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_adj_list, dim3((p.n_pairs + 255) / 256), dim3(256), 0, s, p, work.flags, work.list);
  }
  if (p.periodic && p.dd_keys == nullptr) {
    const hipError_t e = launch_dedup_table(p, dedup_keys, dedup_first, 0, s);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, s, grad, 3 * p.n);
  // no pairs: the scanned total flags[0] is not written; the list is empty
  if (p.n_pairs <= 0) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, work.flags, 1);
  }
  hipLaunchKernelGGL(k_adjoint, dim3(kAdjBlocks), dim3(kAdjBlock), 0, s, p, types, W, grad, part, T, work.list,
                     work.flags + (p.n_pairs > 0 ? p.n_pairs : 0));
  return hipGetLastError();
}

}  // namespace vltr
