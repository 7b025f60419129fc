// Cells (replaces tessellation.rs:607-658 and types.rs:587-676) and output
// compaction.
//
// The reference accumulates every valid pair summary into the cells of both
// ends in pair order (ascending first ball, then neighbor order). Ball b
// therefore first receives, as the second end, the pairs (a, b) with a < b in
// ascending a, and then its own pairs in order. k_cells reproduces exactly
// that order per ball, so sums are bit-identical to the sequential reference
// (no atomics, no reduction-order drift).
#include "vltr_internal.h"
#include "voronota_ltr.h"

namespace vltr {
namespace {

__device__ __forceinline__ bool same_group(const Pipeline& p, int64_t a, int64_t b) {
  if (p.groups != nullptr && a < p.n_groups && b < p.n_groups) return p.groups[a] == p.groups[b];
  return false;
}
__device__ __forceinline__ bool pair_qualifies(const Pipeline& p, int64_t a, int64_t c) {
  const int64_t cc = p.periodic ? c % p.n : c;
  if (same_group(p, a, cc)) return false;
  return (p.periodic && c >= p.n) || a < c;
}

struct Acc {
  double area = 0.0, arc = 0.0, sa_pos = 0.0, sa_neg = 0.0, pv_pos = 0.0, pv_neg = 0.0;
  int added = 0;
  // types.rs:589-610
  __device__ __forceinline__ void add(const PairRec& r, bool as_first) {
    if (!(r.flags & kRecValid) || !(r.area > 0.0)) return;
    added = 1;
    area = area + r.area;
    arc = arc + r.arc_length;
    const double sa = as_first ? r.solid_a : r.solid_b;
    const double pv = as_first ? r.pyr_a : r.pyr_b;
    sa_pos = sa_pos + fmax(sa, 0.0);
    sa_neg = sa_neg - fmin(sa, 0.0);
    pv_pos = pv_pos + fmax(pv, 0.0);
    pv_neg = pv_neg - fmin(pv, 0.0);
  }
};

__global__ void k_cells(Pipeline p) {
  const int64_t b = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (b >= p.n) return;
  const int cap = p.nb_cap;
  const int32_t* lb = p.nb + b * static_cast<int64_t>(cap);
  const int nb = p.nb_cnt[b];
  Acc acc;
  // b as second end: pairs (a, b), a < b canonical, ascending a.
  int64_t prev = -1;
  for (;;) {
    int64_t best = INT64_MAX;
    for (int k = 0; k < nb; ++k) {
      const int64_t c = lb[k];
      if (c < b && c > prev && c < best && pair_qualifies(p, c, b)) best = c;
    }
    if (best == INT64_MAX) break;
    const int32_t* la = p.nb + best * static_cast<int64_t>(cap);
    const int na = p.nb_cnt[best];
    int64_t idx = p.pair_cnt[best];
    for (int k = 0; k < na; ++k) {
      const int64_t c = la[k];
      if (!pair_qualifies(p, best, c)) continue;
      if (c == b) {
        acc.add(p.rec[idx], false);
        break;
      }
      ++idx;
    }
    prev = best;
  }
  // b as first end: its own pairs in neighbor order.
  const int64_t p0 = p.pair_cnt[b], p1 = p.pair_cnt[b + 1];
  for (int64_t t = p0; t < p1; ++t) acc.add(p.rec[t], true);

  CellRec out;
  out.sas = 0.0;
  out.volume = 0.0;
  out.stage = 0;
  out.pad = 0;
  const double4 s4 = p.sph[b];
  const double r = s4.w;
  if (acc.added) {
    // types.rs:624-663
    const double diff = fabs(acc.sa_pos - acc.sa_neg);
    double sas = 0.0, vol = 0.0;
    if (acc.arc > 0.0 && diff > 1e-10) {
      if (acc.sa_pos > acc.sa_neg) {
        const double ad = acc.sa_pos - acc.sa_neg;
        sas = fma(4.0, kPi, -fmax(ad, 0.0)) * r * r;
      } else {
        const double ad = acc.sa_neg - acc.sa_pos;
        sas = fmax(ad, 0.0) * r * r;
      }
      vol = (sas * r / 3.0) + acc.pv_pos - acc.pv_neg;
      const double full = 4.0 / 3.0 * kPi * r * r * r;
      if (vol > full) {
        sas = 0.0;
        vol = acc.pv_pos - acc.pv_neg;
      }
    } else {
      vol = acc.pv_pos - acc.pv_neg;
    }
    out.sas = sas;
    out.volume = vol;
    out.stage = 2;
  } else if (nb == 0) {
    // types.rs:666-675 detached ball (no colliding neighbor)
    out.sas = 4.0 * kPi * r * r;
    out.volume = out.sas * r / 3.0;
    out.stage = 2;
  }
  p.cell[b] = out;
}

__global__ void k_flag_contacts(const PairRec* rec, int64_t n, uint32_t* flag) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  flag[i] = (rec[i].flags & kRecValid) && rec[i].area > 0.0 ? 1u : 0u;
}

__global__ void k_emit_contacts(Pipeline p, const uint32_t* pos, vltr_contact* out) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= p.n_pairs) return;
  const uint32_t k = pos[i];
  if (pos[i + 1] == k) return;
  const PairRec r = p.rec[i];
  vltr_contact c;
  c.id_a = static_cast<uint64_t>(p.pair_a[i]);
  c.id_b = static_cast<uint64_t>(p.pair_b[i]);
  c.area = r.area;
  c.arc_length = r.arc_length;
  c.central = (r.flags & kRecCentral) ? 1 : 0;
  for (int j = 0; j < 7; ++j) c.reserved[j] = 0;
  out[k] = c;
}

__global__ void k_flag_cells(const CellRec* cell, int64_t n, uint32_t* flag) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  flag[i] = cell[i].stage == 2 ? 1u : 0u;
}

__global__ void k_emit_cells(const CellRec* cell, int64_t n, const uint32_t* pos, vltr_cell* out) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = pos[i];
  if (pos[i + 1] == k) return;
  vltr_cell c;
  c.index = static_cast<uint64_t>(i);
  c.sas_area = cell[i].sas;
  c.volume = cell[i].volume;
  out[k] = c;
}

inline int64_t blocks_for(int64_t n, int64_t per) { return (n + per - 1) / per; }

}  // namespace

hipError_t launch_cells(const Pipeline& p, hipStream_t s) {
  if (p.n > 0) hipLaunchKernelGGL(k_cells, dim3(blocks_for(p.n, 128)), dim3(128), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_flag_contacts(const Pipeline& p, uint32_t* flag, hipStream_t s) {
  if (p.n_pairs > 0)
    hipLaunchKernelGGL(k_flag_contacts, dim3(blocks_for(p.n_pairs, 256)), dim3(256), 0, s, p.rec, p.n_pairs, flag);
  return hipGetLastError();
}
hipError_t launch_emit_contacts(const Pipeline& p, const uint32_t* pos, vltr_contact* out, hipStream_t s) {
  if (p.n_pairs > 0) hipLaunchKernelGGL(k_emit_contacts, dim3(blocks_for(p.n_pairs, 256)), dim3(256), 0, s, p, pos, out);
  return hipGetLastError();
}
hipError_t launch_flag_cells(const Pipeline& p, uint32_t* flag, hipStream_t s) {
  if (p.n > 0) hipLaunchKernelGGL(k_flag_cells, dim3(blocks_for(p.n, 256)), dim3(256), 0, s, p.cell, p.n, flag);
  return hipGetLastError();
}
hipError_t launch_emit_cells(const Pipeline& p, const uint32_t* pos, vltr_cell* out, hipStream_t s) {
  if (p.n > 0) hipLaunchKernelGGL(k_emit_cells, dim3(blocks_for(p.n, 256)), dim3(256), 0, s, p.cell, p.n, pos, out);
  return hipGetLastError();
}

}  // namespace vltr
