// Cells (replaces tessellation.rs:607-658 and types.rs:587-676) and output
// compaction.
//
// The reference accumulates every valid pair summary into the cells of both
// ends in pair order (ascending first ball, then neighbor order). Ball b
// therefore first receives, as the second end, the pairs (a, b) with a < b in
// ascending a, and then its own pairs in order. k_cells reproduces exactly
// that order per ball, so sums are bit-identical to the sequential reference
// (no atomics, no reduction-order drift).
#include "dedup.h"
#include "vltr_internal.h"
#include "voronota_ltr.h"

namespace vltr {
namespace {

// The sums a cell accumulates over its contacts (types.rs:589-610).
struct Acc {
  double area = 0.0, arc = 0.0, sa_pos = 0.0, sa_neg = 0.0, pv_pos = 0.0, pv_neg = 0.0;
  int added = 0;
};

constexpr int kWave = 64;
constexpr int kCellsBlock = 64;  // one wavefront per block

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A group of kCellLanes lanes per ball b, kCellBalls balls per wavefront.
// The ball's contributions arrive in the reference's order: first the pairs
// (a, b) with a < b by ascending pair index (= ascending a), then b's own
// pairs. The group's lanes fetch records in parallel into a kCellLanes-entry
// window of shared memory in that order; then its lanes 0-5 each keep one of
// the six sums (types.rs:589-610: area, arc length, positive and negative
// solid angles, positive and negative pyramid volumes) and add the window's
// contributions to it in order, skipping records without a contact. Each sum
// sees the reference's sequence of additions, so it is bit-identical; the six
// run side by side instead of one after another, and four balls' chains of
// loads and additions overlap in one wavefront.
constexpr int kCellLanes = 16;
constexpr int kCellBalls = kWave / kCellLanes;

__device__ __forceinline__ double cell_term(const PairRec& r, bool as_first, int f) {
  if (f == 0) return r.area;
  if (f == 1) return r.arc_length;
  const double v = f < 4 ? (as_first ? r.solid_a : r.solid_b) : (as_first ? r.pyr_a : r.pyr_b);
  // x - fmin(v, 0) is x + (-fmin(v, 0)) exactly
  return (f & 1) ? -fmin(v, 0.0) : fmax(v, 0.0);
}

__global__ __launch_bounds__(kCellsBlock) void k_cells(Pipeline p) {
  __shared__ PairRec win[kCellBalls][kCellLanes];
  const int lane = threadIdx.x & (kWave - 1);
  const int g = lane / kCellLanes;          // the group (ball) of this lane
  const int j = lane & (kCellLanes - 1);    // the lane within its group
  const int gbase = g * kCellLanes;
  if (pairs_overflowed(p)) return;
  // one structure: balls in cell order, like the neighbor and clip kernels (a
  // ball's incoming records are those of its neighbors, whose own records the
  // same XCD reads at about the same time); batches and periodic runs: index
  // order (each XCD's range then holds whole structures)
  const int64_t q = xcd_remap(blockIdx.x, gridDim.x) * kCellBalls + g;
  // whole groups leave together (wave-private shared memory, no block barrier)
  if (q >= p.n) return;
  const int64_t b = (p.periodic || p.n_segs > 1) ? q : static_cast<int64_t>(p.sorted_id[q]);
  PairRec* wg = win[g];
  // the group's share of a wavefront ballot
  auto gballot = [&](bool v) -> uint32_t {
    return static_cast<uint32_t>((__ballot(v) >> gbase) & ((1ull << kCellLanes) - 1ull));
  };
  const int cap = p.nb_cap;
  const int nb = p.nb_cnt[b];
  const int nin = min(static_cast<int>(p.in_cnt[b]), cap);
  const int32_t* inc = p.incoming + b * static_cast<int64_t>(cap);
  const int f = j < 6 ? j : 0;  // the sum this lane keeps (lanes 0-5 of the group)
  double sum = 0.0;
  bool added = false;
  // dedup cells: an own record through an image counts only when it is the
  // first of its canonical pair (the first records of other rows came in
  // through the incoming list)
  const DedupTable T{p.dd_keys, p.dd_first, p.dd_keys != nullptr ? dedup_slots_mask(p.n_pairs) : 0ull};
  // adds the window's records with a contact to this lane's sum, in order;
  // own_base >= 0: the window holds own records own_base, own_base + 1, ...
  auto add_window = [&](int m, bool as_first, int64_t own_base) {
    bool valid = false;
    if (j < m) {
      const PairRec& r = wg[j];
      valid = (r.flags & kRecValid) && r.area > 0.0;
      if (valid && own_base >= 0 && p.dd_keys != nullptr && p.pair_b[own_base + j] >= p.n)
        valid = dedup_first(p, T, own_base + j);
    }
    uint32_t vm = gballot(valid);
    added = added || vm != 0u;
    while (vm != 0u) {
      const int k = __builtin_ctz(vm);
      vm &= vm - 1u;
      sum = sum + cell_term(wg[k], as_first, f);
    }
  };
  const int64_t own0 = p.pair_cnt[b];
  const int nown = static_cast<int>(static_cast<int64_t>(p.pair_cnt[b + 1]) - own0);
  // the first kCellLanes own records are requested first: their loads overlap
  // the incoming part (as 8-byte words: a PairRec local would live in scratch
  // memory); own records without a contact, as pair_contact marks them, are
  // not loaded: their window entries get flags 0
  constexpr int kRecWords = static_cast<int>(sizeof(PairRec) / 8);
  static_assert(sizeof(PairRec) % 8 == 0, "records are moved as 8-byte words");
  const uint8_t* pc = p.pair_contact;
  uint64_t own_first[kRecWords] = {};
  if (j < nown && (pc == nullptr || pc[own0 + j] != 0)) {
#pragma unroll
    for (int k = 0; k < kRecWords; ++k) own_first[k] = reinterpret_cast<const uint64_t*>(p.rec + own0 + j)[k];
  }
  if (nin <= kCellLanes) {
    // one entry per lane, ranked by comparisons with the group's others
    const int32_t pe = j < nin ? inc[j] : 0;
    int rank = 0;
    for (int k = 0; k < nin; ++k) rank += __shfl(pe, k, kCellLanes) < pe ? 1 : 0;
    if (j < nin) wg[rank] = p.rec[pe];
    wave_sync();
    add_window(nin, false, -1);
    wave_sync();
  } else {
    for (int c0 = 0; c0 < nin; c0 += kCellLanes) {
      for (int e = j; e < nin; e += kCellLanes) {
        const int32_t pe = inc[e];
        int rank = 0;
        for (int k = 0; k < nin; ++k) rank += inc[k] < pe ? 1 : 0;
        if (rank >= c0 && rank < c0 + kCellLanes) wg[rank - c0] = p.rec[pe];
      }
      wave_sync();
      add_window(min(kCellLanes, nin - c0), false, -1);
      wave_sync();
    }
  }
  for (int c0 = 0; c0 < nown; c0 += kCellLanes) {
    if (c0 == 0) {
      if (j < nown) {
#pragma unroll
        for (int k = 0; k < kRecWords; ++k) reinterpret_cast<uint64_t*>(&wg[j])[k] = own_first[k];
      }
    } else if (c0 + j < nown) {
      if (pc == nullptr || pc[own0 + c0 + j] != 0)
        wg[j] = p.rec[own0 + c0 + j];
      else
        wg[j].flags = 0u;
    }
    wave_sync();
    add_window(min(kCellLanes, nown - c0), true, own0 + c0);
    wave_sync();
  }
  Acc acc;
  acc.area = __shfl(sum, 0, kCellLanes);
  acc.arc = __shfl(sum, 1, kCellLanes);
  acc.sa_pos = __shfl(sum, 2, kCellLanes);
  acc.sa_neg = __shfl(sum, 3, kCellLanes);
  acc.pv_pos = __shfl(sum, 4, kCellLanes);
  acc.pv_neg = __shfl(sum, 5, kCellLanes);
  acc.added = added ? 1 : 0;
  if (j != 0) return;

  CellRec out;
  out.sas = 0.0;
  out.volume = 0.0;
  out.stage = 0;
  out.pad = 0;
  const double4 s4 = p.sph[b];
  const double r = s4.w;
  // adjoint seed dE/d(solid angle) (0 outside the standard branch)
  double seed = 0.0;
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
      seed = -r * r * (p.sas_weight + p.volume_weight * r / 3.0);
      if (vol > full) {
        sas = 0.0;
        vol = acc.pv_pos - acc.pv_neg;
        seed = 0.0;
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
  if (p.seed != nullptr) p.seed[b] = seed;
}

// Dedup cells: the first record of a canonical pair through an image also
// counts for its second end's canonical ball (updateable.rs:565-630 stores it
// in both balls' lists), which k_cells reads from its incoming list.
__global__ void k_dedup_incoming(Pipeline p, uint64_t mask) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= p.n_pairs) return;
  const int64_t b = p.pair_b[t];
  if (b < p.n) return;
  if (!has_contact(p, t)) return;
  const int64_t cb = canon_id(p, b);
  if (cb == p.pair_a[t] || !dedup_first(p, DedupTable{p.dd_keys, p.dd_first, mask}, t)) return;
  const uint32_t slot = atomicAdd(&p.in_cnt[cb], 1u);
  if (slot < static_cast<uint32_t>(p.nb_cap))
    p.incoming[cb * p.nb_cap + slot] = static_cast<int32_t>(t);
  else
    atomicOr(p.status, kStatNbOverflow);
}

// Records with a contact (area > 0). With a canonical-pair table (T.keys
// non-null, periodic runs) only the first record in pair order of each
// canonical pair the table holds is kept: full runs keep every contact
// between two canonical balls and the first of each pair through an image
// (tessellation.rs:527-601), contacts-only runs the first of each canonical
// pair over all contacts (tessellation.rs:172-188).
__global__ void k_flag_contacts(Pipeline p, DedupTable T, uint32_t* flag) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= p.n_pairs) return;
  // graph replays: pairs past the device-side total hold no record of this run
  if (p.n_pairs_dev != nullptr && i >= static_cast<int64_t>(*p.n_pairs_dev)) {
    flag[i] = 0u;
    return;
  }
  bool keep = has_contact(p, i);
  if (keep && T.keys != nullptr && (p.mode != 0 || p.pair_b[i] >= p.n)) keep = dedup_first(p, T, i);
  flag[i] = keep ? 1u : 0u;
}

__global__ void k_emit_contacts(Pipeline p, const uint32_t* pos, vltr_contact* out) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= p.n_pairs) return;
  const uint32_t k = pos[i];
  if (pos[i + 1] == k || (p.contacts_cap > 0 && static_cast<int64_t>(k) >= p.contacts_cap)) return;
  const PairRec r = p.rec[i];
  // ids local to the ball's structure (batched runs)
  const int32_t a = p.pair_a[i];
  const int64_t off = p.n_segs > 1 ? p.segs[p.seg_of[a]].ball_off : 0;
  vltr_contact c;
  if (p.periodic) {
    // canonical ids; contacts-only runs order them (tessellation.rs:176-186)
    int64_t ca = canon_id(p, a), cb = canon_id(p, p.pair_b[i]);
    if (p.mode != 0 && ca > cb) {
      const int64_t t = ca;
      ca = cb;
      cb = t;
    }
    c.id_a = static_cast<uint64_t>(ca);
    c.id_b = static_cast<uint64_t>(cb);
  } else {
    c.id_a = static_cast<uint64_t>(a - off);
    c.id_b = static_cast<uint64_t>(p.pair_b[i] - off);
  }
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

__global__ void k_emit_cells(Pipeline p, const uint32_t* pos, vltr_cell* out) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= p.n) return;
  const uint32_t k = pos[i];
  if (pos[i + 1] == k) return;
  const CellRec* cell = p.cell;
  vltr_cell c;
  c.index = static_cast<uint64_t>(i - (p.n_segs > 1 ? p.segs[p.seg_of[i]].ball_off : 0));
  c.sas_area = cell[i].sas;
  c.volume = cell[i].volume;
  out[k] = c;
}

__global__ void k_emit_net(Pipeline p, const uint32_t* pos, ContactNet* out) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= p.n_pairs) return;
  const uint32_t k = pos[i];
  if (pos[i + 1] == k) return;
  const ContactNet& src = p.net[i];
  ContactNet& dst = out[k];
  const int m = src.m;
  dst.m = m;
  dst.r_ic = src.r_ic;
  for (int t = 0; t < m && t < kTopoCap; ++t) {
    dst.l[t] = src.l[t];
    dst.r[t] = src.r[t];
    dst.x[t] = src.x[t];
    dst.y[t] = src.y[t];
    dst.z[t] = src.z[t];
    dst.ang[t] = src.ang[t];
  }
}

__global__ void k_seg_offsets(Pipeline p, const uint32_t* pos, int what, uint32_t* out) {
  const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (k > p.n_segs) return;
  const int64_t ball = k < p.n_segs ? p.segs[k].ball_off : p.n;
  // contacts: the first pair of the structure's first ball (within the
  // scanned range; an overflowed graph replay is rerun by the host)
  const int64_t first = static_cast<int64_t>(p.pair_cnt[ball]);
  out[k] = what == 0 ? pos[first < p.n_pairs ? first : p.n_pairs] : pos[ball];
}

// updateable.rs:633-727 with the container's exclusion statuses excl
// (spheres_container.rs: hidden balls, and balls excluded by the caller).
// After init / full reinit (full != 0) a ball without contacts is detached
// only if it is not excluded and has no colliding neighbor; after an
// incremental update an affected ball without contacts is detached whenever
// it is not excluded, and unaffected balls keep their cells. refresh != 0
// (init and updates, not set_exclusion) resets the exclusion status of every
// recomputed ball to its hidden status, as the collision update does. A ball
// that joined the affected set during the run (a new neighbor of a moved
// ball) still carried its old status when the pairs were counted; when that
// status was excluded and is now cleared, the run's pairs miss this ball's
// and the report word kReportExclFlip asks the host for a second pass.
__global__ void k_updateable_cells(Pipeline p, CellRec* state, uint8_t* excl, const uint8_t* affected, int full,
                                   int refresh) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= p.n) return;
  if (!full && !affected[i]) {
    p.cell[i] = state[i];
    return;
  }
  if (refresh) {
    const uint8_t hidden = p.excluded[i];
    if (excl[i] != 0 && hidden == 0) p.status[kReportExclFlip] = 1u;
    excl[i] = hidden;
  }
  const CellRec rec = p.cell[i];
  const bool from_contacts = rec.stage == 2 && p.nb_cnt[i] > 0;
  CellRec out = rec;
  if (!from_contacts) {
    out.sas = 0.0;
    out.volume = 0.0;
    out.stage = 0;
    if (!excl[i] && (!full || p.nb_cnt[i] == 0)) {
      // types.rs:666-675
      const double r = p.sph[i].w;
      out.sas = 4.0 * kPi * r * r;
      out.volume = out.sas * r / 3.0;
      out.stage = 2;
    }
  }
  p.cell[i] = out;
  state[i] = out;
}

// Balls marked 1 (old neighbors of changed balls) and 3 (changed) for the
// affected-set threshold of spheres_container.rs:105-124; one atomic per
// wavefront and count.
__global__ void k_count_marks(const uint8_t* mask, int64_t n, uint32_t* counts) {
  uint32_t c1 = 0, c3 = 0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const uint8_t m = mask[i];
    c1 += m == 1 ? 1u : 0u;
    c3 += m == 3 ? 1u : 0u;
  }
  for (int off = 32; off > 0; off >>= 1) {
    c1 += __shfl_xor(c1, off, 64);
    c3 += __shfl_xor(c3, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (c1) atomicAdd(&counts[0], c1);
    if (c3) atomicAdd(&counts[1], c3);
  }
}
__global__ void k_zero_two(uint32_t* counts) {
  if (threadIdx.x < 2) counts[threadIdx.x] = 0u;
}

// spheres_container.rs:105-124 and 150-170: the colliding neighbors
// (canonical ids) of the changed balls join the affected set; a ball not yet
// marked takes the mark `value` (every writer of one ball writes the same).
__global__ void k_mark_neighbors(Pipeline p, const int32_t* changed, int64_t k, uint8_t* affected, uint8_t value) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= k * p.nb_cap) return;
  const int64_t r = t / p.nb_cap, j = t - r * p.nb_cap;
  const int32_t id = changed[r];
  if (id < 0 || id >= p.n || j >= p.nb_cnt[id]) return;
  uint8_t* m = affected + canon_id(p, p.nb[static_cast<int64_t>(id) * p.nb_cap + j]);
  if (*m == 0) *m = value;
}

__global__ void k_mark_ids(uint8_t* mask, int64_t n, const int32_t* ids, int64_t k, uint8_t value) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t < k && ids[t] >= 0 && ids[t] < n) mask[ids[t]] = value;
}

__global__ void k_fill_u8(uint8_t* mask, int64_t n, uint8_t value) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n) mask[i] = value;
}

__global__ void k_scatter_balls(vltr_ball* balls, const int32_t* ids, const vltr_ball* vals, int64_t count) {
  const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (k < count) balls[ids[k]] = vals[k];
}

__global__ void k_mask_clear(uint8_t* mask, const uint8_t* where, int64_t n) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n && where[i]) mask[i] = 0;
}

inline int64_t blocks_for(int64_t n, int64_t per) { return (n + per - 1) / per; }

}  // namespace

hipError_t launch_dedup_incoming(const Pipeline& p, hipStream_t s) {
  if (p.n_pairs > 0)
    hipLaunchKernelGGL(k_dedup_incoming, dim3(blocks_for(p.n_pairs, 256)), dim3(256), 0, s, p,
                       dedup_slots_mask(p.n_pairs));
  return hipGetLastError();
}

hipError_t launch_cells(const Pipeline& p, hipStream_t s) {
  if (p.n > 0) hipLaunchKernelGGL(k_cells, dim3(blocks_for(p.n, kCellBalls)), dim3(kCellsBlock), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_flag_contacts(const Pipeline& p, unsigned long long* dd_keys, uint32_t* dd_first, uint32_t* flag,
                                hipStream_t s) {
  const DedupTable T{dd_keys, dd_first, dd_keys != nullptr ? dedup_slots_mask(p.n_pairs) : 0ull};
  if (p.n_pairs > 0)
    hipLaunchKernelGGL(k_flag_contacts, dim3(blocks_for(p.n_pairs, 256)), dim3(256), 0, s, p, T, flag);
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
  if (p.n > 0) hipLaunchKernelGGL(k_emit_cells, dim3(blocks_for(p.n, 256)), dim3(256), 0, s, p, pos, out);
  return hipGetLastError();
}

hipError_t launch_updateable_cells(const Pipeline& p, CellRec* state, uint8_t* excl, const uint8_t* affected, int full,
                                   int refresh, hipStream_t s) {
  if (p.n > 0)
    hipLaunchKernelGGL(k_updateable_cells, dim3(blocks_for(p.n, 256)), dim3(256), 0, s, p, state, excl, affected, full,
                       refresh);
  return hipGetLastError();
}
hipError_t launch_mark_neighbors(const Pipeline& p, const int32_t* changed, int64_t k, uint8_t* affected,
                                 uint8_t value, hipStream_t s) {
  if (k > 0 && p.n > 0)
    hipLaunchKernelGGL(k_mark_neighbors, dim3(blocks_for(k * p.nb_cap, 256)), dim3(256), 0, s, p, changed, k, affected,
                       value);
  return hipGetLastError();
}
hipError_t launch_mark_ids(uint8_t* mask, int64_t n, const int32_t* ids, int64_t k, uint8_t value, hipStream_t s) {
  if (k > 0) hipLaunchKernelGGL(k_mark_ids, dim3(blocks_for(k, 256)), dim3(256), 0, s, mask, n, ids, k, value);
  return hipGetLastError();
}
hipError_t launch_count_marks(const uint8_t* mask, int64_t n, uint32_t* counts, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_two, dim3(1), dim3(64), 0, s, counts);
  const int64_t b = blocks_for(n, 256);
  if (n > 0) hipLaunchKernelGGL(k_count_marks, dim3(b < 1024 ? b : 1024), dim3(256), 0, s, mask, n, counts);
  return hipGetLastError();
}
hipError_t launch_fill_u8(uint8_t* mask, int64_t n, uint8_t value, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill_u8, dim3(blocks_for(n, 256)), dim3(256), 0, s, mask, n, value);
  return hipGetLastError();
}
hipError_t launch_scatter_balls(vltr_ball* balls, const int32_t* ids, const vltr_ball* vals, int64_t count,
                                hipStream_t s) {
  if (count > 0) hipLaunchKernelGGL(k_scatter_balls, dim3(blocks_for(count, 256)), dim3(256), 0, s, balls, ids, vals, count);
  return hipGetLastError();
}

hipError_t launch_mask_clear(uint8_t* mask, const uint8_t* where, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_mask_clear, dim3(blocks_for(n, 256)), dim3(256), 0, s, mask, where, n);
  return hipGetLastError();
}

hipError_t launch_emit_net(const Pipeline& p, const uint32_t* pos, ContactNet* out, hipStream_t s) {
  if (p.n_pairs > 0) hipLaunchKernelGGL(k_emit_net, dim3(blocks_for(p.n_pairs, 256)), dim3(256), 0, s, p, pos, out);
  return hipGetLastError();
}

hipError_t launch_seg_offsets(const Pipeline& p, const uint32_t* pos, int what, uint32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_seg_offsets, dim3(blocks_for(p.n_segs + 1, 256)), dim3(256), 0, s, p, pos, what, out);
  return hipGetLastError();
}

}  // namespace vltr
