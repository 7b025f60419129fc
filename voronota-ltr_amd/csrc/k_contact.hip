// Contact construction (replaces src/contact.rs): per candidate pair (a, b),
// the intersection circle of the two probe-expanded balls is cut by the
// radical planes of a's neighbors in their sorted order, restricted to the
// circle, and reduced to area, arc length, centrality, the two solid angles
// and the two pyramid volumes.
//
// k_clip: one wavefront per ball a, in two phases.
//  1. Classification. Lane k holds neighbor k of a (sphere, radical plane
//     with a) in registers and in the wavefront's shared-memory plane table.
//     For each pair (a, b) of a every lane classifies its neighbor
//     (contact.rs:224-276; the answer depends on (a, b, c) only, never on the
//     contour). A killing neighbor means no contact; no cutting neighbor
//     means the uncut full circle, finished one pair per lane afterwards;
//     otherwise the pair and its 64-bit cutting mask go to the wavefront's
//     list.
//  2. Clipping. The wavefront splits into groups of 8 (or 16) lanes; each
//     group clips one listed pair at a time with group_clip (contour points in
//     lanes, planes read from the shared table), eight (four) pairs side by
//     side. The group size and the wavefronts per block are template
//     parameters of k_clip (its launch shapes, chosen by the autotuner).
// Pairs of balls with more than 64 neighbors, and pairs whose contour
// outgrows its lane group, go to k_clip_wide (16 lanes per pair); what
// outgrows that goes to k_cut_slow: one lane per pair, the one-lane code of
// contour.h with the contour in a global workspace. All paths produce
// identical records.
//
// Updateable steps (p.dirty set) clip only the pairs whose two balls are
// both affected; every other pair keeps the record of the previous run
// (updateable.rs:476-512). A wavefront whose ball is not affected copies its
// records as one block, and an affected ball looks up each kept pair in its
// previous pair list.
#include "contour.h"
#include "group_clip.h"
#include "vltr_internal.h"
#include "wave_group.h"

namespace vltr {

// sin/cos of half the five accumulated hexagon rotation angles
// (contact.rs:455-469), computed once on the host with its libm.
__constant__ double c_hex_s[5];
__constant__ double c_hex_c[5];

namespace {

double h_hex_s[5], h_hex_c[5];

constexpr int kWave = 64;
constexpr int kClipSlots = 64;  // neighbor slots handled by k_clip
// Launch shape of k_clip, picked per context by the autotuner among the
// variants of kClipVariants: lanes per pair in phase 2 (8: eight pairs side
// by side, contours beyond 8 points, about 1% of the clipped pairs, go to
// k_clip_wide's 16-lane groups; 16: four pairs side by side), wavefronts per
// block, the occupancy target in wavefronts per SIMD (4 caps registers at 128
// per lane, with some spills; 3 allows 168) and whether the planes' unit
// normals stay in shared memory (9.7 KB per wavefront without them, 11.3 KB
// with them).

// Unit normals of the planes: kept in shared memory (NN) or recomputed each
// time a plane is applied (the shared memory then allows four wavefronts per
// SIMD).
template <bool NN>
struct UnitNormals {
  __device__ void put(int, V3) {}
  __device__ V3 get(int, V3 pn) const { return unit(pn); }
};
template <>
struct UnitNormals<true> {
  double x[kClipSlots], y[kClipSlots], z[kClipSlots];
  __device__ void put(int k, V3 nn) {
    x[k] = nn.x;
    y[k] = nn.y;
    z[k] = nn.z;
  }
  __device__ V3 get(int k, V3) const { return V3{x[k], y[k], z[k]}; }
};

// Restricted contour capacity of a clip group: twice the group size, except
// 11 for 8-lane groups (contours of 9-11 points are 0.02% of the clipped
// pairs on protein-like input; longer ones go to k_clip_wide), which keeps the
// wavefront's shared memory under 10 KB with the unit normals and the circle
// radii (four wavefronts per SIMD).
template <int G>
constexpr int restrict_cap() {
  return G == 8 ? 11 : 2 * G;
}

// Shared memory of one k_clip wavefront.
template <int G, bool NN>
struct ClipLdsT {
  double pcx[kClipSlots], pcy[kClipSlots], pcz[kClipSlots];
  double pnx[kClipSlots], pny[kClipSlots], pnz[kClipSlots];
  UnitNormals<NN> nn;
  double thr[kClipSlots];  // squared distance from a's centre to the plane point
  // radius of the circle of the pair (a, neighbor k): with the plane point
  // (the circle's centre) and the plane normal (the pair's axis) the pair's
  // geometry, read in phase 2 instead of computed again
  double ric[kClipSlots];
  int32_t cid[kClipSlots];
  // listed pairs: (pair offset within a << 8) | slot of b; pairs to clip
  // from the front (with their cutting masks), uncut pairs from the back
  int32_t item[kClipSlots];
  uint64_t mask[kClipSlots];
  GroupLdsT<restrict_cap<G>()> grp[kWave / G];
};

// Value of v in lane k (k uniform over the wavefront).
__device__ __forceinline__ double readlane_d(double v, int k) {
  const uint64_t u = static_cast<uint64_t>(__double_as_longlong(v));
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(u & 0xffffffffu), k));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(u >> 32), k));
  return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}

__device__ __forceinline__ void push_slow(const Pipeline& p, int64_t pidx) {
  const uint32_t k = atomicAdd(p.overflow_cnt, 1u);
  p.overflow_list[k] = static_cast<int32_t>(pidx);
  atomicOr(p.status, kStatContourOverflow);
}
__device__ __forceinline__ void push_wide(const Pipeline& p, int64_t pidx) {
  const uint32_t k = atomicAdd(p.wide_cnt, 1u);
  p.wide_list[k] = static_cast<int32_t>(pidx);
}

// ---- updateable steps: pairs that keep their previous record
constexpr int kRecWords = static_cast<int>(sizeof(PairRec) / 8);
constexpr int kTopoWords = static_cast<int>(sizeof(ContactTopo) / 8);
static_assert(sizeof(PairRec) % 8 == 0 && sizeof(ContactTopo) % 8 == 0, "records are copied as 8-byte words");
static_assert(kTopoWords <= kWave, "one topology's words fit one wavefront");

__device__ __forceinline__ unsigned long long lanes_below(int lane) {
  return lane == 0 ? 0ull : (~0ull >> (kWave - lane));
}

__device__ __forceinline__ bool is_contact(const PairRec& r) { return (r.flags & kRecValid) && r.area > 0.0; }

// One lane copies a stored topology (word by word: a whole-record copy
// would hold it in registers).
__device__ __forceinline__ void copy_topo(ContactTopo* dst, const ContactTopo* src) {
  const uint64_t* s = reinterpret_cast<const uint64_t*>(src);
  uint64_t* d = reinterpret_cast<uint64_t*>(dst);
#pragma unroll 1
  for (int k = 0; k < kTopoWords; ++k) d[k] = s[k];
}

// A record as four 16-byte streaming stores (the line is written whole and
// is not read again by this kernel, so it need not stay in the L2).
__device__ __forceinline__ void store_record(PairRec* dst, const PairRec& r) {
  using U4 = unsigned int __attribute__((ext_vector_type(4)));
  auto lo = [](double x) { return static_cast<unsigned int>(static_cast<uint64_t>(__double_as_longlong(x)) & 0xffffffffu); };
  auto hi = [](double x) { return static_cast<unsigned int>(static_cast<uint64_t>(__double_as_longlong(x)) >> 32); };
  U4* d = reinterpret_cast<U4*>(dst);
  __builtin_nontemporal_store(U4{lo(r.area), hi(r.area), lo(r.arc_length), hi(r.arc_length)}, d + 0);
  __builtin_nontemporal_store(U4{lo(r.solid_a), hi(r.solid_a), lo(r.solid_b), hi(r.solid_b)}, d + 1);
  __builtin_nontemporal_store(U4{lo(r.pyr_a), hi(r.pyr_a), lo(r.pyr_b), hi(r.pyr_b)}, d + 2);
  __builtin_nontemporal_store(U4{r.flags, r.pad, 0u, 0u}, d + 3);
}
// The record of pair pidx, stored when it holds a contact or when the run
// keeps every pair's record (updateable steps: the next step copies the
// records of the pairs it does not clip). Where a run keeps contact bytes the
// readers ask those which pairs have a contact (has_contact), so the records
// of the other pairs are not written.
__device__ __forceinline__ void put_record(const Pipeline& p, int64_t pidx, const PairRec& rec) {
  if (p.records_all || p.pair_contact == nullptr || is_contact(rec)) store_record(p.rec + pidx, rec);
}

// k topology records of the region of this block's XCD: the first of their
// indices within the region (counted past the region's end).
__device__ __forceinline__ uint32_t topo_reserve(const Pipeline& p, uint32_t k) {
  return atomicAdd(p.topo_counts + (blockIdx.x % kTopoRegions) * kClipQueueStride, k);
}
// The store record of index l within this block's region, or kTopoSlotFull
// past its end (the run is then redone with larger regions).
__device__ __forceinline__ uint32_t topo_record(const Pipeline& p, uint32_t l) {
  return static_cast<int64_t>(l) < p.topo_region
             ? static_cast<uint32_t>((blockIdx.x % kTopoRegions) * p.topo_region + l)
             : kTopoSlotFull;
}

// Position of partner b in the previous run's pair list [ob, ob + on) of the
// same ball, or -1. Every lane of the wavefront takes part (b uniform).
__device__ int64_t prev_find_wave(const Pipeline& p, int64_t ob, int on, int32_t b, int lane) {
  for (int j0 = 0; j0 < on; j0 += kWave) {
    const int j = j0 + lane;
    const uint64_t hit = __ballot(j < on && p.prev_pair_b[ob + j] == b);
    if (hit != 0ull) return ob + j0 + __builtin_ctzll(hit);
  }
  return -1;
}
__device__ int64_t prev_find_lane(const Pipeline& p, int64_t ob, int on, int32_t b) {
  for (int j = 0; j < on; ++j)
    if (p.prev_pair_b[ob + j] == b) return ob + j;
  return -1;
}

// Full runs: a contact of pair pidx = (a, b), clipped or kept from the
// previous run, joins b's incoming list when b is canonical
// (tessellation.rs:624-631: an image end receives nothing; the incoming order
// is free, k_cells puts each list in pair order) and is marked in the pair's
// contact byte. Pairs without a contact are not listed: the cells and the
// solvent kernels read contacts only.
__device__ __forceinline__ void push_incoming(const Pipeline& p, int64_t pidx, int32_t b, const PairRec& rec) {
  if (!is_contact(rec)) return;
  if (p.pair_contact != nullptr) p.pair_contact[pidx] = 1;
  if (p.mode != 0 || b >= p.n) return;
  const uint32_t slot = atomicAdd(&p.in_cnt[b], 1u);
  if (slot < static_cast<uint32_t>(p.nb_cap))
    p.incoming[static_cast<int64_t>(b) * p.nb_cap + slot] = static_cast<int32_t>(pidx);
  else
    atomicOr(p.status, kStatNbOverflow);
}

// Pair `to` = (a, b) takes the previous run's record (and topology) of pair
// `from`, or an empty record when from < 0; the wavefront's lanes copy one
// word each. A kept contact joins b's incoming list, and its stored topology
// is copied to a new record of this run's store.
template <bool kTopo>
__device__ void keep_pair_wave(const Pipeline& p, int64_t from, int64_t to, int32_t b, int lane) {
  if (from < 0) {
    if (lane == 0) p.rec[to] = empty_record();
    return;
  }
  if (lane < kRecWords)
    reinterpret_cast<uint64_t*>(p.rec + to)[lane] = reinterpret_cast<const uint64_t*>(p.prev_rec + from)[lane];
  if (lane == 0) push_incoming(p, to, b, p.prev_rec[from]);
  if (!kTopo || p.topo == nullptr || !is_contact(p.prev_rec[from])) return;
  const uint32_t ps = p.prev_topo_slot[from];
  uint32_t slot = ps;
  if (topo_slot_stored(ps)) {
    uint32_t got = 0u;
    if (lane == 0) got = topo_reserve(p, 1u);
    slot = topo_record(p, static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(got))));
    if (slot != kTopoSlotFull && lane < kTopoWords)
      reinterpret_cast<uint64_t*>(p.topo + slot)[lane] = reinterpret_cast<const uint64_t*>(p.prev_topo + ps)[lane];
  }
  if (lane == 0) p.topo_slot[to] = slot;
}
template <bool kTopo>
__device__ void keep_pair_lane(const Pipeline& p, int64_t from, int64_t to, int32_t b) {
  if (from < 0) {
    p.rec[to] = empty_record();
    return;
  }
  const PairRec rec = p.prev_rec[from];
  p.rec[to] = rec;
  push_incoming(p, to, b, rec);
  if (!kTopo || p.topo == nullptr || !is_contact(rec)) return;
  const uint32_t ps = p.prev_topo_slot[from];
  uint32_t slot = ps;
  if (topo_slot_stored(ps)) {
    slot = topo_record(p, topo_reserve(p, 1u));
    if (slot != kTopoSlotFull) copy_topo(p.topo + slot, p.prev_topo + ps);
  }
  p.topo_slot[to] = slot;
}

// Ball a is not dirty: all its pairs keep their previous records. When its
// pair list is the previous one (the usual case: no neighbor of a moved),
// the records are copied as one block, and the stored topologies of 64
// contacts at a time take consecutive new records (one atomic per 64).
template <bool kTopo>
__device__ void keep_ball(const Pipeline& p, int64_t base, int npairs, int64_t ob, int on, int lane) {
  bool diff = on != npairs;
  if (!diff)
    for (int j = lane; j < npairs; j += kWave) diff = diff || p.pair_b[base + j] != p.prev_pair_b[ob + j];
  if (__ballot(diff) == 0ull) {
    const uint64_t* rs = reinterpret_cast<const uint64_t*>(p.prev_rec + ob);
    uint64_t* rd = reinterpret_cast<uint64_t*>(p.rec + base);
    for (int t = lane; t < npairs * kRecWords; t += kWave) rd[t] = rs[t];
    for (int j = lane; j < npairs; j += kWave) push_incoming(p, base + j, p.prev_pair_b[ob + j], p.prev_rec[ob + j]);
    if (kTopo && p.topo != nullptr) {
      for (int j0 = 0; j0 < npairs; j0 += kWave) {
        const int j = j0 + lane;
        uint32_t ps = kTopoSlotUncut;
        bool stored = false;
        if (j < npairs && is_contact(p.prev_rec[ob + j])) {
          ps = p.prev_topo_slot[ob + j];
          stored = topo_slot_stored(ps);
          if (!stored) p.topo_slot[base + j] = ps;
        }
        const unsigned long long m = __ballot(stored);
        if (m == 0ull) continue;
        uint32_t got = 0u;
        if (lane == 0) got = topo_reserve(p, static_cast<uint32_t>(__popcll(m)));
        got = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(got)));
        if (stored) {
          const uint32_t slot = topo_record(p, got + static_cast<uint32_t>(__popcll(m & lanes_below(lane))));
          if (slot != kTopoSlotFull) copy_topo(p.topo + slot, p.prev_topo + ps);
          p.topo_slot[base + j] = slot;
        }
      }
    }
    return;
  }
  for (int j = 0; j < npairs; ++j) {
    const int32_t b = p.pair_b[base + j];
    keep_pair_wave<kTopo>(p, prev_find_wave(p, ob, on, b, lane), base + j, b, lane);
  }
}

// The pair list of ball a (tessellation.rs:298-319: its qualifying neighbors
// in list order, from index base). The qualifying entries of 64 list slots
// take consecutive indices by a ballot: one coalesced store per 64 slots,
// streamed past the L2 like the records.
__device__ __forceinline__ void write_pairs(const Pipeline& p, int64_t a, int64_t base, int cnt, const int32_t* nbl,
                                            int lane) {
  int64_t j = 0;
  for (int k0 = 0; k0 < cnt; k0 += kWave) {
    const int k = k0 + lane;
    int32_t c = 0;
    bool q = false;
    if (k < cnt) {
      c = nbl[k];
      q = pair_qualifies(p, a, c);
    }
    const unsigned long long m = __ballot(q);
    if (q) {
      const int64_t idx = base + j + __popcll(m & lanes_below(lane));
      if (idx < p.pair_cap) {
        __builtin_nontemporal_store(static_cast<int32_t>(a), p.pair_a + idx);
        __builtin_nontemporal_store(c, p.pair_b + idx);
      } else {
        atomicOr(p.status, kStatPairOverflow);
      }
    }
    j += __popcll(m);
  }
}

__device__ __forceinline__ bool is_dirty(const Pipeline& p, int64_t id) {
  return p.dirty[p.periodic ? canon_id(p, id) : id] != 0;
}

// Geometry of the pair (a, neighbor k) from the plane table: the
// intersection circle's centre is the radical plane's point (circle_center
// and intersection_circle take the same steps), its radius is the table's,
// and the pair's axis is the plane's normal (both unit(c - a)); the pair is
// known to have a circle.
template <class L>
__device__ __forceinline__ PairGeom table_geom(const L& w, int k, const Sph& A, const Sph& B) {
  PairGeom g;
  g.a = A;
  g.b = B;
  g.ic = Sph{V3{w.pcx[k], w.pcy[k], w.pcz[k]}, w.ric[k]};
  g.axis = V3{w.pnx[k], w.pny[k], w.pnz[k]};
  g.ic_dir = V3{0.0, 0.0, 0.0};  // used by the classification only
  g.axis_unit = w.nn.get(k, g.axis);
  return g;
}

// Solid angles and pyramid volumes of a round's contacts (contact.rs:797-847
// and the record's last two terms), the per-vertex turns of every group and
// both spheres spread over the whole wavefront: task q is vertex t of group
// h's contour for sphere a (the group's first c_h tasks) or b (its next c_h),
// groups in order, so a round needs about one pass however the contour sizes
// fall among its groups. The contours stay in the groups' shared buffers.
// Lanes 0 and 1 of a group then add the turns of sphere a and b in vertex
// order, as the one-lane code does. c = the contour size of this lane's group
// when its pair has a contact, else 0; B and icc = its pair's second sphere
// and circle centre; A = the wavefront's ball.
template <int GS, class GL>
__device__ __forceinline__ void wave_solid(const GL* grp, const int lane, const Sph& A, const Sph& B, const V3 icc,
                                           const int c, PairRec* rec) {
  constexpr int NG = kWave / GS;
  int pref[NG + 1], cg[NG];
  pref[0] = 0;
  int cmax = 0;
#pragma unroll
  for (int h = 0; h < NG; ++h) {
    cg[h] = __builtin_amdgcn_readlane(c, h * GS);
    pref[h + 1] = pref[h] + 2 * cg[h];
    cmax = cg[h] > cmax ? cg[h] : cmax;
  }
  const int T = pref[NG];
  if (T == 0) return;
  wave_sync();  // the groups' contours are in shared memory
  const int j = lane & (GS - 1);
  const int hg = lane / GS;
  // lane j of a group offers sphere b's values for odd j, sphere a's for even
  const bool jb = (j & 1) != 0;
  const V3 oc = jb ? B.c : A.c;
  const V3 qc = jb ? A.c : B.c;
  const double nrm = norm(icc - oc);
  const double ratio = nrm / (jb ? B.r : A.r);
  int p_own = 0;
#pragma unroll
  for (int h = 0; h < NG; ++h) p_own = hg == h ? pref[h] : p_own;
  // lanes 0 and 1 of the group sum the tasks [lo, hi) of sphere a and b
  const int my_lo = p_own + (jb ? c : 0);
  const int my_hi = j < 2 ? my_lo + c : my_lo;
  double acc = 0.0;
  for (int base = 0; base < T; base += kWave) {
    const int q = base + lane;
    // the group whose tasks hold q: the last one starting at or before q (an
    // empty group starts where the next one does)
    int h = 0;
#pragma unroll
    for (int k = 1; k < NG; ++k) h = q >= pref[k] ? k : h;
    int ph = 0, ch = 0;
#pragma unroll
    for (int k = 0; k < NG; ++k) {
      ph = h == k ? pref[k] : ph;
      ch = h == k ? cg[k] : ch;
    }
    const int rel = q - ph;
    const bool sb = rel >= ch;
    const int t = sb ? rel - ch : rel;
    const int src = q < T ? h * GS + (sb ? 1 : 0) : lane;
    const V3 bc{__shfl(B.c.x, src, kWave), __shfl(B.c.y, src, kWave), __shfl(B.c.z, src, kWave)};
    const V3 ic{__shfl(icc.x, src, kWave), __shfl(icc.y, src, kWave), __shfl(icc.z, src, kWave)};
    const double r_src = __shfl(ratio, src, kWave);
    double t1 = 0.0, t2 = 0.0;
    if (q < T) turn_terms(sb ? bc : A.c, sb ? A.c : bc, ic, r_src, grp + h, t, ch, &t1, &t2);
    const int lo = my_lo > base ? my_lo : base;
    const int hi = my_hi < base + kWave ? my_hi : base + kWave;
    for (int i = 0; i < cmax; ++i) {
      const int sq = lo + i;
      const int sl = (sq - base) & (kWave - 1);
      const double v1 = __shfl(t1, sl, kWave), v2 = __shfl(t2, sl, kWave);
      if (sq < hi) {
        acc = acc + v1;
        acc = acc + v2;
      }
    }
  }
  const double solid = signed_solid(acc, oc, qc, icc);
  const double pyr = nrm * rec->area / 3.0 * (solid < 0.0 ? -1.0 : 1.0);
  const int src_b = (lane & ~(GS - 1)) | 1;
  const double solid_b = __shfl(solid, src_b, kWave), pyr_b = __shfl(pyr, src_b, kWave);
  if (c > 0 && j == 0) {
    rec->solid_a = solid;
    rec->pyr_a = pyr;
    rec->solid_b = solid_b;
    rec->pyr_b = pyr_b;
  }
  wave_sync();  // read before the next round's groups write their contours
}

// One clip task of a wavefront: ball q (in cell order), share wsub of its
// qualifying pairs when kWavesPerBall waves split the ball (pair jj of the
// ball goes to wave jj % kWavesPerBall; each of them builds the plane table,
// classifies and clips only its own pairs). Updateable steps (p.dirty set)
// run whole balls in share 0.
template <int kClipGroup, bool kKeepNormals, int kWavesPerBall, bool kOut>
__device__ __forceinline__ void clip_ball(const Pipeline& p, ClipLdsT<kClipGroup, kKeepNormals>& w, const int lane,
                                          const int64_t q, const int wsub) {
  using ClipGroup = WaveGroupT<kClipGroup>;
  using ClipGroupLds = GroupLdsT<restrict_cap<kClipGroup>()>;
  using ClipLds = ClipLdsT<kClipGroup, kKeepNormals>;
  const int wpb = p.dirty != nullptr ? 1 : kWavesPerBall;
  if (wsub >= wpb) return;
  const int64_t a = p.periodic ? q : static_cast<int64_t>(p.sorted_id[q]);
  const int64_t base = p.pair_cnt[a];
  const int npairs = static_cast<int>(static_cast<int64_t>(p.pair_cnt[a + 1]) - base);
  if (npairs <= 0) return;
  const int cnt = p.nb_cnt[a];
  const int32_t* nbl = p.nb + a * static_cast<int64_t>(p.nb_cap);
  const double* sph4 = reinterpret_cast<const double*>(p.sph);
  const bool incremental = p.dirty != nullptr;
  if (wsub == 0) {
    write_pairs(p, a, base, cnt, nbl, lane);
    if (incremental) {
      // the keep paths below read this ball's pair list back: its stores
      // complete before those loads
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
  }
  int64_t ob = 0;
  int on = 0;
  int32_t prev_b = -1;
  if (incremental) {
    ob = p.prev_pair_cnt[a];
    on = static_cast<int>(static_cast<int64_t>(p.prev_pair_cnt[a + 1]) - ob);
    if (!p.dirty[a]) {
      keep_ball<kOut>(p, base, npairs, ob, on, lane);
      return;
    }
    // the previous partners, one per lane, for the kept pairs' look-ups
    if (lane < on) prev_b = p.prev_pair_b[ob + lane];
  }
  if (cnt > kClipSlots) {
    if (wsub != 0) return;
    for (int j = lane; j < npairs; j += kWave) {
      const int32_t b = p.pair_b[base + j];
      if (incremental && !is_dirty(p, b))
        keep_pair_lane<kOut>(p, prev_find_lane(p, ob, on, b), base + j, b);
      else
        push_wide(p, base + j);
    }
    return;
  }

  // ---- phase 1: classification
  // Lane k holds neighbor k and the geometry of the pair (a, neighbor k)
  // (contact.rs:202-218); each pair's geometry is broadcast from its lane.
  const Sph A = load_sph(sph4, a);
  int32_t c0 = -1;
  Sph C{};
  Plane P{};
  bool c_has_a = false;
  PairGeom gk{};
  bool gk_ok = false;
  if (lane < cnt) {
    c0 = nbl[lane];
    C = load_sph(sph4, c0);
    double thr = 0.0;
    gk_ok = plane_pair_setup(A, C, &P, &gk, &c_has_a, &thr);
    w.pcx[lane] = P.pc.x;
    w.pcy[lane] = P.pc.y;
    w.pcz[lane] = P.pc.z;
    w.pnx[lane] = P.pn.x;
    w.pny[lane] = P.pn.y;
    w.pnz[lane] = P.pn.z;
    if (kKeepNormals) w.nn.put(lane, unit(P.pn));
    w.thr[lane] = thr;
    w.ric[lane] = gk.ic.r;
    w.cid[lane] = c0;
  }
  // the rest of the pair geometries, staged in the clip groups' shared
  // memory (unused until phase 2) for uniform-address reads in the loop
  // below; the circle is the table's plane point and radius
  // (intersection_circle and circle_center take the same steps)
  static_assert(sizeof(double) * 6 * kWave <= sizeof(w.grp), "pair geometry fits the clip groups' buffers");
  double* pg = reinterpret_cast<double*>(&w.grp[0]);
  if (lane < cnt) {
    pg[0 * kWave + lane] = gk.ic_dir.x;
    pg[1 * kWave + lane] = gk.ic_dir.y;
    pg[2 * kWave + lane] = gk.ic_dir.z;
    pg[3 * kWave + lane] = C.c.x;
    pg[4 * kWave + lane] = C.c.y;
    pg[5 * kWave + lane] = C.c.z;
  }
  wave_sync();
  const uint64_t ok_mask = __ballot(gk_ok);
  // The ball's pairs in list order (tessellation.rs:298-319): slot k holds a
  // pair when its neighbor qualifies; the pair's offset jj within a's pairs
  // is the number of qualifying slots before k. Each lane tests its own
  // neighbor, so the loops below walk wavefront-uniform masks.
  bool qualifies = false;
  if (lane < cnt) qualifies = pair_qualifies(p, a, c0);
  const uint64_t qual = __ballot(qualifies);
  const int jj_lane = __popcll(qual & lanes_below(lane));
  const bool mine = qualifies && (wpb == 1 || jj_lane % wpb == wsub);
  const uint64_t kept = incremental ? __ballot(mine && !is_dirty(p, c0)) : 0ull;
  const uint64_t todo = __ballot(mine) & ~kept;
  // pairs of an updateable step whose two balls are not both affected keep
  // their previous records
  for (uint64_t r = kept; r != 0ull; r &= r - 1) {
    const int k = __builtin_ctzll(r);
    const int jj = __popcll(qual & lanes_below(k));
    const int32_t b = __builtin_amdgcn_readlane(c0, k);
    int64_t from;
    if (on <= kWave) {
      const uint64_t hit = __ballot(prev_b == b);
      from = hit != 0ull ? ob + __builtin_ctzll(hit) : -1;
    } else {
      from = prev_find_wave(p, ob, on, b, lane);
    }
    keep_pair_wave<kOut>(p, from, base + jj, b, lane);
  }
  // pairs that cannot have a contact (no intersection circle)
  if (((todo & ~ok_mask) >> lane) & 1ull) put_record(p, base + jj_lane, empty_record());
  int n_clip = 0, n_uncut = 0;
  for (uint64_t r = todo & ok_mask; r != 0ull; r &= r - 1) {
    const int k = __builtin_ctzll(r);
    const int jj = __popcll(qual & lanes_below(k));
    // classify reads the circle, its direction from a, and b
    PairGeom g;
    g.ic = Sph{V3{w.pcx[k], w.pcy[k], w.pcz[k]}, w.ric[k]};
    g.ic_dir = V3{pg[0 * kWave + k], pg[1 * kWave + k], pg[2 * kWave + k]};
    g.b = Sph{V3{pg[3 * kWave + k], pg[4 * kWave + k], pg[5 * kWave + k]}, readlane_d(C.r, k)};
    const int cl = (lane < cnt && lane != k) ? classify(g, C, P, c_has_a) : kSkip;
    const bool kill = __ballot(cl == kKill) != 0ull;
    const uint64_t M = __ballot(cl == kCut);
    if (kill) {
      if (lane == 0) put_record(p, base + jj, empty_record());
    } else if (M == 0ull) {
      if (lane == 0) w.item[kClipSlots - 1 - n_uncut] = (jj << 8) | k;
      ++n_uncut;
    } else {
      if (lane == 0) {
        w.item[n_clip] = (jj << 8) | k;
        w.mask[n_clip] = M;
      }
      ++n_clip;
    }
  }
  wave_sync();

  // uncut pairs: the full circle, one pair per lane
  if (lane < n_uncut) {
    const int it = w.item[kClipSlots - 1 - lane];
    const int32_t b = w.cid[it & 0xff];
    const PairGeom g = table_geom(w, it & 0xff, A, load_sph(sph4, b));
    const PairRec rec = record_uncut(g, p.mode);
    store_record(p.rec + base + (it >> 8), rec);
    push_incoming(p, base + (it >> 8), b, rec);
    if (kOut && p.topo != nullptr) p.topo_slot[base + (it >> 8)] = kTopoSlotUncut;
    if (kOut && p.net != nullptr) p.net[base + (it >> 8)].m = 0;
  }

  // Pairs to clip in decreasing cutting-set size, so that the four pairs
  // clipped side by side need about as many cuts (any order gives the same
  // records).
  if (n_clip > 1) {
    int it = 0, rank = 0;
    uint64_t mk = 0;
    if (lane < n_clip) {
      it = w.item[lane];
      mk = w.mask[lane];
      const int pc = __popcll(mk);
      for (int i = 0; i < n_clip; ++i) {
        const int pi = __popcll(w.mask[i]);
        rank += (pi > pc || (pi == pc && i < lane)) ? 1 : 0;
      }
    }
    wave_sync();
    if (lane < n_clip) {
      w.item[rank] = it;
      w.mask[rank] = mk;
    }
    wave_sync();
  }

  // one topology record of this XCD's region per pair to clip, reserved at
  // once (adjoint runs)
  uint32_t tbase = 0u;
  if (kOut && p.topo != nullptr && n_clip > 0) {
    uint32_t got = 0u;
    if (lane == 0) got = topo_reserve(p, static_cast<uint32_t>(n_clip));
    tbase = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(got)));
  }

  // ---- phase 2: kWave / kClipGroup pairs at a time, kClipGroup lanes each
  const int grp = lane / kClipGroup;
  ClipGroup G{grp * kClipGroup, lane & (kClipGroup - 1)};
  ClipGroupLds* gs = &w.grp[grp];
  // unit normals: from shared memory, or recomputed per applied plane
  struct {
    const ClipLds& w;
    __device__ double thr(int s) const { return w.thr[s]; }
    __device__ CutPlane plane(int s) const {
      const V3 pn{w.pnx[s], w.pny[s], w.pnz[s]};
      return CutPlane{V3{w.pcx[s], w.pcy[s], w.pcz[s]}, pn, w.nn.get(s, pn), w.cid[s]};
    }
  } fetch{w};
  for (int r0 = 0; r0 < n_clip; r0 += kWave / kClipGroup) {
    const int idx = r0 + grp;
    const bool act = idx < n_clip;
    PairRec rec = empty_record();
    int st = kNone, m = 0;
    int32_t b = static_cast<int32_t>(a);
    int64_t pidx = 0;
    int slot = 0;
    if (act) {
      const int it = w.item[idx];
      const uint64_t M = w.mask[idx];
      slot = it & 0xff;
      b = w.cid[slot];
      const PairGeom g = table_geom(w, slot, A, load_sph(sph4, b));
      pidx = base + (it >> 8);
      const TopoOut tout{p.topo, p.topo != nullptr ? p.topo_slot + pidx : nullptr,
                         p.topo != nullptr ? topo_record(p, tbase + static_cast<uint32_t>(idx)) : kTopoSlotFull, sph4,
                         p.net != nullptr ? p.net + pidx : nullptr};
      const bool want = kOut && (p.topo != nullptr || p.net != nullptr);
      rec = group_clip(G, gs, g, static_cast<int32_t>(a), M, 0ull, fetch, c_hex_s, c_hex_c, p.mode, &st,
                       want ? &tout : nullptr, &m);
    }
    // the pair's second sphere and circle centre read again rather than
    // kept in registers through the clip
    if (p.mode == 0)
      wave_solid<kClipGroup>(w.grp, lane, A, load_sph(sph4, b), V3{w.pcx[slot], w.pcy[slot], w.pcz[slot]},
                             st == kContact ? m : 0, &rec);
    if (act && G.lane() == 0) {
      if (st == kOverflow) {
        push_wide(p, pidx);
      } else {
        put_record(p, pidx, rec);
        push_incoming(p, pidx, b, rec);
      }
    }
  }
}

// One wavefront per clip task, tasks dealt by block index (remapped so
// that each XCD walks a contiguous range of balls).
template <int kClipGroup, int kClipWaves, int kWavesPerEu, bool kKeepNormals, int kWavesPerBall, bool kOut>
__global__ __launch_bounds__(kClipWaves* kWave) __attribute__((amdgpu_waves_per_eu(kWavesPerEu, kWavesPerEu))) void k_clip(Pipeline p) {
  __shared__ ClipLdsT<kClipGroup, kKeepNormals> lds[kClipWaves];
  const int wid = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t t = xcd_remap(blockIdx.x, gridDim.x) * kClipWaves + wid;
  if (t >= p.n * kWavesPerBall || pairs_overflowed(p)) return;
  clip_ball<kClipGroup, kKeepNormals, kWavesPerBall, kOut>(p, lds[wid], lane, t / kWavesPerBall,
                                                     static_cast<int>(t % kWavesPerBall));
}

// Persistent form: as many single-wavefront blocks as the GPU holds at once,
// each pulling clip tasks from a device work queue until it is empty. The
// tasks (all balls of all structures of the call, times kWavesPerBall) are
// split into kClipQueues contiguous ranges, one per XCD, each with its own
// counter on its own cache line (block b runs on XCD b % 8): a wavefront
// takes one task at a time from its XCD's range, then from the other ranges
// in turn. Every wavefront leaves once all ranges are taken.
template <int kClipGroup, int kWavesPerEu, bool kKeepNormals, int kWavesPerBall, bool kOut>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(kWavesPerEu, kWavesPerEu))) void k_clip_queue(Pipeline p) {
  __shared__ ClipLdsT<kClipGroup, kKeepNormals> lds;
  const int lane = threadIdx.x;
  if (pairs_overflowed(p)) return;
  const int64_t tasks = p.n * kWavesPerBall;
  // one task per claim; the only state carried across a task is the queue
  // being drained (v steps through the XCD's own range, then the others)
  int v = 0;
  while (v < kClipQueues) {
    const int x = (static_cast<int>(blockIdx.x % kClipQueues) + v) % kClipQueues;
    const int64_t lo = tasks * x / kClipQueues, hi = tasks * (x + 1) / kClipQueues;
    uint32_t got = 0;
    if (lane == 0) got = atomicAdd(p.clip_queue + x * kClipQueueStride, 1u);
    const int64_t t = lo + static_cast<int64_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(got)));
    if (t >= hi) {
      ++v;
      continue;
    }
    clip_ball<kClipGroup, kKeepNormals, kWavesPerBall, kOut>(p, lds, lane, t / kWavesPerBall,
                                                       static_cast<int>(t % kWavesPerBall));
    wave_sync();  // the table of this task is read before the next one's is written
  }
}

// Pairs listed by k_clip (balls with more than 64 neighbors, and contours
// that outgrew its 8-lane groups): one 16-lane group per pair, four groups per
// single-wavefront block. The group classifies the pair's neighbors itself,
// 16 at a time (the first killing neighbor ends the pair), and keeps each
// cutting neighbor's plane (point, normal, threshold, id) in its shared
// memory, compacted in list order, so the clip walks a dense mask over that
// list without touching global memory. Pairs with more than kWidePlanes
// cutting neighbors classify them again into a cutting mask (one bit per
// neighbor slot, up to kNbMax, in the same shared memory) and compute each
// plane when it is needed.
// Contours beyond 16 points go to k_cut_slow. A fixed grid walks the
// device-side count.
constexpr int kWideBlocks = 2048;
constexpr int kWideGroups = kWave / kGroup;  // 16-lane groups per block
constexpr int kWidePlanes = 48;              // cutting planes a group keeps
constexpr int kWideWords = kNbMax / 64;

struct WideMask {
  const uint64_t* w;
  __device__ uint64_t operator()(int k) const { return w[k]; }
};

union WidePlanes {
  struct {
    double pcx[kWidePlanes], pcy[kWidePlanes], pcz[kWidePlanes];
    double pnx[kWidePlanes], pny[kWidePlanes], pnz[kWidePlanes];
    double thr[kWidePlanes];
    int32_t cid[kWidePlanes];
  };
  uint64_t cw[kWideWords];  // cutting mask of a pair with more cutting neighbors
};
static_assert(sizeof(uint64_t) * kWideWords <= sizeof(double) * 7 * kWidePlanes, "the mask overlays the planes");

__global__ __launch_bounds__(kWave) void k_clip_wide(Pipeline p) {
  __shared__ GroupLds grp[kWideGroups];
  __shared__ WidePlanes planes[kWideGroups];
  const int gl = threadIdx.x / kGroup;
  const int ln = threadIdx.x & (kGroup - 1);
  WaveGroup G{gl * kGroup, ln};
  GroupLds* gs = &grp[gl];
  WidePlanes& pl = planes[gl];
  uint64_t* cw = pl.cw;
  const double* sph4 = reinterpret_cast<const double*>(p.sph);
  const int64_t n_wide = *p.wide_cnt;
  const int64_t groups = static_cast<int64_t>(gridDim.x) * kWideGroups;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kWideGroups + gl; t < n_wide; t += groups) {
    const int64_t pidx = p.wide_list[t];
    const int32_t a = p.pair_a[pidx], b = p.pair_b[pidx];
    const int cnt = p.nb_cnt[a];
    const int32_t* nbl = p.nb + static_cast<int64_t>(a) * p.nb_cap;
    if (cnt > kNbMax) {
      if (ln == 0) push_slow(p, pidx);
      continue;
    }
    const Sph A = load_sph(sph4, a);
    PairGeom g;
    if (!pair_setup(A, load_sph(sph4, b), &g)) {
      if (ln == 0) put_record(p, pidx, empty_record());
      continue;
    }
    bool kill = false;
    int ncut = 0;  // cutting neighbors so far (those beyond kWidePlanes are counted, not kept)
    for (int s0 = 0; s0 < cnt; s0 += kGroup) {
      const int sl = s0 + ln;
      int cl = kSkip;
      Plane P{};
      int32_t c = -1;
      if (sl < cnt) {
        c = nbl[sl];
        if (c != b) {
          const Sph C = load_sph(sph4, c);
          P = plane_of(A, C);
          cl = classify(g, C, P, contains(C, A));
        }
      }
      if (G.ballot(cl == kKill) != 0u) {
        kill = true;
        break;
      }
      const uint32_t bits = G.ballot(cl == kCut);
      if (cl == kCut) {
        const int idx = ncut + popc32(bits & ((1u << ln) - 1u));
        if (idx < kWidePlanes) {
          pl.pcx[idx] = P.pc.x;
          pl.pcy[idx] = P.pc.y;
          pl.pcz[idx] = P.pc.z;
          pl.pnx[idx] = P.pn.x;
          pl.pny[idx] = P.pn.y;
          pl.pnz[idx] = P.pn.z;
          pl.thr[idx] = norm2(P.pc - A.c);
          pl.cid[idx] = c;
        }
      }
      ncut += popc32(bits);
    }
    G.sync();
    if (kill) {
      if (ln == 0) put_record(p, pidx, empty_record());
      continue;
    }
    if (ncut == 0) {
      if (ln == 0) {
        const PairRec rec = record_uncut(g, p.mode);
        p.rec[pidx] = rec;
        push_incoming(p, pidx, b, rec);
        if (p.topo != nullptr) p.topo_slot[pidx] = kTopoSlotUncut;
        if (p.net != nullptr) p.net[pidx].m = 0;
      }
      continue;
    }
    // the kept planes in list order: bit k of the dense mask = kept plane k
    struct {
      const WidePlanes& pl;
      __device__ double thr(int k) const { return pl.thr[k]; }
      __device__ CutPlane plane(int k) const {
        const V3 pn{pl.pnx[k], pl.pny[k], pl.pnz[k]};
        return CutPlane{V3{pl.pcx[k], pl.pcy[k], pl.pcz[k]}, pn, unit(pn), pl.cid[k]};
      }
    } fetch{pl};
    // more cutting neighbors than kept: slot k of a's list, its plane computed
    struct {
      const double* sph4;
      const int32_t* nbl;
      Sph A;
      __device__ double thr(int k) const { return norm2(circle_center(A, load_sph(sph4, nbl[k])) - A.c); }
      __device__ CutPlane plane(int k) const {
        const Plane q = plane_of(A, load_sph(sph4, nbl[k]));
        return CutPlane{q.pc, q.pn, unit(q.pn), nbl[k]};
      }
    } fetch_slots{sph4, nbl, A};
    const bool kept = ncut <= kWidePlanes;
    if (!kept) {
      // the cutting mask over a's slots, in place of the planes (no neighbor
      // kills: the pass above found none)
      uint64_t word = 0;
      for (int s0 = 0; s0 < cnt; s0 += kGroup) {
        const int sl = s0 + ln;
        int cl = kSkip;
        if (sl < cnt && nbl[sl] != b) {
          const Sph C = load_sph(sph4, nbl[sl]);
          cl = classify(g, C, plane_of(A, C), contains(C, A));
        }
        word |= static_cast<uint64_t>(G.ballot(cl == kCut)) << (s0 & 63);
        if ((s0 & 63) + kGroup == 64 || s0 + kGroup >= cnt) {
          if (ln == 0) cw[s0 >> 6] = word;
          word = 0;
        }
      }
      G.sync();
    }
    const uint64_t dense = ncut >= 64 ? ~0ull : ((1ull << ncut) - 1ull);
    int st = kNone;
    uint32_t trec = kTopoSlotFull;
    if (p.topo != nullptr) {
      uint32_t got = 0u;
      if (ln == 0) got = topo_reserve(p, 1u);
      trec = topo_record(p, static_cast<uint32_t>(G.shfl(static_cast<int32_t>(got), 0)));
    }
    const TopoOut tout{p.topo, p.topo != nullptr ? p.topo_slot + pidx : nullptr, trec, sph4,
                       p.net != nullptr ? p.net + pidx : nullptr};
    const bool want = p.topo != nullptr || p.net != nullptr;
    const PairRec rec =
        kept ? group_clip(G, gs, g, a, dense, 0ull, fetch, c_hex_s, c_hex_c, p.mode, &st, want ? &tout : nullptr)
             : group_clip_words(G, gs, g, a, (cnt + 63) / 64, WideMask{cw}, fetch_slots, c_hex_s, c_hex_c, p.mode,
                                &st, want ? &tout : nullptr);
    G.sync();  // the group's planes, mask and contour are read before the next pair writes them
    if (ln == 0) {
      if (st == kOverflow) {
        push_slow(p, pidx);
      } else {
        put_record(p, pidx, rec);
        push_incoming(p, pidx, b, rec);
      }
    }
  }
}

// One lane per pair, the one-lane clip (contour.h) with the contour in a
// global workspace region per lane, sized for the longest possible contour
// (6 + nb_cap points while cutting, twice that after the restriction). A
// fixed grid of p.slow_ws_pairs lanes walks the device-side count.
__global__ __launch_bounds__(64) void k_cut_slow(Pipeline p) {
  const int64_t t0 = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  const int64_t count = *p.overflow_cnt;
  const int cc = 6 + p.nb_cap;
  const int cr = 2 * cc;
  double* ws = p.slow_ws + t0 * static_cast<int64_t>(5 * cc + 6 * cr);
  GlobalCut ct{cc, ws, ws + cc, ws + 2 * cc, reinterpret_cast<int32_t*>(ws + 3 * cc),
               reinterpret_cast<int32_t*>(ws + 4 * cc)};
  double* w2 = ws + 5 * cc;
  GlobalArc out{cr, w2, w2 + cr, w2 + 2 * cr, w2 + 3 * cr, reinterpret_cast<int32_t*>(w2 + 4 * cr),
                reinterpret_cast<int32_t*>(w2 + 5 * cr)};
  for (int64_t t = t0; t < count; t += p.slow_ws_pairs) {
    const int64_t pi = p.overflow_list[t];
    const int32_t a = p.pair_a[pi], b = p.pair_b[pi];
    const int32_t* nbl = p.nb + static_cast<int64_t>(a) * p.nb_cap;
    const double* sph4 = reinterpret_cast<const double*>(p.sph);
    ClipOut co;
    const int st = clip(sph4, a, b, nbl, p.nb_cnt[a], c_hex_s, c_hex_c, ct, out, &co);
    const PairRec rec = record_of(st, co, load_sph(sph4, a), load_sph(sph4, b), out, p.mode);
    put_record(p, pi, rec);
    push_incoming(p, pi, b, rec);
    if (p.topo != nullptr && st == kContact) {
      const uint32_t slot = co.m > kTopoCap ? kTopoSlotTooLong : topo_record(p, topo_reserve(p, 1u));
      p.topo_slot[pi] = slot;
      if (topo_slot_stored(slot)) {
        ContactTopo& tp = p.topo[slot];
        uint32_t roots = 0u;
        for (int k = 0; k < co.m; ++k) {
          tp.l[k] = out.l[k];
          tp.r[k] = out.r[k];
          if (out.l[k] == a || out.r[k] == a) {
            const int32_t c = out.l[k] == a ? out.r[k] : out.l[k];
            if (circle_root(out.p(k), load_sph(sph4, a).c, load_sph(sph4, b).c, load_sph(sph4, c).c)) roots |= 1u << k;
          }
        }
        tp.roots = roots;
        tp.m = co.m;
      }
    }
    if (p.net != nullptr && st == kContact) {
      ContactNet& nt = p.net[pi];
      if (co.m > kTopoCap) {
        nt.m = kTopoTooLong;
      } else {
        for (int k = 0; k < co.m; ++k) {
          nt.l[k] = out.l[k];
          nt.r[k] = out.r[k];
          nt.x[k] = out.x[k];
          nt.y[k] = out.y[k];
          nt.z[k] = out.z[k];
          nt.ang[k] = out.ang[k];
        }
        nt.m = co.m;
        nt.r_ic = co.ic.r;
      }
    }
  }
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

namespace {
struct ClipVariant {
  const char* name;
  int waves;   // wavefronts per block
  bool queue;  // persistent blocks on the device work queue
  int wpb;     // wavefronts per ball
  void (*kernel)(Pipeline);
  // the same launch shape writing cut topologies and contours (adjoint and
  // network runs); the plain kernel carries none of that code
  void (*kernel_out)(Pipeline);
};
// Launch shapes: lanes per pair in phase 2 (g8/g16), wavefronts per block
// (w), the occupancy target in wavefronts per SIMD (o), unit normals kept in
// shared memory (n), wavefronts per ball (b2), and the persistent form on the
// device work queue (q). The first is the default: the fastest in the
// autotuner's passes on the 100k-ball plain workload (pipeline 2.93 ms
// against 3.05 for g8w1o4q and 3.06 for g8w1o4n, the fastest one-shot
// launch) and level with g8w1o4q on the mixed batch of 4,096 200-20000-ball
// structures (458 ms). One wavefront per block lets each ball's wavefront
// release its shared memory as soon as it ends; keeping the unit normals
// costs 1.5 KB of it (3.5 wavefronts per SIMD) and saves their recomputation
// at every applied plane; two wavefronts per ball build the plane table twice
// and leave the clip rounds half full (3.7 ms).
const ClipVariant kClipVariants[] = {
    {"g8w1o4nq", 1, true, 1, k_clip_queue<8, 4, true, 1, false>, k_clip_queue<8, 4, true, 1, true>},
    {"g8w1o4q", 1, true, 1, k_clip_queue<8, 4, false, 1, false>, k_clip_queue<8, 4, false, 1, true>},
    {"g8w1o4n", 1, false, 1, k_clip<8, 1, 4, true, 1, false>, k_clip<8, 1, 4, true, 1, true>},
    {"g8w1o4", 1, false, 1, k_clip<8, 1, 4, false, 1, false>, k_clip<8, 1, 4, false, 1, true>},
    {"g8w2o4", 2, false, 1, k_clip<8, 2, 4, false, 1, false>, k_clip<8, 2, 4, false, 1, true>},
    {"g8w4o4", 4, false, 1, k_clip<8, 4, 4, false, 1, false>, k_clip<8, 4, 4, false, 1, true>},
    {"g8w4o3", 4, false, 1, k_clip<8, 4, 3, false, 1, false>, k_clip<8, 4, 3, false, 1, true>},
    {"g16w2o4", 2, false, 1, k_clip<16, 2, 4, false, 1, false>, k_clip<16, 2, 4, false, 1, true>},
    {"g16w4o4", 4, false, 1, k_clip<16, 4, 4, false, 1, false>, k_clip<16, 4, 4, false, 1, true>},
    {"g8w1o3n", 1, false, 1, k_clip<8, 1, 3, true, 1, false>, k_clip<8, 1, 3, true, 1, true>},
    {"g8w1o4b2", 1, false, 2, k_clip<8, 1, 4, false, 2, false>, k_clip<8, 1, 4, false, 2, true>},
    {"g8w1o4qb2", 1, true, 2, k_clip_queue<8, 4, false, 2, false>, k_clip_queue<8, 4, false, 2, true>},
    {"g8w1o3q", 1, true, 1, k_clip_queue<8, 3, false, 1, false>, k_clip_queue<8, 3, false, 1, true>},
    {"g8w1o3nq", 1, true, 1, k_clip_queue<8, 3, true, 1, false>, k_clip_queue<8, 3, true, 1, true>},
};
constexpr int kNumClipVariants = static_cast<int>(sizeof(kClipVariants) / sizeof(kClipVariants[0]));

// Blocks of a persistent variant: what the device holds at once (occupancy
// of the kernel times the compute units), computed once per variant.
int64_t resident_blocks(int v, bool out) {
  static int64_t cache[2][kNumClipVariants] = {};
  int64_t& c = cache[out ? 1 : 0][v];
  if (c > 0) return c;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  const void* k = reinterpret_cast<const void*>(out ? kClipVariants[v].kernel_out : kClipVariants[v].kernel);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kClipVariants[v].waves * kWave, 0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 16;
  c = static_cast<int64_t>(cus) * per_cu;
  return c;
}
}  // namespace

int clip_variants() { return kNumClipVariants; }
const char* clip_variant_name(int v) { return v >= 0 && v < kNumClipVariants ? kClipVariants[v].name : nullptr; }

hipError_t launch_clip(const Pipeline& p, hipStream_t s) {
  if (p.n <= 0 || p.n_pairs <= 0) return hipSuccess;
  const int vi = p.clip_variant >= 0 && p.clip_variant < kNumClipVariants ? p.clip_variant : 0;
  const ClipVariant& v = kClipVariants[vi];
  const int64_t tasks = p.n * v.wpb;
  int64_t blocks = (tasks + v.waves - 1) / v.waves;
  const bool out = p.topo != nullptr || p.net != nullptr;
  if (v.queue) blocks = blocks < resident_blocks(vi, out) ? blocks : resident_blocks(vi, out);
  hipLaunchKernelGGL(out ? v.kernel_out : v.kernel, dim3(blocks), dim3(v.waves * kWave), 0, s, p);
  return hipGetLastError();
}

int64_t slow_lanes(int nb_cap) {
  // workspace of at most 2^23 doubles (64 MiB), 256..4096 lanes
  const int64_t per = 5 * (6 + nb_cap) + 6 * 2 * (6 + nb_cap);
  int64_t lanes = (int64_t(1) << 23) / per;
  lanes = lanes < 256 ? 256 : (lanes > 4096 ? 4096 : lanes);
  return lanes / 64 * 64;
}

hipError_t launch_fallbacks(const Pipeline& p, hipStream_t s) {
  if (p.n <= 0 || p.n_pairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_clip_wide, dim3(kWideBlocks), dim3(kWave), 0, s, p);
  hipLaunchKernelGGL(k_cut_slow, dim3(p.slow_ws_pairs / 64), dim3(64), 0, s, p);
  return hipGetLastError();
}

}  // namespace vltr
