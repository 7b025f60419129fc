// Cell list and neighbor lists (replaces src/spheres_searcher.rs).
//
// The reference bins spheres in a uniform grid of edge max(2R + 0.25)
// (spheres_searcher.rs:62-101) and, per ball, scans the 27 surrounding cells,
// keeps intersecting candidates (hidden-ball rules at :283-310) and sorts them
// by (distance to the intersection-circle centre, index) (:253-326,
// types.rs:148-184). Any two intersecting spheres are less than one cell edge
// apart, so the candidate set does not depend on how cells are stored; here
// cells are hashed: a sphere's bucket is the Morton code of its cell
// coordinates taken modulo 2^k per axis (k >= 2, so the 27 cells around any
// cell fall in 27 distinct buckets). Buckets are built with a counting sort
// (atomic histogram, exclusive scan, scatter) so each bucket is a contiguous,
// spatially coherent run of the sorted sphere array.
#include "vltr_internal.h"

namespace vltr {
namespace {

constexpr int kWave = 64;
constexpr int kNbBlock = 64;                  // one wavefront per block
constexpr int kNbWaves = kNbBlock / kWave;
constexpr int kListCap = 256;                 // per-wave LDS list (4 KB: eight wavefronts per SIMD)
// Lists up to this length are put in order by counting ranks (about eight
// instructions per entry per lane); longer ones of up to a wavefront by a
// bitonic sort across the lanes (21 compare-exchange steps)
constexpr int kRankMax = 24;

__device__ __forceinline__ unsigned long long lanemask_lt() {
  const int lane = threadIdx.x & (kWave - 1);
  return (lane == 0) ? 0ull : ((~0ull) >> (kWave - lane));
}

// Rust `f64 as i32` of an already floored value: saturating, NaN -> 0.
__device__ __forceinline__ int32_t sat_i32(double v) {
  if (v != v) return 0;
  if (v >= 2147483647.0) return 2147483647;
  if (v <= -2147483648.0) return (-2147483647 - 1);
  return static_cast<int32_t>(v);
}

__device__ __forceinline__ uint32_t spread3(uint32_t x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}
__device__ __forceinline__ uint32_t morton(int32_t x, int32_t y, int32_t z, int32_t kbits) {
  const uint32_t m = (1u << kbits) - 1u;
  return spread3(static_cast<uint32_t>(x) & m) | (spread3(static_cast<uint32_t>(y) & m) << 1) |
         (spread3(static_cast<uint32_t>(z) & m) << 2);
}

__device__ __forceinline__ double box_size(const unsigned long long* bits) {
  return __longlong_as_double(static_cast<long long>(*bits));
}

__device__ __forceinline__ int64_t xcd_block() { return xcd_remap(blockIdx.x, gridDim.x); }

// Doubles as 64-bit keys in their numeric order (negative values reversed),
// for atomic min / max of coordinates.
__device__ __forceinline__ unsigned long long ordered_key(double v) {
  const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
  return (u >> 63) ? ~u : (u | (1ull << 63));
}
__device__ __forceinline__ double ordered_value(unsigned long long k) {
  const unsigned long long u = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
  return __longlong_as_double(static_cast<long long>(u));
}
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off, kWave);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off, kWave);
    v = o > v ? o : v;
  }
  return v;
}

// Bucket key of a periodic image that no central ball can touch: its centre
// lies more than one box size (at least the largest 2R) outside the central
// balls' bounding box on some axis. Such images stay out of the cell lists.
constexpr uint32_t kCulled = 0xffffffffu;

__device__ __forceinline__ int32_t seg_of_ball(const GridSeg* segs, int32_t n_segs, int64_t i) {
  int32_t lo = 0, hi = n_segs - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (segs[mid].ball_off <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ void k_init(Pipeline p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *p.box_bits = static_cast<unsigned long long>(__double_as_longlong(1.0));
    for (int d = 0; d < 3; ++d) {
      p.bbox[d * kBboxStride] = ~0ull;
      p.bbox[(3 + d) * kBboxStride] = 0ull;
    }
    for (int x = 0; x < kClipQueues; ++x) p.clip_queue[x * kClipQueueStride] = 0u;
    for (int x = 0; x < kTopoRegions; ++x) p.topo_counts[x * kClipQueueStride] = 0u;
    *p.status = 0u;
    *p.overflow_cnt = 0u;
    *p.wide_cnt = 0u;
    *p.nb_big_cnt = 0u;
    p.status[kReportExclFlip] = 0u;
  }
}

__global__ void k_copy_strided(uint32_t* dst, const uint32_t* src, int count, int stride) {
  if (static_cast<int>(threadIdx.x) < count) dst[threadIdx.x] = src[threadIdx.x * stride];
}

__global__ void k_fill_u32(uint32_t* a, int64_t n, uint32_t v) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    a[i] = v;
}

// types.rs:58-63 (R = r + probe) and spheres_searcher.rs:74-77 (box size).
// The box size (and, with periodic images, the central bounding box) is
// reduced over the wavefront, then over the block in shared memory: one
// atomic per block and quantity (atomics on one address serialise at about
// 13 ns each, so one per wavefront cost 20 us at 100k balls).
constexpr int kPrepBlock = 1024;
__global__ __launch_bounds__(kPrepBlock) void k_prepare(Pipeline p) {
  __shared__ unsigned long long red[7][kPrepBlock / kWave];
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  unsigned long long v[7] = {0ull, ~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};  // box size, min x y z, max x y z
  if (i < p.n) {
    const double4 b = p.balls_in[i];
    const double R = b.w + p.probe;
    p.sph[i] = make_double4(b.x, b.y, b.z, R);
    p.seg_of[i] = p.n_segs > 1 ? seg_of_ball(p.segs, p.n_segs, i) : 0;
    const double bs = fma(R, 2.0, 0.25);
    if (bs > 0.0) v[0] = static_cast<unsigned long long>(__double_as_longlong(bs));  // positive: bits order
    v[1] = v[4] = ordered_key(b.x);
    v[2] = v[5] = ordered_key(b.y);
    v[3] = v[6] = ordered_key(b.z);
  }
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const int nq = p.periodic ? 7 : 1;
  for (int q = 0; q < nq; ++q) {
    const unsigned long long r = (q >= 1 && q <= 3) ? wave_min_u64(v[q]) : wave_max_u64(v[q]);
    if (lane == 0) red[q][wv] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < nq; ++q) {
      unsigned long long r = red[q][0];
      for (int k = 1; k < kPrepBlock / kWave; ++k) {
        const unsigned long long x = red[q][k];
        r = (q >= 1 && q <= 3) ? (x < r ? x : r) : (x > r ? x : r);
      }
      if (q == 0) {
        if (r != 0ull) atomicMax(p.box_bits, r);
      } else if (q <= 3) {
        atomicMin(p.bbox + (q - 1) * kBboxStride, r);
      } else {
        atomicMax(p.bbox + (q - 1) * kBboxStride, r);
      }
    }
  }
}

// types.rs:362-375: image g of ball i is sphere (g + 1) n + i.
__global__ void k_populate(Pipeline p) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= p.n * 26) return;
  const int64_t g = t / p.n, i = t - g * p.n;
  int sx, sy, sz;
  lattice_shift(static_cast<int>(g), &sx, &sy, &sz);
  const double4 s = p.sph[i];
  const V3 c = shift_point(p.pbox, V3{s.x, s.y, s.z}, double(sx), double(sy), double(sz));
  p.sph[(g + 1) * p.n + i] = make_double4(c.x, c.y, c.z, s.w);
  p.seg_of[(g + 1) * p.n + i] = p.seg_of[i];
}

__device__ __forceinline__ void cell_of(double4 s, double box, int32_t* cx, int32_t* cy, int32_t* cz) {
  *cx = sat_i32(floor(s.x / box));
  *cy = sat_i32(floor(s.y / box));
  *cz = sat_i32(floor(s.z / box));
}

__global__ void k_keys(Pipeline p) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= p.n_pop) return;
  const double box = box_size(p.box_bits);
  const double4 s = p.sph[i];
  if (p.periodic && i >= p.n) {
    // an image farther than one box size outside the central bounding box
    // on some axis is farther than R_i + R_j from every central ball
    const unsigned long long* bb = p.bbox;
    const bool far = s.x < ordered_value(bb[0]) - box || s.x > ordered_value(bb[3 * kBboxStride]) + box ||
                     s.y < ordered_value(bb[kBboxStride]) - box || s.y > ordered_value(bb[4 * kBboxStride]) + box ||
                     s.z < ordered_value(bb[2 * kBboxStride]) - box ||
                     s.z > ordered_value(bb[5 * kBboxStride]) + box;
    if (far) {
      p.bucket_key[i] = kCulled;
      return;
    }
  }
  int32_t cx, cy, cz;
  cell_of(s, box, &cx, &cy, &cz);
  const GridSeg sg = p.segs[p.seg_of[i]];
  const uint32_t b = static_cast<uint32_t>(sg.tab_off) + morton(cx, cy, cz, sg.kbits);
  p.bucket_key[i] = b;
  p.bucket_rank[i] = atomicAdd(&p.bucket_cnt[b], 1u);
}

__global__ void k_scatter(Pipeline p) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= p.n_pop) return;
  const uint32_t key = p.bucket_key[i];
  if (key == kCulled) return;
  const uint32_t pos = p.bucket_cnt[key] + p.bucket_rank[i];
  p.sorted_id[pos] = static_cast<int32_t>(i);
  p.sorted_sph[pos] = p.sph[i];
}

struct NbEntry {
  double key;
  int32_t id;
};

__device__ __forceinline__ bool nb_less(const NbEntry& x, const NbEntry& y) {
  if (x.key < y.key) return true;
  if (x.key > y.key) return false;
  return x.id < y.id;
}

// One wavefront per query ball: spheres_searcher.rs:253-326.
__global__ __launch_bounds__(kNbBlock) void k_neighbors(Pipeline p) {
  __shared__ NbEntry list[kNbWaves][kListCap];
  __shared__ uint32_t r_start[kNbWaves][32];
  __shared__ uint32_t r_pre[kNbWaves][32];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int64_t q0 = xcd_block() * kNbWaves + w;
  // Waves past the end still reach every block barrier below.
  const bool active = q0 < p.n;
  const int64_t q = active ? q0 : p.n - 1;
  // Non-periodic: walk balls in cell order for locality; periodic: the
  // canonical balls are the first n spheres, taken in index order.
  const int64_t a = p.periodic ? q : static_cast<int64_t>(p.sorted_id[q]);
  const double4 A4 = p.sph[a];
  const Sph A{V3{A4.x, A4.y, A4.z}, A4.w};
  const double box = box_size(p.box_bits);
  int32_t cx, cy, cz;
  cell_of(A4, box, &cx, &cy, &cz);
  const GridSeg sg = p.segs[p.seg_of[a]];

  uint32_t rs = 0, rc = 0;
  if (active && lane < 27) {
    const int dx = lane / 9 - 1, dy = (lane / 3) % 3 - 1, dz = lane % 3 - 1;
    const uint32_t b = static_cast<uint32_t>(sg.tab_off) + morton(cx + dx, cy + dy, cz + dz, sg.kbits);
    rs = p.bucket_cnt[b];
    rc = p.bucket_cnt[b + 1] - rs;
  }
  // inclusive prefix of rc over lanes 0..26
  uint32_t inc = rc;
  for (int off = 1; off < 32; off <<= 1) {
    const uint32_t v = __shfl_up(inc, off, kWave);
    if (lane >= off) inc += v;
  }
  const uint32_t total = __shfl(inc, 26, kWave);
  if (lane < 27) {
    r_start[w][lane] = rs;
    r_pre[w][lane] = inc - rc;
  }
  __syncthreads();

  // candidate j's id and sphere: its cell t is the last of the 27 whose prefix
  // it has reached (prefixes never decrease; an empty cell has its
  // successor's prefix), found by a binary search over the prefixes in shared
  // memory. The loops fetch the next chunk's candidates before they test the
  // current one, so the loads overlap the tests.
  auto fetch = [&](uint32_t j, int32_t* c, double4* C4) {
    if (j >= total) return;
    int t = 0;
#pragma unroll
    for (int step = 16; step > 0; step >>= 1)
      if (t + step < 27 && r_pre[w][t + step] <= j) t += step;
    const uint32_t pos = r_start[w][t] + (j - r_pre[w][t]);
    *c = p.sorted_id[pos];
    *C4 = p.sorted_sph[pos];
  };
  // Pass 1: the candidates that intersect a (spheres_searcher.rs:281-290),
  // listed in any order.
  int32_t c_next = 0;
  double4 C_next = make_double4(0.0, 0.0, 0.0, 0.0);
  fetch(lane, &c_next, &C_next);
  int cnt1 = 0;
  for (uint32_t base = 0; base < total; base += kWave) {
    const uint32_t j = base + lane;
    const int32_t c = c_next;
    const double4 C4 = C_next;
    fetch(j + kWave, &c_next, &C_next);
    const bool hit = j < total && c != a && intersects(A, Sph{V3{C4.x, C4.y, C4.z}, C4.w});
    const unsigned long long m = __ballot(hit);
    if (hit) {
      const int slot = cnt1 + __popcll(m & lanemask_lt());
      if (slot < kListCap) list[w][slot].id = c;
    }
    cnt1 += __popcll(m);
  }
  __syncthreads();
  if (!active) return;
  bool hidden = false;
  int cnt = 0;
  if (cnt1 > kListCap) {
    // more intersecting candidates than the list holds: the hidden test over
    // all of them (spheres_searcher.rs:295-303), then k_neighbors_big lists
    // the ball
    fetch(lane, &c_next, &C_next);
    for (uint32_t base = 0; base < total; base += kWave) {
      const uint32_t j = base + lane;
      const int32_t c = c_next;
      const double4 C4 = C_next;
      fetch(j + kWave, &c_next, &C_next);
      if (j < total && c != a) {
        const Sph Cs{V3{C4.x, C4.y, C4.z}, C4.w};
        if (intersects(A, Cs) && contains(Cs, A) && (!equals(Cs, A) || a > c)) hidden = true;
      }
    }
    if (__ballot(hidden) == 0ull) {
      if (lane == 0) p.nb_big[atomicAdd(p.nb_big_cnt, 1u)] = static_cast<int32_t>(a);
      return;
    }
  } else {
    // Pass 2 over the intersecting candidates (spheres_searcher.rs:295-310):
    // a candidate containing a (the equal one of lower index) hides a; one
    // that a contains is skipped; the others stay in the list, compacted in
    // place, with their sort keys.
    for (int e0 = 0; e0 < cnt1; e0 += kWave) {
      const int e = e0 + lane;
      bool keep = false;
      int32_t c = 0;
      double key = 0.0;
      if (e < cnt1) {
        c = list[w][e].id;
        const double4 C4 = p.sph[c];
        const Sph Cs{V3{C4.x, C4.y, C4.z}, C4.w};
        if (contains(Cs, A) && (!equals(Cs, A) || a > c)) {
          hidden = true;
        } else if (!contains(A, Cs)) {
          keep = true;
          key = circle_center_distance(A, Cs);
        }
      }
      const unsigned long long m = __ballot(keep);
      if (keep) {
        const int slot = cnt + __popcll(m & lanemask_lt());
        list[w][slot].id = c;
        list[w][slot].key = key;
      }
      cnt += __popcll(m);
    }
  }
  if (__ballot(hidden) != 0ull) {
    if (lane == 0) {
      p.nb_cnt[a] = 0;
      p.excluded[a] = 1;
      p.pair_cnt[a] = 0u;
    }
    return;
  }
  const int stored = cnt;
  // the wavefront's list is private to it: a wavefront-level barrier
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int cap = p.nb_cap;
  int32_t* out = p.nb + a * static_cast<int64_t>(cap);
  uint32_t pairs = 0;
  if (stored > kRankMax && stored <= kWave) {
    // a bitonic sort across the lanes (entry e in lane e, the free lanes
    // holding entries after every real one); ids are distinct, so the order
    // is total and equals the rank order below
    NbEntry me{__builtin_huge_val(), INT32_MAX};
    if (lane < stored) me = list[w][lane];
    for (int k = 2; k <= kWave; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        const NbEntry other{__shfl_xor(me.key, j, kWave), __shfl_xor(me.id, j, kWave)};
        const bool ascending = (lane & k) == 0, lower = (lane & j) == 0;
        if ((lower == ascending) == nb_less(other, me)) me = other;
      }
    }
    if (lane < stored && lane < cap) {
      out[lane] = me.id;
      if (pair_qualifies(p, a, me.id)) pairs += 1;
    }
  } else {
    for (int e = lane; e < stored; e += kWave) {
      const NbEntry me = list[w][e];
      int rank = 0;
      for (int k = 0; k < stored; ++k) rank += nb_less(list[w][k], me) ? 1 : 0;
      if (rank < cap) {
        out[rank] = me.id;
        if (pair_qualifies(p, a, me.id)) pairs += 1;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) pairs += __shfl_xor(pairs, off, kWave);
  if (lane == 0) {
    p.nb_cnt[a] = stored < cap ? stored : cap;
    p.excluded[a] = 0;
    p.pair_cnt[a] = pairs;
    if (cnt > cap) atomicOr(p.status, kStatNbOverflow);
  }
}

// Balls with more than kListCap candidates: one block per ball (grid-stride
// over the device-side list), candidates collected in LDS in any order and
// put in (key, id) order by a bitonic sort, which is the order the wavefront
// kernel's rank sort gives. None of these balls is hidden (the wavefront
// kernel saw all its candidates and kept hidden balls itself).
constexpr int kBigBlock = 256;
constexpr int kBigGrid = 1024;

__global__ __launch_bounds__(kBigBlock) void k_neighbors_big(Pipeline p) {
  __shared__ double key[kNbMax];
  __shared__ int32_t ids[kNbMax];
  __shared__ uint32_t r_start[32], r_pre[32];
  __shared__ uint32_t n_cand, total_s;
  __shared__ uint32_t pairs_s;
  const int t = threadIdx.x;
  const uint32_t nbig = *p.nb_big_cnt;
  for (uint32_t item = blockIdx.x; item < nbig; item += gridDim.x) {
    const int64_t a = p.nb_big[item];
    const double4 A4 = p.sph[a];
    const Sph A{V3{A4.x, A4.y, A4.z}, A4.w};
    if (t < 64) {
      const int lane = t;
      int32_t cx, cy, cz;
      cell_of(A4, box_size(p.box_bits), &cx, &cy, &cz);
      const GridSeg sg = p.segs[p.seg_of[a]];
      uint32_t rs = 0, rc = 0;
      if (lane < 27) {
        const int dx = lane / 9 - 1, dy = (lane / 3) % 3 - 1, dz = lane % 3 - 1;
        const uint32_t b = static_cast<uint32_t>(sg.tab_off) + morton(cx + dx, cy + dy, cz + dz, sg.kbits);
        rs = p.bucket_cnt[b];
        rc = p.bucket_cnt[b + 1] - rs;
      }
      uint32_t inc = rc;
      for (int off = 1; off < 32; off <<= 1) {
        const uint32_t v = __shfl_up(inc, off, kWave);
        if (lane >= off) inc += v;
      }
      if (lane < 27) {
        r_start[lane] = rs;
        r_pre[lane] = inc - rc;
      }
      if (lane == 26) total_s = inc;
      if (lane == 0) {
        n_cand = 0u;
        pairs_s = 0u;
      }
    }
    __syncthreads();
    const uint32_t total = total_s;
    for (uint32_t j = t; j < total; j += kBigBlock) {
      int r = 0;
      for (int k = 1; k < 27; ++k)
        if (r_pre[k] <= j) r = k;
      const uint32_t pos = r_start[r] + (j - r_pre[r]);
      const int32_t c = p.sorted_id[pos];
      if (c == a) continue;
      const double4 C4 = p.sorted_sph[pos];
      const Sph Cs{V3{C4.x, C4.y, C4.z}, C4.w};
      if (intersects(A, Cs) && !(contains(Cs, A) && (!equals(Cs, A) || a > c)) && !contains(A, Cs)) {
        const uint32_t slot = atomicAdd(&n_cand, 1u);
        if (slot < static_cast<uint32_t>(kNbMax)) {
          key[slot] = circle_center_distance(A, Cs);
          ids[slot] = c;
        }
      }
    }
    __syncthreads();
    const uint32_t cnt = n_cand;
    if (cnt > static_cast<uint32_t>(kNbMax)) {
      if (t == 0) atomicOr(p.status, kStatNbTooLong);
      __syncthreads();
      continue;
    }
    uint32_t len = 1;
    while (len < cnt) len <<= 1;
    for (uint32_t j = cnt + t; j < len; j += kBigBlock) {
      key[j] = __longlong_as_double(0x7ff0000000000000ll);  // +inf: after every finite key
      ids[j] = 0x7fffffff;
    }
    __syncthreads();
    // bitonic sort of len entries, ascending by (key, id)
    for (uint32_t size = 2; size <= len; size <<= 1) {
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (uint32_t i = t; i < len / 2; i += kBigBlock) {
          const uint32_t lo = 2 * i - (i & (stride - 1));
          const uint32_t hi = lo + stride;
          const bool up = (lo & size) == 0;
          const NbEntry x{key[lo], ids[lo]}, y{key[hi], ids[hi]};
          if (nb_less(y, x) == up) {
            key[lo] = y.key;
            ids[lo] = y.id;
            key[hi] = x.key;
            ids[hi] = x.id;
          }
        }
        __syncthreads();
      }
    }
    const int cap = p.nb_cap;
    const int kept = static_cast<int>(cnt) < cap ? static_cast<int>(cnt) : cap;
    int32_t* out = p.nb + a * static_cast<int64_t>(cap);
    uint32_t pairs = 0;
    for (int e = t; e < kept; e += kBigBlock) {
      out[e] = ids[e];
      if (pair_qualifies(p, a, ids[e])) ++pairs;
    }
    if (pairs) atomicAdd(&pairs_s, pairs);
    __syncthreads();
    if (t == 0) {
      p.nb_cnt[a] = kept;
      p.excluded[a] = 0;
      p.pair_cnt[a] = pairs_s;
      if (static_cast<int>(cnt) > cap) atomicOr(p.status, kStatNbOverflow);
    }
    __syncthreads();
  }
}

__global__ void k_gather_rows(Pipeline p, const int32_t* ids, int64_t k, int32_t* cnt, int32_t* rows) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= k * p.nb_cap) return;
  const int64_t r = t / p.nb_cap, j = t - r * p.nb_cap;
  const int32_t id = ids[r];
  if (id < 0 || id >= p.n) {
    if (j == 0) cnt[r] = 0;
    return;
  }
  const int c = p.nb_cnt[id];
  if (j == 0) cnt[r] = c;
  if (j < c) rows[t] = p.nb[static_cast<int64_t>(id) * p.nb_cap + j];
}

// ------------------------------------------------------------------ scan ----
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total, uint32_t* sh) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  uint32_t inc = v;
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t t = __shfl_up(inc, off, kWave);
    if (lane >= off) inc += t;
  }
  if (lane == kWave - 1) sh[w] = inc;
  __syncthreads();
  uint32_t wpre = 0, tot = 0;
  for (int k = 0; k < kScanBlock / kWave; ++k) {
    if (k < w) wpre += sh[k];
    tot += sh[k];
  }
  __syncthreads();
  *total = tot;
  return wpre + inc - v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t* in, int64_t n, uint32_t* block_sums) {
  __shared__ uint32_t sh[kScanBlock / kWave];
  const int64_t base = blockIdx.x * static_cast<int64_t>(kScanTile);
  uint32_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + static_cast<int64_t>(k) * kScanBlock + threadIdx.x;
    if (i < n) s += in[i];
  }
  uint32_t tot;
  (void)block_exclusive_scan(s, &tot, sh);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// The totals of an empty scan.
__global__ void k_scan_empty(uint32_t* counts, uint32_t* tmp, uint32_t* total_out, uint32_t* beyond_out) {
  if (threadIdx.x != 0) return;
  counts[0] = 0u;
  tmp[0] = 0u;
  tmp[1] = 0u;
  if (total_out != nullptr) *total_out = 0u;
  if (beyond_out != nullptr) *beyond_out = 0u;
}

// Exclusive scan of tile `blockIdx.x`, offset by the sum of the tile sums
// before it, which the block adds up itself (at most a few thousand words).
// The last tile also writes the total modulo 2^32 to data[n] and sums[nb],
// and 1 to sums[nb + 1] when the exact total exceeds INT32_MAX (the same two
// words to *total_out and *beyond_out when given: a run's report words). A
// tile sums at most 2048 counts of at most 4096, so its sum is exact in 32
// bits; the sums of tiles are added in 64 bits (ScanTotal).
__global__ __launch_bounds__(kScanBlock) void k_scan_apply(uint32_t* data, int64_t n, uint32_t* sums, int64_t nb,
                                                           uint32_t* total_out, uint32_t* beyond_out) {
  __shared__ uint32_t sh[kScanBlock / kWave];
  __shared__ unsigned long long sh64[kScanBlock / kWave];
  const int64_t tile = blockIdx.x;
  const bool last = tile == static_cast<int64_t>(gridDim.x) - 1;
  unsigned long long before = 0;
  for (int64_t k = threadIdx.x; k < tile; k += kScanBlock) before += sums[k];
  for (int off = kWave / 2; off > 0; off >>= 1) before += __shfl_xor(before, off, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) sh64[threadIdx.x / kWave] = before;
  __syncthreads();
  before = 0;
  for (int k = 0; k < kScanBlock / kWave; ++k) before += sh64[k];
  const int64_t base = tile * static_cast<int64_t>(kScanTile) + static_cast<int64_t>(threadIdx.x) * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    v[k] = i < n ? data[i] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = static_cast<uint32_t>(before) + block_exclusive_scan(s, &tot, sh);
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    if (i < n) data[i] = run;
    run += v[k];
  }
  if (last && threadIdx.x == 0) {
    ScanTotal t;
    t.lo = static_cast<uint32_t>(before);
    t.full = before;
    t.add(tot);
    data[n] = t.lo;
    sums[nb] = t.lo;
    sums[nb + 1] = t.beyond_int32();
    if (total_out != nullptr) *total_out = t.lo;
    if (beyond_out != nullptr) *beyond_out = t.beyond_int32();
  }
}

inline int64_t blocks_for(int64_t n, int64_t per) { return (n + per - 1) / per; }

}  // namespace

int64_t scan_tmp_words(int64_t n) { return blocks_for(n, kScanTile) + 2; }

hipError_t launch_scan_u32(uint32_t* counts, int64_t n, uint32_t* tmp, hipStream_t s, uint32_t* total_out,
                           uint32_t* beyond_out) {
  if (n <= 0) {
    hipLaunchKernelGGL(k_scan_empty, dim3(1), dim3(64), 0, s, counts, tmp, total_out, beyond_out);
    return hipGetLastError();
  }
  const int64_t nb = blocks_for(n, kScanTile);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, s, counts, n, tmp);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanBlock), 0, s, counts, n, tmp, nb, total_out, beyond_out);
  return hipGetLastError();
}

hipError_t launch_prepare(const Pipeline& p, hipStream_t s) {
  hipLaunchKernelGGL(k_init, dim3(1), dim3(64), 0, s, p);
  if (p.n > 0) hipLaunchKernelGGL(k_prepare, dim3(blocks_for(p.n, kPrepBlock)), dim3(kPrepBlock), 0, s, p);
  if (p.periodic && p.n > 0) hipLaunchKernelGGL(k_populate, dim3(blocks_for(p.n * 26, 256)), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_gather_rows(const Pipeline& p, const int32_t* ids, int64_t k, int32_t* cnt, int32_t* rows,
                              hipStream_t s) {
  if (k > 0)
    hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for(k * p.nb_cap, 256)), dim3(256), 0, s, p, ids, k, cnt, rows);
  return hipGetLastError();
}

hipError_t launch_copy_strided(uint32_t* dst, const uint32_t* src, int count, int stride, hipStream_t s) {
  if (count > 0) hipLaunchKernelGGL(k_copy_strided, dim3(1), dim3(64), 0, s, dst, src, count, stride);
  return hipGetLastError();
}

hipError_t launch_fill_words(uint32_t* a, int64_t n, uint32_t v, hipStream_t s) {
  if (n > 0) {
    const int64_t b = blocks_for(n, 256) < 4096 ? blocks_for(n, 256) : 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3(b), dim3(256), 0, s, a, n, v);
  }
  return hipGetLastError();
}


hipError_t launch_grid(const Pipeline& p, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(p.n_buckets + 1, 256) < 4096 ? blocks_for(p.n_buckets + 1, 256) : 4096),
                     dim3(256), 0, s, p.bucket_cnt, p.n_buckets + 1, 0u);
  if (p.n_pop > 0) hipLaunchKernelGGL(k_keys, dim3(blocks_for(p.n_pop, 256)), dim3(256), 0, s, p);
  // scan scratch lives after the pair counts' own scratch: callers size it.
  return hipGetLastError();
}

hipError_t launch_scatter(const Pipeline& p, hipStream_t s) {
  if (p.n_pop > 0) hipLaunchKernelGGL(k_scatter, dim3(blocks_for(p.n_pop, 256)), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_neighbors(const Pipeline& p, hipStream_t s) {
  if (p.n > 0) {
    hipLaunchKernelGGL(k_neighbors, dim3(blocks_for(p.n, kNbWaves)), dim3(kNbBlock), 0, s, p);
    const int64_t g = blocks_for(p.n, 64) < kBigGrid ? blocks_for(p.n, 64) : kBigGrid;
    hipLaunchKernelGGL(k_neighbors_big, dim3(g), dim3(kBigBlock), 0, s, p);
  }
  return hipGetLastError();
}

hipError_t launch_pairs(const Pipeline& p, hipStream_t s) {
  // the clip kernels write the pair lists and the incoming lists; the
  // incoming counts and the contact marks start at 0
  if (p.n > 0 && p.mode == 0) {
    const int64_t fb = blocks_for(p.n, 256) < 4096 ? blocks_for(p.n, 256) : 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3(fb), dim3(256), 0, s, p.in_cnt, p.n, 0u);
  }
  if (p.n > 0) {
    if (p.pair_contact != nullptr && p.n_pairs > 0) {
      const int64_t words = (p.n_pairs + 3) / 4;
      const int64_t fw = blocks_for(words, 256) < 4096 ? blocks_for(words, 256) : 4096;
      hipLaunchKernelGGL(k_fill_u32, dim3(fw), dim3(256), 0, s, reinterpret_cast<uint32_t*>(p.pair_contact), words,
                         0u);
    }
  }
  return hipGetLastError();
}

}  // namespace vltr
