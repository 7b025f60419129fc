// Tessellation network on the device (replaces contact.rs:120-177 and
// tessellation.rs:334-401) and the ball-order sort of a periodic updateable's
// contact summary (updateable.rs:333-376).
//
// Records are generated in contact order: per contact with a restricted
// contour of m points, m edges (a, b, right neighbor, or none for an arc) and
// m vertices (a, b, left neighbor, right neighbor, none where the contour
// follows the circle); an uncut contact is one arc edge. Ids become
// canonical (periodic), then sorted within each record. The records are then
// ordered by their ids and the first of each run of equal ids is kept.
//
// Ids are packed into one 2-word key per record: each id takes b bits, where
// b holds n (the code of "none"), most significant id first, so the key order
// is the lexicographic order of the id tuple. The order is a stable LSD
// radix sort of the key's low word and then, for keys wider than 64 bits,
// of its high word; a stable sort keeps the first record in contact order of
// each run (the reference's sort_unstable compares ids only and leaves that
// choice open).
#include <rocprim/device/device_radix_sort.hpp>

#include "vltr_internal.h"
#include "voronota_ltr.h"

namespace vltr {
namespace {

inline int64_t blocks_for(int64_t n, int64_t per) { return (n + per - 1) / per; }

__device__ __forceinline__ bool net_contact(const Pipeline& p, int64_t i) {
  if (p.n_pairs_dev != nullptr && i >= static_cast<int64_t>(*p.n_pairs_dev)) return false;
  return has_contact(p, i);
}

__global__ void k_net_count(Pipeline p, uint32_t* ecnt, uint32_t* vcnt) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= p.n_pairs) return;
  uint32_t e = 0u, v = 0u;
  if (net_contact(p, i)) {
    const int m = p.net[i].m;
    if (m == kTopoTooLong || m < 0 || m > kTopoCap) {
      atomicOr(p.status, kStatNetTooLong);
    } else {
      e = m > 0 ? static_cast<uint32_t>(m) : 1u;
      v = static_cast<uint32_t>(m);
    }
  }
  ecnt[i] = e;
  vcnt[i] = v;
}

// Places the b-bit code c at bit position shift of the key (lo, hi).
__device__ __forceinline__ void put_code(uint64_t c, int shift, uint64_t& lo, uint64_t& hi) {
  if (shift >= 64) {
    hi |= c << (shift - 64);
  } else {
    lo |= c << shift;
    if (shift > 0) hi |= c >> (64 - shift);
  }
}

__device__ __forceinline__ uint64_t get_code(uint64_t lo, uint64_t hi, int shift, int b) {
  const uint64_t mask = (uint64_t(1) << b) - 1;
  if (shift >= 64) return (hi >> (shift - 64)) & mask;
  uint64_t v = lo >> shift;
  if (shift > 0) v |= hi << (64 - shift);
  return v & mask;
}

// Code of an id (negative: none): canonical, with none = n above every id.
__device__ __forceinline__ uint64_t id_code(int64_t id, const NetLayout& L) {
  if (id < 0) return static_cast<uint64_t>(L.n);
  return static_cast<uint64_t>(L.periodic ? id % L.n : id);
}

// Sorted codes of K ids packed into (lo, hi).
template <int K>
__device__ __forceinline__ void pack(const int64_t (&ids)[K], const NetLayout& L, unsigned long long* lo_out,
                                     unsigned long long* hi_out) {
  uint64_t c[K];
  for (int k = 0; k < K; ++k) c[k] = id_code(ids[k], L);
  for (int k = 1; k < K; ++k)
    for (int j = k; j > 0 && c[j - 1] > c[j]; --j) {
      const uint64_t t = c[j - 1];
      c[j - 1] = c[j];
      c[j] = t;
    }
  uint64_t lo = 0ull, hi = 0ull;
  for (int k = 0; k < K; ++k) put_code(c[k], (K - 1 - k) * L.bits, lo, hi);
  *lo_out = lo;
  *hi_out = hi;
}

// The records of pair i at its scanned offsets: keys, their own index (the
// sort's values) and values (edge: length; vertex: x, y, z).
__global__ void k_net_records(Pipeline p, const uint32_t* eoff, const uint32_t* voff, NetLayout L, NetRecs E,
                              NetRecs V) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= p.n_pairs) return;
  int64_t e = eoff[i];
  if (static_cast<int64_t>(eoff[i + 1]) == e) return;
  int64_t v = voff[i];
  const int64_t a = p.pair_a[i], b = p.pair_b[i];
  const ContactNet& nt = p.net[i];
  const int m = nt.m;
  if (m == 0) {
    const int64_t ids[3] = {a, b, -1};
    pack<3>(ids, L, E.lo + e, E.hi + e);
    E.idx[e] = static_cast<uint32_t>(e);
    E.val[e] = p.rec[i].arc_length;
    return;
  }
  const double r_ic = nt.r_ic;
  for (int t = 0; t < m; ++t) {
    const int u = t + 1 < m ? t + 1 : 0;
    const int64_t rt = nt.r[t], lt = nt.l[t];
    const bool arc = rt == a;
    double len;
    if (arc) {
      len = nt.ang[t] * r_ic;
    } else {
      const double dx = nt.x[t] - nt.x[u], dy = nt.y[t] - nt.y[u], dz = nt.z[t] - nt.z[u];
      len = sqrt((dx * dx + dy * dy) + dz * dz);
    }
    const int64_t eids[3] = {a, b, arc ? -1 : rt};
    pack<3>(eids, L, E.lo + e, E.hi + e);
    E.idx[e] = static_cast<uint32_t>(e);
    E.val[e] = len;
    ++e;
    const int64_t vids[4] = {a, b, lt == a ? -1 : lt, arc ? -1 : rt};
    pack<4>(vids, L, V.lo + v, V.hi + v);
    V.idx[v] = static_cast<uint32_t>(v);
    V.val[3 * v] = nt.x[t];
    V.val[3 * v + 1] = nt.y[t];
    V.val[3 * v + 2] = nt.z[t];
    ++v;
  }
}

__global__ void k_gather_u64(const unsigned long long* src, const uint32_t* perm, int64_t count,
                             unsigned long long* dst) {
  const int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (j < count) dst[j] = src[perm[j]];
}

// flag[j] = 1 where sorted record j starts a run of equal ids.
__global__ void k_net_runs(NetRecs R, const uint32_t* perm, int64_t count, uint32_t* flag) {
  const int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (j >= count) return;
  bool first = j == 0;
  if (!first) {
    const uint32_t x = perm[j], y = perm[j - 1];
    first = R.lo[x] != R.lo[y] || R.hi[x] != R.hi[y];
  }
  flag[j] = first ? 1u : 0u;
}

__device__ __forceinline__ uint64_t out_id(uint64_t code, const NetLayout& L) {
  return code == static_cast<uint64_t>(L.n) ? VLTR_NULL_ID : code;
}

__global__ void k_net_emit_edges(NetRecs R, const uint32_t* perm, const uint32_t* pos, int64_t count, NetLayout L,
                                 vltr_cell_edge* out) {
  const int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (j >= count || pos[j + 1] == pos[j]) return;
  const uint32_t x = perm[j];
  vltr_cell_edge o;
  for (int k = 0; k < 3; ++k) o.ball_indices[k] = out_id(get_code(R.lo[x], R.hi[x], (2 - k) * L.bits, L.bits), L);
  o.length = R.val[x];
  out[pos[j]] = o;
}

__global__ void k_net_emit_vertices(NetRecs R, const uint32_t* perm, const uint32_t* pos, int64_t count, NetLayout L,
                                    vltr_cell_vertex* out) {
  const int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (j >= count || pos[j + 1] == pos[j]) return;
  const uint32_t x = perm[j];
  vltr_cell_vertex o;
  for (int k = 0; k < 4; ++k) o.ball_indices[k] = out_id(get_code(R.lo[x], R.hi[x], (3 - k) * L.bits, L.bits), L);
  o.x = R.val[3 * x];
  o.y = R.val[3 * x + 1];
  o.z = R.val[3 * x + 2];
  out[pos[j]] = o;
}

// Periodic updateable summary: key = the smaller id of each contact.
__global__ void k_contacts_order(const vltr_contact* c, int64_t count, uint32_t* key, uint32_t* idx) {
  const int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (j >= count) return;
  const uint64_t a = c[j].id_a, b = c[j].id_b;
  key[j] = static_cast<uint32_t>(a < b ? a : b);
  idx[j] = static_cast<uint32_t>(j);
}

// dst[j] = src[perm[j]] with its ids ordered (smaller first).
__global__ void k_contacts_gather(const vltr_contact* src, const uint32_t* perm, int64_t count, vltr_contact* dst) {
  const int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (j >= count) return;
  vltr_contact x = src[perm[j]];
  if (x.id_a > x.id_b) {
    const uint64_t t = x.id_a;
    x.id_a = x.id_b;
    x.id_b = t;
  }
  dst[j] = x;
}

int bit_width(uint64_t v) {
  int b = 0;
  while (v >> b) ++b;
  return b;
}

}  // namespace

NetLayout net_layout(int64_t n, int periodic) {
  NetLayout L;
  L.n = n;
  L.periodic = periodic;
  L.bits = bit_width(static_cast<uint64_t>(n > 0 ? n : 1));
  return L;
}

hipError_t launch_net_count(const Pipeline& p, uint32_t* ecnt, uint32_t* vcnt, hipStream_t s) {
  if (p.n_pairs > 0)
    hipLaunchKernelGGL(k_net_count, dim3(blocks_for(p.n_pairs, 256)), dim3(256), 0, s, p, ecnt, vcnt);
  return hipGetLastError();
}

hipError_t launch_net_records(const Pipeline& p, const uint32_t* eoff, const uint32_t* voff, const NetLayout& L,
                              const NetRecs& E, const NetRecs& V, hipStream_t s) {
  if (p.n_pairs > 0)
    hipLaunchKernelGGL(k_net_records, dim3(blocks_for(p.n_pairs, 256)), dim3(256), 0, s, p, eoff, voff, L, E, V);
  return hipGetLastError();
}

// rocprim's radix sort over (64-bit key, 32-bit value) pairs, bits [0, end_bit).
static hipError_t sort_pairs64(void* tmp, size_t& tmp_bytes, const unsigned long long* keys_in,
                               unsigned long long* keys_out, const uint32_t* vals_in, uint32_t* vals_out, int64_t count,
                               int end_bit, hipStream_t s) {
  return rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out,
                                   static_cast<unsigned int>(count), 0u, static_cast<unsigned int>(end_bit), s);
}

size_t net_sort_tmp_bytes(int64_t count, int key_bits) {
  size_t need = 0;
  if (count <= 0) return 0;
  for (int pass = 0; pass < 2; ++pass) {
    const int end_bit = pass == 0 ? (key_bits < 64 ? key_bits : 64) : key_bits - 64;
    if (end_bit <= 0) continue;
    size_t bytes = 0;
    if (sort_pairs64(nullptr, bytes, nullptr, nullptr, nullptr, nullptr, count, end_bit, nullptr) != hipSuccess)
      return SIZE_MAX;
    need = bytes > need ? bytes : need;
  }
  return need;
}

hipError_t launch_net_sort(const NetRecs& R, int64_t count, int key_bits, const NetSortWork& W, const uint32_t** perm,
                           hipStream_t s) {
  *perm = W.perm1;
  if (count <= 0) return hipSuccess;
  size_t bytes = W.tmp_bytes;
  hipError_t e = sort_pairs64(W.tmp, bytes, R.lo, W.keys_out, R.idx, W.perm1, count,
                              key_bits < 64 ? key_bits : 64, s);
  if (e != hipSuccess || key_bits <= 64) return e;
  hipLaunchKernelGGL(k_gather_u64, dim3(blocks_for(count, 256)), dim3(256), 0, s, R.hi, W.perm1, count, W.keys_mid);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  bytes = W.tmp_bytes;
  e = sort_pairs64(W.tmp, bytes, W.keys_mid, W.keys_out, W.perm1, W.perm2, count, key_bits - 64, s);
  *perm = W.perm2;
  return e;
}

hipError_t launch_net_runs(const NetRecs& R, const uint32_t* perm, int64_t count, uint32_t* flag, hipStream_t s) {
  if (count > 0) hipLaunchKernelGGL(k_net_runs, dim3(blocks_for(count, 256)), dim3(256), 0, s, R, perm, count, flag);
  return hipGetLastError();
}

hipError_t launch_net_emit_edges(const NetRecs& R, const uint32_t* perm, const uint32_t* pos, int64_t count,
                                 const NetLayout& L, vltr_cell_edge* out, hipStream_t s) {
  if (count > 0)
    hipLaunchKernelGGL(k_net_emit_edges, dim3(blocks_for(count, 256)), dim3(256), 0, s, R, perm, pos, count, L, out);
  return hipGetLastError();
}

hipError_t launch_net_emit_vertices(const NetRecs& R, const uint32_t* perm, const uint32_t* pos, int64_t count,
                                    const NetLayout& L, vltr_cell_vertex* out, hipStream_t s) {
  if (count > 0)
    hipLaunchKernelGGL(k_net_emit_vertices, dim3(blocks_for(count, 256)), dim3(256), 0, s, R, perm, pos, count, L,
                       out);
  return hipGetLastError();
}

size_t contacts_sort_tmp_bytes(int64_t count) {
  if (count <= 0) return 0;
  size_t bytes = 0;
  if (rocprim::radix_sort_pairs(nullptr, bytes, static_cast<const uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                static_cast<const uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                static_cast<unsigned int>(count), 0u, 32u, nullptr) != hipSuccess)
    return SIZE_MAX;
  return bytes;
}

hipError_t launch_contacts_by_first(const vltr_contact* c, int64_t count, int64_t n, const ContactSortWork& W,
                                    hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_contacts_order, dim3(blocks_for(count, 256)), dim3(256), 0, s, c, count, W.key, W.idx);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t bytes = W.tmp_bytes;
  e = rocprim::radix_sort_pairs(W.tmp, bytes, W.key, W.key_out, W.idx, W.perm, static_cast<unsigned int>(count), 0u,
                                static_cast<unsigned int>(bit_width(static_cast<uint64_t>(n))), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_contacts_gather, dim3(blocks_for(count, 256)), dim3(256), 0, s, c, W.perm, count, W.out);
  return hipGetLastError();
}

}  // namespace vltr
