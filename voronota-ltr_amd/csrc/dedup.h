// Canonical pairs of a periodic run (tessellation.rs:556-601,
// updateable.rs:565-630): among the contacts through a periodic image, the
// first record in pair order of each unordered canonical pair (contacts
// between two canonical balls are always kept). An open-addressing table
// keyed by the canonical pair holds the smallest pair index with a contact;
// launch_dedup_table (k_adjoint.hip) fills it after the clip.
#pragma once

#include "vltr_internal.h"

namespace vltr {

// Key lo * n + hi + 1 (0 = empty), first = smallest pair index with that key.
// Sized to at least twice the candidate pairs (dedup_slots), so every probe
// sequence reaches its key or an empty slot.
struct DedupTable {
  unsigned long long* keys;
  uint32_t* first;
  uint64_t mask;
};

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

__device__ __forceinline__ unsigned long long pair_key(const Pipeline& p, int64_t t) {
  const int64_t ca = canon_id(p, p.pair_a[t]), cb = canon_id(p, p.pair_b[t]);
  const int64_t lo = ca < cb ? ca : cb, hi = ca < cb ? cb : ca;
  return static_cast<unsigned long long>(lo) * static_cast<unsigned long long>(p.n) + hi + 1ull;
}

// The table's slot mask for a run of n_pairs candidate pairs: dedup_slots
// (a power of two of at least twice the pairs, and at least 1024) minus one.
__host__ __device__ __forceinline__ uint64_t dedup_slots_mask(int64_t n_pairs) {
  uint64_t h = 1024;
  while (h < 2 * static_cast<uint64_t>(n_pairs > 0 ? n_pairs : 1)) h <<= 1;
  return h - 1;
}

// Whether pair t (a contact through an image, inserted in the table) is the
// first of its canonical pair.
__device__ __forceinline__ bool dedup_first(const Pipeline& p, const DedupTable& T, int64_t t) {
  const unsigned long long key = pair_key(p, t);
  uint64_t slot = mix64(key) & T.mask;
  while (T.keys[slot] != key) slot = (slot + 1) & T.mask;
  return T.first[slot] == static_cast<uint32_t>(t);
}

}  // namespace vltr
