// Totals of the exclusive scans over per-ball and per-pair counts, and the
// guard on the candidate-pair total (device and host).
//
// The scans keep 32-bit offsets and pair indices are int32 throughout (pair
// lists, incoming lists, fallback work lists). Each scan tile sums at most
// 2048 counts of at most 4096 (one ball's pairs), so a tile sum always fits
// 32 bits; the running total over the tiles is also kept in 64 bits, and a
// total beyond INT32_MAX is reported instead of wrapping. The host then
// refuses the run before any pair buffer is sized from the wrapped total.
#pragma once

#include <cstdint>

#include "vltr_geom.h"

namespace vltr {

constexpr uint64_t kMaxPairs = 0x7fffffffull;  // pair indices are int32

struct ScanTotal {
  uint32_t lo = 0;    // the total modulo 2^32 (what the 32-bit offsets hold)
  uint64_t full = 0;  // the exact total
  VLTR_HD void add(uint32_t tile_sum) {
    lo += tile_sum;
    full += tile_sum;
  }
  // 1 when the exact total does not fit an int32 pair index
  VLTR_HD uint32_t beyond_int32() const { return full > kMaxPairs ? 1u : 0u; }
};

// The run's pair total as the report block holds it (the 32-bit total and the
// scan's beyond-int32 word): the total, or -1 when the pairs cannot be
// indexed by int32.
VLTR_HD int64_t checked_pair_total(uint32_t total_word, uint32_t beyond_word) {
  if (beyond_word != 0u || static_cast<uint64_t>(total_word) > kMaxPairs) return -1;
  return static_cast<int64_t>(total_word);
}

}  // namespace vltr
