// Internal declarations shared by the kernel translation units and the host
// orchestration (not part of the public C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "contour.h"
#include "group_clip.h"
#include "scan_total.h"
#include "vltr_geom.h"
#include "voronota_ltr.h"

namespace vltr {

// Per-ball cell accumulator result (types.rs:431-454): stage 0 = no cell,
// 2 = SAS computed (cells are emitted for stage 2 only).
struct CellRec {
  double sas;
  double volume;
  int32_t stage;
  int32_t pad;
};

// Geometry of the uniform cell hash: structure s owns buckets
// [tab_off[s], tab_off[s] + 8^kbits[s]); a ball's bucket is the Morton code of
// its cell coordinates modulo 2^kbits[s] per axis.
struct GridSeg {
  int64_t ball_off;
  int64_t tab_off;
  int32_t kbits;
  int32_t n;
};

// Device-global status word (atomically OR-ed by kernels).
constexpr uint32_t kStatNbOverflow = 1u;    // a neighbor list exceeded its capacity
constexpr uint32_t kStatPairOverflow = 2u;  // more pairs than the pair buffers hold
constexpr uint32_t kStatContourOverflow = 4u;
constexpr uint32_t kStatNbTooLong = 8u;        // a ball has more than kNbMax neighbors
constexpr uint32_t kStatNetTooLong = 16u;      // a network contour has more than kTopoCap points
constexpr int kNbMax = 4096;
constexpr int kBboxStride = 32;  // 256 bytes
// Work-queue counters of the persistent clip kernel: one per XCD, each on its
// own 128-byte line.
constexpr int kClipQueues = 8;
constexpr int kClipQueueStride = 32;  // 32 words
// The run's small device block: the box size word, the report words (from
// byte 8), the bounding box (from byte 256, kBboxStride words apart), the
// clip queue counters (from kSmallQueueOff) and the topology store counters
// (from kSmallTopoOff).
constexpr size_t kSmallQueueOff = 256 + 6 * kBboxStride * sizeof(unsigned long long);
// Adjoint runs: the topology store is split into kTopoRegions equal regions,
// one per XCD, each with its record counter on its own line (from
// kSmallTopoOff, kClipQueueStride words apart), so that the clip kernels'
// reservations of different XCDs do not meet on one address.
constexpr int kTopoRegions = 8;
constexpr size_t kSmallTopoOff = kSmallQueueOff + kClipQueues * kClipQueueStride * sizeof(uint32_t);
constexpr size_t kSmallBytes = kSmallTopoOff + kTopoRegions * kClipQueueStride * sizeof(uint32_t);

// The run's report block: consecutive device words from p.status, copied to
// the host in one transfer. The totals are copied in by k_copy_word after the
// scans that produce them.
enum ReportWord : int {
  kReportStatus = 0,
  kReportOverflow = 1,  // pairs listed for the one-lane kernel
  kReportWide = 2,      // pairs listed for the 16-lane kernel
  kReportNbBig = 3,     // balls listed for the block-per-ball neighbor kernel
  kReportPairs = 4,     // candidate pairs (the pair count scan's total)
  kReportContacts = 5,  // contacts (the contact flag scan's total)
  kReportCells = 6,     // cells (the cell flag scan's total)
  kReportExclFlip = 7,  // nonzero: an updateable run cleared a stale exclusion status
                        // after the pairs were counted (k_updateable_cells)
  kReportMarksOld = 8,      // updateable step: balls marked as old neighbors (k_count_marks,
  kReportMarksChanged = 9,  // written before the run) and as changed
  kReportPairsBeyond = 10,  // 1: the exact pair total exceeds INT32_MAX (kReportPairs wrapped)
  kReportEdges = 11,        // network runs: edge records (the edge count scan's total)
  kReportVerts = 12,        // network runs: vertex records (the vertex count scan's total)
  kReportEdgesBeyond = 13,  // 1: the exact edge total exceeds INT32_MAX
  kReportVertsBeyond = 14,  // 1: the exact vertex total exceeds INT32_MAX
  kReportTopos = 15,        // adjoint runs: records reserved in each of the kTopoRegions
                            // regions of the topology store (words 15 .. 22; may exceed a region)
  kReportWords = kReportTopos + kTopoRegions
};

// Launch parameters of the tessellation pipeline for one call (one or many
// structures). All pointers are device pointers.
struct Pipeline {
  int64_t n;         // balls (canonical)
  double inv_n;      // 1 / n (canon_id), set with n by set_balls
  int64_t n_pop;     // spheres searched (n, or 27 n with periodic images)
  int32_t n_segs;
  int32_t periodic;  // 1: sph holds 27 n images, ids >= n are images
  int32_t nb_cap;    // neighbor slots per ball
  int32_t mode;      // 0 full tessellation, 1 contacts only
  int64_t pair_cap;  // capacity of the pair buffers
  int64_t n_pairs;   // pairs to clip (set once the pair counts are known)
  double probe;
  PBox pbox;

  const double4* balls_in;  // n (x, y, z, r)
  const int32_t* groups;    // may be null
  int64_t n_groups;
  const uint8_t* user_excl;  // n, may be null: balls excluded by the caller (updateable.rs:225-272)

  double4* sph;          // n_pop (x, y, z, R)
  int32_t* seg_of;       // n_pop (segment of each sphere)
  const GridSeg* segs;   // n_segs
  int64_t n_buckets;     // total buckets over all segments
  uint32_t* bucket_key;  // n_pop
  uint32_t* bucket_rank; // n_pop
  uint32_t* bucket_cnt;  // n_buckets + 1 (scanned in place to starts)
  int32_t* sorted_id;    // n_pop
  double4* sorted_sph;   // n_pop
  unsigned long long* box_bits;  // max over 2R + 0.25 as double bits
  unsigned long long* bbox;      // min x, y, z then max x, y, z of the central centres (order-preserving
                                 // keys), one every kBboxStride words so their atomics hit different lines
  uint32_t* clip_queue;          // kClipQueues counters, kClipQueueStride words apart (cleared by k_init)

  int32_t* nb;           // n * nb_cap neighbor ids (sorted by key, id)
  int32_t* nb_cnt;       // n
  uint8_t* excluded;     // n (hidden balls)
  uint32_t* pair_cnt;    // n + 1 (scanned in place to bases)
  int32_t* pair_a;       // pair_cap
  int32_t* pair_b;       // pair_cap
  PairRec* rec;          // pair_cap
  // pair_cap (rounded up to 4) or null: 1 where the pair has a contact,
  // written by the clip kernels as they write or keep the records; the
  // readers ask it which pairs have a contact (has_contact), k_cells loads
  // only those own records
  uint8_t* pair_contact;
  // 1: every pair's record is written (updateable steps: the next step
  // copies the records of the pairs it does not clip); 0: with contact bytes,
  // only the records of pairs with a contact
  int32_t records_all;
  CellRec* cell;         // n
  uint32_t* status;      // 1 word; the first of the run's report block (kReport* words)
  int32_t* overflow_list;  // pair_cap (pairs needing the slow kernel)
  uint32_t* overflow_cnt;  // 1
  int32_t* wide_list;      // pair_cap (pairs of balls with more than 64 neighbors)
  uint32_t* wide_cnt;      // 1
  int32_t* nb_big;         // n: balls with more candidates than one wavefront's list holds
  uint32_t* nb_big_cnt;    // 1
  double* slow_ws;         // workspace of the slow kernel
  int64_t slow_ws_pairs;   // its lanes (one workspace region each)

  // Adjoint runs: the cut topologies of the contacts in a store of
  // kTopoRegions regions of topo_region records (records reserved through
  // the region counters topo_counts, kClipQueueStride words apart: a clip
  // wavefront reserves one record per pair it clips, of its XCD's region),
  // found through the pair's slot (topo_slot, pair_cap; kTopoSlot* values
  // need no record). Null topo: no topologies.
  ContactTopo* topo;
  uint32_t* topo_slot;
  uint32_t* topo_counts;
  int64_t topo_region;
  ContactNet* net;         // pair_cap or null: restricted contours (network output)
  double* seed;            // n or null: dE/d(solid angle) per cell (adjoint runs)
  double sas_weight, volume_weight;  // alpha, beta of the adjoint's energy
  // Incoming lists (full runs): the contacts (a, b), a < b, received by ball
  // b as their second end, appended by the clip kernels as they write or keep
  // each contact's record (the pair lists are written by the clip kernels
  // too, as they walk each ball's neighbors).
  uint32_t* in_cnt;        // n: list lengths
  int32_t* incoming;       // n * nb_cap: their pair indices (any order)

  // Updateable steps (updateable.rs:476-512): a pair is clipped only when
  // both of its balls are marked in dirty; every other pair keeps its record
  // (and cut topology) of the previous run, found by its partner in that
  // run's pair list of the same ball, or gets an empty record when it was no
  // pair then. Null dirty: every pair is clipped.
  const uint8_t* dirty;           // n
  const uint32_t* prev_pair_cnt;  // n + 1: pair bases of the previous run
  const int32_t* prev_pair_b;     // its pair partners
  const PairRec* prev_rec;        // its records
  const ContactTopo* prev_topo;   // its topology store and slots (read when topo is set)
  const uint32_t* prev_topo_slot;

  int32_t clip_variant;  // launch shape of k_clip (index into its variant table)
  // Graph replays: the pair buffers hold pair_cap pairs (a capacity from an
  // earlier run), kernels over pairs run to n_pairs = pair_cap and stop at
  // the device-side total *n_pairs_dev. Null in direct runs (n_pairs exact).
  const uint32_t* n_pairs_dev;
  // Graph replays: the pair scan's beyond-int32 word (nonzero: *n_pairs_dev
  // wrapped and no pair offset can be trusted).
  const uint32_t* n_pairs_beyond_dev;

  // Dedup cells (incremental steps of a periodic updateable,
  // updateable.rs:476-512 with 691-727): a cell sums the contacts of the
  // deduplicated per-ball lists, i.e. of each canonical pair through an image
  // only the first record in pair order, for both of its ends. Non-null: the
  // canonical-pair table (launch_dedup_table) of this run's records; such
  // first records were appended to the second end's incoming list
  // (launch_dedup_incoming). Null: every ball sums its own records
  // (tessellation.rs:607-658).
  unsigned long long* dd_keys;
  uint32_t* dd_first;

  // Contacts the output buffer holds (0: one per candidate pair, the
  // context's own buffer); contacts beyond it are counted, not written.
  int64_t contacts_cap;
};

// A graph replay whose pair total outgrew the capacity it was captured with
// (or INT32_MAX, where the 32-bit total wraps): the pair kernels do nothing
// and the host reruns without the graph.
__device__ __forceinline__ bool pairs_overflowed(const Pipeline& p) {
  return p.n_pairs_dev != nullptr &&
         (static_cast<int64_t>(*p.n_pairs_dev) > p.pair_cap ||
          (p.n_pairs_beyond_dev != nullptr && *p.n_pairs_beyond_dev != 0u));
}

// tessellation.rs:323-328
__device__ __forceinline__ bool same_group(const Pipeline& p, int64_t a, int64_t b) {
  if (p.groups != nullptr && a < p.n_groups && b < p.n_groups) return p.groups[a] == p.groups[b];
  return false;
}
// tessellation.rs:298-319 (and updateable.rs:516-561 without involvement):
// (a, c) is a candidate pair when not in the same group and a < c, or c is a
// periodic image.
// Whether pair t has a contact: its contact byte where the run keeps them,
// else its record.
__device__ __forceinline__ bool has_contact(const Pipeline& p, int64_t t) {
  if (p.pair_contact != nullptr) return p.pair_contact[t] != 0;
  const PairRec& r = p.rec[t];
  return (r.flags & kRecValid) && r.area > 0.0;
}

// p.n and its reciprocal (canon_id).
inline void set_balls(Pipeline& p, int64_t n) {
  p.n = n;
  p.inv_n = n > 0 ? 1.0 / static_cast<double>(n) : 0.0;
}

// c mod n for a sphere id c (0 <= c < 27 n < 2^31: a ball or one of its
// images) without an integer division: the quotient from the reciprocal,
// which is off by at most one either way (c / n <= 27 carries a relative
// error of a few ulp), then corrected.
__device__ __forceinline__ int64_t canon_mod(int64_t c, int64_t n64, double inv_n) {
  const int32_t n = static_cast<int32_t>(n64);
  const int32_t ci = static_cast<int32_t>(c);
  const int32_t q = static_cast<int32_t>(static_cast<double>(ci) * inv_n);
  int32_t r = ci - q * n;
  r = r < 0 ? r + n : (r >= n ? r - n : r);
  return r;
}
__device__ __forceinline__ int64_t canon_id(const Pipeline& p, int64_t c) { return canon_mod(c, p.n, p.inv_n); }

__device__ __forceinline__ bool pair_qualifies(const Pipeline& p, int64_t a, int64_t c) {
  const int64_t cc = p.periodic ? canon_id(p, c) : c;
  if (same_group(p, a, cc)) return false;
  if (p.user_excl != nullptr && (p.user_excl[a] || p.user_excl[cc])) return false;
  return (p.periodic && c >= p.n) || a < c;
}

// Host-side launchers (k_grid.hip).
hipError_t launch_prepare(const Pipeline& p, hipStream_t s);
hipError_t launch_grid(const Pipeline& p, hipStream_t s);
hipError_t launch_scatter(const Pipeline& p, hipStream_t s);
hipError_t launch_neighbors(const Pipeline& p, hipStream_t s);
hipError_t launch_pairs(const Pipeline& p, hipStream_t s);
// Exclusive scan of counts[0..n) in place into counts[0..n], counts[n] = total
// (modulo 2^32); the total and a word that is 1 when the exact total exceeds
// INT32_MAX (ScanTotal), else 0, also go to *total_out and *beyond_out when
// given.
hipError_t launch_scan_u32(uint32_t* counts, int64_t n, uint32_t* tmp, hipStream_t s, uint32_t* total_out = nullptr,
                           uint32_t* beyond_out = nullptr);
// a[0 .. n) = v.
hipError_t launch_fill_words(uint32_t* a, int64_t n, uint32_t v, hipStream_t s);
// dst[k] = src[k * stride], k < count <= 64.
hipError_t launch_copy_strided(uint32_t* dst, const uint32_t* src, int count, int stride, hipStream_t s);
int64_t scan_tmp_words(int64_t n);

// vltr_noalloc.hip: the fixed device workspace of the allocation-free entry
// point, kNoAllocScratchBytes per GPU. noalloc_scratch_setup takes its
// address on the current device (once per device; the context creation
// calls it); noalloc_scratch returns that address (null before the setup).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
hipError_t noalloc_scratch_setup(int device);
void* noalloc_scratch(int device);

// k_contact.hip
hipError_t launch_clip(const Pipeline& p, hipStream_t s);
// Launch shapes of k_clip (same records from each): their number and names.
int clip_variants();
const char* clip_variant_name(int v);
// k_clip_wide and k_cut_slow over the pairs k_clip listed (device-side
// counts; no host synchronisation). p.slow_ws must hold slow_lanes(nb_cap)
// regions.
hipError_t launch_fallbacks(const Pipeline& p, hipStream_t s);
int64_t slow_lanes(int nb_cap);
void set_hexagon_constants(const double* s_half5, const double* c_half5);
hipError_t upload_hexagon_constants(hipStream_t s);

// k_adjoint.hip: grad (3 n, zeroed here) += dE/d(centres) for
// E = sum_contacts W[t_a][t_b] area + alpha sum sas + beta sum volume, from the
// records, topologies and seeds of the last full run; partial energy sums
// per block into part[0 .. adjoint_blocks()).
// Slots of the canonical-pair table (periodic runs): a power of two >= 2 n_pairs.
uint64_t dedup_slots(int64_t n_pairs);
// Fills the canonical-pair table (dedup_slots(p.n_pairs) slots) from this
// run's records: per canonical pair, the smallest pair index with a contact,
// among the contacts through an image (all == 0) or among all contacts
// (all != 0, the contacts-only output of a periodic run).
hipError_t launch_dedup_table(const Pipeline& p, unsigned long long* keys, uint32_t* first, int all, hipStream_t s);
// Scratch of the adjoint: flags (n_pairs + 1 words) and scan_tmp as for
// launch_scan_u32 over n_pairs, list (n_pairs entries): the contacts the
// adjoint visits.
struct AdjointWork {
  uint32_t* flags;
  uint32_t* scan_tmp;
  int32_t* list;
};
hipError_t launch_adjoint(const Pipeline& p, const int32_t* types, const double* w16, double* grad, double* part,
                          unsigned long long* dedup_keys, uint32_t* dedup_first, const AdjointWork& work,
                          hipStream_t s);
int64_t adjoint_blocks();

// k_solvent.hip: emit == 0 counts the kept icosahedron samples (K unit
// vectors in sih) of every exposed ball into cnt; emit != 0 (cnt scanned)
// writes them as balls of radius sr at comb[n + k] with their parents.
hipError_t launch_solvent_images(const Pipeline& p, uint32_t* ioff, uint32_t* cur, int32_t* img, uint32_t* scan_tmp,
                                 hipStream_t s);
hipError_t launch_solvent(const Pipeline& p, const uint32_t* ioff, const int32_t* img, const double* sih, int K,
                          double probe, double sr, uint32_t* cnt, int emit, double4* comb, uint32_t* parent,
                          hipStream_t s);
hipError_t launch_solvent_atoms(const double4* balls, int64_t n, double probe, double volume_probe, double4* comb,
                                hipStream_t s);

// k_cells.hip
hipError_t launch_cells(const Pipeline& p, hipStream_t s);
// Dedup cells: appends every first record through an image to the incoming
// list of its second end's canonical ball (p.dd_keys set, table filled).
hipError_t launch_dedup_incoming(const Pipeline& p, hipStream_t s);
// flag[i] = 1 for each pair with a contact, n_pairs words; with a
// canonical-pair table (dd_keys non-null, periodic runs: launch_dedup_table
// with all = p.mode) only for the first record of each canonical pair it holds.
hipError_t launch_flag_contacts(const Pipeline& p, unsigned long long* dd_keys, uint32_t* dd_first, uint32_t* flag,
                                hipStream_t s);
// out[pos[i]] = the contact of pair i (pos = scanned flags); periodic runs
// report canonical ids, ordered in contacts-only runs.
hipError_t launch_emit_contacts(const Pipeline& p, const uint32_t* pos, vltr_contact* out, hipStream_t s);
hipError_t launch_flag_cells(const Pipeline& p, uint32_t* flag, hipStream_t s);
hipError_t launch_emit_cells(const Pipeline& p, const uint32_t* pos, vltr_cell* out, hipStream_t s);
// Batched runs: out[k] = position of structure k's first contact (what = 0,
// pos = scanned contact flags) or first cell (what = 1, pos = scanned cell
// flags) in the compacted output, k = 0..n_segs (out[n_segs] = total).
hipError_t launch_seg_offsets(const Pipeline& p, const uint32_t* pos, int what, uint32_t* out, hipStream_t s);
// out[pos[i]] = p.net[i] for every pair i with a contact (pos = scanned contact
// flags).
hipError_t launch_emit_net(const Pipeline& p, const uint32_t* pos, ContactNet* out, hipStream_t s);
// k_network.hip: the tessellation network (contact.rs:120-177,
// tessellation.rs:334-401) from the contours in p.net.
// Id packing of the network records: `bits` bits per id, the code n for none.
struct NetLayout {
  int64_t n;
  int32_t periodic;
  int32_t bits;
};
NetLayout net_layout(int64_t n, int periodic);
// One kind of record (edges: 3 ids, vertices: 4): packed id keys, the
// record's own index, values (edge: length; vertex: x, y, z).
struct NetRecs {
  unsigned long long* lo;
  unsigned long long* hi;
  uint32_t* idx;
  double* val;
};
// Scratch of launch_net_sort for `count` records (net_sort_tmp_bytes bytes
// of tmp).
struct NetSortWork {
  void* tmp;
  size_t tmp_bytes;
  unsigned long long* keys_out;  // count
  unsigned long long* keys_mid;  // count
  uint32_t* perm1;               // count
  uint32_t* perm2;               // count
};
// ecnt[i], vcnt[i] = edge and vertex records of pair i (0 without a contact),
// n_pairs words each; a contour past kTopoCap points sets kStatNetTooLong.
hipError_t launch_net_count(const Pipeline& p, uint32_t* ecnt, uint32_t* vcnt, hipStream_t s);
// The records of every contact at the scanned offsets eoff, voff.
hipError_t launch_net_records(const Pipeline& p, const uint32_t* eoff, const uint32_t* voff, const NetLayout& L,
                              const NetRecs& E, const NetRecs& V, hipStream_t s);
// Bytes of rocprim scratch for sorting count keys of key_bits bits
// (SIZE_MAX: the query failed).
size_t net_sort_tmp_bytes(int64_t count, int key_bits);
// Stable order of the records by their keys (key_bits = ids x bits):
// *perm = the record indices in that order (W.perm1 or W.perm2).
hipError_t launch_net_sort(const NetRecs& R, int64_t count, int key_bits, const NetSortWork& W, const uint32_t** perm,
                           hipStream_t s);
// flag[j] = 1 where the j-th record in perm order starts a run of equal ids.
hipError_t launch_net_runs(const NetRecs& R, const uint32_t* perm, int64_t count, uint32_t* flag, hipStream_t s);
// out[pos[j]] = record perm[j] for each run start j (pos = scanned flags).
hipError_t launch_net_emit_edges(const NetRecs& R, const uint32_t* perm, const uint32_t* pos, int64_t count,
                                 const NetLayout& L, vltr_cell_edge* out, hipStream_t s);
hipError_t launch_net_emit_vertices(const NetRecs& R, const uint32_t* perm, const uint32_t* pos, int64_t count,
                                    const NetLayout& L, vltr_cell_vertex* out, hipStream_t s);
// Scratch of launch_contacts_by_first for `count` contacts.
struct ContactSortWork {
  void* tmp;  // contacts_sort_tmp_bytes(count) bytes
  size_t tmp_bytes;
  uint32_t* key;      // count
  uint32_t* key_out;  // count
  uint32_t* idx;      // count
  uint32_t* perm;     // count
  vltr_contact* out;  // count: the result
};
size_t contacts_sort_tmp_bytes(int64_t count);
// Periodic updateable summary (updateable.rs:333-376): W.out = the contacts
// c (ids < n) with their ids ordered, stably sorted by the smaller id.
hipError_t launch_contacts_by_first(const vltr_contact* c, int64_t count, int64_t n, const ContactSortWork& W,
                                    hipStream_t s);

// Cells under the updateable's rules (updateable.rs:633-727), in place in
// p.cell and copied to state: full = 1 after init or a full reinit, else only
// balls with affected[i] != 0 take the new record (the rest keep state[i]);
// excl holds the exclusion statuses (reset to the hidden status of each
// recomputed ball when refresh != 0).
hipError_t launch_updateable_cells(const Pipeline& p, CellRec* state, uint8_t* excl, const uint8_t* affected, int full,
                                   int refresh, hipStream_t s);
// balls[ids[k]] = vals[k], k < count.
hipError_t launch_scatter_balls(vltr_ball* balls, const int32_t* ids, const vltr_ball* vals, int64_t count,
                                hipStream_t s);
// affected[nb % n] = value, where still 0, for every neighbor nb of the balls
// changed[0..k) in p's neighbor lists.
hipError_t launch_mark_neighbors(const Pipeline& p, const int32_t* changed, int64_t k, uint8_t* affected,
                                 uint8_t value, hipStream_t s);
// mask[ids[t]] = value, t < k (ids outside [0, n) skipped).
hipError_t launch_mark_ids(uint8_t* mask, int64_t n, const int32_t* ids, int64_t k, uint8_t value, hipStream_t s);
// counts[0] = balls marked 1, counts[1] = balls marked 3 (counts zeroed here first).
hipError_t launch_count_marks(const uint8_t* mask, int64_t n, uint32_t* counts, hipStream_t s);
hipError_t launch_fill_u8(uint8_t* mask, int64_t n, uint8_t value, hipStream_t s);
// mask[i] = 0 where where[i] != 0.
hipError_t launch_mask_clear(uint8_t* mask, const uint8_t* where, int64_t n, hipStream_t s);
// rows[k * cap + j] = neighbor j of ball ids[k] (cnt[k] entries).
hipError_t launch_gather_rows(const Pipeline& p, const int32_t* ids, int64_t k, int32_t* cnt, int32_t* rows,
                              hipStream_t s);

}  // namespace vltr
