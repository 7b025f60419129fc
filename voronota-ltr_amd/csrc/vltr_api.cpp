// Host orchestration and C ABI (include/voronota_ltr.h).
//
// Replaces the drivers of src/tessellation.rs: compute_tessellation
// (:37-48, :199-292, :406-520) and compute_contacts_only (:125-188). Each call
// runs the device pipeline
//   prepare -> cell hash -> neighbor lists -> pair list -> clip -> cells ->
//   compaction
// on the context's stream; the host reads back three small words (pair total,
// status, contour overflows) and the compacted outputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <map>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <unordered_map>
#include <memory>
#include <mutex>
#include <vector>

#include "vltr_host.h"
#include "vltr_internal.h"
#include "voronota_ltr.h"

using namespace vltr;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define VLTR_HIP(expr)                                                                                   \
  do {                                                                                                   \
    hipError_t e_ = (expr);                                                                              \
    if (e_ != hipSuccess)                                                                                \
      return fail(e_ == hipErrorOutOfMemory ? VLTR_ERR_OUT_OF_MEMORY : VLTR_ERR_DEVICE,                  \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                                    \
  } while (0)

// Grow-only device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t want) {
    if (want <= bytes) return hipSuccess;
    if (ptr) {
      hipError_t e = hipFree(ptr);
      if (e != hipSuccess) return e;
      ptr = nullptr;
      bytes = 0;
    }
    const size_t sz = std::max<size_t>(want + want / 4, 256);
    hipError_t e = hipMalloc(&ptr, sz);
    if (e != hipSuccess) return e;
    bytes = sz;
    return hipSuccess;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

struct HostPinned {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t want) {
    if (want <= bytes) return hipSuccess;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
    const size_t sz = std::max<size_t>(want + want / 4, 4096);
    hipError_t e = hipHostMalloc(&ptr, sz, hipHostMallocDefault);
    if (e != hipSuccess) return e;
    bytes = sz;
    return hipSuccess;
  }
  void release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
};

constexpr int kPhases = 7;
// Regions of one device buffer, 256-byte aligned, taken in order (the
// allocation-free entry point's workspace, the network build's scratch).
struct Arena {
  unsigned char* base;  // null: only the bytes taken are counted
  size_t cap;
  size_t used = 0;
  template <class T>
  T* take(size_t count) {
    const size_t off = (used + 255) & ~size_t(255);
    used = off + count * sizeof(T);
    return (base != nullptr && used <= cap) ? reinterpret_cast<T*>(base + off) : nullptr;
  }
  bool fits() const { return used <= cap; }
};

// Neighbor slots per ball a run starts with (doubled when a list outgrows them).
constexpr int kNbCapDefault = 96;
// Bytes of neighbor and incoming lists up to which a call keeps the slot
// count an earlier call grew to.
constexpr double kNbListBudget = 2.0 * (1u << 30);
// Pinned host block of a context: the run's report words at offset 0, the
// allocation-free entry point's cell-hash segment staged at kPinnedSegOff.
constexpr size_t kPinnedBytes = 256;
constexpr size_t kPinnedSegOff = 128;

}  // namespace

struct vltr_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[kPhases + 1] = {};
  double timings[kPhases] = {};
  int nb_cap = kNbCapDefault;
  // device buffers
  DevBuf balls, groups, sph, seg_of, segs, bkey, brank, bcnt, sorted_id, sorted_sph, small, nb, nb_cnt, excluded,
      pair_cnt, pair_a, pair_b, rec, cell, overflow_list, wide_list, flags, scan_tmp, out_contacts, out_cells, slow_ws,
      in_cnt, incoming, seg_out, topo, seed, energy_part, types, grad, net, net_out, solv_cnt, solv_comb, solv_parent,
      solv_sih, solv_ioff, solv_cur, solv_img, nb_big,
      dedup_keys, dedup_first, adj_list, net_eoff, net_voff, net_work, pair_contact, topo_slot;
  int64_t topo_cap = 0;  // records of the topology store (adjoint runs)
  HostPinned pinned;
  // last device-resident outputs
  size_t last_contacts = 0, last_cells = 0;
  size_t last_n = 0;
  size_t last_structs = 0;  // structures of the last batched run (0: not batched)
  int last_nb_cap = 0;
  uint64_t counts[8] = {};
  int clip_variant = 0;  // launch shape of k_clip (set by the autotuner or the caller)
  // Batched device runs captured in a HIP graph (vltr_context_set_graphs):
  // replayed while the layout, input pointer, probe, mode, launch shape and
  // buffers stay those of the capture.
  bool use_graphs = false;
  hipGraphExec_t graph = nullptr;
  std::vector<uint64_t> graph_offsets;
  const void* graph_src = nullptr;
  double graph_probe = 0.0;
  int graph_mode = 0, graph_variant = 0, graph_nb_cap = 0;
  int64_t graph_cap = 0, graph_buckets = 0;
  std::vector<const void*> graph_bufs;
  uint64_t graph_replays = 0;
  // The shape of the last direct run without extras and its pair capacity
  // (its pair total with a margin): a call of the same shape enqueues the
  // back half without waiting for the pair total (run_pipeline).
  struct {
    bool valid = false;
    int64_t n = 0, n_structs = 0, n_groups = 0;
    double probe = 0.0;
    int mode = 0;
    bool periodic = false;
    const int32_t* groups = nullptr;
    int64_t cap = 0;
  } spec;
  uint64_t spec_runs = 0, spec_misses = 0;
};

struct vltr_solvent_result {
  std::vector<vltr_solvent_sphere> spheres;
};

struct vltr_batch_result {
  size_t n_structs = 0;
  std::vector<uint64_t> ball_off, contact_off, cell_off;  // n_structs + 1 each
  std::vector<vltr_contact> contacts;
  std::vector<vltr_cell> cells;
};


namespace {

// Hexagon rotation constants: the reference accumulates FRAC_PI_3 while the
// angle stays below TAU (contact.rs:455-469) and rotates by each angle with a
// unit quaternion built from sin/cos of half the angle (nalgebra).
void init_hexagon_constants() {
  static bool done = false;
  if (done) return;
  double s[5] = {}, c[5] = {};
  double rot = kFracPi3;
  int k = 0;
  while (rot < kTau && k < 5) {
    s[k] = std::sin(rot / 2.0);
    c[k] = std::cos(rot / 2.0);
    rot += kFracPi3;
    ++k;
  }
  set_hexagon_constants(s, c);
  done = true;
}

int kbits_for(int64_t n) {
  // 2^k cells per axis per structure: about cbrt(n / 8) (a compact structure
  // at protein density spans fewer cells than that), at least 4 per axis.
  int k = 2;
  while (k < 9 && (int64_t(1) << (3 * k)) * 8 < n) ++k;
  return k;
}

// The run's report block (kReport* words) in the context's small buffer.
uint32_t* report_words(vltr_context* ctx) { return reinterpret_cast<uint32_t*>(ctx->small.as<unsigned char>() + 8); }

void fill_pointers(vltr_context* ctx, Pipeline& p) {
  p.balls_in = ctx->balls.as<double4>();
  p.sph = ctx->sph.as<double4>();
  p.seg_of = ctx->seg_of.as<int32_t>();
  p.segs = ctx->segs.as<GridSeg>();
  p.bucket_key = ctx->bkey.as<uint32_t>();
  p.bucket_rank = ctx->brank.as<uint32_t>();
  p.bucket_cnt = ctx->bcnt.as<uint32_t>();
  p.sorted_id = ctx->sorted_id.as<int32_t>();
  p.sorted_sph = ctx->sorted_sph.as<double4>();
  unsigned char* small = ctx->small.as<unsigned char>();
  p.box_bits = reinterpret_cast<unsigned long long*>(small);
  // the report block: status, then the counters and totals (kReport*)
  p.status = report_words(ctx);
  p.overflow_cnt = p.status + kReportOverflow;
  p.wide_cnt = p.status + kReportWide;
  p.nb_big_cnt = p.status + kReportNbBig;
  p.bbox = reinterpret_cast<unsigned long long*>(small + 256);  // 6 words, kBboxStride apart
  p.clip_queue = reinterpret_cast<uint32_t*>(small + kSmallQueueOff);
  p.topo_counts = reinterpret_cast<uint32_t*>(small + kSmallTopoOff);
  p.nb_big = ctx->nb_big.as<int32_t>();
  p.nb = ctx->nb.as<int32_t>();
  p.nb_cnt = ctx->nb_cnt.as<int32_t>();
  p.excluded = ctx->excluded.as<uint8_t>();
  p.pair_cnt = ctx->pair_cnt.as<uint32_t>();
  p.pair_a = ctx->pair_a.as<int32_t>();
  p.pair_b = ctx->pair_b.as<int32_t>();
  p.rec = ctx->rec.as<PairRec>();
  p.pair_contact = ctx->pair_contact.as<uint8_t>();
  p.cell = ctx->cell.as<CellRec>();
  p.overflow_list = ctx->overflow_list.as<int32_t>();
  p.wide_list = ctx->wide_list.as<int32_t>();
  p.slow_ws = ctx->slow_ws.as<double>();
  p.in_cnt = ctx->in_cnt.as<uint32_t>();
  p.incoming = ctx->incoming.as<int32_t>();
}

// Runs the device pipeline on ctx->balls (already on the device). On return
// the compacted contacts/cells are in ctx->out_contacts / out_cells and their
// counts in *n_contacts / *n_cells.
// Extra inputs of the updateable path: balls excluded by the caller, and a
// step run on the cell records before they are compacted.
struct RunExtras {
  const uint8_t* user_excl = nullptr;
  std::function<hipError_t(const Pipeline&, hipStream_t)> post_cells;
  // adjoint runs: the cut topologies (a store of *topo_cap records, one per
  // contact with a stored topology, and a slot per pair; sized here once the
  // pair count is known, the capacity raised and the run redone when the
  // store is too small), cell seeds (n) and the energy weights
  DevBuf* topo_buf = nullptr;
  DevBuf* topo_slot_buf = nullptr;
  int64_t* topo_cap = nullptr;
  // network output: restricted contours per pair, compacted per contact
  DevBuf* net_buf = nullptr;
  DevBuf* net_out = nullptr;
  double* seed = nullptr;
  double sas_weight = 1.0, volume_weight = 1.0;
  // updateable steps: a step run once the neighbor lists are known (before
  // the clip), and the dirty mask with the previous run's pair lists,
  // records and topologies (null dirty: every pair is clipped)
  std::function<hipError_t(const Pipeline&, hipStream_t)> pre_clip;
  const uint8_t* dirty = nullptr;
  const uint32_t* prev_pair_cnt = nullptr;
  const int32_t* prev_pair_b = nullptr;
  const PairRec* prev_rec = nullptr;
  const ContactTopo* prev_topo = nullptr;
  const uint32_t* prev_topo_slot = nullptr;
  // every pair's record written (updateable runs)
  bool records_all = false;
  // called with the report block once the neighbor lists are complete; false
  // ends the run there (run_pipeline returns kRunAborted)
  std::function<bool(const uint32_t*)> front_check;
  // dedup cells (incremental steps of a periodic updateable): the
  // canonical-pair table's buffers (Pipeline::dd_keys)
  DevBuf* dd_keys = nullptr;
  DevBuf* dd_first = nullptr;
};

// Records per region of a topology store of at least tcap records.
int64_t topo_region_for(int64_t tcap) { return (tcap + kTopoRegions - 1) / kTopoRegions; }

// run_pipeline's return value when RunExtras::front_check ended the run.
constexpr int kRunAborted = -1000;

// Capture of a batched run into a HIP graph (see run_batched): the pair
// capacity it is captured with and the caller's device balls, which the
// graph's first node copies into the context.
struct GraphCapture {
  int64_t cap = 0;
  const void* src = nullptr;
  size_t src_bytes = 0;
  hipGraphExec_t exec = nullptr;  // out
};

void drop_graph(vltr_context* ctx);

// Back-half buffers for `cap` pairs (pairs, records, fallback lists, flags,
// compacted outputs) and the extras' per-pair buffers.
int ensure_back(vltr_context* ctx, Pipeline& p, int64_t cap, int64_t n, const RunExtras* extras) {
  p.pair_cap = std::max<int64_t>(cap, 1);
  VLTR_HIP(ctx->pair_a.ensure(p.pair_cap * sizeof(int32_t)));
  VLTR_HIP(ctx->pair_b.ensure(p.pair_cap * sizeof(int32_t)));
  VLTR_HIP(ctx->rec.ensure(p.pair_cap * sizeof(PairRec)));
  VLTR_HIP(ctx->pair_contact.ensure((p.pair_cap + 3) & ~int64_t(3)));
  VLTR_HIP(ctx->overflow_list.ensure(p.pair_cap * sizeof(int32_t)));
  VLTR_HIP(ctx->wide_list.ensure(p.pair_cap * sizeof(int32_t)));
  p.slow_ws_pairs = slow_lanes(p.nb_cap);
  {
    const int64_t cc = 6 + p.nb_cap, cr = 2 * cc;
    VLTR_HIP(ctx->slow_ws.ensure(static_cast<size_t>(p.slow_ws_pairs) * (5 * cc + 6 * cr) * sizeof(double)));
  }
  VLTR_HIP(ctx->flags.ensure((std::max<int64_t>(cap, n) + 2) * sizeof(uint32_t)));
  VLTR_HIP(ctx->scan_tmp.ensure(scan_tmp_words(std::max<int64_t>(cap, n) + 64) * sizeof(uint32_t) * 2));
  VLTR_HIP(ctx->out_contacts.ensure(p.pair_cap * sizeof(vltr_contact)));
  VLTR_HIP(ctx->out_cells.ensure(std::max<int64_t>(n, 1) * sizeof(vltr_cell)));
  fill_pointers(ctx, p);
  if (extras && extras->topo_buf) {
    // one record per clipped pair (contacts are about 30% of the candidate
    // pairs); a region too small makes the run start again (run_pipeline)
    int64_t& tcap = *extras->topo_cap;
    if (tcap < p.pair_cap / 2 + 8192) tcap = p.pair_cap / 2 + 8192;
    p.topo_region = topo_region_for(tcap);
    VLTR_HIP(extras->topo_buf->ensure(kTopoRegions * p.topo_region * sizeof(ContactTopo)));
    VLTR_HIP(extras->topo_slot_buf->ensure(p.pair_cap * sizeof(uint32_t)));
    p.topo = extras->topo_buf->as<ContactTopo>();
    p.topo_slot = extras->topo_slot_buf->as<uint32_t>();
  }
  if (extras && extras->dd_keys && extras->dd_first) {
    const uint64_t slots = dedup_slots(p.pair_cap);
    VLTR_HIP(extras->dd_keys->ensure(slots * sizeof(unsigned long long)));
    VLTR_HIP(extras->dd_first->ensure(slots * sizeof(uint32_t)));
  }
  if (p.periodic) {
    // the canonical-pair table of the periodic contact output
    const uint64_t slots = dedup_slots(p.pair_cap);
    VLTR_HIP(ctx->dedup_keys.ensure(slots * sizeof(unsigned long long)));
    VLTR_HIP(ctx->dedup_first.ensure(slots * sizeof(uint32_t)));
  }
  if (extras && extras->net_buf) {
    VLTR_HIP(extras->net_buf->ensure(p.pair_cap * sizeof(ContactNet)));
    p.net = extras->net_buf->as<ContactNet>();
    VLTR_HIP(ctx->net_eoff.ensure((p.pair_cap + 1) * sizeof(uint32_t)));
    VLTR_HIP(ctx->net_voff.ensure((p.pair_cap + 1) * sizeof(uint32_t)));
    if (extras->net_out) VLTR_HIP(extras->net_out->ensure(p.pair_cap * sizeof(ContactNet)));
  }
  return VLTR_OK;
}

// Scan scratch, contact flags and the compacted outputs of one run: the
// context's buffers, or regions of the fixed workspace and the caller's
// output buffers (the allocation-free entry point).
struct RunBufs {
  uint32_t* scan_tmp;
  uint32_t* flags;
  vltr_contact* contacts;
  vltr_cell* cells;
  uint32_t* seg_out;  // batched runs only
  // periodic runs: the canonical-pair table of the contact output
  unsigned long long* dd_keys;
  uint32_t* dd_first;
};
RunBufs ctx_bufs(vltr_context* ctx) {
  return RunBufs{ctx->scan_tmp.as<uint32_t>(), ctx->flags.as<uint32_t>(), ctx->out_contacts.as<vltr_contact>(),
                 ctx->out_cells.as<vltr_cell>(), ctx->seg_out.as<uint32_t>(),
                 ctx->dedup_keys.as<unsigned long long>(), ctx->dedup_first.as<uint32_t>()};
}

// Front half: cell-list build and neighbor lists with their scanned pair
// counts (pair_cnt[n] = the pair total).
int enqueue_front(vltr_context* ctx, const Pipeline& p, const RunBufs& b, bool events, hipStream_t s) {
  VLTR_HIP(launch_prepare(p, s));
  VLTR_HIP(launch_grid(p, s));
  VLTR_HIP(launch_scan_u32(p.bucket_cnt, p.n_buckets, b.scan_tmp, s));
  VLTR_HIP(launch_scatter(p, s));
  if (events) VLTR_HIP(hipEventRecord(ctx->ev[1], s));
  VLTR_HIP(launch_neighbors(p, s));
  // the pair total and whether its exact value exceeds INT32_MAX go to the
  // report block, read before any pair buffer is sized from them
  VLTR_HIP(launch_scan_u32(p.pair_cnt, p.n, b.scan_tmp, s, p.status + kReportPairs, p.status + kReportPairsBeyond));
  return VLTR_OK;
}

// The pair total of a run whose front half has completed (report block on
// the host), or an error when the pairs cannot be indexed by int32.
int checked_pairs(const uint32_t* host_words, int64_t* P) {
  *P = checked_pair_total(host_words[kReportPairs], host_words[kReportPairsBeyond]);
  if (*P < 0)
    return fail(VLTR_ERR_INVALID_ARGUMENT,
                "more than 2147483647 candidate pairs in one call (pair indices are 32-bit): split the input "
                "into smaller calls or batches");
  return VLTR_OK;
}

// Back half: pairs, clip, cells, compaction of contacts and cells; the run's
// report block (status, slow-path counts, pair, contact and cell totals) goes
// to host_words in one transfer.
int enqueue_back(vltr_context* ctx, Pipeline& p, const RunBufs& b, const RunExtras* extras, bool batched, bool events,
                 uint32_t* host_words, hipStream_t s) {
  const int64_t n = p.n;
  p.records_all = (extras != nullptr && extras->records_all) ? 1 : 0;
  VLTR_HIP(launch_pairs(p, s));
  if (extras && extras->pre_clip) VLTR_HIP(extras->pre_clip(p, s));
  if (extras && extras->dirty) {
    p.dirty = extras->dirty;
    p.prev_pair_cnt = extras->prev_pair_cnt;
    p.prev_pair_b = extras->prev_pair_b;
    p.prev_rec = extras->prev_rec;
    p.prev_topo = extras->prev_topo;
    p.prev_topo_slot = extras->prev_topo_slot;
  }
  // a contact whose topology no clip path records has none (no gradient)
  if (p.topo != nullptr) VLTR_HIP(launch_fill_words(p.topo_slot, p.n_pairs, kTopoSlotTooLong, s));
  if (events) VLTR_HIP(hipEventRecord(ctx->ev[3], s));
  VLTR_HIP(launch_clip(p, s));
  if (events) VLTR_HIP(hipEventRecord(ctx->ev[4], s));
  VLTR_HIP(launch_fallbacks(p, s));
  if (events) VLTR_HIP(hipEventRecord(ctx->ev[5], s));
  if (p.mode == 0 && p.periodic && extras && extras->dd_keys && extras->dd_first) {
    p.dd_keys = extras->dd_keys->as<unsigned long long>();
    p.dd_first = extras->dd_first->as<uint32_t>();
    VLTR_HIP(launch_dedup_table(p, p.dd_keys, p.dd_first, 0, s));
    VLTR_HIP(launch_dedup_incoming(p, s));
  }
  if (p.mode == 0) VLTR_HIP(launch_cells(p, s));
  if (p.mode == 0 && extras && extras->post_cells) VLTR_HIP(extras->post_cells(p, s));
  if (p.net != nullptr) {
    // network records per contact, counted over every contact: the
    // reference assembles the network before the periodic dedup
    // (tessellation.rs:474-478); the network is built from them afterwards
    uint32_t* ecnt = ctx->net_eoff.as<uint32_t>();
    uint32_t* vcnt = ctx->net_voff.as<uint32_t>();
    VLTR_HIP(launch_net_count(p, ecnt, vcnt, s));
    VLTR_HIP(launch_scan_u32(ecnt, p.n_pairs, b.scan_tmp, s, p.status + kReportEdges, p.status + kReportEdgesBeyond));
    VLTR_HIP(launch_scan_u32(vcnt, p.n_pairs, b.scan_tmp, s, p.status + kReportVerts, p.status + kReportVertsBeyond));
  }
  // periodic runs: of each canonical pair only its first record in pair order
  // (through an image in full runs, any in contacts-only runs) is reported
  unsigned long long* out_keys = nullptr;
  uint32_t* out_first = nullptr;
  if (p.periodic) {
    if (p.mode == 0 && p.dd_keys != nullptr) {
      out_keys = p.dd_keys;  // built before the cells, from the same records
      out_first = p.dd_first;
    } else {
      if (b.dd_keys == nullptr || b.dd_first == nullptr) return fail(VLTR_ERR_INTERNAL, "no canonical-pair table");
      out_keys = b.dd_keys;
      out_first = b.dd_first;
      VLTR_HIP(launch_dedup_table(p, out_keys, out_first, p.mode != 0 ? 1 : 0, s));
    }
  }
  uint32_t* flags = b.flags;
  VLTR_HIP(launch_flag_contacts(p, out_keys, out_first, flags, s));
  VLTR_HIP(launch_scan_u32(flags, p.n_pairs, b.scan_tmp, s, p.status + kReportContacts));
  VLTR_HIP(launch_emit_contacts(p, flags, b.contacts, s));
  uint32_t* seg_out = b.seg_out;
  if (batched) VLTR_HIP(launch_seg_offsets(p, flags, 0, seg_out, s));
  if (p.net != nullptr && extras->net_out != nullptr)
    VLTR_HIP(launch_emit_net(p, flags, extras->net_out->as<ContactNet>(), s));
  if (p.mode == 0) {
    VLTR_HIP(launch_flag_cells(p, flags, s));
    VLTR_HIP(launch_scan_u32(flags, n, b.scan_tmp, s, p.status + kReportCells));
    VLTR_HIP(launch_emit_cells(p, flags, b.cells, s));
    if (batched) VLTR_HIP(launch_seg_offsets(p, flags, 1, seg_out + p.n_segs + 1, s));
  }
  if (p.topo != nullptr)
    VLTR_HIP(launch_copy_strided(p.status + kReportTopos, p.topo_counts, kTopoRegions, kClipQueueStride, s));
  VLTR_HIP(hipMemcpyAsync(host_words, p.status, kReportWords * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (events) VLTR_HIP(hipEventRecord(ctx->ev[6], s));
  return VLTR_OK;
}

// Bookkeeping of a completed run (direct or graph replay).
void finish_run(vltr_context* ctx, const Pipeline& p, int64_t n_pairs, size_t n_structs, size_t* n_contacts,
                size_t* n_cells) {
  const uint32_t* host_words = static_cast<const uint32_t*>(ctx->pinned.ptr);
  *n_contacts = host_words[kReportContacts];
  *n_cells = p.mode == 0 ? host_words[kReportCells] : 0;
  ctx->last_n = static_cast<size_t>(p.n);
  ctx->last_structs = n_structs;
  ctx->last_nb_cap = p.nb_cap;
  ctx->counts[0] = static_cast<uint64_t>(p.n);
  ctx->counts[1] = static_cast<uint64_t>(p.n_pop);
  ctx->counts[2] = static_cast<uint64_t>(n_pairs);
  ctx->counts[3] = *n_contacts;
  ctx->counts[4] = *n_cells;
  ctx->counts[5] = static_cast<uint64_t>(host_words[kReportOverflow]) + host_words[kReportWide];
  ctx->counts[6] = static_cast<uint64_t>(p.nb_cap);
  ctx->counts[7] = static_cast<uint64_t>(p.n_buckets);
}

int run_pipeline(vltr_context* ctx, int64_t n, double probe, const vltr_periodic_box* pbox, const int32_t* d_groups,
                 int64_t n_groups, int mode, size_t* n_contacts, size_t* n_cells, const uint64_t* seg_offsets = nullptr,
                 int64_t n_structs = 0, const RunExtras* extras = nullptr, GraphCapture* gc = nullptr) {
  hipStream_t s = ctx->stream;
  const bool capture = gc != nullptr;
  // a direct run may move or resize the buffers a captured graph uses
  if (!capture) {
    drop_graph(ctx);
    // The neighbor slots per ball grow when a list overflows and are kept
    // for later calls, but only while the lists (and the incoming lists of
    // the same size) of this call stay within kNbListBudget: a large call
    // after a dense small one starts again from the default.
    while (ctx->nb_cap > kNbCapDefault && static_cast<double>(n) * ctx->nb_cap * 8.0 > kNbListBudget)
      ctx->nb_cap /= 2;
  }
  const bool periodic = pbox != nullptr;
  const int64_t n_pop = periodic ? 27 * n : n;
  Pipeline p{};
  set_balls(p, n);
  p.n_pop = n_pop;
  const bool batched = seg_offsets != nullptr && n_structs > 0;
  p.n_segs = batched ? static_cast<int32_t>(n_structs) : 1;
  p.periodic = periodic ? 1 : 0;
  p.mode = mode;
  p.probe = probe;
  if (pbox) {
    p.pbox.a = V3{pbox->shift_a[0], pbox->shift_a[1], pbox->shift_a[2]};
    p.pbox.b = V3{pbox->shift_b[0], pbox->shift_b[1], pbox->shift_b[2]};
    p.pbox.c = V3{pbox->shift_c[0], pbox->shift_c[1], pbox->shift_c[2]};
  }
  p.groups = d_groups;
  p.n_groups = n_groups;
  p.user_excl = extras ? extras->user_excl : nullptr;
  p.seed = extras ? extras->seed : nullptr;
  p.sas_weight = extras ? extras->sas_weight : 1.0;
  p.volume_weight = extras ? extras->volume_weight : 1.0;
  p.clip_variant = ctx->clip_variant;

  // One cell-hash segment per structure: its own bucket range, sized for it.
  std::vector<GridSeg> segs(static_cast<size_t>(p.n_segs));
  int64_t tab = 0;
  for (int32_t k = 0; k < p.n_segs; ++k) {
    GridSeg& sg = segs[k];
    const int64_t nk = batched ? static_cast<int64_t>(seg_offsets[k + 1] - seg_offsets[k]) : n_pop;
    sg.ball_off = batched ? static_cast<int64_t>(seg_offsets[k]) : 0;
    sg.tab_off = tab;
    sg.kbits = kbits_for(nk);
    sg.n = static_cast<int32_t>(nk);
    tab += int64_t(1) << (3 * sg.kbits);
  }
  if (tab >= (int64_t(1) << 32)) return fail(VLTR_ERR_INVALID_ARGUMENT, "too many cells");
  p.n_buckets = tab;

  VLTR_HIP(ctx->sph.ensure(n_pop * sizeof(double4)));
  VLTR_HIP(ctx->seg_of.ensure(n_pop * sizeof(int32_t)));
  VLTR_HIP(ctx->segs.ensure(segs.size() * sizeof(GridSeg)));
  VLTR_HIP(ctx->seg_out.ensure(2 * (segs.size() + 1) * sizeof(uint32_t)));
  VLTR_HIP(ctx->bkey.ensure(n_pop * sizeof(uint32_t)));
  VLTR_HIP(ctx->brank.ensure(n_pop * sizeof(uint32_t)));
  VLTR_HIP(ctx->bcnt.ensure((p.n_buckets + 1) * sizeof(uint32_t)));
  VLTR_HIP(ctx->sorted_id.ensure(n_pop * sizeof(int32_t)));
  VLTR_HIP(ctx->sorted_sph.ensure(n_pop * sizeof(double4)));
  VLTR_HIP(ctx->small.ensure(kSmallBytes));
  VLTR_HIP(ctx->nb_cnt.ensure(n * sizeof(int32_t)));
  VLTR_HIP(ctx->nb_big.ensure(n * sizeof(int32_t)));
  VLTR_HIP(ctx->excluded.ensure(n));
  VLTR_HIP(ctx->pair_cnt.ensure((n + 1) * sizeof(uint32_t)));
  VLTR_HIP(ctx->cell.ensure(n * sizeof(CellRec)));
  VLTR_HIP(ctx->in_cnt.ensure(n * sizeof(uint32_t)));
  VLTR_HIP(ctx->scan_tmp.ensure(scan_tmp_words(std::max<int64_t>(p.n_buckets, 1) + n + 64) * sizeof(uint32_t) * 2));
  VLTR_HIP(ctx->pinned.ensure(kPinnedBytes));
  uint32_t* host_words = static_cast<uint32_t*>(ctx->pinned.ptr);

  if (capture) {
    // every buffer at its capacity first (no allocation under capture); the
    // segment table is the one the direct run of this layout uploaded
    p.nb_cap = ctx->nb_cap;
    VLTR_HIP(ctx->nb.ensure(static_cast<size_t>(n) * p.nb_cap * sizeof(int32_t)));
    VLTR_HIP(ctx->incoming.ensure(static_cast<size_t>(n) * p.nb_cap * sizeof(int32_t)));
    VLTR_HIP(ctx->balls.ensure(gc->src_bytes));
    fill_pointers(ctx, p);
    int rc = ensure_back(ctx, p, gc->cap, n, extras);
    if (rc != VLTR_OK) return rc;
    p.n_pairs = p.pair_cap;
    p.n_pairs_dev = p.pair_cnt + n;
    p.n_pairs_beyond_dev = p.status + kReportPairsBeyond;
    VLTR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = [&]() -> int {
      VLTR_HIP(hipMemcpyAsync(ctx->balls.ptr, gc->src, gc->src_bytes, hipMemcpyDeviceToDevice, s));
      int r = enqueue_front(ctx, p, ctx_bufs(ctx), false, s);
      if (r != VLTR_OK) return r;
      return enqueue_back(ctx, p, ctx_bufs(ctx), extras, batched, false, host_words, s);
    }();
    hipGraph_t graph = nullptr;
    const hipError_t e_end = hipStreamEndCapture(s, &graph);
    if (rc != VLTR_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    VLTR_HIP(e_end);
    const hipError_t e_inst = hipGraphInstantiate(&gc->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    VLTR_HIP(e_inst);
    return VLTR_OK;
  }

  VLTR_HIP(hipMemcpyAsync(ctx->segs.ptr, segs.data(), segs.size() * sizeof(GridSeg), hipMemcpyHostToDevice, s));
  // A call repeating the last plain call's shape: the back half follows the
  // front half at once, its pair buffers sized from the last pair total with
  // a margin and the pair kernels stopping at this run's total on the device
  // (as a graph replay does), so the run synchronizes once. A total beyond
  // the buffers, or neighbor lists beyond their slots, turn the pair kernels
  // into no-ops and the run is redone below, synchronizing after the front.
  const bool speculate = extras == nullptr && ctx->spec.valid && ctx->spec.n == n &&
                         ctx->spec.n_structs == (batched ? n_structs : 0) && ctx->spec.probe == probe &&
                         ctx->spec.mode == mode && ctx->spec.periodic == periodic && ctx->spec.groups == d_groups &&
                         ctx->spec.n_groups == n_groups;
  if (speculate) {
    p.nb_cap = ctx->nb_cap;
    VLTR_HIP(ctx->nb.ensure(static_cast<size_t>(n) * p.nb_cap * sizeof(int32_t)));
    VLTR_HIP(ctx->incoming.ensure(static_cast<size_t>(n) * p.nb_cap * sizeof(int32_t)));
    fill_pointers(ctx, p);
    int rc = ensure_back(ctx, p, ctx->spec.cap, n, nullptr);
    if (rc != VLTR_OK) return rc;
    p.n_pairs = p.pair_cap;
    p.n_pairs_dev = p.pair_cnt + n;
    p.n_pairs_beyond_dev = p.status + kReportPairsBeyond;
    VLTR_HIP(hipEventRecord(ctx->ev[0], s));
    rc = enqueue_front(ctx, p, ctx_bufs(ctx), true, s);
    if (rc != VLTR_OK) return rc;
    VLTR_HIP(hipEventRecord(ctx->ev[2], s));
    rc = enqueue_back(ctx, p, ctx_bufs(ctx), nullptr, batched, true, host_words, s);
    if (rc != VLTR_OK) return rc;
    VLTR_HIP(hipStreamSynchronize(s));
    ++ctx->spec_runs;
    const int64_t total = static_cast<int64_t>(host_words[kReportPairs]);
    if (!(host_words[kReportStatus] & (kStatNbOverflow | kStatNbTooLong | kStatPairOverflow)) &&
        host_words[kReportPairsBeyond] == 0u && total <= p.pair_cap) {
      finish_run(ctx, p, total, batched ? static_cast<size_t>(n_structs) : 0, n_contacts, n_cells);
      float ms = 0.0f;
      for (int k = 0; k < kPhases - 1; ++k) {
        if (hipEventElapsedTime(&ms, ctx->ev[k], ctx->ev[k + 1]) == hipSuccess) ctx->timings[k] = ms;
      }
      if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[kPhases - 1]) == hipSuccess) ctx->timings[kPhases - 1] = ms;
      return VLTR_OK;
    }
    ++ctx->spec_misses;
    p.n_pairs_dev = nullptr;
    p.n_pairs_beyond_dev = nullptr;
  }
 restart:
  VLTR_HIP(hipEventRecord(ctx->ev[0], s));
  for (;;) {
    p.nb_cap = ctx->nb_cap;
    VLTR_HIP(ctx->nb.ensure(static_cast<size_t>(n) * p.nb_cap * sizeof(int32_t)));
    VLTR_HIP(ctx->incoming.ensure(static_cast<size_t>(n) * p.nb_cap * sizeof(int32_t)));
    fill_pointers(ctx, p);
    const int rc = enqueue_front(ctx, p, ctx_bufs(ctx), true, s);
    if (rc != VLTR_OK) return rc;
    VLTR_HIP(hipMemcpyAsync(host_words, p.status, kReportWords * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    VLTR_HIP(hipEventRecord(ctx->ev[2], s));
    VLTR_HIP(hipStreamSynchronize(s));
    if (host_words[kReportStatus] & kStatNbTooLong)
      return fail(VLTR_ERR_INTERNAL, "neighbor list longer than 4096 entries");
    if (host_words[kReportStatus] & kStatNbOverflow) {
      if (ctx->nb_cap >= kNbMax) return fail(VLTR_ERR_INTERNAL, "neighbor list longer than 4096 entries");
      ctx->nb_cap *= 2;
      continue;
    }
    break;
  }
  if (extras && extras->front_check && !extras->front_check(host_words)) return kRunAborted;
  int64_t P = 0;
  int rc = checked_pairs(host_words, &P);
  if (rc != VLTR_OK) return rc;
  rc = ensure_back(ctx, p, P, n, extras);
  if (rc != VLTR_OK) return rc;
  p.n_pairs = P;
  rc = enqueue_back(ctx, p, ctx_bufs(ctx), extras, batched, true, host_words, s);
  if (rc != VLTR_OK) return rc;
  VLTR_HIP(hipStreamSynchronize(s));
  if (host_words[kReportStatus] & kStatNbOverflow) {
    // an incoming-pair list outgrew nb_cap (a hidden ball with many
    // neighbors): redo everything with longer lists
    if (ctx->nb_cap >= kNbMax) return fail(VLTR_ERR_INTERNAL, "neighbor list longer than 4096 entries");
    ctx->nb_cap *= 2;
    goto restart;
  }
  if (host_words[kReportStatus] & kStatPairOverflow)
    return fail(VLTR_ERR_INTERNAL, "pair lists outgrew the buffers sized from the pair total");
  if (p.topo != nullptr) {
    int64_t most = 0;
    for (int x = 0; x < kTopoRegions; ++x) most = std::max<int64_t>(most, host_words[kReportTopos + x]);
    if (most > p.topo_region) {
      // a region of the topology store was too small: redo with larger ones
      *extras->topo_cap = kTopoRegions * (most + most / 8 + 1024);
      goto restart;
    }
  }
  finish_run(ctx, p, P, batched ? static_cast<size_t>(n_structs) : 0, n_contacts, n_cells);
  ctx->spec.valid = extras == nullptr;
  ctx->spec.n = n;
  ctx->spec.n_structs = batched ? n_structs : 0;
  ctx->spec.probe = probe;
  ctx->spec.mode = mode;
  ctx->spec.periodic = periodic;
  ctx->spec.groups = d_groups;
  ctx->spec.n_groups = n_groups;
  ctx->spec.cap = P + P / 4 + 4096;
  float ms = 0.0f;
  for (int k = 0; k < kPhases - 1; ++k) {
    if (hipEventElapsedTime(&ms, ctx->ev[k], ctx->ev[k + 1]) == hipSuccess) ctx->timings[k] = ms;
  }
  if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[kPhases - 1]) == hipSuccess) ctx->timings[kPhases - 1] = ms;
  return VLTR_OK;
}

// The tessellation network (contact.rs:120-177, tessellation.rs:334-401) of
// the run just completed on ctx, built on the device: the contours in
// ctx->net and the record counts per pair scanned by enqueue_back become
// edge and vertex records with canonical sorted ids, ordered by their ids
// (stable), the first of each run kept; then two transfers to the host.
int device_network(vltr_context* ctx, int64_t n, bool periodic, vltr_result* r) {
  hipStream_t s = ctx->stream;
  uint32_t* hw = static_cast<uint32_t*>(ctx->pinned.ptr);
  if (hw[kReportStatus] & kStatNetTooLong)
    return fail(VLTR_ERR_INTERNAL, "network output: a contour has more than 32 points");
  const int64_t ne = checked_pair_total(hw[kReportEdges], hw[kReportEdgesBeyond]);
  const int64_t nv = checked_pair_total(hw[kReportVerts], hw[kReportVertsBeyond]);
  if (ne < 0 || nv < 0)
    return fail(VLTR_ERR_INVALID_ARGUMENT, "more than 2147483647 network edges or vertices in one call: split the input");
  r->has_network = true;
  r->edges.clear();
  r->vertices.clear();
  if (ne == 0) return VLTR_OK;  // every contact has an edge: no contacts
  const NetLayout L = net_layout(n, periodic ? 1 : 0);
  const int64_t m = std::max(ne, nv);
  const size_t tmp_e = net_sort_tmp_bytes(ne, 3 * L.bits), tmp_v = net_sort_tmp_bytes(nv, 4 * L.bits);
  if (tmp_e == SIZE_MAX || tmp_v == SIZE_MAX) return fail(VLTR_ERR_DEVICE, "network sort scratch query");
  NetRecs E{}, V{};
  NetSortWork W{};
  uint32_t* flag = nullptr;
  vltr_cell_edge* out_e = nullptr;
  vltr_cell_vertex* out_v = nullptr;
  auto carve = [&](Arena& A) {
    E = NetRecs{A.take<unsigned long long>(ne), A.take<unsigned long long>(ne), A.take<uint32_t>(ne),
                A.take<double>(ne)};
    V = NetRecs{A.take<unsigned long long>(nv), A.take<unsigned long long>(nv), A.take<uint32_t>(nv),
                A.take<double>(3 * nv)};
    W.tmp_bytes = std::max(tmp_e, tmp_v);
    W.tmp = A.take<unsigned char>(W.tmp_bytes);
    W.keys_out = A.take<unsigned long long>(m);
    W.keys_mid = A.take<unsigned long long>(m);
    W.perm1 = A.take<uint32_t>(m);
    W.perm2 = A.take<uint32_t>(m);
    flag = A.take<uint32_t>(m + 1);
    out_e = A.take<vltr_cell_edge>(ne);
    out_v = A.take<vltr_cell_vertex>(nv);
  };
  Arena size{nullptr, SIZE_MAX};
  carve(size);
  VLTR_HIP(ctx->net_work.ensure(size.used));
  Arena A{ctx->net_work.as<unsigned char>(), ctx->net_work.bytes};
  carve(A);
  if (!A.fits()) return fail(VLTR_ERR_INTERNAL, "network scratch");
  VLTR_HIP(ctx->scan_tmp.ensure(scan_tmp_words(m + 64) * sizeof(uint32_t) * 2));
  uint32_t* scan_tmp = ctx->scan_tmp.as<uint32_t>();

  Pipeline p{};
  fill_pointers(ctx, p);
  set_balls(p, n);
  p.n_pairs = static_cast<int64_t>(ctx->counts[2]);
  p.periodic = periodic ? 1 : 0;
  p.net = ctx->net.as<ContactNet>();
  VLTR_HIP(launch_net_records(p, ctx->net_eoff.as<uint32_t>(), ctx->net_voff.as<uint32_t>(), L, E, V, s));
  const uint32_t* perm = nullptr;
  VLTR_HIP(launch_net_sort(E, ne, 3 * L.bits, W, &perm, s));
  VLTR_HIP(launch_net_runs(E, perm, ne, flag, s));
  VLTR_HIP(launch_scan_u32(flag, ne, scan_tmp, s, p.status + kReportEdges));
  VLTR_HIP(launch_net_emit_edges(E, perm, flag, ne, L, out_e, s));
  if (nv > 0) {
    VLTR_HIP(launch_net_sort(V, nv, 4 * L.bits, W, &perm, s));
    VLTR_HIP(launch_net_runs(V, perm, nv, flag, s));
    VLTR_HIP(launch_scan_u32(flag, nv, scan_tmp, s, p.status + kReportVerts));
    VLTR_HIP(launch_net_emit_vertices(V, perm, flag, nv, L, out_v, s));
  }
  VLTR_HIP(hipMemcpyAsync(hw, p.status, kReportWords * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  VLTR_HIP(hipStreamSynchronize(s));
  const size_t ue = hw[kReportEdges], uv = nv > 0 ? hw[kReportVerts] : 0;
  if (ue > static_cast<size_t>(ne) || uv > static_cast<size_t>(nv)) return fail(VLTR_ERR_INTERNAL, "network totals");
  r->edges.resize(ue);
  r->vertices.resize(uv);
  if (ue) VLTR_HIP(hipMemcpyAsync(r->edges.data(), out_e, ue * sizeof(vltr_cell_edge), hipMemcpyDeviceToHost, s));
  if (uv) VLTR_HIP(hipMemcpyAsync(r->vertices.data(), out_v, uv * sizeof(vltr_cell_vertex), hipMemcpyDeviceToHost, s));
  VLTR_HIP(hipStreamSynchronize(s));
  return VLTR_OK;
}

int host_entry(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe, const vltr_periodic_box* pbox,
               const int32_t* groups, size_t n_groups, int with_vertices, int mode, vltr_result** out,
               std::vector<ContactNet>* keep_nets = nullptr) {
  if (!ctx || !out || (n > 0 && !balls)) return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  if (n > (size_t(1) << 31) / 27) return fail(VLTR_ERR_INVALID_ARGUMENT, "too many balls");
  *out = nullptr;
  auto* r = new (std::nothrow) vltr_result();
  if (!r) return fail(VLTR_ERR_OUT_OF_MEMORY, "result allocation");
  r->num_balls = mode == 0 ? n : n;
  if (n == 0) {
    *out = r;
    return VLTR_OK;
  }
  if (hipSetDevice(ctx->device) != hipSuccess) {
    delete r;
    return fail(VLTR_ERR_DEVICE, "hipSetDevice failed");
  }
  hipStream_t s = ctx->stream;
  int rc = VLTR_OK;
  do {
    if (ctx->balls.ensure(n * sizeof(vltr_ball)) != hipSuccess) {
      rc = fail(VLTR_ERR_OUT_OF_MEMORY, "ball buffer");
      break;
    }
    if (hipMemcpyAsync(ctx->balls.ptr, balls, n * sizeof(vltr_ball), hipMemcpyHostToDevice, s) != hipSuccess) {
      rc = fail(VLTR_ERR_DEVICE, "ball upload");
      break;
    }
    const int32_t* dg = nullptr;
    if (groups && n_groups > 0) {
      if (ctx->groups.ensure(n_groups * sizeof(int32_t)) != hipSuccess ||
          hipMemcpyAsync(ctx->groups.ptr, groups, n_groups * sizeof(int32_t), hipMemcpyHostToDevice, s) != hipSuccess) {
        rc = fail(VLTR_ERR_DEVICE, "group upload");
        break;
      }
      dg = ctx->groups.as<int32_t>();
    }
    size_t nc = 0, ncell = 0;
    RunExtras ex;
    const bool net = with_vertices && mode == 0;
    if (net) {
      ex.net_buf = &ctx->net;
      // the contours compacted per contact only for callers that keep them
      if (keep_nets) ex.net_out = &ctx->net_out;
    }
    rc = run_pipeline(ctx, static_cast<int64_t>(n), probe, pbox, dg, dg ? static_cast<int64_t>(n_groups) : 0, mode, &nc,
                      &ncell, nullptr, 0, net ? &ex : nullptr);
    if (rc != VLTR_OK) break;
    r->contacts.resize(nc);
    r->cells.resize(ncell);
    if (nc && hipMemcpyAsync(r->contacts.data(), ctx->out_contacts.ptr, nc * sizeof(vltr_contact),
                             hipMemcpyDeviceToHost, s) != hipSuccess) {
      rc = fail(VLTR_ERR_DEVICE, "contact download");
      break;
    }
    if (ncell && hipMemcpyAsync(r->cells.data(), ctx->out_cells.ptr, ncell * sizeof(vltr_cell), hipMemcpyDeviceToHost,
                                s) != hipSuccess) {
      rc = fail(VLTR_ERR_DEVICE, "cell download");
      break;
    }
    std::vector<ContactNet> nets;
    if (net && keep_nets && nc) {
      nets.resize(nc);
      if (hipMemcpyAsync(nets.data(), ctx->net_out.ptr, nc * sizeof(ContactNet), hipMemcpyDeviceToHost, s) !=
          hipSuccess) {
        rc = fail(VLTR_ERR_DEVICE, "contour download");
        break;
      }
    }
    if (hipStreamSynchronize(s) != hipSuccess) {
      rc = fail(VLTR_ERR_DEVICE, "stream synchronize");
      break;
    }
    if (net) {
      rc = device_network(ctx, static_cast<int64_t>(n), pbox != nullptr, r);
      if (rc != VLTR_OK) break;
      if (keep_nets) keep_nets->swap(nets);
    }
  } while (false);
  if (rc != VLTR_OK) {
    delete r;
    return rc;
  }
  *out = r;
  return VLTR_OK;
}

// A double as Rust's Display prints it: the shortest digits that read back
// to the same value, in positional notation; "NaN", "inf", "-inf".
std::string rust_display(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
  char buf[64];
  int prec = 1;
  for (; prec < 17; ++prec) {
    std::snprintf(buf, sizeof buf, "%.*e", prec - 1, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  std::snprintf(buf, sizeof buf, "%.*e", prec - 1, v);
  // buf = [-]d.ddde[+-]X: place the point X digits to the right
  std::string str(buf);
  const bool neg = str[0] == '-';
  if (neg) str.erase(0, 1);
  const size_t epos = str.find('e');
  const int ex = std::atoi(str.c_str() + epos + 1);
  std::string digits = str.substr(0, epos);
  digits.erase(std::remove(digits.begin(), digits.end(), '.'), digits.end());
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  std::string out;
  const int point = ex + 1;  // digits before the decimal point
  if (point <= 0) {
    out = "0." + std::string(static_cast<size_t>(-point), '0') + digits;
  } else if (point >= static_cast<int>(digits.size())) {
    out = digits + std::string(static_cast<size_t>(point) - digits.size(), '0');
  } else {
    out = digits.substr(0, static_cast<size_t>(point)) + "." + digits.substr(static_cast<size_t>(point));
  }
  if (out == "0") return neg ? "-0" : "0";
  return (neg ? "-" : "") + out;
}

// subdivided_icosahedron.rs:51-131: unit vertices in order of first
// appearance (base vertices, then edge midpoints as the faces are walked).
std::vector<double> icosahedron(int depth) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  const double base[12][3] = {{t, 1.0, 0.0},  {-t, 1.0, 0.0},  {t, -1.0, 0.0}, {-t, -1.0, 0.0},
                              {1.0, 0.0, t},  {1.0, 0.0, -t},  {-1.0, 0.0, t}, {-1.0, 0.0, -t},
                              {0.0, t, 1.0},  {0.0, -t, 1.0},  {0.0, t, -1.0}, {0.0, -t, -1.0}};
  std::vector<V3> v;
  for (const auto& b : base) v.push_back(unit(V3{b[0], b[1], b[2]}));
  std::vector<std::array<int, 3>> tri = {{0, 8, 4},  {1, 10, 7}, {2, 9, 11}, {7, 3, 1},  {0, 5, 10},
                                         {3, 9, 6},  {3, 11, 9}, {8, 6, 4},  {2, 4, 9},  {3, 7, 11},
                                         {4, 2, 0},  {9, 4, 6},  {2, 11, 5}, {0, 10, 8}, {5, 0, 2},
                                         {10, 5, 7}, {1, 6, 8},  {1, 8, 10}, {6, 1, 3},  {11, 7, 5}};
  for (int d = 0; d < depth; ++d) {
    std::vector<std::array<int, 3>> next;
    std::map<std::pair<int, int>, int> mids;
    for (const auto& f : tri) {
      const std::pair<int, int> edges[3] = {{f[1], f[2]}, {f[0], f[2]}, {f[0], f[1]}};
      int mid[3];
      for (int e = 0; e < 3; ++e) {
        const int a = edges[e].first, b = edges[e].second;
        const auto key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
        auto it = mids.find(key);
        if (it == mids.end()) {
          v.push_back(unit((v[a] + v[b]) * 0.5));
          it = mids.emplace(key, static_cast<int>(v.size()) - 1).first;
        }
        mid[e] = it->second;
      }
      next.push_back({f[0], mid[1], mid[2]});
      next.push_back({f[1], mid[0], mid[2]});
      next.push_back({f[2], mid[0], mid[1]});
      next.push_back({mid[0], mid[1], mid[2]});
    }
    tri.swap(next);
  }
  std::vector<double> out;
  for (const auto& p : v) {
    out.push_back(p.x);
    out.push_back(p.y);
    out.push_back(p.z);
  }
  return out;
}

// solvent_spheres.rs:131-168 on the device: stage 1, sampling of the exposed
// atoms (k_solvent), stage 2 on atoms + samples, weights = stage-2 volumes.
int solvent_spheres(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe, double volume_probe, int depth,
                    const vltr_periodic_box* pbox, std::vector<vltr_solvent_sphere>* out) {
  out->clear();
  if (n == 0) return VLTR_OK;
  VLTR_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const std::vector<double> sih = icosahedron(depth);
  const int K = static_cast<int>(sih.size() / 3);
  VLTR_HIP(ctx->solv_sih.ensure(sih.size() * sizeof(double)));
  VLTR_HIP(hipMemcpyAsync(ctx->solv_sih.ptr, sih.data(), sih.size() * sizeof(double), hipMemcpyHostToDevice, s));
  VLTR_HIP(ctx->solv_cnt.ensure((n + 1) * sizeof(uint32_t)));
  VLTR_HIP(ctx->balls.ensure(n * sizeof(vltr_ball)));
  VLTR_HIP(hipMemcpyAsync(ctx->balls.ptr, balls, n * sizeof(vltr_ball), hipMemcpyHostToDevice, s));
  const double sr = probe + volume_probe;
  Pipeline saved{};
  RunExtras ex;
  uint32_t* cnt = ctx->solv_cnt.as<uint32_t>();
  const double* d_sih = ctx->solv_sih.as<double>();
  const uint32_t* ioff = nullptr;
  const int32_t* img = nullptr;
  ex.post_cells = [&](const Pipeline& p, hipStream_t st) -> hipError_t {
    saved = p;
    uint32_t* tmp = ctx->scan_tmp.as<uint32_t>();
    hipError_t e = hipSuccess;
    if (p.periodic) {
      if ((e = ctx->solv_ioff.ensure((n + 1) * sizeof(uint32_t))) != hipSuccess) return e;
      if ((e = ctx->solv_cur.ensure(n * sizeof(uint32_t))) != hipSuccess) return e;
      if ((e = ctx->solv_img.ensure(std::max<int64_t>(p.n_pairs, 1) * sizeof(int32_t))) != hipSuccess) return e;
      ioff = ctx->solv_ioff.as<uint32_t>();
      img = ctx->solv_img.as<int32_t>();
      e = launch_solvent_images(p, ctx->solv_ioff.as<uint32_t>(), ctx->solv_cur.as<uint32_t>(),
                                ctx->solv_img.as<int32_t>(), tmp, st);
      if (e != hipSuccess) return e;
    }
    e = launch_solvent(p, ioff, img, d_sih, K, probe, sr, cnt, 0, nullptr, nullptr, st);
    if (e != hipSuccess) return e;
    return launch_scan_u32(cnt, static_cast<int64_t>(n), tmp, st);
  };
  size_t nc = 0, ncell = 0;
  int rc = run_pipeline(ctx, static_cast<int64_t>(n), probe, pbox, nullptr, 0, 0, &nc, &ncell, nullptr, 0, &ex);
  if (rc != VLTR_OK) return rc;
  uint32_t m = 0;
  VLTR_HIP(hipMemcpy(&m, cnt + n, sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (m == 0) return VLTR_OK;
  const size_t total = n + m;
  if (total > (size_t(1) << 31) / 27) return fail(VLTR_ERR_INVALID_ARGUMENT, "too many solvent spheres");
  VLTR_HIP(ctx->solv_comb.ensure(total * sizeof(vltr_ball)));
  VLTR_HIP(ctx->solv_parent.ensure(m * sizeof(uint32_t)));
  double4* comb = ctx->solv_comb.as<double4>();
  VLTR_HIP(launch_solvent_atoms(saved.balls_in, static_cast<int64_t>(n), probe, volume_probe, comb, s));
  VLTR_HIP(launch_solvent(saved, ioff, img, d_sih, K, probe, sr, cnt, 1, comb, ctx->solv_parent.as<uint32_t>(), s));
  std::vector<vltr_ball> pts(m);
  std::vector<uint32_t> parents(m);
  VLTR_HIP(hipMemcpyAsync(pts.data(), comb + n, m * sizeof(vltr_ball), hipMemcpyDeviceToHost, s));
  VLTR_HIP(hipMemcpyAsync(parents.data(), ctx->solv_parent.ptr, m * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  // stage 2 (probe 0) on atoms + samples; the stage-1 buffers are reused
  VLTR_HIP(hipStreamSynchronize(s));
  VLTR_HIP(ctx->balls.ensure(total * sizeof(vltr_ball)));
  VLTR_HIP(hipMemcpyAsync(ctx->balls.ptr, comb, total * sizeof(vltr_ball), hipMemcpyDeviceToDevice, s));
  rc = run_pipeline(ctx, static_cast<int64_t>(total), 0.0, pbox, nullptr, 0, 0, &nc, &ncell);
  if (rc != VLTR_OK) return rc;
  std::vector<CellRec> cells(m);
  VLTR_HIP(hipMemcpy(cells.data(), ctx->cell.as<CellRec>() + n, m * sizeof(CellRec), hipMemcpyDeviceToHost));
  out->resize(m);
  for (uint32_t k = 0; k < m; ++k) {
    vltr_solvent_sphere& o = (*out)[k];
    o.x = pts[k].x;
    o.y = pts[k].y;
    o.z = pts[k].z;
    o.radius = probe;
    o.weight = cells[k].stage == 2 ? cells[k].volume : 0.0;
    o.parent_index = parents[k];
  }
  return VLTR_OK;
}

int check_offsets(const uint64_t* offsets, size_t n_structs) {
  if (!offsets || n_structs == 0) return fail(VLTR_ERR_INVALID_ARGUMENT, "no structures");
  if (offsets[0] != 0) return fail(VLTR_ERR_INVALID_ARGUMENT, "offsets[0] must be 0");
  for (size_t k = 0; k < n_structs; ++k)
    if (offsets[k + 1] < offsets[k]) return fail(VLTR_ERR_INVALID_ARGUMENT, "offsets must not decrease");
  if (offsets[n_structs] > (uint64_t(1) << 31) - 1) return fail(VLTR_ERR_INVALID_ARGUMENT, "too many balls");
  if (n_structs > (size_t(1) << 24)) return fail(VLTR_ERR_INVALID_ARGUMENT, "too many structures");
  return VLTR_OK;
}

// Batched run on ctx->balls (already on the device); per-structure offsets of
// the outputs are left in ctx->seg_out.
void drop_graph(vltr_context* ctx) {
  if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
  ctx->graph = nullptr;
}

// Addresses of every buffer a captured pipeline reads or writes.
std::vector<const void*> graph_buffers(const vltr_context* ctx) {
  std::vector<const void*> v;
  for (const DevBuf* b : {&ctx->balls, &ctx->sph, &ctx->seg_of, &ctx->segs, &ctx->seg_out, &ctx->bkey, &ctx->brank,
                          &ctx->bcnt, &ctx->sorted_id, &ctx->sorted_sph, &ctx->small, &ctx->nb, &ctx->nb_cnt,
                          &ctx->nb_big, &ctx->excluded, &ctx->pair_cnt, &ctx->pair_a, &ctx->pair_b, &ctx->rec,
                          &ctx->cell, &ctx->overflow_list, &ctx->wide_list, &ctx->slow_ws, &ctx->flags, &ctx->scan_tmp,
                          &ctx->out_contacts, &ctx->out_cells, &ctx->in_cnt, &ctx->incoming})
    v.push_back(b->ptr);
  v.push_back(ctx->pinned.ptr);
  return v;
}

int run_batched(vltr_context* ctx, const uint64_t* offsets, size_t n_structs, double probe, int mode, size_t* nc,
                size_t* ncell);

// Batched run on device balls through the context's HIP graph: a call whose
// layout matches the capture replays it (one launch, no host synchronisation
// inside); a call that repeats the previous one's layout but has no graph yet
// runs directly, is captured with a pair capacity 25% above the pairs seen,
// and the new graph runs; any other call runs directly.
int run_batched_graph(vltr_context* ctx, const vltr_ball* d_balls, const uint64_t* offsets, size_t n_structs,
                      double probe, int mode, size_t* nc, size_t* ncell) {
  const size_t n = static_cast<size_t>(offsets[n_structs]);
  const size_t bytes = n * sizeof(vltr_ball);
  hipStream_t s = ctx->stream;
  const bool same_call = ctx->graph_src == d_balls && ctx->graph_probe == probe && ctx->graph_mode == mode &&
                         ctx->graph_variant == ctx->clip_variant && ctx->graph_offsets.size() == n_structs + 1 &&
                         std::equal(offsets, offsets + n_structs + 1, ctx->graph_offsets.begin());
  const bool match = ctx->graph != nullptr && same_call && ctx->graph_nb_cap == ctx->nb_cap &&
                     ctx->graph_bufs == graph_buffers(ctx);
  if (!match) {
    VLTR_HIP(ctx->balls.ensure(bytes));
    VLTR_HIP(hipMemcpyAsync(ctx->balls.ptr, d_balls, bytes, hipMemcpyDeviceToDevice, s));
    int rc = run_batched(ctx, offsets, n_structs, probe, mode, nc, ncell);
    if (rc != VLTR_OK) return rc;
    if (!same_call) {
      // a call not seen just before: run directly and capture only when the
      // next call repeats it
      ctx->graph_offsets.assign(offsets, offsets + n_structs + 1);
      ctx->graph_src = d_balls;
      ctx->graph_probe = probe;
      ctx->graph_mode = mode;
      ctx->graph_variant = ctx->clip_variant;
      return VLTR_OK;
    }
    GraphCapture gc;
    const int64_t P = static_cast<int64_t>(ctx->counts[2]);
    gc.cap = P + P / 4 + 4096;
    gc.src = d_balls;
    gc.src_bytes = bytes;
    size_t unused_c = 0, unused_l = 0;
    rc = run_pipeline(ctx, static_cast<int64_t>(n), probe, nullptr, nullptr, 0, mode, &unused_c, &unused_l, offsets,
                      static_cast<int64_t>(n_structs), nullptr, &gc);
    if (rc != VLTR_OK) return rc;
    ctx->graph = gc.exec;
    ctx->graph_nb_cap = ctx->nb_cap;
    ctx->graph_cap = gc.cap;
    ctx->graph_buckets = static_cast<int64_t>(ctx->counts[7]);
    ctx->graph_bufs = graph_buffers(ctx);
  }
  VLTR_HIP(hipEventRecord(ctx->ev[0], s));
  VLTR_HIP(hipGraphLaunch(ctx->graph, s));
  VLTR_HIP(hipEventRecord(ctx->ev[kPhases - 1], s));
  VLTR_HIP(hipStreamSynchronize(s));
  const uint32_t* hw = static_cast<const uint32_t*>(ctx->pinned.ptr);
  if (static_cast<int64_t>(hw[kReportPairs]) > ctx->graph_cap || hw[kReportPairsBeyond] != 0u ||
      (hw[kReportStatus] & (kStatNbOverflow | kStatNbTooLong | kStatPairOverflow))) {
    // more pairs or longer neighbor lists than the capture holds: a direct
    // run (which drops the graph; the next call captures again)
    VLTR_HIP(hipMemcpyAsync(ctx->balls.ptr, d_balls, bytes, hipMemcpyDeviceToDevice, s));
    return run_batched(ctx, offsets, n_structs, probe, mode, nc, ncell);
  }
  Pipeline p{};
  set_balls(p, static_cast<int64_t>(n));
  p.n_pop = p.n;
  p.mode = mode;
  p.nb_cap = ctx->graph_nb_cap;
  p.n_buckets = ctx->graph_buckets;
  finish_run(ctx, p, hw[kReportPairs], n_structs, nc, ncell);
  float ms = 0.0f;
  for (int k = 0; k < kPhases - 1; ++k) ctx->timings[k] = 0.0;
  if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[kPhases - 1]) == hipSuccess) ctx->timings[kPhases - 1] = ms;
  ++ctx->graph_replays;
  return VLTR_OK;
}

int run_batched(vltr_context* ctx, const uint64_t* offsets, size_t n_structs, double probe, int mode, size_t* nc,
                size_t* ncell) {
  const int64_t n = static_cast<int64_t>(offsets[n_structs]);
  *nc = *ncell = 0;
  if (n == 0) {
    ctx->last_structs = n_structs;
    ctx->last_n = 0;
    ctx->last_contacts = ctx->last_cells = 0;
    return VLTR_OK;
  }
  return run_pipeline(ctx, n, probe, nullptr, nullptr, 0, mode, nc, ncell, offsets, static_cast<int64_t>(n_structs));
}

}  // namespace

namespace vltr {
int host_fail(int code, const std::string& msg) { return fail(code, msg); }
int host_tessellate(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe, const vltr_periodic_box* pbox,
                    const int32_t* groups, size_t n_groups, int with_vertices, vltr_result** out,
                    std::vector<ContactNet>* nets) {
  return host_entry(ctx, balls, n, probe, pbox, groups, n_groups, with_vertices, 0, out, nets);
}
}  // namespace vltr

extern "C" {

int vltr_compute_tessellation_batched(vltr_context* ctx, const vltr_ball* balls, const uint64_t* offsets,
                                      size_t n_structs, double probe, int contacts_only, vltr_batch_result** out) {
  if (!ctx || !out) return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  const int chk = check_offsets(offsets, n_structs);
  if (chk != VLTR_OK) return chk;
  const size_t n = static_cast<size_t>(offsets[n_structs]);
  if (n > 0 && !balls) return fail(VLTR_ERR_INVALID_ARGUMENT, "null balls");
  auto r = std::make_unique<vltr_batch_result>();
  r->n_structs = n_structs;
  r->ball_off.assign(offsets, offsets + n_structs + 1);
  r->contact_off.assign(n_structs + 1, 0);
  r->cell_off.assign(n_structs + 1, 0);
  const int mode = contacts_only ? 1 : 0;
  if (n > 0) {
    VLTR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    VLTR_HIP(ctx->balls.ensure(n * sizeof(vltr_ball)));
    VLTR_HIP(hipMemcpyAsync(ctx->balls.ptr, balls, n * sizeof(vltr_ball), hipMemcpyHostToDevice, s));
    size_t nc = 0, ncell = 0;
    const int rc = run_batched(ctx, offsets, n_structs, probe, mode, &nc, &ncell);
    if (rc != VLTR_OK) return rc;
    r->contacts.resize(nc);
    r->cells.resize(ncell);
    std::vector<uint32_t> so(2 * (n_structs + 1), 0u);
    if (nc) VLTR_HIP(hipMemcpyAsync(r->contacts.data(), ctx->out_contacts.ptr, nc * sizeof(vltr_contact),
                                    hipMemcpyDeviceToHost, s));
    if (ncell)
      VLTR_HIP(hipMemcpyAsync(r->cells.data(), ctx->out_cells.ptr, ncell * sizeof(vltr_cell), hipMemcpyDeviceToHost, s));
    VLTR_HIP(hipMemcpyAsync(so.data(), ctx->seg_out.ptr, so.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    VLTR_HIP(hipStreamSynchronize(s));
    for (size_t k = 0; k <= n_structs; ++k) {
      r->contact_off[k] = so[k];
      r->cell_off[k] = mode == 0 ? so[n_structs + 1 + k] : 0;
    }
  }
  *out = r.release();
  return VLTR_OK;
}

int vltr_compute_tessellation_batched_device(vltr_context* ctx, const vltr_ball* d_balls, const uint64_t* offsets,
                                             size_t n_structs, double probe, int contacts_only) {
  if (!ctx) return fail(VLTR_ERR_INVALID_ARGUMENT, "null context");
  const int chk = check_offsets(offsets, n_structs);
  if (chk != VLTR_OK) return chk;
  const size_t n = static_cast<size_t>(offsets[n_structs]);
  if (n > 0 && !d_balls) return fail(VLTR_ERR_INVALID_ARGUMENT, "null balls");
  ctx->last_contacts = ctx->last_cells = 0;
  if (n == 0) {
    ctx->last_structs = n_structs;
    ctx->last_n = 0;
    return VLTR_OK;
  }
  VLTR_HIP(hipSetDevice(ctx->device));
  size_t nc = 0, ncell = 0;
  int rc = VLTR_OK;
  if (ctx->use_graphs) {
    rc = run_batched_graph(ctx, d_balls, offsets, n_structs, probe, contacts_only ? 1 : 0, &nc, &ncell);
  } else {
    VLTR_HIP(ctx->balls.ensure(n * sizeof(vltr_ball)));
    VLTR_HIP(hipMemcpyAsync(ctx->balls.ptr, d_balls, n * sizeof(vltr_ball), hipMemcpyDeviceToDevice, ctx->stream));
    rc = run_batched(ctx, offsets, n_structs, probe, contacts_only ? 1 : 0, &nc, &ncell);
  }
  if (rc != VLTR_OK) return rc;
  ctx->last_contacts = nc;
  ctx->last_cells = ncell;
  return VLTR_OK;
}

int vltr_context_set_graphs(vltr_context* ctx, int enable) {
  if (!ctx) return fail(VLTR_ERR_INVALID_ARGUMENT, "null context");
  ctx->use_graphs = enable != 0;
  if (!ctx->use_graphs) {
    VLTR_HIP(hipSetDevice(ctx->device));
    drop_graph(ctx);
  }
  return VLTR_OK;
}
uint64_t vltr_context_graph_replays(const vltr_context* ctx) { return ctx ? ctx->graph_replays : 0; }
uint64_t vltr_context_speculative_runs(const vltr_context* ctx, uint64_t* misses) {
  if (misses) *misses = ctx ? ctx->spec_misses : 0;
  return ctx ? ctx->spec_runs : 0;
}

int vltr_clip_variant_count(void) { return vltr::clip_variants(); }
const char* vltr_clip_variant_name(int variant) { return vltr::clip_variant_name(variant); }
int vltr_context_set_clip_variant(vltr_context* ctx, int variant) {
  if (!ctx) return fail(VLTR_ERR_INVALID_ARGUMENT, "null context");
  if (variant < 0 || variant >= vltr::clip_variants())
    return fail(VLTR_ERR_INVALID_ARGUMENT, "clip variant out of range");
  ctx->clip_variant = variant;
  return VLTR_OK;
}
int vltr_context_clip_variant(const vltr_context* ctx) { return ctx ? ctx->clip_variant : -1; }

int vltr_context_autotune(vltr_context* ctx, const vltr_ball* d_balls, const uint64_t* offsets, size_t n_structs,
                          double probe, int contacts_only, int repeats, double* ms, int max_entries, int* chosen) {
  if (!ctx || !chosen) return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  if (repeats < 1) repeats = 1;
  const int nv = vltr::clip_variants();
  const int before = ctx->clip_variant;
  // direct runs only: each shape would otherwise be captured in a graph of its own
  const bool graphs = ctx->use_graphs;
  ctx->use_graphs = false;
  // an untimed round over every shape, then `repeats` timed rounds, each
  // over every shape in turn, so that no shape is timed while the device is
  // still settling or always right after the same neighbor
  std::vector<std::vector<double>> t(static_cast<size_t>(nv));
  for (int r = 0; r <= repeats; ++r) {
    for (int v = 0; v < nv; ++v) {
      ctx->clip_variant = v;
      const int rc = vltr_compute_tessellation_batched_device(ctx, d_balls, offsets, n_structs, probe, contacts_only);
      if (rc != VLTR_OK) {
        ctx->clip_variant = before;
        ctx->use_graphs = graphs;
        return rc;
      }
      if (r > 0) t[v].push_back(ctx->timings[kPhases - 1]);
    }
  }
  double best = 0.0;
  int best_v = before;
  for (int v = 0; v < nv; ++v) {
    std::vector<double>& x = t[v];
    std::sort(x.begin(), x.end());
    const size_t k = x.size();
    const double med = k % 2 ? x[k / 2] : 0.5 * (x[k / 2 - 1] + x[k / 2]);
    if (ms && v < max_entries) ms[v] = med;
    if (v == 0 || med < best) {
      best = med;
      best_v = v;
    }
  }
  ctx->clip_variant = best_v;
  ctx->use_graphs = graphs;
  *chosen = best_v;
  return VLTR_OK;
}

int vltr_context_batch_offsets(const vltr_context* ctx, size_t n_structs, uint64_t* contact_off, uint64_t* cell_off) {
  if (!ctx || ctx->last_structs != n_structs || n_structs == 0)
    return fail(VLTR_ERR_INVALID_ARGUMENT, "no batched run with that many structures on this context");
  std::vector<uint32_t> so(2 * (n_structs + 1), 0u);
  if (ctx->last_n > 0) {
    VLTR_HIP(hipSetDevice(ctx->device));
    VLTR_HIP(hipMemcpy(so.data(), ctx->seg_out.ptr, so.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  }
  for (size_t k = 0; k <= n_structs; ++k) {
    if (contact_off) contact_off[k] = so[k];
    if (cell_off) cell_off[k] = ctx->last_cells > 0 ? so[n_structs + 1 + k] : 0;
  }
  return VLTR_OK;
}

// solvent_spheres.rs:131-201 (validation messages as its Display impl).
int vltr_compute_solvent_spheres(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe,
                                 double volume_probe, int subdivision_depth, const vltr_periodic_box* periodic_box,
                                 vltr_solvent_result** out) {
  if (!ctx || !out || (n > 0 && !balls)) return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (!std::isfinite(probe) || probe < 0.0)
    return fail(VLTR_ERR_INVALID_ARGUMENT,
                "invalid probe radius: " + rust_display(probe) + " (must be non-negative and finite)");
  if (!std::isfinite(volume_probe) || volume_probe < 0.0)
    return fail(VLTR_ERR_INVALID_ARGUMENT,
                "invalid volume probe: " + rust_display(volume_probe) + " (must be non-negative and finite)");
  if (subdivision_depth < 0) return fail(VLTR_ERR_INVALID_ARGUMENT, "subdivision depth must be 0..4");
  for (size_t i = 0; i < n; ++i) {
    const vltr_ball& b = balls[i];
    if (!std::isfinite(b.x) || !std::isfinite(b.y) || !std::isfinite(b.z))
      return fail(VLTR_ERR_INVALID_ARGUMENT, "invalid ball at index " + std::to_string(i) + ": coordinates must be finite");
    if (!std::isfinite(b.r) || b.r <= 0.0)
      return fail(VLTR_ERR_INVALID_ARGUMENT,
                  "invalid ball at index " + std::to_string(i) + ": radius must be positive and finite");
  }
  auto r = std::make_unique<vltr_solvent_result>();
  const int rc = solvent_spheres(ctx, balls, n, probe, volume_probe, std::min(subdivision_depth, 4), periodic_box,
                                 &r->spheres);
  if (rc != VLTR_OK) return rc;
  *out = r.release();
  return VLTR_OK;
}
size_t vltr_solvent_count(const vltr_solvent_result* r) { return r ? r->spheres.size() : 0; }
const vltr_solvent_sphere* vltr_solvent_spheres(const vltr_solvent_result* r) {
  return r && !r->spheres.empty() ? r->spheres.data() : nullptr;
}
void vltr_solvent_free(vltr_solvent_result* r) { delete r; }

// Energy and its gradient (forward + adjoint) on device-resident balls.
int vltr_energy_gradient_device(vltr_context* ctx, const vltr_ball* d_balls, size_t n, double probe,
                                const vltr_periodic_box* periodic_box, const int32_t* d_types, const double* w16,
                                double sas_weight, double volume_weight, double* d_grad, double* energy) {
  if (!ctx || !w16 || !energy || (n > 0 && (!d_balls || !d_types || !d_grad)))
    return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  *energy = 0.0;
  ctx->last_contacts = ctx->last_cells = 0;
  if (n == 0) return VLTR_OK;
  VLTR_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  VLTR_HIP(ctx->balls.ensure(n * sizeof(vltr_ball)));
  VLTR_HIP(hipMemcpyAsync(ctx->balls.ptr, d_balls, n * sizeof(vltr_ball), hipMemcpyDeviceToDevice, s));
  VLTR_HIP(ctx->seed.ensure(n * sizeof(double)));
  VLTR_HIP(ctx->energy_part.ensure(adjoint_blocks() * sizeof(double)));
  RunExtras ex;
  ex.topo_buf = &ctx->topo;
  ex.topo_slot_buf = &ctx->topo_slot;
  ex.topo_cap = &ctx->topo_cap;
  ex.seed = ctx->seed.as<double>();
  ex.sas_weight = sas_weight;
  ex.volume_weight = volume_weight;
  double* part = ctx->energy_part.as<double>();
  std::vector<double> w(w16, w16 + 16);
  ex.post_cells = [=](const Pipeline& p, hipStream_t st) -> hipError_t {
    unsigned long long* keys = nullptr;
    uint32_t* first = nullptr;
    if (p.periodic) {
      const uint64_t slots = dedup_slots(p.n_pairs);
      hipError_t e = ctx->dedup_keys.ensure(slots * sizeof(unsigned long long));
      if (e == hipSuccess) e = ctx->dedup_first.ensure(slots * sizeof(uint32_t));
      if (e != hipSuccess) return e;
      keys = ctx->dedup_keys.as<unsigned long long>();
      first = ctx->dedup_first.as<uint32_t>();
    }
    hipError_t e = ctx->adj_list.ensure(std::max<int64_t>(p.n_pairs, 1) * sizeof(int32_t));
    if (e != hipSuccess) return e;
    // the contact flags and their scan scratch are free until the contacts are compacted
    const AdjointWork work{ctx->flags.as<uint32_t>(), ctx->scan_tmp.as<uint32_t>(), ctx->adj_list.as<int32_t>()};
    return launch_adjoint(p, d_types, w.data(), d_grad, part, keys, first, work, st);
  };
  size_t nc = 0, ncell = 0;
  const int rc =
      run_pipeline(ctx, static_cast<int64_t>(n), probe, periodic_box, nullptr, 0, 0, &nc, &ncell, nullptr, 0, &ex);
  if (rc != VLTR_OK) return rc;
  std::vector<double> parts(static_cast<size_t>(adjoint_blocks()));
  VLTR_HIP(hipMemcpy(parts.data(), part, parts.size() * sizeof(double), hipMemcpyDeviceToHost));
  double e = 0.0;
  for (double v : parts) e += v;
  *energy = e;
  ctx->last_contacts = nc;
  ctx->last_cells = ncell;
  return VLTR_OK;
}

int vltr_energy_gradient(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe,
                         const vltr_periodic_box* periodic_box, const int32_t* types, const double* w16,
                         double sas_weight, double volume_weight, double* energy, double* grad) {
  if (!ctx || !w16 || !energy || (n > 0 && (!balls || !types || !grad)))
    return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  *energy = 0.0;
  if (n == 0) return VLTR_OK;
  for (size_t i = 0; i < n; ++i)
    if (types[i] < 0 || types[i] > 3) return fail(VLTR_ERR_INVALID_ARGUMENT, "types must be in 0..3");
  VLTR_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  DevBuf db;
  VLTR_HIP(db.ensure(n * sizeof(vltr_ball)));
  VLTR_HIP(ctx->types.ensure(n * sizeof(int32_t)));
  VLTR_HIP(ctx->grad.ensure(3 * n * sizeof(double)));
  VLTR_HIP(hipMemcpyAsync(db.ptr, balls, n * sizeof(vltr_ball), hipMemcpyHostToDevice, s));
  VLTR_HIP(hipMemcpyAsync(ctx->types.ptr, types, n * sizeof(int32_t), hipMemcpyHostToDevice, s));
  int rc = vltr_energy_gradient_device(ctx, db.as<vltr_ball>(), n, probe, periodic_box, ctx->types.as<int32_t>(), w16,
                                       sas_weight, volume_weight, ctx->grad.as<double>(), energy);
  if (rc == VLTR_OK) {
    if (hipMemcpy(grad, ctx->grad.ptr, 3 * n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(VLTR_ERR_DEVICE, "gradient download");
  }
  db.release();
  return rc;
}

size_t vltr_batch_num_structs(const vltr_batch_result* r) { return r ? r->n_structs : 0; }
const vltr_contact* vltr_batch_contacts(const vltr_batch_result* r, size_t k, size_t* count) {
  if (!r || k >= r->n_structs) {
    if (count) *count = 0;
    return nullptr;
  }
  if (count) *count = static_cast<size_t>(r->contact_off[k + 1] - r->contact_off[k]);
  return r->contacts.data() + r->contact_off[k];
}
const vltr_cell* vltr_batch_cells(const vltr_batch_result* r, size_t k, size_t* count) {
  if (!r || k >= r->n_structs) {
    if (count) *count = 0;
    return nullptr;
  }
  if (count) *count = static_cast<size_t>(r->cell_off[k + 1] - r->cell_off[k]);
  return r->cells.data() + r->cell_off[k];
}
void vltr_batch_free(vltr_batch_result* r) { delete r; }

const char* vltr_last_error(void) { return g_last_error.c_str(); }

int vltr_context_create(int device, vltr_context** out) {
  if (!out) return fail(VLTR_ERR_INVALID_ARGUMENT, "null out");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return fail(VLTR_ERR_DEVICE, "no HIP device available (this library runs on the GPU only)");
  if (device < 0 || device >= count) return fail(VLTR_ERR_INVALID_ARGUMENT, "device index out of range");
  VLTR_HIP(hipSetDevice(device));
  auto* ctx = new (std::nothrow) vltr_context();
  if (!ctx) return fail(VLTR_ERR_OUT_OF_MEMORY, "context allocation");
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return fail(VLTR_ERR_DEVICE, "stream creation failed");
  }
  for (auto& e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      vltr_context_destroy(ctx);
      return fail(VLTR_ERR_DEVICE, "event creation failed");
    }
  init_hexagon_constants();
  if (upload_hexagon_constants(ctx->stream) != hipSuccess) {
    vltr_context_destroy(ctx);
    return fail(VLTR_ERR_DEVICE, "constant upload failed");
  }
  // the run's report block on the host, and the allocation-free entry
  // point's workspace on this device (its address taken once per device)
  if (ctx->pinned.ensure(kPinnedBytes) != hipSuccess) {
    vltr_context_destroy(ctx);
    return fail(VLTR_ERR_OUT_OF_MEMORY, "pinned report block");
  }
  const hipError_t es = noalloc_scratch_setup(device);
  if (es != hipSuccess) {
    vltr_context_destroy(ctx);
    return fail(VLTR_ERR_DEVICE, std::string("workspace address: ") + hipGetErrorString(es));
  }
  *out = ctx;
  return VLTR_OK;
}

void vltr_context_destroy(vltr_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  drop_graph(ctx);
  for (DevBuf* b : {&ctx->balls, &ctx->groups, &ctx->sph, &ctx->seg_of, &ctx->segs, &ctx->bkey, &ctx->brank, &ctx->bcnt,
                    &ctx->sorted_id, &ctx->sorted_sph, &ctx->small, &ctx->nb, &ctx->nb_cnt, &ctx->excluded,
                    &ctx->pair_cnt, &ctx->pair_a, &ctx->pair_b, &ctx->rec, &ctx->cell, &ctx->overflow_list, &ctx->wide_list,
                    &ctx->flags, &ctx->scan_tmp, &ctx->out_contacts, &ctx->out_cells, &ctx->slow_ws,
                    &ctx->in_cnt, &ctx->incoming, &ctx->seg_out, &ctx->topo, &ctx->seed, &ctx->energy_part,
                    &ctx->types, &ctx->grad, &ctx->net, &ctx->net_out, &ctx->solv_cnt, &ctx->solv_comb,
                    &ctx->solv_parent, &ctx->solv_sih, &ctx->solv_ioff, &ctx->solv_cur, &ctx->solv_img,
                    &ctx->nb_big, &ctx->dedup_keys, &ctx->dedup_first, &ctx->adj_list, &ctx->net_eoff,
                    &ctx->net_voff, &ctx->net_work, &ctx->pair_contact, &ctx->topo_slot})
    b->release();
  ctx->pinned.release();
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int vltr_context_last_timings(const vltr_context* ctx, double* ms, int max_entries) {
  if (!ctx || !ms) return 0;
  const int k = std::min(max_entries, kPhases);
  for (int i = 0; i < k; ++i) ms[i] = ctx->timings[i];
  return k;
}

int vltr_context_last_counts(const vltr_context* ctx, uint64_t* out, int max_entries) {
  if (!ctx || !out) return 0;
  const int k = std::min(max_entries, 8);
  for (int i = 0; i < k; ++i) out[i] = ctx->counts[i];
  return k;
}

void vltr_periodic_box_from_corners(const double mn[3], const double mx[3], vltr_periodic_box* out) {
  if (!out) return;
  *out = vltr_periodic_box{{mx[0] - mn[0], 0.0, 0.0}, {0.0, mx[1] - mn[1], 0.0}, {0.0, 0.0, mx[2] - mn[2]}};
}

void vltr_periodic_box_from_vectors(const double a[3], const double b[3], const double c[3], vltr_periodic_box* out) {
  if (!out) return;
  *out = vltr_periodic_box{{a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]}};
}

int vltr_compute_tessellation(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe,
                              const vltr_periodic_box* periodic_box, const int32_t* groups, size_t n_groups,
                              int with_cell_vertices, vltr_result** out) {
  return host_entry(ctx, balls, n, probe, periodic_box, groups, n_groups, with_cell_vertices, 0, out);
}

int vltr_compute_contacts_only(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe,
                               const vltr_periodic_box* periodic_box, const int32_t* groups, size_t n_groups,
                               vltr_result** out) {
  return host_entry(ctx, balls, n, probe, periodic_box, groups, n_groups, 0, 1, out);
}

int vltr_compute_tessellation_device(vltr_context* ctx, const vltr_ball* d_balls, size_t n, double probe,
                                     const int32_t* d_groups, size_t n_groups, int contacts_only) {
  if (!ctx || (n > 0 && !d_balls)) return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  ctx->last_contacts = ctx->last_cells = 0;
  if (n == 0) return VLTR_OK;
  VLTR_HIP(hipSetDevice(ctx->device));
  VLTR_HIP(ctx->balls.ensure(n * sizeof(vltr_ball)));
  VLTR_HIP(hipMemcpyAsync(ctx->balls.ptr, d_balls, n * sizeof(vltr_ball), hipMemcpyDeviceToDevice, ctx->stream));
  size_t nc = 0, ncell = 0;
  const int rc = run_pipeline(ctx, static_cast<int64_t>(n), probe, nullptr, d_groups,
                              d_groups ? static_cast<int64_t>(n_groups) : 0, contacts_only ? 1 : 0, &nc, &ncell);
  if (rc != VLTR_OK) return rc;
  ctx->last_contacts = nc;
  ctx->last_cells = ncell;
  return VLTR_OK;
}

}  // extern "C"

namespace {

static_assert(kReportWords * sizeof(uint32_t) <= kPinnedSegOff && kPinnedSegOff + sizeof(GridSeg) <= kPinnedBytes,
              "the pinned block holds the report words and one segment");


// Lanes of the one-lane fallback kernel in the workspace (fewer than the
// context's: the kernel walks its work list with a stride of its lane count).
constexpr int64_t kNoAllocSlowLanes = 256;

// One calls at a time per device: calls share the device's workspace.
std::mutex g_noalloc_lock[64];

// vltr_compute_tessellation_device on the fixed workspace: the same kernels
// and launch order, buffers carved from the workspace, outputs written to the
// caller's buffers.
int run_noalloc(vltr_context* ctx, const vltr_ball* d_balls, int64_t n, double probe, const int32_t* d_groups,
                int64_t n_groups, int mode, vltr_contact* d_contacts, size_t contacts_cap, vltr_cell* d_cells,
                size_t* n_contacts, size_t* n_cells) {
  unsigned char* ws = static_cast<unsigned char*>(noalloc_scratch(ctx->device));
  if (ws == nullptr) return fail(VLTR_ERR_DEVICE, "no workspace address for this device (context not set up)");
  hipStream_t s = ctx->stream;
  uint32_t* host_words = static_cast<uint32_t*>(ctx->pinned.ptr);
  Pipeline p{};
  set_balls(p, n);
  p.n_pop = n;
  p.n_segs = 1;
  p.mode = mode;
  p.probe = probe;
  p.groups = d_groups;
  p.n_groups = n_groups;
  p.clip_variant = ctx->clip_variant;
  p.balls_in = reinterpret_cast<const double4*>(d_balls);
  p.contacts_cap = static_cast<int64_t>(contacts_cap);
  GridSeg* seg_host = reinterpret_cast<GridSeg*>(static_cast<unsigned char*>(ctx->pinned.ptr) + kPinnedSegOff);
  seg_host->ball_off = 0;
  seg_host->tab_off = 0;
  seg_host->kbits = kbits_for(n);
  seg_host->n = static_cast<int32_t>(n);
  p.n_buckets = int64_t(1) << (3 * seg_host->kbits);
  int nb_cap = kNbCapDefault;
  for (;;) {
    Arena A{ws, kNoAllocScratchBytes};
    p.nb_cap = nb_cap;
    unsigned char* small = A.take<unsigned char>(kSmallBytes);
    GridSeg* segs = A.take<GridSeg>(1);
    p.sph = A.take<double4>(n);
    p.seg_of = A.take<int32_t>(n);
    p.bucket_key = A.take<uint32_t>(n);
    p.bucket_rank = A.take<uint32_t>(n);
    p.bucket_cnt = A.take<uint32_t>(p.n_buckets + 1);
    p.sorted_id = A.take<int32_t>(n);
    p.sorted_sph = A.take<double4>(n);
    p.nb = A.take<int32_t>(static_cast<size_t>(n) * nb_cap);
    p.nb_cnt = A.take<int32_t>(n);
    p.nb_big = A.take<int32_t>(n);
    p.excluded = A.take<uint8_t>(n);
    p.pair_cnt = A.take<uint32_t>(n + 1);
    p.cell = A.take<CellRec>(n);
    p.in_cnt = A.take<uint32_t>(n);
    p.incoming = mode == 0 ? A.take<int32_t>(static_cast<size_t>(n) * nb_cap) : nullptr;
    uint32_t* scan_front = A.take<uint32_t>(scan_tmp_words(std::max<int64_t>(p.n_buckets, n) + 64) + 64);
    if (!A.fits())
      return fail(VLTR_ERR_INVALID_ARGUMENT, "the run needs more than the " + std::to_string(kNoAllocScratchBytes >> 20) +
                                                 " MiB workspace of the allocation-free entry point (" +
                                                 std::to_string(n) + " balls): use vltr_compute_tessellation_device");
    p.segs = segs;
    p.box_bits = reinterpret_cast<unsigned long long*>(small);
    p.status = reinterpret_cast<uint32_t*>(small + 8);
    p.overflow_cnt = p.status + kReportOverflow;
    p.wide_cnt = p.status + kReportWide;
    p.nb_big_cnt = p.status + kReportNbBig;
    p.bbox = reinterpret_cast<unsigned long long*>(small + 256);
    p.clip_queue = reinterpret_cast<uint32_t*>(small + kSmallQueueOff);
    p.topo_counts = reinterpret_cast<uint32_t*>(small + kSmallTopoOff);
    VLTR_HIP(hipMemcpyAsync(segs, seg_host, sizeof(GridSeg), hipMemcpyHostToDevice, s));
    const RunBufs front{scan_front, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int rc = enqueue_front(ctx, p, front, false, s);
    if (rc != VLTR_OK) return rc;
    VLTR_HIP(hipMemcpyAsync(host_words, p.status, kReportWords * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    VLTR_HIP(hipStreamSynchronize(s));
    if (host_words[kReportStatus] & kStatNbTooLong)
      return fail(VLTR_ERR_INTERNAL, "neighbor list longer than 4096 entries");
    if (host_words[kReportStatus] & kStatNbOverflow) {
      if (nb_cap >= kNbMax) return fail(VLTR_ERR_INTERNAL, "neighbor list longer than 4096 entries");
      nb_cap *= 2;
      continue;
    }
    int64_t P = 0;
    rc = checked_pairs(host_words, &P);
    if (rc != VLTR_OK) return rc;
    p.pair_cap = std::max<int64_t>(P, 1);
    p.n_pairs = P;
    p.pair_a = A.take<int32_t>(p.pair_cap);
    p.pair_b = A.take<int32_t>(p.pair_cap);
    p.rec = A.take<PairRec>(p.pair_cap);
    p.pair_contact = A.take<uint8_t>((p.pair_cap + 3) & ~int64_t(3));
    p.overflow_list = A.take<int32_t>(p.pair_cap);
    p.wide_list = A.take<int32_t>(p.pair_cap);
    p.slow_ws_pairs = kNoAllocSlowLanes;
    const int64_t cc = 6 + nb_cap, cr = 2 * cc;
    p.slow_ws = A.take<double>(static_cast<size_t>(kNoAllocSlowLanes) * (5 * cc + 6 * cr));
    uint32_t* flags = A.take<uint32_t>(std::max<int64_t>(P, n) + 2);
    uint32_t* scan_back = A.take<uint32_t>(scan_tmp_words(std::max<int64_t>(P, n) + 64) + 64);
    if (!A.fits())
      return fail(VLTR_ERR_INVALID_ARGUMENT, "the run needs more than the " + std::to_string(kNoAllocScratchBytes >> 20) +
                                                 " MiB workspace of the allocation-free entry point (" +
                                                 std::to_string(n) + " balls, " + std::to_string(P) +
                                                 " candidate pairs): use vltr_compute_tessellation_device");
    const RunBufs back{scan_back, flags, d_contacts, d_cells, nullptr, nullptr, nullptr};
    rc = enqueue_back(ctx, p, back, nullptr, false, false, host_words, s);
    if (rc != VLTR_OK) return rc;
    VLTR_HIP(hipStreamSynchronize(s));
    if (host_words[kReportStatus] & kStatNbOverflow) {
      // an incoming-pair list outgrew nb_cap: the whole run with longer lists
      if (nb_cap >= kNbMax) return fail(VLTR_ERR_INTERNAL, "neighbor list longer than 4096 entries");
      nb_cap *= 2;
      continue;
    }
    if (host_words[kReportStatus] & kStatPairOverflow)
      return fail(VLTR_ERR_INTERNAL, "pair lists outgrew the buffers sized from the pair total");
    *n_contacts = host_words[kReportContacts];
    *n_cells = mode == 0 ? host_words[kReportCells] : 0;
    ctx->counts[0] = ctx->counts[1] = static_cast<uint64_t>(n);
    ctx->counts[2] = static_cast<uint64_t>(P);
    ctx->counts[3] = *n_contacts;
    ctx->counts[4] = *n_cells;
    ctx->counts[5] = static_cast<uint64_t>(host_words[kReportOverflow]) + host_words[kReportWide];
    ctx->counts[6] = static_cast<uint64_t>(nb_cap);
    ctx->counts[7] = static_cast<uint64_t>(p.n_buckets);
    if (*n_contacts > contacts_cap)
      return fail(VLTR_ERR_INVALID_ARGUMENT, "the run has " + std::to_string(*n_contacts) +
                                                 " contacts, more than contacts_cap (" + std::to_string(contacts_cap) +
                                                 "); only the first contacts_cap were written");
    return VLTR_OK;
  }
}

}  // namespace

extern "C" {

int vltr_compute_tessellation_device_noalloc(vltr_context* ctx, const vltr_ball* d_balls, size_t n, double probe,
                                             const int32_t* d_groups, size_t n_groups, int contacts_only,
                                             vltr_contact* d_contacts, size_t contacts_cap, vltr_cell* d_cells,
                                             size_t cells_cap, size_t* n_contacts, size_t* n_cells) {
  if (!ctx || !n_contacts || !n_cells || (n > 0 && !d_balls) || (n > 0 && !d_contacts) ||
      (n > 0 && !contacts_only && !d_cells))
    return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  *n_contacts = *n_cells = 0;
  if (n == 0) return VLTR_OK;
  if (contacts_cap == 0) return fail(VLTR_ERR_INVALID_ARGUMENT, "contacts_cap must be positive");
  if (n > kNoAllocScratchBytes / 32)
    return fail(VLTR_ERR_INVALID_ARGUMENT, "too many balls for the " + std::to_string(kNoAllocScratchBytes >> 20) +
                                               " MiB workspace of the allocation-free entry point");
  if (!contacts_only && cells_cap < n)
    return fail(VLTR_ERR_INVALID_ARGUMENT, "cells_cap must be at least the ball count (one cell per ball at most)");
  if (ctx->device < 0 || ctx->device >= 64) return fail(VLTR_ERR_INVALID_ARGUMENT, "device index out of range");
  std::lock_guard<std::mutex> lock(g_noalloc_lock[ctx->device]);
  VLTR_HIP(hipSetDevice(ctx->device));
  return run_noalloc(ctx, d_balls, static_cast<int64_t>(n), probe, d_groups, d_groups ? static_cast<int64_t>(n_groups) : 0,
                     contacts_only ? 1 : 0, d_contacts, contacts_cap, d_cells, n_contacts, n_cells);
}

int vltr_context_device_outputs(const vltr_context* ctx, const vltr_contact** d_contacts, size_t* n_contacts,
                                const vltr_cell** d_cells, size_t* n_cells) {
  if (!ctx) return fail(VLTR_ERR_INVALID_ARGUMENT, "null context");
  if (d_contacts) *d_contacts = ctx->out_contacts.as<const vltr_contact>();
  if (n_contacts) *n_contacts = ctx->last_contacts;
  if (d_cells) *d_cells = ctx->out_cells.as<const vltr_cell>();
  if (n_cells) *n_cells = ctx->last_cells;
  return VLTR_OK;
}

int64_t vltr_context_neighbor_lists(const vltr_context* ctx, size_t num_balls, int32_t* counts, int32_t* ids,
                                    size_t max_ids) {
  if (!ctx || !counts || ctx->last_n != num_balls || ctx->nb.ptr == nullptr) {
    fail(VLTR_ERR_INVALID_ARGUMENT, "no pipeline run with that many balls on this context");
    return -1;
  }
  if (num_balls == 0) return 0;
  if (hipSetDevice(ctx->device) != hipSuccess) return -1;
  std::vector<int32_t> all(num_balls * static_cast<size_t>(ctx->last_nb_cap));
  if (hipMemcpy(counts, ctx->nb_cnt.ptr, num_balls * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(all.data(), ctx->nb.ptr, all.size() * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess) {
    fail(VLTR_ERR_DEVICE, "neighbor list download");
    return -1;
  }
  int64_t total = 0;
  for (size_t i = 0; i < num_balls; ++i) {
    for (int k = 0; k < counts[i]; ++k) {
      if (ids && static_cast<size_t>(total) < max_ids) ids[total] = all[i * ctx->last_nb_cap + k];
      ++total;
    }
  }
  return total;
}

size_t vltr_result_num_balls(const vltr_result* r) { return r ? r->num_balls : 0; }
size_t vltr_result_num_contacts(const vltr_result* r) { return r ? r->contacts.size() : 0; }
const vltr_contact* vltr_result_contacts(const vltr_result* r) { return r ? r->contacts.data() : nullptr; }
size_t vltr_result_num_cells(const vltr_result* r) { return r ? r->cells.size() : 0; }
const vltr_cell* vltr_result_cells(const vltr_result* r) { return r ? r->cells.data() : nullptr; }
int vltr_result_has_cell_vertices(const vltr_result* r) { return r && r->has_network ? 1 : 0; }
size_t vltr_result_num_cell_vertices(const vltr_result* r) { return r ? r->vertices.size() : 0; }
const vltr_cell_vertex* vltr_result_cell_vertices(const vltr_result* r) { return r ? r->vertices.data() : nullptr; }
size_t vltr_result_num_cell_edges(const vltr_result* r) { return r ? r->edges.size() : 0; }
const vltr_cell_edge* vltr_result_cell_edges(const vltr_result* r) { return r ? r->edges.data() : nullptr; }

// types.rs:231-252: plain sums in vector order.
double vltr_result_total_sas_area(const vltr_result* r) {
  double s = 0.0;
  if (r)
    for (const auto& c : r->cells) s += c.sas_area;
  return s;
}
double vltr_result_total_volume(const vltr_result* r) {
  double s = 0.0;
  if (r)
    for (const auto& c : r->cells) s += c.volume;
  return s;
}
double vltr_result_total_contact_area(const vltr_result* r) {
  double s = 0.0;
  if (r)
    for (const auto& c : r->contacts) s += c.area;
  return s;
}
// types.rs:206-228
void vltr_result_sas_areas(const vltr_result* r, double* out) {
  if (!r || !out) return;
  for (size_t i = 0; i < r->num_balls; ++i) out[i] = std::nan("");
  for (const auto& c : r->cells)
    if (c.index < r->num_balls) out[c.index] = c.sas_area;
}
void vltr_result_volumes(const vltr_result* r, double* out) {
  if (!r || !out) return;
  for (size_t i = 0; i < r->num_balls; ++i) out[i] = std::nan("");
  for (const auto& c : r->cells)
    if (c.index < r->num_balls) out[c.index] = c.volume;
}
void vltr_result_free(vltr_result* r) { delete r; }

}  // extern "C"

// ------------------------------------------------------------ updateable ----
// UpdateableTessellation (updateable.rs:154-742, spheres_container.rs:28-405)
// on the device. Every state transition (changed and affected sets, the
// full-reinit thresholds, exclusion statuses, backup and restore) follows the
// reference; the tessellation itself is recomputed for all pairs on each
// step. That is exact: a pair whose ends are both unaffected has the same
// inputs as before, so its record is bit-identical to the summary the
// reference keeps, while pairs with two affected ends are re-clipped in both.
// Cells follow the reference's incremental rules (k_updateable_cells): only
// affected balls take new records.
struct vltr_updateable {
  vltr_context* ctx = nullptr;
  bool backup_enabled = false;
  struct State {
    double probe = 0.0;
    bool periodic = false;
    vltr_periodic_box pbox{};
    std::vector<vltr_ball> balls;  // the container's balls
    std::vector<uint64_t> changed, affected;
    bool was_full_reinit = true;
    bool empty = true;  // updateable.rs result_empty(): no balls or no cells
    size_t n_contacts = 0, n_cells = 0;
  };
  State st, bk;
  bool has_backup = false;
  // cells and exclusion statuses of the current state and of the backup
  DevBuf cells_state, cells_backup, excl, excl_backup, affected_mask, ids, rows, rows_cnt, moved_vals;
  HostPinned stage;  // staging of moved balls and affected marks
  std::vector<uint64_t> aff_scratch;  // the affected list under construction
  // energy and gradient computed in the pipeline pass of every state change
  bool energy_on = false, grad_valid = false;
  double energy = 0.0, sas_weight = 1.0, volume_weight = 1.0;
  std::vector<double> w16;
  size_t n_types = 0;
  DevBuf types, grad, part, topo, topo_slot, seed, dedup_keys, dedup_first;
  int64_t topo_cap = 0;
  // pair lists, records and topologies (store, slots, records in the store)
  // of the last run: the next step's pairs that are not clipped again (one
  // end not affected) take them
  DevBuf prev_pair_cnt, prev_pair_b, prev_rec, prev_topo, prev_topo_slot;
  bool prev_ok = false, prev_topo_ok = false;
  size_t prev_n = 0;
  int64_t prev_pairs = 0, prev_topos = 0;
  // the same for the backup state: a restore keeps every pair's record
  DevBuf bk_pair_cnt, bk_pair_b, bk_rec, bk_topo, bk_topo_slot;
  bool bk_ok = false, bk_topo_ok = false;
  size_t bk_n = 0;
  int64_t bk_pairs = 0, bk_topos = 0;
};

namespace {

using Upd = vltr_updateable;

// sphere_equals_sphere (geometry.rs:69-72) on Sphere::from_ball (types.rs:58-63)
bool same_sphere(const vltr_ball& a, const vltr_ball& b, double probe) {
  return equals(Sph{V3{a.x, a.y, a.z}, a.r + probe}, Sph{V3{b.x, b.y, b.z}, b.r + probe});
}

// Canonical colliding neighbors of the balls ids, from the last pipeline run.
int upd_rows(Upd* u, const std::vector<uint64_t>& ids, std::vector<std::vector<uint64_t>>* out) {
  vltr_context* ctx = u->ctx;
  const size_t n = u->st.balls.size();
  out->assign(ids.size(), {});
  if (ids.empty() || n == 0 || ctx->last_n != n) return VLTR_OK;
  const int cap = ctx->last_nb_cap;
  std::vector<int32_t> id32(ids.begin(), ids.end());
  VLTR_HIP(u->ids.ensure(id32.size() * sizeof(int32_t)));
  VLTR_HIP(u->rows.ensure(id32.size() * cap * sizeof(int32_t)));
  VLTR_HIP(u->rows_cnt.ensure(id32.size() * sizeof(int32_t)));
  hipStream_t s = ctx->stream;
  VLTR_HIP(hipMemcpyAsync(u->ids.ptr, id32.data(), id32.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  Pipeline p{};
  set_balls(p, static_cast<int64_t>(n));
  p.nb = ctx->nb.as<int32_t>();
  p.nb_cnt = ctx->nb_cnt.as<int32_t>();
  p.nb_cap = cap;
  VLTR_HIP(launch_gather_rows(p, u->ids.as<int32_t>(), static_cast<int64_t>(id32.size()), u->rows_cnt.as<int32_t>(),
                              u->rows.as<int32_t>(), s));
  std::vector<int32_t> cnt(id32.size()), rows(id32.size() * cap);
  VLTR_HIP(hipMemcpyAsync(cnt.data(), u->rows_cnt.ptr, cnt.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  VLTR_HIP(hipMemcpyAsync(rows.data(), u->rows.ptr, rows.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  VLTR_HIP(hipStreamSynchronize(s));
  for (size_t k = 0; k < ids.size(); ++k)
    for (int j = 0; j < cnt[k]; ++j) (*out)[k].push_back(static_cast<uint64_t>(rows[k * cap + j]) % n);
  return VLTR_OK;
}

// Options of one updateable run.
struct RunOpts {
  bool full = false;         // init / full reinit rules: every ball takes a new cell
  bool refresh = false;      // affected balls get their hidden status back as exclusion status
  bool incremental = false;  // clip only pairs with both balls affected; the others keep the
                             // records of the last run (updateable.rs:476-512), which must be
                             // of the state this run starts from
  const std::vector<uint64_t>* aff = nullptr;  // affected balls, uploaded as the mask
  bool mask_ready = false;   // the affected mask is on the device already
  int64_t n_moved = 0;       // balls rewritten in the device copy of the previous balls
                             // (ids in u->ids, values in u->moved_vals, on the device)
  bool more = false;         // the moved balls' new neighbors join the affected set (mark 2)
  uint8_t* mask_out = nullptr;  // pinned host copy of the final affected mask (n bytes)
  std::function<bool(const uint32_t*)> front_check;  // RunExtras::front_check
};

// The ids i < n with mask[i] != 0, in order. Branch-free: about 40% of the
// balls are marked in a dynamics step, which a data-dependent branch per id
// mispredicts often enough to cost 100 us at 50k balls.
void marked_ids(const uint8_t* mask, size_t n, std::vector<uint64_t>* out) {
  out->resize(n);
  uint64_t* o = out->data();
  size_t c = 0;
  for (size_t i = 0; i < n; ++i) {
    o[c] = i;
    c += mask[i] != 0;
  }
  out->resize(c);
}

// Pinned staging of at least `bytes`, after the stream is idle when it grows.
int upd_stage(Upd* u, size_t bytes) {
  if (u->stage.bytes < bytes) VLTR_HIP(hipStreamSynchronize(u->ctx->stream));
  VLTR_HIP(u->stage.ensure(bytes));
  return VLTR_OK;
}

// One tessellation of the current balls; without full, only affected balls
// take new cells.
int upd_run(Upd* u, const RunOpts& o) {
  vltr_context* ctx = u->ctx;
  auto& st = u->st;
  const size_t n = st.balls.size();
  const bool keep = o.incremental && !o.full && u->prev_ok && u->prev_n == n;
  const bool scatter = o.n_moved > 0 && ctx->last_n == n;
  u->prev_ok = false;
  ctx->last_n = 0;  // until this run completes
  st.n_contacts = st.n_cells = 0;
  if (n == 0) {
    st.empty = true;
    return VLTR_OK;
  }
  hipStream_t s = ctx->stream;
  VLTR_HIP(hipSetDevice(ctx->device));
  if (scatter) {
    // the device copy holds the previous balls: write only the moved ones
    VLTR_HIP(launch_scatter_balls(ctx->balls.as<vltr_ball>(), u->ids.as<int32_t>(), u->moved_vals.as<vltr_ball>(),
                                  o.n_moved, s));
  } else {
    VLTR_HIP(ctx->balls.ensure(n * sizeof(vltr_ball)));
    VLTR_HIP(hipMemcpyAsync(ctx->balls.ptr, st.balls.data(), n * sizeof(vltr_ball), hipMemcpyHostToDevice, s));
  }
  VLTR_HIP(u->cells_state.ensure(n * sizeof(CellRec)));
  VLTR_HIP(u->excl.ensure(n));
  VLTR_HIP(u->affected_mask.ensure(n));
  uint8_t* am = u->affected_mask.as<uint8_t>();
  if (o.full) {
    VLTR_HIP(launch_fill_u8(u->excl.as<uint8_t>(), static_cast<int64_t>(n), 0, s));
  } else {
    if (!o.mask_ready) {
      const int rc = upd_stage(u, n);
      if (rc != VLTR_OK) return rc;
      uint8_t* hm = static_cast<uint8_t*>(u->stage.ptr);
      std::memset(hm, 0, n);
      if (o.aff)
        for (uint64_t id : *o.aff)
          if (id < n) hm[id] = 1;
      VLTR_HIP(hipMemcpyAsync(am, hm, n, hipMemcpyHostToDevice, s));
    }
    if (o.refresh) VLTR_HIP(launch_mask_clear(u->excl.as<uint8_t>(), am, static_cast<int64_t>(n), s));
  }
  RunExtras ex;
  ex.records_all = true;  // the next step copies the records of the pairs it does not clip
  ex.user_excl = o.full ? nullptr : u->excl.as<uint8_t>();
  ex.front_check = o.front_check;
  CellRec* state = u->cells_state.as<CellRec>();
  uint8_t* excl = u->excl.as<uint8_t>();
  const int32_t* moved_ids = u->ids.as<int32_t>();
  const int64_t n_more = o.more ? o.n_moved : 0;
  const bool full = o.full, refresh = o.refresh;
  uint8_t* mask_out = o.mask_out;
  // the energy's adjoint rides on the same pass (cut topologies and cell
  // seeds written by the clip and cell kernels)
  const bool adj = u->energy_on && u->n_types == n;
  u->grad_valid = false;
  const int32_t* d_types = nullptr;
  double *d_grad = nullptr, *d_part = nullptr;
  if (adj) {
    VLTR_HIP(u->seed.ensure(n * sizeof(double)));
    VLTR_HIP(u->grad.ensure(3 * n * sizeof(double)));
    VLTR_HIP(u->part.ensure(adjoint_blocks() * sizeof(double)));
    ex.topo_buf = &u->topo;
    ex.topo_slot_buf = &u->topo_slot;
    ex.topo_cap = &u->topo_cap;
    ex.seed = u->seed.as<double>();
    ex.sas_weight = u->sas_weight;
    ex.volume_weight = u->volume_weight;
    d_types = u->types.as<int32_t>();
    d_grad = u->grad.as<double>();
    d_part = u->part.as<double>();
  }
  const double* w16 = u->w16.data();
  // the moved balls' new neighbors join the affected set before the clip
  if (n_more > 0)
    ex.pre_clip = [=](const Pipeline& p, hipStream_t st2) -> hipError_t {
      return launch_mark_neighbors(p, moved_ids, n_more, am, 2, st2);
    };
  if (keep && (!adj || u->prev_topo_ok)) {
    ex.dirty = am;
    ex.prev_pair_cnt = u->prev_pair_cnt.as<uint32_t>();
    ex.prev_pair_b = u->prev_pair_b.as<int32_t>();
    ex.prev_rec = u->prev_rec.as<PairRec>();
    ex.prev_topo = adj ? u->prev_topo.as<ContactTopo>() : nullptr;
    ex.prev_topo_slot = adj ? u->prev_topo_slot.as<uint32_t>() : nullptr;
  }
  // updateable.rs:691-727: an incremental step recomputes the affected cells
  // from the deduplicated per-ball contact lists (in a periodic box, of each
  // canonical pair through an image only its first record in pair order);
  // init and full reinits sum every ball's own records (updateable.rs:633-689)
  if (!full && st.periodic) {
    ex.dd_keys = &u->dedup_keys;
    ex.dd_first = &u->dedup_first;
  }
  ex.post_cells = [=](const Pipeline& p, hipStream_t st2) -> hipError_t {
    hipError_t e = launch_updateable_cells(p, state, excl, full ? nullptr : am, full ? 1 : 0, refresh ? 1 : 0, st2);
    if (e != hipSuccess) return e;
    // the final marks travel back with the run's own synchronisation
    if (mask_out != nullptr && (e = hipMemcpyAsync(mask_out, am, n, hipMemcpyDeviceToHost, st2)) != hipSuccess)
      return e;
    if (!adj) return hipSuccess;
    unsigned long long* keys = nullptr;
    uint32_t* first = nullptr;
    if (p.periodic) {
      const uint64_t slots = dedup_slots(p.n_pairs);
      if ((e = u->dedup_keys.ensure(slots * sizeof(unsigned long long))) != hipSuccess) return e;
      if ((e = u->dedup_first.ensure(slots * sizeof(uint32_t))) != hipSuccess) return e;
      keys = u->dedup_keys.as<unsigned long long>();
      first = u->dedup_first.as<uint32_t>();
    }
    vltr_context* c = u->ctx;
    if ((e = c->adj_list.ensure(std::max<int64_t>(p.n_pairs, 1) * sizeof(int32_t))) != hipSuccess) return e;
    // the contact flags and their scan scratch are free until the contacts are compacted
    const AdjointWork work{c->flags.as<uint32_t>(), c->scan_tmp.as<uint32_t>(), c->adj_list.as<int32_t>()};
    return launch_adjoint(p, d_types, w16, d_grad, d_part, keys, first, work, st2);
  };
  size_t nc = 0, ncell = 0;
  int rc = run_pipeline(ctx, static_cast<int64_t>(n), st.probe, st.periodic ? &st.pbox : nullptr, nullptr, 0, 0, &nc,
                        &ncell, nullptr, 0, &ex);
  if (rc != VLTR_OK) return rc;
  if (static_cast<const uint32_t*>(ctx->pinned.ptr)[kReportExclFlip] != 0u) {
    // a new neighbor of a moved ball lost a stale exclusion (it was hidden
    // by a ball that has moved away, or excluded by the caller) after the
    // pairs were counted without it: the same run again, with the affected
    // set and exclusion statuses it left (spheres_container.rs:150-170
    // refreshes these balls before the contacts are computed)
    ex.pre_clip = nullptr;
    ex.front_check = nullptr;
    rc = run_pipeline(ctx, static_cast<int64_t>(n), st.probe, st.periodic ? &st.pbox : nullptr, nullptr, 0, 0, &nc,
                      &ncell, nullptr, 0, &ex);
    if (rc != VLTR_OK) return rc;
  }
  // this run's pair lists and records become the previous run's (the
  // context's buffers are scratch of the next run)
  std::swap(ctx->pair_cnt, u->prev_pair_cnt);
  std::swap(ctx->pair_b, u->prev_pair_b);
  std::swap(ctx->rec, u->prev_rec);
  if (adj) {
    std::swap(u->topo, u->prev_topo);
    std::swap(u->topo_slot, u->prev_topo_slot);
    u->prev_topos = kTopoRegions * topo_region_for(u->topo_cap);
  }
  u->prev_ok = true;
  u->prev_topo_ok = adj;
  u->prev_n = n;
  u->prev_pairs = static_cast<int64_t>(ctx->counts[2]);
  if (adj) {
    std::vector<double> parts(static_cast<size_t>(adjoint_blocks()));
    VLTR_HIP(hipMemcpy(parts.data(), d_part, parts.size() * sizeof(double), hipMemcpyDeviceToHost));
    double e = 0.0;
    for (double v : parts) e += v;
    u->energy = e;
    u->grad_valid = true;
  }
  st.n_contacts = nc;
  st.n_cells = ncell;
  st.empty = ncell == 0;
  return VLTR_OK;
}

// updateable.rs:404-420
int upd_prepare(Upd* u) {
  if (u->backup_enabled && !u->st.empty) {
    u->bk = u->st;
    const size_t n = u->st.balls.size();
    hipStream_t s = u->ctx->stream;
    VLTR_HIP(u->cells_backup.ensure(n * sizeof(CellRec)));
    VLTR_HIP(u->excl_backup.ensure(n));
    VLTR_HIP(hipMemcpyAsync(u->cells_backup.ptr, u->cells_state.ptr, n * sizeof(CellRec), hipMemcpyDeviceToDevice, s));
    VLTR_HIP(hipMemcpyAsync(u->excl_backup.ptr, u->excl.ptr, n, hipMemcpyDeviceToDevice, s));
    // the pair records the backup state was made of (the next run reads the
    // current ones as its previous records, so they are copied, not moved)
    u->bk_ok = u->prev_ok;
    u->bk_topo_ok = u->prev_topo_ok;
    u->bk_n = u->prev_n;
    u->bk_pairs = u->prev_pairs;
    u->bk_topos = u->prev_topos;
    if (u->prev_ok) {
      const size_t np = static_cast<size_t>(u->prev_pairs);
      VLTR_HIP(u->bk_pair_cnt.ensure((u->prev_n + 1) * sizeof(uint32_t)));
      VLTR_HIP(hipMemcpyAsync(u->bk_pair_cnt.ptr, u->prev_pair_cnt.ptr, (u->prev_n + 1) * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, s));
      if (np > 0) {
        VLTR_HIP(u->bk_pair_b.ensure(np * sizeof(int32_t)));
        VLTR_HIP(u->bk_rec.ensure(np * sizeof(PairRec)));
        VLTR_HIP(hipMemcpyAsync(u->bk_pair_b.ptr, u->prev_pair_b.ptr, np * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        VLTR_HIP(hipMemcpyAsync(u->bk_rec.ptr, u->prev_rec.ptr, np * sizeof(PairRec), hipMemcpyDeviceToDevice, s));
        if (u->prev_topo_ok) {
          const size_t nt = static_cast<size_t>(u->prev_topos);
          VLTR_HIP(u->bk_topo_slot.ensure(np * sizeof(uint32_t)));
          VLTR_HIP(hipMemcpyAsync(u->bk_topo_slot.ptr, u->prev_topo_slot.ptr, np * sizeof(uint32_t),
                                  hipMemcpyDeviceToDevice, s));
          if (nt > 0) {
            VLTR_HIP(u->bk_topo.ensure(nt * sizeof(ContactTopo)));
            VLTR_HIP(hipMemcpyAsync(u->bk_topo.ptr, u->prev_topo.ptr, nt * sizeof(ContactTopo),
                                    hipMemcpyDeviceToDevice, s));
          }
        }
      }
    }
    u->has_backup = true;
  }
  u->st.changed.clear();
  u->st.affected.clear();
  u->st.was_full_reinit = false;
  return VLTR_OK;
}

int upd_full(Upd* u, const vltr_ball* balls, size_t n) {
  u->st.balls.assign(balls, balls + n);
  u->st.changed.clear();
  u->st.affected.clear();
  u->st.was_full_reinit = true;
  RunOpts o;
  o.full = true;
  o.refresh = true;
  return upd_run(u, o);
}

// updateable.rs:422-447 with spheres_container.rs:71-182
int upd_update(Upd* u, const vltr_ball* balls, size_t nb, const uint64_t* changed_ids, size_t n_changed, int* ok) {
  *ok = 0;
  int rc = upd_prepare(u);
  if (rc != VLTR_OK) return rc;
  auto& st = u->st;
  vltr_context* ctx = u->ctx;
  // a changed ball count, or no complete last run to take neighbors from
  if (nb != st.balls.size() || ctx->last_n != nb) {
    rc = upd_full(u, balls, nb);
    if (rc == VLTR_OK) *ok = 1;
    return rc;
  }
  const size_t n = nb;
  const size_t threshold = std::max<size_t>(n / 2, 10);
  std::vector<uint64_t> changed;
  if (changed_ids) {
    for (size_t k = 0; k < n_changed; ++k) {
      const uint64_t id = changed_ids[k];
      if (id < n && !same_sphere(balls[id], st.balls[id], st.probe)) changed.push_back(id);
    }
  } else {
    for (size_t i = 0; i < n && changed.size() < threshold + 1; ++i)
      if (!same_sphere(balls[i], st.balls[i], st.probe)) changed.push_back(i);
  }
  if (changed.empty()) return VLTR_OK;
  if (changed.size() > threshold) {
    rc = upd_full(u, balls, nb);
    if (rc == VLTR_OK) *ok = 1;
    return rc;
  }
  // Affected marks on the device: 3 changed, 1 old colliding neighbor (the
  // last run's lists), 2 new colliding neighbor (marked in the run). The
  // changed ids and their new balls go up once, through pinned staging.
  const size_t k = changed.size();
  const size_t vals_off = (k * sizeof(int32_t) + 255) / 256 * 256;
  const size_t mask_off = vals_off + (k * sizeof(vltr_ball) + 255) / 256 * 256;
  if ((rc = upd_stage(u, mask_off + n)) != VLTR_OK) return rc;
  unsigned char* stage = static_cast<unsigned char*>(u->stage.ptr);
  int32_t* h_ids = reinterpret_cast<int32_t*>(stage);
  vltr_ball* h_vals = reinterpret_cast<vltr_ball*>(stage + vals_off);
  const uint8_t* h_mask = stage + mask_off;
  for (size_t j = 0; j < k; ++j) {
    h_ids[j] = static_cast<int32_t>(changed[j]);
    h_vals[j] = balls[changed[j]];
  }
  hipStream_t s = ctx->stream;
  VLTR_HIP(hipSetDevice(ctx->device));
  VLTR_HIP(u->ids.ensure(k * sizeof(int32_t)));
  VLTR_HIP(u->moved_vals.ensure(k * sizeof(vltr_ball)));
  VLTR_HIP(u->affected_mask.ensure(n));
  int32_t* d_ids = u->ids.as<int32_t>();
  uint8_t* am = u->affected_mask.as<uint8_t>();
  VLTR_HIP(hipMemcpyAsync(d_ids, h_ids, k * sizeof(int32_t), hipMemcpyHostToDevice, s));
  VLTR_HIP(hipMemcpyAsync(u->moved_vals.ptr, h_vals, k * sizeof(vltr_ball), hipMemcpyHostToDevice, s));
  VLTR_HIP(launch_fill_u8(am, static_cast<int64_t>(n), 0, s));
  VLTR_HIP(launch_mark_ids(am, static_cast<int64_t>(n), d_ids, static_cast<int64_t>(k), 3, s));
  Pipeline last{};
  set_balls(last, static_cast<int64_t>(n));
  last.nb = ctx->nb.as<int32_t>();
  last.nb_cnt = ctx->nb_cnt.as<int32_t>();
  last.nb_cap = ctx->last_nb_cap;
  VLTR_HIP(launch_mark_neighbors(last, d_ids, static_cast<int64_t>(k), am, 1, s));
  // the marks are counted on the device and reach the host with the run's
  // first report, where the threshold is checked before the pairs are made
  VLTR_HIP(launch_count_marks(am, static_cast<int64_t>(n), report_words(ctx) + kReportMarksOld, s));
  for (uint64_t id : changed) st.balls[id] = balls[id];
  // the reference inserts the old neighbors one by one into the sorted
  // changed list and gives up (full reinit) when one more would pass the
  // threshold, i.e. when |changed| + |new ids| > threshold
  // (spheres_container.rs:105-124); a changed id listed twice stays twice
  // until the final sort + unique
  bool repeats = false;
  RunOpts o;
  o.refresh = true;
  o.incremental = true;
  o.mask_ready = true;
  o.n_moved = static_cast<int64_t>(k);
  o.more = true;
  o.mask_out = stage + mask_off;
  o.front_check = [&](const uint32_t* hw) {
    repeats = hw[kReportMarksChanged] != k;
    return k + hw[kReportMarksOld] <= threshold;
  };
  rc = upd_run(u, o);
  if (rc == kRunAborted) {
    rc = upd_full(u, balls, nb);
    if (rc == VLTR_OK) *ok = 1;
    return rc;
  }
  if (rc != VLTR_OK) return rc;
  // the sorted affected list: the marked ids in order, with the new neighbors
  // of the changed balls (spheres_container.rs:150-170); without new ones, a
  // repeated changed id stays repeated
  std::vector<uint64_t>& affected = u->aff_scratch;
  bool more = false;
  for (size_t i = 0; i < n && !more; ++i) more = h_mask[i] == 2;
  if (repeats && !more) {
    affected = changed;
    for (size_t i = 0; i < n; ++i)
      if (h_mask[i] == 1) affected.push_back(i);
    std::sort(affected.begin(), affected.end());
  } else {
    marked_ids(h_mask, n, &affected);
  }
  st.changed.swap(changed);
  st.affected.swap(affected);  // the old list's storage is the next step's scratch
  st.was_full_reinit = false;
  *ok = 1;
  return VLTR_OK;
}

}  // namespace

extern "C" {

int vltr_updateable_create(int device, int backup_enabled, vltr_updateable** out) {
  if (!out) return fail(VLTR_ERR_INVALID_ARGUMENT, "null out");
  *out = nullptr;
  auto u = std::make_unique<vltr_updateable>();
  const int rc = vltr_context_create(device, &u->ctx);
  if (rc != VLTR_OK) return rc;
  u->backup_enabled = backup_enabled != 0;
  *out = u.release();
  return VLTR_OK;
}

void vltr_updateable_destroy(vltr_updateable* u) {
  if (!u) return;
  if (u->ctx) {
    (void)hipSetDevice(u->ctx->device);
    (void)hipStreamSynchronize(u->ctx->stream);
    for (DevBuf* b : {&u->cells_state, &u->cells_backup, &u->excl, &u->excl_backup, &u->affected_mask, &u->ids, &u->rows,
                      &u->rows_cnt, &u->moved_vals, &u->types, &u->grad, &u->part, &u->topo, &u->seed, &u->dedup_keys,
                      &u->dedup_first, &u->prev_pair_cnt, &u->prev_pair_b, &u->prev_rec, &u->prev_topo, &u->topo_slot,
                      &u->prev_topo_slot, &u->bk_topo_slot,
                      &u->bk_pair_cnt, &u->bk_pair_b, &u->bk_rec, &u->bk_topo})
      b->release();
    u->stage.release();
    vltr_context_destroy(u->ctx);
  }
  delete u;
}

// updateable.rs:187-202
int vltr_updateable_init(vltr_updateable* u, const vltr_ball* balls, size_t n, double probe,
                         const vltr_periodic_box* periodic_box, int* ok) {
  if (!u || !ok || (n > 0 && !balls)) return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  if (n > (size_t(1) << 31) / 27) return fail(VLTR_ERR_INVALID_ARGUMENT, "too many balls");
  *ok = 0;
  int rc = upd_prepare(u);
  if (rc != VLTR_OK) return rc;
  u->st.probe = probe;
  u->st.periodic = periodic_box != nullptr;
  if (periodic_box) u->st.pbox = *periodic_box;
  rc = upd_full(u, balls, n);
  if (rc != VLTR_OK) return rc;
  *ok = u->st.empty ? 0 : 1;
  return VLTR_OK;
}

// updateable.rs:208-213
int vltr_updateable_update(vltr_updateable* u, const vltr_ball* balls, size_t n, int* ok) {
  if (!u || !ok || (n > 0 && !balls)) return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  return upd_update(u, balls, n, nullptr, 0, ok);
}

// updateable.rs:215-220
int vltr_updateable_update_with_changed(vltr_updateable* u, const vltr_ball* balls, size_t n, const uint64_t* changed,
                                        size_t n_changed, int* ok) {
  if (!u || !ok || (n > 0 && !balls) || (n_changed > 0 && !changed))
    return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  *ok = 0;
  if (n_changed == 0) return VLTR_OK;
  return upd_update(u, balls, n, changed, n_changed, ok);
}

// updateable.rs:225-272 with spheres_container.rs:186-208
int vltr_updateable_set_exclusion(vltr_updateable* u, size_t id, int excluded, int* ok) {
  if (!u || !ok) return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  *ok = 0;
  auto& st = u->st;
  const size_t n = st.balls.size();
  if (st.empty || id >= n) return VLTR_OK;
  vltr_context* ctx = u->ctx;
  hipStream_t s = ctx->stream;
  VLTR_HIP(hipSetDevice(ctx->device));
  uint8_t cur = 0;
  VLTR_HIP(hipMemcpyAsync(&cur, u->excl.as<uint8_t>() + id, 1, hipMemcpyDeviceToHost, s));
  VLTR_HIP(hipStreamSynchronize(s));
  if ((cur != 0) == (excluded != 0)) return VLTR_OK;
  // affected: the ball and its contact partners (exclusion) or its colliding
  // neighbors (inclusion), read before the state changes
  std::vector<uint64_t> affected{static_cast<uint64_t>(id)};
  std::vector<uint64_t> partners;
  if (excluded) {
    std::vector<vltr_contact> cs(st.n_contacts);
    if (!cs.empty()) {
      VLTR_HIP(hipMemcpyAsync(cs.data(), ctx->out_contacts.ptr, cs.size() * sizeof(vltr_contact),
                              hipMemcpyDeviceToHost, s));
      VLTR_HIP(hipStreamSynchronize(s));
    }
    for (const auto& c : cs) {
      const uint64_t a = c.id_a % n, b = c.id_b % n;
      if (a == id) partners.push_back(b);
      if (b == id) partners.push_back(a);
    }
  } else {
    std::vector<std::vector<uint64_t>> rows;
    const int rc = upd_rows(u, {static_cast<uint64_t>(id)}, &rows);
    if (rc != VLTR_OK) return rc;
    partners = rows[0];
  }
  for (uint64_t nb : partners) {
    auto it = std::lower_bound(affected.begin(), affected.end(), nb);
    if (it == affected.end() || *it != nb) affected.insert(it, nb);
  }
  int rc = upd_prepare(u);
  if (rc != VLTR_OK) return rc;
  const uint8_t v = excluded ? 1 : 0;
  VLTR_HIP(hipMemcpyAsync(u->excl.as<uint8_t>() + id, &v, 1, hipMemcpyHostToDevice, s));
  st.changed = {static_cast<uint64_t>(id)};
  st.affected = affected;
  st.was_full_reinit = false;
  RunOpts o;
  o.incremental = true;
  o.aff = &affected;
  rc = upd_run(u, o);
  if (rc != VLTR_OK) return rc;
  *ok = 1;
  return VLTR_OK;
}

// updateable.rs:277-322: the state before the last update, exclusion change
// or init comes back whole.
int vltr_updateable_restore(vltr_updateable* u, int* ok) {
  if (!u || !ok) return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  *ok = 0;
  if (!u->backup_enabled || !u->has_backup) return VLTR_OK;
  u->st = u->bk;
  u->has_backup = false;
  const size_t n = u->st.balls.size();
  hipStream_t s = u->ctx->stream;
  VLTR_HIP(hipSetDevice(u->ctx->device));
  VLTR_HIP(u->cells_state.ensure(n * sizeof(CellRec)));
  VLTR_HIP(u->excl.ensure(n));
  VLTR_HIP(hipMemcpyAsync(u->cells_state.ptr, u->cells_backup.ptr, n * sizeof(CellRec), hipMemcpyDeviceToDevice, s));
  VLTR_HIP(hipMemcpyAsync(u->excl.ptr, u->excl_backup.ptr, n, hipMemcpyDeviceToDevice, s));
  const auto kept = u->st;
  // the backup state's pair records come back as the previous run's; a run
  // with no ball affected then gives every pair of the restored balls the
  // record it had (the reference restores the contact lists themselves,
  // updateable.rs:296-309, which may hold contacts a recompute would not
  // give, e.g. of a ball whose exclusion is stale) and every cell its
  // restored record
  u->prev_ok = false;
  if (u->bk_ok) {
    std::swap(u->prev_pair_cnt, u->bk_pair_cnt);
    std::swap(u->prev_pair_b, u->bk_pair_b);
    std::swap(u->prev_rec, u->bk_rec);
    std::swap(u->prev_topo, u->bk_topo);
    std::swap(u->prev_topo_slot, u->bk_topo_slot);
    u->prev_topos = u->bk_topos;
    u->prev_ok = true;
    u->prev_topo_ok = u->bk_topo_ok;
    u->prev_n = u->bk_n;
    u->prev_pairs = u->bk_pairs;
    u->bk_ok = false;
  }
  RunOpts o;
  o.incremental = true;
  const int rc = upd_run(u, o);
  if (rc != VLTR_OK) return rc;
  u->st.changed = kept.changed;
  u->st.affected = kept.affected;
  u->st.was_full_reinit = kept.was_full_reinit;
  *ok = 1;
  return VLTR_OK;
}

int vltr_updateable_set_energy(vltr_updateable* u, const int32_t* types, size_t n, const double* w16,
                               double sas_weight, double volume_weight) {
  if (!u || !w16 || (n > 0 && !types)) return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  for (size_t i = 0; i < n; ++i)
    if (types[i] < 0 || types[i] > 3) return fail(VLTR_ERR_INVALID_ARGUMENT, "types must be in 0..3");
  VLTR_HIP(hipSetDevice(u->ctx->device));
  VLTR_HIP(u->types.ensure(std::max<size_t>(n, 1) * sizeof(int32_t)));
  if (n > 0) VLTR_HIP(hipMemcpy(u->types.ptr, types, n * sizeof(int32_t), hipMemcpyHostToDevice));
  u->w16.assign(w16, w16 + 16);
  u->sas_weight = sas_weight;
  u->volume_weight = volume_weight;
  u->n_types = n;
  u->energy_on = true;
  u->grad_valid = false;
  return VLTR_OK;
}

namespace {
// The energy of the current state: from the last pass, or from a pass over
// the unchanged state (contacts recomputed, cells kept) when the energy was
// switched on after it.
int upd_energy(vltr_updateable* u) {
  if (!u->energy_on) return fail(VLTR_ERR_INVALID_ARGUMENT, "no energy set (vltr_updateable_set_energy)");
  if (u->n_types != u->st.balls.size()) return fail(VLTR_ERR_INVALID_ARGUMENT, "one type per ball required");
  if (!u->grad_valid && !u->st.balls.empty()) {
    const auto kept = u->st;
    RunOpts o;
    o.incremental = true;
    const int rc = upd_run(u, o);
    if (rc != VLTR_OK) return rc;
    u->st.changed = kept.changed;
    u->st.affected = kept.affected;
    u->st.was_full_reinit = kept.was_full_reinit;
  }
  return VLTR_OK;
}
}  // namespace

int vltr_updateable_energy_gradient(vltr_updateable* u, double* energy, double* grad) {
  if (!u || !energy || (!grad && !u->st.balls.empty())) return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  const int rc = upd_energy(u);
  if (rc != VLTR_OK) return rc;
  const size_t n = u->st.balls.size();
  *energy = n ? u->energy : 0.0;
  if (n) VLTR_HIP(hipMemcpy(grad, u->grad.ptr, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
  return VLTR_OK;
}

int vltr_updateable_energy_gradient_device(vltr_updateable* u, double* energy, const double** d_grad) {
  if (!u || !energy || !d_grad) return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  const int rc = upd_energy(u);
  if (rc != VLTR_OK) return rc;
  *energy = u->st.balls.empty() ? 0.0 : u->energy;
  *d_grad = u->st.balls.empty() ? nullptr : u->grad.as<double>();
  return VLTR_OK;
}

// updateable.rs:333-376: contacts listed from each ball's view with
// id_a == ball (periodic: the first of each canonical pair, ids ordered).
int vltr_updateable_summary(const vltr_updateable* u, vltr_result** out) {
  if (!u || !out) return fail(VLTR_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  auto r = std::make_unique<vltr_result>();
  const size_t n = u->st.balls.size();
  r->num_balls = n;
  if (n > 0) {
    vltr_context* ctx = u->ctx;
    hipStream_t s = ctx->stream;
    VLTR_HIP(hipSetDevice(ctx->device));
    r->contacts.resize(u->st.n_contacts);
    r->cells.resize(u->st.n_cells);
    const vltr_contact* src = ctx->out_contacts.as<const vltr_contact>();
    const int64_t nc = static_cast<int64_t>(u->st.n_contacts);
    if (u->st.periodic && nc > 0) {
      // the run reported the first record of each canonical pair with
      // canonical ids: order each pair's ids and list the contacts by their
      // first ball, in pair order within a ball
      const size_t tmp_bytes = contacts_sort_tmp_bytes(nc);
      if (tmp_bytes == SIZE_MAX) return fail(VLTR_ERR_DEVICE, "contact sort scratch query");
      ContactSortWork W{};
      auto carve = [&](Arena& A) {
        W.tmp_bytes = tmp_bytes;
        W.tmp = A.take<unsigned char>(tmp_bytes);
        W.key = A.take<uint32_t>(nc);
        W.key_out = A.take<uint32_t>(nc);
        W.idx = A.take<uint32_t>(nc);
        W.perm = A.take<uint32_t>(nc);
        W.out = A.take<vltr_contact>(nc);
      };
      Arena size{nullptr, SIZE_MAX};
      carve(size);
      VLTR_HIP(ctx->net_work.ensure(size.used));
      Arena A{ctx->net_work.as<unsigned char>(), ctx->net_work.bytes};
      carve(A);
      if (!A.fits()) return fail(VLTR_ERR_INTERNAL, "contact sort scratch");
      VLTR_HIP(launch_contacts_by_first(src, nc, static_cast<int64_t>(n), W, s));
      src = W.out;
    }
    if (!r->contacts.empty())
      VLTR_HIP(hipMemcpyAsync(r->contacts.data(), src, r->contacts.size() * sizeof(vltr_contact),
                              hipMemcpyDeviceToHost, s));
    if (!r->cells.empty())
      VLTR_HIP(hipMemcpyAsync(r->cells.data(), ctx->out_cells.ptr, r->cells.size() * sizeof(vltr_cell),
                              hipMemcpyDeviceToHost, s));
    VLTR_HIP(hipStreamSynchronize(s));
  }
  *out = r.release();
  return VLTR_OK;
}

size_t vltr_updateable_changed_ids(const vltr_updateable* u, uint64_t* out, size_t max_ids) {
  if (!u) return 0;
  const auto& v = u->st.changed;
  for (size_t k = 0; out && k < v.size() && k < max_ids; ++k) out[k] = v[k];
  return v.size();
}
size_t vltr_updateable_affected_ids(const vltr_updateable* u, uint64_t* out, size_t max_ids) {
  if (!u) return 0;
  const auto& v = u->st.affected;
  for (size_t k = 0; out && k < v.size() && k < max_ids; ++k) out[k] = v[k];
  return v.size();
}
int vltr_updateable_last_update_was_full_reinit(const vltr_updateable* u) { return u && u->st.was_full_reinit ? 1 : 0; }
int vltr_updateable_backup_enabled(const vltr_updateable* u) { return u && u->backup_enabled ? 1 : 0; }
int vltr_updateable_last_timings(const vltr_updateable* u, double* ms, int max_entries) {
  return u ? vltr_context_last_timings(u->ctx, ms, max_entries) : 0;
}

}  // extern "C"
