// C ABI over the CPU oracle (TEST INFRASTRUCTURE ONLY; see vltr_oracle.hpp).
// Loaded with ctypes by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg. Inputs are plain arrays; results live behind an opaque
// handle read through accessors and released with oracle_result_free.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "vltr_oracle.hpp"

using namespace vltr_oracle;

namespace {

struct OracleResult {
  size_t num_balls = 0;
  std::vector<Contact<double>> contacts;
  std::vector<Cell<double>> cells;
  bool has_network = false;
  std::vector<CellVertexOut> vertices;
  std::vector<CellEdgeOut> edges;
  // CSR neighbor lists (collision search)
  std::vector<uint64_t> nb_offsets;
  std::vector<uint64_t> nb_index;
  std::vector<double> nb_value;
  std::vector<uint8_t> excluded;
  // pair list (a, b) of the collision step, in reference order
  std::vector<uint64_t> pairs;
};

template <class R>
std::vector<Ball<R>> load_balls(const double* balls, size_t n) {
  std::vector<Ball<R>> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = Ball<R>{R(balls[4 * i]), R(balls[4 * i + 1]), R(balls[4 * i + 2]), R(balls[4 * i + 3])};
  return v;
}

template <class R>
PeriodicBox<R> load_pbox(const double* p) {
  PeriodicBox<R> b;
  b.shift_a = V3<R>{R(p[0]), R(p[1]), R(p[2])};
  b.shift_b = V3<R>{R(p[3]), R(p[4]), R(p[5])};
  b.shift_c = V3<R>{R(p[6]), R(p[7]), R(p[8])};
  return b;
}

void fill_result(OracleResult* r, TessellationResult<double>&& t) {
  r->num_balls = t.num_balls;
  r->contacts = std::move(t.contacts);
  r->cells = std::move(t.cells);
  r->has_network = t.has_network;
  r->vertices = std::move(t.vertices);
  r->edges = std::move(t.edges);
}

// Energy used by the differentiation workload:
//   E = sum_contacts W[t_a][t_b] * area + sum_cells sas_area + sum_cells volume
template <class R>
R energy_of(const TessellationResult<R>& t, const int32_t* types, const double* W) {
  R e = R(0.0);
  for (const auto& c : t.contacts) e = e + R(W[4 * types[c.id_a] + types[c.id_b]]) * c.area;
  for (const auto& c : t.cells) e = e + c.sas_area + c.volume;
  return e;
}

}  // namespace

extern "C" {

typedef struct OracleResult oracle_result;

// tessellation.rs:37-48
int oracle_compute_tessellation(const double* balls, size_t n, double probe, const double* pbox9, const int32_t* groups,
                                size_t n_groups, int with_vertices, int threads, oracle_result** out) {
  auto b = load_balls<double>(balls, n);
  PeriodicBox<double> pb;
  if (pbox9) pb = load_pbox<double>(pbox9);
  auto* r = new OracleResult();
  fill_result(r, compute_tessellation<double>(b, probe, pbox9 ? &pb : nullptr, groups, n_groups, with_vertices != 0, threads));
  *out = r;
  return 0;
}

// tessellation.rs:125-166
int oracle_compute_contacts_only(const double* balls, size_t n, double probe, const double* pbox9, const int32_t* groups,
                                 size_t n_groups, int threads, oracle_result** out) {
  auto b = load_balls<double>(balls, n);
  PeriodicBox<double> pb;
  if (pbox9) pb = load_pbox<double>(pbox9);
  auto* r = new OracleResult();
  r->num_balls = n;
  r->contacts = compute_contacts_only<double>(b, probe, pbox9 ? &pb : nullptr, groups, n_groups, threads);
  *out = r;
  return 0;
}

// Collision step only (tessellation.rs:63-111): per-ball sorted neighbor lists
// (index, key), the hidden flag and the collected pair list.
int oracle_collisions(const double* balls, size_t n, double probe, const double* pbox9, const int32_t* groups,
                      size_t n_groups, int threads, oracle_result** out) {
  auto b = load_balls<double>(balls, n);
  PeriodicBox<double> pb;
  if (pbox9) pb = load_pbox<double>(pbox9);
  auto* r = new OracleResult();
  r->num_balls = n;
  if (n > 0) {
    std::vector<Sphere<double>> sp;
    for (const auto& x : b) sp.push_back(sphere_from_ball(x, probe));
    SpheresSearcher<double> s(pbox9 ? populate_periodic_spheres(pb, sp) : sp);
    std::vector<std::vector<ValuedId<double>>> all(n);
    r->excluded.assign(n, 0);
    parallel_for(n, threads, [&](size_t i) {
      auto cr = s.find_colliding_ids(i, true);
      all[i] = std::move(cr.colliding_ids);
      r->excluded[i] = cr.excluded ? 1 : 0;
    });
    r->nb_offsets.push_back(0);
    for (size_t i = 0; i < n; ++i) {
      for (const auto& v : all[i]) {
        r->nb_index.push_back(v.index);
        r->nb_value.push_back(v.value);
      }
      r->nb_offsets.push_back(r->nb_index.size());
    }
    for (const auto& p : collect_collision_pairs(all, groups, n_groups, pbox9 ? n : 0)) {
      r->pairs.push_back(p.first);
      r->pairs.push_back(p.second);
    }
  } else {
    r->nb_offsets.push_back(0);
  }
  *out = r;
  return 0;
}

// Per-pair summaries of the collision step's pair list, in pair order
// (contact.rs:347-404): out P x 6 (area, arc, solid_a, solid_b, pyr_a, pyr_b),
// flags P (bit0 contact exists, bit1 central). mode 0 full, 1 area only.
int oracle_pair_records(const double* balls, size_t n, double probe, const double* pbox9, const int32_t* groups,
                        size_t n_groups, int mode, int threads, double* out, uint32_t* flags) {
  auto b = load_balls<double>(balls, n);
  PeriodicBox<double> pb;
  if (pbox9) pb = load_pbox<double>(pbox9);
  if (n == 0) return 0;
  auto ctx = make_collision_context<double>(b, probe, pbox9 ? &pb : nullptr, groups, n_groups, threads);
  parallel_for(ctx.pairs.size(), threads, [&](size_t k) {
    const size_t a = ctx.pairs[k].first, bb = ctx.pairs[k].second;
    for (int j = 0; j < 6; ++j) out[6 * k + j] = 0.0;
    flags[k] = 0;
    if (mode == 0) {
      auto cd = construct_contact_descriptor(ctx.searcher.spheres(), a, bb, ctx.all_collisions[a]);
      if (!cd) return;
      const auto s = cd->to_summary();
      out[6 * k + 0] = s.area;
      out[6 * k + 1] = s.arc_length;
      out[6 * k + 2] = s.solid_angle_a;
      out[6 * k + 3] = s.solid_angle_b;
      out[6 * k + 4] = s.pyramid_volume_a;
      out[6 * k + 5] = s.pyramid_volume_b;
      flags[k] = 1u | (s.central ? 2u : 0u);
    } else {
      auto ao = construct_contact_area(ctx.searcher.spheres(), a, bb, ctx.all_collisions[a]);
      if (!ao) return;
      out[6 * k + 0] = ao->area;
      out[6 * k + 1] = ao->arc_length;
      flags[k] = 1u | (ao->central ? 2u : 0u);
    }
  });
  return 0;
}

// Clipped contact polygons for graphics.rs:45-84 (non-periodic): for each
// pair with area > 0, in pair order, 14 doubles (a, b, ic centre xyz, ic r,
// axis xyz, barycenter xyz, m) then m x 6 doubles (x, y, z, angle, left id,
// right id). Returns the number of doubles (out may be NULL to size it).
int64_t oracle_contact_polygons(const double* balls, size_t n, double probe, const int32_t* groups, size_t n_groups,
                                double* out, int64_t max_out) {
  auto b = load_balls<double>(balls, n);
  if (n == 0) return 0;
  auto ctx = make_collision_context<double>(b, probe, nullptr, groups, n_groups, 1);
  std::vector<double> v;
  for (const auto& pr : ctx.pairs) {
    const size_t a = pr.first, bb = pr.second;
    auto cc = clip_contact(ctx.searcher.spheres(), a, bb, ctx.all_collisions[a]);
    if (!cc || !(cc->area > 0.0)) continue;
    const double head[14] = {double(a), double(bb), cc->ic.center.x, cc->ic.center.y, cc->ic.center.z, cc->ic.r,
                             cc->axis.x, cc->axis.y, cc->axis.z, cc->barycenter.x, cc->barycenter.y,
                             cc->barycenter.z, double(cc->contour.size()), 0.0};
    v.insert(v.end(), head, head + 14);
    for (const auto& cp : cc->contour) {
      const double pt[6] = {cp.p.x, cp.p.y, cp.p.z, cp.angle, double(cp.left_id), double(cp.right_id)};
      v.insert(v.end(), pt, pt + 6);
    }
  }
  if (out) std::copy(v.begin(), v.begin() + std::min<int64_t>(max_out, static_cast<int64_t>(v.size())), out);
  return static_cast<int64_t>(v.size());
}

size_t oracle_result_num_balls(const oracle_result* r) { return r->num_balls; }
size_t oracle_result_num_contacts(const oracle_result* r) { return r->contacts.size(); }
size_t oracle_result_num_cells(const oracle_result* r) { return r->cells.size(); }
int oracle_result_has_network(const oracle_result* r) { return r->has_network ? 1 : 0; }
size_t oracle_result_num_vertices(const oracle_result* r) { return r->vertices.size(); }
size_t oracle_result_num_edges(const oracle_result* r) { return r->edges.size(); }
size_t oracle_result_num_neighbors(const oracle_result* r) { return r->nb_index.size(); }
size_t oracle_result_num_pairs(const oracle_result* r) { return r->pairs.size() / 2; }

void oracle_result_contacts(const oracle_result* r, uint64_t* ids, double* area, double* arc, uint8_t* central) {
  for (size_t i = 0; i < r->contacts.size(); ++i) {
    const auto& c = r->contacts[i];
    ids[2 * i] = c.id_a;
    ids[2 * i + 1] = c.id_b;
    area[i] = c.area;
    arc[i] = c.arc_length;
    central[i] = c.central ? 1 : 0;
  }
}
void oracle_result_cells(const oracle_result* r, uint64_t* index, double* sas, double* vol) {
  for (size_t i = 0; i < r->cells.size(); ++i) {
    index[i] = r->cells[i].index;
    sas[i] = r->cells[i].sas_area;
    vol[i] = r->cells[i].volume;
  }
}
void oracle_result_vertices(const oracle_result* r, uint64_t* ids4, double* xyz) {
  for (size_t i = 0; i < r->vertices.size(); ++i) {
    for (int k = 0; k < 4; ++k) ids4[4 * i + k] = r->vertices[i].ids[k];
    xyz[3 * i] = r->vertices[i].x;
    xyz[3 * i + 1] = r->vertices[i].y;
    xyz[3 * i + 2] = r->vertices[i].z;
  }
}
void oracle_result_edges(const oracle_result* r, uint64_t* ids3, double* len) {
  for (size_t i = 0; i < r->edges.size(); ++i) {
    for (int k = 0; k < 3; ++k) ids3[3 * i + k] = r->edges[i].ids[k];
    len[i] = r->edges[i].length;
  }
}
void oracle_result_neighbors(const oracle_result* r, uint64_t* offsets, uint64_t* index, double* value, uint8_t* excluded) {
  std::memcpy(offsets, r->nb_offsets.data(), r->nb_offsets.size() * 8);
  if (!r->nb_index.empty()) {
    std::memcpy(index, r->nb_index.data(), r->nb_index.size() * 8);
    std::memcpy(value, r->nb_value.data(), r->nb_value.size() * 8);
  }
  if (excluded && !r->excluded.empty()) std::memcpy(excluded, r->excluded.data(), r->excluded.size());
}
void oracle_result_pairs(const oracle_result* r, uint64_t* pairs) {
  if (!r->pairs.empty()) std::memcpy(pairs, r->pairs.data(), r->pairs.size() * 8);
}
void oracle_result_free(oracle_result* r) { delete r; }

// Algorithmic work of one call, counted with the Counted scalar
// (mode 0: compute_tessellation, 1: compute_contacts_only, 2: collision search only).
int oracle_count_flops(const double* balls, size_t n, double probe, const double* pbox9, const int32_t* groups,
                       size_t n_groups, int mode, uint64_t* flops, uint64_t* tops) {
  auto b = load_balls<Counted>(balls, n);
  PeriodicBox<Counted> pb;
  if (pbox9) pb = load_pbox<Counted>(pbox9);
  flop_counters() = FlopCounters{};
  if (mode == 0) {
    (void)compute_tessellation<Counted>(b, Counted(probe), pbox9 ? &pb : nullptr, groups, n_groups, false, 1);
  } else if (mode == 1) {
    (void)compute_contacts_only<Counted>(b, Counted(probe), pbox9 ? &pb : nullptr, groups, n_groups, 1);
  } else {
    (void)make_collision_context<Counted>(b, Counted(probe), pbox9 ? &pb : nullptr, groups, n_groups, 1);
  }
  *flops = flop_counters().flops;
  *tops = flop_counters().tops;
  return 0;
}

// Energy of the differentiation workload (no groups; pbox9 may be NULL).
int oracle_energy(const double* balls, size_t n, double probe, const double* pbox9, const int32_t* types,
                  const double* W16, int threads, double* energy) {
  auto b = load_balls<double>(balls, n);
  PeriodicBox<double> pb;
  if (pbox9) pb = load_pbox<double>(pbox9);
  auto t = compute_tessellation<double>(b, probe, pbox9 ? &pb : nullptr, nullptr, 0, false, threads);
  *energy = energy_of(t, types, W16);
  return 0;
}

// Forward-mode derivative of the energy along coordinate `axis` (0,1,2) of
// ball `ball_id`, with the contour combinatorics of the unperturbed input.
int oracle_energy_dual(const double* balls, size_t n, double probe, const double* pbox9, const int32_t* types,
                       const double* W16, size_t ball_id, int axis, double* energy, double* denergy) {
  auto b = load_balls<Dual>(balls, n);
  PeriodicBox<Dual> pb;
  if (pbox9) pb = load_pbox<Dual>(pbox9);
  if (ball_id < n) {
    if (axis == 0) b[ball_id].x.d = 1.0;
    if (axis == 1) b[ball_id].y.d = 1.0;
    if (axis == 2) b[ball_id].z.d = 1.0;
  }
  auto t = compute_tessellation<Dual>(b, Dual(probe), pbox9 ? &pb : nullptr, nullptr, 0, false, 1);
  const Dual e = energy_of(t, types, W16);
  *energy = e.v;
  *denergy = e.d;
  return 0;
}

// Per-term energy contributions (contacts keyed by pair, cells by index) so a
// caller can difference term by term without cancellation in a large total.
int oracle_energy_terms(const double* balls, size_t n, double probe, const int32_t* types, const double* W16, int threads,
                        oracle_result** out) {
  auto b = load_balls<double>(balls, n);
  auto* r = new OracleResult();
  fill_result(r, compute_tessellation<double>(b, probe, nullptr, nullptr, 0, false, threads));
  for (auto& c : r->contacts) c.area = W16[4 * types[c.id_a] + types[c.id_b]] * c.area;
  *out = r;
  return 0;
}

// ----------------------------------------------------------- updateable ----
typedef struct OracleUpd oracle_updateable;
struct OracleUpd {
  UpdateableTessellation<double> t;
  explicit OracleUpd(bool backup, int threads) : t(backup, threads) {}
};

oracle_updateable* oracle_updateable_new(int with_backup, int threads) { return new OracleUpd(with_backup != 0, threads); }
void oracle_updateable_free(oracle_updateable* u) { delete u; }
// Moved spheres leave their old grid boxes (off: the reference's behaviour).
void oracle_updateable_set_remove_from_old_box(oracle_updateable* u, int on) { u->t.set_remove_from_old_box(on != 0); }
int oracle_updateable_init(oracle_updateable* u, const double* balls, size_t n, double probe, const double* pbox9) {
  auto b = load_balls<double>(balls, n);
  PeriodicBox<double> pb;
  if (pbox9) pb = load_pbox<double>(pbox9);
  return u->t.init(b, probe, pbox9 ? &pb : nullptr) ? 1 : 0;
}
int oracle_updateable_update(oracle_updateable* u, const double* balls, size_t n) {
  return u->t.update(load_balls<double>(balls, n)) ? 1 : 0;
}
int oracle_updateable_update_with_changed(oracle_updateable* u, const double* balls, size_t n, const uint64_t* ids,
                                          size_t n_ids) {
  std::vector<size_t> ch(ids, ids + n_ids);
  return u->t.update_with_changed(load_balls<double>(balls, n), ch) ? 1 : 0;
}
int oracle_updateable_set_exclusion(oracle_updateable* u, size_t id, int excluded) {
  return u->t.set_exclusion(id, excluded != 0) ? 1 : 0;
}
int oracle_updateable_restore(oracle_updateable* u) { return u->t.restore() ? 1 : 0; }
int oracle_updateable_last_update_was_full_reinit(const oracle_updateable* u) {
  return u->t.last_update_was_full_reinit() ? 1 : 0;
}
size_t oracle_updateable_num_changed(const oracle_updateable* u) { return u->t.changed_ids().size(); }
size_t oracle_updateable_num_affected(const oracle_updateable* u) { return u->t.affected_ids().size(); }
void oracle_updateable_changed_ids(const oracle_updateable* u, uint64_t* out) {
  for (size_t i = 0; i < u->t.changed_ids().size(); ++i) out[i] = u->t.changed_ids()[i];
}
void oracle_updateable_affected_ids(const oracle_updateable* u, uint64_t* out) {
  for (size_t i = 0; i < u->t.affected_ids().size(); ++i) out[i] = u->t.affected_ids()[i];
}
int oracle_updateable_summary(const oracle_updateable* u, oracle_result** out) {
  auto* r = new OracleResult();
  fill_result(r, u->t.summary());
  *out = r;
  return 0;
}

// Wall-clock of compute_tessellation repeated `reps` times (CPU baseline).
double oracle_time_tessellation(const double* balls, size_t n, double probe, const double* pbox9, const int32_t* groups,
                                size_t n_groups, int contacts_only, int threads, int reps) {
  auto b = load_balls<double>(balls, n);
  PeriodicBox<double> pb;
  if (pbox9) pb = load_pbox<double>(pbox9);
  const auto t0 = std::chrono::steady_clock::now();
  size_t sink = 0;
  for (int k = 0; k < reps; ++k) {
    if (contacts_only)
      sink += compute_contacts_only<double>(b, probe, pbox9 ? &pb : nullptr, groups, n_groups, threads).size();
    else
      sink += compute_tessellation<double>(b, probe, pbox9 ? &pb : nullptr, groups, n_groups, false, threads).contacts.size();
  }
  const auto t1 = std::chrono::steady_clock::now();
  if (sink == static_cast<size_t>(-1)) return -1.0;
  return std::chrono::duration<double>(t1 - t0).count() / (reps > 0 ? reps : 1);
}

int oracle_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// solvent_spheres.rs:131-168. out6: per sphere (x, y, z, radius, weight,
// parent); returns the count (call with out6 = NULL to size).
int64_t oracle_solvent_spheres(const double* balls, size_t n, double probe, double volume_probe, int depth,
                               const double* pbox9, int threads, double* out6, int64_t max_out) {
  auto b = load_balls<double>(balls, n);
  PeriodicBox<double> pb;
  if (pbox9) pb = load_pbox<double>(pbox9);
  const auto s = compute_solvent_spheres(b, probe, volume_probe, depth, pbox9 ? &pb : nullptr, threads);
  for (size_t i = 0; out6 && i < s.size() && static_cast<int64_t>(i) < max_out; ++i) {
    out6[6 * i + 0] = s[i].x;
    out6[6 * i + 1] = s[i].y;
    out6[6 * i + 2] = s[i].z;
    out6[6 * i + 3] = s[i].radius;
    out6[6 * i + 4] = s[i].weight;
    out6[6 * i + 5] = static_cast<double>(s[i].parent);
  }
  return static_cast<int64_t>(s.size());
}

int64_t oracle_icosahedron(int depth, double* out3, int64_t max_out) {
  const auto v = subdivided_icosahedron(depth);
  for (size_t i = 0; out3 && i < v.size() && static_cast<int64_t>(i) < max_out; ++i) {
    out3[3 * i] = v[i].x;
    out3[3 * i + 1] = v[i].y;
    out3[3 * i + 2] = v[i].z;
  }
  return static_cast<int64_t>(v.size());
}

}  // extern "C"
