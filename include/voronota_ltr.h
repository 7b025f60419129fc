/*
 * voronota_ltr.h - C ABI of the MI355X tessellation library.
 *
 * Drop-in boundary for voronota-ltr's tessellation hot path: the crate's
 * public Rust API (src/lib.rs:59-63) and the PyO3 module (src/python.rs).
 * Plain C types only; results live in library-owned handles released with the
 * matching *_free call. Every function returns an int status (VLTR_OK = 0)
 * unless documented otherwise, never aborts across the ABI, and leaves a
 * message for vltr_last_error() on failure.
 */
#ifndef VORONOTA_LTR_H_
#define VORONOTA_LTR_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status ---- */
#define VLTR_OK 0
#define VLTR_ERR_INVALID_ARGUMENT 1
#define VLTR_ERR_DEVICE 2
#define VLTR_ERR_OUT_OF_MEMORY 3
#define VLTR_ERR_INTERNAL 4

/* Sentinel for "no ball" in cell vertex / edge indices (types.rs:98 NULL_ID). */
#define VLTR_NULL_ID UINT64_MAX

/* ------------------------------------------------------------- types ---- */

/* Ball: centre and radius (types.rs:12-21). 32 bytes, AoS. */
typedef struct vltr_ball {
  double x, y, z, r;
} vltr_ball;

/* Periodic box as three lattice vectors (types.rs:290-300 PeriodicBox). */
typedef struct vltr_periodic_box {
  double shift_a[3];
  double shift_b[3];
  double shift_c[3];
} vltr_periodic_box;

/* Contact between two balls (types.rs:66-83). In non-periodic results
 * id_a < id_b; periodic results carry canonical ids as the reference does
 * (tessellation.rs:488-497), which need not be ordered. */
typedef struct vltr_contact {
  uint64_t id_a;
  uint64_t id_b;
  double area;
  double arc_length;
  uint8_t central;
  uint8_t reserved[7];
} vltr_contact;

/* Cell of one ball (types.rs:85-95). */
typedef struct vltr_cell {
  uint64_t index;
  double sas_area;
  double volume;
} vltr_cell;

/* Tessellation network vertex / edge (types.rs:100-144); VLTR_NULL_ID marks
 * the solvent-accessible surface. */
typedef struct vltr_cell_vertex {
  uint64_t ball_indices[4];
  double x, y, z;
} vltr_cell_vertex;

typedef struct vltr_cell_edge {
  uint64_t ball_indices[3];
  double length;
} vltr_cell_edge;

typedef struct vltr_context vltr_context;
typedef struct vltr_result vltr_result;
typedef struct vltr_batch_result vltr_batch_result;
typedef struct vltr_updateable vltr_updateable;
typedef struct vltr_solvent_result vltr_solvent_result;

/* Pseudo-solvent sphere (solvent_spheres.rs:15-32). */
typedef struct vltr_solvent_sphere {
  double x, y, z;
  double radius;
  double weight;
  uint64_t parent_index;
} vltr_solvent_sphere;

/* ------------------------------------------------------------ context --- */

/* Binds a HIP device and a default-priority stream; owns device buffers that
 * are reused across calls. Calls on distinct contexts may run concurrently;
 * calls on one context must be serialized by the caller. No reference
 * counterpart (the reference runs on the Rayon pool, main.rs:139-145). */
int vltr_context_create(int device, vltr_context** out);
void vltr_context_destroy(vltr_context* ctx);

/* Message of the last failed call on this thread ("" if none). */
const char* vltr_last_error(void);

/* Per-phase device time of the last pipeline run on ctx, in milliseconds,
 * measured with HIP events on the context stream. Order: grid (prepare,
 * cell keys, counting sort and scatter of the spheres), neighbors, pairs,
 * clip (classification and contour clipping of every candidate pair), slow
 * (the pairs of balls with more than 64 neighbors and contours beyond 8
 * points), cells+compaction, total. Returns the count written. */
int vltr_context_last_timings(const vltr_context* ctx, double* ms, int max_entries);

/* Work counts of the last pipeline run on ctx: out[0] balls, [1] searched
 * spheres (27 n when periodic), [2] candidate pairs clipped, [3] contacts,
 * [4] cells, [5] pairs clipped outside the main clip kernel (first ball with
 * more than 64 neighbors, or contour beyond 8 points), [6] neighbor slots
 * per ball, [7] cell-hash buckets. Returns the count written. */
int vltr_context_last_counts(const vltr_context* ctx, uint64_t* out, int max_entries);

/* ------------------------------------------------------- periodic box --- */

/* PeriodicBox::from_corners (types.rs:302-310). */
void vltr_periodic_box_from_corners(const double min_corner[3], const double max_corner[3], vltr_periodic_box* out);
/* PeriodicBox::from_vectors (types.rs:312-320). */
void vltr_periodic_box_from_vectors(const double a[3], const double b[3], const double c[3], vltr_periodic_box* out);

/* ------------------------------------------------------- tessellation --- */

/* compute_tessellation (tessellation.rs:37-48): contacts, per-ball SAS areas
 * and volumes of the radical tessellation constrained inside the SAS.
 * periodic_box and groups may be NULL; groups shorter than n only apply to
 * the balls they cover (tessellation.rs:323-328). */
int vltr_compute_tessellation(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe,
                              const vltr_periodic_box* periodic_box, const int32_t* groups, size_t n_groups,
                              int with_cell_vertices, vltr_result** out);

/* compute_contacts_only (tessellation.rs:125-166): contacts without solid
 * angles, volumes or cells. */
int vltr_compute_contacts_only(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe,
                               const vltr_periodic_box* periodic_box, const int32_t* groups, size_t n_groups,
                               vltr_result** out);

/* Device-resident variant of vltr_compute_tessellation / contacts_only
 * (contacts_only != 0): d_balls is a device pointer to n vltr_ball; results
 * stay on the device in buffers owned by ctx (valid until the next call on
 * ctx) and are read with vltr_context_device_outputs. Non-periodic, no cell
 * vertices. No reference counterpart: it is the host entry point minus the
 * host<->device copies. */
int vltr_compute_tessellation_device(vltr_context* ctx, const vltr_ball* d_balls, size_t n, double probe,
                                     const int32_t* d_groups, size_t n_groups, int contacts_only);
/* Device pointers and counts of the last device-resident call. */
int vltr_context_device_outputs(const vltr_context* ctx, const vltr_contact** d_contacts, size_t* n_contacts,
                                const vltr_cell** d_cells, size_t* n_cells);

/* Allocation-free variant of vltr_compute_tessellation_device. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Every intermediate buffer comes from a
 * fixed 64 MiB device workspace per GPU, a file-scope device array whose
 * address is taken when the first context on that GPU is created; the call
 * itself allocates nothing. Outputs go to the caller's device buffers: at
 * most contacts_cap contacts (the count is *n_contacts; beyond the cap the
 * call fails after writing the first contacts_cap) and, in full mode, the
 * cells (cells_cap must be at least n). Runs that do not fit the workspace
 * (about 20,000 balls at protein density in full mode) fail with
 * VLTR_ERR_INVALID_ARGUMENT and a message naming the limit. Calls on one GPU
 * share its workspace: the library serializes them (a lock held until the
 * call's results are complete), and the workspace keeps the last call's
 * contents. Non-periodic; ctx supplies the stream. */
int vltr_compute_tessellation_device_noalloc(vltr_context* ctx, const vltr_ball* d_balls, size_t n, double probe,
                                             const int32_t* d_groups, size_t n_groups, int contacts_only,
                                             vltr_contact* d_contacts, size_t contacts_cap, vltr_cell* d_cells,
                                             size_t cells_cap, size_t* n_contacts, size_t* n_cells);

/* Neighbor lists of the last pipeline run on ctx (diagnostics and parity
 * tests): for each of the num_balls balls, its colliding spheres sorted by
 * (distance from the ball's centre to the intersection-circle centre, index),
 * the list SpheresSearcher::find_colliding_ids returns
 * (spheres_searcher.rs:253-326). counts[i] is list i's length (0 for hidden
 * balls), ids holds the lists back to back (at most max_ids entries written).
 * Returns the total number of entries. */
int64_t vltr_context_neighbor_lists(const vltr_context* ctx, size_t num_balls, int32_t* counts, int32_t* ids,
                                    size_t max_ids);

/* ------------------------------------------------------------ batched --- */

/* Many independent structures in one pipeline pass. Structure k is
 * balls[offsets[k] .. offsets[k+1]) (offsets[0] = 0, n_structs + 1 entries,
 * non-decreasing); its result equals compute_tessellation (or, with
 * contacts_only, compute_contacts_only) of those balls alone
 * (tessellation.rs:37-48, 125-166), with structure-local indices.
 * Non-periodic, no groups, no cell vertices. No reference counterpart: the
 * reference runs structures one call at a time. */
int vltr_compute_tessellation_batched(vltr_context* ctx, const vltr_ball* balls, const uint64_t* offsets,
                                      size_t n_structs, double probe, int contacts_only, vltr_batch_result** out);
size_t vltr_batch_num_structs(const vltr_batch_result* r);
/* Structure k's contacts / cells (count written to *count). */
const vltr_contact* vltr_batch_contacts(const vltr_batch_result* r, size_t k, size_t* count);
const vltr_cell* vltr_batch_cells(const vltr_batch_result* r, size_t k, size_t* count);
void vltr_batch_free(vltr_batch_result* r);

/* Device-resident batched run: d_balls is a device pointer, offsets a host
 * array. Outputs stay in ctx (vltr_context_device_outputs: all structures
 * back to back); vltr_context_batch_offsets copies the n_structs + 1
 * per-structure start positions of contacts and cells (NULL to skip). */
int vltr_compute_tessellation_batched_device(vltr_context* ctx, const vltr_ball* d_balls, const uint64_t* offsets,
                                             size_t n_structs, double probe, int contacts_only);
int vltr_context_batch_offsets(const vltr_context* ctx, size_t n_structs, uint64_t* contact_off, uint64_t* cell_off);

/* HIP-graph replays of the device-resident batched run. No reference
 * counterpart. With graphs on, vltr_compute_tessellation_batched_device
 * captures its whole pipeline in a HIP graph (pair buffers sized 25% above
 * the pair count of a direct run) once a call repeats the previous call's
 * d_balls pointer, offsets, probe, mode and clip launch shape, and replays
 * it, one launch and one synchronisation per call, while later calls keep
 * them and no other run has used the context in between (the balls' values
 * may change). Other calls, and a replay that finds more pairs or longer
 * neighbor lists than the capture holds, run directly. Results are identical
 * either way; vltr_context_last_timings then reports only the total. */
int vltr_context_set_graphs(vltr_context* ctx, int enable);
uint64_t vltr_context_graph_replays(const vltr_context* ctx);

/* Direct runs that skipped the synchronisation on the pair total. No
 * reference counterpart. A call without extras (plain, contacts-only,
 * batched, host or device balls) that repeats the last such call's ball
 * count, structure count, probe, mode, periodicity and groups enqueues the
 * whole pipeline at once, its pair buffers sized 25% above the last pair
 * total; a run that finds more pairs or longer neighbor lists than that is
 * redone with the synchronisation (counted in *misses). Results are
 * identical either way. */
uint64_t vltr_context_speculative_runs(const vltr_context* ctx, uint64_t* misses);

/* Launch shapes of the contact clip kernel and their autotuner. No reference
 * counterpart: the tuning surface of this implementation. Every shape gives
 * the same records; shapes differ in lanes per contact (8 or 16),
 * wavefronts per block and occupancy target. vltr_clip_variant_name returns
 * e.g. "g8w4o4" (NULL out of range); shape 0 is the default of a new
 * context. */
int vltr_clip_variant_count(void);
const char* vltr_clip_variant_name(int variant);
int vltr_context_set_clip_variant(vltr_context* ctx, int variant);
int vltr_context_clip_variant(const vltr_context* ctx);
/* Times every shape on a device-resident batch (arguments as
 * vltr_compute_tessellation_batched_device; one structure is a batch of one):
 * one untimed round over all shapes, then `repeats` rounds each running
 * every shape once; the median pipeline time of a shape (HIP events, ms)
 * goes to ms[variant] (max_entries entries, ms may be NULL). The fastest
 * shape stays set on the context and is returned in *chosen. */
int vltr_context_autotune(vltr_context* ctx, const vltr_ball* d_balls, const uint64_t* offsets, size_t n_structs,
                          double probe, int contacts_only, int repeats, double* ms, int max_entries, int* chosen);

/* ------------------------------------------------------------ adjoint --- */

/* Energy E = sum_contacts W[t_a][t_b] * area + sas_weight * sum_cells sas_area
 * + volume_weight * sum_cells volume of compute_tessellation
 * (tessellation.rs:37-48) and its gradient with respect to the ball centres
 * (radii and probe fixed), holding the forward pass's contour combinatorics
 * fixed. types[i] in 0..3, w16 = row-major 4x4 W; periodic_box may be NULL
 * (with a box the contact sum runs over the deduplicated contact list, as
 * compute_tessellation reports it). No groups.
 * grad receives 3 n doubles (x, y, z per ball). No reference counterpart:
 * the reference has no derivatives. */
int vltr_energy_gradient(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe,
                         const vltr_periodic_box* periodic_box, const int32_t* types, const double* w16,
                         double sas_weight, double volume_weight, double* energy, double* grad);
/* Same on device memory: d_balls, d_types, d_grad are device pointers; w16 and
 * energy are host memory. */
int vltr_energy_gradient_device(vltr_context* ctx, const vltr_ball* d_balls, size_t n, double probe,
                                const vltr_periodic_box* periodic_box, const int32_t* d_types, const double* w16,
                                double sas_weight, double volume_weight, double* d_grad, double* energy);

/* ----------------------------------------------------- solvent spheres --- */

/* compute_solvent_spheres (solvent_spheres.rs:131-168): subdivision_depth
 * 0..4 (12 .. 2562 samples per exposed atom; larger values act as 4);
 * periodic_box may be NULL. Invalid probe, volume probe or ball give
 * VLTR_ERR_INVALID_ARGUMENT with the reference's message. */
int vltr_compute_solvent_spheres(vltr_context* ctx, const vltr_ball* balls, size_t n, double probe,
                                 double volume_probe, int subdivision_depth, const vltr_periodic_box* periodic_box,
                                 vltr_solvent_result** out);
size_t vltr_solvent_count(const vltr_solvent_result* r);
const vltr_solvent_sphere* vltr_solvent_spheres(const vltr_solvent_result* r);
void vltr_solvent_free(vltr_solvent_result* r);

/* --------------------------------------------------------- updateable --- */

/* UpdateableTessellation (updateable.rs:154-742) on the device: a
 * tessellation kept up to date as balls move. Each handle owns its own
 * context (device buffers, stream); calls on one handle must be serialized.
 * Functions returning an int status report the reference's bool in *ok. */
/* UpdateableTessellation::new / with_backup (updateable.rs:164-182) */
int vltr_updateable_create(int device, int backup_enabled, vltr_updateable** out);
void vltr_updateable_destroy(vltr_updateable* u);
/* init (updateable.rs:187-202); *ok = a non-empty result was produced */
int vltr_updateable_init(vltr_updateable* u, const vltr_ball* balls, size_t n, double probe,
                         const vltr_periodic_box* periodic_box, int* ok);
/* update (updateable.rs:208-213): changed balls found by comparison
 * (sphere_equals_sphere); *ok = something changed */
int vltr_updateable_update(vltr_updateable* u, const vltr_ball* balls, size_t n, int* ok);
/* update_with_changed (updateable.rs:215-220): only the listed ids are
 * compared; an empty list is "no update" */
int vltr_updateable_update_with_changed(vltr_updateable* u, const vltr_ball* balls, size_t n, const uint64_t* changed,
                                        size_t n_changed, int* ok);
/* set_exclusion (updateable.rs:225-272) */
int vltr_updateable_set_exclusion(vltr_updateable* u, size_t id, int excluded, int* ok);
/* restore (updateable.rs:277-322): back to the state before the last
 * init / update / set_exclusion; *ok = 0 without backup */
int vltr_updateable_restore(vltr_updateable* u, int* ok);
/* Energy of the updateable's state, E = sum_contacts W[t_a][t_b] area +
 * sas_weight * sum SAS + volume_weight * sum volume (vltr_energy_gradient's
 * energy, with the state's box and exclusions), and its gradient with respect
 * to the centres, computed in the same device pass as every later state
 * change (init, update, set_exclusion, restore). types: one entry in 0..3 per
 * ball (the state's ball count must match). No reference counterpart: the
 * reference has no derivatives. */
int vltr_updateable_set_energy(vltr_updateable* u, const int32_t* types, size_t n, const double* w16,
                               double sas_weight, double volume_weight);
/* E and dE/dcentres (3 n doubles, host) of the current state. */
int vltr_updateable_energy_gradient(vltr_updateable* u, double* energy, double* grad);
/* Same, the gradient left on the device (valid until the next call on u). */
int vltr_updateable_energy_gradient_device(vltr_updateable* u, double* energy, const double** d_grad);

/* summary (updateable.rs:333-376) as a result handle (free with
 * vltr_result_free); contacts listed per first ball, periodic ids canonical
 * and ordered */
int vltr_updateable_summary(const vltr_updateable* u, vltr_result** out);
/* changed_ids / affected_ids (updateable.rs:378-392): total count returned,
 * at most max_ids written */
size_t vltr_updateable_changed_ids(const vltr_updateable* u, uint64_t* out, size_t max_ids);
size_t vltr_updateable_affected_ids(const vltr_updateable* u, uint64_t* out, size_t max_ids);
int vltr_updateable_last_update_was_full_reinit(const vltr_updateable* u);
int vltr_updateable_backup_enabled(const vltr_updateable* u);
/* Phase timings of the handle's last pipeline run (as
 * vltr_context_last_timings). */
int vltr_updateable_last_timings(const vltr_updateable* u, double* ms, int max_entries);

/* ------------------------------------------------------------ results --- */

/* Results trait (types.rs:187-253) over TessellationResult (types.rs:256-287). */
size_t vltr_result_num_balls(const vltr_result* r);
size_t vltr_result_num_contacts(const vltr_result* r);
const vltr_contact* vltr_result_contacts(const vltr_result* r);
size_t vltr_result_num_cells(const vltr_result* r);
const vltr_cell* vltr_result_cells(const vltr_result* r);
int vltr_result_has_cell_vertices(const vltr_result* r);
size_t vltr_result_num_cell_vertices(const vltr_result* r);
const vltr_cell_vertex* vltr_result_cell_vertices(const vltr_result* r);
size_t vltr_result_num_cell_edges(const vltr_result* r);
const vltr_cell_edge* vltr_result_cell_edges(const vltr_result* r);
double vltr_result_total_sas_area(const vltr_result* r);
double vltr_result_total_volume(const vltr_result* r);
double vltr_result_total_contact_area(const vltr_result* r);
/* sas_areas()/volumes(): num_balls values, NaN where the reference has None. */
void vltr_result_sas_areas(const vltr_result* r, double* out);
void vltr_result_volumes(const vltr_result* r, double* out);
void vltr_result_free(vltr_result* r);

#ifdef __cplusplus
}
#endif

#endif /* VORONOTA_LTR_H_ */
