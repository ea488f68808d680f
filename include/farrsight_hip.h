/* farrsight_hip.h - C ABI of the MI355X (gfx950) implementation of the
 * FARRSIGHT hot path: the periodic regularized E-field (direct sum and
 * barycentric treecode) and the remesh interpolation of f onto the panel
 * tree, plus the device-resident step loop that calls them.
 *
 * Conventions (those of the reference's ElectricField::operator(),
 * inc/FieldStructure.hpp:87-93): the caller owns every buffer, outputs are
 * overwritten, counts are element counts. Every entry point returns 0 on
 * success or a negative FS_E_* code; fs_last_error() describes the last
 * failure on the calling thread. Entry points named *_dev take device
 * pointers and enqueue on the given HIP stream (NULL = the context's own
 * stream) without synchronizing; the others take host pointers and are
 * synchronous.
 */
#ifndef FARRSIGHT_HIP_H
#define FARRSIGHT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FS_OK = 0,
    FS_E_INVALID = -1,   /* bad argument (null pointer, negative size, L <= 0) */
    FS_E_NOMEM = -2,     /* device allocation failed */
    FS_E_HIP = -3,       /* a HIP runtime call failed */
    FS_E_RANGE = -4,     /* input outside what the algorithm accepts
                            ("too many shifts!" of AMRStructure_interp.cpp:469) */
    FS_E_INTERNAL = -5   /* a device-side bounded search did not finish */
};

typedef struct fs_ctx fs_ctx;

/* Context: one HIP device, one default-priority stream, grow-only device
 * scratch. Not thread-safe; use one context per host thread. */
int fs_ctx_create(fs_ctx** out, int device);
int fs_ctx_destroy(fs_ctx* ctx);
/* The context's HIP stream (a hipStream_t). */
void* fs_ctx_stream(fs_ctx* ctx);
int fs_ctx_synchronize(fs_ctx* ctx);
const char* fs_last_error(void);

/* ---- Part A: periodic regularized E-field ------------------------------- */

/* Replaces E_MQ_DirectSum::operator() (src/FieldStructure.cpp:106-156):
 *   es[i] = sum_j q_ws[j] * (-z L) (1/sqrt(L^2+4eps^2) - 1/sqrt(4 z^2 L^2 + 4 eps^2)),
 *   z = (targets[i]-sources[j])/L wrapped to [-1/2, 1/2].
 * FP64 throughout, compensated accumulation. Host pointers, synchronous. */
int fs_emq_periodic_f64(fs_ctx* ctx, double* es, const double* targets, int64_t nt,
                        const double* sources, const double* q_ws, int64_t ns,
                        double L, double epsilon);

/* Same operator on device-resident arrays, enqueued on `stream`. */
int fs_emq_periodic_f64_dev(fs_ctx* ctx, double* d_es, const double* d_targets,
                            int64_t nt, const double* d_sources, const double* d_q_ws,
                            int64_t ns, double L, double epsilon, void* stream);

/* Allocation-free form of fs_emq_periodic_f64_dev, with the same arguments
 * and results (bit for bit: the pair-once partials run in passes within the
 * work space, which changes no result). No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Its work
 * space is a fixed 256 MiB array per device, held until the process ends;
 * the pair-once sum (targets = sources from 16384 points) takes about 150-210 B
 * per point (its partials and the sort of the points), so it accepts at most
 * about 1.6e6 such points and refuses larger
 * ones with FS_E_RANGE before any launch. Calls overlapping on one device
 * race for that array. Calls no hipMalloc, hipFree or allocating routine. */
int fs_emq_periodic_f64_noalloc(fs_ctx* ctx, double* d_es, const double* d_targets, int64_t nt,
                                const double* d_sources, const double* d_q_ws, int64_t ns,
                                double L, double epsilon, void* stream);

/* Replaces E_MQ_Treecode::operator() (src/FieldStructure.cpp:230-1050):
 * the periodic barycentric-Lagrange treecode (bisection tree with leaves of
 * fewer than max_source points, MAC theta = mac with the periodic distance,
 * Chebyshev degree `degree` <= 15), evaluated at the sources like the
 * reference (es[0:ns]). Sort, tree, interaction lists, moments and far/near
 * sums all on the device. evals (may be NULL) receives {far, near}
 * kernel-evaluation counts. Host pointers, synchronous. FS_E_RANGE if more
 * than max_source points share one coordinate (the tree cannot separate
 * them). */
int fs_emq_treecode_f64(fs_ctx* ctx, double* es, const double* sources, const double* q_ws,
                        int64_t ns, double L, double epsilon, double mac, int degree,
                        int max_source, int64_t* evals);

/* The tree and interaction lists fs_emq_treecode_f64 builds on the device
 * for the points `sources` (E_MQ_Treecode's build_tree_init,
 * build_tree_1D_Recursive and split_2, src/FieldStructure.cpp:506-632, and
 * build_interaction_list, :663-702), copied out for inspection. stats:
 * [0] nodes, [1] leaves, [2] far entries, [3] near entries. A node is given
 * as three int32 (first, last, depth): its member range in tree order and
 * its depth below the root (0). perm[ns]: the input index of each point in
 * tree order (leaves in increasing x, each leaf's points in input order);
 * leaf_nodes[3 * leaves] in increasing x; far_off and near_off[leaves + 1]
 * are offsets into far_nodes[3 * far entries] and near_nodes[3 * near
 * entries], each leaf's entries in the reference's list order. Any array may
 * be NULL (stats alone sizes the others). Host pointers, synchronous. */
int fs_treecode_tree(fs_ctx* ctx, const double* sources, int64_t ns, double L, double mac,
                     int degree, int max_source, int64_t stats[4], int32_t* perm,
                     int32_t* leaf_nodes, int32_t* far_off, int32_t* far_nodes,
                     int32_t* near_off, int32_t* near_nodes);

/* Replaces E_MQ_DirectSum_openbcs::operator() (src/FieldStructure.cpp:166-191):
 *   es[i] = sum_j q_ws[j] * 0.5 * z / sqrt(z^2 + eps^2) - targets[i],
 *   z = targets[i] - sources[j]
 * (open boundaries, no wrap; -targets[i] is the neutralizing background of
 * unit density). FP64, compensated accumulation; requires epsilon > 0. With
 * the targets being the sources (the same pointer, or for this host entry
 * point two arrays of identical contents) and 16384 points or more, each pair
 * is evaluated once and given to both points (the kernel is odd). Host
 * pointers, synchronous. */
int fs_emq_open_f64(fs_ctx* ctx, double* es, const double* targets, int64_t nt,
                    const double* sources, const double* q_ws, int64_t ns, double epsilon);

/* Same operator on device-resident arrays, enqueued on `stream`. */
int fs_emq_open_f64_dev(fs_ctx* ctx, double* d_es, const double* d_targets, int64_t nt,
                        const double* d_sources, const double* d_q_ws, int64_t ns,
                        double epsilon, void* stream);

/* Replaces ExternalTanh / ExternalSine / ExternalLogistic::operator()
 * (src/FieldStructure.cpp:17-139, constants inc/FieldStructure.hpp:49-84):
 *   es[i] += a(t) * sin(k * targets[i] - omega * t),
 * type 1 the sine ramp, 2 the logistic ramp, any other value the tanh drive
 * (as the deck loader's default branch, AMRSimulation_load.cpp:110-124). es is
 * read and written. Host pointers, synchronous. */
int fs_external_field_add(fs_ctx* ctx, double* es, const double* targets, int64_t nt, double t,
                          int type, double Am, double k, double omega);

/* Same on device-resident arrays, enqueued on `stream`. */
int fs_external_field_add_dev(fs_ctx* ctx, double* d_es, const double* d_targets, int64_t nt,
                              double t, int type, double Am, double k, double omega,
                              void* stream);

/* Tuning knobs of the direct-sum kernels (0 = automatic). targets_per_thread
 * in {1,2,4,8} (8: the FP64 pair-once form only, its default from 2^19 points;
 * the mixed-precision kernels always take 4: each wavefront makes its
 * near/far choices for its 256 consecutive sorted targets); source_splits >=
 * 1. Applies to later calls on this context. */
int fs_emq_set_tuning(fs_ctx* ctx, int targets_per_thread, int source_splits);

/* FP64 direct sum with the targets being the sources (nt == ns and the same
 * pointer; for the host entry point also two arrays of identical contents,
 * as the reference passes them): the kernel is odd, so each pair can be
 * evaluated once and given to both points. mode 0 (default) = from 16384
 * points up, 1 = never, 2 = at every size. Periodic and open-boundary sums
 * alike (targets_per_thread and source_splits of fs_emq_set_tuning also
 * shape the pair-once form of both; the open-boundary full sum has a fixed
 * launch shape). No counterpart in the reference
 * (its E_MQ_DirectSum evaluates every ordered pair); the result agrees with
 * the full sum to rounding. Applies to later calls on this context. */
int fs_emq_set_symmetric(fs_ctx* ctx, int mode);

/* Bound on the pair-once form's partial sums (row and column partials and
 * the per-point state carried between passes), in bytes; 0 (the default) =
 * 4 GiB. The column partials of a whole evaluation take about n^2 / 512
 * bytes (2 GiB at n = 2^20, 32 GiB at 2^22); beyond the bound the source
 * distances are evaluated in passes, each folded into the per-point sums in
 * the same order, so the result does not depend on the bound. At least one
 * pass of 16 source blocks is always taken (about 160 B per point). Applies
 * to later periodic and open-boundary calls on this context. No counterpart
 * in the reference. */
int fs_emq_set_pair_once_budget(fs_ctx* ctx, int64_t bytes);

/* Host-only: the pair-once plan of n points (targets = sources) under a
 * budget (0 = the default): device scratch in bytes, passes, and kernel
 * evaluations. Any output may be NULL. No counterpart in the reference. */
int fs_emq_pair_once_plan(int64_t n, int64_t budget_bytes, int64_t* scratch_bytes, int* passes,
                          int64_t* evals);

/* Kernel evaluations (pairs, padding of the last block or tile included) of
 * the last FP64 direct sum launched on this context (periodic or open):
 * nt * ns for the full sum, about n^2 / 2 for the pair-once form. No counterpart in the reference
 * (for throughput accounting). */
int fs_emq_f64_pair_evals(fs_ctx* ctx, int64_t* evals);

/* Mixed-precision form of fs_emq_periodic_f64 (no counterpart in the
 * reference; the FP64 entry points stay the drop-in). Positions are wrapped
 * in FP64 and differenced exactly in 64-bit fixed point; the step and
 * neutralizing part of the kernel is summed in FP64 from sorted prefix sums,
 * the regularization remainder in FP64 within B = max(2^-6, 16 d, 4000 d^2)
 * of z = 0 (d = eps/L; every pair in FP64 when B >= 1/4; compact one-sided
 * ranges of sorted sources from a Legendre expansion of 6 to 12 terms by
 * their distance, the rest pair by pair) and in FP32 beyond, where the
 * largest range of 2048, 512, 128 or 32 sorted sources that is compact
 * against its distance is summed from an 8-moment expansion about its
 * centre, a range compact enough for the FP64 expansion from that, and the
 * others pair by pair; compensated FP64 accumulation per target.
 * Relative error max_i|E_i - E_i^FP64| / max_i|E_i^FP64|: 1e-10 to 1e-8 at
 * that band on the tested decks; with fs_emq_set_mixed_target each call
 * measures the error of candidate bands first and takes the narrowest one
 * within the given target. Requires
 * epsilon > 0. When targets == sources and nt == ns the sort is shared.
 * Host pointers, synchronous. */
int fs_emq_periodic_mixed(fs_ctx* ctx, double* es, const double* targets, int64_t nt,
                          const double* sources, const double* q_ws, int64_t ns,
                          double L, double epsilon);

/* Error target of the mixed entry points on this context (no counterpart in
 * the reference). 0 (the default): the minimum band B above. > 0: before
 * its pair sums, each call measures on up to 128 probed targets the E error
 * that each of 8 candidate bands (B/2 to 4 sqrt(2) B) would give against the
 * exact FP64 field, and takes the narrowest one within rel_err * max|E|
 * (every pair in FP64 when none is). At most 1e-3. */
int fs_emq_set_mixed_target(fs_ctx* ctx, double rel_err);

/* Size below which the mixed entry points return the plain FP64 sum (no
 * counterpart in the reference). A call with nt * ns < min_pairs evaluates
 * through fs_emq_periodic_f64(_dev) and returns its result bit for bit: at
 * such sizes the mixed form's fixed costs (sort, expansions, prefix sums,
 * finalize: some 20 short launches) exceed the whole FP64 sum. The rule
 * depends on the call's own sizes only. 0 sends every call through the
 * mixed kernels; the default is FS_MIXED_MIN_PAIRS_DEFAULT. Step loops and
 * servers keep their own precision (fs_batch_create, fs_sim_set_options). */
#define FS_MIXED_MIN_PAIRS_DEFAULT (INT64_C(1) << 29)
int fs_emq_set_mixed_min_pairs(fs_ctx* ctx, int64_t min_pairs);

/* Launch shape (targets per thread, persistent workgroups per CU) that the
 * last call through the mixed kernels used, as chosen by its autotuner.
 * Informational. */
int fs_emq_mixed_shape(fs_ctx* ctx, int* targets_per_thread, int* wgs_per_cu);

/* Same on device-resident arrays, enqueued on `stream`. */
int fs_emq_periodic_mixed_dev(fs_ctx* ctx, double* d_es, const double* d_targets,
                              int64_t nt, const double* d_sources, const double* d_q_ws,
                              int64_t ns, double L, double epsilon, void* stream);

/* ---- Part B: remesh interpolation -------------------------------------- */

/* Panel records: FS_PANEL_INTS int32 per panel, the fields of struct Panel
 * (inc/Panel.hpp:41-56):
 *   [0] panel_ind, [1..9] point indices (Panel.hpp:57-75 order), [10] level,
 *   [11] parent, [12] which_child, [13] left, [14] top, [15] right,
 *   [16] bottom neighbour (-1: use the parent's, -2: boundary),
 *   [17] flags (1 left boundary, 2 right boundary, 4 needs refinement,
 *        8 refined in x and p, 16 refined in p only), [18] child_inds_start,
 *   [19] unused. Panel 0 is the root. */
#define FS_PANEL_INTS 20

/* Replaces AMRStructure::interpolate_to_initial_xps
 * (src/structure_files/AMRStructure_interp.cpp:493-793): f of the old mesh
 * (panel records, advected old points old_xs/old_ps with values old_fs)
 * interpolated onto the nx*np points new_xs/new_ps of a base grid (np rows
 * of equal p with nx points each, any order). Periodic pre-shift, row sort,
 * neighbour-walk leaf search and 9-point interpolant as the reference;
 * points found beyond a p boundary get f_beyond. Host pointers. */
int fs_remesh_interp_grid(fs_ctx* ctx, const int32_t* panels, int64_t n_panels,
                          const double* old_xs, const double* old_ps, const double* old_fs,
                          int64_t n_old, const double* new_xs, const double* new_ps, int nx,
                          int np, double x_min, double x_max, int periodic, double f_beyond,
                          double* f_out);

/* Replaces AMRStructure::interpolate_from_mesh(x, p)
 * (AMRStructure_interp.cpp:1094-1143) at n points: periodic pre-shift,
 * root descent, single-point interpolant. Host pointers. */
int fs_remesh_interp_points(fs_ctx* ctx, const int32_t* panels, int64_t n_panels,
                            const double* old_xs, const double* old_ps, const double* old_fs,
                            int64_t n_old, const double* x, const double* p, int64_t n,
                            double x_min, double x_max, int periodic, double f_beyond,
                            double* f_out);

/* ---- The step loop: AMRSimulation on the device ------------------------- */

/* One species of a deck (the per-species keys of the JSON deck,
 * src/simulation_files/AMRSimulation_load.cpp:139-254). ics_type as the
 * reference (1/2 Landau, 3 strong two-stream, 4 colder two-stream,
 * 9/10 ion acoustic); 11 is a bump-on-tail (bulk Maxwellian pth centred at
 * pstr plus a beam_frac beam at beam_v with width beam_vth), which the
 * reference does not have. */
typedef struct fs_species_params {
    double q, m;
    double p_min, p_max;
    int initial_height, p_height, max_height;
    int do_adaptively_refine;
    int n_amr_epsilons;
    double amr_epsilons[5];
    int ics_type;
    double normalized_wavenumber, amp, pth, pstr;
    double beam_frac, beam_v, beam_vth;
} fs_species_params;

/* The deck's top-level keys (AMRSimulation_load.cpp:14-136). The field
 * follows make_field_return_ptr (:49-96): periodic decks take the direct sum
 * or, with use_treecode, the treecode; open decks (bcs 1) the open-boundary
 * direct sum, or with use_treecode the periodic direct sum (the reference's
 * selection). Open decks run in FP64 only. The external drive is the deck's
 * "external_field" block (:99-136): when use_external_field is set,
 * a(t) sin(k x - omega t) is added to every field evaluation at the
 * reference's stage times (ExternalSine for external_type 1,
 * ExternalLogistic for 2, ExternalTanh otherwise). */
typedef struct fs_sim_params {
    double x_min, x_max;
    int bcs;            /* 0 periodic, 1 open */
    int quadrature;     /* 0 trapezoid, 1 Simpson */
    double greens_epsilon;
    int use_treecode;
    double mac;
    int degree, max_source;
    double dt;
    int remesh_period;
    int use_external_field;
    int external_type;
    double Am, k, omega;
} fs_sim_params;

typedef struct fs_sim fs_sim;

/* Host-only: the initial mesh of one species (AMRStructure ctor,
 * src/structure_files/AMRStructure.cpp:91-126: base mesh, f0, adaptive
 * refinement, quadrature weights). Sizes into n_points/n_panels; any output
 * array may be NULL (call once with NULLs to size the buffers). */
int fs_mesh_build(const fs_sim_params* sim, const fs_species_params* species,
                  int64_t* n_points, int64_t* n_panels, int32_t* panels, double* xs,
                  double* ps, double* fs, double* q_ws);

/* AMRSimulation ctor (src/simulation_files/AMRSimulation.cpp:5-93): builds
 * every species' mesh from f0 and evaluates E on the reduced grid. The
 * simulation's state stays in device memory. */
int fs_sim_create(fs_ctx* ctx, const fs_sim_params* sim, const fs_species_params* species,
                  int n_species, fs_sim** out);
/* Destroy every simulation and server of a context before the context. */
int fs_sim_destroy(fs_sim* sim);
/* AMRSimulation::step (AMRSimulation_step.cpp:5-99) without file output:
 * one RK4 step (three field evaluations), then a remesh with the reduced-x
 * field evaluation, or a full field evaluation. Synchronous. */
int fs_sim_step(fs_sim* sim);
int fs_sim_iter(fs_sim* sim, int* iter);
int fs_sim_species_counts(fs_sim* sim, int species, int64_t* n_points, int64_t* n_panels);
/* Copies one species' points (any pointer may be NULL); es is E at the points
 * after the last field evaluation. */
int fs_sim_get_points(fs_sim* sim, int species, double* xs, double* ps, double* fs,
                      double* q_ws, double* es);
int fs_sim_get_panels(fs_sim* sim, int species, int32_t* panels);
/* out[0] charge, [1] momentum, [2] kinetic energy, [3] electric energy
 * (standard_interface/diagnostic_plot.py:116-131); the first four values of
 * fs_sim_diagnostics_full. */
int fs_sim_diagnostics(fs_sim* sim, int species, double out[4]);

/* The per-species diagnostics of standard_interface/diagnostic_plot.py:
 * 116-143, reduced on the device: [0] charge sum q_w, [1] momentum
 * sum p q_w / q, [2] kinetic energy sum p^2 q_w / (2 q/m), [3] electric
 * energy 1/2 sum dx_k E_k^2 over the distinct x in [0, L) (E at the first
 * occurrence, periodic midpoint spacing), [4] entropy
 * sum_{f>0} f log10(f) q_w / q, [5] L1 sum |q_w| / |q|, [6] L2 sum q_w f / q,
 * [7] min f, [8] max f, [9] negative area sum_{f<0} q_w / q, [10] number of
 * points with f < 0, [11] their fraction, [12] negative mass
 * |sum_{f<0} q_w| / sum |q_w|, [13] number of points. Sums in a fixed order
 * (they depend on the species' points alone). n_out >= FS_DIAG_COUNT. */
#define FS_DIAG_COUNT 14
int fs_sim_diagnostics_full(fs_sim* sim, int species, double* out, int n_out);

/* ---- Batches of independent decks (no counterpart in the reference) -----
 * One fs_sim may hold many decks; each deck's field sees only its own
 * charges. All decks step together on one stream: every field evaluation of
 * a step is one batched launch over all decks, every remesh kernel one
 * launch over all species, and the whole static step is captured once in a
 * HIP graph and replayed. Species are numbered over the batch, deck by deck
 * (fs_sim_deck_info gives each deck's first species). A deck's results do
 * not depend on which other decks share the batch. */
enum { FS_PRECISION_F64 = 0, FS_PRECISION_MIXED = 1 };

typedef struct fs_deck {
    const fs_sim_params* sim;
    const fs_species_params* species;
    int n_species;
} fs_deck;

/* Times the remesh's device work (the grid interpolation of every species
 * and the reset of the points to the grid: shift, row sort, walks,
 * coefficients, evaluation; not the reduced-x field) with HIP events on the
 * step's stream. on = 1 runs later steps eagerly (no graph replay) and
 * records the events; fs_sim_remesh_ms gives the last step's figure when
 * that step remeshed. No counterpart in the reference. */
int fs_sim_set_timing(fs_sim* sim, int on);
int fs_sim_remesh_ms(fs_sim* sim, double* ms);

/* Where adaptive remeshes refine their trees (the adaptive loop of
 * generate_mesh, AMRStructure_amr.cpp:901-941, with refine_panels,
 * test_panel and set_leaves_weights): mode 0 on the device (default; the
 * passes run in parallel and give the sequential order's tree bit for bit),
 * 1 on the host (the sequential builder; also what a device remesh falls
 * back to if its passes meet a neighbour link they do not reproduce).
 * fs_sim_refinement_counts gives the remeshes refined each way. No
 * counterpart in the reference. */
int fs_sim_set_refinement(fs_sim* sim, int mode);
int fs_sim_refinement_counts(fs_sim* sim, int64_t* on_device, int64_t* on_host);

/* Steps taken by replaying a captured HIP graph and steps run eagerly (a step
 * in which an adaptive deck remeshes, any step of a batch with a treecode
 * deck, and every step with use_graph 0 or timing on). No counterpart in the
 * reference. */
int fs_sim_step_counts(fs_sim* sim, int64_t* replayed, int64_t* eager);

/* Host-only: the device bytes a deck is estimated to take in a batch or a
 * server (state, RK4 work, field partials -- in FP64 the pair-once partials of
 * its full field --, sorted copies in mixed precision, panel records); the
 * server admits decks against its budget with this estimate. Builds the
 * deck's meshes on the host. No counterpart in the reference. */
int fs_deck_footprint(const fs_deck* deck, int precision, int64_t* bytes, int64_t* n_points);

/* ---- The reference's on-disk state (no counterpart needed: these are the
 * reference's own formats, AMRStructure_io.cpp:3-101,
 * AMRStructure_restart.cpp:33-98, AMRSimulation_field.cpp:81-111) -----------
 * Per species under <sim_dir>/simulation_output/<species_name>/:
 * xs/xs_<iter>, ps/ps_<iter>, fs/fs_<iter>, qws/qws_<iter>, es/es_<iter> (raw
 * doubles, one per point), panels/leaf_point_inds_<iter> (9 int32 per leaf,
 * depth-first leaf order) and panels/full_tree_<iter> (int64 panel count,
 * then one 80-byte record per panel: panel_ind, level, parent_ind,
 * which_child, point_inds[9], left, top, right, bottom, child_inds_start as
 * int32, then is_left_bdry, is_right_bdry, is_refined_xp, is_refined_p,
 * needs_refinement as int8 and 3 bytes of padding). Missing directories are
 * created. */

/* Host-only: one species' files from host arrays (q_ws or es may be NULL:
 * those files are not written). panels: FS_PANEL_INTS int32 per panel. */
int fs_write_species_output(const char* sim_dir, const char* species_name, int iter, int64_t n,
                            const double* xs, const double* ps, const double* fs,
                            const double* q_ws, const double* es, int64_t n_panels,
                            const int32_t* panels);
/* Host-only: reads xs, ps, fs and full_tree of one species back (any output
 * may be NULL; call with NULL arrays to size them). */
int fs_read_species_output(const char* sim_dir, const char* species_name, int iter,
                           int64_t* n, double* xs, double* ps, double* fs, int64_t* n_panels,
                           int32_t* panels);

/* Creates a batch. Each deck remeshes at its own remesh_period (counted from
 * its own first step): a step where only some decks remesh evaluates the
 * full field for every deck and the remesh and reduced-x field for those;
 * precision FS_PRECISION_F64
 * (the reference's method) or FS_PRECISION_MIXED (fs_emq_periodic_mixed's
 * method for every field evaluation). fs_sim_create is a batch of one deck
 * in FP64. */
int fs_batch_create(fs_ctx* ctx, const fs_deck* decks, int n_decks, int precision,
                    fs_sim** out);
/* Continuous batching: a deck joins (built on the host, uploaded; the others
 * keep their device state and field) or leaves between steps, whatever the
 * decks' remesh periods. */
int fs_sim_add_deck(fs_sim* sim, const fs_deck* deck, int* index);
int fs_sim_remove_deck(fs_sim* sim, int deck);
/* A batch with no decks yet, for decks that join
 * through fs_sim_add_deck or fs_sim_restart_deck. */
int fs_batch_create_empty(fs_ctx* ctx, int precision, fs_sim** out);
/* AMRSimulation::write_to_file (AMRSimulation_io.cpp:7-17) and the reduced
 * set of AMRSimulation_field.cpp:81-111 for one deck at the step it reached:
 * each species' files (names: one per species of the deck), and
 * <sim_dir>/simulation_output/reduce_xs/xs_<iter> and reduced_es/es_<iter>
 * (the deck's reduced x, species after species, and E there from the last
 * reduced evaluation). */
int fs_sim_write_output(fs_sim* sim, int deck, const char* sim_dir,
                        const char* const* species_names);
/* The restart of AMRSimulation.cpp:65-121: a deck joins the batch from the
 * files of step iter instead of f0 -- its meshes are built from the deck,
 * then every species' points, f and panel tree replaced by the files, the
 * weights recomputed from the tree and E evaluated on the reduced set. The
 * files must hold a state written right after a remesh (points on their
 * mesh: the weights of moved points are not in the files; refused
 * otherwise). Continuing a deck this
 * way gives bitwise the states of the run that wrote the files. */
int fs_sim_restart_deck(fs_sim* sim, const fs_deck* deck, const char* sim_dir,
                        const char* const* species_names, int iter, int* index);
int fs_sim_num_decks(fs_sim* sim, int* n_decks);
/* Any output may be NULL. iter counts the steps the deck has taken. */
int fs_sim_deck_info(fs_sim* sim, int deck, int* first_species, int* n_species, int* iter,
                     int64_t* n_points);
/* precision as above; use_graph 0 runs every step eagerly (steps in which an
 * adaptive deck remeshes, and batches with a treecode deck, always do). */
int fs_sim_set_options(fs_sim* sim, int precision, int use_graph);
/* Mixed precision: what the accuracy governor measured in the deck's last
 * measured full-field evaluation (RK4 stages 2 and 4 measure their own error
 * before their pair sums -- up to 128 probed targets against the exact FP64
 * field; stage 3 and the full field after a step without remesh keep the
 * band of the stage before; every reduced-x evaluation measures its own). max_e = max|E| over the probed targets; errs[q] = the largest E
 * error there when pairs beyond bands[q] (fraction of the period; >= 1/4
 * means every pair in FP64) take the FP32 path, q = 0..7. The evaluation
 * took the narrowest band with errs[q] <= 4e-9 max_e. Any output may be
 * NULL (errs, bands: 8 doubles). */
int fs_sim_field_governor(fs_sim* sim, int deck, double* max_e, double* errs, double* bands);

/* ---- Serving loop: continuous batching (no counterpart in the reference) --
 * A server owns a context's device and an initially empty batch. Requests
 * (a deck and a step count) may be submitted from any thread at any time;
 * a builder thread makes the deck's meshes on the host, and a worker thread
 * admits built decks between steps in arrival order while a slot is free and
 * the estimated device footprint of the resident decks stays within
 * hbm_budget_bytes, steps all resident decks together, and retires finished
 * ones with their final state copied to the host. While a server runs, its
 * context must not be used by other calls. */
enum {
    FS_REQ_QUEUED = 0,
    FS_REQ_BUILDING = 1,
    FS_REQ_READY = 2,    /* built, waiting for a slot or device memory */
    FS_REQ_ACTIVE = 3,   /* resident, stepping */
    FS_REQ_DONE = 4,
    FS_REQ_FAILED = -1
};

typedef struct fs_server_params {
    double hbm_budget_bytes;
    int max_active;      /* resident decks at most (0 = no limit) */
    int precision;       /* FS_PRECISION_F64 or FS_PRECISION_MIXED */
    int use_graph;
} fs_server_params;

typedef struct fs_server_stats {
    int64_t submitted, completed, failed, active, queued;
    int64_t steps;           /* batch steps taken */
    int64_t point_steps;     /* sum over steps of the resident points */
    double busy_s;           /* time spent in steps */
    double span_s;           /* first admission to last retirement */
    double point_steps_per_s;  /* point_steps / span_s */
    double step_p50_ms, step_p99_ms;  /* per-step latency of the resident decks */
    int64_t hbm_in_use;      /* estimated bytes of the resident decks */
    int64_t steps_replayed;  /* steps taken by a captured-graph replay (steps in
                                which no adaptive deck remeshed) */
} fs_server_stats;

typedef struct fs_server fs_server;

int fs_server_create(fs_ctx* ctx, const fs_server_params* params, fs_server** out);
int fs_server_submit(fs_server* server, const fs_deck* deck, int n_steps, int64_t* request_id);
int fs_server_status(fs_server* server, int64_t request_id, int* state, int* steps_done);
/* Blocks until the request is done or failed (timeout_s <= 0: no limit). */
int fs_server_wait(fs_server* server, int64_t request_id, double timeout_s);
/* Final state of one species of a finished request (any output may be NULL;
 * n_points first). diag as fs_sim_diagnostics. */
int fs_server_result(fs_server* server, int64_t request_id, int species, int64_t* n_points,
                     double* xs, double* ps, double* fs, double* q_ws, double* es,
                     double diag[4]);
int fs_server_get_stats(fs_server* server, fs_server_stats* out);
/* The final state of a finished request written in the reference's files
 * (fs_sim_write_output's per-species files, at the step the deck reached):
 * the checkpoint of a served deck. species_names: one per species. */
int fs_server_write_result(fs_server* server, int64_t request_id, const char* sim_dir,
                           const char* const* species_names);
/* A request that resumes a deck from such files (fs_sim_restart_deck's
 * semantics) and takes n_steps more steps. */
int fs_server_submit_restart(fs_server* server, const fs_deck* deck, const char* sim_dir,
                             const char* const* species_names, int iter, int n_steps,
                             int64_t* request_id);
/* Stops the threads; requests not finished are dropped. */
int fs_server_destroy(fs_server* server);

#ifdef __cplusplus
}
#endif
#endif /* FARRSIGHT_HIP_H */
