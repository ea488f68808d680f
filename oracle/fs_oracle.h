/* Test infrastructure only: C API of the FP64 CPU oracle (oracle/liboracle.so).
 *
 * The oracle restates the reference algorithm of FARRSIGHT (shujhan/
 * multi_species) for the hot path: the periodic regularized E-field (direct
 * sum and barycentric treecode), the panel mesh with adaptive refinement, the
 * remesh interpolation, the reduced-x field utilities, RK4 and the step loop.
 * Each function cites the reference file:line it follows. Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, as the
 * checker; the product (multi_species_amd/) never links it.
 */
#ifndef FS_ORACLE_H
#define FS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- field (part A) ---------------------------------------------------- */

/* E_MQ_DirectSum::operator() (src/FieldStructure.cpp:106-156), same
 * expression order, OpenMP over targets. */
int orc_direct_sum(double* es, const double* targets, int64_t nt,
                   const double* sources, const double* q_ws, int64_t ns,
                   double L, double epsilon);

/* The same sum evaluated in x87 extended precision (64-bit significand) with
 * compensated accumulation: the "exact" value the FP64 sums are measured
 * against. */
int orc_direct_sum_ext(double* es, const double* targets, int64_t nt,
                       const double* sources, const double* q_ws, int64_t ns,
                       double L, double epsilon);

/* E_MQ_DirectSum_openbcs::operator() (src/FieldStructure.cpp:166-191), same
 * expression order, OpenMP over targets; and its extended-precision form. */
int orc_direct_sum_open(double* es, const double* targets, int64_t nt, const double* sources,
                        const double* q_ws, int64_t ns, double epsilon);
int orc_direct_sum_open_ext(double* es, const double* targets, int64_t nt, const double* sources,
                            const double* q_ws, int64_t ns, double epsilon);

/* ExternalTanh / ExternalSine / ExternalLogistic::operator()
 * (src/FieldStructure.cpp:17-139): es[i] += a(t) sin(k x_i - omega t); type 1
 * sine, 2 logistic, otherwise tanh. orc_external_amplitude gives a(t). */
int orc_external_field_add(double* es, const double* targets, int64_t nt, double t, int type,
                           double Am, double k, double omega);
double orc_external_amplitude(int type, double t, double Am, double k);

/* E_MQ_Treecode::operator() (src/FieldStructure.cpp:230-1020): evaluates at
 * the sources, writes es[0:ns]. */
int orc_treecode(double* es, const double* sources, const double* q_ws,
                 int64_t ns, double L, double epsilon, double mac, int degree,
                 int max_source);

/* The tree and interaction lists of E_MQ_Treecode for `sources`
 * (src/FieldStructure.cpp:506-702) in the layout of the product's
 * fs_treecode_tree: stats {nodes, leaves, far entries, near entries}; a node
 * as (first, last, depth) in tree order; perm = input index of each
 * tree-order point; leaves in depth-first (increasing x) order; CSR lists.
 * -2 when more than max_source points share a coordinate. */
int orc_treecode_tree(const double* sources, int64_t ns, double L, double mac, int max_source,
                      int64_t stats[4], int32_t* perm, int32_t* leaf_nodes, int32_t* far_off,
                      int32_t* far_nodes, int32_t* near_off, int32_t* near_nodes);

/* Tree statistics of the last orc_treecode call on this thread:
 * out[0]=node count, out[1]=leaf count, out[2]=far kernel evaluations,
 * out[3]=near kernel evaluations. */
void orc_treecode_last_stats(int64_t out[4]);

/* ---- decks --------------------------------------------------------------- */

typedef struct orc_species_params {
    double q, m;
    double p_min, p_max;
    int initial_height, p_height, max_height;
    int do_adaptively_refine;
    int n_amr_epsilons;
    double amr_epsilons[5];
    int ics_type;               /* AMRSimulation_load.cpp:149-221 (+11 bump-on-tail) */
    double normalized_wavenumber, amp, pth, pstr;
    double beam_frac, beam_v, beam_vth;  /* ics_type 11 only */
} orc_species_params;

typedef struct orc_sim_params {
    double x_min, x_max;
    int bcs;                    /* 0 periodic, 1 open */
    int quadrature;             /* 0 trap, 1 Simpson */
    double greens_epsilon;
    int use_treecode;
    double mac;
    int degree, max_source;
    double dt;
    int remesh_period;
    int use_external_field;     /* AMRSimulation_load.cpp:99-136 */
    int external_type;          /* 1 sine, 2 logistic, else tanh */
    double Am, k, omega;
} orc_sim_params;

/* ---- one species' mesh (part B) ------------------------------------------ */

typedef struct orc_species orc_species;

/* AMRStructure ctor (src/structure_files/AMRStructure.cpp:91-126): builds the
 * initial mesh from f0. */
orc_species* orc_species_create(const orc_sim_params* sim,
                                const orc_species_params* sp);
void orc_species_destroy(orc_species* s);
int64_t orc_species_num_points(const orc_species* s);
int64_t orc_species_num_panels(const orc_species* s);
/* Copies out the point arrays (any pointer may be NULL). */
void orc_species_get_points(const orc_species* s, double* xs, double* ps,
                            double* fs, double* q_ws);
/* Overwrites xs/ps (the advected positions, same count). */
int orc_species_set_points(orc_species* s, const double* xs, const double* ps);
/* Panel records, ORC_PANEL_INTS int32 per panel:
 * [panel_ind, pt0..pt8, level, parent, which_child, left, top, right, bottom,
 *  flags(bit0 left_bdry, bit1 right_bdry, bit2 needs_refinement,
 *        bit3 refined_xp, bit4 refined_p), child_inds_start, 0] */
#define ORC_PANEL_INTS 20
void orc_species_get_panels(const orc_species* s, int32_t* rec);
int64_t orc_species_num_leaves(const orc_species* s);
void orc_species_get_leaves(const orc_species* s, int32_t* leaf_inds);
double orc_species_f_beyond_boundary(const orc_species* s);
/* AMRStructure::remesh (AMRStructure_amr.cpp:1156-1202). Returns 0, or -1 if
 * the reference would have thrown ("too many shifts!"). */
int orc_species_remesh(orc_species* s);
/* Freeze the current mesh as the "old" mesh (AMRStructure_amr.cpp:1167-1175)
 * and evaluate interpolate_from_mesh (AMRStructure_interp.cpp:1094-1143) at
 * n points. */
int orc_species_copy_to_old(orc_species* s);
int orc_species_interp_from_mesh(orc_species* s, const double* x,
                                 const double* p, int64_t n, double* f_out);
/* Overwrites fs (same count) -- test input for the interpolation pins. */
int orc_species_set_fs(orc_species* s, const double* fs);
/* The leaf of the old mesh that interpolate_from_mesh (AMRStructure_interp.cpp:
 * 1094-1143) interpolates from at each point: periodic pre-shift, root
 * descent, 0 beyond a p boundary. */
int orc_species_descend_points(orc_species* s, const double* x, const double* p,
                               int64_t n, int32_t* leaf_out);
/* interpolate_to_initial_xps (AMRStructure_interp.cpp:493-793) of the old
 * mesh onto arbitrary base-grid points (n = nx*np); also returns the leaf
 * each sorted point was assigned (0 = beyond boundary) in sorted order and
 * the sort permutation. */
int orc_species_interp_to_grid(orc_species* s, const double* xs,
                               const double* ps, int nx, int np, double* f_out,
                               int32_t* leaf_sorted, int32_t* sort_perm);

/* ---- whole simulation -------------------------------------------------- */

typedef struct orc_sim orc_sim;
orc_sim* orc_sim_create(const orc_sim_params* sim,
                        const orc_species_params* sp, int n_species);
void orc_sim_destroy(orc_sim* s);
/* AMRSimulation::step (src/simulation_files/AMRSimulation_step.cpp:5-99). */
int orc_sim_step(orc_sim* s);
int orc_sim_num_species(const orc_sim* s);
orc_species* orc_sim_species(orc_sim* s, int i);
/* Per-species E at the species' points (after the step's last evaluation). */
void orc_sim_get_es(orc_sim* s, int i, double* es);
/* Per-species diagnostics (standard_interface/diagnostic_plot.py:116-131):
 * out[0]=charge, [1]=momentum, [2]=kinetic energy, [3]=electric energy. */
void orc_sim_diagnostics(orc_sim* s, int i, double out[4]);
/* The whole per-species set of standard_interface/diagnostic_plot.py:116-143
 * (ORC_DIAG_COUNT values): [0] charge sum q_w, [1] momentum sum p q_w / q,
 * [2] kinetic energy sum p^2 q_w / (2 q/m), [3] electric energy
 * 1/2 sum dx_k E_k^2 over the unique x in [0, L) (first occurrence, periodic
 * midpoint spacing), [4] entropy sum_{f>0} f log10(f) q_w / q, [5] L1
 * sum |q_w| / |q|, [6] L2 sum q_w f / q, [7] min f, [8] max f, [9] negative
 * area sum_{f<0} q_w / q, [10] number of f < 0, [11] their fraction,
 * [12] negative mass |sum_{f<0} q_w| / sum |q_w|, [13] number of points.
 * orc_diagnostics is the same on plain arrays. */
#define ORC_DIAG_COUNT 14
void orc_sim_diagnostics_full(orc_sim* s, int i, double* out);
void orc_diagnostics(int64_t n, const double* xs, const double* ps, const double* fs,
                     const double* q_ws, const double* es, double q, double qm, double L,
                     double* out);
int orc_sim_iter(const orc_sim* s);

#ifdef __cplusplus
}
#endif
#endif
