// Test infrastructure only: internal C++ types shared by the oracle's TUs.
#ifndef FS_ORACLE_INTERNAL_H
#define FS_ORACLE_INTERNAL_H

#include <cstdint>
#include <functional>
#include <vector>

#include "fs_oracle.h"

namespace orc {

// inc/Panel.hpp:41-56. Point layout (Panel.hpp:57-75):
//   2 5 8
//   1 4 7      children [1][3] / [0][2]; a p-only refined panel has
//   0 3 6      children [0] (bottom) and [1] (top).
struct Panel {
    int panel_ind = 0;
    int pt[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
    int level = 0;
    int parent = -1;
    int which_child = -1;
    int left = 0, top = -2, right = 0, bottom = -2;
    bool is_left_bdry = false, is_right_bdry = false;
    bool needs_refinement = false;
    bool refined_xp = false, refined_p = false;
    int child_start = -1;
};

// f0 of AMRSimulation_load.cpp:139-224 / initial_distributions.cpp:12-152.
double f0_eval(const orc_species_params& sp, double kx, double x, double p);

struct Species {
    // AMRStructure.hpp:56-120, the members the hot path reads.
    double q = -1, qm = -1;
    double x_min = 0, x_max = 0, p_min = 0, p_max = 0, Lx = 0, Lp = 0;
    bool periodic = true;
    bool simpson = false;
    int initial_height = 0, p_height = 0, max_height = 0;
    int npanels_x = 0, npanels_p = 0;
    double initial_dx = 0, initial_dp = 0;
    bool adaptive = false;
    std::vector<double> amr_eps;
    int minimum_unrefined_index = 0;
    bool need_further_refinement = false;
    double f_beyond_boundary = 0;
    orc_species_params params{};
    double kx = 0;

    std::vector<Panel> panels, old_panels;
    std::vector<int> leaf_inds;
    std::vector<double> xs, ps, fs, q_ws, es;
    std::vector<double> old_xs, old_ps, old_fs;

    // reduced-x utilities (AMRStructure.cpp:147-196)
    std::vector<double> reduced_xs, reduced_ws;
    std::vector<size_t> inv_inds_reduced_xs;

    void init(const orc_sim_params& sim, const orc_species_params& sp);

    // AMRStructure_amr.cpp
    void create_prerefined_mesh();
    void refine_panels(const std::function<double(double, double)>& f);
    void refine_panels_refine_v(const std::function<double(double, double)>& f);
    void generate_mesh(const std::function<double(double, double)>& f,
                       bool adaptive_now, bool initial);
    void test_panel(int panel_ind);
    void set_leaves_weights();
    void remesh();

    // AMRStructure_interp.cpp
    void shift_xs(std::vector<double>& shifted, const std::vector<double>& x,
                  const std::vector<double>& p) const;
    int descend(double& x, double p, bool& beyond, int panel_ind) const;
    int walk(double& tx, double tp, bool& beyond, int leaf_ind) const;
    void interpolate_to_initial_xps(std::vector<double>& f_out,
                                    const std::vector<double>& x,
                                    const std::vector<double>& p, int nx,
                                    int np, std::vector<int>* leaf_sorted,
                                    std::vector<int>* perm) const;
    void interpolate_from_panel_to_points(std::vector<double>& values,
                                          const std::vector<double>& x,
                                          const std::vector<double>& p,
                                          const std::vector<int>& idx,
                                          int panel_ind) const;
    double interpolate_from_panel(double x, double p, int panel_ind) const;
    double interpolate_from_mesh(double x, double p) const;

    void get_reduced_xs_ws();
    void get_reduced_es(const double* reduced_es);
};

// Solves the 9x9 monomial system of AMRStructure_interp.cpp:923-935 by LU
// with partial pivoting (row of the first largest |entry|).
void solve9(const double A[81], const double b[9], double c[9]);

// Direct sum / treecode used by the simulation (same as the C API).
void direct_sum(double* es, const double* t, int64_t nt, const double* s,
                const double* w, int64_t ns, double L, double eps);
// false when more than max_source points share a coordinate
bool treecode(double* es, const double* s, const double* w, int64_t ns,
              double L, double eps, double mac, int degree, int max_source);
void direct_sum_open(double* es, const double* t, int64_t nt, const double* s,
                     const double* w, int64_t ns, double eps);
double external_amplitude(int type, double t, double Am, double k);
void external_add(double* es, const double* targets, int64_t nt, double t, int type, double Am,
                  double k, double omega);
void external_add(double* es, const double* x, int64_t nt, double t, int type, double Am,
                  double k, double omega);

}  // namespace orc

#endif
