// Host-side panel tree of one species (the structure half of AMRStructure,
// reference src/structure_files/AMRStructure_amr.cpp). Panels live directly
// in the int32 record layout the device kernels and the C ABI use, so a
// mesh is uploaded with one copy. Internal to libfarrsight_hip.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

namespace fs {
namespace mesh {

constexpr int kInts = 20;
// record fields
enum Field : int {
    IND = 0, PT = 1, LEVEL = 10, PARENT = 11, WHICH = 12,
    LEFT = 13, TOP = 14, RIGHT = 15, BOTTOM = 16, FLAGS = 17, CHILD = 18,
};
enum Flag : int {
    LEFT_BDRY = 1, RIGHT_BDRY = 2, NEEDS_REF = 4, REFINED_XP = 8, REFINED_P = 16,
};

// f at a batch of new points (n values into f).
using BatchEval = std::function<void(const double* x, const double* p, size_t n, double* f)>;

struct Spec {
    double x_min = 0, x_max = 0, p_min = 0, p_max = 0;
    int initial_height = 1, p_height = 0, max_height = 1;
    bool periodic = true;
    std::vector<double> amr_eps;
};

class PanelTree {
public:
    std::vector<int32_t> rec;
    std::vector<double> xs, ps, fs;
    int min_unrefined = 0;
    bool need_further = false;

    int size() const { return (int)(rec.size() / kInts); }
    int32_t& at(int panel, int field) { return rec[(size_t)panel * kInts + field]; }
    int32_t at(int panel, int field) const { return rec[(size_t)panel * kInts + field]; }
    bool has(int panel, int flag) const { return (at(panel, FLAGS) & flag) != 0; }

    // create_prerefined_mesh (AMRStructure_amr.cpp:127-230): the uniform base
    // mesh; f of every point is 1 until the caller overwrites it.
    void build_base(const Spec& s);
    // refine_panels (:506-839): every flagged panel from min_unrefined on
    // splits in x and p; new points' f from one batched call.
    void refine_xp(const BatchEval& f);
    // refine_panels_refine_v (:233-504): flagged panels split in p only.
    void refine_p(const BatchEval& f);
    // test_panel (:971-1081)
    void test_panel(int panel, const Spec& s);
    // the adaptive loop of generate_mesh (:901-941)
    void refine_adaptively(const Spec& s, const BatchEval& f);
    // set_leaves_weights (:1084-1151) without the final *f: per-point
    // quadrature weight q dx dp {..}/4 or /9 accumulated in depth-first leaf
    // order; also returns the leaves in that order.
    void leaf_weights(double q, bool simpson, std::vector<double>& w,
                      std::vector<int32_t>& leaves) const;

private:
    void reserve_pass(int kids, int pts, std::vector<double>& nx, std::vector<double>& np);
    int append_panel(int parent, int which, const int (&pts)[9], int l, int t, int r, int b,
                     bool lb, bool rb);
    bool periodic_ = true;
};

// Rows of the base grid: point indices grouped by p ascending, nx per row.
std::vector<int32_t> grid_rows(const std::vector<double>& xs, const std::vector<double>& ps,
                               int nx, int np);

// The reduced-x map of AMRStructure::get_reduced_xs_ws (AMRStructure.cpp:
// 147-181): unique x values ascending, and for each the point indices in
// increasing index order (CSR: offsets[u]..offsets[u+1]).
struct ReducedMap {
    std::vector<double> ux;
    std::vector<int32_t> offsets;
    std::vector<int32_t> members;
    std::vector<int32_t> unique_of;  // per point
};
ReducedMap reduced_map(const std::vector<double>& xs);
// The same into m, reusing its vectors' storage.
void reduced_map(const std::vector<double>& xs, ReducedMap& m);

}  // namespace mesh
}  // namespace fs
