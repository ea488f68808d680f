// Device data layouts and launchers of the remesh interpolation
// (fs_remesh.hip). Internal to libfarrsight_hip.
#pragma once

#include "fs_internal.h"

namespace fs {
namespace remesh {

// A panel is kPanelInts int32 (the layout of the C ABI's panel records):
//   [0] panel_ind, [1..9] point indices (Panel.hpp:57-75 order), [10] level,
//   [11] parent, [12] which_child, [13] left, [14] top, [15] right,
//   [16] bottom (-1: use the parent's, -2: boundary), [17] flags,
//   [18] child_inds_start, [19] unused.
constexpr int kPanelInts = 20;
enum : int {
    kFlagLeftBdry = 1,
    kFlagRightBdry = 2,
    kFlagNeedsRefinement = 4,
    kFlagRefinedXP = 8,
    kFlagRefinedP = 16,
};

// Interpolant of one old panel: coefficients of
// (1, dx, dx dp, dp, dx^2, dx^2 dp, dx^2 dp^2, dx dp^2, dp^2) about point 4.
struct CoefRec {
    double c[9];
    double x4, p4;
    double pad;
};

// One neighbour-walk step's inputs: advected corners 0, 2, 6, 8 and the x of
// point 4; the panel to step to on the right, top, bottom and left.
enum : int {
    kWalkBoundaryRight = 1,
    kWalkBoundaryTop = 2,
    kWalkBoundaryBottom = 4,
    kWalkBoundaryLeft = 8,
    kWalkLeftBdry = 16,
    kWalkRightBdry = 32,
};
struct WalkRec {
    double xbl, pbl, xtl, ptl, xmid, xbr, pbr, xtr, ptr;
    int next[4];
    int flags;
    int pad[3];
};

// One root-descent step's inputs: all nine advected points, the children
// and the children's neighbours the descent may step to.
enum : int {
    kDescRefinedXP = 1,
    kDescRefinedP = 2,
    kDescLeftBdry = 4,
    kDescRightBdry = 8,
    kDescNbrBoundary = 16,  // << d, d = 0 right, 1 top, 2 bottom, 3 left
};
struct DescRec {
    double x[9], p[9];
    int child_start;
    int flags;
    int cn[8];  // child1 top/left, child3 top/right, child0 bottom/left, child2 bottom/right
};

// Advected corners of the root panel (old points 0, 2, 6, 8).
struct RootCorners {
    double xbl, pbl, xtl, ptl, xbr, pbr, xtr, ptr;
};

enum : int { kStatusTooManyShifts = 1, kStatusSearchFailed = 2 };

struct OldMesh {
    CoefRec* coef = nullptr;
    WalkRec* walk = nullptr;
    DescRec* desc = nullptr;
    const double* ox = nullptr;  // advected old points (root corners are 0, 2, 6, 8)
    const double* op = nullptr;
    int npanels = 0;
    static size_t bytes(int npanels);
    void carve(void* base, int npanels);
};

struct GridWork {
    double* shifted = nullptr;
    double* sorted_x = nullptr;
    double* sorted_p = nullptr;
    int32_t* sorted_idx = nullptr;
    int32_t* leaf_of = nullptr;
    int32_t* first_leaf = nullptr;
    int* status = nullptr;
    static size_t bytes(int n, int np);
    void carve(void* base, int n, int np);
};

// Builds coefficient, walk and descent records of an old mesh.
int prepare_old_mesh(const int32_t* d_rec, int npanels, const double* d_ox, const double* d_op,
                     const double* d_of, OldMesh& m, hipStream_t stream);
// interpolate_to_initial_xps: d_row_pts lists the np rows (ascending p) of nx
// point indices each. Writes d_f[0:nx*np]; w.status gets error bits.
int interp_to_grid(const OldMesh& m, const double* d_x,
                   const double* d_p, const int32_t* d_row_pts, int nx, int np, double Lx,
                   bool periodic, double f_beyond, double* d_f, GridWork& w, hipStream_t stream);
// interpolate_from_mesh at n points.
int interp_points(const OldMesh& m, const double* d_x, const double* d_p,
                  int n, double Lx, bool periodic, double f_beyond, double* d_f, int* d_status,
                  hipStream_t stream);
// Maps status bits to an FS_E_* code (and the error text).
int status_to_error(int status);

}  // namespace remesh
}  // namespace fs
