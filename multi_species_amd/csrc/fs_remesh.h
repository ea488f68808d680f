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

// Largest grid row / column handled by one workgroup of the search kernels.
constexpr int kMaxLine = 1024;

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
    kWalkRefined = 64,  // the panel is refined (the root descent goes on)
};
struct WalkRec {
    double xbl, pbl, xtl, ptl, xmid, xbr, pbr, xtr, ptr;
    int next[4];
    int flags;
    int pad[3];
};

// One root-descent step's inputs: all nine advected points, the children
// and the children's neighbours the descent may step to, and which of those
// are refined themselves. Written for refined panels only: a descent that
// reaches a leaf tests the leaf's WalkRec (the same corners and boundary
// flags), so leaves need no descent record.
enum : int {
    kDescRefinedXP = 1,
    kDescRefinedP = 2,
    kDescLeftBdry = 4,
    kDescRightBdry = 8,
    kDescNbrBoundary = 16,  // << d, d = 0 right, 1 top, 2 bottom, 3 left
    // << t: step target t is refined; t = 0..3 the children in the order
    // the descent names them (cs, cs+1, and cs+2, cs+3 or for a p-only split
    // cs, cs+1), t = 4..11 the panels cn[0..7]
    kDescTargetRefined = 256,
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

// One interpolation job: an old (advected) mesh and the points to
// interpolate onto. Grid jobs: np rows of nx points listed by `rows`; point
// jobs: n = nx points x, p (np = 1, rows unused). Every kernel of this module
// takes an array of jobs and works on job blockIdx.y, so all species (and
// all decks of a batch) share one launch per kernel.
struct Job {
    // old mesh
    const int32_t* rec;
    int npanels;
    const double* ox;
    const double* op;
    const double* of;
    CoefRec* coef;
    WalkRec* walk;
    DescRec* desc;
    // target points
    const double* x;
    const double* p;
    const int32_t* rows;
    int nx, np;
    double Lx;
    int periodic;
    double f_beyond;
    // work
    double* shifted;
    double* sorted_x;
    double* sorted_p;
    int32_t* sorted_idx;
    int32_t* leaf_of;
    int32_t* first_leaf;
    // output
    double* f;
    int* status;
    // when set, the job runs only while *on != 0 (a batch whose decks
    // remesh at different periods switches its species on and off here)
    const int* on;
};

// Sizes of the per-job work arrays.
size_t mesh_bytes(int npanels);
size_t grid_bytes(int n, int np);
// Points a job's mesh and grid work arrays into the given device blocks.
void carve_mesh(Job& j, void* base, int npanels);
void carve_grid(Job& j, void* base, int n, int np);

struct Launch {
    int njobs = 0;
    int max_panels = 0;
    int max_points = 0;  // max nx*np over jobs
    int max_rows = 0;    // max np
    int max_nx = 0;
};

// Coefficient, walk and descent records of every job's old mesh.
int prepare_old_meshes(const Job* d_jobs, const Launch& l, hipStream_t stream);
// interpolate_to_initial_xps for every (grid) job.
int interp_grids(const Job* d_jobs, const Launch& l, hipStream_t stream);
// interpolate_from_mesh for every (point) job.
int interp_points(const Job* d_jobs, const Launch& l, hipStream_t stream);
// Clears the status word of every job.
int clear_status(const Job* d_jobs, const Launch& l, hipStream_t stream);
// Maps status bits to an FS_E_* code (and the error text).
int status_to_error(int status);

}  // namespace remesh
}  // namespace fs
