// Adaptive refinement of a species' panel tree on the device: the adaptive
// loop of generate_mesh (reference src/structure_files/AMRStructure_amr.cpp:
// 901-941) with refine_panels (:506-839), test_panel (:971-1081),
// set_leaves_weights (:1084-1151) and the reduced-x map of
// get_reduced_xs_ws (AMRStructure.cpp:147-181). The result is the tree the
// host builder (fs_mesh.cpp) makes, record for record and bit for bit.
// Internal to libfarrsight_hip.
#pragma once

#include <vector>

#include "fs_internal.h"
#include "fs_remesh.h"

namespace fs {
namespace amr {

// Device counters of one refining species.
struct Counters {
    int P;        // panels
    int V;        // points
    int flagged;  // panels in [mu, P) that still need refinement
    int nR;       // panels split by the last pass
    int newV;     // points added by the last pass
    int status;   // kStat* bits
    int P0;       // panels before the last pass
    int U;        // distinct x of the final tree
};

enum : int {
    kStatCapacity = 1,  // a pass outgrew the buffers sized for it
    kStatTopology = 2,  // a neighbour link the reference's walk would not follow
    kStatSlots = 4,     // a point on more leaves than the weight slots hold
    kStatDepth = 8,     // a parent chain longer than the tree can be deep
};

// Buffers of one species, kept from remesh to remesh (grow only). The tree
// (records and points) is also the species' current mesh between remeshes.
struct Work {
    DevBuf rec, x, p, f;       // tree: cap_p panels, cap_v points
    DevBuf chain, cnt, off;    // pass state by panel (kept when grown)
    DevBuf tmp;                // pass temporaries by panel
    DevBuf fin;                // finishing work space by point
    DevBuf out;                // weights and the reduced-x map
    int cap_p = 0, cap_v = 0;
    // blocks come from and go back to this pool when set (the buffers are
    // used on the pool's stream only)
    BlockPool* pool = nullptr;
    int ensure(DevBuf& b, size_t bytes, hipStream_t st) {
        return pool ? pool->ensure(b, bytes, st) : b.ensure(bytes, st);
    }
    void put(DevBuf& b) {
        if (pool) pool->put(b);
        else b.release();
    }
    void release(BlockPool* to) {
        for (DevBuf* b : {&rec, &x, &p, &f, &chain, &cnt, &off, &tmp, &fin, &out}) {
            if (to) to->put(*b);
            else b->release();
        }
        cap_p = cap_v = 0;
    }
};

// One refining species.
struct Input {
    // base mesh: panel records on the device (uploaded once) and its grid
    // points with f from the grid interpolation (nb = nx np points)
    const int32_t* base_rec;
    int base_panels;
    int mu;  // the base's first leaf (min_unrefined)
    const double* gx;
    const double* gp;
    const double* gf;
    int nb;
    // the old (advected) mesh the new points' f is interpolated from
    remesh::Job old;
    int max_height;
    bool periodic;
    std::vector<double> eps;  // amr_epsilons (at most 5)
    double q;
    bool simpson;
    Work* w;
};

// The refined tree: counts, and device views into the species' Work.
struct Output {
    int P = 0, V = 0, U = 0;
    const int32_t* rec = nullptr;
    const double *x = nullptr, *p = nullptr, *f = nullptr;
    const double* bw = nullptr;  // per-point quadrature weight (before * f)
    const double* ux = nullptr;  // distinct x ascending
    const int32_t* roff = nullptr;     // U + 1 offsets into members
    const int32_t* members = nullptr;  // point indices, by x then index
    const int32_t* uof = nullptr;      // distinct-x index of each point
};

// Refines every species of `in` (independent trees, one launch per kernel
// for all of them). One host synchronization per refinement pass and one at
// the end; the grid interpolation of the base points must be enqueued on
// `stream` before. On failure the species' Work holds no valid tree.
// jobs_buf holds the launches' job tables and counters, sort_buf the
// finishing sort of all species at once.
int refine(std::vector<Input>& in, std::vector<Output>& out, DevBuf& jobs_buf,
           DevBuf& sort_buf, hipStream_t stream);

}  // namespace amr
}  // namespace fs
