// Remesh interpolation of f onto the new panel mesh (gfx950).
//
// Replaces the numerical core of AMRStructure::remesh (reference
// src/structure_files/AMRStructure_amr.cpp:1156-1202): interpolate_to_initial_xps
// (AMRStructure_interp.cpp:493-793) and interpolate_from_mesh
// (AMRStructure_interp.cpp:1094-1143). Kernels:
//   coef     one thread per old panel: the 9x9 monomial system of
//            AMRStructure_interp.cpp:923-935 solved by LU with partial
//            pivoting (the matrix in registers: every index static); stores
//            the 9 coefficients and the panel centre.
//   records  one thread per old panel: packs what one leaf-search step reads
//            (advected corners, the neighbour to step to in each direction,
//            boundary flags) into one walk record, and for refined panels
//            the root-descent record (all 9 points, the children, which
//            step targets are refined), so every search step is a single
//            dependent load (a descent ends on the leaf's walk record).
//   shift    one thread per new point: the periodic pre-shift of
//            AMRStructure_interp.cpp:429-491.
//   rowsort  one workgroup per grid row: stable sort of the row by shifted x
//            (a merge of two ascending runs by binary search, else a bitonic
//            network in LDS).
//   column   one workgroup: the walk down the first column (:610-650).
//   rows     one workgroup per row: the walk along the row (:696-722).
//            The reference walks each line sequentially, every point
//            starting from its predecessor's leaf. Here every point first
//            gets a guess (in a row every 4th a root descent and a walk, the
//            others a walk from that one's guess), then walks again from its
//            predecessor's guess, all in parallel; where every guess up
//            to a point equals the leaf the walk from the predecessor's guess
//            returned, that walk started from the true predecessor leaf and
//            is exactly the reference's. From the first mismatch on the line
//            is walked sequentially as in the reference.
//   eval     one thread per point: the interpolant (:943-962).
//   points   one thread per point: shift, root descent, single-point
//            interpolant (interpolate_from_mesh, for AMR refinement).
// The leaf search keeps the reference's tie rules, step order and periodic
// shifts; floating-point contraction is off in this file, so every product,
// comparison and sum rounds exactly as in the reference's expressions (and
// bit-for-bit like this repository's CPU oracle).
#pragma clang fp contract(off)

#include "fs_remesh.h"

namespace fs {
namespace remesh {

// Panel record layout (int32 x kPanelInts): see fs_remesh.h.
constexpr int kLineThreads = 256;
constexpr int kLinePerThread = kMaxLine / kLineThreads;
constexpr int kMaxWalkSteps = 4096;  // a walk's steps (the reference's has no bound)
constexpr int kMaxHistory = 16;      // visited panels kept in registers
constexpr int kMaxDescend = 256;
// line elements per root descent in the guess phase of a row (the first
// column, one workgroup for the whole line, descends for every element)
constexpr int kRowAnchor = 4;

// Workgroups go to the chip's 8 XCDs (each with its own L2) round-robin by
// workgroup id. With the grid's x size a multiple of 8 (the launches of the
// row walks and the grid evaluation pad it), block x of a job runs on XCD
// x % 8; this maps a job's n blocks so that each XCD takes a contiguous
// range of them -- neighbouring rows and points, which read the same walk
// and coefficient records, share an L2 -- and the XCDs stay balanced within
// every job. -1 for the padding blocks (the whole workgroup returns). (The
// per-panel kernels ran slower mapped this way: their neighbour reads are
// not local in the panel order.)
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int x, int n) {
    const int xcd = x % kXcds, k = x / kXcds;
    const int q = n / kXcds, r = n % kXcds;
    if (k >= q + (xcd < r ? 1 : 0)) return -1;
    return xcd * q + min(xcd, r) + k;
}

__global__ __launch_bounds__(64) void coef_kernel(const Job* __restrict__ jobs) {
    const Job& J = jobs[blockIdx.y];
    if (J.on && *J.on == 0) return;
    const int pi = blockIdx.x * 64 + threadIdx.x;
    if (pi >= J.npanels) return;
    const int32_t* __restrict__ rec = J.rec;
    const double* __restrict__ ox = J.ox;
    const double* __restrict__ op = J.op;
    const double* __restrict__ of = J.of;
    CoefRec* __restrict__ coef = J.coef;
    const int32_t* r = rec + (size_t)pi * kPanelInts;
    // the 9x9 system in registers: every index below is static (loops fully
    // unrolled), the pivot row swap a select per candidate row
    double px[9], pp[9], bb[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int q = r[1 + i];
        px[i] = ox[q];
        pp[i] = op[q];
        bb[i] = of[q];
    }
    double a[9][9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const double dx = px[i] - px[4], dp = pp[i] - pp[4];
        a[i][0] = 1;
        a[i][1] = dx;
        a[i][2] = dx * dp;
        a[i][3] = dp;
        a[i][4] = dx * dx;
        a[i][5] = dx * dx * dp;
        a[i][6] = dx * dx * dp * dp;
        a[i][7] = dx * dp * dp;
        a[i][8] = dp * dp;
    }
    int perm[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) perm[i] = i;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        // partial pivoting: the first row of largest |a[i][k]|, i >= k
        int piv = k;
        double best = fabs(a[k][k]);
#pragma unroll
        for (int i = k + 1; i < 9; ++i) {
            const double v = fabs(a[i][k]);
            if (v > best) {
                best = v;
                piv = i;
            }
        }
#pragma unroll
        for (int i = k + 1; i < 9; ++i) {
            const bool sw = (i == piv);
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const double t = a[k][j];
                a[k][j] = sw ? a[i][j] : t;
                a[i][j] = sw ? t : a[i][j];
            }
            const int tp = perm[k];
            perm[k] = sw ? perm[i] : tp;
            perm[i] = sw ? tp : perm[i];
        }
        const double d = a[k][k];
        if (d != 0.0) {
#pragma unroll
            for (int i = k + 1; i < 9; ++i) a[i][k] = a[i][k] / d;
        }
#pragma unroll
        for (int i = k + 1; i < 9; ++i) {
            const double l = a[i][k];
#pragma unroll
            for (int j = k + 1; j < 9; ++j) a[i][j] = a[i][j] - l * a[k][j];
        }
    }
    double y[9], c[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        double b = 0.0;
#pragma unroll
        for (int m = 0; m < 9; ++m)
            if (perm[i] == m) b = bb[m];
        double s = b;
#pragma unroll
        for (int j = 0; j < i; ++j) s = s - a[i][j] * y[j];
        y[i] = s;
    }
#pragma unroll
    for (int i = 8; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int j = i + 1; j < 9; ++j) s = s - a[i][j] * c[j];
        c[i] = s / a[i][i];
    }
    CoefRec out;
#pragma unroll
    for (int i = 0; i < 9; ++i) out.c[i] = c[i];
    out.x4 = px[4];
    out.p4 = pp[4];
    coef[pi] = out;
}

__global__ __launch_bounds__(256) void records_kernel(const Job* __restrict__ jobs) {
    const Job& J = jobs[blockIdx.y];
    if (J.on && *J.on == 0) return;
    const int pi = blockIdx.x * 256 + threadIdx.x;
    if (pi >= J.npanels) return;
    const int32_t* __restrict__ rec = J.rec;
    const double* __restrict__ ox = J.ox;
    const double* __restrict__ op = J.op;
    WalkRec* __restrict__ walk = J.walk;
    DescRec* __restrict__ desc = J.desc;
    const int32_t* r = rec + (size_t)pi * kPanelInts;
    const int flags = r[17];
    const int nb[4] = {r[15], r[14], r[16], r[13]};  // right, top, bottom, left
    const int parent = r[11];
    WalkRec w;
    w.xbl = ox[r[1 + 0]];
    w.pbl = op[r[1 + 0]];
    w.xtl = ox[r[1 + 2]];
    w.ptl = op[r[1 + 2]];
    w.xmid = ox[r[1 + 4]];
    w.xbr = ox[r[1 + 6]];
    w.pbr = op[r[1 + 6]];
    w.xtr = ox[r[1 + 8]];
    w.ptr = op[r[1 + 8]];
    int wf = 0;
    // child offset entered when stepping into a refined (x and p) neighbour:
    // right -> child 0, top -> child 0, bottom -> child 1, left -> child 2
    // (AMRStructure_interp.cpp:91-92, :116-117, :141-142, :176-177)
    const int enter[4] = {0, 0, 1, 2};
    // parent's neighbour slot for the -1 ("use parent's") case
    const int pslot[4] = {15, 14, 16, 13};
    for (int d = 0; d < 4; ++d) {
        int nxt = -3;
        if (nb[d] == -2) {
            wf |= 1 << d;
        } else if (nb[d] == -1) {
            nxt = (parent >= 0) ? rec[(size_t)parent * kPanelInts + pslot[d]] : -3;
        } else {
            const int32_t* q = rec + (size_t)nb[d] * kPanelInts;
            nxt = (q[17] & kFlagRefinedXP) ? q[18] + enter[d] : nb[d];
        }
        w.next[d] = nxt;
    }
    if (flags & kFlagLeftBdry) wf |= kWalkLeftBdry;
    if (flags & kFlagRightBdry) wf |= kWalkRightBdry;
    const bool refined = (flags & (kFlagRefinedXP | kFlagRefinedP)) != 0;
    if (refined) wf |= kWalkRefined;
    w.flags = wf;
    w.pad[0] = w.pad[1] = w.pad[2] = 0;
    walk[pi] = w;
    if (!refined) return;  // a leaf: the descent tests its WalkRec

    auto is_refined = [&](int q) {
        return q >= 0 && (rec[(size_t)q * kPanelInts + 17] & (kFlagRefinedXP | kFlagRefinedP)) != 0;
    };
    DescRec dr;
    for (int i = 0; i < 9; ++i) {
        dr.x[i] = ox[r[1 + i]];
        dr.p[i] = op[r[1 + i]];
    }
    const int cs = r[18];
    dr.child_start = cs;
    int df = 0;
    if (flags & kFlagRefinedXP) df |= kDescRefinedXP;
    if (flags & kFlagRefinedP) df |= kDescRefinedP;
    if (flags & kFlagLeftBdry) df |= kDescLeftBdry;
    if (flags & kFlagRightBdry) df |= kDescRightBdry;
    for (int d = 0; d < 4; ++d)
        if (nb[d] == -2) df |= kDescNbrBoundary << d;
    dr.flags = df;
    for (int i = 0; i < 8; ++i) dr.cn[i] = 0;
    if (flags & (kFlagRefinedXP | kFlagRefinedP)) {
        const bool ponly = (flags & kFlagRefinedP) != 0;
        const int32_t* c0 = rec + (size_t)cs * kPanelInts;
        const int32_t* c1 = rec + (size_t)(cs + 1) * kPanelInts;
        const int32_t* c2 = rec + (size_t)(ponly ? cs : cs + 2) * kPanelInts;
        const int32_t* c3 = rec + (size_t)(ponly ? cs + 1 : cs + 3) * kPanelInts;
        dr.cn[0] = c1[14];  // child 1 top
        dr.cn[1] = c1[13];  // child 1 left
        dr.cn[2] = c3[14];  // child 3 top
        dr.cn[3] = c3[15];  // child 3 right
        dr.cn[4] = c0[16];  // child 0 bottom
        dr.cn[5] = c0[13];  // child 0 left
        dr.cn[6] = c2[16];  // child 2 bottom
        dr.cn[7] = c2[15];  // child 2 right
        const int tgt[4] = {cs, cs + 1, ponly ? cs : cs + 2, ponly ? cs + 1 : cs + 3};
        for (int t = 0; t < 4; ++t)
            if (is_refined(tgt[t])) df |= kDescTargetRefined << t;
        for (int i = 0; i < 8; ++i)
            if (is_refined(dr.cn[i])) df |= kDescTargetRefined << (4 + i);
        dr.flags = df;
    }
    desc[pi] = dr;
}

// shift_xs for one point (AMRStructure_interp.cpp:447-490). Returns false
// when the reference would throw "too many shifts!".
__device__ bool shift_one(double& xt, double p, const RootCorners& rc, double Lx) {
    bool left_ok = (rc.xtl - rc.xbl) * (p - rc.pbl) <= (rc.ptl - rc.pbl) * (xt - rc.xbl);
    int cnt = 0;
    while (!left_ok) {
        xt += Lx;
        left_ok = (rc.xtl - rc.xbl) * (p - rc.pbl) <= (rc.ptl - rc.pbl) * (xt - rc.xbl);
        if (++cnt > 10) return false;
    }
    bool right_ok = (rc.xtr - rc.xbr) * (p - rc.pbr) > (rc.ptr - rc.pbr) * (xt - (rc.xbr));
    cnt = 0;
    while (!right_ok) {
        xt -= Lx;
        right_ok = (rc.xtr - rc.xbr) * (p - rc.pbr) > (rc.ptr - rc.pbr) * (xt - (rc.xbr));
        if (++cnt > 10) return false;
    }
    return true;
}

// find_leaf_containing_point_from_neighbor (AMRStructure_interp.cpp:12-230)
// as a loop over precomputed walk records. Returns the leaf (0 when the walk
// reaches the root sentinel). The walk stops at the first panel it visits a
// second time, as the reference's history set makes it: the first
// kMaxHistory visited panels are kept in registers, later ones in ext (LDS,
// kMaxWalkSteps entries) when given. err = 1 on a runaway walk; err = 2 when
// the history outgrew the registers and no ext was given (the caller redoes
// that walk with one).
__device__ int walk_one(const WalkRec* __restrict__ W, double& tx, double tp, bool& beyond,
                        int start, double Lx, bool periodic, int& err, int* ext) {
    int hist[kMaxHistory];
    int nh = 1;
    hist[0] = start;
    int cur = start;
    for (int it = 0; it < kMaxWalkSteps; ++it) {
        if (cur == 0) return 0;
        if (cur < 0) {
            err = 1;
            return 0;
        }
        const WalkRec w = W[cur];
        if (tx - w.xmid >= Lx / 2 && periodic) tx -= Lx;
        if (tx - w.xmid < -Lx / 2 && periodic) tx += Lx;
        const bool r = (w.xtr - w.xbr) * (tp - w.pbr) > (w.ptr - w.pbr) * (tx - (w.xbr));
        const bool l = (w.xtl - w.xbl) * (tp - w.pbl) <= (w.ptl - w.pbl) * (tx - w.xbl);
        const bool t = (w.xtr - w.xtl) * (tp - w.ptl) < (w.ptr - w.ptl) * (tx - w.xtl);
        const bool b = (w.xbr - w.xbl) * (tp - w.pbl) >= (w.pbr - w.pbl) * (tx - w.xbl);
        const int f = w.flags;
        int nxt = cur;
        if (!r && !(f & kWalkBoundaryRight)) {
            if ((f & kWalkRightBdry) && periodic) tx -= Lx;
            nxt = w.next[0];
        } else if (!t && !(f & kWalkBoundaryTop)) {
            nxt = w.next[1];
        } else if (!b && !(f & kWalkBoundaryBottom)) {
            nxt = w.next[2];
        } else if (!l && !(f & kWalkBoundaryLeft)) {
            if ((f & kWalkLeftBdry) && periodic) tx += Lx;
            nxt = w.next[3];
        }
        if ((!r && (f & kWalkBoundaryRight)) || (!l && (f & kWalkBoundaryLeft)) ||
            (!t && (f & kWalkBoundaryTop)) || (!b && (f & kWalkBoundaryBottom)))
            beyond = true;
        const int nreg = nh < kMaxHistory ? nh : kMaxHistory;
        for (int k = 0; k < nreg; ++k)
            if (hist[k] == nxt) return nxt;
        if (nh > kMaxHistory)
            for (int k = 0; k < nh - kMaxHistory; ++k)
                if (ext[k] == nxt) return nxt;
        if (nh < kMaxHistory) {
            hist[nh] = nxt;
        } else if (ext) {
            ext[nh - kMaxHistory] = nxt;
        } else {
            err = 2;
            return nxt;
        }
        ++nh;
        cur = nxt;
    }
    err = 1;
    return cur;
}

// find_leaf_containing_xp_recursively (AMRStructure_interp.cpp:232-426).
// Refined panels step through their DescRec; the leaf reached is tested with
// its WalkRec (whose corners and boundary bits are the DescRec's).
__device__ int descend_one(const DescRec* __restrict__ D, const WalkRec* __restrict__ W,
                           double& x, double p, bool& beyond, int pi, double Lx, bool periodic,
                           int& err) {
    bool refined = pi >= 0 && (W[pi].flags & kWalkRefined) != 0;
    for (int it = 0; it < kMaxDescend; ++it) {
        if (pi < 0) {
            err = 1;
            return 0;
        }
        if (!refined) {
            const WalkRec& w = W[pi];
            const int f = w.flags;
            const bool r = (w.xtr - w.xbr) * (p - w.pbr) > (w.ptr - w.pbr) * (x - (w.xbr));
            const bool l = (w.xtl - w.xbl) * (p - w.pbl) <= (w.ptl - w.pbl) * (x - w.xbl);
            const bool t = (w.xtr - w.xtl) * (p - w.ptl) < (w.ptr - w.ptl) * (x - w.xtl);
            const bool b = (w.xbr - w.xbl) * (p - w.pbl) >= (w.pbr - w.pbl) * (x - w.xbl);
            if ((!l && (f & kWalkBoundaryLeft)) || (!r && (f & kWalkBoundaryRight)) ||
                (!t && (f & kWalkBoundaryTop)) || (!b && (f & kWalkBoundaryBottom)))
                beyond = true;
            return pi;
        }
        const DescRec& d = D[pi];
        const int f = d.flags;
        const double* X = d.x;
        const double* P = d.p;
        const int cs = d.child_start;
        const bool ponly = (f & kDescRefinedP) != 0;
        const double x_bl = X[0], p_bl = P[0], x_ml = X[1], p_ml = P[1];
        const double x_tl = X[2], p_tl = P[2], x_bm = X[3], p_bm = P[3];
        const double x_mm = X[4], p_mm = P[4], x_tm = X[5], p_tm = P[5];
        const double x_br = X[6], p_br = P[6], x_mr = X[7], p_mr = P[7];
        const double x_tr = X[8], p_tr = P[8];
        const bool b1 = (x_mm - x_ml) * (p - p_ml) >= (p_mm - p_ml) * (x - x_ml);
        const bool r1 = (x_tm - x_mm) * (p - p_mm) >= (p_tm - p_mm) * (x - x_mm);
        const bool b3 = (x_mr - x_mm) * (p - p_mm) >= (p_mr - p_mm) * (x - x_mm);
        int sub, tgt;
        if (b1 && r1) {
            const bool t1 = (x_tm - x_tl) * (p - p_tl) <= (p_tm - p_tl) * (x - x_tl);
            if (t1 || d.cn[0] < 0) {
                const bool l1 = (x_tl - x_ml) * (p - p_ml) <= (p_tl - p_ml) * (x - x_ml);
                if (l1 || d.cn[1] < 0) {
                    sub = cs + 1;
                    tgt = 1;
                } else {
                    sub = d.cn[1];
                    tgt = 4 + 1;
                    if ((f & kDescLeftBdry) && periodic) x += Lx;
                }
            } else {
                sub = d.cn[0];
                tgt = 4 + 0;
            }
        } else if (b3 && !r1) {
            const bool t3 = (x_tr - x_tm) * (p - p_tm) <= (p_tr - p_tm) * (x - x_tm);
            if (t3 || d.cn[2] < 0) {
                const bool r3 = (x_tr - x_mr) * (p - p_mr) <= (p_tr - p_mr) * (x - x_mr);
                if (!r3 || d.cn[3] < 0) {
                    sub = ponly ? cs + 1 : cs + 3;
                    tgt = 3;
                } else {
                    sub = d.cn[3];
                    tgt = 4 + 3;
                    if ((f & kDescRightBdry) && periodic) x -= Lx;
                }
            } else {
                sub = d.cn[2];
                tgt = 4 + 2;
            }
        } else {
            const bool r0 = (x_mm - x_bm) * (p - p_bm) >= (p_mm - p_bm) * (x - x_bm);
            if (r0) {
                const bool b0 = (x_bm - x_bl) * (p - p_bl) >= (p_bm - p_bl) * (x - x_bl);
                if (b0 || d.cn[4] < 0) {
                    const bool l0 = (x_ml - x_bl) * (p - p_bl) <= (p_ml - p_bl) * (x - x_bl);
                    if (l0 || d.cn[5] < 0) {
                        sub = cs;
                        tgt = 0;
                    } else {
                        sub = d.cn[5];
                        tgt = 4 + 5;
                        if ((f & kDescLeftBdry) && periodic) x += Lx;
                    }
                } else {
                    sub = d.cn[4];
                    tgt = 4 + 4;
                }
            } else {
                const bool b2 = (x_br - x_bm) * (p - p_bm) >= (p_br - p_bm) * (x - x_bm);
                if (b2 || d.cn[6] < 0) {
                    const bool r2 = (x_mr - x_br) * (p - p_br) <= (p_mr - p_br) * (x - x_br);
                    if (!r2 || d.cn[7] < 0) {
                        sub = ponly ? cs : cs + 2;
                        tgt = 2;
                    } else {
                        sub = d.cn[7];
                        tgt = 4 + 7;
                        if ((f & kDescRightBdry) && periodic) x -= Lx;
                    }
                } else {
                    sub = d.cn[6];
                    tgt = 4 + 6;
                }
            }
        }
        refined = (f & (kDescTargetRefined << tgt)) != 0;
        pi = sub;
    }
    err = 1;
    return 0;
}

// The advected corners of the root panel are old points 0, 2, 6 and 8
// (AMRStructure_interp.cpp:436-439); read on the device so no host round trip
// is needed between the step and the remesh.
__device__ __forceinline__ RootCorners load_corners(const double* __restrict__ ox,
                                                    const double* __restrict__ op) {
    RootCorners rc;
    rc.xbl = ox[0];
    rc.pbl = op[0];
    rc.xtl = ox[2];
    rc.ptl = op[2];
    rc.xbr = ox[6];
    rc.pbr = op[6];
    rc.xtr = ox[8];
    rc.ptr = op[8];
    return rc;
}


__global__ __launch_bounds__(256) void shift_kernel(const Job* __restrict__ jobs) {
    const Job& J = jobs[blockIdx.y];
    if (J.on && *J.on == 0) return;
    const int n = J.nx * J.np;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double xt = J.x[i];
    if (J.periodic && !shift_one(xt, J.p[i], load_corners(J.ox, J.op), J.Lx))
        atomicOr(J.status, kStatusTooManyShifts);
    J.shifted[i] = xt;
}

// One workgroup per row: the row sorted by (shifted x, position in the row
// list), a stable sort by shifted x. A grid row is ascending in x and the
// periodic pre-shift moves a run at one end by a period, so a row is nearly
// always at most two ascending runs: then each point's rank is its index in
// its run plus a binary search in the other (points of the first run before
// equal keys of the second). Other rows take a bitonic network over the row
// padded to a power of two (+inf keys past the row sort last). Rows are
// mapped to XCDs like the row walks: neighbouring rows read the same lines of
// the strided point arrays, and each XCD's L2 holds the rows it walks next.
__global__ __launch_bounds__(256) void rowsort_kernel(const Job* __restrict__ jobs) {
    __shared__ double keys[kMaxLine];
    __shared__ int pos[kMaxLine];
    __shared__ int s_desc, s_at;
    const Job& J = jobs[blockIdx.y];
    if (J.on && *J.on == 0) return;
    const int row = xcd_block(blockIdx.x, J.np);
    if (row < 0) return;
    const int nx = J.nx;
    const int32_t* rp = J.rows + (size_t)row * nx;
    if (threadIdx.x == 0) {
        s_desc = 0;
        s_at = nx;
    }
    for (int k = threadIdx.x; k < nx; k += blockDim.x) keys[k] = J.shifted[rp[k]];
    __syncthreads();
    for (int k = threadIdx.x + 1; k < nx; k += blockDim.x)
        if (keys[k] < keys[k - 1]) {
            atomicAdd(&s_desc, 1);
            atomicMin(&s_at, k);
        }
    __syncthreads();
    const size_t ob = (size_t)row * nx;
    if (s_desc <= 1) {
        const int d = s_at;  // runs [0, d) and [d, nx)
        for (int k = threadIdx.x; k < nx; k += blockDim.x) {
            const double kk = keys[k];
            int lo, hi, rank;
            if (k < d) {
                lo = d;  // first of the second run with key >= kk
                hi = nx;
                while (lo < hi) {
                    const int m = (lo + hi) >> 1;
                    if (keys[m] < kk) lo = m + 1;
                    else hi = m;
                }
                rank = k + (lo - d);
            } else {
                lo = 0;  // first of the first run with key > kk
                hi = d;
                while (lo < hi) {
                    const int m = (lo + hi) >> 1;
                    if (keys[m] <= kk) lo = m + 1;
                    else hi = m;
                }
                rank = (k - d) + lo;
            }
            const int32_t idx = rp[k];
            J.sorted_x[ob + rank] = kk;
            J.sorted_p[ob + rank] = J.p[idx];
            J.sorted_idx[ob + rank] = idx;
        }
        return;
    }
    int n2 = 1;
    while (n2 < nx) n2 <<= 1;
    for (int k = threadIdx.x; k < n2; k += blockDim.x) {
        if (k >= nx) keys[k] = __builtin_inf();
        pos[k] = k;
    }
    __syncthreads();
    for (int size = 2; size <= n2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < n2 / 2; t += blockDim.x) {
                const int i = 2 * t - (t & (stride - 1)), j = i + stride;
                const double ki = keys[i], kj = keys[j];
                const int pi = pos[i], pj = pos[j];
                const bool greater = ki > kj || (ki == kj && pi > pj);
                if (greater == ((i & size) == 0)) {
                    keys[i] = kj;
                    keys[j] = ki;
                    pos[i] = pj;
                    pos[j] = pi;
                }
            }
            __syncthreads();
        }
    }
    for (int k = threadIdx.x; k < nx; k += blockDim.x) {
        const int32_t idx = rp[pos[k]];
        J.sorted_x[ob + k] = keys[k];
        J.sorted_p[ob + k] = J.p[idx];
        J.sorted_idx[ob + k] = idx;
    }
}

// The walk along one line of points, each starting from its predecessor's
// leaf (see the file comment). Line element k is sorted point base + k*stride.
// first_start: the exact leaf the walk of element 0 starts from; element 0's
// x may already have been moved by the root descent (x0_override).
// Results: leaf per element (walk result, also the next start), beyond flag,
// the element's final x.
struct LineOut {
    int leaf[kLinePerThread];
    double tx[kLinePerThread];
    bool beyond[kLinePerThread];
    bool failed[kLinePerThread];
};

__device__ void walk_line(const Job& J, size_t base, size_t stride, int len, int first_start,
                          bool element0_exact_from_start, int anchor, int* s_guess, int* s_leaf,
                          int* s_first, int* s_fail, int* ext, LineOut& out, int& err) {
    const bool periodic = J.periodic != 0;
    const int tid = threadIdx.x;
    // A: a guess per element (a failed or wrong guess only costs the
    // sequential fallback of phase C, so its error flag is dropped). Anchors,
    // every anchor-th element, descend from the root and walk; the elements
    // between walk from their anchor's guess (a few steps along the line
    // instead of a descent through every level). Element 0 of the first
    // column walks exactly from its given start.
    int guess_err = 0;
#pragma unroll
    for (int q = 0; q < kLinePerThread; ++q) {
        const int k = tid + q * kLineThreads;
        if (k >= len || k % anchor != 0) continue;
        double x = J.sorted_x[base + (size_t)k * stride];
        const double p = J.sorted_p[base + (size_t)k * stride];
        bool bb = false;
        if (k == 0 && element0_exact_from_start) {
            s_guess[0] = walk_one(J.walk, x, p, bb, first_start, J.Lx, periodic, err, ext);
            continue;
        }
        const int d = descend_one(J.desc, J.walk, x, p, bb, 0, J.Lx, periodic, guess_err);
        bb = false;
        s_guess[k] = walk_one(J.walk, x, p, bb, d, J.Lx, periodic, guess_err, nullptr);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kLinePerThread; ++q) {
        const int k = tid + q * kLineThreads;
        if (k >= len || k % anchor == 0) continue;
        double x = J.sorted_x[base + (size_t)k * stride];
        const double p = J.sorted_p[base + (size_t)k * stride];
        bool bb = false;
        s_guess[k] = walk_one(J.walk, x, p, bb, s_guess[k - k % anchor], J.Lx, periodic,
                              guess_err, nullptr);
    }
    __syncthreads();
    // B: walk every element from its predecessor's guess (element 0: from the
    // exact start)
#pragma unroll
    for (int q = 0; q < kLinePerThread; ++q) {
        const int k = tid + q * kLineThreads;
        if (k >= len) continue;
        const int start = (k == 0) ? first_start : s_guess[k - 1];
        double tx = J.sorted_x[base + (size_t)k * stride];
        bool bb = false;
        int werr = 0;
        const int lf = walk_one(J.walk, tx, J.sorted_p[base + (size_t)k * stride], bb, start,
                                J.Lx, periodic, werr, nullptr);
        out.failed[q] = werr != 0;
        out.leaf[q] = lf;
        out.tx[q] = tx;
        out.beyond[q] = bb;
        s_leaf[k] = lf;
    }
    __syncthreads();
    // C: elements 0..m are exact, m = the first element whose exact walk
    // disagrees with the guess its successor started from, or the element
    // before the first walk that could not finish here (it is redone with the
    // full history); -1 when element 0 itself must be redone
    if (tid == 0) {
        *s_first = len;
        *s_fail = len;
    }
    __syncthreads();
    for (int k = tid; k < len - 1; k += kLineThreads)
        if (s_leaf[k] != s_guess[k]) atomicMin(s_first, k);
#pragma unroll
    for (int q = 0; q < kLinePerThread; ++q) {
        const int k = tid + q * kLineThreads;
        if (k < len && out.failed[q]) atomicMin(s_fail, k);
    }
    __syncthreads();
    if (tid == 0 && *s_fail < len) *s_first = min(*s_first, *s_fail - 1);
    __syncthreads();
}

__global__ __launch_bounds__(kLineThreads) void column_kernel(const Job* __restrict__ jobs) {
    __shared__ int s_guess[kMaxLine], s_leaf[kMaxLine], s_first, s_fail;
    __shared__ int ext[kMaxWalkSteps];
    const Job& J = jobs[blockIdx.y];
    if (J.on && *J.on == 0) return;
    const bool periodic = J.periodic != 0;
    const int np = J.np, nx = J.nx;
    int err = 0;
    // the exact start: root descent of sorted point 0 (it may move x0)
    __shared__ int s_start;
    if (threadIdx.x == 0) {
        double x0 = J.sorted_x[0];
        bool bb = false;
        s_start = descend_one(J.desc, J.walk, x0, J.sorted_p[0], bb, 0, J.Lx, periodic, err);
        J.sorted_x[0] = x0;
    }
    __syncthreads();
    LineOut out;
    walk_line(J, 0, (size_t)nx, np, s_start, true, 1, s_guess, s_leaf, &s_first, &s_fail, ext,
              out, err);
    const int m = s_first;  // rows 0..m exact
#pragma unroll
    for (int q = 0; q < kLinePerThread; ++q) {
        const int r = threadIdx.x + q * kLineThreads;
        if (r >= np || r > m) continue;
        const size_t k = (size_t)r * nx;
        J.sorted_x[k] = out.tx[q];
        J.first_leaf[r] = out.leaf[q];
        J.leaf_of[k] = out.beyond[q] ? 0 : out.leaf[q];
    }
    __syncthreads();
    if (threadIdx.x == 0 && m < np - 1) {
        int leaf = m >= 0 ? s_leaf[m] : s_start;
        for (int r = m + 1; r < np; ++r) {
            const size_t k = (size_t)r * nx;
            double tx = J.sorted_x[k];
            bool bb = false;
            leaf = walk_one(J.walk, tx, J.sorted_p[k], bb, leaf, J.Lx, periodic, err, ext);
            J.sorted_x[k] = tx;
            J.first_leaf[r] = leaf;
            J.leaf_of[k] = bb ? 0 : leaf;
        }
    }
    if (err) atomicOr(J.status, kStatusSearchFailed);
}

__global__ __launch_bounds__(kLineThreads) void rows_kernel(const Job* __restrict__ jobs) {
    __shared__ int s_guess[kMaxLine], s_leaf[kMaxLine], s_first, s_fail;
    __shared__ int ext[kMaxWalkSteps];
    const Job& J = jobs[blockIdx.y];
    if (J.on && *J.on == 0) return;
    const int r = xcd_block(blockIdx.x, J.np);
    if (r < 0) return;
    const bool periodic = J.periodic != 0;
    const int nx = J.nx;
    int err = 0;
    // elements 1..nx-1 of the row, starting from the first column's leaf
    const size_t base = (size_t)r * nx + 1;
    LineOut out;
    const int start = J.first_leaf[r];
    walk_line(J, base, 1, nx - 1, start, false, kRowAnchor, s_guess, s_leaf, &s_first, &s_fail,
              ext, out, err);
    const int m = s_first;
#pragma unroll
    for (int q = 0; q < kLinePerThread; ++q) {
        const int k = threadIdx.x + q * kLineThreads;
        if (k >= nx - 1 || k > m) continue;
        J.sorted_x[base + k] = out.tx[q];
        J.leaf_of[base + k] = out.beyond[q] ? 0 : out.leaf[q];
    }
    __syncthreads();
    if (threadIdx.x == 0 && m < nx - 2) {
        int leaf = m >= 0 ? s_leaf[m] : start;
        for (int k = m + 1; k < nx - 1; ++k) {
            double tx = J.sorted_x[base + k];
            bool bb = false;
            leaf = walk_one(J.walk, tx, J.sorted_p[base + k], bb, leaf, J.Lx, periodic, err, ext);
            J.sorted_x[base + k] = tx;
            J.leaf_of[base + k] = bb ? 0 : leaf;
        }
    }
    if (err) atomicOr(J.status, kStatusSearchFailed);
}

// The batched evaluation of AMRStructure_interp.cpp:943-962: row of
// monomials times coefficients, summed in order.
__global__ __launch_bounds__(256) void eval_grid_kernel(const Job* __restrict__ jobs) {
    const Job& J = jobs[blockIdx.y];
    if (J.on && *J.on == 0) return;
    const int n = J.nx * J.np;
    const int bx = xcd_block(blockIdx.x, (n + 255) / 256);
    if (bx < 0) return;
    const int k = bx * 256 + threadIdx.x;
    if (k >= n) return;
    const int leaf = J.leaf_of[k];
    double v;
    if (leaf == 0) {
        v = J.f_beyond;
    } else {
        const CoefRec& c = J.coef[leaf];
        const double dx = J.sorted_x[k] - c.x4, dp = J.sorted_p[k] - c.p4;
        const double dxsq = dx * dx, dpsq = dp * dp;
        const double m[9] = {1, dx, dx * dp, dp, dxsq, dxsq * dp, dxsq * dpsq, dx * dpsq, dpsq};
        v = 0.0;
#pragma unroll
        for (int j = 0; j < 9; ++j) v = v + m[j] * c.c[j];
    }
    J.f[J.sorted_idx[k]] = v;
}

// interpolate_from_mesh (AMRStructure_interp.cpp:1094-1143) at n = nx points.
__global__ __launch_bounds__(256) void points_kernel(const Job* __restrict__ jobs) {
    const Job& J = jobs[blockIdx.y];
    if (J.on && *J.on == 0) return;
    const int n = J.nx;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double xt = J.x[i];
    const double pt = J.p[i];
    if (J.periodic && !shift_one(xt, pt, load_corners(J.ox, J.op), J.Lx)) {
        atomicOr(J.status, kStatusTooManyShifts);
        J.f[i] = 0.0;
        return;
    }
    bool beyond = false;
    int err = 0;
    int leaf = descend_one(J.desc, J.walk, xt, pt, beyond, 0, J.Lx, J.periodic != 0, err);
    if (err) atomicOr(J.status, kStatusSearchFailed);
    if (beyond) leaf = 0;
    double v;
    if (leaf == 0) {
        v = J.f_beyond;
    } else {
        const CoefRec& c = J.coef[leaf];
        const double* cc = c.c;
        const double dx = xt - c.x4, dp = pt - c.p4;
        v = cc[0] + cc[1] * dx + cc[2] * dx * dp + cc[3] * dp + cc[4] * dx * dx +
            cc[5] * dx * dx * dp + cc[6] * dx * dx * dp * dp + cc[7] * dx * dp * dp +
            cc[8] * dp * dp;
    }
    J.f[i] = v;
}

__global__ void clear_status_kernel(const Job* __restrict__ jobs, int njobs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < njobs) *jobs[j].status = 0;
}

// ---------------------------------------------------------------- host side

size_t mesh_bytes(int npanels) {
    return align_up(sizeof(CoefRec) * npanels, 256) + align_up(sizeof(WalkRec) * npanels, 256) +
           align_up(sizeof(DescRec) * npanels, 256);
}

void carve_mesh(Job& j, void* base, int np_) {
    char* s = static_cast<char*>(base);
    j.coef = reinterpret_cast<CoefRec*>(s);
    s += align_up(sizeof(CoefRec) * np_, 256);
    j.walk = reinterpret_cast<WalkRec*>(s);
    s += align_up(sizeof(WalkRec) * np_, 256);
    j.desc = reinterpret_cast<DescRec*>(s);
}

size_t grid_bytes(int n, int np) {
    return align_up(sizeof(double) * n, 256) * 3 + align_up(sizeof(int32_t) * n, 256) * 2 +
           align_up(sizeof(int32_t) * np, 256) + 256;
}

void carve_grid(Job& j, void* base, int n, int np) {
    char* s = static_cast<char*>(base);
    j.shifted = reinterpret_cast<double*>(s);
    s += align_up(sizeof(double) * n, 256);
    j.sorted_x = reinterpret_cast<double*>(s);
    s += align_up(sizeof(double) * n, 256);
    j.sorted_p = reinterpret_cast<double*>(s);
    s += align_up(sizeof(double) * n, 256);
    j.sorted_idx = reinterpret_cast<int32_t*>(s);
    s += align_up(sizeof(int32_t) * n, 256);
    j.leaf_of = reinterpret_cast<int32_t*>(s);
    s += align_up(sizeof(int32_t) * n, 256);
    j.first_leaf = reinterpret_cast<int32_t*>(s);
    s += align_up(sizeof(int32_t) * np, 256);
    j.status = reinterpret_cast<int*>(s);
}

static unsigned cdiv(int a, int b) { return (unsigned)((a + b - 1) / b); }
// grid x sizes of the XCD-mapped kernels: a multiple of the XCD count
static unsigned pad8(unsigned a) { return (a + kXcds - 1) / kXcds * kXcds; }

int clear_status(const Job* d_jobs, const Launch& l, hipStream_t stream) {
    clear_status_kernel<<<cdiv(l.njobs, 64), 64, 0, stream>>>(d_jobs, l.njobs);
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

int prepare_old_meshes(const Job* d_jobs, const Launch& l, hipStream_t stream) {
    if (l.njobs <= 0 || l.max_panels <= 0) return fail(FS_E_INVALID, "remesh: empty old mesh");
    coef_kernel<<<dim3(cdiv(l.max_panels, 64), l.njobs), 64, 0, stream>>>(d_jobs);
    FS_HIP_CHECK(hipGetLastError());
    records_kernel<<<dim3(cdiv(l.max_panels, 256), l.njobs), 256, 0, stream>>>(d_jobs);
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

int interp_grids(const Job* d_jobs, const Launch& l, hipStream_t stream) {
    if (l.max_nx > kMaxLine || l.max_rows > kMaxLine)
        return fail(FS_E_INVALID, "remesh: grid lines longer than %d points", kMaxLine);
    const unsigned nj = (unsigned)l.njobs;
    shift_kernel<<<dim3(cdiv(l.max_points, 256), nj), 256, 0, stream>>>(d_jobs);
    rowsort_kernel<<<dim3(pad8((unsigned)l.max_rows), nj), 256, 0, stream>>>(d_jobs);
    column_kernel<<<dim3(1, nj), kLineThreads, 0, stream>>>(d_jobs);
    rows_kernel<<<dim3(pad8((unsigned)l.max_rows), nj), kLineThreads, 0, stream>>>(d_jobs);
    eval_grid_kernel<<<dim3(pad8(cdiv(l.max_points, 256)), nj), 256, 0, stream>>>(d_jobs);
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

int interp_points(const Job* d_jobs, const Launch& l, hipStream_t stream) {
    if (l.max_points <= 0) return FS_OK;
    points_kernel<<<dim3(cdiv(l.max_points, 256), (unsigned)l.njobs), 256, 0, stream>>>(d_jobs);
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

int status_to_error(int status) {
    if (status & kStatusTooManyShifts)
        return fail(FS_E_RANGE, "remesh: a point is more than 10 periods away (too many shifts!)");
    if (status & kStatusSearchFailed)
        return fail(FS_E_INTERNAL, "remesh: leaf search did not terminate");
    return FS_OK;
}

}  // namespace remesh
}  // namespace fs
