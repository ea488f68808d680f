// Per-thread steps of the device refinement passes (fs_amr.hip): each
// function is one thread's work of one launch, written once for the device
// kernels and the host check that replays the launches in order
// (tests/amr_replay.cpp). Atomic updates go through `Ops` (device atomics,
// or plain updates where the replay runs threads one at a time). The
// reasoning behind the passes is at the top of fs_amr.hip. Internal to
// libfarrsight_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "fs_amr.h"
#include "fs_remesh.h"

#define FS_HD __host__ __device__ __forceinline__

namespace fs {
namespace amr {
namespace pass {

constexpr int kI = remesh::kPanelInts;
enum : int {
    PT = 1, LEVEL = 10, PARENT = 11, WHICH = 12, LEFT = 13, TOP = 14, RIGHT = 15, BOTTOM = 16,
    FLAGS = 17, CHILD = 18,
};
enum : int {
    LB = remesh::kFlagLeftBdry,
    RB = remesh::kFlagRightBdry,
    NEEDS = remesh::kFlagNeedsRefinement,
    RXP = remesh::kFlagRefinedXP,
    RP = remesh::kFlagRefinedP,
};
constexpr int kPlanInts = 12;  // member, flags, 4 side kinds, 4 side targets, 2 spare
constexpr int kSlots = 6;      // leaves whose nine points hold one point (<= 4)
constexpr int kMaxDepth = 32;  // levels a depth-first path key holds (2 bits each)
constexpr int kChainStack = 24;

// Side plans of a member (what the link step does on that side).
enum : int {
    kKindFresh = 0,      // fresh points, nothing to link
    kKindFlag = 1,       // fresh points; flag the target for the next pass
    kKindRewire = 2,     // neighbour split before the pass: link its children
    kKindShareNew = 3,   // neighbour split earlier in the pass: its children's points
    kKindFreshPts = 8,   // bit: this side's points are fresh
};

// One side of a panel being split (the host builder's table, with the child
// point slot the side's two edge points take; the opposite side is 3 - si).
struct SideDef {
    int field, ca, cb, na, nb, nfield, shared, slot, ax, ap, bx, bp, bdry;
};
// left, bottom, top, right: the order refine_panels numbers their points in
FS_HD SideDef side_def(int si) {
    switch (si) {
        case 0: return SideDef{LEFT, 0, 1, 2, 3, RIGHT, 7, 1, 0, 1, 0, 3, LB};
        case 1: return SideDef{BOTTOM, 0, 2, 1, 3, TOP, 5, 3, 1, 0, 3, 0, 0};
        case 2: return SideDef{TOP, 1, 3, 0, 2, BOTTOM, 3, 5, 1, 4, 3, 4, 0};
        default: return SideDef{RIGHT, 2, 3, 0, 1, LEFT, 1, 7, 4, 1, 4, 3, RB};
    }
}

struct Job {
    int32_t* rec;
    double* x;
    double* p;
    double* f;
    int cap_p, cap_v;
    int mu, max_height, periodic, n_eps;
    double eps[5];
    double q;
    int simpson;
    Counters* ctr;
    // pass state (panel indexed from mu; chain by panel index)
    int32_t* chain;
    int2* cnt;
    int2* off;
    // pass temporaries
    int32_t* plan;  // by rank
    int2* bsum;
    // finishing, by point
    int* nslot;
    uint64_t* spath;  // depth-first path key of each slot's leaf
    double* sval;
    uint64_t* key;
    uint64_t* skey;
    int32_t* perm;
    int2* vcnt;
    int2* voff;
    // outputs
    double* bw;
    double* ux;
    int32_t* roff;
    int32_t* members;
    int32_t* uof;
    remesh::Job* pj;  // this species' point job on its old mesh
};

FS_HD int2 add2(int2 a, int2 b) { return make_int2(a.x + b.x, a.y + b.y); }
FS_HD int2 sub2(int2 a, int2 b) { return make_int2(a.x - b.x, a.y - b.y); }

FS_HD bool in_r(const Job& J, int b, int P) {
    return b >= J.mu && b < P && ((J.rec[(size_t)b * kI + FLAGS] & NEEDS) || J.chain[b] != 0);
}

FS_HD bool split_before(const Job& J, int b) {
    return (J.rec[(size_t)b * kI + FLAGS] & RXP) != 0;
}

// Field of side si of panel a when the sequential pass reaches panel t: a -1
// becomes a child of the coarser neighbour G if G is split in the pass before
// t (G links the children of the panel it faces). New panels are numbered
// P + 4 rank. -4: a link the passes do not reproduce.
FS_HD int field_at(const Job& J, int a, int si, int t, int P) {
    const SideDef sd = side_def(si);
    const int32_t* ra = J.rec + (size_t)a * kI;
    const int v = ra[sd.field];
    if (v != -1) return v;
    const int par = ra[PARENT];
    if (par < 0) return -1;
    if (par >= P) return -4;
    const int g = J.rec[(size_t)par * kI + sd.field];
    if (g < 0 || !(in_r(J, g, P) && g < t)) return -1;
    // G's side facing a's parent is the opposite one; a is its child na or nb
    // there
    const SideDef so = side_def(3 - si);
    if (J.rec[(size_t)g * kI + so.field] != par) return -4;
    const int c = a - J.rec[(size_t)par * kI + CHILD];
    const int gbase = P + 4 * J.off[g - J.mu].x;
    return gbase + (c == so.na ? so.ca : so.cb);
}

// test_panel (AMRStructure_amr.cpp:971-1081), the host builder's arithmetic.
FS_HD bool needs_refinement(const Job& J, const int32_t* r) {
    double f9[9];
    for (int i = 0; i < 9; ++i) f9[i] = J.f[r[PT + i]];
    const int ne = J.n_eps;
    bool crit = false;
    if (ne > 0) {
        double mx = f9[0], mn = f9[0];
        for (int i = 1; i < 9; ++i) {
            mx = (mx < f9[i]) ? f9[i] : mx;
            mn = (mn > f9[i]) ? f9[i] : mn;
        }
        crit = crit || (mx - mn > J.eps[0]);
        if (ne > 1) crit = crit || (mx - mn) / f9[4] > J.eps[1];
        if (ne > 2) crit = crit || mx / mn > J.eps[2];
    }
    if (ne > 3) {
        const double dx = J.x[r[PT + 3]] - J.x[r[PT + 0]];
        double m = fabs((f9[3] - f9[0]) / dx);
        for (int i = 1; i < 6; ++i) {
            const double a = fabs((f9[3 + i] - f9[i]) / dx);
            if (m < a) m = a;
        }
        crit = crit || (m > J.eps[3]);
    }
    if (ne > 4) {
        const double dp = J.p[r[PT + 1]] - J.p[r[PT + 0]];
        const double a0 = fabs((f9[7] - f9[6]) / dp), a1 = fabs((f9[8] - f9[7]) / dp);
        double m = a0;
        if (m < a1) m = a1;
        if (m < 0.0) m = 0.0;
        crit = crit || (m > J.eps[4]);
    }
    return r[LEVEL] < J.max_height && crit;
}

// Panel k (>= mu): tests it if the last pass made it ([P0, P)); returns 1
// if it is flagged.
template <class Ops>
FS_HD int test_one(const Job& J, int k, const Ops& ops) {
    const int P = J.ctr->P, P0 = J.ctr->P0, V = J.ctr->V;
    if (k >= P) return 0;
    int32_t* r = J.rec + (size_t)k * kI;
    int fl = r[FLAGS];
    if (k >= P0 && !(fl & RXP)) {
        bool ok = true;
        for (int i = 0; i < 9; ++i) ok = ok && r[PT + i] >= 0 && r[PT + i] < V;
        if (!ok) {
            ops.or_(&J.ctr->status, kStatTopology);
        } else if (needs_refinement(J, r)) {
            fl |= NEEDS;
            r[FLAGS] = fl;
        }
    }
    return (fl & NEEDS) ? 1 : 0;
}

FS_HD void chain_clear_one(const Job& J, int k) {
    if (k < J.ctr->P) J.chain[k] = 0;
}

// R beyond the flagged panels: from flagged panel k, the parents'
// neighbours it and the panels it pulls in flag ahead of themselves.
template <class Ops>
FS_HD void chain_one(const Job& J, int k, const Ops& ops) {
    const int P = J.ctr->P;
    if (k >= P || !(J.rec[(size_t)k * kI + FLAGS] & NEEDS)) return;
    int stack[kChainStack];
    int sp = 0;
    stack[sp++] = k;
    while (sp > 0) {
        const int a = stack[--sp];
        const int32_t* ra = J.rec + (size_t)a * kI;
        const int par = ra[PARENT];
        for (int si = 0; si < 4; ++si) {
            const int fld = side_def(si).field;
            if (ra[fld] != -1 || par < 0 || par >= P) continue;
            const int q = J.rec[(size_t)par * kI + fld];
            if (q <= a || q >= P) continue;
            if ((J.rec[(size_t)q * kI + FLAGS] & (RXP | NEEDS)) || J.chain[q] != 0) continue;
            J.chain[q] = 1;
            if (sp < kChainStack)
                stack[sp++] = q;
            else
                ops.or_(&J.ctr->status, kStatDepth);
        }
    }
}

// Is side si of member jj fresh (two new points)? From the pass-start state.
FS_HD bool side_fresh(const Job& J, int jj, const int32_t* rj, int si, int P) {
    const SideDef sd = side_def(si);
    const int nb = rj[sd.field];
    if (nb < 0 || nb >= P) return true;
    const bool split = split_before(J, nb) || (in_r(J, nb, P) && nb < jj);
    if (!split) return true;
    return sd.bdry != 0 && (rj[FLAGS] & sd.bdry) && J.periodic;
}

// (member of R, new points) of panel k in [mu, P), for the scan.
FS_HD void count_one(const Job& J, int k) {
    const int P = J.ctr->P;
    if (k >= P) return;
    int2 c = make_int2(0, 0);
    if (in_r(J, k, P)) {
        const int32_t* rk = J.rec + (size_t)k * kI;
        int pts = 8;
        for (int si = 0; si < 4; ++si) pts += side_fresh(J, k, rk, si, P) ? 2 : 0;
        c = make_int2(1, pts);
    }
    J.cnt[k - J.mu] = c;
}

// Member jj: its new records and fresh points from the pass-start state, and
// its side plans for the link step.
template <class Ops>
FS_HD void build_one(const Job& J, int jj, const Ops& ops) {
    const int P = J.ctr->P, V = J.ctr->V;
    if (jj >= P || J.cnt[jj - J.mu].x == 0) return;
    const int2 o = J.off[jj - J.mu];
    const int base = P + 4 * o.x;
    int nv = V + o.y;
    int32_t* pl = J.plan + (size_t)o.x * kPlanInts;
    pl[0] = -1;  // (the link step skips a member whose build stopped)
    if (base + 4 > J.cap_p || nv + J.cnt[jj - J.mu].y > J.cap_v) {
        ops.or_(&J.ctr->status, kStatCapacity);
        return;
    }
    int32_t R[kI];
    for (int i = 0; i < kI; ++i) R[i] = J.rec[(size_t)jj * kI + i];
    bool ok = R[PARENT] >= 0 && R[PARENT] < P;
    for (int i = 0; i < 9; ++i) ok = ok && R[PT + i] >= 0 && R[PT + i] < V;
    if (!ok) {
        ops.or_(&J.ctr->status, kStatTopology);
        return;
    }
    const double x0 = J.x[R[PT + 0]], p0 = J.p[R[PT + 0]];
    const double hx = 0.5 * (J.x[R[PT + 3]] - x0), hp = 0.5 * (J.p[R[PT + 1]] - p0);
    double sx[5], sp[5];
    for (int i = 0; i < 5; ++i) {
        sx[i] = x0 + hx * i;
        sp[i] = p0 + hp * i;
    }
    int edge[4][2];
    int cn[4][4];
    for (int c = 0; c < 4; ++c)
        for (int s = 0; s < 4; ++s) cn[c][s] = -1;
    int kind[4], tgt[4];
    int status = 0;
    for (int si = 0; si < 4; ++si) {
        const SideDef sd = side_def(si);
        const int nb = field_at(J, jj, si, jj, P);
        kind[si] = kKindFresh | kKindFreshPts;
        tgt[si] = -1;
        bool fresh = true;
        if (nb < -2) {
            status |= kStatTopology;
        } else if (nb == -2) {
            cn[sd.ca][si] = cn[sd.cb][si] = -2;
        } else if (nb == -1) {
            // the parent's neighbour there, when the pass reaches jj
            const int q = field_at(J, R[PARENT], si, jj, P);
            if (q < 0) {
                status |= kStatTopology;
            } else if (q >= P) {
                kind[si] = kKindFlag | kKindFreshPts;  // a child made in this pass
                tgt[si] = q;
            } else if (!split_before(J, q) && !in_r(J, q, P)) {
                kind[si] = kKindFlag | kKindFreshPts;
                tgt[si] = q;
            }
        } else if (nb >= P) {
            // linked in this pass to a coarser neighbour's new child: unsplit
        } else {
            const bool before = split_before(J, nb);
            const bool nbr = in_r(J, nb, P);
            if (nbr && J.rec[(size_t)nb * kI + side_def(3 - si).field] != jj) status |= kStatTopology;
            if (before || (nbr && nb < jj)) {
                const int cb = before ? J.rec[(size_t)nb * kI + CHILD] : P + 4 * J.off[nb - J.mu].x;
                // (a split panel's children lie below P)
                const bool cb_ok = !before || (cb >= 0 && cb + 4 <= P);
                if (!cb_ok) status |= kStatTopology;
                cn[sd.ca][si] = cb + sd.na;
                cn[sd.cb][si] = cb + sd.nb;
                const bool seam = sd.bdry != 0 && (R[FLAGS] & sd.bdry) && J.periodic;
                if (before) {
                    kind[si] = cb_ok ? (kKindRewire | (seam ? kKindFreshPts : 0)) : kKindFresh;
                    tgt[si] = cb;
                    if (!seam) {
                        edge[si][0] = cb_ok ? J.rec[(size_t)(cb + sd.na) * kI + PT + sd.shared] : 0;
                        edge[si][1] = cb_ok ? J.rec[(size_t)(cb + sd.nb) * kI + PT + sd.shared] : 0;
                        fresh = false;
                    }
                } else if (!seam) {
                    kind[si] = kKindShareNew;
                    tgt[si] = nb;
                    edge[si][0] = edge[si][1] = -3;  // filled by the link step
                    fresh = false;
                }
            } else if (nbr && nb > jj) {
                // nb is split later in the pass and links these children back
                const int cb = P + 4 * J.off[nb - J.mu].x;
                cn[sd.ca][si] = cb + sd.na;
                cn[sd.cb][si] = cb + sd.nb;
            } else if (si == 3 && nb == jj) {
                // a panel that is its own right neighbour (the periodic root)
                cn[2][3] = base;
                cn[0][0] = base + 2;
                cn[3][3] = base + 1;
                cn[1][0] = base + 3;
            }
        }
        if (fresh) {
            edge[si][0] = nv;
            edge[si][1] = nv + 1;
            J.x[nv] = sx[sd.ax];
            J.p[nv] = sp[sd.ap];
            J.x[nv + 1] = sx[sd.bx];
            J.p[nv + 1] = sp[sd.bp];
            nv += 2;
        }
    }
    // interior points: column x1 (p1, p2, p3), x2 (p1, p3), x3 (p1, p2, p3)
    const int i12 = nv, i16 = i12 + 3, i19 = i16 + 2;
    for (int i = 0; i < 3; ++i) {
        J.x[i12 + i] = sx[1];
        J.p[i12 + i] = sp[1 + i];
        J.x[i19 + i] = sx[3];
        J.p[i19 + i] = sp[1 + i];
    }
    J.x[i16] = sx[2];
    J.p[i16] = sp[1];
    J.x[i16 + 1] = sx[2];
    J.p[i16 + 1] = sp[3];
    if (status) ops.or_(&J.ctr->status, status);

    const int* q = R + PT;
    const int pts[4][9] = {
        {q[0], edge[0][0], q[1], edge[1][0], i12, i12 + 1, q[3], i16, q[4]},
        {q[1], edge[0][1], q[2], i12 + 1, i12 + 2, edge[2][0], q[4], i16 + 1, q[5]},
        {q[3], i16, q[4], edge[1][1], i19, i19 + 1, q[6], edge[3][0], q[7]},
        {q[4], i16 + 1, q[5], i19 + 1, i19 + 2, edge[2][1], q[7], edge[3][1], q[8]},
    };
    // (left, top, right, bottom) of each child; cn[child][side], side 0 left,
    // 1 bottom, 2 top, 3 right
    const int nbr[4][4] = {
        {cn[0][0], base + 1, base + 2, cn[0][1]},
        {cn[1][0], cn[1][2], base + 3, base},
        {base, base + 3, cn[2][3], cn[2][1]},
        {base + 1, cn[3][2], cn[3][3], base + 2},
    };
    const bool lb = R[FLAGS] & LB, rb = R[FLAGS] & RB;
    for (int c = 0; c < 4; ++c) {
        int32_t* w = J.rec + (size_t)(base + c) * kI;
        w[0] = base + c;
        for (int i = 0; i < 9; ++i) w[PT + i] = pts[c][i];
        w[LEVEL] = R[LEVEL] + 1;
        w[PARENT] = jj;
        w[WHICH] = c;
        w[LEFT] = nbr[c][0];
        w[TOP] = nbr[c][1];
        w[RIGHT] = nbr[c][2];
        w[BOTTOM] = nbr[c][3];
        w[FLAGS] = c < 2 ? (lb ? LB : 0) : (rb ? RB : 0);
        w[CHILD] = -1;
        w[19] = 0;
    }
    pl[1] = R[FLAGS];
    for (int si = 0; si < 4; ++si) {
        pl[2 + si] = kind[si];
        pl[6 + si] = tgt[si];
    }
    pl[0] = jj;
}

// Member jj: points shared with children made in this pass, links of older
// children to the new ones, its own record, flags for the next pass.
template <class Ops>
FS_HD void link_one(const Job& J, int jj, const Ops& ops) {
    const int P = J.ctr->P;
    if (jj >= P || J.cnt[jj - J.mu].x == 0) return;
    const int r = J.off[jj - J.mu].x;
    const int base = P + 4 * r;
    if (base + 4 > J.cap_p) return;  // reported by the build step
    const int32_t* pl = J.plan + (size_t)r * kPlanInts;
    if (pl[0] != jj) return;  // its build stopped (status set there)
    const int nR = J.ctr->nR;
    for (int si = 0; si < 4; ++si) {
        const SideDef sd = side_def(si);
        const int kind = pl[2 + si] & 7, t = pl[6 + si];
        const bool t_ok = kind == kKindFlag    ? (t >= 0 && t < P + 4 * nR)
                          : kind == kKindRewire ? (t >= 0 && t + 4 <= P)
                          : kind == kKindShareNew ? (t >= J.mu && t < P && J.cnt[t - J.mu].x == 1)
                                                  : true;
        if (!t_ok) {
            ops.or_(&J.ctr->status, kStatTopology);
            continue;
        }
        if (kind == kKindFlag) {
            ops.or_(&J.rec[(size_t)t * kI + FLAGS], NEEDS);
        } else if (kind == kKindRewire) {
            J.rec[(size_t)(t + sd.na) * kI + sd.nfield] = base + sd.ca;
            J.rec[(size_t)(t + sd.nb) * kI + sd.nfield] = base + sd.cb;
        } else if (kind == kKindShareNew) {
            const int tr = J.off[t - J.mu].x;
            // the neighbour made fresh points on its side facing jj
            if (!(J.plan[(size_t)tr * kPlanInts + 2 + (3 - si)] & kKindFreshPts)) {
                ops.or_(&J.ctr->status, kStatTopology);
                continue;
            }
            const int cb = P + 4 * tr;
            const int a = J.rec[(size_t)(cb + sd.na) * kI + PT + sd.shared];
            const int b = J.rec[(size_t)(cb + sd.nb) * kI + PT + sd.shared];
            J.rec[(size_t)(base + sd.ca) * kI + PT + sd.slot] = a;
            J.rec[(size_t)(base + sd.cb) * kI + PT + sd.slot] = b;
        }
    }
    J.rec[(size_t)jj * kI + FLAGS] = (pl[1] | RXP) & ~NEEDS;
    J.rec[(size_t)jj * kI + CHILD] = base;
}

// Counters after the pass; the point job of the new points.
FS_HD void commit_one(const Job& J) {
    Counters& C = *J.ctr;
    const int P0 = C.P, V0 = C.V;
    C.P0 = P0;
    C.P = P0 + 4 * C.nR;
    C.V = V0 + C.newV;
    C.flagged = 0;
    remesh::Job& pj = *J.pj;
    pj.x = J.x + V0;
    pj.p = J.p + V0;
    pj.f = J.f + V0;
    pj.nx = C.newV;
    pj.np = 1;
}

// ---- weights and the reduced-x map ------------------------------------------

FS_HD bool is_leaf(int flags) { return (flags & (RXP | RP)) == 0; }

FS_HD void leaf_init_one(const Job& J, int k) {
    if (k < J.ctr->V) J.nslot[k] = 0;
}

// Leaf k's nine quadrature terms q dx dp / 4 {1,2,1,2,4,2,1,2,1} or / 9
// {1,4,1,4,16,4,1,4,1}, each with the leaf's depth-first path key: the child
// index at every level from the root, 2 bits a level, left-aligned. No leaf's
// path is a prefix of another's, so the keys order the leaves as the
// depth-first walk of set_leaves_weights does (children in index order).
template <class Ops>
FS_HD void leaf_terms_one(const Job& J, int k, const Ops& ops) {
    const int P = J.ctr->P, V = J.ctr->V;
    if (k >= P) return;
    const int32_t* rk = J.rec + (size_t)k * kI;
    if (!is_leaf(rk[FLAGS])) return;
    uint64_t path = 0;
    int a = k, depth = 0;
    while (a != 0) {
        const int b = J.rec[(size_t)a * kI + PARENT];
        const int cs = b >= 0 && b < P ? J.rec[(size_t)b * kI + CHILD] : -1;
        if (b < 0 || b >= P || cs < 0 || cs > a || a - cs > 3 || ++depth > kMaxDepth) {
            ops.or_(&J.ctr->status, kStatDepth);
            return;
        }
        path = (path >> 2) | ((uint64_t)(a - cs) << 62);
        a = b;
    }
    bool ok = true;
    for (int i = 0; i < 9; ++i) ok = ok && rk[PT + i] >= 0 && rk[PT + i] < V;
    if (!ok) {
        ops.or_(&J.ctr->status, kStatTopology);
        return;
    }
    const double dx = J.x[rk[PT + 3]] - J.x[rk[PT + 0]];
    const double dp = J.p[rk[PT + 1]] - J.p[rk[PT + 0]];
    const double w = J.simpson ? J.q * dx * dp / 9.0 : J.q * dx * dp / 4.0;
    for (int i = 0; i < 9; ++i) {
        const double kt = J.simpson ? (i == 4 ? 16.0 : (i & 1) ? 4.0 : 1.0)
                                    : (i == 4 ? 4.0 : (i & 1) ? 2.0 : 1.0);
        const int pt = rk[PT + i];
        const int s = ops.add(&J.nslot[pt], 1);
        if (s >= kSlots) {
            ops.or_(&J.ctr->status, kStatSlots);
            continue;
        }
        J.spath[(size_t)pt * kSlots + s] = path;
        J.sval[(size_t)pt * kSlots + s] = w * kt;
    }
}

// Point i: its leaves' terms added in depth-first leaf order from 0.
FS_HD void weight_one(const Job& J, int i) {
    if (i >= J.ctr->V) return;
    const int n = J.nslot[i] < kSlots ? J.nslot[i] : kSlots;
    uint64_t rk[kSlots];
    double v[kSlots];
    for (int s = 0; s < n; ++s) {
        rk[s] = J.spath[(size_t)i * kSlots + s];
        v[s] = J.sval[(size_t)i * kSlots + s];
    }
    for (int s = 1; s < n; ++s) {
        const uint64_t r = rk[s];
        const double t = v[s];
        int u = s - 1;
        while (u >= 0 && rk[u] > r) {
            rk[u + 1] = rk[u];
            v[u + 1] = v[u];
            --u;
        }
        rk[u + 1] = r;
        v[u + 1] = t;
    }
    double w = 0.0;
    for (int s = 0; s < n; ++s) w += v[s];
    J.bw[i] = w;
}

// Order-preserving image of a double (+0 and -0 one value).
FS_HD uint64_t order_key(double v) {
    if (v == 0.0) v = 0.0;
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

FS_HD void key_one(const Job& J, int i) {
    if (i < J.ctr->V) J.key[i] = order_key(J.x[i]);
}

FS_HD void head_one(const Job& J, int k) {
    if (k >= J.ctr->V) return;
    J.vcnt[k] = make_int2((k == 0 || J.skey[k] != J.skey[k - 1]) ? 1 : 0, 0);
}

// Distinct x ascending, members by x then index (the sort is stable), and
// each point's distinct-x index.
FS_HD void map_one(const Job& J, int k) {
    const int V = J.ctr->V;
    if (k >= V) return;
    const int h = J.vcnt[k].x;
    const int run = J.voff[k].x + h - 1;
    const int pi = J.perm[k];
    J.members[k] = pi;
    J.uof[pi] = run;
    if (h) {
        J.ux[run] = J.x[pi];
        J.roff[run] = k;
    }
    if (k == 0) J.roff[J.ctr->U] = V;
}

}  // namespace pass
}  // namespace amr
}  // namespace fs
