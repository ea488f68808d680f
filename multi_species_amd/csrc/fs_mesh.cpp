// Host-side panel tree (see fs_mesh.h). The panel and point insertion order,
// the neighbour bookkeeping and the test/quadrature arithmetic follow the
// reference exactly, because they define the bit-exact tree that the
// device interpolation and the field see. Parent fields are read before the
// children are appended (the reference reads them through a pointer that
// push_back may invalidate; this is its evident intent). In test_panel the
// p-derivative criterion uses the two values the reference's loop leaves
// set and zero for the four it leaves uninitialised.
#pragma clang fp contract(off)

#include "fs_mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>

namespace fs {
namespace mesh {

int PanelTree::append_panel(int parent, int which, const int (&pts)[9], int l, int t, int r,
                            int b, bool lb, bool rb) {
    const int id = size();
    int32_t q[kInts] = {};
    q[IND] = id;
    for (int k = 0; k < 9; ++k) q[PT + k] = pts[k];
    q[LEVEL] = parent >= 0 ? at(parent, LEVEL) + 1 : 0;
    q[PARENT] = parent;
    q[WHICH] = which;
    q[LEFT] = l;
    q[TOP] = t;
    q[RIGHT] = r;
    q[BOTTOM] = b;
    q[FLAGS] = (lb ? LEFT_BDRY : 0) | (rb ? RIGHT_BDRY : 0);
    q[CHILD] = -1;
    rec.insert(rec.end(), q, q + kInts);
    return id;
}

void PanelTree::build_base(const Spec& s) {
    if (s.initial_height < 1) throw std::invalid_argument("height must be greater than 1");
    periodic_ = s.periodic;
    const double dx = (s.x_max - s.x_min) / 4, dp = (s.p_max - s.p_min) / 4;
    double gx[5], gp[5];
    for (int i = 0; i < 5; ++i) {
        gx[i] = s.x_min + i * dx;
        gp[i] = s.p_min + i * dp;
    }
    rec.clear();
    xs.clear();
    ps.clear();
    // 25 points: the 3x3 corners/midpoints of the root, then the columns
    // x1 (ml, mr of the left children) and x3, then the right edge
    // (AMRStructure_amr.cpp:145-162)
    for (int i = 0; i < 5; i += 2)
        for (int j = 0; j < 5; j += 2) {
            xs.push_back(gx[i]);
            ps.push_back(gp[j]);
        }
    for (int i = 0; i < 2; ++i) {
        const double xe = gx[2 * i], xc = gx[1 + 2 * i];
        xs.push_back(xe), ps.push_back(gp[1]);
        xs.push_back(xe), ps.push_back(gp[3]);
        for (int j = 0; j < 5; ++j) xs.push_back(xc), ps.push_back(gp[j]);
    }
    for (int j = 1; j < 5; j += 2) xs.push_back(gx[4]), ps.push_back(gp[j]);
    fs.assign(xs.size(), 1.0);

    // root: Panel::Panel() (Panel.cpp:15-24): neighbours left/right 0,
    // top/bottom -2, points 0..8; then the four level-1 children
    const int root_pts[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
    const bool per = s.periodic;
    append_panel(-1, -1, root_pts, per ? 0 : -2, -2, per ? 0 : -2, -2, per, per);
    const int c[4][9] = {{0, 9, 1, 11, 12, 13, 3, 16, 4},
                         {1, 10, 2, 13, 14, 15, 4, 17, 5},
                         {3, 16, 4, 18, 19, 20, 6, 23, 7},
                         {4, 17, 5, 20, 21, 22, 7, 24, 8}};
    // neighbours (left, top, right, bottom) of the level-1 panels
    // (AMRStructure_amr.cpp:173-192)
    const int nper[4][4] = {{3, 2, 3, -2}, {4, -2, 4, 1}, {1, 4, 1, -2}, {2, -2, 2, 3}};
    const int nopen[4][4] = {{-2, 2, 3, -2}, {-2, -2, 4, 1}, {1, 4, -2, -2}, {2, -2, -2, 3}};
    for (int k = 0; k < 4; ++k) {
        const int* nb = per ? nper[k] : nopen[k];
        append_panel(0, k, c[k], nb[0], nb[1], nb[2], nb[3], per && k < 2, per && k >= 2);
        at(1 + k, FLAGS) |= NEEDS_REF;
    }
    at(0, FLAGS) |= REFINED_XP;
    at(0, CHILD) = 1;
    min_unrefined = 1;

    auto one = [](const double*, const double*, size_t n, double* f) {
        std::fill(f, f + n, 1.0);
    };
    for (int lev = 1; lev < s.initial_height; ++lev) {
        const int before = size();
        for (int k = min_unrefined; k < size(); ++k) at(k, FLAGS) |= NEEDS_REF;
        refine_xp(one);
        min_unrefined = before;
    }
    for (int lev = 0; lev < s.p_height; ++lev) {
        const int before = size();
        for (int k = min_unrefined; k < size(); ++k) at(k, FLAGS) |= NEEDS_REF;
        refine_p(one);
        min_unrefined = before;
    }
}

namespace {

// One side of a panel being split in x and p.
struct Side {
    int field;          // neighbour field of the panel
    int ca, cb;         // the two children on this side
    int na, nb;         // the neighbour's children facing them (offsets)
    int nfield;         // field of those neighbour children that points back
    int shared_pt;      // their point that this side's new points coincide with
    int ax, ap, bx, bp; // sub-grid coordinates of the two new edge points
    int bdry_flag;      // panel flag that forces fresh points (periodic seam)
};
// Processed in this order: left, bottom, top, right (it numbers the points).
const Side kSides[4] = {
    {LEFT, 0, 1, 2, 3, RIGHT, 7, 0, 1, 0, 3, LEFT_BDRY},
    {BOTTOM, 0, 2, 1, 3, TOP, 5, 1, 0, 3, 0, 0},
    {TOP, 1, 3, 0, 2, BOTTOM, 3, 1, 4, 3, 4, 0},
    {RIGHT, 2, 3, 0, 1, LEFT, 1, 4, 1, 4, 3, RIGHT_BDRY},
};

}  // namespace

// Room for a pass that splits every panel flagged now into `kids` children
// with at most `pts` new points each (balance flags may add more panels to
// the same pass; the vectors then grow as usual).
void PanelTree::reserve_pass(int kids, int pts, std::vector<double>& nx,
                             std::vector<double>& np) {
    size_t flagged = 0;
    for (int k = min_unrefined; k < size(); ++k) flagged += has(k, NEEDS_REF) ? 1 : 0;
    rec.reserve(rec.size() + (flagged * kids + 16) * kInts);
    nx.reserve(flagged * pts + 16);
    np.reserve(flagged * pts + 16);
    xs.reserve(xs.size() + flagged * pts + 16);
    ps.reserve(ps.size() + flagged * pts + 16);
    fs.reserve(fs.size() + flagged * pts + 16);
}

// The new points of a pass, kept per thread so that their pages stay mapped
// from pass to pass.
struct PassBuffers {
    std::vector<double> nx, np, nf;
};
static PassBuffers& pass_buffers() {
    static thread_local PassBuffers b;
    b.nx.clear();
    b.np.clear();
    return b;
}

void PanelTree::refine_xp(const BatchEval& f) {
    PassBuffers& pb = pass_buffers();
    std::vector<double>& nx = pb.nx;
    std::vector<double>& np = pb.np;
    reserve_pass(4, 16, nx, np);
    int nv = (int)xs.size();
    const int before = size();
    for (int jj = min_unrefined; jj < before; ++jj) {
        if (!has(jj, NEEDS_REF)) continue;
        int32_t P[kInts];
        std::copy(&rec[(size_t)jj * kInts], &rec[(size_t)jj * kInts] + kInts, P);
        const double x0 = xs[P[PT + 0]], p0 = ps[P[PT + 0]];
        const double hx = 0.5 * (xs[P[PT + 3]] - x0), hp = 0.5 * (ps[P[PT + 1]] - p0);
        double sx[5], sp[5];
        for (int i = 0; i < 5; ++i) {
            sx[i] = x0 + hx * i;
            sp[i] = p0 + hp * i;
        }
        const int base = size();
        int edge[4][2];          // new/shared point indices per side
        int cn[4][4];            // child neighbour per child per field slot
        for (auto& row : cn)
            for (int& v : row) v = -1;
        for (int si = 0; si < 4; ++si) {
            const Side& sd = kSides[si];
            const int nb = P[sd.field];
            auto fresh = [&]() {
                edge[si][0] = nv++;
                edge[si][1] = nv++;
                nx.push_back(sx[sd.ax]), np.push_back(sp[sd.ap]);
                nx.push_back(sx[sd.bx]), np.push_back(sp[sd.bp]);
            };
            if (nb == -2) {
                cn[sd.ca][si] = cn[sd.cb][si] = -2;
                fresh();
            } else if (nb == -1) {
                const int q = at(P[PARENT], sd.field);
                if (!has(q, REFINED_XP)) {
                    at(q, FLAGS) |= NEEDS_REF;
                    need_further = true;
                }
                fresh();
            } else if (!has(nb, REFINED_XP)) {
                fresh();
                if (si == 3 && nb == jj) {  // a panel that is its own right neighbour
                    cn[2][3] = base;
                    cn[0][0] = base + 2;
                    cn[3][3] = base + 1;
                    cn[1][0] = base + 3;
                }
            } else {
                const int qa = at(nb, CHILD) + sd.na, qb = at(nb, CHILD) + sd.nb;
                cn[sd.ca][si] = qa;
                cn[sd.cb][si] = qb;
                at(qa, sd.nfield) = base + sd.ca;
                at(qb, sd.nfield) = base + sd.cb;
                if (sd.bdry_flag && (P[FLAGS] & sd.bdry_flag) && periodic_) {
                    fresh();
                } else {
                    edge[si][0] = at(qa, PT + sd.shared_pt);
                    edge[si][1] = at(qb, PT + sd.shared_pt);
                }
            }
        }
        // interior points: column x1 (p1,p2,p3), x2 (p1,p3), x3 (p1,p2,p3)
        const int i12 = nv, i16 = i12 + 3, i19 = i16 + 2;
        for (int i = 0; i < 3; ++i) nx.push_back(sx[1]), np.push_back(sp[1 + i]);
        for (int i = 0; i < 2; ++i) nx.push_back(sx[2]), np.push_back(sp[1 + 2 * i]);
        for (int i = 0; i < 3; ++i) nx.push_back(sx[3]), np.push_back(sp[1 + i]);
        nv += 8;

        at(jj, FLAGS) = (at(jj, FLAGS) | REFINED_XP) & ~NEEDS_REF;
        at(jj, CHILD) = base;
        const int* q = P + PT;
        const int L9 = edge[0][0], L10 = edge[0][1], B11 = edge[1][0], B18 = edge[1][1];
        const int T15 = edge[2][0], T22 = edge[2][1], R23 = edge[3][0], R24 = edge[3][1];
        const bool lb = P[FLAGS] & LEFT_BDRY, rb = P[FLAGS] & RIGHT_BDRY;
        // cn[child][side]: side 0 left, 1 bottom, 2 top, 3 right
        const int c0[9] = {q[0], L9, q[1], B11, i12, i12 + 1, q[3], i16, q[4]};
        const int c1[9] = {q[1], L10, q[2], i12 + 1, i12 + 2, T15, q[4], i16 + 1, q[5]};
        const int c2[9] = {q[3], i16, q[4], B18, i19, i19 + 1, q[6], R23, q[7]};
        const int c3[9] = {q[4], i16 + 1, q[5], i19 + 1, i19 + 2, T22, q[7], R24, q[8]};
        append_panel(jj, 0, c0, cn[0][0], base + 1, base + 2, cn[0][1], lb, false);
        append_panel(jj, 1, c1, cn[1][0], cn[1][2], base + 3, base, lb, false);
        append_panel(jj, 2, c2, base, base + 3, cn[2][3], cn[2][1], false, rb);
        append_panel(jj, 3, c3, base + 1, cn[3][2], cn[3][3], base + 2, false, rb);
    }
    std::vector<double>& nf = pb.nf;
    nf.resize(nx.size());
    if (!nx.empty()) f(nx.data(), np.data(), nx.size(), nf.data());
    xs.insert(xs.end(), nx.begin(), nx.end());
    ps.insert(ps.end(), np.begin(), np.end());
    fs.insert(fs.end(), nf.begin(), nf.end());
}

void PanelTree::refine_p(const BatchEval& f) {
    PassBuffers& pb = pass_buffers();
    std::vector<double>& nx = pb.nx;
    std::vector<double>& np = pb.np;
    reserve_pass(2, 6, nx, np);
    int nv = (int)xs.size();
    const int before = size();
    auto refined = [&](int k) { return has(k, REFINED_XP) || has(k, REFINED_P); };
    for (int jj = min_unrefined; jj < before; ++jj) {
        if (!has(jj, NEEDS_REF)) continue;
        int32_t P[kInts];
        std::copy(&rec[(size_t)jj * kInts], &rec[(size_t)jj * kInts] + kInts, P);
        const double x0 = xs[P[PT + 0]], p0 = ps[P[PT + 0]];
        const double hx = 0.5 * (xs[P[PT + 3]] - x0), hp = 0.5 * (ps[P[PT + 1]] - p0);
        double sx[5], sp[5];
        for (int i = 0; i < 5; ++i) {
            sp[i] = p0 + hp * i;
            sx[i] = x0 + hx * i;
        }
        const int base = size();
        // children's left/right default to themselves until the neighbour on
        // that side is refined (AMRStructure_amr.cpp:284-289)
        int c0L = base, c0R = base, c1L = base + 1, c1R = base + 1, c0B = -1, c1T = -1;
        int e[2][2];  // left (9, 10), right (13, 14)
        for (int side = 0; side < 2; ++side) {
            const bool left = side == 0;
            const int fld = left ? LEFT : RIGHT;
            const int nb = P[fld];
            const double xe = left ? sx[0] : sx[4];
            auto fresh = [&]() {
                e[side][0] = nv++;
                e[side][1] = nv++;
                nx.push_back(xe), np.push_back(sp[1]);
                nx.push_back(xe), np.push_back(sp[3]);
            };
            int *ca = left ? &c0L : &c0R, *cb = left ? &c1L : &c1R;
            if (nb == -2) {
                *ca = *cb = -2;
                fresh();
            } else if (nb == -1) {
                const int q = at(P[PARENT], fld);
                if (!refined(q)) {
                    at(q, FLAGS) |= NEEDS_REF;
                    need_further = true;
                }
                fresh();
            } else if (!refined(nb)) {
                fresh();
            } else {
                const int off = (left && has(nb, REFINED_XP)) ? 2 : 0;
                *ca = at(nb, CHILD) + off;
                *cb = at(nb, CHILD) + off + 1;
                at(*ca, left ? RIGHT : LEFT) = base;
                at(*cb, left ? RIGHT : LEFT) = base + 1;
                const int bflag = left ? LEFT_BDRY : RIGHT_BDRY;
                if ((P[FLAGS] & bflag) && periodic_) {
                    fresh();
                } else {
                    const int spt = left ? 7 : 1;
                    e[side][0] = at(*ca, PT + spt);
                    e[side][1] = at(*cb, PT + spt);
                }
            }
            if (left) {
                // bottom then top are checked between left and right
                const int bn = P[BOTTOM];
                if (bn == -2) {
                    c0B = -2;
                } else if (bn == -1) {
                    std::fprintf(stderr, "not allowed to refine in v if panel doesn't have bottom neighbor!\n");
                } else if (has(bn, REFINED_XP)) {
                    std::fprintf(stderr, "Shouldn't be allowed to call refine in v if bottom neighbor is refined in x and v!\n");
                } else if (!has(bn, REFINED_P)) {
                    c0B = bn;
                    at(bn, TOP) = base;
                } else {
                    c0B = at(bn, CHILD) + 1;
                    at(c0B, TOP) = base;
                }
                const int tn = P[TOP];
                if (tn == -2) {
                    c1T = -2;
                } else if (tn == -1) {
                    std::fprintf(stderr, "not allowed to refine in v if panel doesn't have top neighbor!\n");
                } else if (has(tn, REFINED_XP)) {
                    std::fprintf(stderr, "Shouldn't be allowed to call refine in v if bottom neighbor is refined in x and v!\n");
                } else if (has(tn, REFINED_P)) {
                    c1T = at(tn, CHILD);
                    at(c1T, BOTTOM) = base + 1;
                }
            }
        }
        const int i11 = nv;
        nx.push_back(sx[2]), np.push_back(sp[1]);
        nx.push_back(sx[2]), np.push_back(sp[3]);
        nv += 2;

        at(jj, FLAGS) = (at(jj, FLAGS) | REFINED_P) & ~NEEDS_REF;
        at(jj, CHILD) = base;
        const int* q = P + PT;
        const bool lb = P[FLAGS] & LEFT_BDRY, rb = P[FLAGS] & RIGHT_BDRY;
        const int c0[9] = {q[0], e[0][0], q[1], q[3], i11, q[4], q[6], e[1][0], q[7]};
        const int c1[9] = {q[1], e[0][1], q[2], q[4], i11 + 1, q[5], q[7], e[1][1], q[8]};
        append_panel(jj, 0, c0, c0L, base + 1, c0R, c0B, lb, rb);
        append_panel(jj, 1, c1, c1L, c1T, c1R, base, lb, rb);
    }
    std::vector<double>& nf = pb.nf;
    nf.resize(nx.size());
    if (!nx.empty()) f(nx.data(), np.data(), nx.size(), nf.data());
    xs.insert(xs.end(), nx.begin(), nx.end());
    ps.insert(ps.end(), np.begin(), np.end());
    fs.insert(fs.end(), nf.begin(), nf.end());
}

void PanelTree::test_panel(int pi, const Spec& s) {
    double f9[9];
    for (int i = 0; i < 9; ++i) f9[i] = fs[at(pi, PT + i)];
    const size_t ne = s.amr_eps.size();
    bool crit = false;
    if (ne > 0) {
        double mx = f9[0], mn = f9[0];
        for (int i = 1; i < 9; ++i) {
            mx = (mx < f9[i]) ? f9[i] : mx;
            mn = (mn > f9[i]) ? f9[i] : mn;
        }
        crit = crit || (mx - mn > s.amr_eps[0]);
        if (ne > 1) crit = crit || (mx - mn) / f9[4] > s.amr_eps[1];
        if (ne > 2) crit = crit || mx / mn > s.amr_eps[2];
    }
    if (ne > 3) {
        const double dx = xs[at(pi, PT + 3)] - xs[at(pi, PT + 0)];
        double m = std::fabs((f9[3] - f9[0]) / dx);
        for (int i = 1; i < 6; ++i) {
            const double a = std::fabs((f9[3 + i] - f9[i]) / dx);
            if (m < a) m = a;
        }
        crit = crit || (m > s.amr_eps[3]);
    }
    if (ne > 4) {
        const double dp = ps[at(pi, PT + 1)] - ps[at(pi, PT + 0)];
        // the reference's loop leaves only its last row's two slopes set
        const double a0 = std::fabs((f9[7] - f9[6]) / dp), a1 = std::fabs((f9[8] - f9[7]) / dp);
        double m = a0;
        if (m < a1) m = a1;
        if (m < 0.0) m = 0.0;
        crit = crit || (m > s.amr_eps[4]);
    }
    if (at(pi, LEVEL) < s.max_height && crit) {
        at(pi, FLAGS) |= NEEDS_REF;
        need_further = true;
    }
}

void PanelTree::refine_adaptively(const Spec& s, const BatchEval& f) {
    for (int k = min_unrefined; k < size(); ++k) test_panel(k, s);
    while (need_further) {
        need_further = false;
        refine_xp(f);
        for (int k = min_unrefined; k < size(); ++k)
            if (!has(k, REFINED_XP)) test_panel(k, s);
    }
}

void PanelTree::leaf_weights(double q, bool simpson, std::vector<double>& w,
                             std::vector<int32_t>& leaves) const {
    w.assign(xs.size(), 0.0);
    leaves.clear();
    static const double kTrap[9] = {1.0, 2.0, 1.0, 2.0, 4.0, 2.0, 1.0, 2.0, 1.0};
    static const double kSimp[9] = {1.0, 4.0, 1.0, 4.0, 16.0, 4.0, 1.0, 4.0, 1.0};
    // depth-first, children in index order
    std::vector<int> stack(1, 0);
    while (!stack.empty()) {
        const int pi = stack.back();
        stack.pop_back();
        const int nch = has(pi, REFINED_P) ? 2 : (has(pi, REFINED_XP) ? 4 : 0);
        if (nch) {
            for (int c = nch - 1; c >= 0; --c) stack.push_back(at(pi, CHILD) + c);
            continue;
        }
        leaves.push_back(pi);
        const double dx = xs[at(pi, PT + 3)] - xs[at(pi, PT + 0)];
        const double dp = ps[at(pi, PT + 1)] - ps[at(pi, PT + 0)];
        const double a = simpson ? q * dx * dp / 9.0 : q * dx * dp / 4.0;
        const double* k = simpson ? kSimp : kTrap;
        for (int i = 0; i < 9; ++i) w[at(pi, PT + i)] += a * k[i];
    }
}

std::vector<int32_t> grid_rows(const std::vector<double>& xs, const std::vector<double>& ps,
                               int nx, int np) {
    const int n = nx * np;
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return ps[a] < ps[b]; });
    (void)xs;
    return order;
}

ReducedMap reduced_map(const std::vector<double>& xs) {
    ReducedMap m;
    reduced_map(xs, m);
    return m;
}

void reduced_map(const std::vector<double>& xs, ReducedMap& m) {
    // unique values by exact equality (as the reference's sort-then-compare
    // does; +0 and -0 are one value), found with an open-addressed hash in
    // point order -- a remeshed species has ~1e3 distinct x among ~1e5..1e6
    // points -- then ranked by sorting only the distinct values
    const size_t n = xs.size();
    m.unique_of.assign(n, 0);
    size_t cap = 1024;
    std::vector<int32_t> slot(cap, -1);  // -> index into `first`
    std::vector<int32_t> first;          // first point of each distinct value
    auto key = [](double v) {
        if (v == 0.0) v = 0.0;
        uint64_t u;
        std::memcpy(&u, &v, sizeof u);
        u ^= u >> 33;
        u *= 0xff51afd7ed558ccdULL;
        u ^= u >> 33;
        return u;
    };
    static thread_local std::vector<int32_t> tag;
    tag.resize(n);
    for (size_t i = 0; i < n; ++i) {
        if (2 * (first.size() + 1) > cap) {  // grow, rehash
            cap *= 2;
            slot.assign(cap, -1);
            for (size_t d = 0; d < first.size(); ++d) {
                size_t h = key(xs[first[d]]) & (cap - 1);
                while (slot[h] >= 0) h = (h + 1) & (cap - 1);
                slot[h] = (int32_t)d;
            }
        }
        size_t h = key(xs[i]) & (cap - 1);
        while (slot[h] >= 0 && !(xs[first[slot[h]]] == xs[i])) h = (h + 1) & (cap - 1);
        if (slot[h] < 0) {
            slot[h] = (int32_t)first.size();
            first.push_back((int32_t)i);
        }
        tag[i] = slot[h];
    }
    const size_t U = first.size();
    std::vector<int32_t> order(U);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(),
              [&](int32_t a, int32_t b) { return xs[first[a]] < xs[first[b]]; });
    std::vector<int32_t> rank(U);
    m.ux.resize(U);
    for (size_t u = 0; u < U; ++u) {
        rank[order[u]] = (int32_t)u;
        m.ux[u] = xs[first[order[u]]];
    }
    for (size_t i = 0; i < n; ++i) m.unique_of[i] = rank[tag[i]];
    m.offsets.assign(U + 1, 0);
    for (size_t i = 0; i < n; ++i) m.offsets[m.unique_of[i] + 1]++;
    for (size_t k = 0; k < U; ++k) m.offsets[k + 1] += m.offsets[k];
    m.members.assign(n, 0);
    std::vector<int32_t> fill(m.offsets.begin(), m.offsets.end() - 1);
    for (size_t i = 0; i < n; ++i) m.members[fill[m.unique_of[i]]++] = (int32_t)i;
}

}  // namespace mesh
}  // namespace fs
