// Test infrastructure only: FP64 CPU restatement of the reference's periodic
// field solvers (src/FieldStructure.cpp). Checked against the reference's
// own compiled solver (oracle/_ref) on seeded inputs, against closed forms and
// against the extended-precision sum below (tests/test_oracle_field.py).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "fs_oracle_internal.h"

namespace orc {

// src/FieldStructure.cpp:106-156. Same association of every product and sum
// as the reference expression, `round` for the periodic wrap, the three
// DBL_MIN skips and a sequential inner sum per target.
void direct_sum(double* es, const double* targets, int64_t nt,
                const double* sources, const double* q_ws, int64_t ns,
                double L, double epsilon) {
    const double norm_const = 1 / std::sqrt(L * L + 4 * epsilon * epsilon);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nt; ++i) {
        const double xi = targets[i];
        double ei = 0.0;
        for (int64_t j = 0; j < ns; ++j) {
            double z = (xi - sources[j]) / L;
            z = z - std::round(z);
            if (std::fabs(z - 0.5) < DBL_MIN || std::fabs(z + 0.5) < DBL_MIN ||
                std::fabs(z) < DBL_MIN)
                continue;
            ei += q_ws[j] * -1.0 * z * L *
                  (norm_const - 1 / std::sqrt(4 * z * z * L * L +
                                              4 * epsilon * epsilon));
        }
        es[i] = ei;
    }
}

// src/FieldStructure.cpp:166-191 (E_MQ_DirectSum_openbcs::operator()): the
// same expression, q_ws[j] * 0.5 * z / sqrt(z*z + eps^2) summed in j order,
// then the -x_i background.
void direct_sum_open(double* es, const double* targets, int64_t nt, const double* sources,
                     const double* q_ws, int64_t ns, double epsilon) {
    const double eps_sq = epsilon * epsilon;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nt; ++i) {
        const double xi = targets[i];
        double ei = 0.0;
        for (int64_t j = 0; j < ns; ++j) {
            const double z = xi - sources[j];
            ei += q_ws[j] * 0.5 * z / std::sqrt(z * z + eps_sq);
        }
        es[i] = ei - xi;
    }
}

// The drive amplitude of src/FieldStructure.cpp:17-139 (ExternalTanh's
// a_t_k, ExternalSine's and ExternalLogistic's a0) with the constants of
// inc/FieldStructure.hpp:49-84; type as the deck loader's switch
// (AMRSimulation_load.cpp:110-124: 1 sine, 2 logistic, otherwise tanh).
double external_amplitude(int type, double t, double Am, double k) {
    if (type == 1) {
        const double t1 = 50, t2 = 150, t3 = 200;
        double a0 = 0;
        if (t < t1)
            a0 = Am * std::sin(t * M_PI / 100);
        else if (t < t2)
            a0 = Am;
        else if (t < t3)
            a0 = Am * std::cos((t - 150) * M_PI / 100);
        return a0;
    }
    if (type == 2) {
        const double t1 = 60;
        if (t < t1) return Am / (1.0 + std::exp(-40.0 * (t - 10.0)));
        return Am * (1.0 - 1.0 / (1.0 + std::exp(-40.0 * (t - 110.0))));
    }
    const double t0 = 0, tL = 69, tR = 307, twL = 20, twR = 20;
    auto drive_fun = [&](double s) {
        return 0.5 * (std::tanh((s - tL) / twL) - std::tanh((s - tR) / twR));
    };
    const double eps = drive_fun(t0);
    return (drive_fun(t) - eps) / (1 - eps) * Am * k;
}

// es[i] += a(t) * sin(k * x_i - omega * t) (the three operator()s' loop).
void external_add(double* es, const double* targets, int64_t nt, double t, int type, double Am,
                  double k, double omega) {
    const double a = external_amplitude(type, t, Am, k);
    const double omt = omega * t;
    for (int64_t i = 0; i < nt; ++i) es[i] += a * std::sin(k * targets[i] - omt);
}

namespace {

// Treecode node (src/FieldStructure.hpp:26-40), with its depth below the
// root for the tree export.
struct TNode {
    size_t m0 = 0, m1 = 0;
    double x0 = 0, x1 = 0, xc = 0;
    int depth = 0;
    std::vector<size_t> children;
};

thread_local int64_t g_tc_stats[4] = {0, 0, 0, 0};

// The tree of E_MQ_Treecode in its tree order: px and lambda reordered,
// old_index the input index of each tree-order point, leaves in depth-first
// order, far/near the interaction lists of each leaf.
struct TTree {
    std::vector<double> px, lambda;
    std::vector<int> old_index;
    std::vector<TNode> tree;
    std::vector<size_t> leaves;
    std::vector<std::vector<size_t>> far, near;
    bool ok = true;
};

// src/FieldStructure.cpp:240-247 (fmod), build_tree_init (:506-527) with
// minval/maxval and their paddings (:1032-1050), build_tree_1D_Recursive and
// split_2 (:530-632), build_interaction_list (:663-702).
TTree build_tree(const double* sources, const double* q_ws, size_t n, double L, double mac,
                 int max_source) {
    TTree T;
    std::vector<double>& px = T.px;
    std::vector<double>& lambda = T.lambda;
    px.resize(n);
    lambda.resize(n);
    for (size_t i = 0; i < n; ++i) {            // :240-247
        px[i] = std::fmod(sources[i], L);
        lambda[i] = q_ws ? q_ws[i] : 0.0;
    }
    const double sq_theta = mac * mac;
    const size_t N0 = (size_t)max_source;

    // minval/maxval with their paddings (:1032-1050)
    double mn = px[0], mx = px[0];
    for (size_t i = 1; i < n; ++i) {
        if (mn > px[i]) mn = px[i];
        if (mx < px[i]) mx = px[i];
    }
    mn -= 0.1234;
    mx += 0.001;

    std::vector<int>& old_index = T.old_index;
    old_index.resize(n);
    for (size_t i = 0; i < n; ++i) old_index[i] = (int)i;

    // build_tree_init (:506-527)
    std::vector<TNode>& tree = T.tree;
    tree.resize(1);
    tree[0].m0 = 0;
    tree[0].m1 = n - 1;
    tree[0].x0 = mn;
    tree[0].x1 = mx;
    tree[0].xc = 0.5 * (mn + mx);
    std::vector<size_t>& leaves = T.leaves;

    // split_2 (:555-632): midpoint bisection, x <= mid goes to child 0, the
    // address-table swaps amount to a stable partition of x, lambda and the
    // original indices; empty children are dropped. A node deeper than 200
    // that must split (more than max_source coincident points) fails.
    std::function<void(size_t)> build = [&](size_t node) {   // :530-545
        if (!T.ok) return;
        const size_t cnt = tree[node].m1 - tree[node].m0 + 1;
        if (cnt >= N0) {
            if (tree[node].depth > 200) {
                T.ok = false;
                return;
            }
            const double a = tree[node].x0, b = tree[node].x1;
            const double mid = 0.5 * (a + b);
            const size_t s = tree[node].m0, e = tree[node].m1;
            std::vector<size_t> side[2];
            for (size_t k = s; k <= e; ++k) side[px[k] <= mid ? 0 : 1].push_back(k);
            std::vector<double> nx, nl;
            std::vector<int> no;
            nx.reserve(cnt);
            nl.reserve(cnt);
            no.reserve(cnt);
            for (int c = 0; c < 2; ++c)
                for (size_t k : side[c]) {
                    nx.push_back(px[k]);
                    nl.push_back(lambda[k]);
                    no.push_back(old_index[k]);
                }
            std::copy(nx.begin(), nx.end(), px.begin() + s);
            std::copy(nl.begin(), nl.end(), lambda.begin() + s);
            std::copy(no.begin(), no.end(), old_index.begin() + s);
            size_t seq = s;
            for (int c = 0; c < 2; ++c) {
                if (side[c].empty()) continue;
                TNode ch;
                ch.x0 = (c == 0) ? a : mid;
                ch.x1 = (c == 0) ? mid : b;
                ch.xc = 0.5 * (ch.x0 + ch.x1);
                ch.m0 = seq;
                ch.m1 = seq + side[c].size() - 1;
                ch.depth = tree[node].depth + 1;
                seq += side[c].size();
                tree[node].children.push_back(tree.size());
                tree.push_back(ch);
            }
            const std::vector<size_t> kids = tree[node].children;
            for (size_t k : kids) build(k);
        } else {
            leaves.push_back(node);
        }
    };
    build(0);
    if (!T.ok) return T;

    // build_interaction_list (:663-702)
    const size_t nleaf = leaves.size();
    T.far.assign(nleaf, {});
    T.near.assign(nleaf, {});
    std::function<void(size_t, size_t)> lists = [&](size_t li, size_t node) {
        const double xcb = tree[leaves[li]].xc, xcp = tree[node].xc;
        const double x_vb = tree[leaves[li]].x1, x_vp = tree[node].x1;
        const double r_dist = std::fabs(xcb - xcp);
        const double dfbpx = std::fmin(r_dist, L - r_dist);
        const double dfbx = x_vb - xcb, dfpx = x_vp - xcp;
        const double sq_R = dfbpx * dfbpx, sq_rb = dfbx * dfbx, sq_rp = dfpx * dfpx;
        if (tree[node].children.empty()) {
            T.near[li].push_back(node);
        } else if ((std::sqrt(sq_rb) + std::sqrt(sq_rp)) / std::sqrt(sq_R) <
                   std::sqrt(sq_theta)) {
            T.far[li].push_back(node);
        } else {
            for (size_t c : tree[node].children) lists(li, c);
        }
    };
    for (size_t li = 0; li < nleaf; ++li) lists(li, 0);
    return T;
}

}  // namespace

// src/FieldStructure.cpp:230-1050 without the heap bookkeeping (and without
// its double frees). Evaluates at the sources. False when more than
// max_source points share a coordinate (the reference recurses without end).
bool treecode(double* es, const double* sources, const double* q_ws,
              int64_t ns_, double L, double epsilon, double mac, int degree,
              int max_source) {
    const size_t n = (size_t)ns_;
    const int P = degree, PP = P + 1;            // :249-256
    const double deno = L * L + 4 * epsilon * epsilon;
    TTree T = build_tree(sources, q_ws, n, L, mac, max_source);
    if (!T.ok) return false;
    const std::vector<double>& px = T.px;
    const std::vector<double>& lambda = T.lambda;
    const std::vector<int>& old_index = T.old_index;
    const std::vector<TNode>& tree = T.tree;
    const std::vector<size_t>& leaves = T.leaves;
    const size_t nleaf = leaves.size();
    const std::vector<std::vector<size_t>>& far = T.far;
    const std::vector<std::vector<size_t>>& near = T.near;

    // alloc_set_cluster_list (:775-866): Chebyshev points of the 2nd kind and
    // barycentric moments for every node but the root.
    const double pi = 3.14159265358979323846;
    const double h = pi / P;
    std::vector<double> t(PP), w1i(PP), dj(PP);
    for (int i = 0; i < PP; ++i) t[i] = std::cos(i * h);
    dj[0] = 0.5;
    dj[P] = 0.5;
    for (int j = 1; j < P; ++j) dj[j] = 1.0;
    for (int j = 0; j < PP; ++j) w1i[j] = ((j % 2 == 0) ? 1 : -1) * dj[j];
    const size_t nnode = tree.size();
    std::vector<double> t1(nnode * PP, 0.0), mom(nnode * PP, 0.0);
    // nodes in parallel: each node's moments are summed over its own points
    // in the reference's order
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t node = 1; node < nnode; ++node) {
        std::vector<double> a1i(PP);
        const double x1 = tree[node].x0, x2 = tree[node].x1;
        double* tn = &t1[node * PP];
        double* mn_ = &mom[node * PP];
        for (int j = 0; j < PP; ++j) tn[j] = x1 + (t[j] + 1.0) * 0.5 * (x2 - x1);
        for (size_t k = tree[node].m0; k <= tree[node].m1; ++k) {
            const double x = px[k];
            int exact = -1;
            double SumA1 = 0.0;
            for (int j = 0; j < PP; ++j) {
                const double dx = x - tn[j];
                if (std::fabs(dx) <= DBL_MIN) {
                    exact = j;
                } else {
                    a1i[j] = w1i[j] / dx;
                    SumA1 += a1i[j];
                }
            }
            if (exact > -1) {
                SumA1 = 1.0;
                for (int j = 0; j < PP; ++j) a1i[j] = 0.0;
                a1i[exact] = 1.0;
            }
            const double D = 1.0 / SumA1;
            for (int i = 0; i < PP; ++i) {
                const double s = a1i[i] * D;
                mn_[i] += s * lambda[k];
            }
        }
    }

    // Call_BL (:904-961) then Call_Ds (:963-1020), accumulated per target in
    // the reordered numbering.
    std::vector<double> keps(n, 0.0);
    int64_t far_evals = 0, near_evals = 0;
    auto kern = [&](double s) {
        return -1.0 * s * (1 - std::sqrt(deno / (4 * s * s * L * L +
                                                 4 * epsilon * epsilon)));
    };
    // leaves in parallel: every target's far and near sums keep the
    // reference's order (each target belongs to one leaf)
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : far_evals)
    for (size_t li = 0; li < nleaf; ++li) {
        const TNode& lf = tree[leaves[li]];
        for (size_t ii = lf.m0; ii <= lf.m1; ++ii) {
            double tempx = 0.0;
            const double p_x = px[ii];
            for (size_t fi : far[li]) {
                for (int kk = 0; kk < PP; ++kk) {
                    double s = (p_x - t1[fi * PP + kk]) / L;
                    s = s - std::round(s);
                    if (std::fabs(s - 0.5) < DBL_MIN || std::fabs(s + 0.5) < DBL_MIN ||
                        std::fabs(s) < DBL_MIN)
                        continue;
                    tempx += mom[fi * PP + kk] * -1.0 * s *
                             (1 - std::sqrt(deno / (4 * s * s * L * L +
                                                    4 * epsilon * epsilon)));
                }
                far_evals += PP;
            }
            keps[ii] += tempx;
        }
    }
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : near_evals)
    for (size_t li = 0; li < nleaf; ++li) {
        const TNode& lf = tree[leaves[li]];
        for (size_t ii = lf.m0; ii <= lf.m1; ++ii) {
            double tempx = 0.0;
            for (size_t ni : near[li]) {
                for (size_t jj = tree[ni].m0; jj <= tree[ni].m1; ++jj) {
                    double s = (px[ii] - px[jj]) / L;
                    s = s - std::round(s);
                    if (std::fabs(s - 0.5) < DBL_MIN || std::fabs(s + 0.5) < DBL_MIN ||
                        std::fabs(s) < DBL_MIN)
                        continue;
                    tempx += lambda[jj] * -1.0 * s *
                             (1 - std::sqrt(deno / (4 * s * s * L * L +
                                                    4 * epsilon * epsilon)));
                }
                near_evals += (int64_t)(tree[ni].m1 - tree[ni].m0 + 1);
            }
            keps[ii] += tempx;
        }
    }
    (void)kern;
    // un-permute (:480-486) and copy out (:278-280)
    for (size_t i = 0; i < n; ++i) es[old_index[i]] = keps[i];
    g_tc_stats[0] = (int64_t)nnode;
    g_tc_stats[1] = (int64_t)nleaf;
    g_tc_stats[2] = far_evals;
    g_tc_stats[3] = near_evals;
    return true;
}

}  // namespace orc

extern "C" {

int orc_direct_sum(double* es, const double* targets, int64_t nt,
                   const double* sources, const double* q_ws, int64_t ns,
                   double L, double epsilon) {
    if (nt < 0 || ns < 0 || !(L > 0)) return -1;
    orc::direct_sum(es, targets, nt, sources, q_ws, ns, L, epsilon);
    return 0;
}

// Extended-precision evaluation of the same mathematical sum: every kernel
// value in long double and a Neumaier-compensated long double accumulator.
int orc_direct_sum_ext(double* es, const double* targets, int64_t nt,
                       const double* sources, const double* q_ws, int64_t ns,
                       double L, double epsilon) {
    if (nt < 0 || ns < 0 || !(L > 0)) return -1;
    const long double Ll = L, el = epsilon;
    const long double c0 = 1.0L / sqrtl(Ll * Ll + 4.0L * el * el);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nt; ++i) {
        long double s = 0.0L, c = 0.0L;
        const long double xi = targets[i];
        for (int64_t j = 0; j < ns; ++j) {
            long double z = (xi - (long double)sources[j]) / Ll;
            z = z - roundl(z);
            if (z == 0.5L || z == -0.5L || z == 0.0L) continue;
            const long double term =
                (long double)q_ws[j] * -z * Ll *
                (c0 - 1.0L / sqrtl(4.0L * z * z * Ll * Ll + 4.0L * el * el));
            const long double tsum = s + term;
            if (fabsl(s) >= fabsl(term))
                c += (s - tsum) + term;
            else
                c += (term - tsum) + s;
            s = tsum;
        }
        es[i] = (double)(s + c);
    }
    return 0;
}

int orc_direct_sum_open(double* es, const double* targets, int64_t nt, const double* sources,
                        const double* q_ws, int64_t ns, double epsilon) {
    if (nt < 0 || ns < 0 || !(epsilon > 0)) return -1;
    orc::direct_sum_open(es, targets, nt, sources, q_ws, ns, epsilon);
    return 0;
}

// Extended-precision evaluation of the open-boundary sum: long double kernel
// values, Neumaier-compensated long double accumulator.
int orc_direct_sum_open_ext(double* es, const double* targets, int64_t nt, const double* sources,
                            const double* q_ws, int64_t ns, double epsilon) {
    if (nt < 0 || ns < 0 || !(epsilon > 0)) return -1;
    const long double e2 = (long double)epsilon * (long double)epsilon;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nt; ++i) {
        long double s = 0.0L, c = 0.0L;
        const long double xi = targets[i];
        for (int64_t j = 0; j < ns; ++j) {
            const long double z = xi - (long double)sources[j];
            const long double term = (long double)q_ws[j] * 0.5L * z / sqrtl(z * z + e2);
            const long double tsum = s + term;
            if (fabsl(s) >= fabsl(term))
                c += (s - tsum) + term;
            else
                c += (term - tsum) + s;
            s = tsum;
        }
        es[i] = (double)((s + c) - xi);
    }
    return 0;
}

int orc_external_field_add(double* es, const double* targets, int64_t nt, double t, int type,
                           double Am, double k, double omega) {
    if (nt < 0) return -1;
    orc::external_add(es, targets, nt, t, type, Am, k, omega);
    return 0;
}

double orc_external_amplitude(int type, double t, double Am, double k) {
    return orc::external_amplitude(type, t, Am, k);
}

int orc_treecode(double* es, const double* sources, const double* q_ws,
                 int64_t ns, double L, double epsilon, double mac, int degree,
                 int max_source) {
    if (ns <= 0 || degree < 1 || max_source < 1 || !(L > 0)) return -1;
    return orc::treecode(es, sources, q_ws, ns, L, epsilon, mac, degree, max_source) ? 0 : -2;
}

int orc_treecode_tree(const double* sources, int64_t ns, double L, double mac, int max_source,
                      int64_t stats[4], int32_t* perm, int32_t* leaf_nodes, int32_t* far_off,
                      int32_t* far_nodes, int32_t* near_off, int32_t* near_nodes) {
    if (ns <= 0 || max_source < 1 || !(L > 0)) return -1;
    orc::TTree T = orc::build_tree(sources, nullptr, (size_t)ns, L, mac, max_source);
    if (!T.ok) return -2;
    int64_t nfar = 0, nnear = 0;
    for (size_t li = 0; li < T.leaves.size(); ++li) {
        nfar += (int64_t)T.far[li].size();
        nnear += (int64_t)T.near[li].size();
    }
    stats[0] = (int64_t)T.tree.size();
    stats[1] = (int64_t)T.leaves.size();
    stats[2] = nfar;
    stats[3] = nnear;
    auto put = [&](int32_t* dst, size_t node) {
        dst[0] = (int32_t)T.tree[node].m0;
        dst[1] = (int32_t)T.tree[node].m1;
        dst[2] = T.tree[node].depth;
    };
    if (perm)
        for (int64_t i = 0; i < ns; ++i) perm[i] = T.old_index[i];
    int64_t fo = 0, no = 0;
    for (size_t li = 0; li < T.leaves.size(); ++li) {
        if (leaf_nodes) put(leaf_nodes + 3 * li, T.leaves[li]);
        if (far_off) far_off[li] = (int32_t)fo;
        if (near_off) near_off[li] = (int32_t)no;
        for (size_t node : T.far[li]) {
            if (far_nodes) put(far_nodes + 3 * fo, node);
            ++fo;
        }
        for (size_t node : T.near[li]) {
            if (near_nodes) put(near_nodes + 3 * no, node);
            ++no;
        }
    }
    if (far_off) far_off[T.leaves.size()] = (int32_t)fo;
    if (near_off) near_off[T.leaves.size()] = (int32_t)no;
    return 0;
}

void orc_treecode_last_stats(int64_t out[4]) {
    for (int i = 0; i < 4; ++i) out[i] = orc::g_tc_stats[i];
}

}  // extern "C"
