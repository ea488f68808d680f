// Host-side treecode structure (see fs_tree.h).
//
// E_MQ_Treecode splits a node at the midpoint of its interval, sends x <= mid
// to child 0, and stably partitions the node's points in place
// (src/FieldStructure.cpp:555-632); a node with at least max_source points is
// split (:530-545), empty children are dropped. Because every partition is
// stable, the final tree order is: leaves in depth-first (increasing x)
// order, each leaf's points in input order. This builder gets there from one
// sort by x: a node is a contiguous range of the sorted points and its split
// is an upper_bound at the midpoint.
#pragma clang fp contract(off)

#include "fs_tree.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <functional>
#include <numeric>

namespace fs {
namespace tree {

bool build(const double* px, int n, double L, double mac, int degree, int max_source, Tree& t) {
    t = Tree();
    if (n <= 0) return true;
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return px[a] < px[b]; });
    std::vector<double> sx(n);
    for (int k = 0; k < n; ++k) sx[k] = px[order[k]];
    return build_sorted(order, sx, L, mac, degree, max_source, t);
}

bool build_sorted(const std::vector<int32_t>& order, const std::vector<double>& sx, double L,
                  double mac, int degree, int max_source, Tree& t) {
    t = Tree();
    const int n = (int)sx.size();
    if (n <= 0) return true;
    // minval/maxval with their paddings (src/FieldStructure.cpp:1032-1050)
    double mn = sx[0] - 0.1234, mx = sx[n - 1] + 0.001;
    // nodes over sorted ranges [s0, s1)
    struct Pending { int node, s0, s1, depth; };
    std::vector<int> sorted_lo, sorted_hi;
    t.nodes.push_back(Node{mn, mx, 0.5 * (mn + mx), 0, 0, {-1, -1}});
    sorted_lo.push_back(0);
    sorted_hi.push_back(n);
    bool ok = true;
    std::function<void(int, int)> rec = [&](int node, int depth) {
        const int s0 = sorted_lo[node], s1 = sorted_hi[node];
        if (s1 - s0 >= max_source) {
            if (depth > 200) {
                ok = false;
                return;
            }
            const double a = t.nodes[node].x0, b = t.nodes[node].x1;
            const double mid = 0.5 * (a + b);
            const int cut = (int)(std::upper_bound(sx.begin() + s0, sx.begin() + s1, mid) - sx.begin());
            const int r0[2] = {s0, cut}, r1[2] = {cut, s1};
            for (int c = 0; c < 2; ++c) {
                if (r1[c] - r0[c] < 1) continue;
                const double c0 = (c == 0) ? a : mid, c1 = (c == 0) ? mid : b;
                t.nodes[node].child[c] = (int)t.nodes.size();
                t.nodes.push_back(Node{c0, c1, 0.5 * (c0 + c1), 0, 0, {-1, -1}});
                sorted_lo.push_back(r0[c]);
                sorted_hi.push_back(r1[c]);
            }
            for (int c = 0; c < 2; ++c)
                if (t.nodes[node].child[c] >= 0) rec(t.nodes[node].child[c], depth + 1);
        } else {
            t.leaves.push_back(node);
        }
    };
    rec(0, 0);
    if (!ok) return false;
    // tree order: leaves in depth-first order, each leaf's points in input
    // order (one pass over the input indices fills every leaf's slots)
    std::vector<int> leaf_start(t.nodes.size(), -1), leaf_end(t.nodes.size(), -1);
    std::vector<int32_t> leaf_of(n);
    int pos = 0;
    for (int leaf : t.leaves) {
        leaf_start[leaf] = pos;
        for (int k = sorted_lo[leaf]; k < sorted_hi[leaf]; ++k) leaf_of[order[k]] = leaf;
        pos += sorted_hi[leaf] - sorted_lo[leaf];
        leaf_end[leaf] = pos - 1;
    }
    t.perm.assign(n, 0);
    {
        std::vector<int> fill(leaf_start);
        for (int i = 0; i < n; ++i) t.perm[fill[leaf_of[i]]++] = i;
    }
    // member ranges of internal nodes: from their first to their last leaf
    std::function<void(int)> ranges = [&](int node) {
        Node& nd = t.nodes[node];
        if (nd.child[0] < 0 && nd.child[1] < 0) {
            nd.m0 = leaf_start[node];
            nd.m1 = leaf_end[node];
            return;
        }
        int lo = INT32_MAX, hi = -1;
        for (int c = 0; c < 2; ++c)
            if (nd.child[c] >= 0) {
                ranges(nd.child[c]);
                lo = std::min(lo, t.nodes[nd.child[c]].m0);
                hi = std::max(hi, t.nodes[nd.child[c]].m1);
            }
        nd.m0 = lo;
        nd.m1 = hi;
    };
    ranges(0);

    // interaction lists (src/FieldStructure.cpp:663-702)
    const double sq_theta = mac * mac;
    const size_t nleaf = t.leaves.size();
    std::vector<std::vector<int32_t>> far(nleaf), near(nleaf);
    std::function<void(size_t, int)> lists = [&](size_t li, int node) {
        const Node& b = t.nodes[t.leaves[li]];
        const Node& p = t.nodes[node];
        const double r_dist = std::fabs(b.xc - p.xc);
        const double dfbpx = std::fmin(r_dist, L - r_dist);
        const double dfbx = b.x1 - b.xc, dfpx = p.x1 - p.xc;
        const double sq_R = dfbpx * dfbpx, sq_rb = dfbx * dfbx, sq_rp = dfpx * dfpx;
        const bool is_leaf = p.child[0] < 0 && p.child[1] < 0;
        if (is_leaf) {
            near[li].push_back(node);
        } else if ((std::sqrt(sq_rb) + std::sqrt(sq_rp)) / std::sqrt(sq_R) < std::sqrt(sq_theta)) {
            far[li].push_back(node);
        } else {
            for (int c = 0; c < 2; ++c)
                if (p.child[c] >= 0) lists(li, p.child[c]);
        }
    };
    for (size_t li = 0; li < nleaf; ++li) lists(li, 0);
    const int PP = degree + 1;
    t.far_off.assign(1, 0);
    t.near_off.assign(1, 0);
    for (size_t li = 0; li < nleaf; ++li) {
        const Node& lf = t.nodes[t.leaves[li]];
        const int64_t cnt = lf.m1 - lf.m0 + 1;
        t.far_list.insert(t.far_list.end(), far[li].begin(), far[li].end());
        t.far_off.push_back((int32_t)t.far_list.size());
        t.near_list.insert(t.near_list.end(), near[li].begin(), near[li].end());
        t.near_off.push_back((int32_t)t.near_list.size());
        t.far_evals += cnt * (int64_t)far[li].size() * PP;
        for (int nn : near[li]) t.near_evals += cnt * (int64_t)(t.nodes[nn].m1 - t.nodes[nn].m0 + 1);
    }
    // Chebyshev points of the 2nd kind mapped to every non-root node
    // (src/FieldStructure.cpp:792-823)
    const double pi = 3.14159265358979323846;
    const double h = pi / degree;
    std::vector<double> tk(PP);
    for (int i = 0; i < PP; ++i) tk[i] = std::cos(i * h);
    t.cheb.assign(t.nodes.size() * PP, 0.0);
    for (size_t node = 1; node < t.nodes.size(); ++node) {
        const double x1 = t.nodes[node].x0, x2 = t.nodes[node].x1;
        for (int j = 0; j < PP; ++j) t.cheb[node * PP + j] = x1 + (tk[j] + 1.0) * 0.5 * (x2 - x1);
    }
    return true;
}

}  // namespace tree
}  // namespace fs
