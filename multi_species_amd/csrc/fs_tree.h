// Host-side treecode structure: the 1-D bisection tree, the MAC interaction
// lists and the Chebyshev points of E_MQ_Treecode (reference
// src/FieldStructure.cpp:384-866). Internal to libfarrsight_hip.
#pragma once

#include <cstdint>
#include <vector>

namespace fs {
namespace tree {

struct Node {
    double x0, x1, xc;  // interval and centre
    int m0, m1;         // members [m0, m1] in tree order
    int child[2];       // -1 when absent
};

struct Tree {
    std::vector<Node> nodes;          // depth-first creation order, root first
    std::vector<int32_t> leaves;      // leaf node ids in depth-first order
    std::vector<int32_t> perm;        // perm[k] = input index of the k-th point in tree order
    std::vector<int32_t> far_off, far_list;    // CSR over leaves
    std::vector<int32_t> near_off, near_list;  // CSR over leaves
    std::vector<double> cheb;         // [node][degree+1] Chebyshev points (root row unused)
    int64_t far_evals = 0, near_evals = 0;
};

// Builds the tree of the points px (already reduced by fmod(x, L) like
// src/FieldStructure.cpp:240-247). Returns false if a node cannot be split
// (more than max_source identical coordinates).
bool build(const double* px, int n, double L, double mac, int degree, int max_source, Tree& t);

// The same from the points already in stable increasing order: order[k] is
// the input index of the k-th smallest, sx[k] its coordinate.
bool build_sorted(const std::vector<int32_t>& order, const std::vector<double>& sx, double L,
                  double mac, int degree, int max_source, Tree& t);

}  // namespace tree
}  // namespace fs
