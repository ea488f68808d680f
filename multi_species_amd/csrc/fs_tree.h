// Device-built treecode structure: the 1-D bisection tree and the MAC
// interaction lists of E_MQ_Treecode (reference src/FieldStructure.cpp:
// 506-702), built from the points' stable sort by fmod(x, L) without leaving
// the device (fs_tree_build.hip). Internal to libfarrsight_hip.
#pragma once

#include <cstdint>

#include "fs_internal.h"

namespace fs {
namespace tree {

// Member points per moment chunk: a node's moments are summed chunk by
// chunk (one workgroup each), then the chunks in order.
constexpr int kMomChunk = 2048;

// A node covers the sorted points [s0, s1) and the interval [x0, x1] (xc its
// centre). Nodes are numbered level by level, the root 0; child c* and parent
// are -1 when absent, leaf is the node's index in the leaf array or -1.
struct Node {
    double x0, x1, xc;
    int s0, s1;
    int c0, c1, parent;
    int leaf;
    int depth;
    int pad;
};

// Sizes the build leaves in device memory (err: 0 ok, 1 a node of more than
// max_source coincident points, 2 node capacity exceeded).
struct Counts {
    int nnodes, nleaf, err;
    int mom_items;  // moment chunks of kMomChunk member points over all non-root nodes
    long long far_total, near_total, blk_total;
    long long far_evals, near_evals;
};

// Device views of one built tree. Tree order is the sorted order with each
// leaf's points put back in input order (the reference's stable partitions
// give that order); a node's members are the tree-order range [s0, s1 - 1].
struct Built {
    Counts c;                      // copied to the host
    const Node* nodes;             // [c.nnodes]
    const int32_t* leaves;         // [c.nleaf] node of each leaf
    const int32_t* perm;           // [n] input index of each tree-order point
    const int2* range;             // [c.nnodes] (s0, s1 - 1)
    const double* cheb;            // [c.nnodes][degree + 1] Chebyshev points
    const int32_t* far_off;        // [c.nleaf + 1] CSR over leaves
    const int32_t* far_list;
    const int32_t* near_off;
    const int32_t* near_list;
    const int4* blocks;            // [c.blk_total] (leaf, first, last, 0)
    const int32_t* mom_off;        // [c.nnodes + 1] first moment chunk of each node
    // the evaluation's work arrays: tree-order positions fmod(x, L) and
    // weights, E in tree order, moments [c.nnodes][degree + 1]
    double* px;
    double* lam;
    double* e_tree;
    double* mom;
    double* mompart;               // [c.mom_items][degree + 1] chunk partial moments
    // per node: (sum m_k, sum m_k t_k, min t_k, max t_k) of its moments m_k
    // at its Chebyshev points t_k; per leaf (at its node): (sum lambda,
    // sum lambda x, min x, max x) of its points
    double4* far_sums;
    double4* leaf_sums;
};

// Builds the tree of the n points d_x (input order) on `stream`: sort, tree,
// lists, tree order and Chebyshev points. One host synchronization (the sizes
// of the lists). work.base holds the build, work.lists the lists and the
// evaluation's work arrays.
int build_device(const double* d_x, int n, double L, double mac, int degree, int max_source,
                 TreeWork& work, Built& out, hipStream_t stream);

}  // namespace tree
}  // namespace fs
