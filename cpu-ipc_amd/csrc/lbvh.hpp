// Stackless traversal of the device LBVH (packed Node records, see common.hpp).
#pragma once
#include "common.hpp"

namespace cipc {

// A query can be split over kParts lanes: the frontier holds up to kParts disjoint
// subtrees that together cover every leaf, in depth-first order (the root's descendants
// down to depth log2(kParts), or a leaf where the tree is shallower), each with the node
// that follows it.  Lane `part` walks only its subtree; the lanes' visits concatenated in
// part order are the whole-tree depth-first visit.
constexpr int kParts = 8;

struct TreeView {
    int n;
    const Node* nodes;
    const int* frontier;   // [kParts roots (-1 = empty part), kParts stops]
};

// Visits, in depth-first (Morton) order, every leaf of the subtree [start, stop) whose box
// overlaps `q` with `gap` (CPULinearBVH::query semantics, LBVH.cpp:252-282, without the
// final sort) and calls visit(primitive).  Internal boxes are exact unions of their
// leaves, so the visited set equals the brute-force set {i : overlaps(q, box_i, gap)}
// restricted to the subtree.
template <typename F>
__device__ inline void lbvh_visit_range(int n, const Node* __restrict__ nd, int start, int stop, const Box& q, double gap,
                                        F&& visit) {
    if (n <= 0 || start < 0) return;
    // Every step moves forward in depth-first order, so 2n-1 steps visit every node once;
    // the bound and the index check keep a malformed tree from looping or reading out of
    // range.
    const int nodes = 2 * n - 1;
    int node = start;
    for (int step = 0; node >= 0 && node < nodes && node != stop && step < 2 * nodes; ++step) {
        const Node r = nd[node];
        if (overlaps(q, Box{r.lo, r.hi}, gap)) {
            if (node >= n - 1) { visit(r.left); node = r.skip; }
            else node = r.left;
        } else {
            node = r.skip;
        }
    }
}
// part in [0, kParts): that part's subtree; part < 0: the whole tree.
template <typename F>
__device__ inline void lbvh_visit(const TreeView& t, int part, const Box& q, double gap, F&& visit) {
    if (part < 0) lbvh_visit_range(t.n, t.nodes, 0, -1, q, gap, visit);
    else lbvh_visit_range(t.n, t.nodes, t.frontier[part], t.frontier[kParts + part], q, gap, visit);
}

}  // namespace cipc
