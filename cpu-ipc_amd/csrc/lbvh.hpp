// Stackless traversal of the device LBVH (packed Node records, see common.hpp).
#pragma once
#include "common.hpp"

namespace cipc {

// A query can be split over kParts lanes.  Part p owns the subtree reached from the root by
// the kPartBits bits of p, most significant first (0 = first child); where a leaf is met
// before the bits run out, the part whose remaining bits are zero owns it and the others
// are empty.  The parts are disjoint, cover every leaf, and in part order their visits
// concatenate to the whole-tree depth-first visit.
constexpr int kPartBits = 3;
constexpr int kParts = 1 << kPartBits;

// Workgroups are dispatched to the 8 XCDs of the device round-robin.  xcd_block maps the
// physical index so that XCD k runs the k-th contiguous slice of the logical range: queries
// taken in Morton order then walk nearby subtrees out of the same L2.
constexpr int kXcds = 8;
__device__ inline int xcd_block(int b, int nb) {
    const int q = nb / kXcds, r = nb % kXcds, k = b % kXcds;
    return k * q + (k < r ? k : r) + b / kXcds;
}

struct TreeView {
    int n;
    const Node* nodes;
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
    if (t.n <= 0) return;
    if (part < 0) { lbvh_visit_range(t.n, t.nodes, 0, -1, q, gap, visit); return; }
    int node = 0;
    for (int l = kPartBits - 1; l >= 0; --l) {
        if (node >= t.n - 1) {
            if (part & ((1 << (l + 1)) - 1)) return;
            break;
        }
        node = ((part >> l) & 1) ? t.nodes[node].right : t.nodes[node].left;
    }
    lbvh_visit_range(t.n, t.nodes, node, t.nodes[node].skip, q, gap, visit);
}

}  // namespace cipc
