// Stackless traversal of the device LBVH (see context.hpp for the node layout).
#pragma once
#include "common.hpp"

namespace cipc {

// Visits, in depth-first (Morton) order, every leaf whose box overlaps `q` with `gap`
// (CPULinearBVH::query semantics, LBVH.cpp:252-282, without the final sort) and calls
// visit(primitive).  A leaf test is the same predicate the reference applies to each
// candidate afterwards, and internal boxes are exact unions, so the visited set equals
// the brute-force set {i : overlaps(q, box_i, gap)}.
template <typename F>
__device__ inline void lbvh_visit(int n, const Box* __restrict__ nb, const int* __restrict__ left,
                                  const int* __restrict__ skip, const Box& q, double gap, F&& visit) {
    if (n <= 0) return;
    // Every step moves forward in depth-first order, so 2n-1 steps visit every node once;
    // the bound and the index check keep a malformed tree from looping or reading out of
    // range.
    const int nodes = 2 * n - 1;
    int node = 0;
    for (int step = 0; node >= 0 && node < nodes && step < 2 * nodes; ++step) {
        if (overlaps(q, nb[node], gap)) {
            if (node >= n - 1) { visit(left[node]); node = skip[node]; }
            else node = left[node];
        } else {
            node = skip[node];
        }
    }
}

}  // namespace cipc
