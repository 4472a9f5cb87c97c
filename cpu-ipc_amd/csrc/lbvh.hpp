// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate expressions of CPU-IPC
// (MPL-2.0) in the reference's operation order; each such part cites the source it follows.
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
    const Node4* wide;
    const int* leafpos = nullptr;   // leaf position of each primitive (the forward walks)
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
// Root of part `part` (in [0, 2^bits)), or -1 for an empty part.
__device__ inline int part_root(const TreeView& t, int part, int bits = kPartBits) {
    if (t.n <= 0) return -1;
    int node = 0;
    for (int l = bits - 1; l >= 0; --l) {
        if (node >= t.n - 1) return (part & ((1 << (l + 1)) - 1)) ? -1 : node;
        node = ((part >> l) & 1) ? t.nodes[node].right : t.nodes[node].left;
    }
    return node;
}
// part in [0, kParts): that part's subtree; part < 0: the whole tree.  Exact.
template <typename F>
__device__ inline void lbvh_visit(const TreeView& t, int part, const Box& q, double gap, F&& visit) {
    if (t.n <= 0) return;
    if (part < 0) { lbvh_visit_range(t.n, t.nodes, 0, -1, q, gap, visit); return; }
    const int node = part_root(t, part);
    if (node >= 0) lbvh_visit_range(t.n, t.nodes, node, t.nodes[node].skip, q, gap, visit);
}

// Box test against a float box: evaluated in double on the outward-rounded bounds, so
// (rounding being monotone) it never rejects what the exact test accepts.
__device__ inline bool overlaps_f(const Box& q, const float* lo, const float* hi, double gap) {
    return !(((double)lo[0] - q.hi.x > gap) || ((double)lo[1] - q.hi.y > gap) || ((double)lo[2] - q.hi.z > gap) ||
             (q.lo.x - (double)hi[0] > gap) || (q.lo.y - (double)hi[1] > gap) || (q.lo.z - (double)hi[2] > gap));
}

// Wide walk of the subtree at binary node `start` with a per-lane stack in LDS (entry i at
// stk[64 i]): one 128-byte record per step, about half the steps of the binary walk.  Calls
// hit(primitive) for a superset of the overlapping leaves (float boxes; a single-leaf
// subtree is tested exactly).  Returns false if the stack or the step bound ran out.
constexpr int kStack = 32;
template <typename F>
__device__ inline bool wide_visit(const TreeView& t, int start, const Box& q, double gap, int* __restrict__ stk, F&& hit) {
    if (start < 0) return true;
    if (start >= t.n - 1) {
        const Node& r = t.nodes[start];
        if (overlaps(q, Box{r.lo, r.hi}, gap)) hit(r.left);
        return true;
    }
    int sp = 1;
    stk[0] = start;
    for (int steps = 0; sp > 0; ++steps) {
        if (steps >= t.n) return false;
        const int node = stk[64 * (--sp)];
        const Node4 r = t.wide[node];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int ch = r.child[k];
            if (ch == kNoChild || !overlaps_f(q, r.lo[k], r.hi[k], gap)) continue;
            if (ch < 0) {
                hit(-ch - 1);
            } else {
                if (sp >= kStack) return false;
                stk[64 * (sp++)] = ch;
            }
        }
    }
    return true;
}
// wide_visit restricted to the leaves after leaf position `after` in depth-first order:
// Node4::pad[k] holds the last leaf position under child k, so whole subtrees at or before
// `after` are skipped without a box test.  For a symmetric pair search in which every
// primitive queries the same tree, each unordered pair is then met once, from the primitive
// whose leaf comes first.
template <typename F>
__device__ inline bool wide_visit_after(const TreeView& t, int start, const Box& q, double gap, int after, int* __restrict__ stk,
                                        F&& hit) {
    if (start < 0) return true;
    if (start >= t.n - 1) {
        if (start - (t.n - 1) <= after) return true;
        const Node& r = t.nodes[start];
        if (overlaps(q, Box{r.lo, r.hi}, gap)) hit(r.left);
        return true;
    }
    int sp = 1;
    stk[0] = start;
    for (int steps = 0; sp > 0; ++steps) {
        if (steps >= t.n) return false;
        const int node = stk[64 * (--sp)];
        const Node4 r = t.wide[node];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int ch = r.child[k];
            if (ch == kNoChild || r.pad[k] <= after || !overlaps_f(q, r.lo[k], r.hi[k], gap)) continue;
            if (ch < 0) {
                hit(-ch - 1);
            } else {
                if (sp >= kStack) return false;
                stk[64 * (sp++)] = ch;
            }
        }
    }
    return true;
}

// One part of a split query: the wide walk, or, should its stack run out, reset() and the
// exact binary walk of the same subtree.  Deterministic, so a second call returns the same
// hits in the same order.
template <typename F, typename R>
__device__ inline void part_walk(const TreeView& t, int part, const Box& q, double gap, int* __restrict__ stk, F&& hit, R&& reset,
                                 int bits = kPartBits) {
    const int root = part_root(t, part, bits);
    if (root < 0) return;
    if (wide_visit(t, root, q, gap, stk, hit)) return;
    reset();
    lbvh_visit_range(t.n, t.nodes, root, t.nodes[root].skip, q, gap, hit);
}
// part_walk over the leaves after position `after` only (wide_visit_after)
template <typename F, typename R>
__device__ inline void part_walk_after(const TreeView& t, int part, const Box& q, double gap, int after, int* __restrict__ stk,
                                       F&& hit, R&& reset, int bits = kPartBits) {
    const int root = part_root(t, part, bits);
    if (root < 0) return;
    if (wide_visit_after(t, root, q, gap, after, stk, hit)) return;
    reset();
    lbvh_visit_range(t.n, t.nodes, root, t.nodes[root].skip, q, gap, [&](int prim) {
        if (t.leafpos[prim] > after) hit(prim);
    });
}

}  // namespace cipc
