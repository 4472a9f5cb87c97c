// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate expressions of CPU-IPC
// (MPL-2.0) in the reference's operation order; each such part cites the source it follows.
// Broad phase and active set on the device.  Compiled with -ffp-contract=off.
//
// Reference: CPULinearBVH::build / query and CPUBroadPhaseBVH::build / buildSwept
// (CPU IPC/LBVH.cpp:163-339), SpatialHash::calculateActivateSet
// (CollisionBroadPhase.cpp:765-1059), Ground::calculateActivateSet (:1076-1092).
//
// Tree topology: scene box reduction -> 30-bit Morton codes of box centres normalised to
// the scene box -> stable radix sort of (code, primitive) -> unique keys code << 32 |
// primitive -> Karras splits, one thread per internal node -> skip pointers for stackless
// depth-first traversal.  Refit (every build): leaf boxes in leaf order -> internal boxes
// as unions of power-of-two leaf ranges (no inter-work-group synchronisation).  Query
// results do not depend on the topology, so it is rebuilt only every kMaxRefits fits.
//
// Active set: each surface vertex (point-triangle) and surface edge (edge-edge) query is
// split over kParts lanes, one subtree each, that buffer their first candidates; a second
// pass per query packs them (re-walking only overflowing subtrees) and sorts them by
// primitive id as the reference's query does; one thread per candidate classifies it.
// Compaction reproduces the reference's serial order: partial-CCD pairs, then Self_ActiveSet =
// point-triangle entries, edge-edge entries, and the point-point/point-edge entries
// deduplicated in std::map key order with multiplicity; Self_EE_ActiveSet for mollified
// entries; Environment_ActiveSet in ascending surface order.
#include "context.hpp"
#include "contact.hpp"
#include "lbvh.hpp"

#include <hipcub/hipcub.hpp>
#include <type_traits>

namespace cipc {

// ---------------------------------------------------------------------------------
// LBVH build
// ---------------------------------------------------------------------------------
__device__ inline unsigned expand_bits(unsigned v) {       // LBVH.cpp:17-24
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
__device__ inline unsigned morton3(double x, double y, double z) {  // LBVH.cpp:26-37
    const double res = 1024.0;
    auto q = [&](double v) { v = fmax(0.0, fmin(v * res, res - 1.0)); return (unsigned)v; };
    return (expand_bits(q(x)) << 2) | (expand_bits(q(y)) << 1) | expand_bits(q(z));
}

constexpr int kRedBlock = 256;
constexpr int kRedParts = 256;
__global__ __launch_bounds__(kRedBlock) void box_reduce_partial(int n, const Box* __restrict__ b, Box* __restrict__ part) {
    __shared__ Box sh[kRedBlock];
    Box acc{{1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308},
            {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308}};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        acc = Box{vmin(acc.lo, b[i].lo), vmax(acc.hi, b[i].hi)};
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            sh[threadIdx.x] = Box{vmin(sh[threadIdx.x].lo, sh[threadIdx.x + s].lo), vmax(sh[threadIdx.x].hi, sh[threadIdx.x + s].hi)};
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void morton_keys(int n, const Box* __restrict__ b, const Box* __restrict__ part, int nparts,
                            unsigned* __restrict__ codes, int* __restrict__ ids) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Box sc = part[0];
    for (int k = 1; k < nparts; ++k) sc = Box{vmin(sc.lo, part[k].lo), vmax(sc.hi, part[k].hi)};
    const d3 ext = sc.hi - sc.lo;  // nparts == 1: the partials were reduced beforehand
    const d3 c = 0.5 * (b[i].lo + b[i].hi);
    const d3 off = c - sc.lo;
    const double x = ext.x > 0.0 ? off.x / ext.x : 0.0;
    const double y = ext.y > 0.0 ? off.y / ext.y : 0.0;
    const double z = ext.z > 0.0 ? off.z / ext.z : 0.0;
    codes[i] = morton3(x, y, z);
    ids[i] = i;
}
// Unique keys (code << 32 | primitive).  The radix sort of (code, primitive) pairs is stable
// and the primitives enter in ascending order, so these keys come out sorted exactly as a
// sort of the 64-bit keys would order them.
__global__ void join_keys(int n, const unsigned* __restrict__ codes, const int* __restrict__ ids, unsigned long long* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = ((unsigned long long)codes[i] << 32) | (unsigned)ids[i];
}
__device__ inline int prefix(const unsigned long long* k, int n, int a, int b) {
    if (b < 0 || b >= n) return -1;
    return __clzll(k[a] ^ k[b]);
}
// Karras 2012 topology (LBVH.cpp:62-119,213-228): internal node i covers the key range
// [first, last] and splits it at the highest differing bit.
__global__ void karras_kernel(int n, const unsigned long long* __restrict__ k, Node* __restrict__ nodes, int* __restrict__ range_first, int* __restrict__ range_last, int* __restrict__ split_right) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int first, last;
    if (i == 0) { first = 0; last = n - 1; }
    else {
        const int lp = prefix(k, n, i, i - 1), rp = prefix(k, n, i, i + 1);
        const int dir = rp > lp ? 1 : -1;
        const int minp = lp < rp ? lp : rp;
        int lmax = 2;
        while (prefix(k, n, i, i + dir * lmax) > minp) lmax <<= 1;
        int len = 0;
        for (int st = lmax >> 1; st > 0; st >>= 1)
            if (prefix(k, n, i, i + dir * (len + st)) > minp) len += st;
        const int other = i + dir * len;
        first = i < other ? i : other;
        last = i < other ? other : i;
    }
    int split;
    const unsigned long long fc = k[first], lc = k[last];
    if (fc == lc) split = (first + last) >> 1;
    else {
        const int common = __clzll(fc ^ lc);
        int s = first, step = last - first;
        do {
            step = (step + 1) >> 1;
            const int cand = s + step;
            if (cand < last && __clzll(fc ^ k[cand]) > common) s = cand;
        } while (step > 1);
        split = s;
    }
    const int l = (split == first) ? split + n - 1 : split;
    const int r = (split + 1 == last) ? split + 1 + n - 1 : split + 1;
    nodes[i].left = l;
    nodes[i].right = r;
    range_first[i] = first;
    range_last[i] = last;
    split_right[split] = r;
}
__global__ void leaf_ids(int n, const int* __restrict__ order, Node* __restrict__ nodes, int* __restrict__ leafpos) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    nodes[n - 1 + s].left = order[s];
    leafpos[order[s]] = s;
}
__global__ void skip_kernel(int n, const int* __restrict__ range_last, const int* __restrict__ split_right, Node* __restrict__ nodes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * n - 1) return;
    const int last = (i < n - 1) ? range_last[i] : i - (n - 1);
    nodes[i].skip = (last >= n - 1) ? -1 : split_right[last];
}
// Internal boxes.  Node i of the Karras tree covers the sorted leaves [first, last], so its
// box is the union of that range.  Power-of-two reduction levels over the sorted leaf boxes
// (level l node j = leaves [j 2^l, (j+1) 2^l)) are built 8 levels per launch in LDS; each
// internal box is then the union of at most 2 log2(n) level nodes.  No inter-work-group
// communication is needed.
constexpr int kSegBlock = 256;
constexpr int kSegLevels = 8;   // log2(kSegBlock)
__device__ inline Box empty_box() {
    return Box{{1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308},
               {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308}};
}
// Primitive boxes: static (CPUBroadPhaseBVH::build, LBVH.cpp:284-305) or swept along -p
// with the step whose bits are at *abits (buildSwept, LBVH.cpp:307-339).  The step is read
// on the device, so building the swept trees needs no host round trip.
__device__ inline double step_of(const unsigned long long* abits) { return __longlong_as_double((long long)*abits); }
// Where a tree's primitive boxes come from: a face or an edge of the surface, static (the
// box of the positions) or swept (the box of x and x - a p, a the scene's step in abits).
struct PrimSrc {
    int kind = 0;   // 0: the prim array holds them; 1: faces; 2: surface edges
    const int4* faces = nullptr;
    const int2* se = nullptr;
    const double* x = nullptr;
    const double* p = nullptr;
    const unsigned long long* abits = nullptr;
    const int* vs = nullptr;
};
__device__ inline Box prim_box(const PrimSrc& s, int i) {
    if (s.kind == 1) {
        const int4 v = s.faces[i];
        if (s.abits == nullptr) return box_tri(ld3(s.x, v.x), ld3(s.x, v.y), ld3(s.x, v.z));
        const double a = step_of(s.vs ? s.abits + kSlots * s.vs[v.x] : s.abits);
        Box b = swept_pt(ld3(s.x, v.x), ld3(s.p, v.x), a);
        b = box_union(b, swept_pt(ld3(s.x, v.y), ld3(s.p, v.y), a));
        b = box_union(b, swept_pt(ld3(s.x, v.z), ld3(s.p, v.z), a));
        return b;
    }
    const int2 v = s.se[i];
    if (s.abits == nullptr) return box_edge(ld3(s.x, v.x), ld3(s.x, v.y));
    const double a = step_of(s.vs ? s.abits + kSlots * s.vs[v.x] : s.abits);
    return box_union(swept_pt(ld3(s.x, v.x), ld3(s.p, v.x), a), swept_pt(ld3(s.x, v.y), ld3(s.p, v.y), a));
}

__global__ void prim_boxes(int n, PrimSrc src, Box* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = prim_box(src, i);
}

// Box of the sorted leaves [a, b]: the union of at most 2 log2(n) level nodes (exact: the
// union is a min / max).
__device__ inline Box range_box(int a, int b, const SegLevels& L, const Box* __restrict__ seg) {
    int l = 0;
    Box acc = empty_box();
    while (a <= b && l < 40) {
        if (a & 1) { const Box x = seg[L.offset[l] + a]; acc = Box{vmin(acc.lo, x.lo), vmax(acc.hi, x.hi)}; ++a; }
        if (a <= b && !(b & 1)) { const Box x = seg[L.offset[l] + b]; acc = Box{vmin(acc.lo, x.lo), vmax(acc.hi, x.hi)}; --b; }
        a >>= 1; b >>= 1; ++l;
    }
    return acc;
}
// A box into its slot of a four-wide record: slot code (record << 2 | k), or -1 for none.
__device__ inline void put_wide_box(Node4* __restrict__ wide, int code, const Box& x) {
    if (code < 0) return;
    Node4& r = wide[code >> 2];
    const int k = code & 3;
    r.lo[k][0] = __double2float_rd(x.lo.x); r.lo[k][1] = __double2float_rd(x.lo.y); r.lo[k][2] = __double2float_rd(x.lo.z);
    r.hi[k][0] = __double2float_ru(x.hi.x); r.hi[k][1] = __double2float_ru(x.hi.y); r.hi[k][2] = __double2float_ru(x.hi.z);
}
// A tree's refit, for one tree or two at once (a work-group range each):
//   leaves: the leaf boxes in leaf order (node records and level 0) and reduction levels
//     1..kSegLevels, each work-group over its kSegBlock leaves; with a primitive source the
//     primitive boxes are computed here too (and stored, for the exact leaf tests).
//   upper: the next kSegLevels levels from the last built, one launch per kSegLevels
//     levels (one for up to 2^16 leaves).
//   internal: each internal node's box, the union of the level boxes that cover its leaf
//     range (range_box).
// No work-group waits for another inside a launch: on this GPU that would take a device-wide
// fence, which writes back and invalidates an XCD's L2.
// Each box also goes, rounded outward to float, into the four-wide record slots that hold
// its node (wslot, set with the topology), so the records need no launch of their own.
struct RefitTree {
    int n = 0, levels = 0, leaf_blocks = 0, int_blocks = 0;
    const int* order = nullptr;
    Box* prim = nullptr;
    PrimSrc src;
    Node* nodes = nullptr;
    Node4* wide = nullptr;
    const int* first = nullptr;
    const int* last = nullptr;
    Box* seg = nullptr;
    const int2* wslot = nullptr;
    SegLevels L;
};
struct RefitSet {
    RefitTree t[2];
    int ntrees = 1;
};
__global__ __launch_bounds__(kSegBlock) void refit_leaves(RefitSet rs) {
    const int which = (rs.ntrees > 1 && (int)blockIdx.x >= rs.t[0].leaf_blocks) ? 1 : 0;
    const RefitTree& t = rs.t[which];
    const int blk = (int)blockIdx.x - (which ? rs.t[0].leaf_blocks : 0);
    __shared__ Box sh[kSegBlock];
    const int n = t.n;
    const int g = blk * kSegBlock + threadIdx.x;
    Box b = empty_box();
    if (g < n) {
        const int q = t.order[g];
        if (t.src.kind != 0) {
            b = prim_box(t.src, q);
            t.prim[q] = b;
        } else {
            b = t.prim[q];
        }
        Node& r = t.nodes[n - 1 + g];
        r.lo = b.lo;
        r.hi = b.hi;
        t.seg[g] = b;
        const int2 ws = t.wslot[n - 1 + g];
        put_wide_box(t.wide, ws.x, b);
        put_wide_box(t.wide, ws.y, b);
    }
    sh[threadIdx.x] = b;
    __syncthreads();
    for (int l = 1; l <= kSegLevels; ++l) {
        const int half = 1 << (l - 1);
        if ((threadIdx.x & ((1 << l) - 1)) == 0) {
            const Box a = sh[threadIdx.x], c = sh[threadIdx.x + half];
            sh[threadIdx.x] = Box{vmin(a.lo, c.lo), vmax(a.hi, c.hi)};
        }
        __syncthreads();
        if (l < 40 && t.L.count[l] > 0 && (threadIdx.x & ((1 << l) - 1)) == 0) {
            const int j = g >> l;
            if (j < t.L.count[l]) t.seg[t.L.offset[l] + j] = sh[threadIdx.x];
        }
    }
}
// Levels l0+1 .. l0+kSegLevels from level l0, for every tree of the set that has them: the
// trees' work-groups one after the other (blocks_for(count[l0], kSegBlock) each).
__device__ inline int upper_blocks(const RefitTree& t, int l0) {
    return t.levels > l0 ? (t.L.count[l0] + kSegBlock - 1) / kSegBlock : 0;
}
__global__ __launch_bounds__(kSegBlock) void refit_upper(RefitSet rs, int l0) {
    const int b0 = upper_blocks(rs.t[0], l0);
    const int which = (rs.ntrees > 1 && (int)blockIdx.x >= b0) ? 1 : 0;
    const RefitTree& t = rs.t[which];
    const int blk = (int)blockIdx.x - (which ? b0 : 0);
    __shared__ Box sh[kSegBlock];
    const int nsrc = t.L.count[l0];
    const Box* src = t.seg + t.L.offset[l0];
    const int g = blk * kSegBlock + threadIdx.x;
    sh[threadIdx.x] = g < nsrc ? src[g] : empty_box();
    __syncthreads();
    for (int l = 1; l <= kSegLevels; ++l) {
        const int half = 1 << (l - 1);
        if ((threadIdx.x & ((1 << l) - 1)) == 0) {
            const Box a = sh[threadIdx.x], c = sh[threadIdx.x + half];
            sh[threadIdx.x] = Box{vmin(a.lo, c.lo), vmax(a.hi, c.hi)};
        }
        __syncthreads();
        const int lv = l0 + l;
        if (lv < 40 && t.L.count[lv] > 0 && (threadIdx.x & ((1 << l) - 1)) == 0) {
            const int j = g >> l;
            if (j < t.L.count[lv]) t.seg[t.L.offset[lv] + j] = sh[threadIdx.x];
        }
    }
}
__global__ void refit_internal(RefitSet rs) {
    const int which = (rs.ntrees > 1 && (int)blockIdx.x >= rs.t[0].int_blocks) ? 1 : 0;
    const RefitTree& t = rs.t[which];
    const int blk = (int)blockIdx.x - (which ? rs.t[0].int_blocks : 0);
    const int i = blk * blockDim.x + threadIdx.x;
    if (i >= t.n - 1) return;
    const Box own = range_box(t.first[i], t.last[i], t.L, t.seg);
    t.nodes[i].lo = own.lo;
    t.nodes[i].hi = own.hi;
    put_wide_box(t.wide, t.wslot[i].y, own);
}

// The parts of the four-wide records that depend on the topology only, once per topology:
// what each slot holds and, for every node, the slots its box goes to (wslot.x: the record
// of its parent, for a leaf child; wslot.y: the record of its grandparent).  A record holds
// its node's grandchildren, or the children themselves where those are leaves; unused slots
// hold an empty box.  pad[k]: the last leaf position under child k (wide_visit_after).
__global__ void fill_int2(int2* __restrict__ p, int n, int2 v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void wide_static(int n, const Node* __restrict__ nodes, const int* __restrict__ range_last, Node4* __restrict__ wide,
                            int2* __restrict__ wslot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int ch[4], via[4], m = 0;
    const int c2[2] = {nodes[i].left, nodes[i].right};
    for (int k = 0; k < 2; ++k) {
        const int c = c2[k];
        if (c < n - 1) {
            ch[m] = nodes[c].left; via[m++] = 1;
            ch[m] = nodes[c].right; via[m++] = 1;
        } else {
            ch[m] = c; via[m++] = 0;
        }
    }
    Node4 r;
    for (int k = 0; k < 4; ++k) {
        r.lo[k][0] = r.lo[k][1] = r.lo[k][2] = __int_as_float(0x7f800000);
        r.hi[k][0] = r.hi[k][1] = r.hi[k][2] = __int_as_float(0xff800000);
        if (k < m) {
            const int c = ch[k];
            r.child[k] = c >= n - 1 ? -(nodes[c].left + 1) : c;
            r.pad[k] = c >= n - 1 ? c - (n - 1) : range_last[c];
            if (via[k]) wslot[c].y = i << 2 | k;
            else wslot[c].x = i << 2 | k;
        } else {
            r.child[k] = kNoChild;
            r.pad[k] = -1;
        }
    }
    wide[i] = r;
}

// Topology from the current primitive boxes.
void build_topology(Context& c, LBVH& t, int n, hipStream_t st, DBuf<unsigned char>& tmp) {
    ++c.topo_gen;
    t.n = n;
    t.age = 0;
    const int nodes = 2 * n - 1;
    t.nodes.resize(nodes); t.wide.resize(std::max(n - 1, 1));
    t.range_first.resize(std::max(n - 1, 1)); t.range_last.resize(std::max(n - 1, 1)); t.split_right.resize(std::max(n - 1, 1));
    t.codes.resize(n); t.codes_sorted.resize(n); t.order.resize(2 * (size_t)n); t.keys_sorted.resize(n);
    // level l of the box reduction holds ceil(n / 2^l) boxes; level 0 is the leaves
    SegLevels L{};
    long long total = 0;
    int levels = 0;
    for (int l = 0; l < 40; ++l) {
        const long long cnt = ((long long)n + (1ll << l) - 1) >> l;
        L.count[l] = (int)cnt;
        L.offset[l] = total;
        total += cnt;
        levels = l;
        if (cnt == 1) break;
    }
    for (int l = levels + 1; l < 40; ++l) { L.count[l] = 0; L.offset[l] = total; }
    t.L = L;
    t.levels = levels;
    t.seg.resize((size_t)total);
    const int parts = std::min<int>(kRedParts, (int)blocks_for(n, kRedBlock));
    t.scene_partial.resize(parts);
    hipLaunchKernelGGL(box_reduce_partial, dim3(parts), dim3(kRedBlock), 0, st, n, t.prim.p, t.scene_partial.p);
    CIPC_CHECK_LAUNCH();
    if (parts > 1) {
        hipLaunchKernelGGL(box_reduce_partial, dim3(1), dim3(kRedBlock), 0, st, parts, t.scene_partial.p, t.scene_partial.p);
        CIPC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(morton_keys, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, t.prim.p, t.scene_partial.p, 1, t.codes.p,
                       t.order.p + n);
    CIPC_CHECK_LAUNCH();
    sort_u32_pairs(t.codes.p, t.codes_sorted.p, t.order.p + n, t.order.p, n, 30, st, tmp);
    hipLaunchKernelGGL(join_keys, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, t.codes_sorted.p, t.order.p, t.keys_sorted.p);
    CIPC_CHECK_LAUNCH();
    if (n > 1) {
        hipLaunchKernelGGL(karras_kernel, dim3(blocks_for(n - 1, 256)), dim3(256), 0, st, n, t.keys_sorted.p, t.nodes.p,
                           t.range_first.p, t.range_last.p, t.split_right.p);
        CIPC_CHECK_LAUNCH();
    }
    t.leafpos.resize(n);
    hipLaunchKernelGGL(leaf_ids, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, t.order.p, t.nodes.p, t.leafpos.p);
    hipLaunchKernelGGL(skip_kernel, dim3(blocks_for(nodes, 256)), dim3(256), 0, st, n, t.range_last.p, t.split_right.p, t.nodes.p);
    t.wslot.resize(nodes);
    hipLaunchKernelGGL(fill_int2, dim3(blocks_for(nodes, 256)), dim3(256), 0, st, t.wslot.p, nodes, make_int2(-1, -1));
    if (n > 1)
        hipLaunchKernelGGL(wide_static, dim3(blocks_for(n - 1, 256)), dim3(256), 0, st, n, t.nodes.p, t.range_last.p, t.wide.p,
                           t.wslot.p);
    CIPC_CHECK_LAUNCH();
}

// Boxes of the current primitive boxes on the existing topologies: one tree, or two in the
// same two launches.
static RefitTree refit_view(LBVH& t, const PrimSrc& src) {
    RefitTree r;
    r.n = t.n; r.levels = t.levels;
    r.leaf_blocks = (int)blocks_for(t.n, kSegBlock);
    r.int_blocks = t.n > 1 ? (int)blocks_for(t.n - 1, 256) : 0;
    r.order = t.order.p; r.prim = t.prim.p; r.src = src; r.nodes = t.nodes.p; r.wide = t.wide.p;
    r.first = t.range_first.p; r.last = t.range_last.p; r.seg = t.seg.p; r.wslot = t.wslot.p; r.L = t.L;
    return r;
}
static void refit_set(const RefitSet& rs, hipStream_t st) {
    int lb = 0, ib = 0, levels = 0;
    for (int k = 0; k < rs.ntrees; ++k) {
        lb += rs.t[k].leaf_blocks;
        ib += rs.t[k].int_blocks;
        levels = std::max(levels, rs.t[k].levels);
    }
    if (lb > 0) hipLaunchKernelGGL(refit_leaves, dim3(lb), dim3(kSegBlock), 0, st, rs);
    for (int l0 = kSegLevels; l0 < levels; l0 += kSegLevels) {
        int ub = 0;
        for (int k = 0; k < rs.ntrees; ++k) ub += rs.t[k].levels > l0 ? (int)blocks_for(rs.t[k].L.count[l0], kSegBlock) : 0;
        if (ub > 0) hipLaunchKernelGGL(refit_upper, dim3(ub), dim3(kSegBlock), 0, st, rs, l0);
    }
    if (ib > 0) hipLaunchKernelGGL(refit_internal, dim3(ib), dim3(256), 0, st, rs);
    CIPC_CHECK_LAUNCH();
}
void refit(Context& c, LBVH& t, const PrimSrc& src, hipStream_t st) {
    (void)c;
    RefitSet rs;
    rs.t[0] = refit_view(t, src);
    rs.ntrees = 1;
    refit_set(rs, st);
    ++t.age;
}
static void refit_two(LBVH& a, const PrimSrc& sa, LBVH& b, const PrimSrc& sb, hipStream_t st) {
    RefitSet rs;
    rs.t[0] = refit_view(a, sa);
    rs.t[1] = refit_view(b, sb);
    rs.ntrees = 2;
    refit_set(rs, st);
    ++a.age;
    ++b.age;
}

// The tree over n > 0 primitives has a topology that the next build refits.
static bool refit_only(const LBVH& t, int n) { return n > 0 && t.n == n && t.age < kMaxRefits; }
// A tree over n primitives whose boxes come from src: a new topology (from boxes computed
// first) when none fits, else a refit that computes the boxes itself.
static void build_tree_from(Context& c, LBVH& t, int n, const PrimSrc& src, hipStream_t st, DBuf<unsigned char>& tmp) {
    if (n <= 0) { t.n = n; return; }
    t.prim.resize(n);
    if (t.n != n || t.age >= kMaxRefits) {
        hipLaunchKernelGGL(prim_boxes, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, src, t.prim.p);
        CIPC_CHECK_LAUNCH();
        build_topology(c, t, n, st, tmp);
        refit(c, t, PrimSrc{}, st);
    } else {
        refit(c, t, src, st);
    }
}
void build_tree(Context& c, LBVH& t, int n) {
    if (n <= 0) { t.n = n; return; }
    if (t.n != n || t.age >= kMaxRefits) build_topology(c, t, n, c.stream, c.tmp);
    refit(c, t, PrimSrc{}, c.stream);
}

// The static face and edge trees, on the context's stream.
void build_static_trees(Context& c) {
    Scene& s = c.sc;
    c.face_tree.prim.resize(std::max(s.F, 1));
    c.edge_tree.prim.resize(std::max(s.Es, 1));
    PrimSrc fs, es;
    fs.kind = 1; fs.faces = s.faces.p; fs.x = s.x.p;
    es.kind = 2; es.se = s.se.p; es.x = s.x.p;
    build_tree_from(c, c.face_tree, s.F, fs, c.stream, c.tmp);
    build_tree_from(c, c.edge_tree, s.Es, es, c.stream, c.tmp);
}
void build_swept_trees(Context& c, const unsigned long long* abits, hipStream_t st) {
    Scene& s = c.sc;
    c.sface_tree.prim.resize(std::max(s.F, 1));
    c.sedge_tree.prim.resize(std::max(s.Es, 1));
    PrimSrc fs, es;
    fs.kind = 1; fs.faces = s.faces.p; fs.x = s.x.p; fs.p = s.dir.p; fs.abits = abits; fs.vs = scene_of_vertex(s);
    es.kind = 2; es.se = s.se.p; es.x = s.x.p; es.p = s.dir.p; es.abits = abits; es.vs = scene_of_vertex(s);
    if (refit_only(c.sface_tree, s.F) && refit_only(c.sedge_tree, s.Es)) {
        refit_two(c.sface_tree, fs, c.sedge_tree, es, st);   // both trees in the same two launches
        return;
    }
    build_tree_from(c, c.sface_tree, s.F, fs, st, c.tmp_sb);
    build_tree_from(c, c.sedge_tree, s.Es, es, st, c.tmp_sb);
}

// ---------------------------------------------------------------------------------
// Active set (CollisionBroadPhase.cpp:765-1059)
// ---------------------------------------------------------------------------------
// candidate categories; kDrop: a box hit that the reference's candidate filters discard
enum : int { kNone = 0, kSelfPT = 1, kMap = 2, kSelfEE = 3, kMolPE = 4, kMolEE = 5, kDrop = 6 };

// Candidate collection.
//   A (traverse): each query is split over kParts lanes, one subtree each; a lane records
//     the first kPartCap leaves whose box overlaps the query box with gap sqrt(Hhat) (for
//     edges only later edges, as queryEdgeForEdgesWithBBoxCheck keeps) and its count; the
//     lanes of a query add their counts.  The walk loads nothing but tree nodes.
//   B (compact, one thread per query): the parts in order into the packed layout,
//     re-walking only parts with more than kPartCap hits; sort by primitive id.
//   C (classify, one thread per hit): the reference's incidence and collider filters
//     (CollisionBroadPhase.cpp:769-776,862-870), then the distance-type classification.
// The leaf boxes are the primitives' own boxes (prim_box), so the leaf test
// is the box test of the reference's candidate filter.  Queries are taken in the Morton
// order of their rest positions (perm) so that the lanes of a wave walk nearby subtrees.
constexpr int kPartCap = 32;
// The split (2^PB lanes per query, PB = 0..4) is a template argument, chosen per context by
// the autotuner; the hits and their order do not depend on it (they are ranked by primitive).
constexpr int kMaxSplitBits = 4;
static_assert(64 % (1 << kMaxSplitBits) == 0, "a query's lanes must share a wavefront");

template <int PB>
__device__ inline int sum_over_parts(int v) {
#pragma unroll
    for (int m = 1; m < (1 << PB); m <<= 1) v += __shfl_xor(v, m, 1 << PB);
    return v;
}

// The point walk also sets each surface vertex's ground flag (distance^2 to the ground below
// Hhat, Ground::calculateActivateSet, CollisionBroadPhase.cpp:1076-1092), for the same scan.
template <int PB>
__device__ inline void pt_traverse_at(int t, int* stk, int Vs, const int* __restrict__ perm, const int* __restrict__ sv,
                                      const double* __restrict__ x, TreeView tv, double sq, int* __restrict__ cnt,
                                      int* __restrict__ cntp, int* __restrict__ buf, int* __restrict__ fg, d3 gn, double gd,
                                      double hhat) {
    constexpr int P = 1 << PB;
    const bool active = t < Vs * P;
    const int s = active ? perm[t >> PB] : 0, part = t & (P - 1);
    int n = 0;
    if (active) {
        const Box qb = box_pt(ld3(x, sv[s]));
        int* b = buf + ((size_t)s * P + part) * kPartCap;
        part_walk(tv, part, qb, sq, stk + threadIdx.x, [&](int f) { if (n < kPartCap) b[n] = f; ++n; }, [&] { n = 0; }, PB);
        cntp[s * P + part] = n;
    }
    const int tot = sum_over_parts<PB>(n);
    if (active && part == 0) {
        cnt[s] = tot;
        const double dist = dot(gn, ld3(x, sv[s])) - gd;
        fg[s] = (dist * dist < hhat) ? 1 : 0;
    }
}
template <int PB>
__global__ __launch_bounds__(64) void pt_traverse(int Vs, const int* __restrict__ perm, const int* __restrict__ sv,
                                                   const double* __restrict__ x, TreeView tv, double sq, int* __restrict__ cnt,
                                                   int* __restrict__ cntp, int* __restrict__ buf, int* __restrict__ fg, d3 gn,
                                                   double gd, double hhat) {
    __shared__ int stk[kStack * 64];
    pt_traverse_at<PB>(xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x, stk, Vs, perm, sv, x, tv, sq, cnt, cntp, buf,
                       fg, gn, gd, hhat);
    if (blockIdx.x == 0 && threadIdx.x == 0) { cnt[Vs] = 0; fg[Vs] = 0; }   // the scans' tails
}

template <int PB>
__device__ inline void pt_compact_at(int t, int* stk, int Vs, const int* __restrict__ perm, const int* __restrict__ sv,
                                     const double* __restrict__ x, TreeView tv, double sq, const int* __restrict__ off,
                                     const int* __restrict__ cntp, const int* __restrict__ buf, int* __restrict__ candU,
                                     int* __restrict__ own, const int* __restrict__ dc) {
    constexpr int P = 1 << PB;
    if (t >= Vs || dc[kDcFlags]) return;   // over capacity: the host builds the set again
    const int s = perm[t];
    const int o = off[s];
    int pos = o;
    for (int part = 0; part < P; ++part) {
        const int c = cntp[s * P + part];
        if (c <= kPartCap) {
            for (int k = 0; k < c; ++k) candU[pos + k] = buf[((size_t)s * P + part) * kPartCap + k];
        } else {
            const Box qb = box_pt(ld3(x, sv[s]));
            int m = 0;
            part_walk(tv, part, qb, sq, stk + threadIdx.x, [&](int f) { if (m < c) candU[pos + m] = f; ++m; }, [&] { m = 0; }, PB);
        }
        pos += c;
    }
    for (int k = o; k < pos; ++k) own[k] = s;
}

template <int PB>
__global__ __launch_bounds__(64) void pt_compact(int Vs, const int* __restrict__ perm, const int* __restrict__ sv, const double* __restrict__ x, TreeView tv,
                           double sq, const int* __restrict__ off, const int* __restrict__ cntp, const int* __restrict__ buf,
                           int* __restrict__ candU, int* __restrict__ own, const int* __restrict__ dc) {
    __shared__ int stk[kStack * 64];
    pt_compact_at<PB>(blockIdx.x * blockDim.x + threadIdx.x, stk, Vs, perm, sv, x, tv, sq, off, cntp, buf, candU, own, dc);
}

// Point-triangle hits: filters, then classification (CollisionBroadPhase.cpp:765-857).
__device__ inline void pt_classify(int u, const int* __restrict__ dc, int F, const int* __restrict__ own, const int* __restrict__ off,
                                   const int* __restrict__ sv, const double* __restrict__ x, const int* __restrict__ bt,
                                   const int4* __restrict__ faces, const int* __restrict__ candU, int* __restrict__ cand, double hhat,
                                   double sq, const int* __restrict__ vs, int4* __restrict__ rec, int* __restrict__ cat) {
    if (dc[kDcFlags] || u >= dc[kDcPT]) return;
    // sorted position: the hit's rank among its query's hits (all distinct)
    const int q = own[u], o = off[q], e = off[q + 1];
    const int f = candU[u];
    int k = o;
    for (int j = o; j < e; ++j) k += candU[j] < f;
    cand[k] = f;
    if ((unsigned)f >= (unsigned)F || k >= e) {   // cannot happen for a consistent walk; never index with it
        rec[k < e ? k : o] = make_int4(0, 0, 0, 0);
        cat[k < e ? k : o] = kDrop;
        return;
    }
    const int vI = sv[q];
    const int4 fv = faces[f];
    if ((is_collider(bt[vI]) && is_collider(bt[fv.x]) && is_collider(bt[fv.y]) && is_collider(bt[fv.z])) || vI == fv.x ||
        vI == fv.y || vI == fv.z || (vs && vs[vI] != vs[fv.x])) {   // (scenes of a batch never touch)
        rec[k] = make_int4(0, 0, 0, 0);
        cat[k] = kDrop;
        return;
    }
    const d3 P = ld3(x, vI);
    {
        const d3 A = ld3(x, fv.x), B = ld3(x, fv.y), C = ld3(x, fv.z);
        if (!overlaps(box_pt(P), box_tri(A, B, C), sq)) {   // the walk's float boxes admit a superset
            rec[k] = make_int4(0, 0, 0, 0);
            cat[k] = kDrop;
            return;
        }
        const int dt = dtype_pt(P, A, B, C);
        int4 r = make_int4(0, 0, 0, 0);
        int ct = kNone;
        double d;
        switch (dt) {
        case 0: d = d_PP(P, A); if (d < hhat) { r = make_int4(-vI - 1, fv.x, -1, -1); ct = kMap; } break;
        case 1: d = d_PP(P, B); if (d < hhat) { r = make_int4(-vI - 1, fv.y, -1, -1); ct = kMap; } break;
        case 2: d = d_PP(P, C); if (d < hhat) { r = make_int4(-vI - 1, fv.z, -1, -1); ct = kMap; } break;
        case 3: d = d_PE(P, A, B); if (d < hhat) { r = make_int4(-vI - 1, fv.x, fv.y, -1); ct = kMap; } break;
        case 4: d = d_PE(P, B, C); if (d < hhat) { r = make_int4(-vI - 1, fv.y, fv.z, -1); ct = kMap; } break;
        case 5: d = d_PE(P, C, A); if (d < hhat) { r = make_int4(-vI - 1, fv.z, fv.x, -1); ct = kMap; } break;
        case 6: d = d_PT(P, A, B, C); if (d < hhat) { r = make_int4(-vI - 1, fv.x, fv.y, fv.z); ct = kSelfPT; } break;
        default: break;
        }
        rec[k] = r;
        cat[k] = ct;
    }
}

template <int PB>
__device__ inline void ee_traverse_at(int t, int* stk, int Es, const int* __restrict__ perm, const int2* __restrict__ se,
                                      const double* __restrict__ x, TreeView tv, double sq, int* __restrict__ cnt,
                                      int* __restrict__ cntp, int* __restrict__ buf) {
    constexpr int P = 1 << PB;
    const bool active = t < Es * P;
    const int eI = active ? perm[t >> PB] : 0, part = t & (P - 1);
    int n = 0;
    if (active) {
        const int2 A = se[eI];
        const Box b1 = box_edge(ld3(x, A.x), ld3(x, A.y));
        int* b = buf + ((size_t)eI * P + part) * kPartCap;
        part_walk(tv, part, b1, sq, stk + threadIdx.x, [&](int eJ) {
            if (eJ > eI) { if (n < kPartCap) b[n] = eJ; ++n; }
        }, [&] { n = 0; }, PB);
        cntp[eI * P + part] = n;
    }
    const int tot = sum_over_parts<PB>(n);
    if (active && part == 0) cnt[eI] = tot;
}
template <int PB>
__global__ __launch_bounds__(64) void ee_traverse(int Es, const int* __restrict__ perm, const int2* __restrict__ se,
                                                   const double* __restrict__ x, TreeView tv, double sq, int* __restrict__ cnt,
                                                   int* __restrict__ cntp, int* __restrict__ buf) {
    __shared__ int stk[kStack * 64];
    ee_traverse_at<PB>(xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x, stk, Es, perm, se, x, tv, sq, cnt, cntp, buf);
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[Es] = 0;   // the scan's tail
}

template <int PB>
__device__ inline void ee_compact_at(int t, int* stk, int Es, const int* __restrict__ perm, const int2* __restrict__ se,
                                     const double* __restrict__ x, TreeView tv, double sq, const int* __restrict__ off,
                                     const int* __restrict__ cntp, const int* __restrict__ buf, int* __restrict__ candU,
                                     int* __restrict__ own, const int* __restrict__ dc) {
    constexpr int P = 1 << PB;
    if (t >= Es || dc[kDcFlags]) return;
    const int eI = perm[t];
    const int o = off[eI] - off[0];   // the joint scan's edge starts begin at the point total
    int pos = o;
    for (int part = 0; part < P; ++part) {
        const int c = cntp[eI * P + part];
        if (c <= kPartCap) {
            for (int k = 0; k < c; ++k) candU[pos + k] = buf[((size_t)eI * P + part) * kPartCap + k];
        } else {
            const int2 A = se[eI];
            const Box b1 = box_edge(ld3(x, A.x), ld3(x, A.y));
            int m = 0;
            part_walk(tv, part, b1, sq, stk + threadIdx.x, [&](int eJ) {
                if (eJ > eI) { if (m < c) candU[pos + m] = eJ; ++m; }
            }, [&] { m = 0; }, PB);
        }
        pos += c;
    }
    for (int k = o; k < pos; ++k) own[k] = eI;
}

template <int PB>
__global__ __launch_bounds__(64) void ee_compact(int Es, const int* __restrict__ perm, const int2* __restrict__ se, const double* __restrict__ x, TreeView tv,
                           double sq, const int* __restrict__ off, const int* __restrict__ cntp, const int* __restrict__ buf,
                           int* __restrict__ candU, int* __restrict__ own, const int* __restrict__ dc) {
    __shared__ int stk[kStack * 64];
    ee_compact_at<PB>(blockIdx.x * blockDim.x + threadIdx.x, stk, Es, perm, se, x, tv, sq, off, cntp, buf, candU, own, dc);
}

// Edge-edge hits: filters, then classification (CollisionBroadPhase.cpp:858-1052).
__device__ inline void ee_classify(int u, const int* __restrict__ dc, int Es, const int* __restrict__ own, const int* __restrict__ off,
                                   const int2* __restrict__ se, const double* __restrict__ x, const double* __restrict__ vrest,
                                   const int* __restrict__ bt, const int* __restrict__ candU, int* __restrict__ cand, double hhat,
                                   double sq, const int* __restrict__ vs, int4* __restrict__ rec, int2* __restrict__ aux,
                                   int* __restrict__ cat) {
    if (dc[kDcFlags] || u >= dc[kDcEE]) return;
    const int eI = own[u], o = off[eI] - off[0], e = off[eI + 1] - off[0];
    const int eJ = candU[u];
    int k = o;
    for (int j = o; j < e; ++j) k += candU[j] < eJ;
    cand[k] = eJ;
    if ((unsigned)eJ >= (unsigned)Es || k >= e) {   // cannot happen for a consistent walk; never index with it
        rec[k < e ? k : o] = make_int4(0, 0, 0, 0);
        aux[k < e ? k : o] = make_int2(-1, -1);
        cat[k < e ? k : o] = kDrop;
        return;
    }
    const int2 A = se[eI], B = se[eJ];
    if ((is_collider(bt[A.x]) && is_collider(bt[A.y]) && is_collider(bt[B.x]) && is_collider(bt[B.y])) || A.x == B.x ||
        A.x == B.y || A.y == B.x || A.y == B.y || (vs && vs[A.x] != vs[B.x])) {
        rec[k] = make_int4(0, 0, 0, 0);
        aux[k] = make_int2(-1, -1);
        cat[k] = kDrop;
        return;
    }
    const d3 a0 = ld3(x, A.x), a1 = ld3(x, A.y);
    {
        const d3 b0 = ld3(x, B.x), b1v = ld3(x, B.y);
        if (!overlaps(box_edge(a0, a1), box_edge(b0, b1v), sq)) {   // the walk's float boxes admit a superset
            rec[k] = make_int4(0, 0, 0, 0);
            aux[k] = make_int2(-1, -1);
            cat[k] = kDrop;
            return;
        }
        const int dt = dtype_ee(a0, a1, b0, b1v);
        const double cr = ee_cross_sqn(a0, a1, b0, b1v);
        const double ex = eps_x_of(ld3(vrest, A.x), ld3(vrest, A.y), ld3(vrest, B.x), ld3(vrest, B.y));
        const int add_e = (cr < ex) ? -eJ - 2 : -1;
        int4 r = make_int4(0, 0, 0, 0);
        bool hit = false;
        double d;
        switch (dt) {
        case 0: d = d_PP(a0, b0); if (d < hhat) { r = make_int4(-A.x - 1, B.x, -1, add_e); hit = true; } break;
        case 1: d = d_PP(a0, b1v); if (d < hhat) { r = make_int4(-A.x - 1, B.y, -1, add_e); hit = true; } break;
        case 2: d = d_PE(a0, b0, b1v); if (d < hhat) { r = make_int4(-A.x - 1, B.x, B.y, add_e); hit = true; } break;
        case 3: d = d_PP(a1, b0); if (d < hhat) { r = make_int4(-A.y - 1, B.x, -1, add_e); hit = true; } break;
        case 4: d = d_PP(a1, b1v); if (d < hhat) { r = make_int4(-A.y - 1, B.y, -1, add_e); hit = true; } break;
        case 5: d = d_PE(a1, b0, b1v); if (d < hhat) { r = make_int4(-A.y - 1, B.x, B.y, add_e); hit = true; } break;
        case 6: d = d_PE(b0, a0, a1); if (d < hhat) { r = make_int4(-B.x - 1, A.x, A.y, add_e); hit = true; } break;
        case 7: d = d_PE(b1v, a0, a1); if (d < hhat) { r = make_int4(-B.y - 1, A.x, A.y, add_e); hit = true; } break;
        case 8:
            d = d_EE(a0, a1, b0, b1v);
            if (d < hhat) {
                r = (add_e <= -2) ? make_int4(A.x, A.y, B.x, -B.y - Es - 2) : make_int4(A.x, A.y, B.x, B.y);
                hit = true;
            }
            break;
        default: break;
        }
        int ct = kNone;
        int2 ax = make_int2(-1, -1);
        if (hit) {
            // serial compaction rules, CollisionBroadPhase.cpp:1028-1052
            if (r.w >= 0) ct = kSelfEE;
            else if (r.w == -1) ct = kMap;
            else if (r.w >= -Es - 1) { ax = make_int2(eI, -r.w - 2); r.w = -1; ct = kMolPE; }
            else { r.w = -r.w - Es - 2; ct = kMolEE; }
        }
        rec[k] = r;
        aux[k] = ax;
        cat[k] = ct;
    }
}

__global__ void fill_tail(int* p, int n) { p[n] = 0; }
// Per-category output positions: one exclusive scan of four 0/1 category counts per hit
// (int4 x, y, z, w).
struct Cnt4Sum {
    __host__ __device__ int4 operator()(const int4& p, const int4& q) const { return make_int4(p.x + q.x, p.y + q.y, p.z + q.z, p.w + q.w); }
};
// Both hit kinds in one scan: items [0, n1) the point-triangle hits (dc[kDcPT] of them), the
// rest the edge-edge hits (dc[kDcEE]), each with its own category masks; nothing is counted
// over capacity.
struct CatFlags {
    const int* catPT;
    const int* catEE;
    const int* dc;
    int n1;
    unsigned a0, a1, a2, a3, b0, b1, b2, b3;
    __host__ __device__ int4 operator()(int i) const {
        if (dc[kDcFlags]) return make_int4(0, 0, 0, 0);
        const bool pt = i < n1;
        const int j = pt ? i : i - n1;
        if (j >= dc[pt ? kDcPT : kDcEE]) return make_int4(0, 0, 0, 0);
        const unsigned bit = 1u << (pt ? catPT : catEE)[j];
        const unsigned m0 = pt ? a0 : b0, m1 = pt ? a1 : b1, m2 = pt ? a2 : b2, m3 = pt ? a3 : b3;
        return make_int4((m0 & bit) ? 1 : 0, (m1 & bit) ? 1 : 0, (m2 & bit) ? 1 : 0, (m3 & bit) ? 1 : 0);
    }
};
// PP/PE map key: (c0 + V, c1, c2 + 1) packed with `bits` per field preserves the signed
// lexicographic order of std::map<EncodedContact, int> (all keys have c3 = -1).
__device__ inline unsigned long long map_key(int4 r, int V, int bits) {
    return ((unsigned long long)(r.x + V) << (2 * bits)) | ((unsigned long long)r.y << bits) | (unsigned long long)(r.z + 1);
}
// Point-triangle hits -> Self_ActiveSet head (a), map keys (b), partial-CCD pairs (c).
__device__ inline void scatter_pt(int i, int nPT, const int* __restrict__ cat, const int4* __restrict__ pos,
                                  const int4* __restrict__ rec, const int* __restrict__ own, const int* __restrict__ cand, int V, int bits,
                                  int4* __restrict__ self_set, unsigned long long* __restrict__ mkeys, int2* __restrict__ ccd) {
    if (i >= nPT) return;
    const int ct = cat[i];
    if (ct == kDrop) return;
    const int4 p = pos[i];
    ccd[p.z] = make_int2(-own[i] - 1, cand[i]);
    if (ct == kSelfPT) self_set[p.x] = rec[i];
    else if (ct == kMap) mkeys[p.y] = map_key(rec[i], V, bits);
}
// Edge-edge hits -> Self_ActiveSet (a), map keys (b), Self_EE_ActiveSet + eIe (c), CCD pairs (d).
// base: the point hits' totals (Self_ActiveSet, map keys, CCD pairs), where the edge hits'
// entries start.
__device__ inline void scatter_ee(int i, int nEE, int4 base, const int* __restrict__ cat, const int4* __restrict__ pos,
                                  const int4* __restrict__ rec, const int2* __restrict__ aux, const int* __restrict__ own,
                                  const int* __restrict__ cand, int V, int bits, int4* __restrict__ self_set,
                                  unsigned long long* __restrict__ mkeys, int4* __restrict__ ee_set, int2* __restrict__ ee_ie,
                                  int2* __restrict__ ccd) {
    if (i >= nEE) return;
    const int base_self = base.x, base_map = base.y, base_ccd = base.z;
    const int ct = cat[i];
    if (ct == kDrop) return;
    const int4 q = pos[i], z = pos[0];   // the joint scan's edge segment starts at the point totals
    const int4 p = make_int4(q.x - z.x, q.y - z.y, q.z - z.z, q.w - z.w);
    ccd[base_ccd + p.w] = make_int2(own[i], cand[i]);
    if (ct == kSelfEE) self_set[base_self + p.x] = rec[i];
    else if (ct == kMap) mkeys[base_map + p.y] = map_key(rec[i], V, bits);
    else if (ct == kMolPE || ct == kMolEE) { ee_set[p.z] = rec[i]; ee_ie[p.z] = aux[i]; }
}
// Counts after the category scans: the per-category totals are the scans' values at the hit
// counts, the ground count the ground scan's total.  The map keys fill [0, nm) of a
// buffer of capM; past nm they are the all-ones key, which sorts last.
__device__ inline void category_totals(int Vs, const int4& a, const int4& b, const int* __restrict__ pg, int capM, int* __restrict__ dc) {
    const int nm = a.y + b.y;
    dc[kDcMapKeys] = nm;
    if (nm > capM) { dc[kDcFlags] |= 4; return; }
    // the output counts are published only for a set that fits: over capacity they stay zero,
    // so that later stages read no entry of an unfinished set
    dc[kDcA] = a.x; dc[kDcA2] = a.y; dc[kDcAK] = a.z;
    dc[kDcB] = b.x; dc[kDcB2] = b.y; dc[kDcMol] = b.z; dc[kDcBK] = b.w;
    dc[kDcCCD] = a.z + b.w;
    dc[kDcEnv] = pg[Vs] - pg[0];
}
__device__ inline void map_tail(int i, int capM, int nm, unsigned long long* __restrict__ mkeys) {
    if (i >= capM) return;
    if (i >= nm) mkeys[i] = ~0ull;
}
// After the run-length encoding of the sorted keys: the runs without the all-ones run (there
// is one exactly when nm < capM) are the merged map entries, decoded after the A and B parts
// of Self_ActiveSet; the first thread publishes the set's size.
__global__ void map_decode(int capM, const int* __restrict__ nruns, int* __restrict__ dc, const unsigned long long* __restrict__ keys,
                           const int* __restrict__ counts, int V, int bits, int4* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (dc[kDcFlags]) return;
    const int base = dc[kDcA] + dc[kDcB];
    const int nM = *nruns - (dc[kDcMapKeys] < capM ? 1 : 0);
    if (i == 0) dc[kDcSelf] = base + nM;
    if (i >= nM) return;
    const unsigned long long k = keys[i], m = (1ull << bits) - 1;
    out[base + i] = make_int4((int)((k >> (2 * bits)) & m) - V, (int)((k >> bits) & m), (int)(k & m) - 1, -counts[i]);
}
// pos: the joint scan's ground starts, offset by its first entry
__device__ inline void ground_scatter(int s, int Vs, const int* __restrict__ sv, const int* __restrict__ flag, const int* __restrict__ pos,
                                      int* __restrict__ out) {
    if (s >= Vs || !flag[s]) return;
    out[pos[s] - pos[0]] = sv[s];
}
// After the category scan, in one launch (the four parts write disjoint entries): work-groups
// [0, bPT) place the point hits, the next bEE the edge hits, the next bG the ground entries,
// the rest write the map keys' all-ones tail.  Every thread reads the per-category totals
// (the scan's values at the hit counts) itself; the first publishes them (category_totals).
// Nothing is placed when a capacity was exceeded: the host builds the set again.
struct ScatterArgs {
    int* dc;
    int V, bits, Vs, capM, bPT, bEE, bG;
    const int *catPT, *ownPT, *candPT, *catEE, *ownEE, *candEE, *sv, *fg, *pg;
    const int4 *posPT, *recPT, *posEE, *recEE;
    const int2* auxEE;
    int4 *self_set, *ee_set;
    int2 *ee_ie, *ccd;
    unsigned long long* mkeys;
    int* env;
};
__global__ void scatter_all(ScatterArgs a) {
    int* dc = a.dc;
    if (dc[kDcFlags]) return;   // the candidates exceeded their capacity (an earlier launch)
    const int nPT = dc[kDcPT], nEE = dc[kDcEE];
    const int4 ta = a.posPT[nPT], q = a.posEE[nEE], z = a.posEE[0];
    const int4 tb = make_int4(q.x - z.x, q.y - z.y, q.z - z.z, q.w - z.w);
    if (blockIdx.x == 0 && threadIdx.x == 0) category_totals(a.Vs, ta, tb, a.pg, a.capM, dc);
    const int nm = ta.y + tb.y;
    if (nm > a.capM) return;
    int b = blockIdx.x;
    if (b < a.bPT) {
        scatter_pt(b * blockDim.x + threadIdx.x, nPT, a.catPT, a.posPT, a.recPT, a.ownPT, a.candPT, a.V, a.bits, a.self_set, a.mkeys,
                   a.ccd);
        return;
    }
    b -= a.bPT;
    if (b < a.bEE) {
        scatter_ee(b * blockDim.x + threadIdx.x, nEE, ta, a.catEE, a.posEE, a.recEE, a.auxEE, a.ownEE, a.candEE, a.V, a.bits,
                   a.self_set, a.mkeys, a.ee_set, a.ee_ie, a.ccd);
        return;
    }
    b -= a.bEE;
    if (b < a.bG) {
        ground_scatter(b * blockDim.x + threadIdx.x, a.Vs, a.sv, a.fg, a.pg, a.env);
        return;
    }
    b -= a.bG;
    map_tail(b * blockDim.x + threadIdx.x, a.capM, nm, a.mkeys);
}
// the tails of the joint scan's point and ground segments when there is no point walk
__global__ void point_tails(int Vs, int* __restrict__ cnt, int* __restrict__ fg) {
    if (blockIdx.x == 0 && threadIdx.x == 0) { cnt[Vs] = 0; fg[Vs] = 0; }
}

__global__ void dc_reset(int* __restrict__ dc) {
    if (threadIdx.x < kDcWords) dc[threadIdx.x] = 0;
}
// The candidate totals (the per-query scans' last entries) against their capacities.
__global__ void candidate_totals(int Vs, int Es, const int* __restrict__ offPT, const int* __restrict__ offEE, int capPT, int capEE,
                                 int* __restrict__ dc) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int nPT = offPT[Vs], nEE = offEE[Es] - offEE[0];
    dc[kDcPT] = nPT;
    dc[kDcEE] = nEE;
    dc[kDcFlags] = (nPT > capPT ? 1 : 0) | (nEE > capEE ? 2 : 0);
}

// Exclusive scan of the four category flags over n items (flags past a kind's count are
// zero, so each kind's totals follow its last hit).
void category_scan(Context& c, const CatFlags& f, int n, DBuf<int4>& pos) {
    pos.resize((size_t)n + 1);
    hipcub::CountingInputIterator<int> idx(0);
    hipcub::TransformInputIterator<int4, CatFlags, hipcub::CountingInputIterator<int>> it(idx, f);
    size_t bytes = 0;
    CIPC_HIP(hipcub::DeviceScan::ExclusiveScan(nullptr, bytes, it, pos.p, Cnt4Sum(), make_int4(0, 0, 0, 0), n + 1, c.stream));
    c.tmp.reserve(bytes);
    CIPC_HIP(hipcub::DeviceScan::ExclusiveScan(c.tmp.p, bytes, it, pos.p, Cnt4Sum(), make_int4(0, 0, 0, 0), n + 1, c.stream));
}

// The walks' per-part buffers for 2^pt and 2^ee lanes per query.
void size_active_walks(Context& c, int pt, int ee) {
    ActiveSetScratch& b = c.as;
    const size_t Vs = std::max(c.sc.Vs, 1), Es = std::max(c.sc.Es, 1);
    b.cntpPT.resize(Vs << pt); b.cntpEE.resize(Es << ee);
    b.bufPT.resize((Vs << pt) * kPartCap); b.bufEE.resize((Es << ee) * kPartCap);
}

// Both walks of the active set in one launch, and both compactions in another: work-groups
// [0, nbp) take point queries with 2^PBP lanes each, the rest edge queries with 2^PBE (a
// work-group is one wavefront, so a wavefront's lanes share a query kind).  The traversal
// maps work-groups to XCDs as the single walks do, over the joint grid.
struct BothWalks {
    int Vs, Es, nbp;
    const int *pperm, *sv, *eperm;
    const int2* se;
    const double* x;
    TreeView ft, et;
    double sq, gd, hhat;
    d3 gn;
    int *cntPT, *cntEE, *fg, *cntpPT, *cntpEE, *bufPT, *bufEE;
    const int *oPT, *oEE;
    int* dc;
    int capPT, capEE;
    int *candUPT, *candUEE, *ownPT, *ownEE;
};
template <int PBP, int PBE>
__global__ __launch_bounds__(64) void traverse_both(BothWalks a) {
    __shared__ int stk[kStack * 64];
    const int mb = xcd_block(blockIdx.x, gridDim.x);
    if (mb < a.nbp)
        pt_traverse_at<PBP>(mb * 64 + threadIdx.x, stk, a.Vs, a.pperm, a.sv, a.x, a.ft, a.sq, a.cntPT, a.cntpPT, a.bufPT, a.fg, a.gn,
                            a.gd, a.hhat);
    else
        ee_traverse_at<PBE>((mb - a.nbp) * 64 + threadIdx.x, stk, a.Es, a.eperm, a.se, a.x, a.et, a.sq, a.cntEE, a.cntpEE, a.bufEE);
    if (blockIdx.x == 0 && threadIdx.x == 0) { a.cntPT[a.Vs] = 0; a.fg[a.Vs] = 0; a.cntEE[a.Es] = 0; }   // the scans' tails
    if (blockIdx.x == 0 && threadIdx.x < kDcWords) a.dc[threadIdx.x] = 0;   // the iteration's device counts start at zero
}
// The compaction also publishes the candidate totals (the per-query scans' last entries)
// against their capacities, as candidate_totals does; every thread reads them from the scan,
// so none waits for that write.
template <int PBP, int PBE>
__global__ __launch_bounds__(64) void compact_both(BothWalks a) {
    __shared__ int stk[kStack * 64];
    const int b = blockIdx.x;
    const int nPT = a.oPT[a.Vs], nEE = a.oEE[a.Es] - a.oEE[0];
    const int flags = (nPT > a.capPT ? 1 : 0) | (nEE > a.capEE ? 2 : 0);
    if (b == 0 && threadIdx.x == 0) {
        a.dc[kDcPT] = nPT;
        a.dc[kDcEE] = nEE;
        a.dc[kDcFlags] = flags;
    }
    if (flags) return;   // over capacity: the host builds the set again
    // (the compactions' own check of dc[kDcFlags] reads zero here, before or after the write)
    if (b < a.nbp)
        pt_compact_at<PBP>(b * 64 + threadIdx.x, stk, a.Vs, a.pperm, a.sv, a.x, a.ft, a.sq, a.oPT, a.cntpPT, a.bufPT, a.candUPT,
                           a.ownPT, a.dc);
    else
        ee_compact_at<PBE>((b - a.nbp) * 64 + threadIdx.x, stk, a.Es, a.eperm, a.se, a.x, a.et, a.sq, a.oEE, a.cntpEE, a.bufEE,
                           a.candUEE, a.ownEE, a.dc);
}
// Both hit kinds' filters and classification in one launch: work-groups [0, nbp) the point
// hits, the rest the edge hits.
struct BothClassify {
    int nbp, F, Es;
    const int *dc, *ownPT, *oPT, *sv, *bt, *candUPT, *ownEE, *oEE, *candUEE, *vs;
    const int4* faces;
    const int2* se;
    const double *x, *vrest;
    double hhat, sq;
    int *candPT, *catPT, *candEE, *catEE;
    int4 *recPT, *recEE;
    int2* auxEE;
};
__global__ void classify_both(BothClassify a) {
    const int b = blockIdx.x;
    if (b < a.nbp)
        pt_classify(b * blockDim.x + threadIdx.x, a.dc, a.F, a.ownPT, a.oPT, a.sv, a.x, a.bt, a.faces, a.candUPT, a.candPT, a.hhat, a.sq,
                    a.vs, a.recPT, a.catPT);
    else
        ee_classify((b - a.nbp) * blockDim.x + threadIdx.x, a.dc, a.Es, a.ownEE, a.oEE, a.se, a.x, a.vrest, a.bt, a.candUEE, a.candEE,
                    a.hhat, a.sq, a.vs, a.recEE, a.auxEE, a.catEE);
}
template <int PBP, typename F>
static void with_pbe(int pbe, F&& f) {
    switch (pbe) {
    case 0: f(std::integral_constant<int, PBP>{}, std::integral_constant<int, 0>{}); break;
    case 1: f(std::integral_constant<int, PBP>{}, std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, PBP>{}, std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, PBP>{}, std::integral_constant<int, 3>{}); break;
    default: f(std::integral_constant<int, PBP>{}, std::integral_constant<int, 4>{}); break;
    }
}
template <typename F>
static void with_splits(int pbp, int pbe, F&& f) {
    switch (pbp) {
    case 0: with_pbe<0>(pbe, f); break;
    case 1: with_pbe<1>(pbe, f); break;
    case 2: with_pbe<2>(pbe, f); break;
    case 3: with_pbe<3>(pbe, f); break;
    default: with_pbe<4>(pbe, f); break;
    }
}
// Both walks (compact = false) or both compactions, with the context's splits, on st; needs
// Vs > 0 and Es > 0.
static void launch_both_walks(Context& c, bool compact, hipStream_t st) {
    Scene& s = c.sc;
    ActiveSetScratch& b = c.as;
    const int pbp = c.tune.as_pt_bits, pbe = c.tune.as_ee_bits;
    BothWalks a;
    a.Vs = s.Vs; a.Es = s.Es;
    a.nbp = (int)blocks_for(compact ? (size_t)s.Vs : (size_t)s.Vs << pbp, 64);
    const int nbe = (int)blocks_for(compact ? (size_t)s.Es : (size_t)s.Es << pbe, 64);
    a.pperm = s.sv_perm.p; a.sv = s.sv.p; a.eperm = s.se_perm.p; a.se = s.se.p; a.x = s.x.p;
    a.ft = TreeView{c.face_tree.n, c.face_tree.nodes.p, c.face_tree.wide.p};
    a.et = TreeView{c.edge_tree.n, c.edge_tree.nodes.p, c.edge_tree.wide.p};
    a.sq = sqrt(s.hhat); a.gd = s.gd; a.hhat = s.hhat; a.gn = s.gn;
    a.cntPT = b.cnt_all.p; a.cntEE = a.cntPT + (s.Vs + 1); a.fg = a.cntEE + (s.Es + 1);
    a.cntpPT = b.cntpPT.p; a.cntpEE = b.cntpEE.p; a.bufPT = b.bufPT.p; a.bufEE = b.bufEE.p;
    a.oPT = b.off_all.p; a.oEE = a.oPT + (s.Vs + 1); a.dc = c.dcnt.p; a.capPT = b.capPT; a.capEE = b.capEE;
    a.candUPT = b.candUPT.p; a.candUEE = b.candUEE.p; a.ownPT = b.ownPT.p; a.ownEE = b.ownEE.p;
    with_splits(pbp, pbe, [&](auto P, auto E) {
        if (compact) hipLaunchKernelGGL((compact_both<decltype(P)::value, decltype(E)::value>), dim3(a.nbp + nbe), dim3(64), 0, st, a);
        else hipLaunchKernelGGL((traverse_both<decltype(P)::value, decltype(E)::value>), dim3(a.nbp + nbe), dim3(64), 0, st, a);
    });
    CIPC_CHECK_LAUNCH();
}

// The active set's walks with 2^pb lanes per query: the traversal (compact = false) or the
// compaction of its hits (compact = true), point-triangle (pt) or edge-edge.
void launch_active_walk(Context& c, bool pt, bool compact, int pb, hipStream_t st) {
    if (!st) st = c.stream;
    Scene& s = c.sc;
    ActiveSetScratch& b = c.as;
    const double sq = sqrt(s.hhat);
    const TreeView ft{c.face_tree.n, c.face_tree.nodes.p, c.face_tree.wide.p};
    const TreeView et{c.edge_tree.n, c.edge_tree.nodes.p, c.edge_tree.wide.p};
    const int* dc = c.dcnt.p;
    const int n = pt ? s.Vs : s.Es;
    if (n <= 0) return;
    const dim3 grid(blocks_for(compact ? (size_t)n : (size_t)n << pb, 64)), block(64);
    // the joint count / start arrays: points (Vs + 1), edges (Es + 1), ground flags (Vs + 1)
    int* cPT = b.cnt_all.p;
    int* cEE = cPT + (s.Vs + 1);
    int* cG = cEE + (s.Es + 1);
    const int* oPT = b.off_all.p;
    const int* oEE = oPT + (s.Vs + 1);
    auto trav_pt = [&](auto k) { hipLaunchKernelGGL(k, grid, block, 0, st, s.Vs, s.sv_perm.p, s.sv.p, s.x.p, ft, sq, cPT,
                                                    b.cntpPT.p, b.bufPT.p, cG, s.gn, s.gd, s.hhat); };
    auto trav_ee = [&](auto k) { hipLaunchKernelGGL(k, grid, block, 0, st, s.Es, s.se_perm.p, s.se.p, s.x.p, et, sq, cEE,
                                                    b.cntpEE.p, b.bufEE.p); };
    auto comp_pt = [&](auto k) { hipLaunchKernelGGL(k, grid, block, 0, st, s.Vs, s.sv_perm.p, s.sv.p, s.x.p, ft, sq, oPT,
                                                    b.cntpPT.p, b.bufPT.p, b.candUPT.p, b.ownPT.p, dc); };
    auto comp_ee = [&](auto k) { hipLaunchKernelGGL(k, grid, block, 0, st, s.Es, s.se_perm.p, s.se.p, s.x.p, et, sq, oEE,
                                                    b.cntpEE.p, b.bufEE.p, b.candUEE.p, b.ownEE.p, dc); };
#define CIPC_SPLIT(F, K)                 \
    switch (pb) {                        \
    case 0: F(K<0>); break;              \
    case 1: F(K<1>); break;              \
    case 2: F(K<2>); break;              \
    case 3: F(K<3>); break;              \
    default: F(K<4>); break;             \
    }
    if (pt && !compact) CIPC_SPLIT(trav_pt, pt_traverse)
    else if (!pt && !compact) CIPC_SPLIT(trav_ee, ee_traverse)
    else if (pt) CIPC_SPLIT(comp_pt, pt_compact)
    else CIPC_SPLIT(comp_ee, ee_compact)
#undef CIPC_SPLIT
    CIPC_CHECK_LAUNCH();
}

// Buffers of the hits and of the outputs are sized by the capacities, so no stage needs a
// count on the host: every kernel reads its count from dcnt, and does nothing once a
// capacity is exceeded (the host then grows it and builds the set again).  Bounds: a
// category of hits is at most the hits of its type, so the outputs are bounded by the hit
// capacities (Self_ActiveSet head nA + nB, map tail nM <= nm <= capM, Self_EE_ActiveSet
// nE <= nEE, CCD pairs nPTk + nEEk <= nPT + nEE, ground nG <= Vs).
void compute_active_set(Context& c, bool publish) {
    Scene& s = c.sc;
    // Everything on the main stream: both static trees refitted by the same launches, both
    // walks in one launch, both compactions in one, both classifications in one.
    const int Vs = s.Vs, Es = s.Es;
    {
        PrimSrc fs, es;
        fs.kind = 1; fs.faces = s.faces.p; fs.x = s.x.p;
        es.kind = 2; es.se = s.se.p; es.x = s.x.p;
        c.face_tree.prim.resize(std::max(s.F, 1));
        c.edge_tree.prim.resize(std::max(Es, 1));
        if (refit_only(c.face_tree, s.F) && refit_only(c.edge_tree, Es)) refit_two(c.face_tree, fs, c.edge_tree, es, c.stream);
        else build_static_trees(c);
    }
    ActiveSetScratch& b = c.as;
    const double sq = sqrt(s.hhat);
    if (b.capPT == 0) b.capPT = std::max(1024, 4 * Vs);
    if (b.capEE == 0) b.capEE = std::max(1024, 4 * Es);
    if (b.capM == 0) b.capM = 1024;
    const int capPT = b.capPT, capEE = b.capEE, capM = b.capM;
    c.dcnt.resize(kDcWords);
    int* dc = c.dcnt.p;
    const bool both = Vs > 0 && Es > 0;
    if (!both) hipLaunchKernelGGL(dc_reset, dim3(1), dim3(64), 0, c.stream, dc);   // else the joint walk clears them
    // hit counts per point and per edge query and the ground flags, in one array for one scan
    const int nall = (Vs + 1) + (Es + 1) + (Vs + 1);
    b.cnt_all.resize(nall); b.off_all.resize(nall);
    int* cPT = b.cnt_all.p;
    int* cEE = cPT + (Vs + 1);
    int* cG = cEE + (Es + 1);
    const int* oPT = b.off_all.p;
    const int* oEE = oPT + (Vs + 1);
    const int* oG = oEE + (Es + 1);
    size_active_walks(c, c.tune.as_pt_bits, c.tune.as_ee_bits);
    // A: one walk per query part (the point walk also sets the ground flags; the walks zero
    // the scan segments' tails)
    if (both) {
        launch_both_walks(c, false, c.stream);
    } else {
        if (Vs > 0) launch_active_walk(c, true, false, c.tune.as_pt_bits);
        if (Es > 0) launch_active_walk(c, false, false, c.tune.as_ee_bits);
        if (Vs == 0) hipLaunchKernelGGL(point_tails, dim3(1), dim3(1), 0, c.stream, Vs, cPT, cG);
        if (Es == 0) hipLaunchKernelGGL(fill_tail, dim3(1), dim3(1), 0, c.stream, cEE, Es);
    }
    CIPC_CHECK_LAUNCH();
    // one scan: the edge and ground segments' starts come out offset by their first entry
    exclusive_scan_int(c, b.cnt_all.p, b.off_all.p, nall);
    if (!both) hipLaunchKernelGGL(candidate_totals, dim3(1), dim3(64), 0, c.stream, Vs, Es, oPT, oEE, capPT, capEE, dc);   // else the joint compaction
    CIPC_CHECK_LAUNCH();
    // B + C: pack, sort, filter and classify the hits
    b.candUPT.resize(capPT + 1); b.candUEE.resize(capEE + 1);
    b.candPT.resize(capPT + 1); b.recPT.resize(capPT + 1); b.catPT.resize(capPT + 1); b.ownPT.resize(capPT + 1);
    b.candEE.resize(capEE + 1); b.recEE.resize(capEE + 1); b.catEE.resize(capEE + 1); b.auxEE.resize(capEE + 1); b.ownEE.resize(capEE + 1);
    BothClassify cl;
    cl.F = s.F; cl.Es = Es;
    cl.dc = dc; cl.ownPT = b.ownPT.p; cl.oPT = oPT; cl.sv = s.sv.p; cl.bt = s.bt.p; cl.candUPT = b.candUPT.p;
    cl.ownEE = b.ownEE.p; cl.oEE = oEE; cl.candUEE = b.candUEE.p; cl.vs = scene_of_vertex(s); cl.faces = s.faces.p; cl.se = s.se.p;
    cl.x = s.x.p; cl.vrest = s.vrest.p; cl.hhat = s.hhat; cl.sq = sq;
    cl.candPT = b.candPT.p; cl.catPT = b.catPT.p; cl.candEE = b.candEE.p; cl.catEE = b.catEE.p;
    cl.recPT = b.recPT.p; cl.recEE = b.recEE.p; cl.auxEE = b.auxEE.p;
    // the point hits' work-groups, then the edge hits' (none of a kind without queries)
    cl.nbp = Vs > 0 ? (int)blocks_for(capPT, 128) : 0;
    const int nbe = Es > 0 ? (int)blocks_for(capEE, 128) : 0;
    if (both) {
        launch_both_walks(c, true, c.stream);
    } else {
        if (Vs > 0) launch_active_walk(c, true, true, c.tune.as_pt_bits);
        if (Es > 0) launch_active_walk(c, false, true, c.tune.as_ee_bits);
    }
    if (cl.nbp + nbe > 0) hipLaunchKernelGGL(classify_both, dim3(cl.nbp + nbe), dim3(128), 0, c.stream, cl);
    CIPC_CHECK_LAUNCH();
    // category positions
    const unsigned keptPT = (1u << kNone) | (1u << kSelfPT) | (1u << kMap);
    const unsigned keptEE = (1u << kNone) | (1u << kSelfEE) | (1u << kMap) | (1u << kMolPE) | (1u << kMolEE);
    // both kinds of hit in one scan: points [0, capPT], edges after them
    b.pos_all.resize((size_t)capPT + 1 + capEE + 1);
    {
        const CatFlags f{b.catPT.p, b.catEE.p, dc, capPT + 1, 1u << kSelfPT, 1u << kMap, keptPT, 0u,
                         1u << kSelfEE, 1u << kMap, (1u << kMolPE) | (1u << kMolEE), keptEE};
        category_scan(c, f, capPT + 1 + capEE + 1, b.pos_all);
    }
    const int4* posPT = b.pos_all.p;
    const int4* posEE = posPT + (capPT + 1);
    // PP/PE map keys need the vertex count to fit three fields in 64 bits
    int bits = 1;
    while ((1ll << bits) < (long long)s.V + 2) ++bits;
    if (3 * bits > 64) throw Error(CIPC_ERR_OVERFLOW, "vertex count too large for the point-point/point-edge key packing");
    // outputs, sized by their bounds (the map tail of Self_ActiveSet has at most capM entries)
    c.self_set.resize((size_t)capPT + capEE + capM + 1); c.ee_set.resize((size_t)capEE + 1); c.ee_ie.resize((size_t)capEE + 1);
    c.ccd_set.resize((size_t)capPT + capEE + 1); c.env_set.resize((size_t)Vs + 1);
    b.mkeys.resize(capM); b.mkeys_sorted.resize(capM); b.ukeys.resize(capM); b.ucounts.resize(capM); b.nruns.resize(1);
    {
        ScatterArgs a;
        a.dc = dc; a.V = s.V; a.bits = bits; a.Vs = Vs; a.capM = capM;
        a.bPT = (int)blocks_for(capPT, 256); a.bEE = (int)blocks_for(capEE, 256); a.bG = (int)blocks_for(Vs, 256);
        a.catPT = b.catPT.p; a.ownPT = b.ownPT.p; a.candPT = b.candPT.p; a.posPT = posPT; a.recPT = b.recPT.p;
        a.catEE = b.catEE.p; a.ownEE = b.ownEE.p; a.candEE = b.candEE.p; a.posEE = posEE; a.recEE = b.recEE.p; a.auxEE = b.auxEE.p;
        a.sv = s.sv.p; a.fg = cG; a.pg = oG; a.self_set = c.self_set.p; a.ee_set = c.ee_set.p; a.ee_ie = c.ee_ie.p;
        a.ccd = c.ccd_set.p; a.mkeys = b.mkeys.p; a.env = c.env_set.p;
        const int blocks = a.bPT + a.bEE + a.bG + (int)blocks_for(capM, 256);
        hipLaunchKernelGGL(scatter_all, dim3(blocks), dim3(256), 0, c.stream, a);
        CIPC_CHECK_LAUNCH();
    }
    // PP/PE map: sort packed keys, run-length encode (std::map order with multiplicity); the
    // all-ones keys past nm form the last run
    sort_u64(c, b.mkeys.p, b.mkeys_sorted.p, capM, 3 * bits);
    size_t bytes = 0;
    CIPC_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, bytes, b.mkeys_sorted.p, b.ukeys.p, b.ucounts.p, b.nruns.p, capM, c.stream));
    c.tmp.reserve(bytes);
    CIPC_HIP(hipcub::DeviceRunLengthEncode::Encode(c.tmp.p, bytes, b.mkeys_sorted.p, b.ukeys.p, b.ucounts.p, b.nruns.p, capM, c.stream));
    hipLaunchKernelGGL(map_decode, dim3(blocks_for(capM, 256)), dim3(256), 0, c.stream, capM, b.nruns.p, dc, b.ukeys.p, b.ucounts.p,
                       s.V, bits, c.self_set.p);
    CIPC_CHECK_LAUNCH();
    // the counts are read with the host's next synchronisation (settle_active_set)
    c.pin_dc.reserve(kDcWords);
    if (publish) CIPC_HIP(hipMemcpyAsync(c.pin_dc.p, dc, kDcWords * sizeof(int), hipMemcpyDeviceToHost, c.stream));
    c.dc_pending = true;
}

namespace {
int grown(int cap, int need) {
    if (need > INT32_MAX / 2) throw Error(CIPC_ERR_OVERFLOW, "active-set count beyond the 32-bit index range");
    return need > cap ? need + need / 4 + 256 : cap;
}
}  // namespace

bool grow_capacities(Context& c) {
    if (!c.dc_pending) return false;
    ActiveSetScratch& b = c.as;
    const int* h = c.pin_dc.p;
    const int flags = h[kDcFlags];
    // the hit counts are final whenever they were counted; the map and contact counts only
    // once the hits fitted
    if (flags & 1) b.capPT = grown(b.capPT, h[kDcPT]);
    if (flags & 2) b.capEE = grown(b.capEE, h[kDcEE]);
    if (flags & 4) b.capM = grown(b.capM, h[kDcMapKeys]);
    if (flags == 0 && h[kDcContactOver]) c.capS = grown(c.capS, h[kDcSelf] + h[kDcMol]);
    return flags != 0 || h[kDcContactOver] != 0;
}

bool settle_active_set(Context& c) {
    if (!c.dc_pending) return false;
    bool rebuilt = false;
    for (int round = 0;; ++round) {
        CIPC_HIP(hipStreamSynchronize(c.stream));
        if (!grow_capacities(c)) break;
        if (round >= 3) throw Error(CIPC_ERR_RUNTIME, "active-set capacities kept overflowing");
        compute_active_set(c);
        rebuilt = true;
    }
    const int* h = c.pin_dc.p;
    c.n_self = h[kDcSelf]; c.n_ee = h[kDcMol]; c.n_ccd = h[kDcCCD]; c.n_env = h[kDcEnv];
    c.dc_pending = false;
    return rebuilt;
}

void active_set_now(Context& c) {
    compute_active_set(c);
    settle_active_set(c);
}

// ---------------------------------------------------------------------------------
// isIntersected (IPCSolver.cpp:476-542): ground gaps of the current
// Environment_ActiveSet, then every surface edge against every face whose static box it
// touches, with IglUtils::segTriIntersect (ContactMechanics.h:47-71).
// ---------------------------------------------------------------------------------
// Eigen 3.4 FullPivLU solve of the 3x3 system of segTriIntersect: complete pivoting on
// the first largest |a_ij| in column-major order, rank threshold 3 eps |max pivot|.
__device__ inline bool seg_tri(d3 e0, d3 e1, d3 t0, d3 t1, d3 t2) {
    const d3 c0 = t1 - t0, c1 = t2 - t0, c2 = e0 - e1;
    const d3 n = cross(c0, c1);
    if (dot(n, e0 - t0) * dot(n, e1 - t0) > 0.0) return false;
    double A[3][3] = {{c0.x, c1.x, c2.x}, {c0.y, c1.y, c2.y}, {c0.z, c1.z, c2.z}};
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])
                     + A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    if (det == 0.0) return false;
    int rt[3] = {0, 1, 2}, ct[3] = {0, 1, 2}, nz = 3;
    double maxpiv = 0.0;
    for (int k = 0; k < 3; ++k) {
        int br = k, bc = k;
        double best = fabs(A[k][k]);
        for (int cc = k; cc < 3; ++cc)
            for (int r = k; r < 3; ++r)
                if (fabs(A[r][cc]) > best) { best = fabs(A[r][cc]); br = r; bc = cc; }
        if (best == 0.0) { nz = k; for (int i = k; i < 3; ++i) { rt[i] = i; ct[i] = i; } break; }
        if (best > maxpiv) maxpiv = best;
        rt[k] = br; ct[k] = bc;
        if (br != k) for (int cc = 0; cc < 3; ++cc) { const double t = A[k][cc]; A[k][cc] = A[br][cc]; A[br][cc] = t; }
        if (bc != k) for (int r = 0; r < 3; ++r) { const double t = A[r][k]; A[r][k] = A[r][bc]; A[r][bc] = t; }
        for (int r = k + 1; r < 3; ++r) A[r][k] /= A[k][k];
        for (int cc = k + 1; cc < 3; ++cc)
            for (int r = k + 1; r < 3; ++r) A[r][cc] -= A[r][k] * A[k][cc];
    }
    int rank = 0;
    const double thr = maxpiv * (2.220446049250313e-16 * 3.0);
    for (int i = 0; i < nz; ++i) rank += fabs(A[i][i]) > thr;
    const d3 rhs = e0 - t0;
    double cv[3] = {rhs.x, rhs.y, rhs.z}, xs[3] = {0, 0, 0};
    if (rank > 0) {
        for (int k = 0; k < 3; ++k) { const double t = cv[k]; cv[k] = cv[rt[k]]; cv[rt[k]] = t; }
        for (int i = 0; i < 3; ++i) if (cv[i] != 0.0) for (int r = i + 1; r < 3; ++r) cv[r] -= cv[i] * A[r][i];
        for (int i = rank - 1; i >= 0; --i) {
            if (cv[i] == 0.0) continue;
            cv[i] /= A[i][i];
            for (int r = 0; r < i; ++r) cv[r] -= cv[i] * A[r][i];
        }
        int q[3] = {0, 1, 2};
        for (int k = 0; k < 3; ++k) { const int t = q[k]; q[k] = q[ct[k]]; q[ct[k]] = t; }
        for (int i = 0; i < 3; ++i) xs[q[i]] = (i < rank) ? cv[i] : 0.0;
    }
    return xs[0] >= 0.0 && xs[1] >= 0.0 && xs[0] + xs[1] <= 1.0 && xs[2] >= 0.0 && xs[2] <= 1.0;
}
// hit[0] (one scene) or hit[scene] (batch) is set for a face that an edge crosses
__global__ void intersect_kernel(int F, const int4* __restrict__ faces, const int2* __restrict__ se, const double* __restrict__ x,
                                 const int* __restrict__ bt, TreeView et, const int* __restrict__ vs, int* __restrict__ hit) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int4 fv = faces[f];
    const d3 t0 = ld3(x, fv.x), t1 = ld3(x, fv.y), t2 = ld3(x, fv.z);
    const Box q = box_tri(t0, t1, t2);
    bool any = false;
    lbvh_visit(et, -1, q, 0.0, [&](int e) {
        if (any) return;
        const int2 E = se[e];
        if (E.x == fv.x || E.x == fv.y || E.x == fv.z || E.y == fv.x || E.y == fv.y || E.y == fv.z) return;
        if (vs && vs[E.x] != vs[fv.x]) return;
        if (is_collider(bt[E.x]) && is_collider(bt[E.y]) && is_collider(bt[fv.x]) && is_collider(bt[fv.y]) && is_collider(bt[fv.z])) return;
        if (seg_tri(ld3(x, E.x), ld3(x, E.y), t0, t1, t2)) any = true;
    });
    if (any) atomicOr(hit + (vs ? vs[fv.x] : 0), 1);
}
__global__ void ground_touch_kernel(int n, const int* __restrict__ env, const double* __restrict__ x, d3 gn, double gd,
                                    const int* __restrict__ vs, int* __restrict__ hit) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = env[i];
    const double dist = dot(gn, ld3(x, v)) - gd;
    if (dist * dist <= 0.0) atomicOr(hit + (vs ? vs[v] : 0), 1);
}
__global__ void zero_int(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void scene_intersections(Context& c, int* out) {
    Scene& s = c.sc;
    const int S = s.S;
    c.counters.resize(2 * (size_t)S);
    hipLaunchKernelGGL(zero_int, dim3(blocks_for(2 * S, 64)), dim3(64), 0, c.stream, c.counters.p, 2 * S);
    if (c.n_env > 0)
        hipLaunchKernelGGL(ground_touch_kernel, dim3(blocks_for(c.n_env, 256)), dim3(256), 0, c.stream, (int)c.n_env, c.env_set.p,
                           s.x.p, s.gn, s.gd, scene_of_vertex(s), c.counters.p);
    build_static_trees(c);
    const TreeView et{c.edge_tree.n, c.edge_tree.nodes.p, c.edge_tree.wide.p};
    if (s.F > 0)
        hipLaunchKernelGGL(intersect_kernel, dim3(blocks_for(s.F, 128)), dim3(128), 0, c.stream, s.F, s.faces.p, s.se.p, s.x.p, s.bt.p,
                           et, scene_of_vertex(s), c.counters.p + S);
    CIPC_CHECK_LAUNCH();
    std::vector<int> h(2 * (size_t)S);
    read_back(c, h.data(), (const int*)c.counters.p, h.size());
    for (int k = 0; k < S; ++k) out[k] = (h[k] != 0 || h[S + k] != 0) ? 1 : 0;
}

bool any_intersection(Context& c) {
    std::vector<int> r(c.sc.S);
    scene_intersections(c, r.data());
    for (int v : r)
        if (v) return true;
    return false;
}

}  // namespace cipc
