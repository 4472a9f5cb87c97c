// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate expressions of CPU-IPC
// (MPL-2.0) in the reference's operation order; each such part cites the source it follows.
// Newton step bound on the device.  Compiled with -ffp-contract=off.
//
// Reference: limitStepByGround (CPU IPC/FeasibleStep.cpp:8-40), Self_largestFeasibleStepSize
// (ContactMechanics.cpp:3704-3752), the CFL bound (IPCSolver.cpp:960-974), the swept
// rebuild SpatialHash::build(mesh, p, alpha) -> CPUBroadPhaseBVH::buildSwept
// (CollisionBroadPhase.cpp:223-237, LBVH.cpp:307-339), Self_largestFeasibleStepSize_CCD
// (ContactMechanics.cpp:3754-3867) and the merge (IPCSolver.cpp:985-989).
//
// Every bound is a min (or max) of per-primitive values; they are reduced with integer
// atomics on the IEEE bit patterns of non-negative doubles, which order like the values,
// so the result is exact and independent of scheduling.
#include "context.hpp"
#include "contact.hpp"
#include "lbvh.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace cipc {

__device__ inline unsigned long long ord_bits(double v) {
    if (v == 0.0) v = 0.0;  // -0 -> +0
    return (unsigned long long)__double_as_longlong(v);
}
__device__ inline void atomic_min_nonneg(unsigned long long* a, double v) { atomicMin(a, ord_bits(v)); }
__device__ inline void atomic_max_nonneg(unsigned long long* a, double v) { atomicMax(a, ord_bits(v)); }
__device__ inline double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }

// slots (kSlots per scene, scene k at k * kSlots): 0 ground min, 1 CFL max |p|, 2 partial
// CCD min, 3 full CCD min, 4 the point-triangle phase's result read by the edge-edge phase,
// 5 the swept trees' step
__global__ void init_bounds(int S, unsigned long long* sl) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S) return;
    unsigned long long* s = sl + kSlots * k;
    s[0] = ord_bits(1.0);
    s[1] = ord_bits(0.0);
    s[2] = ord_bits(1.0);
    s[3] = ord_bits(1.0);
}
__device__ inline int scene_of(const int* vs, int v) { return vs ? vs[v] : 0; }
__device__ inline double wave_min(double v) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) v = fmin(v, __shfl_xor(v, m));
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) v = fmax(v, __shfl_xor(v, m));
    return v;
}
constexpr int kNoScene = -1;   // a lane without a value (past the end of its range)
// min (Max: max) of v into slot k of the lane's scene.  One atomic per wavefront when all of
// its lanes with a value share the scene (always for one scene; the usual case in a batch,
// whose primitives are ordered by scene), else one per lane.  Every lane of the wavefront
// calls it; a lane without a value passes kNoScene and the neutral value.
template <bool Max>
__device__ inline void scene_atomic(unsigned long long* sl, int scene, int k, double v) {
    const unsigned long long has = __ballot(scene != kNoScene);
    if (!has) return;
    const int s0 = __shfl(scene, __ffsll((long long)has) - 1);
    if (__all(scene == s0 || scene == kNoScene)) {
        const double w = Max ? wave_max(v) : wave_min(v);
        if ((threadIdx.x & 63) == (unsigned)(__ffsll((long long)has) - 1)) {
            if (Max) atomic_max_nonneg(sl + kSlots * s0 + k, w);
            else atomic_min_nonneg(sl + kSlots * s0 + k, w);
        }
    } else if (scene != kNoScene) {
        if (Max) atomic_max_nonneg(sl + kSlots * scene + k, v);
        else atomic_min_nonneg(sl + kSlots * scene + k, v);
    }
}

__global__ void ground_cfl_kernel(int Vs, const int* __restrict__ sv, const double* __restrict__ x, const double* __restrict__ p,
                                  d3 gn, double gd, double slack, const int* __restrict__ vs, unsigned long long* __restrict__ sl) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double best = 1.0, move = 0.0;
    int scene = kNoScene;
    if (s < Vs) {
        const int v = sv[s];
        scene = scene_of(vs, v);
        const d3 pv = ld3(p, v);
        const double nm = dot(gn, -pv);
        if (!(nm >= 0.0)) {
            const double dist = dot(gn, ld3(x, v)) - gd;
            const double cand = -dist / nm * slack;
            if (cand >= 0.0) best = cand;
        }
        move = sqrt(sqn(pv));
    }
    scene_atomic<false>(sl, scene, 0, best);
    scene_atomic<true>(sl, scene, 1, move);
}

// The pairs are the active set's CCD pairs: their count is the device count n_dev.
__global__ void partial_ccd_kernel(const int* __restrict__ n_dev, const int2* __restrict__ pairs, const int* __restrict__ sv, const int4* __restrict__ faces,
                                   const int2* __restrict__ se, const double* __restrict__ x, const double* __restrict__ p, double eta,
                                   const int* __restrict__ vs, unsigned long long* __restrict__ sl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = *n_dev;
    double a = 1.0;
    int scene = kNoScene;
    if (i < n) {
        const int2 pr = pairs[i];
        int v0;
        if (pr.x < 0) {
            v0 = sv[-pr.x - 1];
        } else {
            v0 = se[pr.x].x;
        }
        scene = scene_of(vs, v0);
        // the merge takes min(ground step, partial CCD): a pair at or past the ground step
        // cannot change it (slot 0 is final: the ground kernel ran before this one)
        double g = bits_to_double(sl[kSlots * scene]);
        g = fmin(1.0, g);
        const double cap = g <= 0.0 ? 1.0 : g;
        if (pr.x < 0) {
            const int pt = v0;
            const int4 f = faces[pr.y];
            a = accd_pt(ld3(x, pt), ld3(x, f.x), ld3(x, f.y), ld3(x, f.z), -ld3(p, pt), -ld3(p, f.x), -ld3(p, f.y), -ld3(p, f.z), eta,
                        0.0, cap);
        } else {
            const int2 e1 = se[pr.x], e2 = se[pr.y];
            a = accd_ee(ld3(x, e1.x), ld3(x, e1.y), ld3(x, e2.x), ld3(x, e2.y), -ld3(p, e1.x), -ld3(p, e1.y), -ld3(p, e2.x),
                        -ld3(p, e2.y), eta, 0.0, cap);
        }
    }
    scene_atomic<false>(sl, scene, 2, a);
}

// Full CCD, point-triangle phase (ContactMechanics.cpp:3764-3816): each surface vertex is
// split over 2^pb lanes, one subtree of the swept face tree each; the step is the min
// over every lane, reduced across the wavefront before one atomic.  The minimum does not
// depend on the split, which the autotuner picks per context.
__global__ __launch_bounds__(64) void full_ccd_pt_kernel(int Vs, const int* __restrict__ perm, const int* __restrict__ sv,
                                                          const int4* __restrict__ faces, const int* __restrict__ bt,
                                                          const double* __restrict__ x, const double* __restrict__ p,
                                                          TreeView tv, unsigned long long* __restrict__ sl, double eta,
                                                          const int* __restrict__ vs, int pb) {
    __shared__ int stk[kStack * 64];
    const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    double best = 1.0;
    int scene = kNoScene;
    if (t < (Vs << pb)) {
        const int s = perm[t >> pb], part = t & ((1 << pb) - 1);
        const int vI = sv[s];
        scene = scene_of(vs, vI);
        const double a = bits_to_double(sl[kSlots * scene + 5]);
        const d3 X = ld3(x, vI), P = ld3(p, vI);
        const Box vb = swept_pt(X, P, a);   // the query box: the point swept with the trees' step
        part_walk(tv, part, vb, 0.0, stk + threadIdx.x, [&](int f) {
            const int4 fv = faces[f];
            if (is_collider(bt[vI]) && is_collider(bt[fv.x]) && is_collider(bt[fv.y]) && is_collider(bt[fv.z])) return;
            if (vI == fv.x || vI == fv.y || vI == fv.z) return;
            if (vs && vs[fv.x] != scene) return;          // scenes of a batch never touch
            const d3 x0 = ld3(x, fv.x), x1 = ld3(x, fv.y), x2 = ld3(x, fv.z);
            const d3 p0 = ld3(p, fv.x), p1 = ld3(p, fv.y), p2 = ld3(p, fv.z);
            const Box fb = box_union(box_union(swept_pt(x0, p0, a), swept_pt(x1, p1, a)), swept_pt(x2, p2, a));
            if (!overlaps(vb, fb, 0.0)) return;
            // capped at the lane's best so far, itself at most the seeded step a: the
            // minimum over all pairs is unchanged
            const double la = accd_pt(X, x0, x1, x2, -P, -p0, -p1, -p2, eta, 0.0, fmin(best, a));
            best = fmin(la, best);
        }, [] {}, pb);
    }
    scene_atomic<false>(sl, scene, 3, best);
}

// Full CCD, edge-edge phase (ContactMechanics.cpp:3818-3866): every unordered pair of
// surface edges whose swept boxes (built with the incoming step) overlap, the pair ordered as
// the reference orders it (I the larger edge id, J the smaller) and filtered with the exact
// swept boxes of the point-triangle-reduced step read from the slots.  Each edge (split over
// 2^pb lanes) walks only the leaves after its own leaf, so each pair is met once.
__global__ __launch_bounds__(64) void full_ccd_ee_kernel(int Es, const int* __restrict__ perm, const int2* __restrict__ se,
                                                          const int* __restrict__ bt, const double* __restrict__ x,
                                                          const double* __restrict__ p, TreeView tv, const Box* __restrict__ ebox,
                                                          unsigned long long* __restrict__ sl, double eta, const int* __restrict__ vs,
                                                          int pb) {
    __shared__ int stk[kStack * 64];
    const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    double best = 1.0;
    int scene = kNoScene;
    if (t < (Es << pb)) {
        const int eQ = perm[t >> pb], part = t & ((1 << pb) - 1);
        const int2 Q = se[eQ];
        scene = scene_of(vs, Q.x);
        const double a = bits_to_double(sl[kSlots * scene + 4]);
        const Box bq = box_union(swept_pt(ld3(x, Q.x), ld3(p, Q.x), a), swept_pt(ld3(x, Q.y), ld3(p, Q.y), a));
        // Query box: bq, swept with the point-triangle-reduced step, which is at most the
        // tree's step.  The swept box of a smaller step lies inside that of a larger one (the
        // rounded end point x - a p moves monotonically with a), so bq lies inside the edge's
        // leaf box and every H's filter box bh inside its leaf box: a pair that passes the
        // final bq/bh test passes the leaf-box test too, and walking with bq finds it.
        part_walk_after(tv, part, bq, 0.0, tv.leafpos[eQ], stk + threadIdx.x, [&](int eH) {
            if (!overlaps(bq, ebox[eH], 0.0)) return;  // exact leaf test (the walk's boxes are float)
            const int eI = eH > eQ ? eH : eQ, eJ = eH > eQ ? eQ : eH;
            const int2 I = se[eI], J = se[eJ];
            if (is_collider(bt[I.x]) && is_collider(bt[I.y]) && is_collider(bt[J.x]) && is_collider(bt[J.y])) return;
            if (I.x == J.x || I.x == J.y || I.y == J.x || I.y == J.y) return;
            const int2 H = se[eH];
            if (vs && vs[H.x] != scene) return;            // scenes of a batch never touch
            const Box bh = box_union(swept_pt(ld3(x, H.x), ld3(p, H.x), a), swept_pt(ld3(x, H.y), ld3(p, H.y), a));
            if (!overlaps(bq, bh, 0.0)) return;
            const d3 i0 = ld3(x, I.x), i1 = ld3(x, I.y), pi0 = ld3(p, I.x), pi1 = ld3(p, I.y);
            const d3 j0 = ld3(x, J.x), j1 = ld3(x, J.y), pj0 = ld3(p, J.x), pj1 = ld3(p, J.y);
            const double la = accd_ee(i0, i1, j0, j1, -pi0, -pi1, -pj0, -pj1, eta, 0.0, fmin(best, a));
            best = fmin(la, best);
        }, [] {}, pb);
    }
    scene_atomic<false>(sl, scene, 3, best);
}
// slot 4 = slot 3 of every scene: the point-triangle result that the edge-edge phase filters with
__global__ void copy_pt_result(int S, unsigned long long* sl) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < S) sl[kSlots * k + 4] = sl[kSlots * k + 3];
}
// Step of the swept trees from the ground bound (IPCSolver.cpp:943-958): min(1, ground), or
// 1 when that is not positive; it seeds the full-CCD minimum (slot 3) and is kept in slot 5.
__global__ void swept_step(int S, int has_surface, unsigned long long* sl) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S) return;
    unsigned long long* s = sl + kSlots * k;
    double a = has_surface ? fmin(1.0, bits_to_double(s[0])) : 1.0;
    if (a <= 0) a = 1;
    s[3] = ord_bits(a);
    s[5] = ord_bits(a);
}

// The merge of IPCSolver.cpp:938-989 from one scene's slots.
static void merge_bounds(const Scene& s, const unsigned long long* h, bool any_ccd, double out[5]) {
    double ground = 1.0;
    std::memcpy(&ground, &h[0], sizeof(double));
    const double aground = s.Vs > 0 ? std::min(1.0, ground) : 1.0;
    const double alpha = aground <= 0 ? 1.0 : aground;
    double maxmove, partial_min, full;
    std::memcpy(&maxmove, &h[1], 8);
    std::memcpy(&partial_min, &h[2], 8);
    std::memcpy(&full, &h[3], 8);
    const double partial = any_ccd ? std::min(alpha, partial_min) : alpha;
    const double acfl = maxmove > 0.0 ? std::sqrt(s.hhat) / (maxmove * 2.0) : 1.0;
    double merged = std::min(alpha, acfl);
    if (partial > 2.0 * acfl) {
        merged = std::min(partial, full);
        merged = std::max(merged, acfl);
    }
    out[0] = merged; out[1] = aground; out[2] = partial; out[3] = acfl; out[4] = full;
}

void step_bound_early(Context& c, hipStream_t st, bool ground) {
    Scene& s = c.sc;
    const int S = s.S;
    const int* vs = scene_of_vertex(s);
    const double slack = 0.8;
    const double eta = 1.0 - slack;
    DBuf<unsigned long long>& sl = c.u64scratch;
    sl.resize((size_t)kSlots * S);
    hipLaunchKernelGGL(init_bounds, dim3(blocks_for(S, 64)), dim3(64), 0, st, S, sl.p);
    if (s.Vs > 0 && ground)
        hipLaunchKernelGGL(ground_cfl_kernel, dim3(blocks_for(s.Vs, 256)), dim3(256), 0, st, s.Vs, s.sv.p, s.x.p, s.dir.p, s.gn,
                           s.gd, slack, vs, sl.p);
    CIPC_CHECK_LAUNCH();
    // full CCD with the swept trees built for each scene's ground-limited step, on the device
    hipLaunchKernelGGL(swept_step, dim3(blocks_for(S, 64)), dim3(64), 0, st, S, s.Vs > 0 ? 1 : 0, sl.p);
    CIPC_CHECK_LAUNCH();
    build_swept_trees(c, sl.p + 5, st);
    LBVH& ft = c.sface_tree;
    LBVH& et = c.sedge_tree;
    const TreeView ftv{ft.n, ft.nodes.p, ft.wide.p}, etv{et.n, et.nodes.p, et.wide.p, et.leafpos.p};
    if (s.Vs > 0 && ft.n > 0) launch_full_ccd_pt(c, st, c.tune.ccd_pt_bits);
    if (s.Es > 0 && et.n > 0) {
        hipLaunchKernelGGL(copy_pt_result, dim3(blocks_for(S, 64)), dim3(64), 0, st, S, sl.p);
        launch_full_ccd_ee(c, st, c.tune.ccd_ee_bits);
    }
    CIPC_CHECK_LAUNCH();
}

// The two full-CCD walks with each surface vertex / edge split over 2^pb lanes.
void launch_full_ccd_pt(Context& c, hipStream_t st, int pb) {
    Scene& s = c.sc;
    const LBVH& ft = c.sface_tree;
    const TreeView ftv{ft.n, ft.nodes.p, ft.wide.p};
    hipLaunchKernelGGL(full_ccd_pt_kernel, dim3(blocks_for((size_t)s.Vs << pb, 64)), dim3(64), 0, st, s.Vs, s.sv_perm.p, s.sv.p,
                       s.faces.p, s.bt.p, s.x.p, s.dir.p, ftv, c.u64scratch.p, 1.0 - 0.8, scene_of_vertex(s), pb);
}
void launch_full_ccd_ee(Context& c, hipStream_t st, int pb) {
    Scene& s = c.sc;
    const LBVH& et = c.sedge_tree;
    const TreeView etv{et.n, et.nodes.p, et.wide.p, et.leafpos.p};
    hipLaunchKernelGGL(full_ccd_ee_kernel, dim3(blocks_for((size_t)s.Es << pb, 64)), dim3(64), 0, st, s.Es, s.se_perm.p, s.se.p, s.bt.p,
                       s.x.p, s.dir.p, etv, et.prim.p, c.u64scratch.p, 1.0 - 0.8, scene_of_vertex(s), pb);
}

// ---- autotuner ----------------------------------------------------------------------
__global__ void copy_slot(int S, int dst, int src, unsigned long long* sl) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < S) sl[kSlots * k + dst] = sl[kSlots * k + src];
}

// Times every split of the two full-CCD walks and of the active set's two walks (1 to 16
// lanes per query) on the current positions and direction, three runs each after one warm-up, with HIP events on the
// context's stream, and keeps the fastest for the context's later step bounds.  Each run
// starts from the minimum the walk starts from in a step bound (slot 3 reseeded from the
// swept step, slot 5, for the point-triangle walk and from the point-triangle result, slot
// 4, for the edge-edge walk), so the pruning is the one a step bound sees.  Each walk's
// result is a minimum over pairs, so every split gives the same step.
int autotune(Context& c, TuneEntry* out, int max_out) {
    Scene& s = c.sc;
    const int S = s.S;
    if (!s.dir.p) throw Error(CIPC_ERR_LOGIC, "the autotuner times the step bound: set a direction first");
    step_bound_early(c, c.stream, true);        // trees, seeded slots, swept boxes
    CIPC_HIP(hipStreamSynchronize(c.stream));
    hipEvent_t e0, e1;
    CIPC_HIP(hipEventCreate(&e0));
    CIPC_HIP(hipEventCreate(&e1));
    int n = 0;
    for (int k = 0; k < 2; ++k) {
        const bool pt = k == 0;
        if (pt ? (s.Vs == 0 || c.sface_tree.n == 0) : (s.Es == 0 || c.sedge_tree.n == 0)) continue;
        TuneEntry e{};
        e.kernel = k;
        e.n = kTuneSplits;
        double best = 1e300;
        for (int pb = 0; pb < kTuneSplits; ++pb) {
            float total = 0.f;
            for (int r = 0; r < 4; ++r) {
                hipLaunchKernelGGL(copy_slot, dim3(blocks_for(S, 64)), dim3(64), 0, c.stream, S, 3, pt ? 5 : 4, c.u64scratch.p);
                CIPC_HIP(hipEventRecord(e0, c.stream));
                if (pt) launch_full_ccd_pt(c, c.stream, pb);
                else launch_full_ccd_ee(c, c.stream, pb);
                CIPC_HIP(hipEventRecord(e1, c.stream));
                CIPC_CHECK_LAUNCH();
                CIPC_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                CIPC_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) total += ms;
            }
            e.ms[pb] = total / 3.0;
            if (e.ms[pb] < best) { best = e.ms[pb]; e.chosen = pb; }
        }
        if (pt) c.tune.ccd_pt_bits = e.chosen;
        else c.tune.ccd_ee_bits = e.chosen;
        if (n < max_out && out) out[n] = e;
        ++n;
    }
    // the active set's walks, on the static trees of the current positions (the set is built
    // once first; each walk writes only its hit counts and buffers, so runs repeat exactly)
    active_set_now(c);
    size_active_walks(c, kTuneSplits - 1, kTuneSplits - 1);
    for (int k = 3; k <= 4; ++k) {
        const bool pt = k == 3;
        if (pt ? (s.Vs == 0 || c.face_tree.n <= 0) : (s.Es == 0 || c.edge_tree.n <= 0)) continue;
        TuneEntry e{};
        e.kernel = k;
        e.n = kTuneSplits;
        double best = 1e300;
        for (int pb = 0; pb < kTuneSplits; ++pb) {
            float total = 0.f;
            for (int r = 0; r < 4; ++r) {
                CIPC_HIP(hipEventRecord(e0, c.stream));
                launch_active_walk(c, pt, false, pb);
                CIPC_HIP(hipEventRecord(e1, c.stream));
                CIPC_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                CIPC_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) total += ms;
            }
            e.ms[pb] = total / 3.0;
            if (e.ms[pb] < best) { best = e.ms[pb]; e.chosen = pb; }
        }
        if (pt) c.tune.as_pt_bits = e.chosen;
        else c.tune.as_ee_bits = e.chosen;
        if (n < max_out && out) out[n] = e;
        ++n;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return n;
}
void step_bound_partial(Context& c, hipStream_t st) {
    Scene& s = c.sc;
    const int S = s.S;
    const double eta = 1.0 - 0.8;
    DBuf<unsigned long long>& sl = c.u64scratch;
    // (its cap is the ground step, slot 0, final once step_bound_early has run)
    // grid: the pairs' capacity (the active set's hit capacities), or the settled count
    const size_t cap = c.dc_pending ? (size_t)c.as.capPT + c.as.capEE : (size_t)c.n_ccd;
    if (cap > 0)
        hipLaunchKernelGGL(partial_ccd_kernel, dim3(blocks_for(cap, 64)), dim3(64), 0, st, c.dcnt.p + kDcCCD, c.ccd_set.p, s.sv.p,
                           s.faces.p, s.se.p, s.x.p, s.dir.p, eta, scene_of_vertex(s), sl.p);
    CIPC_CHECK_LAUNCH();
    c.pin_slots.reserve((size_t)kSlots * S);
    CIPC_HIP(hipMemcpyAsync(c.pin_slots.p, sl.p, (size_t)kSlots * S * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
}
void step_bound_launch(Context& c, hipStream_t st) {
    step_bound_early(c, st);
    step_bound_partial(c, st);
}
void step_bound_finish(Context& c, double* out) {
    const Scene& s = c.sc;
    for (int k = 0; k < s.S; ++k) merge_bounds(s, c.pin_slots.p + (size_t)kSlots * k, c.n_ccd > 0, out + 5 * (size_t)k);
}
void step_bound_scenes(Context& c, double* out) {
    step_bound_launch(c, c.stream);
    CIPC_HIP(hipStreamSynchronize(c.stream));
    step_bound_finish(c, out);
}

// Self_largestFeasibleStepSize_CCD seeded with step 1 (the animated Dirichlet update,
// IPCSolver.cpp:1131-1134): the swept trees and the full CCD of step_bound_early with the
// ground slot left at its initial 1.
double full_ccd_bound(Context& c) {
    Scene& s = c.sc;
    if (s.S != 1) throw Error(CIPC_ERR_LOGIC, "a batch has one step bound per scene");
    step_bound_early(c, c.stream, false);
    c.pin_slots.reserve(kSlots);
    CIPC_HIP(hipMemcpyAsync(c.pin_slots.p, c.u64scratch.p, kSlots * sizeof(unsigned long long), hipMemcpyDeviceToHost, c.stream));
    CIPC_HIP(hipStreamSynchronize(c.stream));
    double full;
    std::memcpy(&full, &c.pin_slots.p[3], 8);
    return full;
}

void step_bound(Context& c, double out[5]) {
    if (c.sc.S != 1) throw Error(CIPC_ERR_LOGIC, "a batch has one step bound per scene");
    step_bound_scenes(c, out);
}

}  // namespace cipc
