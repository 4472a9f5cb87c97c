// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate expressions of CPU-IPC
// (MPL-2.0) in the reference's operation order; each such part cites the source it follows.
// Contact stage of the Newton assembly: ground barrier, self-contact barriers and the
// mollified edge-edge barrier.  Compiled with -ffp-contract=off.
//
// Reference: ground loop (CPU IPC/IPCSolver.cpp:270-293), compute_g_dpt
// (ContactMechanics.cpp:2970-3032), compute_g_dee (:3034-3080), compute_H_dpt
// (:3466-3573), compute_H_dee (:3575-3700).
//
// Each contact becomes a local Hessian over its 2-4 vertices (kernel 1), is projected to
// PSD by register-resident Jacobi (kernel 2), and is scattered as lower-triangle 3x3
// block records keyed (column vertex << kb | row vertex, 2^kb > V) plus per-vertex gradient
// records (kernel 3).  The records are sorted and summed by key in sparse.hip, so the
// result does not depend on thread scheduling.
#include "context.hpp"
#include "contact.hpp"

namespace cipc {

constexpr int kHU = 78;                  // upper triangle of a 12x12
constexpr unsigned long long kNoKey = ~0ull;

__device__ inline int nv_of(int type) { return type == 0 ? 2 : (type == 1 ? 3 : 4); }

// Decode an EncodedContact (EncodedContact.h; SelfConstraintVal, ContactMechanics.cpp:2478-2503):
// type 0 PP, 1 PE, 2 PT, 3 EE; vertices in the reference's argument order.
__device__ inline int decode(int4 c, int v[4]) {
    if (c.x >= 0) { v[0] = c.x; v[1] = c.y; v[2] = c.z; v[3] = c.w; return 3; }
    v[0] = -c.x - 1; v[1] = c.y; v[2] = c.z; v[3] = c.w;
    if (c.z < 0) return 0;
    if (c.w < 0) return 1;
    return 2;
}
__device__ inline double self_value(int type, const d3* p) {
    switch (type) {
    case 0: return d_PP(p[0], p[1]);
    case 1: return d_PE(p[0], p[1], p[2]);
    case 2: return d_PT(p[0], p[1], p[2], p[3]);
    default: return d_EE(p[0], p[1], p[2], p[3]);
    }
}
// Gradient and full Hessian (N x N row-major) of the distance of `type` at packed points.
__device__ inline void dist_gh(int type, const double* q, double* g, double* H) {
    if (type == 0) {
        const double e[3] = {q[0] - q[3], q[1] - q[4], q[2] - q[5]};
        for (int i = 0; i < 3; ++i) { g[i] = 2.0 * e[i]; g[3 + i] = -g[i]; }
        for (int i = 0; i < 36; ++i) H[i] = 0.0;
        for (int i = 0; i < 6; ++i) { H[i * 6 + i] = 2.0; }
        for (int i = 0; i < 3; ++i) { H[i * 6 + 3 + i] = -2.0; H[(3 + i) * 6 + i] = -2.0; }
    } else if (type == 1) {
        dist_pe_gh(q, g, H);
    } else if (type == 2) {
        dist_pt_gh(q, g, H);
    } else {
        dist_ee_gh(q, g, H);
    }
}

// Kernel 1: local Hessian (upper triangle, SoA) and local gradient of every contact, in
// two launches over the contact range [0, n_self + n_ee):
//   contact_self_kernel, c < n_self: Self_ActiveSet entry, compute_H_dpt / compute_g_dpt
//   contact_ee_kernel, otherwise   : Self_EE_ActiveSet entry, compute_H_dee / compute_g_dee
// The self form is instantiated per distance type, so that its arrays have indices fixed at
// compile time and stay in registers.
template <int Type>
__device__ inline void self_local(int c, int n, int4 e, const int (&v)[4], const double* __restrict__ x, double hhat, double kappa,
                                  double* __restrict__ Hs, double* __restrict__ Gs) {
    constexpr int nv = Type == 0 ? 2 : (Type == 1 ? 3 : 4), N = 3 * nv;
    double H[144], g[12];       // indices fixed at compile time: only the first N (N x N) live
    d3 p[4];
    double q[12];
#pragma unroll
    for (int k = 0; k < nv; ++k) { p[k] = ld3(x, v[k]); q[3 * k] = p[k].x; q[3 * k + 1] = p[k].y; q[3 * k + 2] = p[k].z; }
    const double d = self_value(Type, p);
    dist_gh(Type, q, g, H);
    const double gb = barrier_g(d, hhat), hb = barrier_h(d, hhat);
    const double coef = (Type <= 1) ? kappa * -e.w : kappa;
    const double cH = coef * hb, cg = coef * gb;
    int idx = 0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = i; j < N; ++j) Hs[(size_t)(idx++) * n + c] = (cH * g[i]) * g[j] + cg * H[i * N + j];
    const double gs = (Type <= 1) ? (kappa * -e.w) * gb : kappa * gb;
#pragma unroll
    for (int i = 0; i < N; ++i) Gs[(size_t)i * n + c] = gs * g[i];
}
// Contacts of the records: Self_ActiveSet entries [0, ns), then Self_EE_ActiveSet entries,
// from the active set's device counts; none when they exceed the records' capacity (the
// host grows it and assembles again).
__device__ inline int record_contacts(const int* __restrict__ dc, int& ns) {
    if (dc[kDcContactOver]) { ns = 0; return 0; }
    ns = dc[kDcSelf];
    return ns + dc[kDcMol];
}
__global__ void contact_totals(int capS, int* __restrict__ dc) {
    if (blockIdx.x == 0 && threadIdx.x == 0) dc[kDcContactOver] = dc[kDcSelf] + dc[kDcMol] > capS ? 1 : 0;
}
__global__ __launch_bounds__(64) void contact_self_kernel(const int* __restrict__ dc, int n, const int4* __restrict__ self_set,
                                                          const double* __restrict__ x, double hhat, Kap kap, double* __restrict__ Hs,
                                                          double* __restrict__ Gs, int4* __restrict__ verts, int* __restrict__ dims) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    int n_self;
    record_contacts(dc, n_self);
    if (c >= n_self) return;
    const int4 e = self_set[c];
    int v[4];
    const int type = decode(e, v);
    const double kappa = kap.at(v[0]);
    switch (type) {
    case 0: self_local<0>(c, n, e, v, x, hhat, kappa, Hs, Gs); break;
    case 1: self_local<1>(c, n, e, v, x, hhat, kappa, Hs, Gs); break;
    case 2: self_local<2>(c, n, e, v, x, hhat, kappa, Hs, Gs); break;
    default: self_local<3>(c, n, e, v, x, hhat, kappa, Hs, Gs); break;
    }
    verts[c] = make_int4(v[0], v[1], type >= 1 ? v[2] : -1, type >= 2 ? v[3] : -1);
    dims[c] = 3 * nv_of(type);
}
__global__ __launch_bounds__(64) void contact_ee_kernel(
    const int* __restrict__ dc, int n, const int4* __restrict__ ee_set,
    const int2* __restrict__ ee_ie, const int2* __restrict__ se, const double* __restrict__ x,
    const double* __restrict__ vrest, double hhat, Kap kap, double* __restrict__ Hs, double* __restrict__ Gs,
    int4* __restrict__ verts, int* __restrict__ dims) {
    int n_self;
    const int nc = record_contacts(dc, n_self);
    const int c = n_self + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= nc) return;
    double H[144];
    {
        const int i = c - n_self;
        const int4 e = ee_set[i];
        int v[4];
        const int type = decode(e, v);
        const double kappa = kap.at(v[0]);
        const int nvd = nv_of(type);
        int e4[4];
        if (e.w >= 0) { e4[0] = e.x; e4[1] = e.y; e4[2] = e.z; e4[3] = e.w; }
        else {
            const int2 ie = ee_ie[i];
            const int2 a = se[ie.x], b = se[ie.y];
            e4[0] = a.x; e4[1] = a.y; e4[2] = b.x; e4[3] = b.y;
        }
        d3 p[4], pe[4];
        for (int k = 0; k < nvd; ++k) p[k] = ld3(x, v[k]);
        for (int k = 0; k < 4; ++k) pe[k] = ld3(x, e4[k]);
        const double d = self_value(type, p);
        const double b = barrier_b(d, hhat), gb = barrier_g(d, hhat), hb = barrier_h(d, hhat);
        const double ex = eps_x_of(ld3(vrest, e4[0]), ld3(vrest, e4[1]), ld3(vrest, e4[2]), ld3(vrest, e4[3]));
        const double cr = ee_cross_sqn(pe[0], pe[1], pe[2], pe[3]);
        const double mol = mol_e(cr, ex);
        // distance gradient/Hessian embedded over e4 (indMap of ContactMechanics.cpp:3627-3685)
        double qd[12];
        for (int k = 0; k < nvd; ++k) { qd[3 * k] = p[k].x; qd[3 * k + 1] = p[k].y; qd[3 * k + 2] = p[k].z; }
        double gl[12];
        dist_gh(type, qd, gl, H);
        int map[4] = {0, 1, 2, 3};
        if (type != 3) {
            for (int k = 0; k < 4; ++k) {
                if (v[0] == e4[k]) map[0] = k;
                else if (v[1] == e4[k]) map[1] = k;
                else if (nvd == 3 && v[2] == e4[k]) map[2] = k;
            }
        }
        const int Nd = 3 * nvd;
        double gd[12];
        for (int k = 0; k < 12; ++k) gd[k] = 0.0;
        for (int a = 0; a < nvd; ++a)
            for (int r = 0; r < 3; ++r) gd[3 * map[a] + r] = gl[3 * a + r];
        // mollifier gradient and Hessian of the cross norm
        double qe[12], gc[12], Hc[144];
        for (int k = 0; k < 4; ++k) { qe[3 * k] = pe[k].x; qe[3 * k + 1] = pe[k].y; qe[3 * k + 2] = pe[k].z; }
        ee_cross_gh(qe, gc, Hc);
        double ge[12], qg = 0.0, qH = 0.0;
        const bool act = cr < ex;
        if (act) {
            const double oe = 1.0 / ex;
            qg = 2.0 * oe * (-oe * cr + 1.0);
            qH = -2.0 / (ex * ex);
        }
        for (int k = 0; k < 12; ++k) ge[k] = act ? gc[k] * qg : 0.0;
        const double kgb = kappa * gb, kb = kappa * b, keh = kappa * mol * hb, keg = kappa * mol * gb;
        int idx = 0;
        for (int r = 0; r < 12; ++r)
            for (int s = r; s < 12; ++s) {
                // embedded distance Hessian entry
                double hd = 0.0;
                const int vr = r / 3, vs = s / 3;
                int lr = -1, ls = -1;
                for (int a = 0; a < nvd; ++a) { if (map[a] == vr) lr = a; if (map[a] == vs) ls = a; }
                if (lr >= 0 && ls >= 0) hd = H[(3 * lr + r % 3) * Nd + 3 * ls + s % 3];
                const double eH = act ? Hc[r * 12 + s] * qg + (qH * gc[r]) * gc[s] : 0.0;
                const double kgg = (kgb * gd[r]) * ge[s];
                const double kggT = (kgb * gd[s]) * ge[r];
                Hs[(size_t)(idx++) * n + c] = kgg + kggT + kb * eH + (keh * gd[r]) * gd[s] + keg * hd;
            }
        const double gsc = kappa * (gb * mol);
        for (int k = 0; k < 12; ++k) Gs[(size_t)k * n + c] = kb * ge[k] + gsc * gd[k];
        verts[c] = make_int4(e4[0], e4[1], e4[2], e4[3]);
        dims[c] = 12;
    }
}

// Kernel 2: makePD of each local Hessian, in place, one launch per vertex count and kind,
// on the translation-free subspace.  Self_ActiveSet blocks (point-triangle, edge-edge,
// point-edge distances) use the structured range basis (project_contact_structured); the
// mollified edge-edge blocks and point-point blocks the general Jacobi (project_contact).
// The packed index of (i, j) is the one kernel 1 writes.
// range 0: the Self_ActiveSet contacts, 1: the Self_EE_ActiveSet ones.
template <int M, bool Structured>
__device__ inline void contact_pd(int t, const int* __restrict__ dc, int range, int n, const int* __restrict__ dims,
                                  double* __restrict__ Hs, const double* __restrict__ Gs) {
    int ns;
    const int nc = record_contacts(dc, ns);
    const int c0 = range == 0 ? 0 : ns, c1 = range == 0 ? ns : nc;
    const int c = c0 + t;
    if (c >= c1 || dims[c] != 3 * M) return;
    double* h = Hs + c;
    auto acc = [&](int k) -> double& { return h[(size_t)k * n]; };
    if constexpr (Structured) {
        double G[3 * M];
#pragma unroll
        for (int k = 0; k < 3 * M; ++k) G[k] = Gs[(size_t)k * n + c];
        project_contact_structured<M>(acc, G);
    } else {
        project_contact<M>(acc);
    }
}
template <int M, bool Structured>
__global__ __launch_bounds__(64) void contact_pd_kernel(const int* __restrict__ dc, int range, int n, const int* __restrict__ dims,
                                                        double* __restrict__ Hs, const double* __restrict__ Gs) {
    contact_pd<M, Structured>(blockIdx.x * blockDim.x + threadIdx.x, dc, range, n, dims, Hs, Gs);
}
// Both 4-vertex kinds in one launch (the same registers either way): row y = 0 of work-groups
// the Self_ActiveSet contacts (structured basis), y = 1 the Self_EE_ActiveSet ones.
__global__ __launch_bounds__(64) void contact_pd4_kernel(const int* __restrict__ dc, int n, const int* __restrict__ dims,
                                                         double* __restrict__ Hs, const double* __restrict__ Gs) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.y == 0) contact_pd<4, true>(t, dc, 0, n, dims, Hs, Gs);
    else contact_pd<4, false>(t, dc, 1, n, dims, Hs, Gs);
}

// Kernel 3: block and gradient records.  Slots: 10 blocks and 4 gradients per contact.  With
// dc, the contacts are the active set's (device count) in slots of n contacts, and the slots
// past the count are written empty; without, n contacts.
__global__ void contact_scatter_kernel(int n, const int* __restrict__ dc, const int4* __restrict__ verts, const int* __restrict__ dims,
                                       const double* __restrict__ Hs, const double* __restrict__ Gs,
                                       const int* __restrict__ bt, int slot0_b, int slot0_g, int kb,
                                       unsigned long long* __restrict__ bkey, int* __restrict__ bidx, double* __restrict__ bval,
                                       int* __restrict__ gkey, int* __restrict__ gidx, double* __restrict__ gval) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    if (dc) {
        int ns;
        if (c >= record_contacts(dc, ns)) {
            for (int k = 0; k < 10; ++k) {
                const int s = slot0_b + 10 * c + k;
                bidx[s] = s;
                bkey[s] = kNoKey;
            }
            for (int k = 0; k < 4; ++k) {
                const int s = slot0_g + 4 * c + k;
                gidx[s] = s;
                gkey[s] = 0x7fffffff;
            }
            return;
        }
    }
    const int N = dims[c], nv = N / 3;
    const int4 vv = verts[c];
    const int v[4] = {vv.x, vv.y, vv.z, vv.w};
    // upper-triangle index of (i, j), i <= j, in an N x N
    auto ux = [N](int i, int j) { return i * N - i * (i - 1) / 2 + (j - i); };
    int slot = 0;
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b <= a; ++b, ++slot) {
            const int s = slot0_b + 10 * c + slot;
            bidx[s] = s;
            if (a >= nv || !is_free(bt[v[a]]) || !is_free(bt[v[b]])) { bkey[s] = kNoKey; continue; }
            // global lower block: rows of the larger vertex id
            const int ra = v[a] >= v[b] ? a : b, cb = v[a] >= v[b] ? b : a;
            const int R = v[ra], Cc = v[cb];
            bkey[s] = ((unsigned long long)(unsigned)Cc << kb) | (unsigned)R;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    const int gi = 3 * ra + i, gj = 3 * cb + j;
                    const int lo = gi < gj ? gi : gj, hi = gi < gj ? gj : gi;
                    bval[(size_t)9 * s + 3 * i + j] = Hs[(size_t)ux(lo, hi) * n + c];
                }
        }
    for (int a = 0; a < 4; ++a) {
        const int s = slot0_g + 4 * c + a;
        gidx[s] = s;
        if (a >= nv) { gkey[s] = 0x7fffffff; continue; }
        gkey[s] = v[a];
        for (int i = 0; i < 3; ++i) gval[(size_t)3 * s + i] = Gs[(size_t)(3 * a + i) * n + c];
    }
}

// Ground (IPCSolver.cpp:270-293): one block and one gradient record per active vertex, in
// slots [0, cap); slots past the device count are written empty.
// Its first thread also checks the contacts against their capacity (contact_totals).
__global__ void ground_kernel(int cap, int* __restrict__ dc, int capS, const int* __restrict__ env, const double* __restrict__ x,
                              const int* __restrict__ bt, d3 gn, double gd, double hhat, Kap kap, int kb,
                              unsigned long long* __restrict__ bkey, int* __restrict__ bidx, double* __restrict__ bval,
                              int* __restrict__ gkey, int* __restrict__ gidx, double* __restrict__ gval) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) dc[kDcContactOver] = dc[kDcSelf] + dc[kDcMol] > capS ? 1 : 0;
    if (i >= cap) return;
    if (i >= dc[kDcEnv]) {
        bidx[i] = i; bkey[i] = kNoKey;
        gidx[i] = i; gkey[i] = 0x7fffffff;
        return;
    }
    const int v = env[i];
    const double kappa = kap.at(v);
    const double dist = dot(gn, ld3(x, v)) - gd;
    const double d = dist * dist;
    const double g_b = barrier_g(d, hhat), H_b = barrier_h(d, hhat);
    const d3 g = (((kappa * g_b) * 2.0) * dist) * gn;
    gkey[i] = v; gidx[i] = i;
    gval[3 * i] = g.x; gval[3 * i + 1] = g.y; gval[3 * i + 2] = g.z;
    const double param = 4.0 * H_b * d + 2.0 * g_b;
    bidx[i] = i;
    if (param > 0 && is_free(bt[v])) {
        bkey[i] = ((unsigned long long)(unsigned)v << kb) | (unsigned)v;
        const double nn[3] = {gn.x, gn.y, gn.z};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) bval[(size_t)9 * i + 3 * a + b] = (kappa * param) * (nn[a] * nn[b]);
    } else {
        bkey[i] = kNoKey;
    }
}

// ---------------------------------------------------------------------------------
// Lagged friction (Friction.cpp; FrictionKinematics.h with SFCLAMPING_ORDER 1)
// ---------------------------------------------------------------------------------
__device__ inline double f0_SF(double x2, double e) { return x2 * (-sqrt(x2) / 3.0 + e) / (e * e) + e / 3.0; }  // :262-265
__device__ inline double f1_SF_div(double x2, double e) { return (-sqrt(x2) + 2.0 * e) / (e * e); }           // :267-270
__device__ inline double f2_SF(double x2, double e) { return 2.0 * (e - sqrt(x2)) / (e * e); }               // :272-275
__device__ inline d3 normalized(d3 v) {   // Eigen normalized(): v / |v|, unchanged when |v| = 0
    const double n2 = sqn(v);
    if (!(n2 > 0.0)) return v;
    const double n = sqrt(n2);
    return d3{v.x / n, v.y / n, v.z / n};
}
// Friction::initialize for one Self_ActiveSet entry (Friction.cpp:88-137): lambda, closest
// point coordinates (computeClosestPoint_*) and tangent basis (computeTangentBasis_*).
__global__ void friction_init_self(int n, const int4* __restrict__ set, const double* __restrict__ x, double hhat, Kap kap,
                                   int4* __restrict__ fset, double* __restrict__ lam, double* __restrict__ coord,
                                   double* __restrict__ basis) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int4 e = set[c];
    int v[4];
    const int type = decode(e, v);
    const double kappa = kap.at(v[0]);
    d3 p[4];
    for (int k = 0; k < nv_of(type); ++k) p[k] = ld3(x, v[k]);
    const double d = self_value(type, p);
    double l = barrier_g(d, hhat);
    l *= -kappa * 2.0 * sqrt(d);
    if (e.w < -1) l *= -e.w;
    double b1 = 0.0, b2 = 0.0;
    d3 B0, B1;
    if (type == 3) {            // edge-edge (FrictionKinematics.h:81-114)
        const d3 e20 = p[0] - p[2], e01 = p[1] - p[0], e23 = p[3] - p[2];
        ldlt2_solve(sqn(e01), -dot(e23, e01), sqn(e23), -dot(e20, e01), dot(e20, e23), b1, b2);
        B0 = normalized(e01);
        B1 = normalized(cross(cross(e01, p[3] - p[2]), e01));
    } else if (type == 0) {     // point-point (:211-227)
        const d3 v01 = p[1] - p[0];
        const d3 xc = cross(d3{1.0, 0.0, 0.0}, v01), yc = cross(d3{0.0, 1.0, 0.0}, v01);
        if (sqn(xc) > sqn(yc)) { B0 = normalized(xc); B1 = normalized(cross(v01, xc)); }
        else { B0 = normalized(yc); B1 = normalized(cross(v01, yc)); }
    } else if (type == 1) {     // point-edge (:155-172)
        const d3 e12 = p[2] - p[1];
        b1 = dot(p[0] - p[1], e12) / sqn(e12);
        B0 = normalized(e12);
        B1 = normalized(cross(e12, p[0] - p[1]));
    } else {                    // point-triangle (:18-41)
        const d3 r0 = p[2] - p[1], r1 = p[3] - p[1], q = p[0] - p[1];
        ldlt2_solve(dot(r0, r0), dot(r1, r0), dot(r1, r1), dot(r0, q), dot(r1, q), b1, b2);
        B0 = normalized(r0);
        B1 = normalized(cross(cross(r0, r1), r0));
    }
    fset[c] = e;
    lam[c] = l;
    coord[2 * c] = b1; coord[2 * c + 1] = b2;
    st3(basis + 6 * (size_t)c, 0, B0);
    st3(basis + 6 * (size_t)c, 1, B1);
}
__global__ void friction_init_ground(int n, const int* __restrict__ env, const double* __restrict__ x, d3 gn, double gd, double hhat,
                                     Kap kap, int* __restrict__ fenv, double* __restrict__ lam) {      // Friction.cpp:141-152
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int v = env[c];
    const double kappa = kap.at(v);
    const double dist = dot(gn, ld3(x, v)) - gd;
    const double d = dist * dist;
    double l = barrier_g(d, hhat);
    l *= -kappa * 2.0 * sqrt(d);
    fenv[c] = v;
    lam[c] = l;
}

// A lagged contact's vertices, relative displacement (computeRelDX_*) and lift
// (liftRelDXTanToMesh_*) of a tangent direction.
struct FrictionPair {
    int type, nv, v[4];
    double b1, b2;
    d3 B0, B1;
    __device__ FrictionPair(int4 e, const double* coord, const double* basis, int c) {
        type = decode(e, v);
        nv = nv_of(type);
        b1 = coord[2 * c]; b2 = coord[2 * c + 1];
        B0 = ld3(basis + 6 * (size_t)c, 0);
        B1 = ld3(basis + 6 * (size_t)c, 1);
    }
    __device__ d3 rel_dx(const double* x, const double* xp) const {
        d3 dx[4];
        for (int k = 0; k < nv; ++k) dx[k] = ld3(x, v[k]) - ld3(xp, v[k]);
        if (type == 3) return dx[0] + b1 * (dx[1] - dx[0]) - (dx[2] + b2 * (dx[3] - dx[2]));          // :117-126
        if (type == 0) return dx[0] - dx[1];                                                        // :229-235
        if (type == 1) return dx[0] - (dx[1] + b1 * (dx[2] - dx[1]));                               // :174-181
        return dx[0] - (dx[1] + b1 * (dx[2] - dx[1]) + b2 * (dx[3] - dx[1]));                        // :43-51
    }
    // L holds all four vertices' rows, those past nv zero (a fixed-size array the compiler
    // keeps in registers).
    __device__ void lift(double t0, double t1, double (&L)[12]) const {
        const d3 T = t0 * B0 + t1 * B1;
        const d3 z{0.0, 0.0, 0.0};
        d3 s[4] = {z, z, z, z};
        if (type == 3) { s[0] = (1.0 - b1) * T; s[1] = b1 * T; s[2] = (b2 - 1.0) * T; s[3] = -b2 * T; }        // :128-139
        else if (type == 0) { s[0] = T; s[1] = -1.0 * T; }                                                   // :237-244
        else if (type == 1) { s[0] = T; s[1] = (b1 - 1.0) * T; s[2] = -b1 * T; }                            // :183-192
        else { s[0] = T; s[1] = (-1 + b1 + b2) * T; s[2] = -b1 * T; s[3] = -b2 * T; }                      // :53-63
#pragma unroll
        for (int k = 0; k < 4; ++k) { L[3 * k] = s[k].x; L[3 * k + 1] = s[k].y; L[3 * k + 2] = s[k].z; }
    }
};

// Friction::addGradient / addHessian for the lagged self contacts (Friction.cpp:158-289,
// 349-457): local gradient and the analytic PSD Hessian (analyticFrictionHessian, :11-64),
// in the same SoA layout as the contact blocks so the contact scatter writes their records.
__global__ __launch_bounds__(64) void friction_local_kernel(int n, const int4* __restrict__ fset, const double* __restrict__ lam,
                                                            const double* __restrict__ coord, const double* __restrict__ basis,
                                                            const double* __restrict__ x, const double* __restrict__ xp, double eps2,
                                                            double coef, double* __restrict__ Hs, double* __restrict__ Gs,
                                                            int4* __restrict__ verts, int* __restrict__ dims) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double eps = sqrt(eps2);
    const FrictionPair fp(fset[c], coord, basis, c);
    const int N = 3 * fp.nv;
    const d3 r3 = fp.rel_dx(x, xp);
    const double u0 = dot(r3, fp.B0), u1 = dot(r3, fp.B1);
    // gradient (:172-190)
    double t0 = u0, t1 = u1;
    const double sq = u0 * u0 + u1 * u1;
    if (sq > eps2) { const double nr = sqrt(sq); t0 /= nr; t1 /= nr; }
    else { const double f1 = f1_SF_div(sq, eps); t0 *= f1; t1 *= f1; }
    double L[12];
    fp.lift(t0, t1, L);
    const double gs = coef * lam[c];
#pragma unroll
    for (int i = 0; i < 12; ++i)
        if (i < N) Gs[(size_t)i * n + c] = L[i] * gs;
    // Hessian: at most two nonnegative eigenpairs in the tangent plane, each w e l l^T with l
    // the lift of its tangent direction; the packed upper triangle is written term by term, the
    // first pair's term then the second's, with every array index fixed at compile time
    const double w = fmax(0.0, coef * lam[c]);
    double la[12], lb[12], sa = 0.0, sb = 0.0;
    bool ha = false, hb = false;
    auto pair = [&](double d0, double d1, double ev, double (&l)[12], double& sc, bool& has) {
        const double e = fmax(0.0, ev);
        has = e != 0.0;
        sc = w * e;
        fp.lift(d0, d1, l);
    };
    if (w != 0.0) {
        const double nr = sqrt(sq);
        if (nr > 0.0) {
            const double p0 = u0 / nr, p1 = u1 / nr;
            if (nr > eps) pair(-p1, p0, 1.0 / nr, la, sa, ha);
            else {
                const double pe = f1_SF_div(sq, eps), pa = f2_SF(sq, eps);
                pair(p0, p1, pa, la, sa, ha);
                pair(-p1, p0, pe, lb, sb, hb);
            }
        } else {
            const double iso = f1_SF_div(0.0, eps);
            pair(1.0, 0.0, iso, la, sa, ha);
            pair(0.0, 1.0, iso, lb, sb, hb);
        }
    }
#pragma unroll
    for (int i = 0; i < 12; ++i)
#pragma unroll
        for (int j = i; j < 12; ++j)
            if (j < N) {
                double h = 0.0;
                if (ha) h += (sa * la[i]) * la[j];
                if (hb) h += (sb * lb[i]) * lb[j];
                const int k = i * N - i * (i - 1) / 2 + (j - i);     // packed (i, j), rows i < N
                Hs[(size_t)k * n + c] = h;
            }
    verts[c] = make_int4(fp.v[0], fp.v[1], fp.nv >= 3 ? fp.v[2] : -1, fp.nv >= 4 ? fp.v[3] : -1);
    dims[c] = N;
}
// Ground friction records (Friction.cpp:283-300,310-347): one gradient and one 3x3 block.
__global__ void friction_ground_kernel(int n, const int* __restrict__ fenv, const double* __restrict__ lam, const double* __restrict__ x,
                                       const double* __restrict__ xp, const int* __restrict__ bt, d3 gn, double eps2, double coef, int kb,
                                       int slot_b, int slot_g, unsigned long long* __restrict__ bkey, int* __restrict__ bidx, double* __restrict__ bval,
                                       int* __restrict__ gkey, int* __restrict__ gidx, double* __restrict__ gval) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double eps = sqrt(eps2);
    const int v = fenv[c];
    const d3 d = ld3(x, v) - ld3(xp, v);
    const d3 pj = d - dot(d, gn) * gn;
    const double mag2 = sqn(pj);
    const d3 g = mag2 > eps2 ? (coef * lam[c] / sqrt(mag2)) * pj : (coef * lam[c] / eps) * pj;
    const int sg = slot_g + c, sb = slot_b + c;
    gkey[sg] = v; gidx[sg] = sg;
    gval[3 * (size_t)sg] = g.x; gval[3 * (size_t)sg + 1] = g.y; gval[3 * (size_t)sg + 2] = g.z;
    // tangent basis of the ground (Friction.cpp:66-76)
    const d3 helper = fabs(gn.x) < 0.9 ? d3{1.0, 0.0, 0.0} : d3{0.0, 1.0, 0.0};
    const d3 gb0 = normalized(helper - dot(helper, gn) * gn);
    const d3 gb1 = normalized(cross(gn, gb0));
    const double mult = fmax(0.0, coef * lam[c]);
    const double r0 = dot(gb0, d), r1 = dot(gb1, d);
    const double nr = sqrt(r0 * r0 + r1 * r1);
    double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (mult > 0.0) {
        const double b0[3] = {gb0.x, gb0.y, gb0.z}, b1[3] = {gb1.x, gb1.y, gb1.z};
        if (nr > eps) {
            const double p0 = r0 / nr, p1 = r1 / nr;
            const d3 l = (-p1) * gb0 + p0 * gb1;
            const double lv[3] = {l.x, l.y, l.z};
            const double sc = mult / nr;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) H[3 * i + j] = (sc * lv[i]) * lv[j];
        } else {
            const double sc = mult / eps;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) H[3 * i + j] = sc * (b0[i] * b0[j] + b1[i] * b1[j]);
        }
    }
    bidx[sb] = sb;
    if (is_free(bt[v])) {
        bkey[sb] = ((unsigned long long)(unsigned)v << kb) | (unsigned)v;
        for (int k = 0; k < 9; ++k) bval[(size_t)9 * sb + k] = H[k];
    } else {
        bkey[sb] = kNoKey;
    }
}

// Friction::energy parts (Friction.cpp:459-545): sum over self contacts of lambda times the
// clamped tangential norm, and sum over ground contacts of mu lambda times theirs.
__device__ inline double friction_self_term(int c, const int4* fset, const double* lam, const double* coord, const double* basis,
                                            const double* x, const double* xp, double eps2, double eps) {
    const FrictionPair fp(fset[c], coord, basis, c);
    const d3 r3 = fp.rel_dx(x, xp);
    const double a = dot(r3, fp.B0), b = dot(r3, fp.B1);
    const double sq = a * a + b * b;
    return sq > eps2 ? lam[c] * sqrt(sq) : lam[c] * f0_SF(sq, eps);
}
__device__ inline double friction_ground_term(int c, const int* fenv, const double* lam, const double* x, const double* xp, d3 gn,
                                              double eps2, double eps, double mu) {
    const int v = fenv[c];
    const d3 d = ld3(x, v) - ld3(xp, v);
    const d3 pj = d - dot(d, gn) * gn;
    const double sq = sqn(pj);
    const double cl = sq > eps2 ? sqrt(sq) - 0.5 * eps : 0.5 * sq / eps;
    return mu * lam[c] * cl;
}
constexpr int kFB = 256;
__global__ __launch_bounds__(kFB) void e_friction(int ns, int ng, const int4* __restrict__ fset, const double* __restrict__ lam,
                                                  const double* __restrict__ coord, const double* __restrict__ basis,
                                                  const int* __restrict__ fenv, const double* __restrict__ glam,
                                                  const double* __restrict__ x, const double* __restrict__ xp, d3 gn, double eps2,
                                                  double mu, double* __restrict__ part, double* __restrict__ es,
                                                  double* __restrict__ eg) {
    // part[blockIdx.x] = self partial, part[gridDim.x + blockIdx.x] = ground partial; with
    // es / eg the per-contact values are stored instead
    __shared__ double sh[2][kFB];
    const double eps = sqrt(eps2);
    double as = 0.0, ag = 0.0;
    for (int c = blockIdx.x * kFB + threadIdx.x; c < ns; c += gridDim.x * kFB) {
        const double v = friction_self_term(c, fset, lam, coord, basis, x, xp, eps2, eps);
        if (es) es[c] = v;
        else as += v;
    }
    for (int c = blockIdx.x * kFB + threadIdx.x; c < ng; c += gridDim.x * kFB) {
        const double v = friction_ground_term(c, fenv, glam, x, xp, gn, eps2, eps, mu);
        if (eg) eg[c] = v;
        else ag += v;
    }
    if (part == nullptr) return;   // per-contact values only (uniform over the work-group)
    sh[0][threadIdx.x] = as;
    sh[1][threadIdx.x] = ag;
    __syncthreads();
    for (int s = kFB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { sh[0][threadIdx.x] += sh[0][threadIdx.x + s]; sh[1][threadIdx.x] += sh[1][threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[blockIdx.x] = sh[0][0]; part[gridDim.x + blockIdx.x] = sh[1][0]; }
}

void friction_update(Context& c, double kappa) {
    Scene& s = c.sc;
    const int ns = (int)c.n_self, ne = (int)c.n_env;
    c.fr_self.resize(std::max(ns, 1)); c.fr_lam.resize(std::max(ns, 1));
    c.fr_coord.resize(2 * (size_t)std::max(ns, 1)); c.fr_basis.resize(6 * (size_t)std::max(ns, 1));
    c.fr_env.resize(std::max(ne, 1)); c.fr_env_lam.resize(std::max(ne, 1));
    if (ns > 0)
        hipLaunchKernelGGL(friction_init_self, dim3(blocks_for(ns, 128)), dim3(128), 0, c.stream, ns, c.self_set.p, s.x.p, s.hhat,
                           scene_kappa(s, kappa), c.fr_self.p, c.fr_lam.p, c.fr_coord.p, c.fr_basis.p);
    if (ne > 0)
        hipLaunchKernelGGL(friction_init_ground, dim3(blocks_for(ne, 256)), dim3(256), 0, c.stream, ne, c.env_set.p, s.x.p, s.gn, s.gd,
                           s.hhat, scene_kappa(s, kappa), c.fr_env.p, c.fr_env_lam.p);
    CIPC_CHECK_LAUNCH();
    c.n_fself = ns;
    c.n_fenv = ne;
}

void friction_records(Context& c, size_t slot_b, size_t slot_g, hipStream_t st) {
    Scene& s = c.sc;
    const int nf = (int)c.n_fself, ngf = (int)c.n_fenv;
    const double eps2 = s.fhat * s.dt * s.dt;
    if (ngf > 0) {
        hipLaunchKernelGGL(friction_ground_kernel, dim3(blocks_for(ngf, 256)), dim3(256), 0, st, ngf, c.fr_env.p, c.fr_env_lam.p,
                           s.x.p, s.vprev.p, s.bt.p, s.gn, eps2, s.fric, s.kb, (int)slot_b, (int)slot_g, c.ckey.p, c.cidx.p, c.cval.p,
                           c.gkey.p, c.gidx.p, c.gval.p);
        CIPC_CHECK_LAUNCH();
    }
    if (nf > 0) {
        double* Hs = c.fhess.p;
        double* Gs = c.fhess.p + (size_t)kHU * nf;
        hipLaunchKernelGGL(friction_local_kernel, dim3(blocks_for(nf, 64)), dim3(64), 0, st, nf, c.fr_self.p, c.fr_lam.p,
                           c.fr_coord.p, c.fr_basis.p, s.x.p, s.vprev.p, eps2, s.fric, Hs, Gs, c.fverts.p, c.fdims.p);
        CIPC_CHECK_LAUNCH();
        hipLaunchKernelGGL(contact_scatter_kernel, dim3(blocks_for(nf, 128)), dim3(128), 0, st, nf, nullptr, c.fverts.p, c.fdims.p, Hs, Gs,
                           s.bt.p, (int)(slot_b + ngf), (int)(slot_g + ngf), s.kb, c.ckey.p, c.cidx.p, c.cval.p, c.gkey.p, c.gidx.p,
                           c.gval.p);
        CIPC_CHECK_LAUNCH();
    }
}

void friction_energy_partials(Context& c, double* part, double* elem_self, double* elem_ground, hipStream_t st) {
    Scene& s = c.sc;
    hipLaunchKernelGGL(e_friction, dim3(kFB), dim3(kFB), 0, st, (int)c.n_fself, (int)c.n_fenv, c.fr_self.p, c.fr_lam.p,
                       c.fr_coord.p, c.fr_basis.p, c.fr_env.p, c.fr_env_lam.p, s.x.p, s.vprev.p, s.gn, s.fhat * s.dt * s.dt, s.fric,
                       part, elem_self, elem_ground);
    CIPC_CHECK_LAUNCH();
}

void size_contact_records(Context& c) {
    const int64_t need = c.n_self + c.n_ee;
    if (need > INT32_MAX / 16) throw Error(CIPC_ERR_OVERFLOW, "too many contacts for the contact records");
    if (c.capS == 0) c.capS = 1024;
    if (need > c.capS) c.capS = (int)(need + need / 4 + 256);
}

// Record slots: ground [0, capG), capG = Vs (at most one entry per surface vertex); the
// contacts' 10 block and 4 gradient slots each for capS contacts; then the lagged friction
// records.  Slots past a device count are written empty, so the counts stay on the device;
// records keep the order ground, contacts, friction, and the sort is stable, so every block
// and gradient is summed in the same order as with packed records.
// Record slots and buffers of the next contact assembly (host side only).  With the counts
// still on the device, the capacity of the last assembly (grown by the host when the device
// reports it exceeded).
void contact_records_prepare(Context& c) {
    Scene& s = c.sc;
    if (!c.dc_pending) size_contact_records(c);
    else if (c.capS == 0) c.capS = 1024;
    const int capS = c.capS, capG = s.Vs;
    const size_t nbc = (size_t)capG + 10 * (size_t)capS, ngc = (size_t)capG + 4 * (size_t)capS;
    const size_t nb = nbc + (size_t)c.n_fenv + 10 * (size_t)c.n_fself, ng = ngc + (size_t)c.n_fenv + 4 * (size_t)c.n_fself;
    c.ckey.resize(std::max<size_t>(nb, 1)); c.cidx.resize(std::max<size_t>(nb, 1));
    c.cval.resize(9 * std::max<size_t>(nb, 1));
    c.gkey.resize(std::max<size_t>(ng, 1)); c.gidx.resize(std::max<size_t>(ng, 1));
    c.gval.resize(3 * std::max<size_t>(ng, 1));
    if (c.n_fself > 0) {
        c.fhess.resize((size_t)(kHU + 12) * c.n_fself);
        c.fverts.resize(c.n_fself);
        c.fdims.resize(c.n_fself);
    }
    c.chess.resize((size_t)(kHU + 12) * capS);
    c.cverts.resize(capS);
    c.cdims.resize(capS);
    c.ckey.n = nb;
    c.gkey.n = ng;
    c.records_capS = capS;
}

// Lagged friction (IPCSolver.cpp:309-317): its records go after the barrier records and depend
// on the lagged state and the positions only, so they can be built on `st` from the start of
// an iteration (slots of contact_records_prepare).
void friction_records_launch(Context& c, hipStream_t st) {
    const size_t nbc = (size_t)c.sc.Vs + 10 * (size_t)c.records_capS, ngc = (size_t)c.sc.Vs + 4 * (size_t)c.records_capS;
    friction_records(c, nbc, ngc, st);
}

void contact_assemble(Context& c, double kappa, bool friction_launched) {
    Scene& s = c.sc;
    if (!friction_launched) contact_records_prepare(c);
    const int capS = c.records_capS, capG = s.Vs;
    const int* dc = c.dcnt.p;
    const size_t nbc = (size_t)capG + 10 * (size_t)capS, ngc = (size_t)capG + 4 * (size_t)capS;
    if (capG == 0) hipLaunchKernelGGL(contact_totals, dim3(1), dim3(64), 0, c.stream, capS, c.dcnt.p);
    CIPC_HIP(hipEventRecord(c.fr_fork, c.stream));
    if (capG > 0) {
        hipLaunchKernelGGL(ground_kernel, dim3(blocks_for(capG, 256)), dim3(256), 0, c.stream, capG, c.dcnt.p, capS, c.env_set.p, s.x.p, s.bt.p,
                           s.gn, s.gd, s.hhat, scene_kappa(s, kappa), s.kb, c.ckey.p, c.cidx.p, c.cval.p, c.gkey.p, c.gidx.p, c.gval.p);
        CIPC_CHECK_LAUNCH();
    }
    double* Hs = c.chess.p;
    double* Gs = c.chess.p + (size_t)kHU * capS;
    hipLaunchKernelGGL(contact_self_kernel, dim3(blocks_for(capS, 64)), dim3(64), 0, c.stream, dc, capS, c.self_set.p, s.x.p, s.hhat,
                       scene_kappa(s, kappa), Hs, Gs, c.cverts.p, c.cdims.p);
    hipLaunchKernelGGL(contact_ee_kernel, dim3(blocks_for(capS, 64)), dim3(64), 0, c.stream, dc, capS, c.ee_set.p, c.ee_ie.p, s.se.p,
                       s.x.p, s.vrest.p, s.hhat, scene_kappa(s, kappa), Hs, Gs, c.cverts.p, c.cdims.p);
    CIPC_CHECK_LAUNCH();
    // Main stream: the projections of 4-vertex contacts.  Second stream (enqueued after them,
    // so that the main stream never waits on the host): the friction records, then the
    // projections of 2- and 3-vertex contacts.  Each projection kernel touches only the
    // contacts of its size.
    CIPC_HIP(hipEventRecord(c.cl_done, c.stream));
    hipLaunchKernelGGL(contact_pd4_kernel, dim3(blocks_for(capS, 64), 2), dim3(64), 0, c.stream, dc, capS, c.cdims.p, Hs, Gs);
    CIPC_CHECK_LAUNCH();
    CIPC_HIP(hipStreamWaitEvent(c.aux, c.fr_fork, 0));
    if (!friction_launched) friction_records(c, nbc, ngc, c.aux);
    CIPC_HIP(hipStreamWaitEvent(c.aux, c.cl_done, 0));
    hipLaunchKernelGGL((contact_pd_kernel<3, true>), dim3(blocks_for(capS, 64)), dim3(64), 0, c.aux, dc, 0, capS, c.cdims.p, Hs, Gs);
    hipLaunchKernelGGL((contact_pd_kernel<2, false>), dim3(blocks_for(capS, 64)), dim3(64), 0, c.aux, dc, 0, capS, c.cdims.p, Hs, Gs);
    CIPC_CHECK_LAUNCH();
    CIPC_HIP(hipEventRecord(c.fr_join, c.aux));
    CIPC_HIP(hipStreamWaitEvent(c.stream, c.fr_join, 0));
    hipLaunchKernelGGL(contact_scatter_kernel, dim3(blocks_for(capS, 128)), dim3(128), 0, c.stream, capS, dc, c.cverts.p, c.cdims.p, Hs,
                       Gs, s.bt.p, capG, capG, s.kb, c.ckey.p, c.cidx.p, c.cval.p, c.gkey.p, c.gidx.p, c.gval.p);
    CIPC_CHECK_LAUNCH();
    // the counts again, now with the contact-capacity check, for the host's next synchronisation
    if (c.dc_pending) CIPC_HIP(hipMemcpyAsync(c.pin_dc.p, c.dcnt.p, kDcWords * sizeof(int), hipMemcpyDeviceToHost, c.stream));
}

void upload_counts(Context& c) {
    int h[kDcWords] = {};
    h[kDcSelf] = (int)c.n_self; h[kDcMol] = (int)c.n_ee; h[kDcCCD] = (int)c.n_ccd; h[kDcEnv] = (int)c.n_env;
    c.dcnt.resize(kDcWords);
    CIPC_HIP(hipMemcpyAsync(c.dcnt.p, h, sizeof h, hipMemcpyHostToDevice, c.stream));
    CIPC_HIP(hipStreamSynchronize(c.stream));
    c.dc_pending = false;
}

// ---------------------------------------------------------------------------------
// Element-level evaluation entry points (tests)
// ---------------------------------------------------------------------------------
__global__ void distance_eval_kernel(int n, int kind, const double* __restrict__ pts, double* __restrict__ d,
                                     double* __restrict__ g, double* __restrict__ Ho) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double q[12], gg[12], H[144];
    for (int i = 0; i < 12; ++i) q[i] = pts[12 * t + i];
    d3 p[4];
    for (int k = 0; k < 4; ++k) p[k] = d3{q[3 * k], q[3 * k + 1], q[3 * k + 2]};
    int N;
    if (kind == 4) { ee_cross_gh(q, gg, H); N = 12; d[t] = ee_cross_sqn(p[0], p[1], p[2], p[3]); }
    else { dist_gh(kind, q, gg, H); N = 3 * nv_of(kind); d[t] = self_value(kind, p); }
    for (int i = 0; i < 12; ++i) g[12 * t + i] = i < N ? gg[i] : 0.0;
    for (int i = 0; i < 144; ++i) Ho[144 * t + i] = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) Ho[144 * t + i * N + j] = H[i * N + j];
}
template <int N>
__global__ __launch_bounds__(64) void make_pd_eval_kernel(int n, double* __restrict__ H) {
    constexpr int NU = N * (N + 1) / 2;
    __shared__ double lds[(NU + N * N) * 64];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double* h = H + (size_t)N * N * t;
    double* m = lds + threadIdx.x;
    for (int i = 0; i < N; ++i)
        for (int j = i; j < N; ++j) m[64 * Sym<N>::ix(i, j)] = h[i * N + j];
    make_pd_lds<N>(m);
    for (int i = 0; i < N; ++i)
        for (int j = i; j < N; ++j) { h[i * N + j] = m[64 * Sym<N>::ix(i, j)]; h[j * N + i] = m[64 * Sym<N>::ix(i, j)]; }
}
template <int M>
__global__ __launch_bounds__(64) void contact_pd_eval_kernel(int n, double* __restrict__ H, const double* __restrict__ G) {
    constexpr int N = 3 * M;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double* h = H + (size_t)N * N * t;
    auto acc = [&](int k) -> double& { return h[Sym<N>::row(k) * N + Sym<N>::col(k)]; };
    bool changed;
    if constexpr (M >= 3) {
        if (G != nullptr) {
            double g[N];
            for (int k = 0; k < N; ++k) g[k] = G[(size_t)N * t + k];
            changed = project_contact_structured<M>(acc, g);
        } else {
            changed = project_contact<M>(acc);
        }
    } else {
        changed = project_contact<M>(acc);
    }
    if (!changed) return;
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j) h[j * N + i] = h[i * N + j];
}
__global__ void accd_eval_kernel(int n, int kind, const double* __restrict__ pts, const double* __restrict__ dirs, double eta,
                                 double th, double* __restrict__ toc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    d3 p[4], dd[4];
    for (int k = 0; k < 4; ++k) { p[k] = ld3(pts + 12 * t, k); dd[k] = ld3(dirs + 12 * t, k); }
    toc[t] = kind == 0 ? accd_pt(p[0], p[1], p[2], p[3], dd[0], dd[1], dd[2], dd[3], eta, th)
                       : accd_ee(p[0], p[1], p[2], p[3], dd[0], dd[1], dd[2], dd[3], eta, th);
}
__global__ void classify_eval_kernel(int n, int kind, const double* __restrict__ pts, int* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    d3 p[4];
    for (int k = 0; k < 4; ++k) p[k] = ld3(pts + 12 * t, k);
    out[t] = kind == 0 ? dtype_pt(p[0], p[1], p[2], p[3]) : dtype_ee(p[0], p[1], p[2], p[3]);
}

template <typename T>
struct DevArr {
    T* p = nullptr;
    explicit DevArr(size_t n) { CIPC_HIP(hipMalloc(&p, sizeof(T) * (n ? n : 1))); }
    ~DevArr() { if (p) (void)hipFree(p); }
    void up(const T* h, size_t n) { CIPC_HIP(hipMemcpy(p, h, sizeof(T) * n, hipMemcpyHostToDevice)); }
    void down(T* h, size_t n) { CIPC_HIP(hipMemcpy(h, p, sizeof(T) * n, hipMemcpyDeviceToHost)); }
};

}  // namespace cipc

using namespace cipc;

extern "C" int cipc_gpu_eval_distance_impl(int n, int kind, const double* pts, double* d, double* g, double* H) {
    DevArr<double> dp(12 * (size_t)n), dd(n), dg(12 * (size_t)n), dh(144 * (size_t)n);
    dp.up(pts, 12 * (size_t)n);
    hipLaunchKernelGGL(distance_eval_kernel, dim3(blocks_for(n, 64)), dim3(64), 0, 0, n, kind, dp.p, dd.p, dg.p, dh.p);
    CIPC_CHECK_LAUNCH();
    dd.down(d, n); dg.down(g, 12 * (size_t)n); dh.down(H, 144 * (size_t)n);
    return 0;
}
extern "C" int cipc_gpu_eval_make_pd_impl(int n, int dim, double* H) {
    DevArr<double> dh((size_t)dim * dim * n);
    dh.up(H, (size_t)dim * dim * n);
    if (dim == 6) hipLaunchKernelGGL(make_pd_eval_kernel<6>, dim3(blocks_for(n, 64)), dim3(64), 0, 0, n, dh.p);
    else if (dim == 9) hipLaunchKernelGGL(make_pd_eval_kernel<9>, dim3(blocks_for(n, 64)), dim3(64), 0, 0, n, dh.p);
    else hipLaunchKernelGGL(make_pd_eval_kernel<12>, dim3(blocks_for(n, 64)), dim3(64), 0, 0, n, dh.p);
    CIPC_CHECK_LAUNCH();
    dh.down(H, (size_t)dim * dim * n);
    return 0;
}
extern "C" int cipc_gpu_eval_contact_pd_impl(int n, int nverts, double* H, const double* G) {
    const size_t dim = 3 * (size_t)nverts;
    DevArr<double> dh(dim * dim * n), dg(G ? dim * n : 1);
    dh.up(H, dim * dim * n);
    if (G) dg.up(G, dim * n);
    const double* gp = G ? dg.p : nullptr;
    if (nverts == 2) hipLaunchKernelGGL(contact_pd_eval_kernel<2>, dim3(blocks_for(n, 64)), dim3(64), 0, 0, n, dh.p, gp);
    else if (nverts == 3) hipLaunchKernelGGL(contact_pd_eval_kernel<3>, dim3(blocks_for(n, 64)), dim3(64), 0, 0, n, dh.p, gp);
    else hipLaunchKernelGGL(contact_pd_eval_kernel<4>, dim3(blocks_for(n, 64)), dim3(64), 0, 0, n, dh.p, gp);
    CIPC_CHECK_LAUNCH();
    dh.down(H, dim * dim * n);
    return 0;
}
extern "C" int cipc_gpu_eval_accd_impl(int n, int kind, const double* pts, const double* dirs, double eta, double th, double* toc) {
    DevArr<double> dp(12 * (size_t)n), dd(12 * (size_t)n), dt(n);
    dp.up(pts, 12 * (size_t)n); dd.up(dirs, 12 * (size_t)n);
    hipLaunchKernelGGL(accd_eval_kernel, dim3(blocks_for(n, 64)), dim3(64), 0, 0, n, kind, dp.p, dd.p, eta, th, dt.p);
    CIPC_CHECK_LAUNCH();
    dt.down(toc, n);
    return 0;
}
extern "C" int cipc_gpu_eval_classify_impl(int n, int kind, const double* pts, int* type) {
    DevArr<double> dp(12 * (size_t)n);
    DevArr<int> dt(n);
    dp.up(pts, 12 * (size_t)n);
    hipLaunchKernelGGL(classify_eval_kernel, dim3(blocks_for(n, 64)), dim3(64), 0, 0, n, kind, dp.p, dt.p);
    CIPC_CHECK_LAUNCH();
    dt.down(type, n);
    return 0;
}
