// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate expressions of CPU-IPC
// (MPL-2.0) in the reference's operation order; each such part cites the source it follows.
// Data-parallel parts of the barrier-stiffness control.  Compiled with -ffp-contract=off.
//
// Reference: initKappa (CPU IPC/IPCSolver.cpp:766-808) with computeEGradient
// (IPCSolver.cpp:74-106), and postLineSearch's close-constraint bookkeeping
// (IPCSolver.cpp:811-866).  The scalar decisions around them stay on the host.
//
// initKappa's two sums need per-vertex vectors: g_E (inertia without drag + tet elastic
// forces times dt^2) and g_c (ground and self barrier gradients with kappa = 1).  Tet
// forces go to the tet slots of the element staging and are gathered per vertex (tets come
// first in every vertex's slot run); contact gradients become vertex-keyed records that are
// sorted and summed in record order, as in the assembly.  Both sums are then reduced in a
// fixed order.
#include "contact.hpp"
#include "context.hpp"
#include "snh.hpp"

namespace cipc {

namespace {

__device__ inline int decode_c(int4 c, int v[4]) {
    if (c.x >= 0) { v[0] = c.x; v[1] = c.y; v[2] = c.z; v[3] = c.w; return 3; }
    v[0] = -c.x - 1; v[1] = c.y; v[2] = c.z; v[3] = c.w;
    if (c.z < 0) return 0;
    if (c.w < 0) return 1;
    return 2;
}
__device__ inline double constraint_value(int4 c, const double* x) {
    int v[4];
    const int type = decode_c(c, v);
    const d3 p = ld3(x, v[0]);
    switch (type) {
    case 0: return d_PP(p, ld3(x, v[1]));
    case 1: return d_PE(p, ld3(x, v[1]), ld3(x, v[2]));
    case 2: return d_PT(p, ld3(x, v[1]), ld3(x, v[2]), ld3(x, v[3]));
    default: return d_EE(p, ld3(x, v[1]), ld3(x, v[2]), ld3(x, v[3]));
    }
}

}  // namespace

// Tet corner forces vol (pa F + pb cof F) B_a into the tet slots (no Hessian, no SVD).
__global__ __launch_bounds__(256) void tet_force_kernel(int T, const int4* __restrict__ tets, const double* __restrict__ x,
                                                        const double* __restrict__ dminv, const double* __restrict__ vol, Lame lame,
                                                        const int* __restrict__ slot_g, double* __restrict__ stagG) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    double u, r;
    lame.at(t, u, r);
    const int4 id = tets[t];
    const d3 p0 = ld3(x, id.x);
    const d3 e1 = ld3(x, id.y) - p0, e2 = ld3(x, id.z) - p0, e3 = ld3(x, id.w) - p0;
    double D[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) D[k][j] = dminv[(size_t)(k + 3 * j) * T + t];
    d3 f[3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
        f[j] = d3{e1.x * D[0][j] + e2.x * D[1][j] + e3.x * D[2][j], e1.y * D[0][j] + e2.y * D[1][j] + e3.y * D[2][j],
                  e1.z * D[0][j] + e2.z * D[1][j] + e3.z * D[2][j]};
    const d3 c0 = cross(f[1], f[2]), c1 = cross(f[2], f[0]), c2 = cross(f[0], f[1]);
    const double I2 = sqn(f[0]) + sqn(f[1]) + sqn(f[2]);
    const double I3 = dot(f[0], c0);
    const double pa = u * (1 - 1 / (I2 + 1)), pb = r * (I3 - 1 - u * 3 / (r * 4));
    const d3 P[3] = {pa * f[0] + pb * c0, pa * f[1] + pb * c1, pa * f[2] + pb * c2};   // columns of PK1
    const double B[4][3] = {{-(D[0][0] + D[1][0] + D[2][0]), -(D[0][1] + D[1][1] + D[2][1]), -(D[0][2] + D[1][2] + D[2][2])},
                            {D[0][0], D[0][1], D[0][2]}, {D[1][0], D[1][1], D[1][2]}, {D[2][0], D[2][1], D[2][2]}};
    const double vt = vol[t];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const d3 fa = vt * (B[a][0] * P[0] + B[a][1] * P[1] + B[a][2] * P[2]);
        st3(stagG, slot_g[4 * t + a], fa);
    }
}

// g_E(v) = m (x - x~) + dt^2 (sum of v's tet-corner forces, in slot order).
__global__ void egrad_gather_kernel(int V, const double* __restrict__ x, const double* __restrict__ xt, const double* __restrict__ mass,
                                    double dt2, const int* __restrict__ vptr, const int* __restrict__ vtet,
                                    const double* __restrict__ stagG, double* __restrict__ gE) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    d3 g = mass[v] * (ld3(x, v) - ld3(xt, v));
    const int q0 = vptr[v];
    for (int q = q0; q < q0 + vtet[v]; ++q) g = g + dt2 * ld3(stagG, q);
    st3(gE, v, g);
}

// Barrier gradient records with kappa = 1: ground constraint i at slot i, self constraint i
// at slots ne + 4 i + k (unused slots keyed past every vertex).
__global__ void kappa_contact_records(int ne, int ns, const int* __restrict__ env, const int4* __restrict__ self_set,
                                      const double* __restrict__ x, d3 gn, double gd, double hhat, int* __restrict__ gkey,
                                      int* __restrict__ gidx, double* __restrict__ gval) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ne) {
        const int v = env[i];
        const double dist = dot(gn, ld3(x, v)) - gd;
        const double gb = barrier_g(dist * dist, hhat);
        st3(gval, i, ((gb * 2.0) * dist) * gn);
        gkey[i] = v;
        gidx[i] = i;
        return;
    }
    const int c = i - ne;
    if (c >= ns) return;
    const int4 e = self_set[c];
    int v[4];
    const int type = decode_c(e, v);
    const int nv = type == 0 ? 2 : (type == 1 ? 3 : 4);
    double q[12], g[12], H[144];
    d3 p[4];
    for (int k = 0; k < nv; ++k) { p[k] = ld3(x, v[k]); q[3 * k] = p[k].x; q[3 * k + 1] = p[k].y; q[3 * k + 2] = p[k].z; }
    double d;
    if (type == 0) {
        d = d_PP(p[0], p[1]);
        for (int k = 0; k < 3; ++k) { g[k] = 2.0 * (q[k] - q[3 + k]); g[3 + k] = -g[k]; }
    } else {
        d = type == 1 ? d_PE(p[0], p[1], p[2]) : (type == 2 ? d_PT(p[0], p[1], p[2], p[3]) : d_EE(p[0], p[1], p[2], p[3]));
        if (type == 1) dist_pe_gh(q, g, H);
        else if (type == 2) dist_pt_gh(q, g, H);
        else dist_ee_gh(q, g, H);
    }
    const double gb = barrier_g(d, hhat);
    const double sc = (type <= 1) ? -e.w * gb : gb;
    for (int k = 0; k < 4; ++k) {
        const int slot = ne + 4 * c + k;
        gidx[slot] = slot;
        if (k < nv) {
            gkey[slot] = v[k];
            st3(gval, slot, d3{sc * g[3 * k], sc * g[3 * k + 1], sc * g[3 * k + 2]});
        } else {
            gkey[slot] = 0x7fffffff;
        }
    }
}

__global__ __launch_bounds__(kEnergyBlock) void kappa_products(int V, const double* __restrict__ gc, const double* __restrict__ gE,
                                                               double* __restrict__ part) {
    double a = 0.0, b = 0.0;
    for (int v = blockIdx.x * kEnergyBlock + threadIdx.x; v < V; v += kEnergyGrid * kEnergyBlock) {
        const d3 c = ld3(gc, v);
        a += dot(c, ld3(gE, v));
        b += sqn(c);
    }
    energy_block_sum(a, part);
    __syncthreads();
    energy_block_sum(b, part + kEnergyGrid);
}
__global__ __launch_bounds__(kEnergyBlock) void sum_two(const double* __restrict__ part, double* __restrict__ out) {
    for (int t = 0; t < 2; ++t) {
        energy_block_sum(part[t * kEnergyGrid + threadIdx.x], out + t);
        __syncthreads();
    }
}

__global__ void zero_vec(double* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

int64_t kappa_gradients(Context& c, double out[2]) {
    Scene& s = c.sc;
    const int ne = (int)c.n_env, ns = (int)c.n_self;
    out[0] = out[1] = 0.0;
    if (ne + ns == 0) return 0;
    const int V = s.V;
    c.kgE.resize(3 * (size_t)V);
    c.kgc.resize(3 * (size_t)V);
    if (s.T > 0) {
        hipLaunchKernelGGL(tet_force_kernel, dim3(blocks_for(s.T, 256)), dim3(256), 0, c.stream, s.T, s.tets.p, s.x.p, s.dminv.p,
                           s.vol.p, scene_lame(s), c.sp.slot_g.p, c.stagG.p);
        CIPC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(egrad_gather_kernel, dim3(blocks_for(V, 256)), dim3(256), 0, c.stream, V, s.x.p, s.xt.p, s.mass.p,
                       s.dt * s.dt, c.sp.vptr.p, c.sp.vtet.p, c.stagG.p, c.kgE.p);
    hipLaunchKernelGGL(zero_vec, dim3(blocks_for(3 * (size_t)V, 256)), dim3(256), 0, c.stream, c.kgc.p, 3 * (size_t)V);
    CIPC_CHECK_LAUNCH();
    const int ng = ne + 4 * ns;
    c.gkey.resize(ng); c.gidx.resize(ng); c.gval.resize(3 * (size_t)ng);
    c.gkey_sorted.resize(ng); c.gidx_sorted.resize(ng);
    hipLaunchKernelGGL(kappa_contact_records, dim3(blocks_for(ne + ns, 64)), dim3(64), 0, c.stream, ne, ns, c.env_set.p, c.self_set.p,
                       s.x.p, s.gn, s.gd, s.hhat, c.gkey.p, c.gidx.p, c.gval.p);
    CIPC_CHECK_LAUNCH();
    sort_i32_pairs(c, c.gkey.p, c.gkey_sorted.p, c.gidx.p, c.gidx_sorted.p, ng, s.kb);
    add_gradient_records_into(c, ng, c.gkey_sorted.p, c.gidx_sorted.p, c.gval.p, c.kgc.p);
    DBuf<double>& part = c.dscratch;
    part.resize(2 * (size_t)kEnergyGrid + 2);
    hipLaunchKernelGGL(kappa_products, dim3(kEnergyGrid), dim3(kEnergyBlock), 0, c.stream, V, c.kgc.p, c.kgE.p, part.p);
    hipLaunchKernelGGL(sum_two, dim3(1), dim3(kEnergyBlock), 0, c.stream, part.p, part.p + 2 * kEnergyGrid);
    CIPC_CHECK_LAUNCH();
    CIPC_HIP(hipMemcpyAsync(out, part.p + 2 * kEnergyGrid, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CIPC_HIP(hipStreamSynchronize(c.stream));
    return ne + ns;
}

// ---------------------------------------------------------------------------------
// The kappa loop's exit test of solveIPCStep (IPCSolver.cpp:1171-1188): the smallest and
// largest constraint value over the ground constraints and then the self constraints.  A
// minimum and a maximum do not depend on the order of their operands, so a grid-strided
// pass with one partial per work-group gives the oracle's values exactly.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(kEnergyBlock) void extrema_partials(int ne, int ns, const int* __restrict__ env,
                                                                 const int4* __restrict__ self_set, const double* __restrict__ x, d3 gn,
                                                                 double gd, double* __restrict__ part) {
    double lo = INFINITY, hi = -INFINITY;
    for (int i = blockIdx.x * kEnergyBlock + threadIdx.x; i < ne + ns; i += kEnergyGrid * kEnergyBlock) {
        double d;
        if (i < ne) {
            const double dist = dot(gn, ld3(x, env[i])) - gd;
            d = dist * dist;
        } else {
            d = constraint_value(self_set[i - ne], x);
        }
        lo = fmin(lo, d);
        hi = fmax(hi, d);
    }
    __shared__ double sl[kEnergyBlock], sh[kEnergyBlock];
    sl[threadIdx.x] = lo;
    sh[threadIdx.x] = hi;
    __syncthreads();
    for (int s = kEnergyBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sl[threadIdx.x] = fmin(sl[threadIdx.x], sl[threadIdx.x + s]);
            sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[blockIdx.x] = sl[0]; part[kEnergyGrid + blockIdx.x] = sh[0]; }
}
__global__ __launch_bounds__(kEnergyBlock) void extrema_final(const double* __restrict__ part, double* __restrict__ out) {
    static_assert(kEnergyGrid == kEnergyBlock, "one partial per thread");
    __shared__ double sl[kEnergyBlock], sh[kEnergyBlock];
    sl[threadIdx.x] = part[threadIdx.x];
    sh[threadIdx.x] = part[kEnergyGrid + threadIdx.x];
    __syncthreads();
    for (int s = kEnergyBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sl[threadIdx.x] = fmin(sl[threadIdx.x], sl[threadIdx.x + s]);
            sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = sl[0]; out[1] = sh[0]; }
}

int64_t constraint_extrema(Context& c, double out[2]) {
    Scene& s = c.sc;
    const int ne = (int)c.n_env, ns = (int)c.n_self;
    out[0] = out[1] = 0.0;
    if (ne + ns == 0) return 0;
    DBuf<double>& part = c.dscratch2;
    part.resize(2 * (size_t)kEnergyGrid + 2);
    hipLaunchKernelGGL(extrema_partials, dim3(kEnergyGrid), dim3(kEnergyBlock), 0, c.stream, ne, ns, c.env_set.p, c.self_set.p, s.x.p,
                       s.gn, s.gd, part.p);
    hipLaunchKernelGGL(extrema_final, dim3(1), dim3(kEnergyBlock), 0, c.stream, part.p, part.p + 2 * kEnergyGrid);
    CIPC_CHECK_LAUNCH();
    read_back(c, out, part.p + 2 * kEnergyGrid, 2);
    return ne + ns;
}

// ---------------------------------------------------------------------------------
// Close constraints (postLineSearch)
// ---------------------------------------------------------------------------------
__global__ void close_check(int nge, const int* __restrict__ cenv, const double* __restrict__ cenv_val, int ngs,
                            const int4* __restrict__ cself, const double* __restrict__ cself_val, const double* __restrict__ x, d3 gn,
                            double gd, int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nge) {
        const double dist = dot(gn, ld3(x, cenv[i])) - gd;
        if (dist * dist <= cenv_val[i]) atomicOr(flag, 1);
        return;
    }
    const int j = i - nge;
    if (j < ngs && constraint_value(cself[j], x) <= cself_val[j]) atomicOr(flag, 1);
}
// flags[i] = value < dtol for ground constraints [0, ne) and self constraints [ne, ne + ns)
// flags[ne + ns] = 0 (the scan's tail)
__global__ void close_flags(int ne, int ns, const int* __restrict__ env, const int4* __restrict__ self_set, const double* __restrict__ x,
                            d3 gn, double gd, double dtol, int* __restrict__ flags, double* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == ne + ns) flags[i] = 0;
    if (i >= ne + ns) return;
    double d;
    if (i < ne) {
        const double dist = dot(gn, ld3(x, env[i])) - gd;
        d = dist * dist;
    } else {
        d = constraint_value(self_set[i - ne], x);
    }
    vals[i] = d;
    flags[i] = d < dtol ? 1 : 0;
}
__global__ void close_scatter(int ne, int ns, const int* __restrict__ env, const int4* __restrict__ self_set, const int* __restrict__ flags,
                              const int* __restrict__ pos, const double* __restrict__ vals, int* __restrict__ cenv,
                              double* __restrict__ cenv_val, int4* __restrict__ cself, double* __restrict__ cself_val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne + ns || !flags[i]) return;
    const int o = pos[i];
    if (i < ne) { cenv[o] = env[i]; cenv_val[o] = vals[i]; }
    else { const int k = o - pos[ne]; cself[k] = self_set[i - ne]; cself_val[k] = vals[i]; }
}

int close_constraints(Context& c, double dtol, int64_t counts[2]) {
    Scene& s = c.sc;
    const int nge = (int)c.n_cenv, ngs = (int)c.n_cself;
    c.counters.resize(4);
    int* flag = c.counters.p;
    CIPC_HIP(hipMemsetAsync(flag, 0, sizeof(int), c.stream));
    if (nge + ngs > 0) {
        hipLaunchKernelGGL(close_check, dim3(blocks_for(nge + ngs, 256)), dim3(256), 0, c.stream, nge, c.cenv.p, c.cenv_val.p, ngs,
                           c.cself.p, c.cself_val.p, s.x.p, s.gn, s.gd, flag);
        CIPC_CHECK_LAUNCH();
    }
    const int ne = (int)c.n_env, ns = (int)c.n_self, n = ne + ns;
    int closer = 0, tot[2] = {0, 0};
    if (n > 0) {
        c.iscratch.resize((size_t)n + 1);
        c.iscratch2.resize((size_t)n + 1);
        c.kvals.resize(n);
        int* flags = c.iscratch.p;
        int* pos = c.iscratch2.p;
        hipLaunchKernelGGL(close_flags, dim3(blocks_for(n + 1, 256)), dim3(256), 0, c.stream, ne, ns, c.env_set.p, c.self_set.p, s.x.p,
                           s.gn, s.gd, dtol, flags, c.kvals.p);
        CIPC_CHECK_LAUNCH();
        exclusive_scan_int(c, flags, pos, n + 1);
        int h[3];
        CIPC_HIP(hipMemcpyAsync(&h[0], pos + ne, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        CIPC_HIP(hipMemcpyAsync(&h[1], pos + n, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        CIPC_HIP(hipMemcpyAsync(&h[2], flag, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        CIPC_HIP(hipStreamSynchronize(c.stream));
        tot[0] = h[0];
        tot[1] = h[1] - h[0];
        closer = h[2];
        c.cenv.resize(std::max(tot[0], 1)); c.cenv_val.resize(std::max(tot[0], 1));
        c.cself.resize(std::max(tot[1], 1)); c.cself_val.resize(std::max(tot[1], 1));
        hipLaunchKernelGGL(close_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, ne, ns, c.env_set.p, c.self_set.p, flags, pos,
                           c.kvals.p, c.cenv.p, c.cenv_val.p, c.cself.p, c.cself_val.p);
        CIPC_CHECK_LAUNCH();
        CIPC_HIP(hipStreamSynchronize(c.stream));
    } else {
        CIPC_HIP(hipMemcpyAsync(&closer, flag, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        CIPC_HIP(hipStreamSynchronize(c.stream));
    }
    c.n_cenv = tot[0];
    c.n_cself = tot[1];
    counts[0] = tot[0];
    counts[1] = tot[1];
    return closer;
}

}  // namespace cipc
