// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate expressions of CPU-IPC
// (MPL-2.0) in the reference's operation order; each such part cites the source it follows.
// Cloth stage of the Newton assembly and energy: Baraff-Witkin triangles, the quadratic
// bending form and hinge bending.
//
// Reference: computeGradientAndHessian, triangle and bending loops (CPU IPC/IPCSolver.cpp:
// 166-267), computeEnergy (IPCSolver.cpp:330-368), hingeBendingGradientAndHessian /
// hingeBendingEnergy (HingeBending.cpp:79-106).
//
// The element kernels write into the same slot-ordered staging as the tet kernel: every
// (element, local block) and (element, corner) has a slot fixed with the static pattern,
// and the block / vertex gathers sum the slots in order.  Element numbering in the slot
// tables: tets [0, T), triangles [T, T+Tr), bending elements [T+Tr, T+Tr+Nb).  Corner forces
// are staged without the dt^2 factor, which the vertex gather applies.
#include "cloth.hpp"
#include "contact.hpp"
#include "context.hpp"

namespace cipc {

// One thread per triangle.  F = [x1 - x0, x2 - x0] Dm^-1; dF/dx_c = B_c with
// B_0 = -(B_1 + B_2), B_1 = row 0 of Dm^-1, B_2 = row 1 (computePFPX32D_double,
// Elasticity.cpp:851-897).  K_ab[i][k] = area dt^2 sum_{j,l} B_a[j] B_b[l] H[3j+i][3l+k].
template <typename HT>
__global__ __launch_bounds__(256) void bw_element_kernel(int Tr, int e0, const int4* __restrict__ tris, const double* __restrict__ x,
                                                         const double* __restrict__ dmt, const double* __restrict__ area,
                                                         double st, double sh, double k, double dt2,
                                                         const int* __restrict__ slot_h, const int* __restrict__ slot_g,
                                                         HT* __restrict__ stagH, double* __restrict__ stagG) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Tr) return;
    const int4 id = tris[t];
    const int vid[3] = {id.x, id.y, id.z};
    const d3 p0 = ld3(x, id.x);
    const d3 e1 = ld3(x, id.y) - p0, e2 = ld3(x, id.z) - p0;
    const double D00 = dmt[t], D10 = dmt[(size_t)Tr + t], D01 = dmt[2 * (size_t)Tr + t], D11 = dmt[3 * (size_t)Tr + t];
    const d3 fu = D00 * e1 + D10 * e2, fv = D01 * e1 + D11 * e2;
    const double B[3][2] = {{-(D00 + D10), -(D01 + D11)}, {D00, D01}, {D10, D11}};
    double P[6];
    bw_pk1(fu, fv, st, sh, k, P);
    const double A = area[t];
    const int e = e0 + t;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const d3 f = d3{A * (B[a][0] * P[0] + B[a][1] * P[3]), A * (B[a][0] * P[1] + B[a][1] * P[4]),
                        A * (B[a][0] * P[2] + B[a][1] * P[5])};
        st3(stagG, slot_g[4 * e + a], f);
    }
    double H[36];
    bw_hessian(fu, fv, st, sh, k, H);
    const double sc = A * dt2;
#pragma unroll
    for (int lb = 0; lb < 10; ++lb) {
        const int a = lb_a(lb), b = lb_b(lb);
        if (a > 2) continue;
        const int j = slot_h[10 * e + lb];
        if (j < 0) continue;
        double K[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                double acc = 0.0;
#pragma unroll
                for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                    for (int ll = 0; ll < 2; ++ll) acc += (B[a][jj] * B[b][ll]) * H[6 * (3 * jj + i) + 3 * ll + kk];
                K[i][kk] = sc * acc;
            }
        // stored block (row vertex max, column vertex min): K_ab, or K_ab^T when vid[a] < vid[b]
        const bool tr = vid[a] < vid[b];
        HT* o = stagH + (size_t)9 * j;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) o[3 * i + kk] = (HT)(tr ? K[kk][i] : K[i][kk]);
    }
}

// One thread per bending element of the quadratic form: forces D (Q (x) I3) x, blocks
// D dt^2 Q_ab I3 (IPCSolver.cpp:197-237).  Q is read as 16 coalesced SoA streams.
template <typename HT>
__global__ __launch_bounds__(256) void bend_quad_kernel(int Nb, int e0, const int4* __restrict__ bv, const double* __restrict__ x,
                                                        const double* __restrict__ bq, double plate, double dt2,
                                                        const int* __restrict__ slot_h, const int* __restrict__ slot_g,
                                                        HT* __restrict__ stagH, double* __restrict__ stagG) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= Nb) return;
    const int4 id = bv[h];
    const int vid[4] = {id.x, id.y, id.z, id.w};
    d3 p[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) p[a] = ld3(x, vid[a]);
    double Q[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) Q[i] = bq[(size_t)i * Nb + h];
    const int e = e0 + h;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        d3 f{0, 0, 0};
#pragma unroll
        for (int b = 0; b < 4; ++b) f = f + (plate * Q[4 * a + b]) * p[b];
        st3(stagG, slot_g[4 * e + a], f);
    }
    const double sc = plate * dt2;
#pragma unroll
    for (int lb = 0; lb < 10; ++lb) {
        const int j = slot_h[10 * e + lb];
        if (j < 0) continue;
        const double q = sc * Q[4 * lb_a(lb) + lb_b(lb)];
        HT* o = stagH + (size_t)9 * j;
#pragma unroll
        for (int i = 0; i < 9; ++i) o[i] = (HT)((i % 4 == 0) ? q : 0.0);
    }
}

// One thread per hinge: E = D w dtheta^2 / 2, dtheta = remainder(theta - rest, 2 pi);
// g = D w dtheta grad theta, H = makePD(D w (grad grad^T + dtheta Hess theta)), both times
// dt^2 (HingeBending.cpp:79-106; IPCSolver.cpp:238-267).  The angle is translation-invariant,
// so the projection runs on the translation-free subspace like the contact blocks.
template <typename HT>
__global__ __launch_bounds__(64) void bend_hinge_kernel(int Nb, int e0, const int4* __restrict__ bv, const double* __restrict__ x,
                                                        const double* __restrict__ rest, const double* __restrict__ weight,
                                                        double plate, double dt2, const int* __restrict__ slot_h,
                                                        const int* __restrict__ slot_g, HT* __restrict__ stagH,
                                                        double* __restrict__ stagG) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= Nb) return;
    const int4 id = bv[h];
    const int vid[4] = {id.x, id.y, id.z, id.w};
    d3 q[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) q[a] = ld3(x, vid[a]);
    double g[12];
    const double th = hinge_angle_grad(q, g);
    const double df = remainder(th - rest[h], 6.28318530717958647692);
    const double s = plate * weight[h];
    const int e = e0 + h;
#pragma unroll
    for (int a = 0; a < 4; ++a)
        st3(stagG, slot_g[4 * e + a], d3{s * df * g[3 * a], s * df * g[3 * a + 1], s * df * g[3 * a + 2]});
    double Hs[78];
    {
        M3x B[10];
        hinge_angle_hess(q, B);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = a; b < 4; ++b)
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = (a == b ? i : 0); j < 3; ++j) {
                        const int r = 3 * a + i, c = 3 * b + j;
                        Hs[Sym<12>::ix(r, c)] = s * (g[r] * g[c] + df * B[ub_index(a, b)].m[i][j]);
                    }
    }
    project_contact<4>([&](int k) -> double& { return Hs[k]; });
#pragma unroll
    for (int lb = 0; lb < 10; ++lb) {
        const int j = slot_h[10 * e + lb];
        if (j < 0) continue;
        const int a = lb_a(lb), b = lb_b(lb);
        const bool tr = vid[a] < vid[b];
        HT* o = stagH + (size_t)9 * j;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk)
                o[3 * i + kk] = (HT)(dt2 * (tr ? Hs[Sym<12>::ix(3 * a + kk, 3 * b + i)] : Hs[Sym<12>::ix(3 * a + i, 3 * b + kk)]));
    }
}

template <typename HT>
static void cloth_assemble_t(Context& c, hipStream_t st, HT* stagH) {
    Scene& s = c.sc;
    const double dt2 = s.dt * s.dt;
    if (s.Tr > 0) {
        hipLaunchKernelGGL(bw_element_kernel<HT>, dim3(blocks_for(s.Tr, 256)), dim3(256), 0, st, s.Tr, s.T, s.tris.p, s.x.p, s.dmt.p,
                           s.area.p, s.stretch, s.shear, s.srate, dt2, c.sp.slot_h.p, c.sp.slot_g.p, stagH, c.stagG.p);
        CIPC_CHECK_LAUNCH();
    }
    if (s.Nb > 0) {
        if (s.bmode == 0)
            hipLaunchKernelGGL(bend_quad_kernel<HT>, dim3(blocks_for(s.Nb, 256)), dim3(256), 0, st, s.Nb, s.T + s.Tr, s.bverts.p, s.x.p,
                               s.bq.p, s.plate, dt2, c.sp.slot_h.p, c.sp.slot_g.p, stagH, c.stagG.p);
        else
            hipLaunchKernelGGL(bend_hinge_kernel<HT>, dim3(blocks_for(s.Nb, 64)), dim3(64), 0, st, s.Nb, s.T + s.Tr, s.bverts.p, s.x.p,
                               s.brest.p, s.bweight.p, s.plate, dt2, c.sp.slot_h.p, c.sp.slot_g.p, stagH, c.stagG.p);
        CIPC_CHECK_LAUNCH();
    }
}
void cloth_assemble(Context& c, hipStream_t st) {
    if (c.hessian_fp32) cloth_assemble_t<float>(c, st, reinterpret_cast<float*>(c.stagH.p));
    else cloth_assemble_t<double>(c, st, c.stagH.p);
}

// ---------------------------------------------------------------------------------
// Energy partials (grid-strided, fixed order; summed with the other terms)
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(kEnergyBlock) void e_bw(int Tr, const int4* __restrict__ tris, const double* __restrict__ x,
                                                     const double* __restrict__ dmt, const double* __restrict__ area, double st,
                                                     double sh, double k, double* __restrict__ part) {
    double acc = 0.0;
    for (int t = blockIdx.x * kEnergyBlock + threadIdx.x; t < Tr; t += gridDim.x * kEnergyBlock) {
        const int4 id = tris[t];
        const d3 p0 = ld3(x, id.x);
        const d3 e1 = ld3(x, id.y) - p0, e2 = ld3(x, id.z) - p0;
        const double D00 = dmt[t], D10 = dmt[(size_t)Tr + t], D01 = dmt[2 * (size_t)Tr + t], D11 = dmt[3 * (size_t)Tr + t];
        acc += bw_energy(D00 * e1 + D10 * e2, D01 * e1 + D11 * e2, st, sh, k) * area[t];
    }
    energy_block_sum(acc, part);
}
// quadratic form: sum x^T (Q (x) I3) x (scaled by D/2 on the host); hinge: sum D w dtheta^2 / 2
__global__ __launch_bounds__(kEnergyBlock) void e_bend(int Nb, int mode, const int4* __restrict__ bv, const double* __restrict__ x,
                                                       const double* __restrict__ bq, const double* __restrict__ rest,
                                                       const double* __restrict__ weight, double plate, double* __restrict__ part) {
    double acc = 0.0;
    for (int h = blockIdx.x * kEnergyBlock + threadIdx.x; h < Nb; h += gridDim.x * kEnergyBlock) {
        const int4 id = bv[h];
        const d3 q[4] = {ld3(x, id.x), ld3(x, id.y), ld3(x, id.z), ld3(x, id.w)};
        if (mode == 0) {
            double e = 0.0;
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) e += bq[(size_t)(4 * a + b) * Nb + h] * dot(q[a], q[b]);
            acc += e;
        } else {
            const double df = remainder(hinge_angle_grad(q, nullptr) - rest[h], 6.28318530717958647692);
            acc += 0.5 * plate * weight[h] * df * df;
        }
    }
    energy_block_sum(acc, part);
}

void cloth_energy_partials(Context& c, double* part_bw, double* part_bend, hipStream_t st) {
    Scene& s = c.sc;
    hipLaunchKernelGGL(e_bw, dim3(kEnergyGrid), dim3(kEnergyBlock), 0, st, s.Tr, s.tris.p, s.x.p, s.dmt.p, s.area.p, s.stretch,
                       s.shear, s.srate, part_bw);
    hipLaunchKernelGGL(e_bend, dim3(kEnergyGrid), dim3(kEnergyBlock), 0, st, s.Nb, s.bmode, s.bverts.p, s.x.p, s.bq.p,
                       s.brest.p, s.bweight.p, s.plate, part_bend);
    CIPC_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------------
// Element-level evaluation (tests)
// ---------------------------------------------------------------------------------
__global__ void bw_eval_kernel(int n, const double* __restrict__ F, double st, double sh, double k, double* __restrict__ P,
                               double* __restrict__ H, double* __restrict__ e) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const d3 fu = ld3(F, 2 * t), fv = ld3(F, 2 * t + 1);
    double p[6], h[36];
    bw_pk1(fu, fv, st, sh, k, p);
    bw_hessian(fu, fv, st, sh, k, h);
    for (int i = 0; i < 6; ++i) P[6 * t + i] = p[i];
    for (int i = 0; i < 36; ++i) H[36 * t + i] = h[i];
    e[t] = bw_energy(fu, fv, st, sh, k);
}
__global__ __launch_bounds__(64) void hinge_eval_kernel(int n, const double* __restrict__ pts, const double* __restrict__ rest,
                                                        const double* __restrict__ weight, double plate, double* __restrict__ G,
                                                        double* __restrict__ H, double* __restrict__ E) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    d3 q[4];
    for (int a = 0; a < 4; ++a) q[a] = ld3(pts, 4 * t + a);
    double g[12];
    const double th = hinge_angle_grad(q, g);
    const double df = remainder(th - rest[t], 6.28318530717958647692);
    const double s = plate * weight[t];
    for (int i = 0; i < 12; ++i) G[12 * t + i] = s * df * g[i];
    double Hs[78];
    {
        M3x B[10];
        hinge_angle_hess(q, B);
        for (int a = 0; a < 4; ++a)
            for (int b = a; b < 4; ++b)
                for (int i = 0; i < 3; ++i)
                    for (int j = (a == b ? i : 0); j < 3; ++j) {
                        const int r = 3 * a + i, c = 3 * b + j;
                        Hs[Sym<12>::ix(r, c)] = s * (g[r] * g[c] + df * B[ub_index(a, b)].m[i][j]);
                    }
    }
    project_contact<4>([&](int k) -> double& { return Hs[k]; });
    for (int r = 0; r < 12; ++r)
        for (int c = 0; c < 12; ++c) H[144 * t + 12 * r + c] = Hs[Sym<12>::ix(r, c)];
    E[t] = 0.5 * plate * weight[t] * df * df;
}

}  // namespace cipc

namespace {
template <typename T>
struct Dev {
    T* p = nullptr;
    explicit Dev(size_t n) { CIPC_HIP(hipMalloc(&p, sizeof(T) * (n ? n : 1))); }
    ~Dev() { if (p) (void)hipFree(p); }
};
}  // namespace

extern "C" int cipc_gpu_eval_cloth_impl(int n, const double* F, double st, double sh, double k, double* P, double* H, double* e) {
    using namespace cipc;
    Dev<double> dF(6 * (size_t)n), dP(6 * (size_t)n), dH(36 * (size_t)n), de(n);
    CIPC_HIP(hipMemcpy(dF.p, F, sizeof(double) * 6 * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(bw_eval_kernel, dim3(blocks_for(n, 64)), dim3(64), 0, 0, n, dF.p, st, sh, k, dP.p, dH.p, de.p);
    CIPC_CHECK_LAUNCH();
    CIPC_HIP(hipMemcpy(P, dP.p, sizeof(double) * 6 * n, hipMemcpyDeviceToHost));
    CIPC_HIP(hipMemcpy(H, dH.p, sizeof(double) * 36 * n, hipMemcpyDeviceToHost));
    CIPC_HIP(hipMemcpy(e, de.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int cipc_gpu_eval_hinge_impl(int n, const double* pts, const double* rest, const double* weight, double plate, double* g,
                                        double* H, double* e) {
    using namespace cipc;
    Dev<double> dp(12 * (size_t)n), dr(n), dw(n), dg(12 * (size_t)n), dH(144 * (size_t)n), de(n);
    CIPC_HIP(hipMemcpy(dp.p, pts, sizeof(double) * 12 * n, hipMemcpyHostToDevice));
    CIPC_HIP(hipMemcpy(dr.p, rest, sizeof(double) * n, hipMemcpyHostToDevice));
    CIPC_HIP(hipMemcpy(dw.p, weight, sizeof(double) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(hinge_eval_kernel, dim3(blocks_for(n, 64)), dim3(64), 0, 0, n, dp.p, dr.p, dw.p, plate, dg.p, dH.p, de.p);
    CIPC_CHECK_LAUNCH();
    CIPC_HIP(hipMemcpy(g, dg.p, sizeof(double) * 12 * n, hipMemcpyDeviceToHost));
    CIPC_HIP(hipMemcpy(H, dH.p, sizeof(double) * 144 * n, hipMemcpyDeviceToHost));
    CIPC_HIP(hipMemcpy(e, de.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return 0;
}
