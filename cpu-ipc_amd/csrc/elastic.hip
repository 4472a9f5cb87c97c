// Elastic stage of the Newton assembly: inertia + stable neo-Hookean tets.
//
// Reference: computeGradientAndHessian, inertia/drag and tet loops (CPU IPC/IPCSolver.cpp:
// 118-164), the mass diagonal of NewtonLinearSystem::assemble (NewtonLinearSystem.cpp:
// 133-141) and the lower-triangle/free-vertex/zero filters of BHessian::toTriplets
// (ContactMechanics.cpp:13-134).
//
// Data flow (deterministic, no atomics):
//   1. snh_element_kernel: one thread per tet, F from x and Dm^-1 (SoA), projected
//      12x12 stiffness.  Every (tet, local block) and (tet, corner) has a slot fixed when
//      the pattern is built: its position in the list of contributions sorted by
//      (block column, block row, tet), or by (vertex, tet).  The kernel writes each 3x3
//      block (9 values, in the stored block's orientation) and each corner force (3
//      values) to its slot.
//   2. block_gather_kernel: one thread per 3x3 block of the static block pattern sums
//      its contiguous run of slots in tet order and records which entries received a
//      nonzero contribution (the reference drops zero entries before summing
//      duplicates); mass_diag_kernel then adds the mass on diagonal blocks.
//   3. vertex_gather_kernel: one thread per vertex, inertia + drag + its run of corner
//      forces.  The gathers stream their slots; only the element kernel scatters.
#include "context.hpp"
#include "snh.hpp"

#include <algorithm>
#include <numeric>

namespace cipc {

// ---------------------------------------------------------------------------------
// Static pattern: block column b holds block rows a >= b; diagonal blocks exist for
// every vertex (mass), off-diagonal ones where an element (tet, triangle or bending
// element) couples two free vertices.
// ---------------------------------------------------------------------------------
void build_static_pattern(Context& c) {
    Scene& s = c.sc;
    const int V = s.V, T = s.T, Tr = s.Tr, Nb = s.Nb;
    const int NE = T + Tr + Nb;
    const std::vector<int>& bt = s.h_bt;
    // vertices of element e (tets, then triangles, then bending elements); nv[e] corners
    auto ev = [&](int e, int a) {
        if (e < T) return s.h_tets[4 * (size_t)e + a];
        if (e < T + Tr) return s.h_tris[3 * (size_t)(e - T) + a];
        return s.h_bverts[4 * (size_t)(e - T - Tr) + a];
    };
    auto env = [&](int e) { return (e >= T && e < T + Tr) ? 3 : 4; };
    // (col, row, src) triples; src = element*16 + local block
    struct Rec { int col, row; long long src; };
    std::vector<Rec> rec;
    rec.reserve((size_t)NE * 10 + V);
    for (int v = 0; v < V; ++v) rec.push_back({v, v, -1});
    for (int e = 0; e < NE; ++e) {
        const int nv = env(e);
        for (int lb = 0; lb < 10; ++lb) {
            if (lb_a(lb) >= nv) continue;
            const int va = ev(e, lb_a(lb)), vb = ev(e, lb_b(lb));
            if (!is_free(bt[va]) || !is_free(bt[vb])) continue;
            rec.push_back({std::min(va, vb), std::max(va, vb), (long long)e * 16 + lb});
        }
    }
    // counting sort by column, then by (row, src) inside a column
    std::vector<int> cnt(V + 1, 0);
    for (const Rec& r : rec) cnt[r.col + 1]++;
    for (int v = 0; v < V; ++v) cnt[v + 1] += cnt[v];
    std::vector<Rec> srt(rec.size());
    {
        std::vector<int> pos(cnt.begin(), cnt.end() - 1);
        for (const Rec& r : rec) srt[pos[r.col]++] = r;
    }
    std::vector<int> col_ptr(V + 1, 0), row, gptr, diag(V, -1);
    std::vector<long long> gsrc;
    row.reserve(rec.size() / 2);
    gptr.reserve(rec.size() / 2);
    gsrc.reserve((size_t)NE * 10);
    for (int v = 0; v < V; ++v) {
        auto b = srt.begin() + cnt[v], e = srt.begin() + cnt[v + 1];
        std::sort(b, e, [](const Rec& x, const Rec& y) {
            if (x.row != y.row) return x.row < y.row;
            return x.src < y.src;  // -1 (mass marker) first, then element order
        });
        for (auto it = b; it != e; ++it) {
            if (it == b || it->row != (it - 1)->row) {
                if (it->row == v) diag[v] = (int)row.size();
                row.push_back(it->row);
                gptr.push_back((int)gsrc.size());
            }
            if (it->src >= 0) gsrc.push_back(it->src);
        }
        col_ptr[v + 1] = (int)row.size();
    }
    gptr.push_back((int)gsrc.size());
    // vertex gradient gather lists: element corners of every vertex in element order
    std::vector<int> vptr(V + 1, 0);
    size_t ncorner = 0;
    for (int e = 0; e < NE; ++e)
        for (int a = 0; a < env(e); ++a) { vptr[ev(e, a) + 1]++; ++ncorner; }
    for (int v = 0; v < V; ++v) vptr[v + 1] += vptr[v];
    std::vector<long long> vsrc(ncorner);
    {
        std::vector<int> pos(vptr.begin(), vptr.end() - 1);
        for (int e = 0; e < NE; ++e)
            for (int a = 0; a < env(e); ++a) vsrc[pos[ev(e, a)]++] = (long long)e * 4 + a;
    }
    StaticPattern& sp = c.sp;
    sp.nblocks = (int)row.size();
    auto up = [&](DBuf<int>& d, const std::vector<int>& h) {
        d.resize(std::max<size_t>(h.size(), 1));
        if (!h.empty()) CIPC_HIP(hipMemcpy(d.p, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
    };
    // slot of every (element, local block) and (element, corner): its position in the sorted lists
    if (gsrc.size() >= (size_t)INT32_MAX / 9 || vsrc.size() >= (size_t)INT32_MAX / 3)
        throw Error(CIPC_ERR_OVERFLOW, "element pattern exceeds 32-bit staging offsets");
    std::vector<int> slot_h((size_t)10 * std::max(NE, 1), -1), slot_g((size_t)4 * std::max(NE, 1), -1);
    for (size_t j = 0; j < gsrc.size(); ++j) slot_h[(size_t)10 * (gsrc[j] >> 4) + (gsrc[j] & 15)] = (int)j;
    for (size_t j = 0; j < vsrc.size(); ++j) slot_g[vsrc[j]] = (int)j;
    up(sp.col_ptr, col_ptr); up(sp.row, row); up(sp.gptr, gptr); up(sp.vptr, vptr); up(sp.diag_block, diag);
    up(sp.slot_h, slot_h); up(sp.slot_g, slot_g);
    std::vector<int> vtet(V, 0);
    for (int i = 0; i < 4 * T; ++i) vtet[s.h_tets[i]]++;
    up(sp.vtet, vtet);
    c.stagH.resize((size_t)9 * std::max<size_t>(gsrc.size(), 1));
    c.stagG.resize((size_t)3 * std::max<size_t>(vsrc.size(), 1));
    c.sbv.resize((size_t)9 * std::max(sp.nblocks, 1));
    c.sbf.resize(std::max(sp.nblocks, 1));
    c.grad.resize((size_t)3 * std::max(V, 1));
}

// ---------------------------------------------------------------------------------
// Per-tet kernel
// ---------------------------------------------------------------------------------
__device__ inline void load_tet(int t, int T, const int4* __restrict__ tets, const double* __restrict__ x,
                                const double* __restrict__ dminv, M33& F, d3 B[4]) {
    const int4 id = tets[t];
    const d3 p0 = ld3(x, id.x);
    const d3 e1 = ld3(x, id.y) - p0, e2 = ld3(x, id.z) - p0, e3 = ld3(x, id.w) - p0;
    double D[3][3];  // D[k][j] = Dm^-1(k, j); stored column-major, SoA over tets
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) D[k][j] = dminv[(size_t)(k + 3 * j) * T + t];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            F.a[i][j] = comp(e1, i) * D[0][j] + comp(e2, i) * D[1][j] + comp(e3, i) * D[2][j];
    B[0] = d3{-(D[0][0] + D[1][0] + D[2][0]), -(D[0][1] + D[1][1] + D[2][1]), -(D[0][2] + D[1][2] + D[2][2])};
    B[1] = d3{D[0][0], D[0][1], D[0][2]};
    B[2] = d3{D[1][0], D[1][1], D[1][2]};
    B[3] = d3{D[2][0], D[2][1], D[2][2]};
}

// Block K_ab (3x3) of the projected element stiffness before the vol dt^2 factor.
__device__ inline void snh_block(const SNH& e, const d3 B[4], const d3 G1[4], const d3 GJ[4],
                                 const double beta[4][3], int a, int b, double K[3][3]) {
    const double bb = dot(B[a], B[b]);
    const d3 w = mvec(e.F, cross(B[a], B[b]));
    const double h[3][3] = {{0, -w.z, w.y}, {w.z, 0, -w.x}, {-w.y, w.x, 0}};
    const double c1 = 4.0 * e.a3;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            K[i][j] = (i == j ? e.A * bb : 0.0) - e.a2 * h[i][j] + c1 * comp(G1[a], i) * comp(G1[b], j)
                    + e.r * comp(GJ[a], i) * comp(GJ[b], j);
    if (e.project) {
#pragma unroll
        for (int m = 0; m < 9; ++m) {
            if (e.lam[m] < 0.0) {
                const d3 qa = snh_mode_apply(e, m, beta[a]);
                const d3 qb = snh_mode_apply(e, m, beta[b]);
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) K[i][j] -= e.lam[m] * comp(qa, i) * comp(qb, j);
            }
        }
    }
}

// H: staging type of the Hessian blocks (double; float for the mixed form)
template <typename H>
__global__ __launch_bounds__(256) void snh_element_kernel(int T, const int4* __restrict__ tets, const double* __restrict__ x,
                                                          const double* __restrict__ dminv, const double* __restrict__ vol,
                                                          Lame lame, double dt2, const int* __restrict__ slot_h,
                                                          const int* __restrict__ slot_g, H* __restrict__ stagH,
                                                          double* __restrict__ stagG) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    SNH e;
    d3 B[4];
    load_tet(t, T, tets, x, dminv, e.F, B);
    const int4 tv = tets[t];
    const int vid[4] = {tv.x, tv.y, tv.z, tv.w};
    double u, r;
    lame.at(t, u, r);
    snh_eval(e, u, r);
    const double vt = vol[t];
    // gradient f_a = vol P B_a (dt^2 is applied by the vertex gather, as IPCSolver.cpp:149-151)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const d3 f = vt * (e.pa * mvec(e.F, B[a]) + e.pb * mvec(e.C, B[a]));
        st3(stagG, slot_g[4 * t + a], f);
    }
    d3 G1[4], GJ[4];
    double beta[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        G1[a] = mvec(e.F, B[a]);
        GJ[a] = mvec(e.C, B[a]);
#pragma unroll
        for (int k = 0; k < 3; ++k) beta[a][k] = dot(mcol(e.V, k), B[a]);
    }
    const double sc = vt * dt2;
#pragma unroll
    for (int lb = 0; lb < 10; ++lb) {
        const int j = slot_h[10 * t + lb];
        if (j < 0) continue;   // a vertex of the block is not free
        const int a = lb_a(lb), b = lb_b(lb);
        double K[3][3];
        snh_block(e, B, G1, GJ, beta, a, b, K);
        // stored block (row vertex max(va, vb), column vertex min): K_ab, or K_ab^T when
        // va < vb; a diagonal block is symmetric and stored whole
        const bool tr = vid[a] < vid[b];
        H* o = stagH + (size_t)9 * j;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k) o[3 * i + k] = (H)(sc * (tr ? K[k][i] : K[i][k]));
    }
}

// ---------------------------------------------------------------------------------
// Gathers
// ---------------------------------------------------------------------------------
template <typename H>
__global__ __launch_bounds__(256) void block_gather_kernel(int nb, const int* __restrict__ gptr, const H* __restrict__ stagH,
                                                           double* __restrict__ sbv, unsigned short* __restrict__ sbf) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    unsigned mask = 0;
    for (int q = gptr[k]; q < gptr[k + 1]; ++q) {
        const H* p = stagH + (size_t)9 * q;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const double v = (double)p[i];
            if (v != 0.0) { acc[i] += v; mask |= 1u << i; }
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) sbv[(size_t)9 * k + i] = acc[i];
    sbf[k] = (unsigned short)mask;
}

// Adds the mass m (1 + drag) on every diagonal DOF (NewtonLinearSystem.cpp:133-141).
__global__ void mass_diag_kernel(int V, const int* __restrict__ diag_block, const double* __restrict__ mass, double drag,
                                 double* __restrict__ sbv, unsigned short* __restrict__ sbf) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int k = diag_block[v];
    const double m = mass[v] * (1.0 + drag);
    sbv[(size_t)9 * k + 0] += m;
    sbv[(size_t)9 * k + 4] += m;
    sbv[(size_t)9 * k + 8] += m;
    sbf[k] |= (1u << 0) | (1u << 4) | (1u << 8);
}

// g = m (x - x~); g += drag m (x - V_prev); g += dt^2 f over elements (IPCSolver.cpp:118-267).
__global__ __launch_bounds__(256) void vertex_gather_kernel(int V, int T, const double* __restrict__ x, const double* __restrict__ xt,
                                                            const double* __restrict__ vprev, const double* __restrict__ mass,
                                                            double drag, double dt2, const int* __restrict__ vptr,
                                                            const double* __restrict__ stagG, double* __restrict__ grad) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const d3 xv = ld3(x, v);
    const double m = mass[v];
    d3 g = m * (xv - ld3(xt, v));
    g = g + (drag * m) * (xv - ld3(vprev, v));
    for (int q = vptr[v]; q < vptr[v + 1]; ++q) g = g + dt2 * ld3(stagG, q);
    st3(grad, v, g);
}

// Soft targets: g += w (x - t), w I on the vertex's diagonal block (addSoftDerivatives,
// BoundaryConditions.cpp:175-208).  Soft vertices are distinct and free.
__global__ void soft_kernel(int n, const int* __restrict__ idx, const double* __restrict__ tgt, double w,
                            const double* __restrict__ x, const int* __restrict__ diag_block, double* __restrict__ grad,
                            double* __restrict__ sbv, unsigned short* __restrict__ sbf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = idx[i];
    st3(grad, v, ld3(grad, v) + w * (ld3(x, v) - ld3(tgt, i)));
    const int k = diag_block[v];
    sbv[(size_t)9 * k + 0] += w;
    sbv[(size_t)9 * k + 4] += w;
    sbv[(size_t)9 * k + 8] += w;
    sbf[k] |= (1u << 0) | (1u << 4) | (1u << 8);
}

void elastic_assemble(Context& c, hipStream_t st) {
    Scene& s = c.sc;
    const double dt2 = s.dt * s.dt;
    if (s.T > 0) {
        if (c.hessian_fp32)
            hipLaunchKernelGGL(snh_element_kernel<float>, dim3(blocks_for(s.T, 256)), dim3(256), 0, st, s.T, s.tets.p, s.x.p,
                               s.dminv.p, s.vol.p, scene_lame(s), dt2, c.sp.slot_h.p, c.sp.slot_g.p,
                               reinterpret_cast<float*>(c.stagH.p), c.stagG.p);
        else
            hipLaunchKernelGGL(snh_element_kernel<double>, dim3(blocks_for(s.T, 256)), dim3(256), 0, st, s.T, s.tets.p, s.x.p,
                               s.dminv.p, s.vol.p, scene_lame(s), dt2, c.sp.slot_h.p, c.sp.slot_g.p, c.stagH.p, c.stagG.p);
    }
    CIPC_CHECK_LAUNCH();
    cloth_assemble(c, st);   // triangles and bending elements into the same staging
    if (c.sp.nblocks > 0) {
        if (c.hessian_fp32)
            hipLaunchKernelGGL(block_gather_kernel<float>, dim3(blocks_for(c.sp.nblocks, 256)), dim3(256), 0, st, c.sp.nblocks,
                               c.sp.gptr.p, reinterpret_cast<const float*>(c.stagH.p), c.sbv.p, c.sbf.p);
        else
            hipLaunchKernelGGL(block_gather_kernel<double>, dim3(blocks_for(c.sp.nblocks, 256)), dim3(256), 0, st, c.sp.nblocks,
                               c.sp.gptr.p, c.stagH.p, c.sbv.p, c.sbf.p);
        CIPC_CHECK_LAUNCH();
        hipLaunchKernelGGL(mass_diag_kernel, dim3(blocks_for(s.V, 256)), dim3(256), 0, st, s.V, c.sp.diag_block.p,
                           s.mass.p, s.drag, c.sbv.p, c.sbf.p);
        CIPC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(vertex_gather_kernel, dim3(blocks_for(s.V, 256)), dim3(256), 0, st, s.V, s.T, s.x.p, s.xt.p,
                       s.vprev.p, s.mass.p, s.drag, dt2, c.sp.vptr.p, c.stagG.p, c.grad.p);
    CIPC_CHECK_LAUNCH();
    if (c.n_soft > 0) {
        hipLaunchKernelGGL(soft_kernel, dim3(blocks_for(c.n_soft, 256)), dim3(256), 0, st, c.n_soft, c.soft_idx.p, c.soft_t.p,
                           c.soft_w, s.x.p, c.sp.diag_block.p, c.grad.p, c.sbv.p, c.sbf.p);
        CIPC_CHECK_LAUNCH();
    }
}

// ---------------------------------------------------------------------------------
// Element-level evaluation (tests): PK1, projected 9x9 Hessian and energy density.
// ---------------------------------------------------------------------------------
__global__ void snh_eval_kernel(int n, const double* __restrict__ Fin, double u, double r, double* __restrict__ P,
                                double* __restrict__ H, double* __restrict__ e) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    SNH s;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) s.F.a[i][j] = Fin[9 * t + i + 3 * j];
    snh_eval(s, u, r);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) P[9 * t + i + 3 * j] = s.pa * s.F.a[i][j] + s.pb * s.C.a[i][j];
    // explicit 9x9 over column-major vec(F): A I + a2 HJ + a3 g1 g1^T + r gJ gJ^T - sum lam q q^T
    for (int p = 0; p < 9; ++p)
        for (int q = 0; q < 9; ++q) {
            const int ip = p % 3, jp = p / 3, iq = q % 3, jq = q / 3;
            // HJ[(jp,ip),(jq,iq)] = -sum_k eps(jp,jq,k) hat(f_k)[ip][iq]
            double hj = 0.0;
            if (jp != jq) {
                const int k = 3 - jp - jq;
                const double eps = ((jp + 1) % 3 == jq) ? 1.0 : -1.0;
                const d3 f = mcol(s.F, k);
                const double hat[3][3] = {{0, -f.z, f.y}, {f.z, 0, -f.x}, {-f.y, f.x, 0}};
                hj = -eps * hat[ip][iq];
            }
            double h = (p == q ? s.A : 0.0) + s.a2 * hj + s.a3 * (2 * s.F.a[ip][jp]) * (2 * s.F.a[iq][jq])
                     + s.r * s.C.a[ip][jp] * s.C.a[iq][jq];
            if (s.project) {
                for (int m = 0; m < 9; ++m) {
                    if (s.lam[m] >= 0.0) continue;
                    double bp[3], bq[3];
                    for (int k = 0; k < 3; ++k) { bp[k] = s.V.a[jp][k]; bq[k] = s.V.a[jq][k]; }
                    const d3 qp = snh_mode_apply(s, m, bp), qq = snh_mode_apply(s, m, bq);
                    h -= s.lam[m] * comp(qp, ip) * comp(qq, iq);
                }
            }
            H[81 * t + 9 * p + q] = h;
        }
    e[t] = snh_energy_density(s.I2, s.I3, u, r);
}

}  // namespace cipc

extern "C" int cipc_gpu_eval_snh_impl(int n, const double* F, double u, double r, double* P, double* H, double* e) {
    using namespace cipc;
    double *dF, *dP, *dH, *de;
    CIPC_HIP(hipMalloc(&dF, sizeof(double) * 9 * n));
    CIPC_HIP(hipMalloc(&dP, sizeof(double) * 9 * n));
    CIPC_HIP(hipMalloc(&dH, sizeof(double) * 81 * n));
    CIPC_HIP(hipMalloc(&de, sizeof(double) * n));
    CIPC_HIP(hipMemcpy(dF, F, sizeof(double) * 9 * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(snh_eval_kernel, dim3(blocks_for(n, 64)), dim3(64), 0, 0, n, dF, u, r, dP, dH, de);
    CIPC_CHECK_LAUNCH();
    CIPC_HIP(hipMemcpy(P, dP, sizeof(double) * 9 * n, hipMemcpyDeviceToHost));
    CIPC_HIP(hipMemcpy(H, dH, sizeof(double) * 81 * n, hipMemcpyDeviceToHost));
    CIPC_HIP(hipMemcpy(e, de, sizeof(double) * n, hipMemcpyDeviceToHost));
    hipFree(dF); hipFree(dP); hipFree(dH); hipFree(de);
    return 0;
}
