// Line-search energy on the device.  Compiled with -ffp-contract=off.
//
// Reference: computeEnergy (CPU IPC/IPCSolver.cpp:322-456):
//   E = dt^2 (E_SNH + E_BW + E_bend) + sum m|x - x~|^2/2 + sum drag m|x - V_prev|^2/2 + E_soft
//       + kappa (sum_ground b + sum_self b * multiplicity + sum_EE e b) + E_friction
// with getObjEnergy_StableNHK2_3D (Elasticity.cpp:695-727) and the barrier sums of
// Evaluate_*ConstraintVals + compute_b (ContactMechanics.cpp:2478-2544,775-778).
//
// Each term is reduced deterministically: a fixed grid of 256 work-groups, each summing a
// grid-strided subset in index order and then a shared-memory tree; the partials are
// summed in order by one work-group.  The reference's TBB reductions have no fixed
// order; results agree to rounding.
#include "context.hpp"
#include "contact.hpp"

namespace cipc {

constexpr int kEB = kEnergyBlock;   // threads per work-group
constexpr int kEG = kEnergyGrid;    // work-groups

template <typename F>
__device__ inline void block_sum_store(double v, double* part, F) { energy_block_sum(v, part); }

__global__ __launch_bounds__(kEB) void e_snh(int T, const int4* __restrict__ tets, const double* __restrict__ x,
                                             const double* __restrict__ dminv, const double* __restrict__ vol, double u, double r,
                                             double* __restrict__ part) {
    double acc = 0.0;
    for (int t = blockIdx.x * kEB + threadIdx.x; t < T; t += kEG * kEB) {
        const int4 id = tets[t];
        const d3 p0 = ld3(x, id.x);
        const d3 e1 = ld3(x, id.y) - p0, e2 = ld3(x, id.z) - p0, e3 = ld3(x, id.w) - p0;
        double D[3][3];
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) D[k][j] = dminv[(size_t)(k + 3 * j) * T + t];
        d3 f[3];
        for (int j = 0; j < 3; ++j) {
            f[j] = d3{e1.x * D[0][j] + e2.x * D[1][j] + e3.x * D[2][j], e1.y * D[0][j] + e2.y * D[1][j] + e3.y * D[2][j],
                      e1.z * D[0][j] + e2.z * D[1][j] + e3.z * D[2][j]};
        }
        const double I2 = sqn(f[0]) + sqn(f[1]) + sqn(f[2]);
        const double I3 = dot(f[0], cross(f[1], f[2]));
        const double k = I3 - 1 - 3 * u / 4 / r;
        acc += (0.5 * u * (I2 - 3) + 0.5 * r * k * k - 0.5 * u * log(I2 + 1)) * vol[t];
    }
    block_sum_store(acc, part, 0);
}
__global__ __launch_bounds__(kEB) void e_inertia(int V, const double* __restrict__ x, const double* __restrict__ xt,
                                                 const double* __restrict__ vp, const double* __restrict__ m, double drag,
                                                 double* __restrict__ part) {
    double acc = 0.0;
    for (int v = blockIdx.x * kEB + threadIdx.x; v < V; v += kEG * kEB) {
        const d3 xv = ld3(x, v);
        acc += sqn(xv - ld3(xt, v)) * m[v] / 2.0;
        acc += (sqn(xv - ld3(vp, v)) * (drag) * m[v] / 2.0);
    }
    block_sum_store(acc, part, 0);
}
__global__ __launch_bounds__(kEB) void e_ground(int n, const int* __restrict__ env, const double* __restrict__ x, d3 gn, double gd,
                                                double hhat, double* __restrict__ part) {
    double acc = 0.0;
    for (int i = blockIdx.x * kEB + threadIdx.x; i < n; i += kEG * kEB) {
        const double dist = dot(gn, ld3(x, env[i])) - gd;
        acc += barrier_b(dist * dist, hhat);
    }
    block_sum_store(acc, part, 0);
}
__device__ inline double self_d(int4 c, const double* x) {
    if (c.x >= 0) return d_EE(ld3(x, c.x), ld3(x, c.y), ld3(x, c.z), ld3(x, c.w));
    const d3 p = ld3(x, -c.x - 1);
    if (c.z < 0) return d_PP(p, ld3(x, c.y));
    if (c.w < 0) return d_PE(p, ld3(x, c.y), ld3(x, c.z));
    return d_PT(p, ld3(x, c.y), ld3(x, c.z), ld3(x, c.w));
}
__global__ __launch_bounds__(kEB) void e_self(int n, const int4* __restrict__ set, const double* __restrict__ x, double hhat,
                                              double* __restrict__ part) {
    double acc = 0.0;
    for (int i = blockIdx.x * kEB + threadIdx.x; i < n; i += kEG * kEB) {
        const int4 c = set[i];
        double b = barrier_b(self_d(c, x), hhat);
        if (c.w < -1) b *= -c.w;
        acc += b;
    }
    block_sum_store(acc, part, 0);
}
__global__ __launch_bounds__(kEB) void e_mollified(int n, const int4* __restrict__ set, const int2* __restrict__ ie,
                                                   const int2* __restrict__ se, const double* __restrict__ x,
                                                   const double* __restrict__ vrest, double hhat, double* __restrict__ part) {
    double acc = 0.0;
    for (int i = blockIdx.x * kEB + threadIdx.x; i < n; i += kEG * kEB) {
        const int4 c = set[i];
        int e4[4];
        if (c.w >= 0) { e4[0] = c.x; e4[1] = c.y; e4[2] = c.z; e4[3] = c.w; }
        else {
            const int2 a = se[ie[i].x], b = se[ie[i].y];
            e4[0] = a.x; e4[1] = a.y; e4[2] = b.x; e4[3] = b.y;
        }
        const double ex = eps_x_of(ld3(vrest, e4[0]), ld3(vrest, e4[1]), ld3(vrest, e4[2]), ld3(vrest, e4[3]));
        const double e = mol_e(ee_cross_sqn(ld3(x, e4[0]), ld3(x, e4[1]), ld3(x, e4[2]), ld3(x, e4[3])), ex);
        acc += barrier_b(self_d(c, x), hhat) * e;
    }
    block_sum_store(acc, part, 0);
}
__global__ __launch_bounds__(kEB) void e_soft(int n, const int* __restrict__ idx, const double* __restrict__ tgt, double w,
                                              const double* __restrict__ x, double* __restrict__ part) {
    double acc = 0.0;
    for (int i = blockIdx.x * kEB + threadIdx.x; i < n; i += kEG * kEB) acc += 0.5 * w * sqn(ld3(x, idx[i]) - ld3(tgt, i));
    block_sum_store(acc, part, 0);
}
__global__ __launch_bounds__(kEB) void sum_partials(const double* __restrict__ part, int nterms, double* __restrict__ out) {
    __shared__ double sh[kEB];
    for (int t = 0; t < nterms; ++t) {
        sh[threadIdx.x] = part[t * kEG + threadIdx.x];
        __syncthreads();
        for (int s = kEB / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[t] = sh[0];
        __syncthreads();
    }
}

void energy_terms(Context& c, double kappa, double out[9]) {
    (void)kappa;
    Scene& s = c.sc;
    static_assert(kEG == kEB, "partials are summed by one work-group");
    DBuf<double>& part = c.dscratch;
    part.resize((size_t)11 * kEG);
    double* P = part.p;
    hipLaunchKernelGGL(e_snh, dim3(kEG), dim3(kEB), 0, c.stream, s.T, s.tets.p, s.x.p, s.dminv.p, s.vol.p, s.u, s.r, P);
    hipLaunchKernelGGL(e_inertia, dim3(kEG), dim3(kEB), 0, c.stream, s.V, s.x.p, s.xt.p, s.vprev.p, s.mass.p, s.drag, P + kEG);
    hipLaunchKernelGGL(e_ground, dim3(kEG), dim3(kEB), 0, c.stream, (int)c.n_env, c.env_set.p, s.x.p, s.gn, s.gd, s.hhat, P + 2 * kEG);
    hipLaunchKernelGGL(e_self, dim3(kEG), dim3(kEB), 0, c.stream, (int)c.n_self, c.self_set.p, s.x.p, s.hhat, P + 3 * kEG);
    hipLaunchKernelGGL(e_mollified, dim3(kEG), dim3(kEB), 0, c.stream, (int)c.n_ee, c.ee_set.p, c.ee_ie.p, s.se.p, s.x.p, s.vrest.p,
                       s.hhat, P + 4 * kEG);
    // lagged friction: self part (lambda times clamped norm) and ground part (mu included)
    friction_energy_partials(c, P + 5 * kEG);
    // cloth triangles and bending (IPCSolver.cpp:330-368)
    cloth_energy_partials(c, P + 7 * kEG, P + 8 * kEG);
    // soft targets (softEnergy, BoundaryConditions.cpp:210-229)
    hipLaunchKernelGGL(e_soft, dim3(kEG), dim3(kEB), 0, c.stream, c.n_soft, c.soft_idx.p, c.soft_t.p, c.soft_w, s.x.p, P + 9 * kEG);
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(kEB), 0, c.stream, P, 10, P + 10 * kEG);
    CIPC_CHECK_LAUNCH();
    double t[10];
    CIPC_HIP(hipMemcpyAsync(t, P + 10 * kEG, 10 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CIPC_HIP(hipStreamSynchronize(c.stream));
    for (int k = 0; k < 5; ++k) out[k] = t[k];
    out[5] = s.fric * t[5] + t[6];   // Friction::energy: mu * sum_self + sum_ground (Friction.cpp:521-544)
    out[6] = t[7];
    out[7] = s.bmode == 0 ? 0.5 * s.plate * t[8] : t[8];   // quadratic form: D/2 x^T (Q (x) I3) x
    out[8] = t[9];
}

// Total as computeEnergy adds it: dt^2 (E_SNH + E_BW + E_bend) + inertia + soft + kappa
// barriers + friction.
double total_energy(const Context& c, double kappa, const double t[9]) {
    double E = t[0];
    E += t[6];
    E += t[7];
    E *= c.sc.dt * c.sc.dt;
    E += t[1];
    E += t[8];
    E += kappa * (t[2] + t[3] + t[4]);
    E += t[5];
    return E;
}

__global__ void step_forward_kernel(int V, const int* __restrict__ bt, const double* __restrict__ x0, const double* __restrict__ p,
                                    double a, double* __restrict__ x) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || !is_free(bt[v])) return;
    st3(x, v, ld3(x0, v) - a * ld3(p, v));
}
void step_forward(Context& c, double alpha) {
    Scene& s = c.sc;
    hipLaunchKernelGGL(step_forward_kernel, dim3(blocks_for(s.V, 256)), dim3(256), 0, c.stream, s.V, s.bt.p, s.x0.p, s.dir.p, alpha, s.x.p);
    CIPC_CHECK_LAUNCH();
}

}  // namespace cipc
