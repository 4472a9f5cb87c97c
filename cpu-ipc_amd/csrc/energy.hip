// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate expressions of CPU-IPC
// (MPL-2.0) in the reference's operation order; each such part cites the source it follows.
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

#include <mutex>

namespace cipc {

constexpr int kEB = kEnergyBlock;   // threads per work-group
constexpr int kEG = kEnergyGrid;    // work-groups

// one partial per work-group; none when the kernel stores per-element values instead (part
// null: the same for every thread of the work-group, so the barriers stay uniform)
template <typename F>
__device__ inline void block_sum_store(double v, double* part, F) {
    if (part) energy_block_sum(v, part);
}

__device__ inline void snh_sum(int T, const int4* __restrict__ tets, const double* __restrict__ x,
                               const double* __restrict__ dminv, const double* __restrict__ vol, Lame lame,
                               double* __restrict__ part, double* __restrict__ elem) {
    double acc = 0.0;
    for (int t = blockIdx.x * kEB + threadIdx.x; t < T; t += gridDim.x * kEB) {
        double u, r;
        lame.at(t, u, r);
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
        const double et = (0.5 * u * (I2 - 3) + 0.5 * r * k * k - 0.5 * u * log(I2 + 1)) * vol[t];
        if (elem) elem[t] = et;
        else acc += et;
    }
    block_sum_store(acc, part, 0);
}
__device__ inline void inertia_sum(int V, const double* __restrict__ x, const double* __restrict__ xt,
                                   const double* __restrict__ vp, const double* __restrict__ m, double drag,
                                   double* __restrict__ part, double* __restrict__ elem) {
    double acc = 0.0;
    for (int v = blockIdx.x * kEB + threadIdx.x; v < V; v += gridDim.x * kEB) {
        const d3 xv = ld3(x, v);
        if (elem) {
            elem[v] = sqn(xv - ld3(xt, v)) * m[v] / 2.0 + (sqn(xv - ld3(vp, v)) * (drag) * m[v] / 2.0);
            continue;
        }
        acc += sqn(xv - ld3(xt, v)) * m[v] / 2.0;
        acc += (sqn(xv - ld3(vp, v)) * (drag) * m[v] / 2.0);
    }
    block_sum_store(acc, part, 0);
}
// The contact terms take their count from device memory (ndev, the active set's device
// count) when it is given, else n.
__device__ inline void ground_sum(int n, const int* __restrict__ ndev, const int* __restrict__ env, const double* __restrict__ x,
                                  d3 gn, double gd, double hhat, double* __restrict__ part, double* __restrict__ elem) {
    double acc = 0.0;
    if (ndev) n = *ndev;
    for (int i = blockIdx.x * kEB + threadIdx.x; i < n; i += gridDim.x * kEB) {
        const double dist = dot(gn, ld3(x, env[i])) - gd;
        const double b = barrier_b(dist * dist, hhat);
        if (elem) elem[i] = b;
        else acc += b;
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
__device__ inline void self_sum(int n, const int* __restrict__ ndev, const int4* __restrict__ set, const double* __restrict__ x,
                                double hhat, double* __restrict__ part, double* __restrict__ elem) {
    double acc = 0.0;
    if (ndev) n = *ndev;
    for (int i = blockIdx.x * kEB + threadIdx.x; i < n; i += gridDim.x * kEB) {
        const int4 c = set[i];
        double b = barrier_b(self_d(c, x), hhat);
        if (c.w < -1) b *= -c.w;
        if (elem) elem[i] = b;
        else acc += b;
    }
    block_sum_store(acc, part, 0);
}
__device__ inline void mollified_sum(int n, const int* __restrict__ ndev, const int4* __restrict__ set,
                                     const int2* __restrict__ ie, const int2* __restrict__ se, const double* __restrict__ x,
                                     const double* __restrict__ vrest, double hhat, double* __restrict__ part,
                                     double* __restrict__ elem) {
    double acc = 0.0;
    if (ndev) n = *ndev;
    for (int i = blockIdx.x * kEB + threadIdx.x; i < n; i += gridDim.x * kEB) {
        const int4 c = set[i];
        int e4[4];
        if (c.w >= 0) { e4[0] = c.x; e4[1] = c.y; e4[2] = c.z; e4[3] = c.w; }
        else {
            const int2 a = se[ie[i].x], b = se[ie[i].y];
            e4[0] = a.x; e4[1] = a.y; e4[2] = b.x; e4[3] = b.y;
        }
        const double ex = eps_x_of(ld3(vrest, e4[0]), ld3(vrest, e4[1]), ld3(vrest, e4[2]), ld3(vrest, e4[3]));
        const double e = mol_e(ee_cross_sqn(ld3(x, e4[0]), ld3(x, e4[1]), ld3(x, e4[2]), ld3(x, e4[3])), ex);
        const double b = barrier_b(self_d(c, x), hhat) * e;
        if (elem) elem[i] = b;
        else acc += b;
    }
    block_sum_store(acc, part, 0);
}
// Per-element values (elem, one grid-strided thread per element) or one partial per
// work-group (part); the batch path takes the first, the single-scene sums the second.
__global__ __launch_bounds__(kEB) void e_snh(int T, const int4* __restrict__ tets, const double* __restrict__ x,
                                             const double* __restrict__ dminv, const double* __restrict__ vol, Lame lame,
                                             double* __restrict__ part, double* __restrict__ elem) {
    snh_sum(T, tets, x, dminv, vol, lame, part, elem);
}
__global__ __launch_bounds__(kEB) void e_inertia(int V, const double* __restrict__ x, const double* __restrict__ xt,
                                                 const double* __restrict__ vp, const double* __restrict__ m, double drag,
                                                 double* __restrict__ part, double* __restrict__ elem) {
    inertia_sum(V, x, xt, vp, m, drag, part, elem);
}
__global__ __launch_bounds__(kEB) void e_ground(int n, const int* __restrict__ ndev, const int* __restrict__ env, const double* __restrict__ x,
                                                d3 gn, double gd, double hhat, double* __restrict__ part, double* __restrict__ elem) {
    ground_sum(n, ndev, env, x, gn, gd, hhat, part, elem);
}
__global__ __launch_bounds__(kEB) void e_self(int n, const int* __restrict__ ndev, const int4* __restrict__ set, const double* __restrict__ x,
                                              double hhat, double* __restrict__ part, double* __restrict__ elem) {
    self_sum(n, ndev, set, x, hhat, part, elem);
}
__global__ __launch_bounds__(kEB) void e_mollified(int n, const int* __restrict__ ndev, const int4* __restrict__ set,
                                                   const int2* __restrict__ ie, const int2* __restrict__ se, const double* __restrict__ x,
                                                   const double* __restrict__ vrest, double hhat, double* __restrict__ part,
                                                   double* __restrict__ elem) {
    mollified_sum(n, ndev, set, ie, se, x, vrest, hhat, part, elem);
}
// The five state and contact terms of one scene in one launch: work-group row y sums term y
// exactly as the term's own kernel on a kEG-wide grid would (the same grid stride, the same
// work-group index), so the partials are the same; the contact counts are the active set's
// device counts.
struct EnergyTermArgs {
    int T, V;
    const int4* tets;
    const double *x, *dminv, *vol, *xt, *vp, *m, *vrest;
    Lame lame;
    double drag, gd, hhat;
    d3 gn;
    const int* dc;
    const int* env;
    const int4 *self_set, *ee_set;
    const int2 *ee_ie, *se;
};
constexpr int kFusedTerms = 5;
__global__ __launch_bounds__(kEB) void e_terms(EnergyTermArgs a, double* __restrict__ P) {
    double* part = P + (size_t)blockIdx.y * kEG;
    switch (blockIdx.y) {
    case 0: snh_sum(a.T, a.tets, a.x, a.dminv, a.vol, a.lame, part, nullptr); break;
    case 1: inertia_sum(a.V, a.x, a.xt, a.vp, a.m, a.drag, part, nullptr); break;
    case 2: ground_sum(0, a.dc + kDcEnv, a.env, a.x, a.gn, a.gd, a.hhat, part, nullptr); break;
    case 3: self_sum(0, a.dc + kDcSelf, a.self_set, a.x, a.hhat, part, nullptr); break;
    default: mollified_sum(0, a.dc + kDcMol, a.ee_set, a.ee_ie, a.se, a.x, a.vrest, a.hhat, part, nullptr); break;
    }
}
__global__ __launch_bounds__(kEB) void e_soft(int n, const int* __restrict__ idx, const double* __restrict__ tgt, double w,
                                              const double* __restrict__ x, double* __restrict__ part) {
    double acc = 0.0;
    for (int i = blockIdx.x * kEB + threadIdx.x; i < n; i += gridDim.x * kEB) acc += 0.5 * w * sqn(ld3(x, idx[i]) - ld3(tgt, i));
    block_sum_store(acc, part, 0);
}
// out[t] = the partials of term t summed in a fixed tree; a term whose bit is clear in
// `present` was not launched (no elements): its sum is zero.
__global__ __launch_bounds__(kEB) void sum_partials(const double* __restrict__ part, int nterms, unsigned present,
                                                   double* __restrict__ out) {
    __shared__ double sh[kEB];
    for (int t = 0; t < nterms; ++t) {
        if (!((present >> t) & 1u)) {
            if (threadIdx.x == 0) out[t] = 0.0;
            continue;
        }
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
    if (c.sc.S != 1) throw Error(CIPC_ERR_LOGIC, "a batch has one energy per scene");
    energy_terms_launch(c, c.stream);
    CIPC_HIP(hipStreamSynchronize(c.stream));
    energy_terms_finish(c, out);
}

void energy_terms_launch(Context& c, hipStream_t st) {
    DBuf<double>& part = c.dscratch;
    part.resize((size_t)kEnergyScratch);
    energy_partials_launch(c, st, part.p);
    c.pin_energy.reserve(10);
    CIPC_HIP(hipMemcpyAsync(c.pin_energy.p, part.p + 10 * kEG, 10 * sizeof(double), hipMemcpyDeviceToHost, st));
}

// The partial sums of the ten energy terms (kEG per term) and, after them, the ten sums,
// into P (kEnergyScratch doubles).  Terms with no elements in the scene (cloth, soft
// targets) are not launched and their partials are neither written nor read; every other
// entry of P is written before it is read.
void energy_partials_launch(Context& c, hipStream_t st, double* P) {
    Scene& s = c.sc;
    static_assert(kEG == kEB, "partials are summed by one work-group");
    EnergyTermArgs a;
    a.T = s.T; a.V = s.V; a.tets = s.tets.p; a.x = s.x.p; a.dminv = s.dminv.p; a.vol = s.vol.p; a.xt = s.xt.p;
    a.vp = s.vprev.p; a.m = s.mass.p; a.vrest = s.vrest.p; a.lame = scene_lame(s); a.drag = s.drag; a.gd = s.gd;
    a.hhat = s.hhat; a.gn = s.gn; a.dc = c.dcnt.p; a.env = c.env_set.p; a.self_set = c.self_set.p; a.ee_set = c.ee_set.p;
    a.ee_ie = c.ee_ie.p; a.se = s.se.p;
    hipLaunchKernelGGL(e_terms, dim3(kEG, kFusedTerms), dim3(kEB), 0, st, a, P);
    unsigned present = (1u << kFusedTerms) - 1;
    // lagged friction: self part (lambda times clamped norm) and ground part (mu included)
    friction_energy_partials(c, P + 5 * kEG, nullptr, nullptr, st);
    present |= 3u << 5;
    // cloth triangles and bending (IPCSolver.cpp:330-368), when the scene has them
    if (s.Tr > 0 || s.Nb > 0) {
        cloth_energy_partials(c, P + 7 * kEG, P + 8 * kEG, st);
        present |= 3u << 7;
    }
    // soft targets (softEnergy, BoundaryConditions.cpp:210-229), when set
    if (c.n_soft > 0) {
        hipLaunchKernelGGL(e_soft, dim3(kEG), dim3(kEB), 0, st, c.n_soft, c.soft_idx.p, c.soft_t.p, c.soft_w, s.x.p, P + 9 * kEG);
        present |= 1u << 9;
    }
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(kEB), 0, st, P, 10, present, P + 10 * kEG);
    CIPC_CHECK_LAUNCH();
}

// energy_terms_finish on the device: the nine terms from the ten sums
__global__ void finish_terms_kernel(const double* __restrict__ t, double fric, int quad_bending, double plate, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 5; ++k) out[k] = t[k];
    out[5] = fric * t[5] + t[6];
    out[6] = t[7];
    out[7] = quad_bending ? 0.5 * plate * t[8] : t[8];
    out[8] = t[9];
}

// ---- the allocation-free energy ------------------------------------------------------
// Its partial sums live in a file-scope device array instead of the context's scratch.  The
// array's host-side name is not a device address: each device's copy is found once, with
// that device current, by hipGetSymbolAddress, and only that address reaches the kernels.
// Every call on a device shares its copy: calls that overlap on one device race.
static_assert(kEnergyScratch == 11 * kEG, "ten terms of kEG partials and the ten sums");
__device__ __align__(256) double scratch[kEnergyScratch];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};
double* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(CIPC_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        CIPC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    return g_scratch_ptr[device];
}
}  // namespace

void energy_terms_noalloc(Context& c, double* out_device) {
    const Scene& s = c.sc;
    if (s.S != 1) throw Error(CIPC_ERR_LOGIC, "a batch has one energy per scene: the allocation-free energy takes one scene");
    double* P = scratch_on(c.device);
    energy_partials_launch(c, c.stream, P);
    hipLaunchKernelGGL(finish_terms_kernel, dim3(1), dim3(64), 0, c.stream, P + 10 * kEG, s.fric, s.bmode == 0 ? 1 : 0, s.plate, out_device);
    CIPC_CHECK_LAUNCH();
    CIPC_HIP(hipStreamSynchronize(c.stream));
}

void energy_terms_finish(Context& c, double out[9]) {
    const Scene& s = c.sc;
    const double* t = c.pin_energy.p;
    for (int k = 0; k < 5; ++k) out[k] = t[k];
    out[5] = s.fric * t[5] + t[6];   // Friction::energy: mu * sum_self + sum_ground (Friction.cpp:521-544)
    out[6] = t[7];
    out[7] = s.bmode == 0 ? 0.5 * s.plate * t[8] : t[8];   // quadratic form: D/2 x^T (Q (x) I3) x
    out[8] = t[9];
}

// ---- per-scene sums of a batch ---------------------------------------------------------
// Each term's per-element values are summed per scene.  Tets and vertices of a scene are
// contiguous ranges.  A contact list is ordered by scene within each of its parts (the
// queries run over surface vertices and edges in scene order), so it splits into runs of
// one scene each: every run is summed by one work-group and the runs of a scene are added
// in list order.  The order of every addition is fixed, so the sums are deterministic.
constexpr int kRunGrid = 1024;
// per-element values (no partial sums): one thread per element, so that a large batch fills
// the GPU (the order of the later sums does not depend on this grid)
inline unsigned elem_grid(int64_t n) { return std::max(1u, blocks_for((size_t)n, kEB)); }
__device__ inline int first_vertex(int4 c) { return c.x >= 0 ? c.x : -c.x - 1; }
__device__ inline int first_vertex(int v) { return v; }
// block sum of one value per thread (all threads get it)
__device__ inline double block_total(double v) {
    __shared__ double sh[kEB];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kEB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}
// out[k * 9 + term] = sum of elem[off[k] .. off[k + 1]) (one work-group per scene)
__global__ __launch_bounds__(kEB) void range_sums(const double* __restrict__ elem, const int* __restrict__ off, int term,
                                                  double* __restrict__ out) {
    const int k = blockIdx.x;
    double acc = 0.0;
    for (int i = off[k] + threadIdx.x; i < off[k + 1]; i += kEB) acc += elem[i];
    const double t = block_total(acc);
    if (threadIdx.x == 0) out[9 * k + term] = t;
}
// head[i] = 1 where the scene of record i differs from record i - 1
template <typename R>
__global__ void scene_run_heads(int n, const R* __restrict__ rec, const int* __restrict__ vs, int* __restrict__ key,
                                int* __restrict__ head) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { head[n] = 0; return; }
    const int k = vs[first_vertex(rec[i])];
    key[i] = k;
    head[i] = (i == 0 || vs[first_vertex(rec[i - 1])] != k) ? 1 : 0;
}
__global__ void run_starts(int n, const int* __restrict__ head, const int* __restrict__ pos, int* __restrict__ start,
                           int* __restrict__ nruns) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { const int r = n > 0 ? pos[n - 1] + head[n - 1] : 0; start[r] = n; *nruns = r; return; }
    if (head[i]) start[pos[i]] = i;
}
// run sums: work-group b sums runs b, b + kRunGrid, ...
__global__ __launch_bounds__(kEB) void run_sums(const double* __restrict__ elem, const int* __restrict__ start,
                                                const int* __restrict__ nruns, double* __restrict__ rsum) {
    const int R = *nruns;
    for (int r = blockIdx.x; r < R; r += gridDim.x) {
        double acc = 0.0;
        for (int i = start[r] + threadIdx.x; i < start[r + 1]; i += kEB) acc += elem[i];
        const double t = block_total(acc);
        if (threadIdx.x == 0) rsum[r] = t;
    }
}
// out[k * 9 + term] = sum over scene k's runs in list order (one work-group per scene)
__global__ __launch_bounds__(kEB) void scene_from_runs(const double* __restrict__ rsum, const int* __restrict__ start,
                                                       const int* __restrict__ key, const int* __restrict__ nruns, int term,
                                                       double* __restrict__ out) {
    const int k = blockIdx.x, R = *nruns;
    double acc = 0.0;
    for (int r = threadIdx.x; r < R; r += kEB)
        if (key[start[r]] == k) acc += rsum[r];
    const double t = block_total(acc);
    if (threadIdx.x == 0) out[9 * k + term] = t;
}
__global__ void zero_term(int S, int term, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < S) out[9 * k + term] = 0.0;
}

namespace {
template <typename R>
void scene_list_sums(Context& c, int n, const R* rec, const double* elem, int term, double* out) {
    Scene& s = c.sc;
    const int S = s.S;
    if (n == 0) {
        hipLaunchKernelGGL(zero_term, dim3(blocks_for(S, 64)), dim3(64), 0, c.stream, S, term, out);
        CIPC_CHECK_LAUNCH();
        return;
    }
    DBuf<int>& key = c.iscratch;
    DBuf<int>& aux = c.iscratch2;   // head (n + 1), pos (n + 1), start (n + 1), count (1)
    key.resize((size_t)n + 1);
    aux.resize(3 * ((size_t)n + 1) + 1);
    int* head = aux.p;
    int* pos = aux.p + (n + 1);
    int* start = aux.p + 2 * (size_t)(n + 1);
    int* cnt = aux.p + 3 * (size_t)(n + 1);
    c.dscratch2.resize((size_t)n + 1);
    hipLaunchKernelGGL(scene_run_heads<R>, dim3(blocks_for((size_t)n + 1, 256)), dim3(256), 0, c.stream, n, rec, s.vscene.p, key.p, head);
    CIPC_CHECK_LAUNCH();
    exclusive_scan_int(c, head, pos, n + 1);
    hipLaunchKernelGGL(run_starts, dim3(blocks_for((size_t)n + 1, 256)), dim3(256), 0, c.stream, n, head, pos, start, cnt);
    hipLaunchKernelGGL(run_sums, dim3(kRunGrid), dim3(kEB), 0, c.stream, elem, start, cnt, c.dscratch2.p);
    hipLaunchKernelGGL(scene_from_runs, dim3(S), dim3(kEB), 0, c.stream, c.dscratch2.p, start, key.p, cnt, term, out);
    CIPC_CHECK_LAUNCH();
}
}  // namespace

void energy_terms_scenes(Context& c, double* out) {
    Scene& s = c.sc;
    const int S = s.S;
    if (S == 1) {
        energy_terms(c, 0.0, out);
        return;
    }
    if (s.Tr > 0 || s.Nb > 0 || c.n_soft > 0) throw Error(CIPC_ERR_LOGIC, "batched scenes hold tets only");
    // per-element values: tets, vertices, ground, self, mollified, friction self, friction ground
    const size_t nT = s.T, nV = s.V, nG = c.n_env, nS = c.n_self, nE = c.n_ee, nFs = c.n_fself, nFg = c.n_fenv;
    DBuf<double>& el = c.dscratch;
    el.resize(nT + nV + nG + nS + nE + nFs + nFg + (size_t)9 * S + 1);
    double* eT = el.p;
    double* eV = eT + nT;
    double* eG = eV + nV;
    double* eS = eG + nG;
    double* eE = eS + nS;
    double* eFs = eE + nE;
    double* eFg = eFs + nFs;
    double* dev_out = eFg + nFg;
    hipLaunchKernelGGL(e_snh, dim3(elem_grid(s.T)), dim3(kEB), 0, c.stream, s.T, s.tets.p, s.x.p, s.dminv.p, s.vol.p, scene_lame(s), nullptr, eT);
    hipLaunchKernelGGL(e_inertia, dim3(elem_grid(s.V)), dim3(kEB), 0, c.stream, s.V, s.x.p, s.xt.p, s.vprev.p, s.mass.p, s.drag, nullptr, eV);
    if (nG) hipLaunchKernelGGL(e_ground, dim3(elem_grid(nG)), dim3(kEB), 0, c.stream, (int)nG, nullptr, c.env_set.p, s.x.p, s.gn, s.gd, s.hhat, nullptr, eG);
    if (nS) hipLaunchKernelGGL(e_self, dim3(elem_grid(nS)), dim3(kEB), 0, c.stream, (int)nS, nullptr, c.self_set.p, s.x.p, s.hhat, nullptr, eS);
    if (nE)
        hipLaunchKernelGGL(e_mollified, dim3(elem_grid(nE)), dim3(kEB), 0, c.stream, (int)nE, nullptr, c.ee_set.p, c.ee_ie.p, s.se.p, s.x.p, s.vrest.p,
                           s.hhat, nullptr, eE);
    if (nFs + nFg) friction_energy_partials(c, nullptr, eFs, eFg, c.stream);
    CIPC_CHECK_LAUNCH();
    // device offsets of tets and vertices follow the output block
    DBuf<int>& off = c.batch_off;
    hipLaunchKernelGGL(range_sums, dim3(S), dim3(kEB), 0, c.stream, eT, off.p, 0, dev_out);
    hipLaunchKernelGGL(range_sums, dim3(S), dim3(kEB), 0, c.stream, eV, off.p + (S + 1), 1, dev_out);
    CIPC_CHECK_LAUNCH();
    scene_list_sums<int>(c, (int)nG, c.env_set.p, eG, 2, dev_out);
    scene_list_sums<int4>(c, (int)nS, c.self_set.p, eS, 3, dev_out);
    scene_list_sums<int4>(c, (int)nE, c.ee_set.p, eE, 4, dev_out);
    scene_list_sums<int4>(c, (int)nFs, c.fr_self.p, eFs, 5, dev_out);
    scene_list_sums<int>(c, (int)nFg, c.fr_env.p, eFg, 6, dev_out);
    std::vector<double> h((size_t)9 * S);
    read_back(c, h.data(), dev_out, h.size());
    for (int k = 0; k < S; ++k) {
        const double* t = h.data() + 9 * (size_t)k;
        double* o = out + 9 * (size_t)k;
        for (int q = 0; q < 5; ++q) o[q] = t[q];
        o[5] = s.fric * t[5] + t[6];   // as energy_terms: mu * sum_self + sum_ground
        o[6] = 0.0;
        o[7] = 0.0;
        o[8] = 0.0;
    }
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
__global__ void step_forward_scenes_kernel(int V, const int* __restrict__ bt, const int* __restrict__ vs, const double* __restrict__ x0,
                                           const double* __restrict__ p, const double* __restrict__ a, double* __restrict__ x) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || !is_free(bt[v])) return;
    st3(x, v, ld3(x0, v) - a[vs[v]] * ld3(p, v));
}
void step_forward_scenes(Context& c, const double* alpha) {
    Scene& s = c.sc;
    if (s.S == 1) { step_forward(c, alpha[0]); return; }
    c.batch_alpha.resize(s.S);
    CIPC_HIP(hipMemcpyAsync(c.batch_alpha.p, alpha, s.S * sizeof(double), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(step_forward_scenes_kernel, dim3(blocks_for(s.V, 256)), dim3(256), 0, c.stream, s.V, s.bt.p, s.vscene.p, s.x0.p,
                       s.dir.p, c.batch_alpha.p, s.x.p);
    CIPC_CHECK_LAUNCH();
    CIPC_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace cipc
