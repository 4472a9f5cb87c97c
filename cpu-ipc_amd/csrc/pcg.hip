// Newton linear solve on the device: H p = rhs for the system of the last assemble
// (NewtonLinearSystem::solve, CPU IPC/NewtonLinearSystem.cpp:21-56; tolerance and iteration
// cap as LinearSolverOptions.h:9-19, relative residual ||rhs - H p|| / ||rhs||).
//
// The matrix is used as a full symmetric 3x3-block matrix (BSR) built from the merged lower
// blocks of the assembly: every block (a, b), a > b, is also placed at (b, a) transposed, and
// the entries are sorted by (block row, block column) with one radix sort.  Zero entries
// that the CSR export drops contribute nothing to a product, so the two agree.
//
// Two forms:
//   precision 0: FP64 conjugate gradients with the 3x3 block-Jacobi preconditioner;
//   precision 1: the same iteration in FP32 (blocks, vectors) as the inner solver of FP64
//                iterative refinement: r = rhs - H x in FP64, solve H d = r in FP32 to a
//                loose relative tolerance, x += d, until the FP64 residual meets the target.
//
// One iteration is three kernels; scalars (alpha, beta, norms) stay on the device and are
// combined from per-work-group partials in a fixed order (deterministic).  The host checks
// convergence every kCheck iterations; kernels launched past convergence return at once.
#include "context.hpp"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>

namespace cipc {

constexpr int kPB = 256;       // threads per work-group (one vertex = 3 DOFs per thread)
constexpr int kCheck = 16;     // iterations between host convergence checks
constexpr int kMaxRefine = 64; // FP64 refinement steps of the mixed form

// device scalars of one solve
struct PcgScalars {
    double rz;        // r . z
    double rr;        // r . r
    double bn;        // ||b||
    double tol;       // absolute residual target
    int done;         // converged (or broken down)
    int iters;
};

// ---- full symmetric BSR from the merged lower blocks -------------------------------
__global__ void bsr_entries(int nm, const int* __restrict__ mrow, const int* __restrict__ mcol, int kb,
                            unsigned long long* __restrict__ key, int* __restrict__ src, const int* __restrict__ nm_ptr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = *nm_ptr;
    if (k >= nm) return;
    if (k >= n) {   // past the merged count: sentinels sorted last
        key[2 * k] = key[2 * k + 1] = ~0ull;
        src[2 * k] = src[2 * k + 1] = -1;
        return;
    }
    const unsigned a = (unsigned)mrow[k], b = (unsigned)mcol[k];
    key[2 * k] = ((unsigned long long)a << kb) | b;          // (a, b) as stored
    src[2 * k] = 2 * k;
    if (a != b) {
        key[2 * k + 1] = ((unsigned long long)b << kb) | a;  // (b, a) = transpose
        src[2 * k + 1] = 2 * k + 1;
    } else {
        key[2 * k + 1] = ~0ull;
        src[2 * k + 1] = -1;
    }
}
template <typename T>
__global__ void bsr_fill(int ne, const unsigned long long* __restrict__ key, const int* __restrict__ src, const double* __restrict__ mval,
                         int kb, int* __restrict__ col, T* __restrict__ val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    const int s = src[i];
    if (s < 0) return;
    col[i] = (int)(key[i] & ((1ull << kb) - 1));
    const double* m = mval + (size_t)9 * (s >> 1);
    const bool tr = s & 1;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) val[(size_t)9 * i + 3 * r + c] = (T)(tr ? m[3 * c + r] : m[3 * r + c]);
}
// row pointers: first entry of every block row among the sorted keys
__global__ void bsr_rowptr(int V, int ne, const unsigned long long* __restrict__ key, int kb, int* __restrict__ rowptr) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a > V) return;
    const unsigned long long k0 = (unsigned long long)a << kb;
    int lo = 0, hi = ne;
    while (lo < hi) { const int m = (lo + hi) >> 1; if (key[m] < k0) lo = m + 1; else hi = m; }
    rowptr[a] = lo;
}
// block-Jacobi: inverse of each diagonal block (adjugate / determinant, symmetric)
template <typename T>
__global__ void diag_inverse(int V, const int* __restrict__ rowptr, const int* __restrict__ col, const T* __restrict__ val,
                             T* __restrict__ minv) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= V) return;
    double d[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int e = rowptr[a]; e < rowptr[a + 1]; ++e)
        if (col[e] == a) { for (int k = 0; k < 9; ++k) d[k] = (double)val[(size_t)9 * e + k]; break; }
    const double c00 = d[4] * d[8] - d[5] * d[7], c01 = d[5] * d[6] - d[3] * d[8], c02 = d[3] * d[7] - d[4] * d[6];
    const double det = d[0] * c00 + d[1] * c01 + d[2] * c02;
    const double id = det != 0.0 ? 1.0 / det : 0.0;
    const double inv[9] = {c00 * id, (d[2] * d[7] - d[1] * d[8]) * id, (d[1] * d[5] - d[2] * d[4]) * id,
                           c01 * id, (d[0] * d[8] - d[2] * d[6]) * id, (d[2] * d[3] - d[0] * d[5]) * id,
                           c02 * id, (d[1] * d[6] - d[0] * d[7]) * id, (d[0] * d[4] - d[1] * d[3]) * id};
    for (int k = 0; k < 9; ++k) minv[(size_t)9 * a + k] = (T)inv[k];
}

// ---- helpers ---------------------------------------------------------------------------
template <typename T>
__device__ inline void block_mv(const T* __restrict__ B, const T* __restrict__ x, double y[3]) {
    y[0] = (double)B[0] * x[0] + (double)B[1] * x[1] + (double)B[2] * x[2];
    y[1] = (double)B[3] * x[0] + (double)B[4] * x[1] + (double)B[5] * x[2];
    y[2] = (double)B[6] * x[0] + (double)B[7] * x[1] + (double)B[8] * x[2];
}
// work-group sum of up to three values; thread 0 stores them at part[k * gridDim.x + blockIdx.x]
template <int K>
__device__ inline void group_partials(double v[K], double* __restrict__ part) {
    __shared__ double sh[K][kPB];
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = kPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int k = 0; k < K; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) part[k * gridDim.x + blockIdx.x] = sh[k][0];
}
// every work-group sums the partials of the previous kernel in the same order
__device__ inline double sum_partials_fixed(const double* __restrict__ part, int n) {
    __shared__ double sh[kPB];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kPB) acc += part[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// ---- PCG kernels (T: storage precision of blocks and vectors; scalars FP64) ------------
// r = b - A x (A x skipped when x is zero), z = M^-1 r, p = z; partials of r.z, r.r, b.b
template <typename T>
__global__ __launch_bounds__(kPB) void pcg_start(int V, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                 const T* __restrict__ val, const T* __restrict__ minv, const double* __restrict__ b,
                                                 const T* __restrict__ x, int x_zero, T* __restrict__ r, T* __restrict__ z,
                                                 T* __restrict__ p, double* __restrict__ part) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[3] = {0, 0, 0};
    if (a < V) {
        double rv[3] = {b[3 * a], b[3 * a + 1], b[3 * a + 2]};
        const double bb = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
        if (!x_zero) {
            for (int e = rowptr[a]; e < rowptr[a + 1]; ++e) {
                double y[3];
                block_mv(val + (size_t)9 * e, x + 3 * (size_t)col[e], y);
                rv[0] -= y[0]; rv[1] -= y[1]; rv[2] -= y[2];
            }
        }
        double zv[3];
        const T rt[3] = {(T)rv[0], (T)rv[1], (T)rv[2]};
        block_mv(minv + (size_t)9 * a, rt, zv);
        for (int k = 0; k < 3; ++k) { r[3 * a + k] = rt[k]; z[3 * a + k] = (T)zv[k]; p[3 * a + k] = (T)zv[k]; }
        acc[0] = (double)rt[0] * (T)zv[0] + (double)rt[1] * (T)zv[1] + (double)rt[2] * (T)zv[2];
        acc[1] = (double)rt[0] * rt[0] + (double)rt[1] * rt[1] + (double)rt[2] * rt[2];
        acc[2] = bb;
    }
    group_partials<3>(acc, part);
}
// scalars from the start partials: rz, rr, ||b|| (or ||r0|| as the reference of an inner solve)
__global__ __launch_bounds__(kPB) void pcg_init_scalars(const double* __restrict__ part, int ng, double rel_tol, int rel_to_r0,
                                                        PcgScalars* __restrict__ sc) {
    const double rz = sum_partials_fixed(part, ng);
    const double rr = sum_partials_fixed(part + ng, ng);
    const double bb = sum_partials_fixed(part + 2 * ng, ng);
    if (threadIdx.x == 0) {
        sc->rz = rz;
        sc->rr = rr;
        sc->bn = sqrt(bb);
        sc->tol = rel_tol * (rel_to_r0 ? sqrt(rr) : sqrt(bb));
        sc->done = !(sqrt(rr) > sc->tol) || !(rz > 0.0);
        sc->iters = 0;
    }
}
// q = A p and partials of p.q
template <typename T>
__global__ __launch_bounds__(kPB) void pcg_spmv(int V, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                const T* __restrict__ val, const T* __restrict__ p, T* __restrict__ q,
                                                const PcgScalars* __restrict__ sc, double* __restrict__ part) {
    if (sc->done) return;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[1] = {0.0};
    if (a < V) {
        double y[3] = {0, 0, 0};
        for (int e = rowptr[a]; e < rowptr[a + 1]; ++e) {
            double t[3];
            block_mv(val + (size_t)9 * e, p + 3 * (size_t)col[e], t);
            y[0] += t[0]; y[1] += t[1]; y[2] += t[2];
        }
        const T yt[3] = {(T)y[0], (T)y[1], (T)y[2]};
        for (int k = 0; k < 3; ++k) q[3 * a + k] = yt[k];
        acc[0] = (double)p[3 * a] * yt[0] + (double)p[3 * a + 1] * yt[1] + (double)p[3 * a + 2] * yt[2];
    }
    group_partials<1>(acc, part);
}
// alpha = rz / p.q; x += alpha p; r -= alpha q; z = M^-1 r; partials of r.z and r.r
template <typename T>
__global__ __launch_bounds__(kPB) void pcg_update(int V, const double* __restrict__ part_pq, int ng, const T* __restrict__ minv,
                                                  const T* __restrict__ p, const T* __restrict__ q, T* __restrict__ x, T* __restrict__ r,
                                                  T* __restrict__ z, const PcgScalars* __restrict__ sc, double* __restrict__ part) {
    if (sc->done) return;
    const double pq = sum_partials_fixed(part_pq, ng);
    const double alpha = sc->rz / pq;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    if (a < V) {
        T rt[3];
        for (int k = 0; k < 3; ++k) {
            x[3 * a + k] = (T)((double)x[3 * a + k] + alpha * (double)p[3 * a + k]);
            rt[k] = (T)((double)r[3 * a + k] - alpha * (double)q[3 * a + k]);
            r[3 * a + k] = rt[k];
        }
        double zv[3];
        block_mv(minv + (size_t)9 * a, rt, zv);
        for (int k = 0; k < 3; ++k) z[3 * a + k] = (T)zv[k];
        acc[0] = (double)rt[0] * (T)zv[0] + (double)rt[1] * (T)zv[1] + (double)rt[2] * (T)zv[2];
        acc[1] = (double)rt[0] * rt[0] + (double)rt[1] * rt[1] + (double)rt[2] * rt[2];
    }
    group_partials<2>(acc, part);
}
// beta = rz' / rz; p = z + beta p.  Every work-group reads the previous r.z here; the
// scalars (and the convergence flag) are replaced by the following one-work-group kernel,
// so no work-group can see a value another one has changed.
template <typename T>
__global__ __launch_bounds__(kPB) void pcg_direction(int V, const double* __restrict__ part, int ng, const T* __restrict__ z,
                                                     T* __restrict__ p, const PcgScalars* __restrict__ sc) {
    if (sc->done) return;
    const double rz1 = sum_partials_fixed(part, ng);
    const double beta = rz1 / sc->rz;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < V)
        for (int k = 0; k < 3; ++k) p[3 * a + k] = (T)((double)z[3 * a + k] + beta * (double)p[3 * a + k]);
}
__global__ void pcg_scalars(const double* __restrict__ part, int ng, PcgScalars* __restrict__ sc, int max_iter) {
    if (sc->done) return;
    const double rz1 = sum_partials_fixed(part, ng);
    const double rr = sum_partials_fixed(part + ng, ng);
    if (threadIdx.x == 0) {
        sc->rz = rz1;
        sc->rr = rr;
        sc->iters += 1;
        if (!(sqrt(rr) > sc->tol) || sc->iters >= max_iter || !(rz1 > 0.0)) sc->done = 1;
    }
}
// x (FP64) += d (FP32), for the refinement
__global__ void add_correction(int n, const float* __restrict__ d, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += (double)d[i];
}
template <typename S, typename D>
__global__ void convert(int n, const S* __restrict__ s, D* __restrict__ d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (D)s[i];
}
__global__ void zero_d(int n, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}
__global__ void zero_f(int n, float* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0f;
}

// ---- host ---------------------------------------------------------------------------
namespace {
PcgScalars read_scalars(Context& c, const PcgScalars* d) {
    PcgScalars h;
    CIPC_HIP(hipMemcpyAsync(&h, d, sizeof(PcgScalars), hipMemcpyDeviceToHost, c.stream));
    CIPC_HIP(hipStreamSynchronize(c.stream));
    return h;
}

// One PCG solve of A x = b (b FP64; A, x and the work vectors in T).  Returns iterations.
template <typename T>
int pcg_run(Context& c, const T* val, const T* minv, const double* b, T* x, int x_zero, double rel_tol, int rel_to_r0, int max_iter,
            double& rr_out, double& bn_out) {
    const int V = c.sc.V;
    const int ng = (int)blocks_for(V, kPB);
    SolveBuffers& w = c.solve;
    T* r = reinterpret_cast<T*>(w.r.p);
    T* z = reinterpret_cast<T*>(w.z.p);
    T* p = reinterpret_cast<T*>(w.p.p);
    T* q = reinterpret_cast<T*>(w.q.p);
    double* part = w.part.p;
    PcgScalars* sc = reinterpret_cast<PcgScalars*>(w.scal.p);
    hipLaunchKernelGGL(pcg_start<T>, dim3(ng), dim3(kPB), 0, c.stream, V, w.rowptr.p, w.col.p, val, minv, b, x, x_zero, r, z, p, part);
    hipLaunchKernelGGL(pcg_init_scalars, dim3(1), dim3(kPB), 0, c.stream, part, ng, rel_tol, rel_to_r0, sc);
    CIPC_CHECK_LAUNCH();
    PcgScalars h = read_scalars(c, sc);
    while (!h.done) {
        for (int k = 0; k < kCheck; ++k) {
            hipLaunchKernelGGL(pcg_spmv<T>, dim3(ng), dim3(kPB), 0, c.stream, V, w.rowptr.p, w.col.p, val, p, q, sc, part);
            hipLaunchKernelGGL(pcg_update<T>, dim3(ng), dim3(kPB), 0, c.stream, V, part, ng, minv, p, q, x, r, z, sc, part + ng);
            hipLaunchKernelGGL(pcg_direction<T>, dim3(ng), dim3(kPB), 0, c.stream, V, part + ng, ng, z, p, sc);
            hipLaunchKernelGGL(pcg_scalars, dim3(1), dim3(kPB), 0, c.stream, part + ng, ng, sc, max_iter);
        }
        CIPC_CHECK_LAUNCH();
        h = read_scalars(c, sc);
    }
    rr_out = h.rr;
    bn_out = h.bn;
    return h.iters;
}
}  // namespace

// Full BSR (FP64 and, for the mixed form, FP32) and block-Jacobi inverses of the last assembly.
void build_solve_matrix(Context& c, bool fp32) {
    Scene& s = c.sc;
    const int V = s.V;
    SolveBuffers& w = c.solve;
    const size_t mcap = c.mrow.n;                 // merged blocks are bounded by the buffers' size
    const size_t ne = 2 * mcap;
    w.key.resize(ne); w.key_sorted.resize(ne); w.src.resize(ne); w.src_sorted.resize(ne);
    w.col.resize(ne); w.val.resize(9 * ne); w.rowptr.resize((size_t)V + 1); w.minv.resize((size_t)9 * V);
    const int* nm_ptr = c.mptr.p + V;
    hipLaunchKernelGGL(bsr_entries, dim3(blocks_for(mcap, 256)), dim3(256), 0, c.stream, (int)mcap, c.mrow.p, c.mcol.p, s.kb, w.key.p,
                       w.src.p, nm_ptr);
    CIPC_CHECK_LAUNCH();
    // keys use the low 2 kb bits; sentinels have them all set and sort last
    sort_u64_pairs(c, w.key.p, w.key_sorted.p, w.src.p, w.src_sorted.p, (int)ne, 2 * s.kb);
    hipLaunchKernelGGL(bsr_fill<double>, dim3(blocks_for(ne, 256)), dim3(256), 0, c.stream, (int)ne, w.key_sorted.p, w.src_sorted.p,
                       c.mval.p, s.kb, w.col.p, w.val.p);
    hipLaunchKernelGGL(bsr_rowptr, dim3(blocks_for((size_t)V + 1, 256)), dim3(256), 0, c.stream, V, (int)ne, w.key_sorted.p, s.kb,
                       w.rowptr.p);
    hipLaunchKernelGGL(diag_inverse<double>, dim3(blocks_for(V, 256)), dim3(256), 0, c.stream, V, w.rowptr.p, w.col.p, w.val.p, w.minv.p);
    CIPC_CHECK_LAUNCH();
    if (fp32) {
        w.valf.resize(9 * ne); w.minvf.resize((size_t)9 * V);
        hipLaunchKernelGGL((convert<double, float>), dim3(blocks_for(9 * ne, 256)), dim3(256), 0, c.stream, (int)(9 * ne), w.val.p,
                           w.valf.p);
        hipLaunchKernelGGL(diag_inverse<float>, dim3(blocks_for(V, 256)), dim3(256), 0, c.stream, V, w.rowptr.p, w.col.p, w.valf.p,
                           w.minvf.p);
        CIPC_CHECK_LAUNCH();
    }
}

void linear_solve(Context& c, int precision, double rel_tol, int max_iter, SolveResult& res) {
    Scene& s = c.sc;
    const int V = s.V, n = 3 * V;
    SolveBuffers& w = c.solve;
    const int ng = (int)blocks_for(V, kPB);
    w.r.resize((size_t)n); w.z.resize((size_t)n); w.p.resize((size_t)n); w.q.resize((size_t)n);
    w.part.resize((size_t)4 * ng);
    w.scal.resize((sizeof(PcgScalars) + 7) / 8);
    w.x.resize((size_t)n);
    build_solve_matrix(c, precision == 1);
    res = SolveResult{};
    if (precision == 0) {
        double rr = 0, bn = 0;
        hipLaunchKernelGGL(zero_d, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, w.x.p);
        res.iterations = pcg_run<double>(c, w.val.p, w.minv.p, c.rhs.p, w.x.p, 1, rel_tol, 0, max_iter, rr, bn);
        res.rel_residual = bn > 0 ? std::sqrt(rr) / bn : 0.0;
    } else {
        // FP64 refinement around FP32 PCG
        w.xf.resize((size_t)n); w.res.resize((size_t)n);
        hipLaunchKernelGGL(zero_d, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, w.x.p);
        const double inner_tol = 1e-3;
        double rr = 0, bn = 0;
        int iters = 0, outer = 0;
        for (;;) {
            // r = b - A x in FP64, ||r|| / ||b||
            res.iterations = iters;
            double* rd = w.res.p;
            {
                // reuse the FP64 start kernel to form r (and its norms) without a solve
                hipLaunchKernelGGL(pcg_start<double>, dim3(ng), dim3(kPB), 0, c.stream, V, w.rowptr.p, w.col.p, w.val.p, w.minv.p,
                                   c.rhs.p, w.x.p, outer == 0 ? 1 : 0, rd, w.q.p, w.p.p, w.part.p);
                hipLaunchKernelGGL(pcg_init_scalars, dim3(1), dim3(kPB), 0, c.stream, w.part.p, ng, rel_tol, 0,
                                   reinterpret_cast<PcgScalars*>(w.scal.p));
                CIPC_CHECK_LAUNCH();
                const PcgScalars h = read_scalars(c, reinterpret_cast<PcgScalars*>(w.scal.p));
                rr = h.rr; bn = h.bn;
                if (!(std::sqrt(rr) > rel_tol * bn) || iters >= max_iter) break;
            }
            // inner FP32 solve of A d = r
            hipLaunchKernelGGL(zero_f, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, w.xf.p);
            double irr = 0, ibn = 0;
            const int it = pcg_run<float>(c, w.valf.p, w.minvf.p, rd, w.xf.p, 1, inner_tol, 1, max_iter - iters, irr, ibn);
            if (it == 0 || outer >= kMaxRefine) break;   // no progress possible (breakdown) or refinement cap
            iters += it;
            hipLaunchKernelGGL(add_correction, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, w.xf.p, w.x.p);
            CIPC_CHECK_LAUNCH();
            ++outer;
        }
        res.iterations = iters;
        res.refinements = outer;
        res.rel_residual = bn > 0 ? std::sqrt(rr) / bn : 0.0;
    }
    // the Newton direction for the step bound and line search
    s.dir.resize((size_t)n);
    CIPC_HIP(hipMemcpyAsync(s.dir.p, w.x.p, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    CIPC_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace cipc
