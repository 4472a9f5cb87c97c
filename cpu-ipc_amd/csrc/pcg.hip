// Newton linear solve on the device: H p = rhs for the system of the last assemble
// (NewtonLinearSystem::solve, CPU IPC/NewtonLinearSystem.cpp:21-56; tolerance and iteration
// cap as LinearSolverOptions.h:9-19, relative residual ||rhs - H p|| / ||rhs||).
//
// The matrix is used as a full symmetric 3x3-block matrix (BSR) built from the merged lower
// blocks of the assembly: every block (a, b), a > b, is also placed at (b, a) transposed, and
// the entries are sorted by (block row, block column) with one radix sort.  Zero entries
// that the CSR export drops contribute nothing to a product, so the two agree.
//
// Two precisions:
//   precision 0: FP64 conjugate gradients;
//   precision 1: the same iteration in FP32 (blocks, vectors) as the inner solver of FP64
//                iterative refinement: r = rhs - H x in FP64, solve H d = r in FP32, x += d,
//                until the FP64 residual meets the target.
// Three preconditioners: 3x3 block-Jacobi; a block lower-triangular factor G with
// M^-1 = G^T G, either the block FSAI of the matrix or the learned factor (learned.hip).
//
// Scalars (alpha, beta, norms) stay on the device and are combined from per-work-group
// partials in a fixed order (deterministic).  kCheck iterations are captured once into a
// HIP graph and replayed; the host checks convergence after each replay and kernels
// launched past convergence return at once.
#include <rocsolver/rocsolver.h>
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
    double pq;        // p . q of the current iteration
    double beta;      // r.z' / r.z of the current iteration
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
// Row-parallel sparse products: kGL lanes per 3x3-block row, so that the lanes of a row read
// consecutive blocks (coalesced), combined with cross-lane shuffles.  kRows rows per work-group.
constexpr int kGL = 8;
constexpr int kRows = kPB / kGL;
template <typename T>
__device__ inline void group_reduce3(double y[3]) {
#pragma unroll
    for (int o = kGL / 2; o > 0; o >>= 1) {
        y[0] += __shfl_xor(y[0], o, kGL);
        y[1] += __shfl_xor(y[1], o, kGL);
        y[2] += __shfl_xor(y[2], o, kGL);
    }
}
// Interleaved copy of the BSR values for the products: entries in groups of kGL, each
// group stored component-major (entry e, component k at (e / kGL) 9 kGL + k kGL + e % kGL), so
// that the kGL lanes of a row, which read consecutive entries, load each component with one
// contiguous access instead of nine 72-byte-strided ones.
__host__ __device__ inline size_t il_index(size_t e, int k) { return (e / kGL) * (9 * kGL) + (size_t)k * kGL + e % kGL; }
template <typename T>
__global__ void interleave_blocks(size_t ne, const T* __restrict__ val, T* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;   // (entry, component)
    if (i >= 9 * ne) return;
    const size_t e = i / 9;
    out[il_index(e, (int)(i % 9))] = val[i];
}
template <typename T, typename X>
__device__ inline void block_mv_il(const T* __restrict__ B, size_t e, const X* __restrict__ x, double y[3]) {
    const T* b = B + (e / kGL) * (9 * kGL) + e % kGL;
    y[0] = (double)b[0] * x[0] + (double)b[kGL] * x[1] + (double)b[2 * kGL] * x[2];
    y[1] = (double)b[3 * kGL] * x[0] + (double)b[4 * kGL] * x[1] + (double)b[5 * kGL] * x[2];
    y[2] = (double)b[6 * kGL] * x[0] + (double)b[7 * kGL] * x[1] + (double)b[8 * kGL] * x[2];
}
// y = (A x)_a over entries rowptr[a] + sub, step kGL (all lanes of the group get the sum);
// val interleaved
template <typename T, typename X>
__device__ inline void row_product(int a, int sub, const int* __restrict__ rowptr, const int* __restrict__ col, const T* __restrict__ val,
                                   const X* __restrict__ x, double y[3]) {
    y[0] = y[1] = y[2] = 0.0;
    for (int e = rowptr[a] + sub; e < rowptr[a + 1]; e += kGL) {
        double t[3];
        block_mv_il(val, (size_t)e, x + 3 * (size_t)col[e], t);
        y[0] += t[0]; y[1] += t[1]; y[2] += t[2];
    }
    group_reduce3<T>(y);
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
// work-group sums of two values stored to two partial arrays
__device__ inline void group_partials2(double v[2], double* __restrict__ pa, double* __restrict__ pb) {
    __shared__ double sh[2][kPB];
    sh[0][threadIdx.x] = v[0];
    sh[1][threadIdx.x] = v[1];
    __syncthreads();
    for (int s = kPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { sh[0][threadIdx.x] += sh[0][threadIdx.x + s]; sh[1][threadIdx.x] += sh[1][threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { pa[blockIdx.x] = sh[0][0]; pb[blockIdx.x] = sh[1][0]; }
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

// ---- work-group layout of the row kernels ----------------------------------------------
// One system: work-group b covers block rows [b kRows, b kRows + kRows) of [0, V).  A batch of
// independent systems (block-diagonal matrix, one block per scene): bl[b] = {scene, first
// row, end row}; the work-groups of a scene are consecutive and never cover two scenes, and
// gb[k] .. gb[k + 1] are scene k's work-groups, so that the partials of one scene are summed
// exactly as those of the same system solved alone.
struct Layout {
    int V = 0, ng = 0;
    const int4* bl = nullptr;
    const int* gb = nullptr;
    __device__ void block(int& scene, int& row0, int& rend) const {
        if (bl) {
            const int4 t = bl[blockIdx.x];
            scene = t.x; row0 = t.y; rend = t.z;
        } else {
            scene = 0; row0 = blockIdx.x * kRows; rend = V;
        }
    }
    __device__ void parts(int scene, int& lo, int& n) const {
        if (gb) { lo = gb[scene]; n = gb[scene + 1] - lo; }
        else { lo = 0; n = ng; }
    }
};

// ---- PCG kernels (T: storage precision of blocks and vectors; scalars FP64) ------------
// Partials layout (ng work-groups each): [0] p.q, [1] r.z, [2] r.r, [3] b.b
// r = b - A x (A x skipped when x is zero); partials of r.r and b.b
// (grid: kRows block rows per work-group, kGL lanes per row)
template <typename T>
__global__ __launch_bounds__(kPB) void pcg_residual(Layout L, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                    const T* __restrict__ val, const double* __restrict__ b, const T* __restrict__ x,
                                                    int x_zero, T* __restrict__ r, double* __restrict__ part_rr,
                                                    double* __restrict__ part_bb) {
    int scene, row0, rend;
    L.block(scene, row0, rend);
    const int a = row0 + threadIdx.x / kGL, sub = threadIdx.x % kGL;
    double acc[2] = {0, 0};
    if (a < rend) {
        double y[3] = {0, 0, 0};
        if (!x_zero) row_product<T, T>(a, sub, rowptr, col, val, x, y);
        if (sub == 0) {
            const double rv[3] = {b[3 * a] - y[0], b[3 * a + 1] - y[1], b[3 * a + 2] - y[2]};
            acc[1] = b[3 * a] * b[3 * a] + b[3 * a + 1] * b[3 * a + 1] + b[3 * a + 2] * b[3 * a + 2];
            for (int k = 0; k < 3; ++k) r[3 * a + k] = (T)rv[k];
            acc[0] = (double)(T)rv[0] * (T)rv[0] + (double)(T)rv[1] * (T)rv[1] + (double)(T)rv[2] * (T)rv[2];
        }
    }
    group_partials2(acc, part_rr, part_bb);
}
// block-Jacobi: z = M^-1 r; partial of r.z; p = z when `init`
template <typename T>
__global__ __launch_bounds__(kPB) void prec_jacobi(Layout L, const T* __restrict__ minv, const T* __restrict__ r, T* __restrict__ z,
                                                   T* __restrict__ p, int init, const PcgScalars* __restrict__ sc,
                                                   double* __restrict__ part_rz) {
    int scene, row0, rend;
    L.block(scene, row0, rend);
    if (sc && sc[scene].done) return;
    const int a = row0 + threadIdx.x / kGL, sub = threadIdx.x % kGL;
    double acc[1] = {0.0};
    if (a < rend && sub == 0) {
        const T rt[3] = {r[3 * a], r[3 * a + 1], r[3 * a + 2]};
        double zv[3];
        block_mv(minv + (size_t)9 * a, rt, zv);
        for (int k = 0; k < 3; ++k) {
            z[3 * a + k] = (T)zv[k];
            if (init) p[3 * a + k] = (T)zv[k];
        }
        acc[0] = (double)rt[0] * (T)zv[0] + (double)rt[1] * (T)zv[1] + (double)rt[2] * (T)zv[2];
    }
    group_partials<1>(acc, part_rz);
}
// factor preconditioner M^-1 = G^T G, G block lower-triangular stored at the lower entries of
// the BSR (the entry of (i, j), j <= i, holds G_ij):  t = G r (G interleaved)
template <typename T>
__global__ __launch_bounds__(kPB) void prec_g1(Layout L, const int* __restrict__ rowptr, const int* __restrict__ col, const T* __restrict__ G,
                                               const T* __restrict__ r, T* __restrict__ t, const PcgScalars* __restrict__ sc) {
    int scene, row0, rend;
    L.block(scene, row0, rend);
    if (sc && sc[scene].done) return;
    const int a = row0 + threadIdx.x / kGL, sub = threadIdx.x % kGL;
    if (a >= rend) return;
    double y[3] = {0, 0, 0};
    for (int e = rowptr[a] + sub; e < rowptr[a + 1]; e += kGL) {
        const int j = col[e];
        if (j > a) break;
        double w[3];
        block_mv_il(G, (size_t)e, r + 3 * (size_t)j, w);
        y[0] += w[0]; y[1] += w[1]; y[2] += w[2];
    }
    group_reduce3<T>(y);
    if (sub == 0)
        for (int k = 0; k < 3; ++k) t[3 * a + k] = (T)y[k];
}
// z = G^T t: row j gathers G_ij^T t_i over its upper entries (i >= j) through the mirror
// index of each entry (G in its plain block layout); partial of r.z; p = z when `init`
template <typename T>
__global__ __launch_bounds__(kPB) void prec_g2(Layout L, const int* __restrict__ rowptr, const int* __restrict__ col,
                                               const int* __restrict__ mirror, const T* __restrict__ G, const T* __restrict__ t,
                                               const T* __restrict__ r, T* __restrict__ z, T* __restrict__ p, int init,
                                               const PcgScalars* __restrict__ sc, double* __restrict__ part_rz) {
    int scene, row0, rend;
    L.block(scene, row0, rend);
    if (sc && sc[scene].done) return;
    const int a = row0 + threadIdx.x / kGL, sub = threadIdx.x % kGL;
    double acc[1] = {0.0};
    if (a < rend) {
        double y[3] = {0, 0, 0};
        for (int e = rowptr[a + 1] - 1 - sub; e >= rowptr[a]; e -= kGL) {
            const int i = col[e];
            if (i < a) break;
            const T* B = G + (size_t)9 * mirror[e];          // G_ia, a <= i
            const T* ti = t + 3 * (size_t)i;
            y[0] += (double)B[0] * ti[0] + (double)B[3] * ti[1] + (double)B[6] * ti[2];
            y[1] += (double)B[1] * ti[0] + (double)B[4] * ti[1] + (double)B[7] * ti[2];
            y[2] += (double)B[2] * ti[0] + (double)B[5] * ti[1] + (double)B[8] * ti[2];
        }
        group_reduce3<T>(y);
        if (sub == 0) {
            const T zt[3] = {(T)y[0], (T)y[1], (T)y[2]};
            for (int k = 0; k < 3; ++k) {
                z[3 * a + k] = zt[k];
                if (init) p[3 * a + k] = zt[k];
            }
            acc[0] = (double)r[3 * a] * zt[0] + (double)r[3 * a + 1] * zt[1] + (double)r[3 * a + 2] * zt[2];
        }
    }
    group_partials<1>(acc, part_rz);
}
// Per-scene controls of a batch solve (null for one system): relative tolerance, iteration
// cap and whether the scene takes part at all.
struct SceneControl {
    const double* rel_tol = nullptr;
    const int* max_iter = nullptr;
    const int* active = nullptr;
};
// scalars from the start partials: rz, rr, ||b|| (or ||r0|| as the reference of an inner
// solve); one work-group per scene
__global__ __launch_bounds__(kPB) void pcg_init_scalars(Layout L, const double* __restrict__ part_rz, const double* __restrict__ part_rr,
                                                        const double* __restrict__ part_bb, double rel_tol, int rel_to_r0,
                                                        SceneControl ctl, PcgScalars* __restrict__ sc) {
    const int k = blockIdx.x;
    int lo, n;
    L.parts(k, lo, n);
    const double rz = sum_partials_fixed(part_rz + lo, n);
    const double rr = sum_partials_fixed(part_rr + lo, n);
    const double bb = sum_partials_fixed(part_bb + lo, n);
    if (threadIdx.x == 0) {
        const double tol = ctl.rel_tol ? ctl.rel_tol[k] : rel_tol;
        PcgScalars& s = sc[k];
        s.rz = rz;
        s.rr = rr;
        s.bn = sqrt(bb);
        s.tol = tol * (rel_to_r0 ? sqrt(rr) : sqrt(bb));
        s.done = !(sqrt(rr) > s.tol) || !(rz > 0.0) || (ctl.active && !ctl.active[k]);
        s.iters = 0;
    }
}
// q = A p and partials of p.q
template <typename T>
__global__ __launch_bounds__(kPB) void pcg_spmv(Layout L, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                const T* __restrict__ val, const T* __restrict__ p, T* __restrict__ q,
                                                const PcgScalars* __restrict__ sc, double* __restrict__ part_pq) {
    int scene, row0, rend;
    L.block(scene, row0, rend);
    if (sc[scene].done) return;
    const int a = row0 + threadIdx.x / kGL, sub = threadIdx.x % kGL;
    double acc[1] = {0.0};
    if (a < rend) {
        double y[3];
        row_product<T, T>(a, sub, rowptr, col, val, p, y);
        if (sub == 0) {
            const T yt[3] = {(T)y[0], (T)y[1], (T)y[2]};
            for (int k = 0; k < 3; ++k) q[3 * a + k] = yt[k];
            acc[0] = (double)p[3 * a] * yt[0] + (double)p[3 * a + 1] * yt[1] + (double)p[3 * a + 2] * yt[2];
        }
    }
    group_partials<1>(acc, part_pq);
}
// p.q of every scene from the product's partials, once (one work-group per scene)
__global__ __launch_bounds__(kPB) void pcg_reduce_pq(Layout L, const double* __restrict__ part_pq, PcgScalars* __restrict__ sc) {
    const int k = blockIdx.x;
    if (sc[k].done) return;
    int lo, n;
    L.parts(k, lo, n);
    const double pq = sum_partials_fixed(part_pq + lo, n);
    if (threadIdx.x == 0) sc[k].pq = pq;
}
// alpha = rz / p.q; x += alpha p; r -= alpha q; partial of r.r; with block-Jacobi (minv set)
// also z = M^-1 r and its r.z partial
template <typename T>
__global__ __launch_bounds__(kPB) void pcg_update(Layout L, const T* __restrict__ minv,
                                                  const T* __restrict__ p, const T* __restrict__ q, T* __restrict__ x, T* __restrict__ r,
                                                  T* __restrict__ z, const PcgScalars* __restrict__ sc, double* __restrict__ part_rz,
                                                  double* __restrict__ part_rr) {
    int scene, row0, rend;
    L.block(scene, row0, rend);
    if (sc[scene].done) return;
    const double alpha = sc[scene].rz / sc[scene].pq;
    const int a = row0 + threadIdx.x / kGL, sub = threadIdx.x % kGL;
    double acc[2] = {0.0, 0.0};
    if (a < rend && sub == 0) {
        T rt[3];
        for (int k = 0; k < 3; ++k) {
            x[3 * a + k] = (T)((double)x[3 * a + k] + alpha * (double)p[3 * a + k]);
            rt[k] = (T)((double)r[3 * a + k] - alpha * (double)q[3 * a + k]);
            r[3 * a + k] = rt[k];
        }
        if (minv) {
            double zv[3];
            block_mv(minv + (size_t)9 * a, rt, zv);
            for (int k = 0; k < 3; ++k) z[3 * a + k] = (T)zv[k];
            acc[0] = (double)rt[0] * (T)zv[0] + (double)rt[1] * (T)zv[1] + (double)rt[2] * (T)zv[2];
        }
        acc[1] = (double)rt[0] * rt[0] + (double)rt[1] * rt[1] + (double)rt[2] * rt[2];
    }
    if (minv) group_partials2(acc, part_rz, part_rr);
    else { double a1[1] = {acc[1]}; group_partials<1>(a1, part_rr); }
}
// p = z + beta p with the beta of pcg_scalars (one DOF per thread); skipped by a scene that
// has just converged, whose direction is no longer used
template <typename T>
__global__ __launch_bounds__(kPB) void pcg_direction(int V, const T* __restrict__ z, T* __restrict__ p, const PcgScalars* __restrict__ sc) {
    if (sc->done) return;
    const double beta = sc->beta;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;       // one DOF per thread
    if (t < 3 * V) p[t] = (T)((double)z[t] + beta * (double)p[t]);
}
// the same on the row layout of a batch: the 3 kRows DOFs of the work-group's rows
template <typename T>
__global__ __launch_bounds__(kPB) void pcg_direction_rows(Layout L, const T* __restrict__ z, T* __restrict__ p,
                                                          const PcgScalars* __restrict__ sc) {
    int scene, row0, rend;
    L.block(scene, row0, rend);
    if (sc[scene].done) return;
    const double beta = sc[scene].beta;
    const int t = 3 * row0 + threadIdx.x;
    if ((int)threadIdx.x < 3 * kRows && t < 3 * rend) p[t] = (T)((double)z[t] + beta * (double)p[t]);
}
// r.z', r.r and beta = r.z' / r.z from the partials, once per iteration (one work-group per
// scene); the convergence test
__global__ void pcg_scalars(Layout L, const double* __restrict__ part_rz, const double* __restrict__ part_rr,
                            PcgScalars* __restrict__ sc, int max_iter, SceneControl ctl) {
    const int k = blockIdx.x;
    if (sc[k].done) return;
    int lo, n;
    L.parts(k, lo, n);
    const double rz1 = sum_partials_fixed(part_rz + lo, n);
    const double rr = sum_partials_fixed(part_rr + lo, n);
    if (threadIdx.x == 0) {
        PcgScalars& s = sc[k];
        const int cap = ctl.max_iter ? ctl.max_iter[k] : max_iter;
        s.beta = rz1 / s.rz;
        s.rz = rz1;
        s.rr = rr;
        s.iters += 1;
        if (!(sqrt(rr) > s.tol) || s.iters >= cap || !(rz1 > 0.0)) s.done = 1;
    }
}
// x (FP64) += d (FP32), for the refinement (scenes with add[scene] != 0 when add is set)
__global__ void add_correction(int n, const float* __restrict__ d, double* __restrict__ x, const int* __restrict__ vs,
                               const int* __restrict__ add) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (!add || add[vs[i / 3]])) x[i] += (double)d[i];
}
template <typename S, typename D>
__global__ void convert(size_t n, const S* __restrict__ s, D* __restrict__ d) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
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

// ---- block FSAI factor ----------------------------------------------------------------
// For block row i with lower pattern P (its entries j <= i, at most kFsaiMax blocks: the
// diagonal and the largest off-diagonal blocks by Frobenius norm), G_{i,P} is the last block
// row of L^-1, L the Cholesky factor of A[P,P]: then (G A G^T)_ii = I and G minimizes the
// Kaporin condition number over the pattern (Kolotilina & Yeremin's factorized sparse
// approximate inverse, in 3x3 blocks).  One wavefront per block row; A[P,P] in LDS.
constexpr int kFsaiMax = 24;                   // blocks per row pattern
constexpr int kFsaiN = 3 * kFsaiMax;           // dense order
constexpr int kFsaiLd = kFsaiN + 1;            // padded leading dimension

__device__ inline int find_col(const int* __restrict__ col, int lo, int hi, int c) {
    while (lo < hi) { const int m = (lo + hi) >> 1; if (col[m] < c) lo = m + 1; else hi = m; }
    return lo;
}

__global__ __launch_bounds__(64) void fsai_rows(int V, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                const double* __restrict__ val, double* __restrict__ G) {
    __shared__ double M[kFsaiN * kFsaiLd];
    __shared__ double Z[kFsaiN * 3];
    __shared__ int P[kFsaiMax], Pe[kFsaiMax];
    __shared__ int np_sh;
    __shared__ double red[64];
    const int lane = threadIdx.x;
    for (int i = blockIdx.x; i < V; i += gridDim.x) {
        const int e0 = rowptr[i], e1 = rowptr[i + 1];
        const int ediag = find_col(col, e0, e1, i);           // diagonal entry of row i (present)
        const int nl = ediag - e0;                              // off-diagonal lower candidates
        // choose the candidates: all, or the kFsaiMax-1 largest by Frobenius norm (ties by column)
        for (int e = e0 + lane; e < e1; e += 64) {              // clear this row's G entries
            for (int k = 0; k < 9; ++k) G[(size_t)9 * e + k] = 0.0;
        }
        __syncthreads();
        if (lane == 0) np_sh = 0;
        __syncthreads();
        for (int c = lane; c < nl; c += 64) {
            bool keep = true;
            if (nl > kFsaiMax - 1) {
                double nc = 0;
                for (int k = 0; k < 9; ++k) nc += val[(size_t)9 * (e0 + c) * 1 + k] * val[(size_t)9 * (e0 + c) + k];
                int rank = 0;
                for (int o = 0; o < nl; ++o) {
                    double no = 0;
                    for (int k = 0; k < 9; ++k) no += val[(size_t)9 * (e0 + o) + k] * val[(size_t)9 * (e0 + o) + k];
                    rank += (no > nc) || (no == nc && o < c);
                }
                keep = rank < kFsaiMax - 1;
            }
            if (keep) {
                const int slot = atomicAdd(&np_sh, 1);
                Pe[slot] = e0 + c;
            }
        }
        __syncthreads();
        const int k = np_sh + 1;
        if (lane == 0) {
            // sort the kept entries by column (insertion sort, k <= kFsaiMax), diagonal last
            for (int a = 1; a < k - 1; ++a) {
                const int v = Pe[a];
                int b = a - 1;
                while (b >= 0 && Pe[b] > v) { Pe[b + 1] = Pe[b]; --b; }
                Pe[b + 1] = v;
            }
            Pe[k - 1] = ediag;
            for (int a = 0; a < k; ++a) P[a] = col[Pe[a]];
        }
        __syncthreads();
        const int n = 3 * k;
        // M = A[P,P] (lower blocks looked up in the block rows of P; upper by symmetry)
        for (int ab = lane; ab < k * k; ab += 64) {
            const int a = ab / k, b = ab % k;
            if (b > a) continue;
            const int ra = P[a];
            const int rs = rowptr[ra], re = rowptr[ra + 1];
            const int f = find_col(col, rs, re, P[b]);
            const bool has = f < re && col[f] == P[b];
            for (int u = 0; u < 3; ++u)
                for (int v = 0; v < 3; ++v) {
                    const double m = has ? val[(size_t)9 * f + 3 * u + v] : 0.0;   // A_{P_a P_b}(u, v)
                    M[(3 * a + u) * kFsaiLd + 3 * b + v] = m;
                    M[(3 * b + v) * kFsaiLd + 3 * a + u] = m;
                }
        }
        __syncthreads();
        // Cholesky M = L L^T in place (lower triangle)
        for (int c = 0; c < n; ++c) {
            if (lane == 0) M[c * kFsaiLd + c] = sqrt(fmax(M[c * kFsaiLd + c], 1e-300));
            __syncthreads();
            const double dc = M[c * kFsaiLd + c];
            for (int r = c + 1 + lane; r < n; r += 64) M[r * kFsaiLd + c] /= dc;
            __syncthreads();
            const int m = n - c - 1;
            for (int idx = lane; idx < m * m; idx += 64) {
                const int r = c + 1 + idx / m, q = c + 1 + idx % m;
                if (q <= r) M[r * kFsaiLd + q] -= M[r * kFsaiLd + c] * M[q * kFsaiLd + c];
            }
            __syncthreads();
        }
        // L^T Z = E (last three unit columns), bottom-up; lanes split the dot products
        for (int r = n - 1; r >= 0; --r) {
            for (int qq = 0; qq < 3; ++qq) {
                double acc = 0.0;
                for (int s2 = r + 1 + lane; s2 < n; s2 += 64) acc += M[s2 * kFsaiLd + r] * Z[s2 * 3 + qq];
                red[lane] = acc;
                __syncthreads();
                for (int w = 32; w > 0; w >>= 1) {
                    if (lane < w) red[lane] += red[lane + w];
                    __syncthreads();
                }
                if (lane == 0) {
                    const double rhs = (r == n - 3 + qq) ? 1.0 : 0.0;
                    Z[r * 3 + qq] = (rhs - red[0]) / M[r * kFsaiLd + r];
                }
                __syncthreads();
            }
        }
        // G_{i, P_a} (row u, column v) = Z[3a + v][u]
        for (int idx = lane; idx < 9 * k; idx += 64) {
            const int a = idx / 9, u = (idx % 9) / 3, v = idx % 3;
            G[(size_t)9 * Pe[a] + 3 * u + v] = Z[(3 * a + v) * 3 + u];
        }
        __syncthreads();
    }
}
// mirror[e] = index of the entry (col(e), row(e))
__global__ void bsr_mirror(int V, const int* __restrict__ rowptr, const int* __restrict__ col, int* __restrict__ mirror) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= V) return;
    for (int e = rowptr[a]; e < rowptr[a + 1]; ++e) {
        const int j = col[e];
        mirror[e] = find_col(col, rowptr[j], rowptr[j + 1], a);
    }
}

// ---- host ---------------------------------------------------------------------------
namespace {
void read_scalars(Context& c, const PcgScalars* d, int S, std::vector<PcgScalars>& h) {
    h.resize(S);
    read_back(c, h.data(), d, (size_t)S);
}
bool all_done(const std::vector<PcgScalars>& h) {
    for (const PcgScalars& s : h)
        if (!s.done) return false;
    return true;
}

// The preconditioner of one solve: block-Jacobi (minv) or the factor G (with mirror).
template <typename T>
struct Prec {
    const T* minv = nullptr;
    const T* G = nullptr;      // plain blocks (transposed products)
    const T* Gi = nullptr;     // interleaved (row products)
    const int* mirror = nullptr;
};

// Work-group layout of the current system: kRows block rows per work-group; for a batch
// the tables of per-scene work-groups (cached per vertex partition).
Layout make_layout(Context& c) {
    Scene& s = c.sc;
    SolveBuffers& w = c.solve;
    Layout L;
    L.V = s.V;
    if (s.S == 1) {
        L.ng = (int)blocks_for(s.V, kRows);
        return L;
    }
    if (w.layout_key != s.h_voff) {
        std::vector<int4> bl;
        std::vector<int> gb(s.S + 1, 0);
        for (int k = 0; k < s.S; ++k) {
            gb[k] = (int)bl.size();
            for (int r = s.h_voff[k]; r < s.h_voff[k + 1]; r += kRows) bl.push_back(make_int4(k, r, s.h_voff[k + 1], 0));
        }
        gb[s.S] = (int)bl.size();
        w.bl.resize(std::max<size_t>(bl.size(), 1));
        w.gb.resize(gb.size());
        CIPC_HIP(hipMemcpyAsync(w.bl.p, bl.data(), bl.size() * sizeof(int4), hipMemcpyHostToDevice, c.stream));
        CIPC_HIP(hipMemcpyAsync(w.gb.p, gb.data(), gb.size() * sizeof(int), hipMemcpyHostToDevice, c.stream));
        CIPC_HIP(hipStreamSynchronize(c.stream));
        w.layout_key = s.h_voff;
        w.layout_ng = (int)bl.size();
    }
    L.ng = w.layout_ng;
    L.bl = w.bl.p;
    L.gb = w.gb.p;
    return L;
}

template <typename T>
void apply_prec(Context& c, const Layout& L, const Prec<T>& M, const T* r, T* z, T* p, int init, const PcgScalars* sc, double* part_rz) {
    SolveBuffers& w = c.solve;
    if (M.G) {
        T* t = reinterpret_cast<T*>(w.t.p);
        hipLaunchKernelGGL(prec_g1<T>, dim3(L.ng), dim3(kPB), 0, c.stream, L, w.rowptr.p, w.col.p, M.Gi, r, t, sc);
        hipLaunchKernelGGL(prec_g2<T>, dim3(L.ng), dim3(kPB), 0, c.stream, L, w.rowptr.p, w.col.p, M.mirror, M.G, t, r, z, p, init, sc,
                           part_rz);
    } else {
        hipLaunchKernelGGL(prec_jacobi<T>, dim3(L.ng), dim3(kPB), 0, c.stream, L, M.minv, r, z, p, init, sc, part_rz);
    }
}

// PCG solves of A x = b (b FP64; A, x and the work vectors in T), one per scene in lock step.
// kCheck iterations are captured once into a HIP graph and replayed until the device flags
// say that every scene has converged.  Returns the scenes' final scalars.
template <typename T>
void pcg_run(Context& c, const Layout& L, int S, const T* val, const Prec<T>& M, const double* b, T* x, int x_zero, double rel_tol,
             int rel_to_r0, int max_iter, SceneControl ctl, std::vector<PcgScalars>& h) {
    const int V = c.sc.V;
    const int ng = L.ng;
    const int ngd = (int)blocks_for(3 * V, kPB);
    SolveBuffers& w = c.solve;
    T* r = reinterpret_cast<T*>(w.r.p);
    T* z = reinterpret_cast<T*>(w.z.p);
    T* p = reinterpret_cast<T*>(w.p.p);
    T* q = reinterpret_cast<T*>(w.q.p);
    double* part_pq = w.part.p;
    double* part_rz = w.part.p + ng;
    double* part_rr = w.part.p + 2 * ng;
    double* part_bb = w.part.p + 3 * ng;
    PcgScalars* sc = reinterpret_cast<PcgScalars*>(w.scal.p);
    hipLaunchKernelGGL(pcg_residual<T>, dim3(ng), dim3(kPB), 0, c.stream, L, w.rowptr.p, w.col.p, val, b, x, x_zero, r, part_rr, part_bb);
    apply_prec<T>(c, L, M, r, z, p, 1, nullptr, part_rz);
    hipLaunchKernelGGL(pcg_init_scalars, dim3(S), dim3(kPB), 0, c.stream, L, part_rz, part_rr, part_bb, rel_tol, rel_to_r0, ctl, sc);
    CIPC_CHECK_LAUNCH();
    read_scalars(c, sc, S, h);
    if (all_done(h)) return;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    CIPC_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < kCheck; ++k) {
        hipLaunchKernelGGL(pcg_spmv<T>, dim3(ng), dim3(kPB), 0, c.stream, L, w.rowptr.p, w.col.p, val, p, q, sc, part_pq);
        hipLaunchKernelGGL(pcg_reduce_pq, dim3(S), dim3(kPB), 0, c.stream, L, part_pq, sc);
        hipLaunchKernelGGL(pcg_update<T>, dim3(ng), dim3(kPB), 0, c.stream, L, M.G ? nullptr : M.minv, p, q, x, r, z, sc,
                           part_rz, part_rr);
        if (M.G) apply_prec<T>(c, L, M, r, z, p, 0, sc, part_rz);
        hipLaunchKernelGGL(pcg_scalars, dim3(S), dim3(kPB), 0, c.stream, L, part_rz, part_rr, sc, max_iter, ctl);
        if (L.bl) hipLaunchKernelGGL(pcg_direction_rows<T>, dim3(ng), dim3(kPB), 0, c.stream, L, z, p, sc);
        else hipLaunchKernelGGL(pcg_direction<T>, dim3(ngd), dim3(kPB), 0, c.stream, V, z, p, sc);
    }
    const hipError_t ce = hipStreamEndCapture(c.stream, &graph);
    CIPC_HIP(ce);
    CIPC_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    try {
        while (!all_done(h)) {
            CIPC_HIP(hipGraphLaunch(exec, c.stream));
            read_scalars(c, sc, S, h);
        }
    } catch (...) {
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        throw;
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
}

// Per-scene state of the FP64 refinement.
struct Refine {
    int iters = 0, outer = 0;
    bool finished = false;
    double rr = 0.0, bn = 0.0;
};
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
        hipLaunchKernelGGL((convert<double, float>), dim3(blocks_for(9 * ne, 256)), dim3(256), 0, c.stream, 9 * ne, w.val.p, w.valf.p);
        hipLaunchKernelGGL(diag_inverse<float>, dim3(blocks_for(V, 256)), dim3(256), 0, c.stream, V, w.rowptr.p, w.col.p, w.valf.p,
                           w.minvf.p);
        CIPC_CHECK_LAUNCH();
    }
    w.ne = ne;
}

// The block FSAI factor of the current BSR (and its FP32 copy for the mixed form).
void build_fsai(Context& c, bool fp32) {
    SolveBuffers& w = c.solve;
    const int V = c.sc.V;
    const size_t ne = w.ne;
    w.G.resize(9 * ne);
    w.mirror.resize(ne);
    hipLaunchKernelGGL(bsr_mirror, dim3(blocks_for(V, 256)), dim3(256), 0, c.stream, V, w.rowptr.p, w.col.p, w.mirror.p);
    const int grid = std::min(V, 256 * 16);
    hipLaunchKernelGGL(fsai_rows, dim3(grid), dim3(64), 0, c.stream, V, w.rowptr.p, w.col.p, w.val.p, w.G.p);
    CIPC_CHECK_LAUNCH();
    if (fp32) {
        w.Gf.resize(9 * ne);
        hipLaunchKernelGGL((convert<double, float>), dim3(blocks_for(9 * ne, 256)), dim3(256), 0, c.stream, 9 * ne, w.G.p, w.Gf.p);
        CIPC_CHECK_LAUNCH();
    }
}

// One solve per scene of the last assembly (one scene: the plain solve).  Each scene runs
// the same iteration as it would alone: its own scalars, convergence test and refinement.
void linear_solve_scenes(Context& c, int precision, int preconditioner, double rel_tol, int max_iter, const int* active,
                         SolveResult* res) {
    Scene& s = c.sc;
    const int V = s.V, n = 3 * V, S = s.S;
    SolveBuffers& w = c.solve;
    const Layout L = make_layout(c);
    const int ng = L.ng;
    w.r.resize((size_t)n); w.z.resize((size_t)n); w.p.resize((size_t)n); w.q.resize((size_t)n); w.t.resize((size_t)n);
    w.part.resize((size_t)4 * ng);
    w.scal.resize((size_t)S * ((sizeof(PcgScalars) + 7) / 8));
    w.x.resize((size_t)n);
    build_solve_matrix(c, precision == 1);
    w.factor = 0;
    if (preconditioner == 1) build_fsai(c, precision == 1);
    else if (preconditioner == 2) build_learned(c, precision == 1);
    w.factor = preconditioner;
    // interleaved copies for the row products
    const size_t ne = w.ne, nil = (ne + kGL - 1) / kGL * 9 * kGL;
    const bool fp32 = precision == 1;
    w.vil.resize(nil);           // FP64: the solve, or the refinement's residuals
    hipLaunchKernelGGL(interleave_blocks<double>, dim3(blocks_for(9 * ne, 256)), dim3(256), 0, c.stream, ne, w.val.p, w.vil.p);
    if (fp32) {
        w.vilf.resize(nil);
        hipLaunchKernelGGL(interleave_blocks<float>, dim3(blocks_for(9 * ne, 256)), dim3(256), 0, c.stream, ne, w.valf.p, w.vilf.p);
    }
    if (preconditioner != 0) {
        if (fp32) {
            w.Gilf.resize(nil);
            hipLaunchKernelGGL(interleave_blocks<float>, dim3(blocks_for(9 * ne, 256)), dim3(256), 0, c.stream, ne, w.Gf.p, w.Gilf.p);
        } else {
            w.Gil.resize(nil);
            hipLaunchKernelGGL(interleave_blocks<double>, dim3(blocks_for(9 * ne, 256)), dim3(256), 0, c.stream, ne, w.G.p, w.Gil.p);
        }
    }
    CIPC_CHECK_LAUNCH();
    Prec<double> Md;
    Prec<float> Mf;
    if (preconditioner == 0) { Md.minv = w.minv.p; Mf.minv = w.minvf.p; }
    else {
        Md.G = w.G.p; Md.mirror = w.mirror.p; Mf.G = w.Gf.p; Mf.mirror = w.mirror.p;
        Md.Gi = fp32 ? nullptr : w.Gil.p;
        Mf.Gi = fp32 ? w.Gilf.p : nullptr;
    }
    // per-scene controls (a batch): participation, inner tolerances and caps, correction mask
    SceneControl ctl;
    std::vector<int> act(S, 1);
    if (active)
        for (int k = 0; k < S; ++k) act[k] = active[k] ? 1 : 0;
    if (S > 1) {
        w.ctl_d.resize((size_t)S);
        w.ctl_i.resize(3 * (size_t)S);
        CIPC_HIP(hipMemcpyAsync(w.ctl_i.p, act.data(), S * sizeof(int), hipMemcpyHostToDevice, c.stream));
        ctl.active = w.ctl_i.p;
    }
    for (int k = 0; k < S; ++k) res[k] = SolveResult{};
    std::vector<PcgScalars> h;
    hipLaunchKernelGGL(zero_d, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, w.x.p);
    if (precision == 0) {
        pcg_run<double>(c, L, S, w.vil.p, Md, c.rhs.p, w.x.p, 1, rel_tol, 0, max_iter, ctl, h);
        for (int k = 0; k < S; ++k) {
            res[k].iterations = h[k].iters;
            res[k].rel_residual = h[k].bn > 0 ? std::sqrt(h[k].rr) / h[k].bn : 0.0;
        }
    } else {
        // FP64 refinement around FP32 PCG; each inner solve aims just past the remaining
        // reduction (floored at the FP32 reach), so one or two refinements usually suffice
        w.xf.resize((size_t)n); w.res.resize((size_t)n);
        double* part_rr = w.part.p + 2 * ng;
        double* part_bb = w.part.p + 3 * ng;
        PcgScalars* sc = reinterpret_cast<PcgScalars*>(w.scal.p);
        std::vector<Refine> st(S);
        for (int k = 0; k < S; ++k) st[k].finished = !act[k];
        std::vector<double> itol(S, 0.0);
        std::vector<int> icap(S, 0), need_inner(S, 0), add(S, 0);   // host sources of async copies: kept alive
        for (int round = 0;; ++round) {
            double* rd = w.res.p;
            hipLaunchKernelGGL(pcg_residual<double>, dim3(ng), dim3(kPB), 0, c.stream, L, w.rowptr.p, w.col.p, w.vil.p, c.rhs.p, w.x.p,
                               round == 0 ? 1 : 0, rd, part_rr, part_bb);
            hipLaunchKernelGGL(pcg_init_scalars, dim3(S), dim3(kPB), 0, c.stream, L, part_rr, part_rr, part_bb, rel_tol, 0, SceneControl{},
                               sc);
            CIPC_CHECK_LAUNCH();
            read_scalars(c, sc, S, h);
            bool any = false;
            for (int k = 0; k < S; ++k) {
                need_inner[k] = 0;
                if (st[k].finished) continue;
                st[k].rr = h[k].rr; st[k].bn = h[k].bn;
                if (!(std::sqrt(h[k].rr) > rel_tol * h[k].bn) || st[k].iters >= max_iter) { st[k].finished = true; continue; }
                const double need = rel_tol * h[k].bn / std::sqrt(h[k].rr);      // reduction still needed
                itol[k] = std::max(0.3 * need, 2e-6);
                icap[k] = max_iter - st[k].iters;
                need_inner[k] = 1;
                any = true;
            }
            if (!any) break;
            SceneControl ictl;
            if (S > 1) {
                CIPC_HIP(hipMemcpyAsync(w.ctl_d.p, itol.data(), S * sizeof(double), hipMemcpyHostToDevice, c.stream));
                CIPC_HIP(hipMemcpyAsync(w.ctl_i.p + S, icap.data(), S * sizeof(int), hipMemcpyHostToDevice, c.stream));
                CIPC_HIP(hipMemcpyAsync(w.ctl_i.p, need_inner.data(), S * sizeof(int), hipMemcpyHostToDevice, c.stream));
                ictl.rel_tol = w.ctl_d.p;
                ictl.max_iter = w.ctl_i.p + S;
                ictl.active = w.ctl_i.p;
            }
            hipLaunchKernelGGL(zero_f, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, w.xf.p);
            std::vector<PcgScalars> hi;
            pcg_run<float>(c, L, S, w.vilf.p, Mf, rd, w.xf.p, 1, itol[0], 1, icap[0], ictl, hi);
            // scenes that made no progress (breakdown) or hit the refinement cap stop here,
            // without their correction
            std::fill(add.begin(), add.end(), 0);
            for (int k = 0; k < S; ++k) {
                if (!need_inner[k]) continue;
                const int it = hi[k].iters;
                if (it == 0 || st[k].outer >= kMaxRefine) { st[k].finished = true; continue; }
                st[k].iters += it;
                ++st[k].outer;
                add[k] = 1;
            }
            if (S > 1) {
                CIPC_HIP(hipMemcpyAsync(w.ctl_i.p + 2 * S, add.data(), S * sizeof(int), hipMemcpyHostToDevice, c.stream));
                hipLaunchKernelGGL(add_correction, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, w.xf.p, w.x.p, s.vscene.p,
                                   w.ctl_i.p + 2 * S);
            } else if (add[0]) {
                hipLaunchKernelGGL(add_correction, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, w.xf.p, w.x.p, nullptr, nullptr);
            }
            CIPC_CHECK_LAUNCH();
        }
        for (int k = 0; k < S; ++k) {
            res[k].iterations = st[k].iters;
            res[k].refinements = st[k].outer;
            res[k].rel_residual = st[k].bn > 0 ? std::sqrt(st[k].rr) / st[k].bn : 0.0;
        }
    }
    // the Newton direction for the step bound and line search
    s.dir.resize((size_t)n);
    CIPC_HIP(hipMemcpyAsync(s.dir.p, w.x.p, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    CIPC_HIP(hipStreamSynchronize(c.stream));
}

// ---- direct tier ------------------------------------------------------------------------
// The host direct solvers (PARDISO, CHOLMOD; NewtonLinearSystem.cpp:150-174) on the GPU: the
// assembled lower CSC scattered into a dense column-major matrix, factored by a dense
// Cholesky (rocSOLVER potrf) and solved (potrs); the FP64 residual of the solution is
// evaluated on the BSR as for PCG.  For systems up to kDirectMaxDofs unknowns.
__global__ void csc_lower_to_dense(int n, const int* __restrict__ outer, const int* __restrict__ inner, const double* __restrict__ val,
                                   double* __restrict__ A) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n) return;
    for (int k = outer[col]; k < outer[col + 1]; ++k) A[(size_t)col * n + inner[k]] = val[k];
}
__global__ void zero_dense(size_t n, double* __restrict__ A) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) A[i] = 0.0;
}

void direct_solve(Context& c, SolveResult& res) {
    Scene& s = c.sc;
    if (s.S != 1) throw Error(CIPC_ERR_LOGIC, "a batch is solved per scene");
    const int V = s.V, n = 3 * V;
    if (n > kDirectMaxDofs)
        throw Error(CIPC_ERR_INVALID_ARGUMENT, "the direct solve takes at most " + std::to_string(kDirectMaxDofs) + " unknowns, the system has " +
                                                   std::to_string(n));
    SolveBuffers& w = c.solve;
    w.dense.resize((size_t)n * n);
    w.x.resize((size_t)n);
    w.info.resize(1);
    if (!c.blas) {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success) throw Error(CIPC_ERR_HIP, "rocblas_create_handle failed");
        c.blas = h;
    }
    rocblas_handle h = static_cast<rocblas_handle>(c.blas);
    if (rocblas_set_stream(h, c.stream) != rocblas_status_success) throw Error(CIPC_ERR_HIP, "rocblas_set_stream failed");
    hipLaunchKernelGGL(zero_dense, dim3(4096), dim3(256), 0, c.stream, (size_t)n * n, w.dense.p);
    hipLaunchKernelGGL(csc_lower_to_dense, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, c.outer.p, c.inner.p, c.val.p, w.dense.p);
    CIPC_HIP(hipMemcpyAsync(w.x.p, c.rhs.p, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    CIPC_CHECK_LAUNCH();
    if (rocsolver_dpotrf(h, rocblas_fill_lower, n, w.dense.p, n, w.info.p) != rocblas_status_success)
        throw Error(CIPC_ERR_HIP, "rocsolver_dpotrf failed");
    int info = 0;
    read_back(c, &info, w.info.p, 1);
    if (info != 0)
        throw Error(CIPC_ERR_RUNTIME, "the Newton system is not positive definite (Cholesky pivot " + std::to_string(info) + ")");
    if (rocsolver_dpotrs(h, rocblas_fill_lower, n, 1, w.dense.p, n, w.x.p, n) != rocblas_status_success)
        throw Error(CIPC_ERR_HIP, "rocsolver_dpotrs failed");
    // FP64 residual on the BSR of the same system
    build_solve_matrix(c, false);
    const Layout L = make_layout(c);
    const int ng = L.ng;
    const size_t ne = w.ne, nil = (ne + kGL - 1) / kGL * 9 * kGL;
    w.vil.resize(nil);
    w.r.resize((size_t)n);
    w.part.resize((size_t)4 * ng);
    w.scal.resize((sizeof(PcgScalars) + 7) / 8);
    hipLaunchKernelGGL(interleave_blocks<double>, dim3(blocks_for(9 * ne, 256)), dim3(256), 0, c.stream, ne, w.val.p, w.vil.p);
    double* part_rr = w.part.p + 2 * ng;
    double* part_bb = w.part.p + 3 * ng;
    PcgScalars* sc = reinterpret_cast<PcgScalars*>(w.scal.p);
    hipLaunchKernelGGL(pcg_residual<double>, dim3(ng), dim3(kPB), 0, c.stream, L, w.rowptr.p, w.col.p, w.vil.p, c.rhs.p, w.x.p, 0, w.r.p,
                       part_rr, part_bb);
    hipLaunchKernelGGL(pcg_init_scalars, dim3(1), dim3(kPB), 0, c.stream, L, part_rr, part_rr, part_bb, 1.0, 0, SceneControl{}, sc);
    CIPC_CHECK_LAUNCH();
    std::vector<PcgScalars> hs;
    read_scalars(c, sc, 1, hs);
    res.iterations = 0;
    res.refinements = 0;
    res.rel_residual = hs[0].bn > 0 ? std::sqrt(hs[0].rr) / hs[0].bn : 0.0;
    s.dir.resize((size_t)n);
    CIPC_HIP(hipMemcpyAsync(s.dir.p, w.x.p, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    CIPC_HIP(hipStreamSynchronize(c.stream));
}

void destroy_direct(Context& c) {
    if (c.blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(c.blas));
    c.blas = nullptr;
}

void linear_solve(Context& c, int precision, int preconditioner, double rel_tol, int max_iter, SolveResult& res) {
    if (c.sc.S != 1) throw Error(CIPC_ERR_LOGIC, "a batch is solved per scene");
    if (preconditioner == kDirectSolve) {
        direct_solve(c, res);
        return;
    }
    linear_solve_scenes(c, precision, preconditioner, rel_tol, max_iter, nullptr, &res);
}

}  // namespace cipc
