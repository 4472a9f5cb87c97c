// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate expressions of CPU-IPC
// (MPL-2.0) in the reference's operation order; each such part cites the source it follows.
// Contact geometry on the device: distance-type classification, squared distances,
// barrier, mollifier, PSD projection and additive CCD.  Included only by translation
// units compiled with -ffp-contract=off so that classification, distances, boxes and
// CCD round exactly as the reference expressions are written (bit-exact pair sets).
#pragma once
#include "common.hpp"

namespace cipc {

#include "distance_gen.inc"

// ---- barrier b(d) = -(d - dHat)^2 ln(d/dHat) and derivatives (ContactMechanics.cpp:775-790)
__device__ inline double barrier_b(double d, double dh) { return -(d - dh) * (d - dh) * log(d / dh); }
__device__ inline double barrier_g(double d, double dh) {
    const double t = d - dh;
    return t * log(d / dh) * -2.0 - (t * t) / d;
}
__device__ inline double barrier_h(double d, double dh) {
    const double t = d - dh;
    return (log(d / dh) * -2.0 - t * 4.0 / d) + 1.0 / (d * d) * (t * t);
}

// ---- squared distances in the reference's operation order
__device__ inline double d_PP(d3 a, d3 b) { return sqn(a - b); }                                 // :282-287
__device__ inline double d_PE(d3 p, d3 e0, d3 e1) { return sqn(cross(e0 - p, e1 - p)) / sqn(e1 - e0); } // :807-813
__device__ inline double d_PT(d3 p, d3 t0, d3 t1, d3 t2) {                                       // :298-306
    const d3 b = cross(t1 - t0, t2 - t0);
    const double aTb = dot(p - t0, b);
    return aTb * aTb / sqn(b);
}
__device__ inline double d_EE(d3 a0, d3 a1, d3 b0, d3 b1) {                                      // :764-773
    const d3 b = cross(a1 - a0, b1 - b0);
    const double aTb = dot(b0 - a0, b);
    return aTb * aTb / sqn(b);
}
__device__ inline double ee_cross_sqn(d3 a0, d3 a1, d3 b0, d3 b1) { return sqn(cross(a1 - a0, b1 - b0)); }

// Eigen 3.4 LDLT solve of a symmetric 2x2 [[a00, a10], [a10, a11]] (diagonal pivoting on
// the larger |diagonal|, unit lower L, D pseudo-inverse below DBL_MIN), as
// (basis * basis^T).ldlt().solve(...) evaluates it in dType_PT.
__device__ inline void ldlt2_solve(double a00, double a10, double a11, double b0, double b1, double& x0, double& x1) {
    bool swapped = fabs(a11) > fabs(a00);
    double d0 = swapped ? a11 : a00, d1 = swapped ? a00 : a11, l = a10;
    if (!(fabs(d0) > 0.0)) {
        swapped = false; d0 = a00; d1 = a11;
    } else {
        l = l / d0;
        const double tmp = d0 * l;
        d1 = d1 - l * tmp;
    }
    double y0 = swapped ? b1 : b0, y1 = swapped ? b0 : b1;
    y1 = y1 - l * y0;
    const double tol = 2.2250738585072014e-308;
    y0 = fabs(d0) > tol ? y0 / d0 : 0.0;
    y1 = fabs(d1) > tol ? y1 / d1 : 0.0;
    y0 = y0 - l * y1;
    if (swapped) { x0 = y1; x1 = y0; } else { x0 = y0; x1 = y1; }
}
__device__ inline void pt_param(d3 r0, d3 n, d3 rhs, double& p0, double& p1) {
    const d3 r1 = cross(r0, n);
    ldlt2_solve(dot(r0, r0), dot(r1, r0), dot(r1, r1), dot(r0, rhs), dot(r1, rhs), p0, p1);
}
// dType_PT (ContactMechanics.cpp:143-193) with (v0 | v1 v2 v3); with the triangle
// passed as (t0 t1 t2) it is also point_triangle_distance_type (AdditiveCCD.cpp:8-62).
__device__ inline int dtype_pt(d3 v0, d3 v1, d3 v2, d3 v3) {
    const d3 n = cross(v2 - v1, v3 - v1);
    double p00, p10, p01, p11, p02, p12;
    pt_param(v2 - v1, n, v0 - v1, p00, p10);
    if (p00 > 0.0 && p00 < 1.0 && p10 >= 0.0) return 3;
    pt_param(v3 - v2, n, v0 - v2, p01, p11);
    if (p01 > 0.0 && p01 < 1.0 && p11 >= 0.0) return 4;
    pt_param(v1 - v3, n, v0 - v3, p02, p12);
    if (p02 > 0.0 && p02 < 1.0 && p12 >= 0.0) return 5;
    if (p00 <= 0.0 && p02 >= 1.0) return 0;
    if (p01 <= 0.0 && p00 >= 1.0) return 1;
    if (p02 <= 0.0 && p01 >= 1.0) return 2;
    return 6;
}
// dType_EE (ContactMechanics.cpp:195-280) == edge_edge_distance_type (AdditiveCCD.cpp:64-139)
__device__ inline int dtype_ee(d3 v0, d3 v1, d3 v2, d3 v3) {
    const d3 u = v1 - v0, v = v3 - v2, w = v0 - v2;
    const double a = sqn(u), b = dot(u, v), c = sqn(v), d = dot(u, w), e = dot(v, w);
    const double D = a * c - b * b;
    double tD = D, sN, tN;
    int def = 8;
    sN = (b * e - c * d);
    if (sN <= 0.0) { tN = e; tD = c; def = 2; }
    else if (sN >= D) { tN = e + b; tD = c; def = 5; }
    else {
        tN = (a * e - b * d);
        const d3 uv = cross(u, v);
        if (tN > 0.0 && tN < tD && (dot(uv, w) == 0.0 || sqn(uv) < 1.0e-20 * a * c)) {
            if (sN < D / 2) { tN = e; tD = c; def = 2; }
            else { tN = e + b; tD = c; def = 5; }
        }
    }
    if (tN <= 0.0) {
        if (-d <= 0.0) return 0;
        else if (-d >= a) return 3;
        else return 6;
    } else if (tN >= tD) {
        if ((-d + b) <= 0.0) return 1;
        else if ((-d + b) >= a) return 4;
        else return 7;
    }
    return def;
}

// ---- mollifier (ContactMechanics.cpp:678-762)
__device__ inline double eps_x_of(d3 ra0, d3 ra1, d3 rb0, d3 rb1) { return 1e-3 * sqn(ra0 - ra1) * sqn(rb0 - rb1); }
__device__ inline double mol_e(double c, double ex) {
    if (!(c < ex)) return 1.0;
    const double t = c / ex;
    return (-t + 2.0) * t;
}

// ---- PSD projection IglUtils::makePD<N> (ContactMechanics.h:27-45): cyclic Jacobi on the
// upper triangle held in registers; the block is rebuilt from the eigenpairs with the
// negative eigenvalues set to zero (when none is negative this reproduces the input up
// to rounding, which is what "return unchanged" means at the 1e-10 parity bar).
template <int N>
struct Sym {
    static constexpr int ix(int i, int j) { return i <= j ? i * N - i * (i - 1) / 2 + (j - i) : j * N - j * (j - 1) / 2 + (i - j); }
    static constexpr int row(int k) { int i = 0; while (k >= N - i) { k -= N - i; ++i; } return i; }
    static constexpr int col(int k) { int i = 0; while (k >= N - i) { k -= N - i; ++i; } return i + k; }
};
// In place on the packed upper triangle A (row-major, N(N+1)/2 entries).
// Z is the eigenvector matrix, row-major N x N, reached through zp(k) (registers, or a
// strided LDS column of the calling thread).
template <int N, typename ZP>
__device__ inline void make_pd_upper_z(double* __restrict__ A, ZP zp) {
#define Z(k) (*zp(k))
#pragma unroll
    for (int i = 0; i < N * N; ++i) Z(i) = (i % (N + 1) == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = 0.0, dg = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            dg += A[Sym<N>::ix(i, i)] * A[Sym<N>::ix(i, i)];
#pragma unroll
            for (int j = i + 1; j < N; ++j) off += A[Sym<N>::ix(i, j)] * A[Sym<N>::ix(i, j)];
        }
        if (off == 0.0 || off <= 1e-30 * dg) break;
#pragma unroll
        for (int p = 0; p < N; ++p) {
#pragma unroll
            for (int q = p + 1; q < N; ++q) {
                const double apq = A[Sym<N>::ix(p, q)];
                if (apq != 0.0) {
                    const double app = A[Sym<N>::ix(p, p)], aqq = A[Sym<N>::ix(q, q)];
                    const double th = (aqq - app) / (2.0 * apq);
                    const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                    const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
                    for (int k = 0; k < N; ++k) {
                        if (k == p || k == q) continue;
                        const double akp = A[Sym<N>::ix(k, p)], akq = A[Sym<N>::ix(k, q)];
                        A[Sym<N>::ix(k, p)] = c * akp - s * akq;
                        A[Sym<N>::ix(k, q)] = s * akp + c * akq;
                    }
                    A[Sym<N>::ix(p, p)] = app - t * apq;
                    A[Sym<N>::ix(q, q)] = aqq + t * apq;
                    A[Sym<N>::ix(p, q)] = 0.0;
#pragma unroll
                    for (int k = 0; k < N; ++k) {
                        const double zkp = Z(k * N + p), zkq = Z(k * N + q);
                        Z(k * N + p) = c * zkp - s * zkq;
                        Z(k * N + q) = s * zkp + c * zkq;
                    }
                }

            }
        }
    }
    double w[N];
#pragma unroll
    for (int i = 0; i < N; ++i) w[i] = fmax(A[Sym<N>::ix(i, i)], 0.0);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = i; j < N; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < N; ++k) acc += Z(i * N + k) * w[k] * Z(j * N + k);
            A[Sym<N>::ix(i, j)] = acc;
        }
#undef Z
}
template <int N>
__device__ inline void make_pd_upper(double* __restrict__ A) {
    double Zr[N * N];
    make_pd_upper_z<N>(A, [&](int k) { return &Zr[k]; });
}
// The same rotations with the packed triangle and the eigenvectors of one thread in a
// strided LDS column (entry k at m[k * 64]): plain loops, a handful of registers, no
// scratch.  Used for the per-contact projection where N reaches 12.
template <int N>
__device__ inline void make_pd_lds(double* __restrict__ m) {
    constexpr int NU = N * (N + 1) / 2;
    auto A = [&](int i, int j) -> double& { return m[64 * Sym<N>::ix(i, j)]; };
    auto Z = [&](int i, int j) -> double& { return m[64 * (NU + i * N + j)]; };
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) Z(i, j) = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = 0.0, dg = 0.0;
        for (int i = 0; i < N; ++i) {
            const double d = A(i, i);
            dg += d * d;
            for (int j = i + 1; j < N; ++j) { const double a = A(i, j); off += a * a; }
        }
        if (off == 0.0 || off <= 1e-30 * dg) break;
        for (int p = 0; p < N; ++p) {
            for (int q = p + 1; q < N; ++q) {
                const double apq = A(p, q);
                if (apq == 0.0) continue;
                const double app = A(p, p), aqq = A(q, q);
                const double th = (aqq - app) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < N; ++k) {
                    if (k == p || k == q) continue;
                    const double akp = A(k, p), akq = A(k, q);
                    A(k, p) = c * akp - s * akq;
                    A(k, q) = s * akp + c * akq;
                }
                A(p, p) = app - t * apq;
                A(q, q) = aqq + t * apq;
                A(p, q) = 0.0;
                for (int k = 0; k < N; ++k) {
                    const double zkp = Z(k, p), zkq = Z(k, q);
                    Z(k, p) = c * zkp - s * zkq;
                    Z(k, q) = s * zkp + c * zkq;
                }
            }
        }
    }
    double w[N];
#pragma unroll
    for (int i = 0; i < N; ++i) w[i] = fmax(A(i, i), 0.0);
    for (int i = 0; i < N; ++i)
        for (int j = i; j < N; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < N; ++k) acc += Z(i, k) * w[k] * Z(j, k);
            A(i, j) = acc;
        }
}
// ---- The same projection for contact Hessians, on the translation-free subspace.
// A contact barrier Hessian H (3M x 3M, M = 2, 3, 4 vertices) is the Hessian of a function of
// vertex differences only, so H (1_M (x) e_k) = 0 for the three rigid translations.  With the
// orthonormal Helmert rows h_0..h_{M-2} of length M (all orthogonal to 1_M), B = Hm (x) I_3 is an
// orthonormal basis of the complement, H = B^T R B with R = B H B^T of size 3(M-1), and the
// eigenpairs of H are those of R lifted by B^T plus three zero eigenvalues.  Hence
//   makePD(H) = H - B^T (sum_{lambda<0} lambda z z^T) B   over the eigenpairs (lambda, z) of R,
// and H is returned unchanged when R has no negative eigenvalue.  Exact up to the rounding of
// H's translation components (relative 1e-16 of |H|).  The 12 x 12 projection becomes 9 x 9,
// 9 x 9 becomes 6 x 6 and 6 x 6 becomes 3 x 3: about 2.5x fewer rotations, and the packed
// triangle plus eigenvectors (126 doubles for M = 4) fit the register file.
__device__ constexpr double helmert(int M, int i, int a) {
    // row i: i+1 entries 1/sqrt((i+1)(i+2)), then -(i+1)/sqrt((i+1)(i+2)), then zeros
    constexpr double r2 = 0.70710678118654752440, r6 = 0.40824829046386301637, r12 = 0.28867513459481288225;
    (void)M;
    if (a <= i) return i == 0 ? r2 : i == 1 ? r6 : r12;
    if (a == i + 1) return i == 0 ? -r2 : i == 1 ? -2.0 * r6 : -3.0 * r12;
    return 0.0;
}
// Circle-method round robin over NP (even) indices: pair i of round r.
__device__ constexpr int rr_slot(int NP, int r, int k) { return k == 0 ? 0 : ((k - 1 + r) % (NP - 1)) + 1; }

// Cyclic-by-rounds Jacobi on the packed upper triangle R (Sym<NR>) with eigenvectors Z
// (row-major NR x NR, columns are eigenvectors).  Each round rotates NP/2 disjoint pairs whose
// angles depend only on entries no other pair of the round touches, so the angles of a round
// are computed together (independent square roots and divisions) before the rotations.
template <int NR>
__device__ inline void jacobi_rounds(double* __restrict__ R, double* __restrict__ Z) {
    constexpr int NP = NR + (NR & 1);
    constexpr int NH = NP / 2;
#pragma unroll
    for (int i = 0; i < NR * NR; ++i) Z[i] = (i % (NR + 1) == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 32; ++sweep) {
        double off = 0.0, dg = 0.0;
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            dg += R[Sym<NR>::ix(i, i)] * R[Sym<NR>::ix(i, i)];
#pragma unroll
            for (int j = i + 1; j < NR; ++j) off += R[Sym<NR>::ix(i, j)] * R[Sym<NR>::ix(i, j)];
        }
        if (off == 0.0 || off <= 1e-30 * dg) break;
#pragma unroll
        for (int r = 0; r < NP - 1; ++r) {
            double cs[NH], sn[NH];
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                const int a = rr_slot(NP, r, h), b = rr_slot(NP, r, NP - 1 - h);
                const int p = a < b ? a : b, q = a < b ? b : a;
                cs[h] = 1.0; sn[h] = 0.0;
                if (q >= NR) continue;
                const double apq = R[Sym<NR>::ix(p, q)];
                const double app = R[Sym<NR>::ix(p, p)], aqq = R[Sym<NR>::ix(q, q)];
                const bool zero = apq == 0.0;
                const double th = (aqq - app) / (zero ? 1.0 : 2.0 * apq);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                t = zero ? 0.0 : t;
                const double c = 1.0 / sqrt(t * t + 1.0);
                cs[h] = c; sn[h] = t * c;
                R[Sym<NR>::ix(p, p)] = app - t * apq;
                R[Sym<NR>::ix(q, q)] = aqq + t * apq;
                R[Sym<NR>::ix(p, q)] = zero ? apq : 0.0;
            }
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                const int a = rr_slot(NP, r, h), b = rr_slot(NP, r, NP - 1 - h);
                const int p = a < b ? a : b, q = a < b ? b : a;
                if (q >= NR) continue;
                const double c = cs[h], s = sn[h];
#pragma unroll
                for (int k = 0; k < NR; ++k) {
                    if (k == p || k == q) continue;
                    const double akp = R[Sym<NR>::ix(k, p)], akq = R[Sym<NR>::ix(k, q)];
                    R[Sym<NR>::ix(k, p)] = c * akp - s * akq;
                    R[Sym<NR>::ix(k, q)] = s * akp + c * akq;
                }
#pragma unroll
                for (int k = 0; k < NR; ++k) {
                    const double zkp = Z[k * NR + p], zkq = Z[k * NR + q];
                    Z[k * NR + p] = c * zkp - s * zkq;
                    Z[k * NR + q] = s * zkp + c * zkq;
                }
            }
        }
    }
}

// R = B H B^T on the packed upper triangles (H: Sym<3M> through h(k) -> double&).
template <int M, typename HA>
__device__ inline void reduce_hessian(HA h, double* __restrict__ R) {
    constexpr int N = 3 * M, NR = 3 * (M - 1);
    double H[N * (N + 1) / 2];
#pragma unroll
    for (int k = 0; k < N * (N + 1) / 2; ++k) H[k] = h(k);
#pragma unroll
    for (int i = 0; i < M - 1; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int j = i; j < M - 1; ++j)
#pragma unroll
                for (int l = (j == i ? k : 0); l < 3; ++l) {
                    double acc = 0.0;
#pragma unroll
                    for (int a = 0; a <= i + 1; ++a)
#pragma unroll
                        for (int b = 0; b <= j + 1; ++b)
                            acc += helmert(M, i, a) * helmert(M, j, b) * H[Sym<N>::ix(3 * a + k, 3 * b + l)];
                    R[Sym<NR>::ix(3 * i + k, 3 * j + l)] = acc;
                }
}
// H -= B^T Nr B  (Nr: Sym<3(M-1)>): H_(a,k),(b,l) -= sum_{i,j} Hm[i][a] Hm[j][b] Nr_(i,k),(j,l)
template <int M, typename HA>
__device__ inline void lift_subtract(HA h, const double* __restrict__ Nr) {
    constexpr int N = 3 * M, NR = 3 * (M - 1);
#pragma unroll
    for (int a = 0; a < M; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int b = a; b < M; ++b)
#pragma unroll
                for (int l = (b == a ? k : 0); l < 3; ++l) {
                    double acc = 0.0;
#pragma unroll
                    for (int i = (a > 0 ? a - 1 : 0); i < M - 1; ++i)
#pragma unroll
                        for (int j = (b > 0 ? b - 1 : 0); j < M - 1; ++j)
                            acc += helmert(M, i, a) * helmert(M, j, b) * Nr[Sym<NR>::ix(3 * i + k, 3 * j + l)];
                    h(Sym<N>::ix(3 * a + k, 3 * b + l)) -= acc;
                }
}

// General contact block: eigenpairs of R by Jacobi.  In place on the packed upper triangle
// (Sym<3M>) reached through h(k) -> double&; returns false (h untouched) when no eigenvalue
// is negative.
template <int M, typename HA>
__device__ inline bool project_contact(HA h) {
    constexpr int NR = 3 * (M - 1), NRU = NR * (NR + 1) / 2;
    double R[NRU];
    reduce_hessian<M>(h, R);
    double Z[NR * NR];
    jacobi_rounds<NR>(R, Z);
    double w[NR];
    bool neg = false;
#pragma unroll
    for (int k = 0; k < NR; ++k) { w[k] = fmin(R[Sym<NR>::ix(k, k)], 0.0); neg |= w[k] < 0.0; }
    if (!neg) return false;
    // negative part on the reduced space, reusing R's registers
#pragma unroll
    for (int p = 0; p < NR; ++p)
#pragma unroll
        for (int q = p; q < NR; ++q) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < NR; ++k) acc += (w[k] * Z[p * NR + k]) * Z[q * NR + k];
            R[Sym<NR>::ix(p, q)] = acc;
        }
    lift_subtract<M>(h, R);
    return true;
}

// ---- Structured contact blocks: a smaller eigenproblem on the block's range.
// For a point-plane (point-triangle) or line-line (edge-edge) squared distance d = s^2 the
// gradient is grad s = c (x) n (c: per-vertex weights summing to zero, n: unit normal), and
// the Hessian of s maps into { w (x) n : sum w = 0 } + { c (x) t : t perpendicular to n }
// (the weights and the normal are all that move to first order).  The barrier block
// kappa (b'' g g^T + b' H_d), g = 2 s grad s, H_d = 2 grad s grad s^T + 2 s Hess s, therefore
// has its range in that 5-dimensional space (3 + 2), with the orthonormal basis
// hm_i (x) n (i = 0..2, the Helmert rows) and c (x) t1, c (x) t2 (c normalised).
// For a point-line (point-edge) squared distance d = |r|^2, grad d = 2 c (x) r and the range
// is { c (x) t : t in R^3 } + { (0,1,-1) (x) r }: 4-dimensional.  c and n (or r) are the
// rank-one factors of the block's own gradient G = coef b' grad d (a 3M vector, M rows of
// 3).  The eigenpairs of H are those of the K x K matrix T = Q^T R Q lifted by Q, plus zeros,
// up to the rounding of H (relative 1e-14 on random blocks).  The gradient vanishes only at
// d = dHat, where b' = b'' = 0 and the block is exactly zero, or when the block is not
// finite; the block is then left as it is.
template <int M>
struct RangeDim { static constexpr int K = (M == 4) ? 5 : 4; };

// The basis is built in a frame V = (u, t1, t2) with u = n (or r): column j of the basis is
// om_j (x) V e_{dir_j}, om_j a unit vector of per-vertex weights orthogonal to the ones vector.
// T_ij = sum_{a,b} om_i[a] om_j[b] (V^T H_ab V)[dir_i][dir_j], so each 3x3 block is rotated
// once into the frame and the vertex weights are contracted afterwards.
template <int M, typename HA>
__device__ inline bool project_contact_structured(HA h, const double* __restrict__ G) {
    static_assert(M == 3 || M == 4, "structured blocks have 3 or 4 vertices");
    constexpr int N = 3 * M, K = RangeDim<M>::K, NB = M * (M + 1) / 2;
    // rank-one factors of G (M rows of 3): direction of the largest row, then the weights
    double best = -1.0, ux = 0.0, uy = 0.0, uz = 0.0;
#pragma unroll
    for (int a = 0; a < M; ++a) {
        const double r2 = G[3 * a] * G[3 * a] + G[3 * a + 1] * G[3 * a + 1] + G[3 * a + 2] * G[3 * a + 2];
        if (r2 > best) { best = r2; ux = G[3 * a]; uy = G[3 * a + 1]; uz = G[3 * a + 2]; }
    }
    if (!(best > 0.0) || !(best < 1.7976931348623157e308)) return false;
    const double iu = 1.0 / sqrt(best);
    ux *= iu; uy *= iu; uz *= iu;
    double c[M], cn = 0.0;
#pragma unroll
    for (int a = 0; a < M; ++a) { c[a] = G[3 * a] * ux + G[3 * a + 1] * uy + G[3 * a + 2] * uz; cn += c[a] * c[a]; }
    const double ic = 1.0 / sqrt(cn);
#pragma unroll
    for (int a = 0; a < M; ++a) c[a] *= ic;
    // frame (Duff et al. 2017, "Building an orthonormal basis, revisited")
    double V[3][3];
    {
        const double sg = uz >= 0.0 ? 1.0 : -1.0;
        const double aa = -1.0 / (sg + uz), bb = ux * uy * aa;
        V[0][0] = ux; V[1][0] = uy; V[2][0] = uz;
        V[0][1] = 1.0 + sg * ux * ux * aa; V[1][1] = sg * bb; V[2][1] = -sg * ux;
        V[0][2] = bb; V[1][2] = sg + uy * uy * aa; V[2][2] = -uy;
    }
    // basis columns: weights om[j] and frame axis dir[j]
    double om[K][M];
    int dir[K];
    if constexpr (M == 4) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
#pragma unroll
            for (int a = 0; a < M; ++a) om[j][a] = helmert(M, j, a);
            dir[j] = 0;
        }
#pragma unroll
        for (int a = 0; a < M; ++a) { om[3][a] = c[a]; om[4][a] = c[a]; }
        dir[3] = 1; dir[4] = 2;
    } else {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
#pragma unroll
            for (int a = 0; a < M; ++a) om[j][a] = c[a];
            dir[j] = j;
        }
        // (0, 1, -1)/sqrt2 made orthogonal to c
        double dv[3] = {0.0, 0.70710678118654752440, -0.70710678118654752440};
        const double dc = dv[0] * c[0] + dv[1] * c[1] + dv[2] * c[2];
        double dn = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) { dv[a] -= dc * c[a]; dn += dv[a] * dv[a]; }
        const double idn = 1.0 / sqrt(dn);
#pragma unroll
        for (int a = 0; a < 3; ++a) om[3][a] = dv[a] * idn;
        dir[3] = 0;
    }
    // blocks in the frame, a <= b
    double P[NB][3][3];
#pragma unroll
    for (int a = 0; a < M; ++a)
#pragma unroll
        for (int b = a; b < M; ++b) {
            double Hb[3][3], tmp[3][3];
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int l = 0; l < 3; ++l) Hb[k][l] = h(Sym<N>::ix(3 * a + k, 3 * b + l));
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int y = 0; y < 3; ++y) tmp[k][y] = Hb[k][0] * V[0][y] + Hb[k][1] * V[1][y] + Hb[k][2] * V[2][y];
#pragma unroll
            for (int x = 0; x < 3; ++x)
#pragma unroll
                for (int y = 0; y < 3; ++y) P[Sym<M>::ix(a, b)][x][y] = V[0][x] * tmp[0][y] + V[1][x] * tmp[1][y] + V[2][x] * tmp[2][y];
        }
    double T[K * (K + 1) / 2];
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = i; j < K; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int a = 0; a < M; ++a)
#pragma unroll
                for (int b = 0; b < M; ++b) {
                    const double sab = a <= b ? P[Sym<M>::ix(a, b)][dir[i]][dir[j]] : P[Sym<M>::ix(b, a)][dir[j]][dir[i]];
                    acc += (om[i][a] * om[j][b]) * sab;
                }
            T[Sym<K>::ix(i, j)] = acc;
        }
    // compiler barrier: the block is re-read for the update below rather than held in
    // registers through the eigensolve
    __asm__ volatile("" ::: "memory");
    double Y[K * K];
    jacobi_rounds<K>(T, Y);
    double w[K];
    bool neg = false;
#pragma unroll
    for (int k = 0; k < K; ++k) { w[k] = fmin(T[Sym<K>::ix(k, k)], 0.0); neg |= w[k] < 0.0; }
    if (!neg) return false;
    // eigenvectors in full coordinates: u_k(a, :) = V (sum_j Y[j][k] om_j[a] e_{dir_j})
    double U[K][N];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int a = 0; a < M; ++a) {
            double f[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int j = 0; j < K; ++j) f[dir[j]] += Y[j * K + k] * om[j][a];
#pragma unroll
            for (int r = 0; r < 3; ++r) U[k][3 * a + r] = V[r][0] * f[0] + V[r][1] * f[1] + V[r][2] * f[2];
        }
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
        for (int s2 = r; s2 < N; ++s2) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < K; ++k) acc += (w[k] * U[k][r]) * U[k][s2];
            h(Sym<N>::ix(r, s2)) -= acc;
        }
    return true;
}

// Full N x N row-major input/output (element-level evaluation).
template <int N>
__device__ inline void make_pd(double* __restrict__ H) {
    double A[N * (N + 1) / 2];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = i; j < N; ++j) A[Sym<N>::ix(i, j)] = H[i * N + j];
    make_pd_upper<N>(A);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = i; j < N; ++j) { H[i * N + j] = A[Sym<N>::ix(i, j)]; H[j * N + i] = A[Sym<N>::ix(i, j)]; }
}

// ---- Additive CCD (AdditiveCCD.cpp:146-329), reference operation order
__device__ inline double ccd_pe(d3 p, d3 e0, d3 e1) { return sqn(cross(e0 - p, e1 - p)) / sqn(e1 - e0); }
__device__ inline double ccd_pt(d3 p, d3 t0, d3 t1, d3 t2) {
    const d3 b = cross(t1 - t0, t2 - t0);
    const double aTb = dot(p - t0, b);
    return aTb * aTb / sqn(b);
}
__device__ inline double ccd_ee(d3 a0, d3 a1, d3 b0, d3 b1) {
    const d3 b = cross(a1 - a0, b1 - b0);
    const double aTb = dot(b0 - a0, b);
    return aTb * aTb / sqn(b);
}
__device__ inline double ee_unclassified(d3 a0, d3 a1, d3 b0, d3 b1) {
    switch (dtype_ee(a0, a1, b0, b1)) {
    case 0: return sqn(a0 - b0);
    case 1: return sqn(a0 - b1);
    case 2: return ccd_pe(a0, b0, b1);
    case 3: return sqn(a1 - b0);
    case 4: return sqn(a1 - b1);
    case 5: return ccd_pe(a1, b0, b1);
    case 6: return ccd_pe(b0, a0, a1);
    case 7: return ccd_pe(b1, a0, a1);
    case 8: return ccd_ee(a0, a1, b0, b1);
    default: return 1.7976931348623157e308;
    }
}
__device__ inline double pt_unclassified(d3 p, d3 t0, d3 t1, d3 t2) {
    switch (dtype_pt(p, t0, t1, t2)) {
    case 0: return sqn(p - t0);
    case 1: return sqn(p - t1);
    case 2: return sqn(p - t2);
    case 3: return ccd_pe(p, t0, t1);
    case 4: return ccd_pe(p, t1, t2);
    case 5: return ccd_pe(p, t2, t0);
    case 6: return ccd_pt(p, t0, t1, t2);
    default: return 1.7976931348623157e308;
    }
}
__device__ inline double clamp01(double toc) { return fmax(0.0, fmin(toc, 1.0)); }
// edge_edge_ccd (AdditiveCCD.cpp:221-279).  `cap`: the caller only needs the result when it
// is below cap (a step bound already known); the advance stops once it reaches cap and
// returns a value >= cap.  Results below cap are those of the uncapped loop.
__device__ inline double accd_ee(d3 ea0, d3 ea1, d3 eb0, d3 eb1, d3 dea0, d3 dea1, d3 deb0, d3 deb1, double eta, double th,
                                 double cap = __builtin_inf()) {
    const d3 mov = (dea0 + dea1 + deb0 + deb1) / 4;
    dea0 = dea0 - mov; dea1 = dea1 - mov; deb0 = deb0 - mov; deb1 = deb1 - mov;
    const double maxd = sqrt(fmax(sqn(dea0), sqn(dea1))) + sqrt(fmax(sqn(deb0), sqn(deb1)));
    if (maxd == 0) return 1.0;
    double d2 = ee_unclassified(ea0, ea1, eb0, eb1);
    double dF = d2 - th * th;
    if (dF <= 0) {
        d2 = fmin(fmin(sqn(ea0 - eb0), sqn(ea0 - eb1)), fmin(sqn(ea1 - eb0), sqn(ea1 - eb1)));
        dF = d2 - th * th;
    }
    double dc = sqrt(d2);
    const double gap = eta * dF / (dc + th);
    double toc = 0.0;
    for (int it = 0; it < 10000; ++it) {
        const double tl = (1 - eta) * dF / ((dc + th) * maxd);
        if (!isfinite(tl) || tl <= 0.0) return clamp01(toc);
        ea0 = ea0 + tl * dea0; ea1 = ea1 + tl * dea1; eb0 = eb0 + tl * deb0; eb1 = eb1 + tl * deb1;
        d2 = ee_unclassified(ea0, ea1, eb0, eb1);
        dF = d2 - th * th;
        if (dF <= 0) {
            d2 = fmin(fmin(sqn(ea0 - eb0), sqn(ea0 - eb1)), fmin(sqn(ea1 - eb0), sqn(ea1 - eb1)));
            dF = d2 - th * th;
        }
        dc = sqrt(d2);
        if (toc != 0.0 && (dF / (dc + th) < gap)) break;
        toc += tl;
        if (toc > 1.0) return 1.0;
        if (toc >= cap) return toc;
    }
    return clamp01(toc);
}
// point_triangle_ccd (AdditiveCCD.cpp:281-329); `cap` as for accd_ee
__device__ inline double accd_pt(d3 p, d3 t0, d3 t1, d3 t2, d3 dp, d3 dt0, d3 dt1, d3 dt2, double eta, double th,
                                 double cap = __builtin_inf()) {
    const d3 mov = (dt0 + dt1 + dt2 + dp) / 4;
    dt0 = dt0 - mov; dt1 = dt1 - mov; dt2 = dt2 - mov; dp = dp - mov;
    const double mx = fmax(fmax(sqn(dt0), sqn(dt1)), sqn(dt2));
    const double maxd = sqrt(sqn(dp)) + sqrt(mx);
    if (maxd == 0) return 1.0;
    double d2 = pt_unclassified(p, t0, t1, t2);
    double dc = sqrt(d2);
    const double gap = eta * (d2 - th * th) / (dc + th);
    double toc = 0.0;
    for (int it = 0; it < 10000; ++it) {
        const double tl = (1 - eta) * (d2 - th * th) / ((dc + th) * maxd);
        if (!isfinite(tl) || tl <= 0.0) return clamp01(toc);
        p = p + tl * dp; t0 = t0 + tl * dt0; t1 = t1 + tl * dt1; t2 = t2 + tl * dt2;
        d2 = pt_unclassified(p, t0, t1, t2);
        dc = sqrt(d2);
        if (toc != 0.0 && ((d2 - th * th) / (dc + th) < gap)) break;
        toc += tl;
        if (toc > 1.0) return 1.0;
        if (toc >= cap) return toc;
    }
    return clamp01(toc);
}

// Swept boxes {x, x - a p} (LBVH.cpp:120-129,400-431).
__device__ inline Box box_pt(d3 p) { return Box{p, p}; }
__device__ inline Box box_add(Box b, d3 p) { return Box{vmin(b.lo, p), vmax(b.hi, p)}; }
__device__ inline Box box_edge(d3 a, d3 b) { return box_add(box_pt(a), b); }
__device__ inline Box box_tri(d3 a, d3 b, d3 c) { return box_add(box_add(box_pt(a), b), c); }
__device__ inline Box box_union(Box a, Box b) { return Box{vmin(a.lo, b.lo), vmax(a.hi, b.hi)}; }
__device__ inline Box swept_pt(d3 p, d3 dir, double a) { return box_add(box_pt(p), p - a * dir); }

}  // namespace cipc
