// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate expressions of CPU-IPC
// (MPL-2.0) in the reference's operation order; each such part cites the source it follows.
// ORACLE — test infrastructure only.  CPU restatement of CPU-IPC's per-Newton-iteration
// path, used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
// checker.  It is never linked into, loaded by, or called from the product library.
//
// Restates the behaviour of MPL-2.0 licensed CPU-IPC sources; each function cites the
// reference file:line it follows (paths relative to the reference root).
//
// Small dense linear algebra shared by the oracle: 3-vectors, the 3x3 rotation-aware SVD
// convention, a cyclic-Jacobi symmetric eigensolver with the reference's makePD rule,
// and a second-order forward-mode automatic-differentiation scalar used to take exact
// gradients and Hessians of the closed-form distance functions.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <utility>
#include <vector>

namespace oracle {

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
// Component-wise arithmetic in the evaluation order Eigen's fixed-size expressions use
// (left to right, no fused multiply-add: the oracle is compiled with -ffp-contract=off).
inline V3 operator+(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(const V3& a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(double s, const V3& a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(const V3& a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double sqn(const V3& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline V3 cross(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 vmin(const V3& a, const V3& b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
inline V3 vmax(const V3& a, const V3& b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }

// Column-major 3x3: m[r + 3c].
struct M3 {
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double operator()(int r, int c) const { return m[r + 3 * c]; }
    double& operator()(int r, int c) { return m[r + 3 * c]; }
};
inline M3 matmul(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r(i, j) = a(i, 0) * b(0, j) + a(i, 1) * b(1, j) + a(i, 2) * b(2, j);
    return r;
}
inline double det3(const M3& a) {
    return a(0, 0) * (a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2))
         - a(1, 0) * (a(0, 1) * a(2, 2) - a(2, 1) * a(0, 2))
         + a(2, 0) * (a(0, 1) * a(1, 2) - a(1, 1) * a(0, 2));
}

// ---------------------------------------------------------------------------------
// 3x3 SVD with the reference's conventions (CPU IPC/RotationAwareSVD.h:11-13;
// CPU IPC/ImplicitQR3x3SVD.h:725-732): A = U diag(s) V^T with U, V proper rotations,
// |s0| >= |s1| >= |s2| and only s2 possibly negative (it carries sign(det A)).
// The reference computes it with the Gast-Fu-Jiang-Teran implicit-shifted QR; the
// oracle uses one-sided Jacobi (Hestenes) instead: the decomposition is unique up to
// rounding for distinct singular values, and the reference consumes it only through
// s (I2 = sum s^2, I3 = prod s, Elasticity.cpp:1147-1150,1343-1345) and through
// S = V diag(s) V^T (Elasticity.cpp:704-709), which are both invariant to the
// remaining sign/rotation freedom.
// ---------------------------------------------------------------------------------
inline void svd3(const M3& A, M3& U, double s[3], M3& V) {
    M3 B = A;
    V = M3();
    V(0, 0) = V(1, 1) = V(2, 2) = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
        for (int k = 0; k < 3; ++k) {
            const int p = P[k], q = Q[k];
            double alpha = 0, beta = 0, gamma = 0;
            for (int i = 0; i < 3; ++i) {
                alpha += B(i, p) * B(i, p);
                beta += B(i, q) * B(i, q);
                gamma += B(i, p) * B(i, q);
            }
            if (gamma == 0.0 || std::fabs(gamma) <= 1e-17 * std::sqrt(alpha * beta)) continue;
            rotated = true;
            const double zeta = (beta - alpha) / (2.0 * gamma);
            const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
            const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
            for (int i = 0; i < 3; ++i) {
                const double bp = B(i, p), bq = B(i, q);
                B(i, p) = c * bp - sn * bq;
                B(i, q) = sn * bp + c * bq;
                const double vp = V(i, p), vq = V(i, q);
                V(i, p) = c * vp - sn * vq;
                V(i, q) = sn * vp + c * vq;
            }
        }
        if (!rotated) break;
    }
    double nrm[3];
    for (int k = 0; k < 3; ++k) nrm[k] = std::sqrt(B(0, k) * B(0, k) + B(1, k) * B(1, k) + B(2, k) * B(2, k));
    int ord[3] = {0, 1, 2};
    std::sort(ord, ord + 3, [&](int a, int b) { return nrm[a] > nrm[b]; });
    M3 Bs, Vs;
    for (int k = 0; k < 3; ++k) {
        s[k] = nrm[ord[k]];
        for (int i = 0; i < 3; ++i) {
            Bs(i, k) = B(i, ord[k]);
            Vs(i, k) = V(i, ord[k]);
        }
    }
    V = Vs;
    U = M3();
    // Left vectors from the orthogonalised columns; complete rank-deficient cases.
    auto col = [&](const M3& M, int k) { return V3(M(0, k), M(1, k), M(2, k)); };
    auto setcol = [&](M3& M, int k, const V3& v) { M(0, k) = v.x; M(1, k) = v.y; M(2, k) = v.z; };
    V3 u0, u1, u2;
    if (s[0] > 0) u0 = col(Bs, 0) / s[0]; else u0 = V3(1, 0, 0);
    if (s[1] > 0) {
        u1 = col(Bs, 1) / s[1];
    } else {
        V3 trial = std::fabs(u0.x) < 0.9 ? V3(1, 0, 0) : V3(0, 1, 0);
        u1 = cross(u0, trial);
        u1 = u1 / std::sqrt(sqn(u1));
    }
    if (s[2] > 0) u2 = col(Bs, 2) / s[2]; else u2 = cross(u0, u1);
    setcol(U, 0, u0); setcol(U, 1, u1); setcol(U, 2, u2);
    if (det3(U) < 0) { setcol(U, 2, -u2); s[2] = -s[2]; }
    if (det3(V) < 0) { setcol(V, 2, -col(V, 2)); s[2] = -s[2]; }
}

// ---------------------------------------------------------------------------------
// Symmetric eigen-decomposition by cyclic Jacobi (rotations in Rutishauser's form).
// A is N x N row-major, overwritten; eigenvalues ascending in w, eigenvectors in the
// columns of Z (row-major Z[i*N + k] = component i of vector k).
// ---------------------------------------------------------------------------------
template <int N>
inline void sym_eig(double* A, double* w, double* Z) {
    for (int i = 0; i < N * N; ++i) Z[i] = 0.0;
    for (int i = 0; i < N; ++i) Z[i * N + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < N; ++i) {
            diag += A[i * N + i] * A[i * N + i];
            for (int j = i + 1; j < N; ++j) off += A[i * N + j] * A[i * N + j];
        }
        if (off == 0.0 || off <= 1e-34 * diag) break;
        for (int p = 0; p < N; ++p) {
            for (int q = p + 1; q < N; ++q) {
                const double apq = A[p * N + q];
                if (apq == 0.0) continue;
                const double app = A[p * N + p], aqq = A[q * N + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < N; ++k) {
                    const double akp = A[k * N + p], akq = A[k * N + q];
                    A[k * N + p] = c * akp - s * akq;
                    A[k * N + q] = s * akp + c * akq;
                }
                for (int k = 0; k < N; ++k) {
                    const double apk = A[p * N + k], aqk = A[q * N + k];
                    A[p * N + k] = c * apk - s * aqk;
                    A[q * N + k] = s * apk + c * aqk;
                }
                A[p * N + q] = A[q * N + p] = 0.0;
                for (int k = 0; k < N; ++k) {
                    const double zkp = Z[k * N + p], zkq = Z[k * N + q];
                    Z[k * N + p] = c * zkp - s * zkq;
                    Z[k * N + q] = s * zkp + c * zkq;
                }
            }
        }
    }
    int ord[N];
    double wl[N];
    for (int i = 0; i < N; ++i) { ord[i] = i; wl[i] = A[i * N + i]; }
    std::sort(ord, ord + N, [&](int a, int b) { return wl[a] < wl[b]; });
    double Zs[N * N];
    for (int k = 0; k < N; ++k) {
        w[k] = wl[ord[k]];
        for (int i = 0; i < N; ++i) Zs[i * N + k] = Z[i * N + ord[k]];
    }
    std::memcpy(Z, Zs, sizeof(double) * N * N);
}

// IglUtils::makePD (CPU IPC/ContactMechanics.h:27-45): if the smallest eigenvalue is
// >= 0 the matrix is returned untouched; otherwise the leading (negative) eigenvalues
// are set to zero and the matrix is rebuilt as Z D Z^T.  H is row-major N x N.
//
// Eigen's SelfAdjointEigenSolver reads only the lower triangle, so the decomposition is
// of the matrix mirrored from H's lower triangle; an unprojected H is returned as given
// (it may be non-symmetric: the hinge-bending Hessian is).
template <int N>
inline bool make_pd(double* H) {
    double A[N * N], w[N], Z[N * N];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j <= i; ++j) A[i * N + j] = A[j * N + i] = H[i * N + j];
    sym_eig<N>(A, w, Z);
    if (w[0] >= 0.0) return false;
    for (int i = 0; i < N; ++i) {
        if (w[i] < 0.0) w[i] = 0.0; else break;
    }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double acc = 0.0;
            for (int k = 0; k < N; ++k) acc += Z[i * N + k] * w[k] * Z[j * N + k];
            H[i * N + j] = acc;
        }
    return true;
}

// ---------------------------------------------------------------------------------
// Second-order forward-mode AD over N independent variables.  Evaluating a closed-form
// scalar expression with this type yields its exact gradient and Hessian (up to
// rounding); the oracle uses it for every distance derivative so that it shares no
// derivative code with the device path.
// ---------------------------------------------------------------------------------
template <int N>
struct HD {
    double v = 0;
    double g[N] = {};
    double h[N][N] = {};
    static HD var(double value, int i) { HD r; r.v = value; r.g[i] = 1.0; return r; }
    static HD cst(double value) { HD r; r.v = value; return r; }
};
template <int N> inline HD<N> operator+(const HD<N>& a, const HD<N>& b) {
    HD<N> r; r.v = a.v + b.v;
    for (int i = 0; i < N; ++i) { r.g[i] = a.g[i] + b.g[i]; for (int j = 0; j < N; ++j) r.h[i][j] = a.h[i][j] + b.h[i][j]; }
    return r;
}
template <int N> inline HD<N> operator-(const HD<N>& a, const HD<N>& b) {
    HD<N> r; r.v = a.v - b.v;
    for (int i = 0; i < N; ++i) { r.g[i] = a.g[i] - b.g[i]; for (int j = 0; j < N; ++j) r.h[i][j] = a.h[i][j] - b.h[i][j]; }
    return r;
}
template <int N> inline HD<N> operator*(const HD<N>& a, const HD<N>& b) {
    HD<N> r; r.v = a.v * b.v;
    for (int i = 0; i < N; ++i) {
        r.g[i] = a.g[i] * b.v + a.v * b.g[i];
        for (int j = 0; j < N; ++j)
            r.h[i][j] = a.h[i][j] * b.v + a.v * b.h[i][j] + a.g[i] * b.g[j] + a.g[j] * b.g[i];
    }
    return r;
}
template <int N> inline HD<N> inv(const HD<N>& b) {
    HD<N> r; const double iv = 1.0 / b.v; r.v = iv;
    for (int i = 0; i < N; ++i) {
        r.g[i] = -b.g[i] * iv * iv;
        for (int j = 0; j < N; ++j)
            r.h[i][j] = 2.0 * b.g[i] * b.g[j] * iv * iv * iv - b.h[i][j] * iv * iv;
    }
    return r;
}
template <int N> inline HD<N> operator/(const HD<N>& a, const HD<N>& b) { return a * inv(b); }

template <int N> struct HV3 { HD<N> x, y, z; };
template <int N> inline HV3<N> hsub(const HV3<N>& a, const HV3<N>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <int N> inline HD<N> hdot(const HV3<N>& a, const HV3<N>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <int N> inline HV3<N> hcross(const HV3<N>& a, const HV3<N>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <int N> inline HV3<N> hvar(const V3& p, int slot) {
    return {HD<N>::var(p.x, 3 * slot), HD<N>::var(p.y, 3 * slot + 1), HD<N>::var(p.z, 3 * slot + 2)};
}

}  // namespace oracle
