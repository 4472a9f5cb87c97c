// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate expressions of CPU-IPC
// (MPL-2.0) in the reference's operation order; each such part cites the source it follows.
// Stable neo-Hookean tet element on the device: energy, PK1 and the PSD-projected element
// stiffness, all in FP64.
//
// Reference: computePEPF_StableNHK3D_2_double (CPU IPC/Elasticity.cpp:1144-1170),
// project_StabbleNHK_2_H_3D active branch + IglUtils::makePD<9> (Elasticity.cpp:1337-1345,
// 1436-1477; ContactMechanics.h:27-45), getObjEnergy_StableNHK2_3D (Elasticity.cpp:695-727),
// element assembly vol dt^2 PFPX^T H PFPX (IPCSolver.cpp:139-164).
//
// Device formulation (same mathematics, different arithmetic):
//  * The reference takes I2 = sum s_i^2 and I3 = prod s_i from its rotation-aware SVD; they
//    equal ||F||_F^2 and det F exactly in real arithmetic, so they are computed directly.
//  * With F = U diag(s) V^T (U, V rotations, s2 signed) the 9x9 Hessian
//      H = A I + a2 HJ + a3 g1 g1^T + r gJ gJ^T,   A = mu Ic/(Ic+1), a2 = r(J - 1 - 3mu/4r)
//    (g1 = 2 vec F, gJ = vec cof F, HJ = Hessian of det) has closed-form eigenpairs: for
//    each axis k with (i,j) the other two,
//      flip  A - a2 s_k, M = (u_i v_j^T + u_j v_i^T)/sqrt2
//      twist A + a2 s_k, M = (u_i v_j^T - u_j v_i^T)/sqrt2
//    and three scaling modes M = U diag(w) V^T, (lambda, w) the eigenpairs of the 3x3
//      S = A I + a2 [[0,s2,s1],[s2,0,s0],[s1,s0,0]] + (2mu/(Ic+1)^2) s s^T + r c c^T,
//      c = (s1 s2, s0 s2, s0 s1).
//    makePD leaves H untouched when no eigenvalue is negative; otherwise it zeroes the
//    negative ones, i.e. H+ = H - sum_{lambda<0} lambda vec(M) vec(M)^T.
//  * PFPX = B^T (x) I3 with B(0,j) = -sum_k Dm^-1(k,j), B(a,j) = Dm^-1(a-1,j), so with the
//    3-vectors B_a the 12x12 stiffness needs no 9x9 matrix at all:
//      K_ab = A (B_a.B_b) I3 - a2 hat(F (B_a x B_b)) + 4 a3 (F B_a)(F B_b)^T
//             + r (cof F B_a)(cof F B_b)^T - sum_{lambda<0} lambda (M B_a)(M B_b)^T
//    with hat(w) the cross-product matrix of w, and M B_a = sum_k (v_k . B_a) (M v_k).
//    Every M B_a lies in the frame of U, so the last sum is U P_ab U^T with a 3x3 P_ab that
//    is bilinear in V^T B_a and V^T B_b (SNH::Sw): no branch per mode.
#pragma once
#include "common.hpp"

namespace cipc {

struct M33 {
    double a[3][3];  // a[row][col]
};

__host__ __device__ inline d3 mcol(const M33& m, int c) { return {m.a[0][c], m.a[1][c], m.a[2][c]}; }
__host__ __device__ inline d3 mvec(const M33& m, d3 v) {
    return {m.a[0][0] * v.x + m.a[0][1] * v.y + m.a[0][2] * v.z,
            m.a[1][0] * v.x + m.a[1][1] * v.y + m.a[1][2] * v.z,
            m.a[2][0] * v.x + m.a[2][1] * v.y + m.a[2][2] * v.z};
}
__host__ __device__ inline double comp(d3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
__host__ __device__ inline void setcol(M33& m, int c, d3 v) { m.a[0][c] = v.x; m.a[1][c] = v.y; m.a[2][c] = v.z; }

// One-sided Jacobi SVD of a 3x3: F = U diag(s) V^T, U and V proper rotations,
// s0 >= s1 >= |s2|, s2 carrying sign(det F).
__device__ inline void svd3(const M33& F, M33& U, double s[3], M33& V) {
    d3 b0 = mcol(F, 0), b1 = mcol(F, 1), b2 = mcol(F, 2);
    d3 w0 = {1, 0, 0}, w1 = {0, 1, 0}, w2 = {0, 0, 1};
    auto rot = [](d3& bp, d3& bq, d3& wp, d3& wq) -> bool {
        const double al = sqn(bp), be = sqn(bq), ga = dot(bp, bq);
        // converged to the rounding level: |ga| <= 4e-16 sqrt(al be), in squares (one sqrt less);
        // a threshold below the FP64 unit roundoff is never met and would run every sweep
        if (ga * ga <= 1.6e-31 * (al * be)) return false;
        // tangent of the smaller rotation angle, tan = 2 ga sgn(d) / (|d| + sqrt(d^2 + 4 ga^2)),
        // d = be - al (the textbook sgn(z)/(|z| + sqrt(1 + z^2)), z = d / 2ga, without its two
        // divisions); one square root, one division and one reciprocal square root
        const double d = be - al, g2 = 2.0 * ga;
        const double t = (d >= 0 ? g2 : -g2) / (fabs(d) + sqrt(d * d + g2 * g2));
        const double c = rsqrt(1.0 + t * t), sn = c * t;
        const d3 np = c * bp - sn * bq, nq = sn * bp + c * bq;
        const d3 vp = c * wp - sn * wq, vq = sn * wp + c * wq;
        bp = np; bq = nq; wp = vp; wq = vq;
        return true;
    };
    for (int sweep = 0; sweep < 12; ++sweep) {
        bool r = rot(b0, b1, w0, w1);
        r |= rot(b0, b2, w0, w2);
        r |= rot(b1, b2, w1, w2);
        if (!r) break;
    }
    double n0 = sqrt(sqn(b0)), n1 = sqrt(sqn(b1)), n2 = sqrt(sqn(b2));
    auto swp = [](d3& x, d3& y, d3& p, d3& q, double& a, double& b) {
        d3 t = x; x = y; y = t; t = p; p = q; q = t; double u = a; a = b; b = u;
    };
    if (n0 < n1) swp(b0, b1, w0, w1, n0, n1);
    if (n1 < n2) swp(b1, b2, w1, w2, n1, n2);
    if (n0 < n1) swp(b0, b1, w0, w1, n0, n1);
    if (dot(cross(w0, w1), w2) < 0) { w2 = -w2; b2 = -b2; }
    d3 u0, u1;
    if (n0 > 0) u0 = b0 / n0; else u0 = d3{1, 0, 0};
    if (n1 > 0) {
        u1 = b1 / n1;
    } else {
        const d3 tr = fabs(u0.x) < 0.9 ? d3{1, 0, 0} : d3{0, 1, 0};
        u1 = cross(u0, tr);
        u1 = u1 / sqrt(sqn(u1));
    }
    const d3 u2 = cross(u0, u1);
    s[0] = n0; s[1] = n1; s[2] = dot(b2, u2);  // F v2 = s2 u2, signed
    setcol(U, 0, u0); setcol(U, 1, u1); setcol(U, 2, u2);
    setcol(V, 0, w0); setcol(V, 1, w1); setcol(V, 2, w2);
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi: S = W diag(l) W^T (columns of W).
__device__ inline void sym3_eig(double S[3][3], double l[3], double W[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) W[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 16; ++sweep) {
        const double off = S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2];
        const double dg = S[0][0] * S[0][0] + S[1][1] * S[1][1] + S[2][2] * S[2][2];
        if (off == 0.0 || off <= 1e-34 * dg) break;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int p = (k == 2) ? 1 : 0, q = (k == 0) ? 1 : 2;
            const double apq = S[p][q];
            if (apq == 0.0) continue;
            // as in svd3: tan = 2 apq sgn(d) / (|d| + sqrt(d^2 + 4 apq^2)), d = S_qq - S_pp
            const double d = S[q][q] - S[p][p], g2 = 2.0 * apq;
            const double t = (d >= 0 ? g2 : -g2) / (fabs(d) + sqrt(d * d + g2 * g2));
            const double c = rsqrt(t * t + 1.0), s = t * c;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const double a = S[m][p], b = S[m][q];
                S[m][p] = c * a - s * b; S[m][q] = s * a + c * b;
            }
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const double a = S[p][m], b = S[q][m];
                S[p][m] = c * a - s * b; S[q][m] = s * a + c * b;
            }
            S[p][q] = S[q][p] = 0.0;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const double a = W[m][p], b = W[m][q];
                W[m][p] = c * a - s * b; W[m][q] = s * a + c * b;
            }
        }
    }
    l[0] = S[0][0]; l[1] = S[1][1]; l[2] = S[2][2];
}

// Everything the gradient, the projected stiffness and the energy need from one F.
struct SNH {
    M33 F, C;             // F and cof F (columns c_j = f_{j+1} x f_{j+2})
    double I2, I3;
    double A, a2, a3, r;  // Hessian coefficients
    double pa, pb;        // PK1 = pa F + pb cof F
    M33 U, V;             // SVD frames
    double s[3];
    double W[3][3];       // scaling-mode eigenvectors (columns)
    double lam[9];        // eigenvalues: flip k (0..2), twist k (3..5), scaling m (6..8); the
                          // scaling ones hold S's diagonal when S is positive definite
    bool project;         // any eigenvalue < 0
    // The projection in the frame of U (set when project): with w = max(-lambda, 0) and
    // beta = V^T B_a, the correction of block (a, b) is U P_ab U^T,
    //   P_ab[i][j] = Sw[i][j] beta_a[i] beta_b[j]                   (scaling modes)
    //   + for axis k, (i, j) its other two axes (flip and twist modes of k):
    //     P[i][i] += pp[k] beta_a[j] beta_b[j],  P[j][j] += pp[k] beta_a[i] beta_b[i],
    //     P[i][j] += pm[k] beta_a[j] beta_b[i],  P[j][i] += pm[k] beta_a[i] beta_b[j],
    //   Sw = W diag(w_scaling) W^T, pp = (w_flip + w_twist)/2, pm = (w_flip - w_twist)/2.
    double Sw[3][3], pp[3], pm[3];
};

// snh_eval in two halves: the invariants, Hessian coefficients and PK1 coefficients (all a
// gradient needs), then the SVD frames, eigenvalues and projection data.
__device__ inline void snh_invariants(SNH& e, double u, double r);
__device__ inline void snh_modes(SNH& e, double u, double r);
__device__ inline void snh_eval(SNH& e, double u, double r) {
    snh_invariants(e, u, r);
    snh_modes(e, u, r);
}

__device__ inline void snh_invariants(SNH& e, double u, double r) {
    const d3 f0 = mcol(e.F, 0), f1 = mcol(e.F, 1), f2 = mcol(e.F, 2);
    setcol(e.C, 0, cross(f1, f2)); setcol(e.C, 1, cross(f2, f0)); setcol(e.C, 2, cross(f0, f1));
    e.I2 = sqn(f0) + sqn(f1) + sqn(f2);
    e.I3 = dot(f0, mcol(e.C, 0));
    const double Ic = e.I2, J = e.I3, mu = u;
    e.A = (Ic * mu) / (Ic + 1);
    e.a2 = r * (J - 1 - (3 * mu) / (4.0 * r));
    e.a3 = mu / (2 * (Ic + 1) * (Ic + 1));
    e.r = r;
    e.pa = u * (1 - 1 / (e.I2 + 1));
    e.pb = r * (e.I3 - 1 - u * 3 / (r * 4));
}

__device__ inline void snh_modes(SNH& e, double u, double r) {
    const double Ic = e.I2, mu = u;
    svd3(e.F, e.U, e.s, e.V);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        e.lam[k] = e.A - e.a2 * e.s[k];
        e.lam[3 + k] = e.A + e.a2 * e.s[k];
    }
    const double cc = 2.0 * mu / ((Ic + 1) * (Ic + 1));
    const double cv[3] = {e.s[1] * e.s[2], e.s[0] * e.s[2], e.s[0] * e.s[1]};
    double S[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double off = (a == b) ? 0.0 : e.s[3 - a - b];
            S[a][b] = (a == b ? e.A : 0.0) + e.a2 * off + cc * e.s[a] * e.s[b] + r * cv[a] * cv[b];
        }
    double l[3];
    // The scaling modes matter only through their negative eigenvalues.  When S is positive
    // definite (Sylvester's criterion: positive leading minors) there are none and their
    // eigenvectors are not needed: l takes S's (positive) diagonal and W the identity.  Near
    // rest and under moderate strain S is positive definite, which skips the Jacobi sweeps.
    const double m2 = S[0][0] * S[1][1] - S[0][1] * S[0][1];
    const double m3 = S[0][0] * (S[1][1] * S[2][2] - S[1][2] * S[1][2]) - S[0][1] * (S[0][1] * S[2][2] - S[1][2] * S[0][2]) +
                      S[0][2] * (S[0][1] * S[1][2] - S[1][1] * S[0][2]);
    if (S[0][0] > 0.0 && m2 > 0.0 && m3 > 0.0) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            l[i] = S[i][i];
#pragma unroll
            for (int j = 0; j < 3; ++j) e.W[i][j] = (i == j) ? 1.0 : 0.0;
        }
    } else {
        sym3_eig(S, l, e.W);
    }
    e.lam[6] = l[0]; e.lam[7] = l[1]; e.lam[8] = l[2];
    bool neg = false;
#pragma unroll
    for (int m = 0; m < 9; ++m) neg |= e.lam[m] < 0.0;
    e.project = neg;
    double w[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) w[m] = e.lam[m] < 0.0 ? -e.lam[m] : 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        e.pp[k] = 0.5 * (w[k] + w[3 + k]);
        e.pm[k] = 0.5 * (w[k] - w[3 + k]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            e.Sw[i][j] = w[6] * e.W[i][0] * e.W[j][0] + w[7] * e.W[i][1] * e.W[j][1] + w[8] * e.W[i][2] * e.W[j][2];
}

// M_m applied to a vector x, given beta_k = v_k . x (mode index m static after unrolling).
__device__ inline d3 snh_mode_apply(const SNH& e, int m, const double beta[3]) {
    const double is2 = 0.70710678118654752440;
    if (m < 6) {
        const int k = m % 3;
        const int i = (k == 0) ? 1 : 0, j = (k == 2) ? 1 : 2;
        const double sg = (m < 3) ? 1.0 : -1.0;
        return (is2 * beta[j]) * mcol(e.U, i) + (sg * is2 * beta[i]) * mcol(e.U, j);
    }
    const int q = m - 6;
    return (e.W[0][q] * beta[0]) * mcol(e.U, 0) + (e.W[1][q] * beta[1]) * mcol(e.U, 1) + (e.W[2][q] * beta[2]) * mcol(e.U, 2);
}

// Energy density psi(F) (Elasticity.cpp:711-714) with I2 = ||F||^2, I3 = det F.
__device__ inline double snh_energy_density(double I2, double I3, double u, double r) {
    const double k = I3 - 1 - 3 * u / 4 / r;
    return 0.5 * u * (I2 - 3) + 0.5 * r * k * k - 0.5 * u * log(I2 + 1);
}

}  // namespace cipc
