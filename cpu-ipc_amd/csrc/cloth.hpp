// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate expressions of CPU-IPC
// (MPL-2.0) in the reference's operation order; each such part cites the source it follows.
// Cloth element math on the device: Baraff-Witkin triangles and hinge angles.
//
// Reference: computePEPF_baraffwitkin_double / project_baraffwitkint_H_3D /
// getObjEnergy_baraffwitkin_3D (CPU IPC/Elasticity.cpp:1173-1205,1481-1544,729-771) with
// anisotropy directions a = (1,0), b = (0,1); edgeTheta and LibShell::angle
// (Elasticity.cpp:12-150).
//
// A 3x2 deformation gradient is held as its two columns fu = F a and fv = F b; vec(F) is
// (fu, fv), six entries, column-major like Eigen's storage.
#pragma once
#include "common.hpp"

namespace cipc {

// PK1 (6 entries, column-major) of the Baraff-Witkin energy: stretch 2 c_u fu, 2 c_v fv with
// c = 1 - 1/s (+ 1.5 k (s + 1/s - 2) in tension), s the column length; shear 2 I6 (fv, fu).
__device__ inline void bw_pk1(d3 fu, d3 fv, double st, double sh, double k, double P[6]) {
    const double I6 = dot(fu, fv);
    const double I5u = dot(fu, fu), I5v = dot(fv, fv);
    const double su = sqrt(I5u), sv = sqrt(I5v);
    double cu = 1.0 - 1 / su, cv = 1.0 - 1 / sv;
    if (I5u > 1) cu += 1.5 * k * (su + 1 / su - 2);
    if (I5v > 1) cv += 1.5 * k * (sv + 1 / sv - 2);
    P[0] = st * (cu * 2. * fu.x) + sh * (2 * I6 * fv.x);
    P[1] = st * (cu * 2. * fu.y) + sh * (2 * I6 * fv.y);
    P[2] = st * (cu * 2. * fu.z) + sh * (2 * I6 * fv.z);
    P[3] = st * (cv * 2. * fv.x) + sh * (2 * I6 * fu.x);
    P[4] = st * (cv * 2. * fv.y) + sh * (2 * I6 * fu.y);
    P[5] = st * (cv * 2. * fv.z) + sh * (2 * I6 * fu.z);
}

__device__ inline d3 unit_or_same(d3 v) {
    const double n2 = sqn(v);
    if (!(n2 > 0.0)) return v;
    const double n = sqrt(n2);
    return v / n;
}

// Positive semi-definite Hessian over vec(F), 6x6 row-major.  Stretch: per column the exact
// Hessian in tension (E'/s on the orthogonal complement, E'' along the column), the
// along-column term 2 f^ f^T alone in compression.  Shear: the PSD form of the I6^2
// Hessian from its eigensystem: |I6| (T - Tq Tq^T / |Tq|^2) + l0 q q^T (times 2), with
// l0 = (I2 + sqrt(I2^2 + 12 I6^2)) / 2, q the unit eigenvector for l0 and T = (I + sgn S)/2,
// S swapping the two columns.
__device__ inline void bw_hessian(d3 fu, d3 fv, double st, double sh, double k, double H[36]) {
#pragma unroll
    for (int i = 0; i < 36; ++i) H[i] = 0.0;
    const double I5u = dot(fu, fu), I5v = dot(fv, fv);
    const double su = sqrt(I5u), sv = sqrt(I5v);
    const double isou = su > 1 ? 2 * (((su - 1) * (3 * su * k - 3 * k + 2)) / (2 * su)) : 0.0;
    const double isov = sv > 1 ? 2 * (((sv - 1) * (3 * sv * k - 3 * k + 2)) / (2 * sv)) : 0.0;
    const d3 nu = unit_or_same(fu), nv = unit_or_same(fv);
    const double uC = (su > 1.0) ? (3 * I5u * k - 3 * k + 2) / su : 2.0;
    const double vC = (sv > 1.0) ? (3 * I5v * k - 3 * k + 2) / sv : 2.0;
    const double un[3] = {nu.x, nu.y, nu.z}, vn[3] = {nv.x, nv.y, nv.z};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            H[6 * i + j] = st * ((i == j ? isou : 0.0) + uC * (un[i] * un[j]));
            H[6 * (3 + i) + 3 + j] = st * ((i == j ? isov : 0.0) + vC * (vn[i] * vn[j]));
        }
    const double I6 = dot(fu, fv);
    const double sg = (I6 >= 0) ? 1.0 : -1.0;
    const double g[6] = {fv.x, fv.y, fv.z, fu.x, fu.y, fu.z};
    const double I2 = I5u + I5v;
    const double l0 = 0.5 * (I2 + sqrt(I2 * I2 + 12.0 * I6 * I6));
    double q[6], qn = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) { q[i] = I6 * g[(i + 3) % 6] + l0 * g[i]; qn += q[i] * q[i]; }
    if (qn > 0) {
        qn = sqrt(qn);
#pragma unroll
        for (int i = 0; i < 6; ++i) q[i] /= qn;
    }
    double Tq[6], tn = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) { Tq[i] = 0.5 * (q[i] + sg * q[(i + 3) % 6]); tn += Tq[i] * Tq[i]; }
    const double aI6 = fabs(I6);
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const double T = 0.5 * ((i == j ? 1.0 : 0.0) + (j == (i + 3) % 6 ? sg : 0.0));
            H[6 * i + j] += sh * ((aI6 * (T - (Tq[i] * Tq[j]) / tn) + l0 * (q[i] * q[j])) * 2);
        }
}

// Energy density (before the area): stretch (s-1)^2 + [s > 1] k (s-1)^3 per column,
// shear I6^2.
__device__ inline double bw_energy(d3 fu, d3 fv, double st, double sh, double k) {
    const double I6 = dot(fu, fv);
    const double nu = sqrt(sqn(fu)), nv = sqrt(sqn(fv));
    const double cu = nu <= 1 ? 0.0 : 1.0, cv = nv <= 1 ? 0.0 : 1.0;
    const double du = nu - 1, dv = nv - 1;
    const double stretch = du * du + cu * k * (du * du * du) + dv * dv + cv * k * (dv * dv * dv);
    return st * stretch + sh * (I6 * I6);
}

// ---------------------------------------------------------------------------------
// Hinge angle.  n0 = (q0-q2) x (q1-q2), n1 = (q1-q3) x (q0-q3), e = q1 - q0;
// theta = 2 atan2((n0 x n1).e / |e|, n0.n1 + |n0||n1|).  The 12x12 second derivative is
// composed from the angle's derivatives in (n0, n1, e) and the cross-product maps of the
// normals; only blocks (a, b) with a <= b are formed (the matrix is symmetric).
// ---------------------------------------------------------------------------------
struct M3x {
    double m[3][3];
};
__device__ inline M3x crossm(d3 v) { return M3x{{{0, -v.z, v.y}, {v.z, 0, -v.x}, {-v.y, v.x, 0}}}; }
__device__ inline M3x mmul(const M3x& A, const M3x& B) {
    M3x R;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) R.m[i][j] = A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j] + A.m[i][2] * B.m[2][j];
    return R;
}
__device__ inline M3x mtmul(const M3x& A, const M3x& B) {   // A^T B
    M3x R;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) R.m[i][j] = A.m[0][i] * B.m[0][j] + A.m[1][i] * B.m[1][j] + A.m[2][i] * B.m[2][j];
    return R;
}
__device__ inline d3 rowmul(d3 r, const M3x& M) {           // r^T M
    return d3{r.x * M.m[0][0] + r.y * M.m[1][0] + r.z * M.m[2][0], r.x * M.m[0][1] + r.y * M.m[1][1] + r.z * M.m[2][1],
              r.x * M.m[0][2] + r.y * M.m[1][2] + r.z * M.m[2][2]};
}

// Angle and gradient (12).
__device__ inline double hinge_angle_grad(const d3 q[4], double g[12]) {
    const d3 v = cross(q[0] - q[2], q[1] - q[2]), w = cross(q[1] - q[3], q[0] - q[3]), ax = q[1] - q[0];
    const double an = sqrt(sqn(ax));
    const double th = 2.0 * atan2(dot(cross(v, w), ax) / an, dot(v, w) + sqrt(sqn(v)) * sqrt(sqn(w)));
    if (g) {
        const double vv = sqn(v), ww = sqn(w);
        const d3 a = -cross(ax, v) / vv / an, b = cross(ax, w) / ww / an;
        const d3 g0 = rowmul(a, crossm(q[2] - q[1])) + rowmul(b, crossm(q[1] - q[3]));
        const d3 g1 = rowmul(a, crossm(q[0] - q[2])) + rowmul(b, crossm(q[3] - q[0]));
        const d3 g2 = rowmul(a, crossm(q[1] - q[0]));
        const d3 g3 = rowmul(b, crossm(q[0] - q[1]));
        const d3 gs[4] = {g0, g1, g2, g3};
#pragma unroll
        for (int c = 0; c < 4; ++c) { g[3 * c] = gs[c].x; g[3 * c + 1] = gs[c].y; g[3 * c + 2] = gs[c].z; }
    }
    return th;
}

// Upper blocks B[a][b] (a <= b) of the angle's second derivative.  Block index of a pair
// (a, b) with a <= b: kUB[a][b].
__device__ constexpr int ub_index(int a, int b) { return a * 4 - a * (a - 1) / 2 + (b - a); }

__device__ inline void hinge_angle_hess(const d3 q[4], M3x B[10]) {
#pragma unroll
    for (int k = 0; k < 10; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) B[k].m[i][j] = 0.0;
    const d3 v = cross(q[0] - q[2], q[1] - q[2]), w = cross(q[1] - q[3], q[0] - q[3]), ax = q[1] - q[0];
    const double an = sqrt(sqn(ax)), vv = sqn(v), ww = sqn(w);
    const d3 av = cross(ax, v), aw = cross(ax, w);
    const M3x cax = crossm(ax);
    M3x A00, A11, dah;
    const double va[3] = {v.x, v.y, v.z}, wa[3] = {w.x, w.y, w.z}, ava[3] = {av.x, av.y, av.z}, awa[3] = {aw.x, aw.y, aw.z};
    const double axa[3] = {ax.x, ax.y, ax.z};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            A00.m[i][j] = 2.0 * ava[i] * va[j] / vv / vv / an + -cax.m[i][j] / vv / an;
            A11.m[i][j] = -2.0 * awa[i] * wa[j] / ww / ww / an + cax.m[i][j] / ww / an;
            dah.m[i][j] = (i == j ? 1.0 : 0.0) / an - axa[i] * axa[j] / an / an / an;
        }
    M3x A02 = mmul(crossm(v), dah), A12 = mmul(crossm(w), dah);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) { A02.m[i][j] /= vv; A12.m[i][j] = -A12.m[i][j] / ww; }
    // normal maps: d n0 / d q_{vi[k]} = vq[k], d n1 / d q_{wi[k]} = wq[k]
    const M3x vq[3] = {crossm(q[0] - q[2]), crossm(q[1] - q[0]), crossm(q[2] - q[1])};
    const M3x wq[3] = {crossm(q[0] - q[1]), crossm(q[1] - q[3]), crossm(q[3] - q[0])};
    constexpr int vi[3] = {1, 2, 0}, wi[3] = {3, 0, 1};
    auto acc = [&](int a, int b, const M3x& M, double s) {
        M3x& T = B[ub_index(a, b)];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) T.m[i][j] += s * M.m[i][j];
    };
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            if (vi[i] <= vi[j]) acc(vi[i], vi[j], mmul(mtmul(vq[i], A00), vq[j]), 1.0);
            if (wi[i] <= wi[j]) acc(wi[i], wi[j], mmul(mtmul(wq[i], A11), wq[j]), 1.0);
        }
        if (vi[i] <= 1) acc(vi[i], 1, mtmul(vq[i], A02), 1.0);
        if (vi[i] == 0) acc(0, 0, mtmul(vq[i], A02), -1.0);
        if (wi[i] <= 1) acc(wi[i], 1, mtmul(wq[i], A12), 1.0);
        if (wi[i] == 0) acc(0, 0, mtmul(wq[i], A12), -1.0);
    }
    const d3 a = -av / vv / an, b = aw / ww / an;
    const M3x D1 = crossm(a), D2 = crossm(b);
    acc(0, 1, D1, -1.0); acc(0, 2, D1, 1.0); acc(1, 2, D1, -1.0);
    acc(1, 3, D2, 1.0); acc(0, 1, D2, 1.0); acc(0, 3, D2, -1.0);
}

}  // namespace cipc
