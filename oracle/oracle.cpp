// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate expressions of CPU-IPC
// (MPL-2.0) in the reference's operation order; each such part cites the source it follows.
// ORACLE — test infrastructure only.  CPU restatement of CPU-IPC's per-Newton-iteration
// path, used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
// checker; never part of the product library.  Restates the behaviour of MPL-2.0
// licensed CPU-IPC sources, cited per function (paths relative to the reference root,
// "CPU IPC/" omitted).  Built with -ffp-contract=off so that every expression rounds in
// the written order.
//
// Parallel structure follows the reference (per-element parallel loops, serial contact
// gradient loops) with OpenMP; every reduction and scatter is done in a fixed order, so
// the oracle is deterministic where the reference's TBB reductions are not.
#include "oracle.h"
#include "oracle_math.hpp"

#include <map>
#include <stdexcept>
#include <omp.h>

using namespace oracle;

namespace {

inline bool isFree(int t) { return t == 0; }                 // BoundaryConditions.h:30-33
inline bool isCollider(int t) { return t >= 2; }             // BoundaryConditions.h:35-38

using I4 = std::array<int, 4>;

struct Blocks {                                              // BHessian, ContactMechanics.h:8-23
    std::vector<std::array<double, 9>> h3;   std::vector<int> d1;
    std::vector<std::array<double, 36>> h6;  std::vector<std::array<int, 2>> d2;
    std::vector<std::array<double, 81>> h9;  std::vector<std::array<int, 3>> d3;
    std::vector<std::array<double, 144>> h12; std::vector<std::array<int, 4>> d4;
};

// ----------------------------------------------------------------------------------
// Axis-aligned boxes and the LBVH (LBVH.cpp:1-431)
// ----------------------------------------------------------------------------------
struct AABB {
    V3 lo{std::numeric_limits<double>::max(), std::numeric_limits<double>::max(), std::numeric_limits<double>::max()};
    V3 hi{std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest()};
    void add(const V3& p) { lo = vmin(lo, p); hi = vmax(hi, p); }           // LBVH.cpp:82-86
    void add(const AABB& b) { lo = vmin(lo, b.lo); hi = vmax(hi, b.hi); }   // LBVH.cpp:88-92
    V3 center() const { return 0.5 * (lo + hi); }                           // LBVH.cpp:94-97
    // LBVH.cpp:157-161: touching counts as overlapping.
    bool overlaps(const AABB& b, double gap) const {
        return !((b.lo.x - hi.x > gap) || (b.lo.y - hi.y > gap) || (b.lo.z - hi.z > gap) ||
                 (lo.x - b.hi.x > gap) || (lo.y - b.hi.y > gap) || (lo.z - b.hi.z > gap));
    }
};
AABB pointBox(const V3& p) { AABB b; b.add(p); return b; }
AABB edgeBox(const V3& a, const V3& c) { AABB b; b.add(a); b.add(c); return b; }
AABB triBox(const V3& a, const V3& c, const V3& d) { AABB b; b.add(a); b.add(c); b.add(d); return b; }
// LBVH.cpp:120-129,400-431: swept boxes contain x and x - alpha * p.
AABB sweptPoint(const V3& p, const V3& dir, double a) { AABB b; b.add(p); b.add(p - a * dir); return b; }
AABB sweptEdge(const V3& p0, const V3& d0, const V3& p1, const V3& d1, double a) {
    AABB b = sweptPoint(p0, d0, a); b.add(sweptPoint(p1, d1, a)); return b;
}
AABB sweptTri(const V3& p0, const V3& d0, const V3& p1, const V3& d1, const V3& p2, const V3& d2, double a) {
    AABB b; b.add(sweptPoint(p0, d0, a)); b.add(sweptPoint(p1, d1, a)); b.add(sweptPoint(p2, d2, a)); return b;
}

uint32_t expandBits(uint32_t v) {                            // LBVH.cpp:17-24
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
uint32_t morton(double x, double y, double z) {              // LBVH.cpp:26-37
    const double res = 1024.0;
    auto q = [&](double v) { v = std::max(0.0, std::min(v * res, res - 1.0)); return (uint32_t)v; };
    return (expandBits(q(x)) << 2) | (expandBits(q(y)) << 1) | expandBits(q(z));
}
int clz64(uint64_t v) { return v == 0 ? 64 : __builtin_clzll(v); }

struct LBVH {                                                // CPULinearBVH, LBVH.h:21-47
    struct Node { uint32_t parent = ~0u, left = ~0u, right = ~0u, prim = ~0u; };
    std::vector<Node> nodes; std::vector<AABB> boxes; std::vector<uint64_t> keys;

    int prefix(int a, int b) const {
        if (b < 0 || b >= (int)keys.size()) return -1;
        return clz64(keys[a] ^ keys[b]);
    }
    std::pair<uint32_t, uint32_t> range(uint32_t idx) const {   // LBVH.cpp:62-93
        const int n = (int)keys.size();
        if (idx == 0) return {0u, (uint32_t)(n - 1)};
        const int cur = (int)idx;
        const int lp = prefix(cur, cur - 1), rp = prefix(cur, cur + 1);
        const int dir = rp > lp ? 1 : -1;
        const int minp = std::min(lp, rp);
        int lmax = 2;
        while (prefix(cur, cur + dir * lmax) > minp) lmax <<= 1;
        int len = 0;
        for (int st = lmax >> 1; st > 0; st >>= 1)
            if (prefix(cur, cur + dir * (len + st)) > minp) len += st;
        const int other = cur + dir * len;
        return {(uint32_t)std::min(cur, other), (uint32_t)std::max(cur, other)};
    }
    uint32_t split(uint32_t first, uint32_t last) const {       // LBVH.cpp:95-119
        const uint64_t fc = keys[first], lc = keys[last];
        if (fc == lc) return (first + last) >> 1;
        const int common = clz64(fc ^ lc);
        uint32_t s = first, step = last - first;
        do {
            step = (step + 1) >> 1;
            const uint32_t cand = s + step;
            if (cand < last && clz64(fc ^ keys[cand]) > common) s = cand;
        } while (step > 1);
        return s;
    }
    void build(const std::vector<AABB>& prims) {                // LBVH.cpp:163-250
        const size_t n = prims.size();
        nodes.clear(); boxes.clear(); keys.clear();
        if (n == 0) return;
        AABB scene;
        for (const auto& b : prims) scene.add(b);
        const V3 ext = scene.hi - scene.lo;
        std::vector<std::pair<uint64_t, uint32_t>> ent(n);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; ++i) {
            const V3 off = prims[i].center() - scene.lo;
            const double x = ext.x > 0.0 ? off.x / ext.x : 0.0;
            const double y = ext.y > 0.0 ? off.y / ext.y : 0.0;
            const double z = ext.z > 0.0 ? off.z / ext.z : 0.0;
            ent[i] = {((uint64_t)morton(x, y, z) << 32) | (uint64_t)i, (uint32_t)i};
        }
        std::sort(ent.begin(), ent.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        nodes.assign(2 * n - 1, Node{});
        boxes.assign(2 * n - 1, AABB{});
        keys.resize(n);
        for (size_t s = 0; s < n; ++s) {
            keys[s] = ent[s].first;
            nodes[s + n - 1].prim = ent[s].second;
            boxes[s + n - 1] = prims[ent[s].second];
        }
        if (n == 1) return;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n - 1; ++i) {
            auto r = range((uint32_t)i);
            uint32_t sp = split(r.first, r.second);
            uint32_t l = sp, rr = sp + 1;
            if (sp == r.first) l += (uint32_t)(n - 1);
            if (sp + 1 == r.second) rr += (uint32_t)(n - 1);
            nodes[i].left = l; nodes[i].right = rr;
            nodes[l].parent = (uint32_t)i; nodes[rr].parent = (uint32_t)i;
        }
        // post-order internal boxes (LBVH.cpp:230-249)
        std::vector<std::pair<uint32_t, bool>> st;
        st.emplace_back(0u, false);
        while (!st.empty()) {
            auto c = st.back(); st.pop_back();
            const Node& nd = nodes[c.first];
            if (nd.prim != ~0u) continue;
            if (c.second) { boxes[c.first] = boxes[nd.left]; boxes[c.first].add(boxes[nd.right]); }
            else { st.emplace_back(c.first, true); st.emplace_back(nd.right, false); st.emplace_back(nd.left, false); }
        }
    }
    void query(const AABB& q, double gap, std::vector<int>& out) const {   // LBVH.cpp:252-282
        out.clear();
        if (nodes.empty()) return;
        uint32_t stack[128]; size_t sz = 0;
        stack[sz++] = 0;
        while (sz > 0) {
            const uint32_t ni = stack[--sz];
            if (!q.overlaps(boxes[ni], gap)) continue;
            const Node& nd = nodes[ni];
            if (nd.prim != ~0u) { out.push_back((int)nd.prim); continue; }
            if (sz + 2 > 128) throw std::runtime_error("CPU LBVH traversal stack overflow");
            stack[sz++] = nd.right; stack[sz++] = nd.left;
        }
        std::sort(out.begin(), out.end());
    }
};

// ----------------------------------------------------------------------------------
// The mesh state (mesh3D subset, SimulationMesh.h:37-129)
// ----------------------------------------------------------------------------------
struct Mesh {
    int V = 0, T = 0, F = 0, Vs = 0, Es = 0;
    std::vector<V3> x, vrest, vprev, xt;
    std::vector<double> mass;  std::vector<int> bt;
    std::vector<I4> tets; std::vector<M3> dminv; std::vector<double> vol;
    // cloth triangles and bending elements (SimulationMesh.h:17-35,79-113)
    int Tr = 0, Nb = 0, bmode = 0;
    std::vector<std::array<int, 3>> tris; std::vector<std::array<double, 4>> dmt; std::vector<double> area;
    double stretchS = 0, shearS = 0, srate = 0, plate = 0;
    std::vector<I4> bverts; std::vector<std::array<double, 16>> bq; std::vector<double> brest, bweight;
    // soft target-position penalty (SoftBoundaryCondition, BoundaryConditions.h:54-67)
    std::vector<int> softIdx; std::vector<V3> softT; double softW = 0;
    std::vector<std::array<int, 3>> faces; std::vector<int> sv; std::vector<std::pair<int, int>> se;
    double dt = 0, u = 0, r = 0, drag = 0, hhat = 0, fhat = 0, fric = 0;
    V3 gn{0, 1, 0}; double gd = -1;
    std::vector<I4> selfAS, eeAS; std::vector<std::pair<int, int>> eeIe, ccdAS; std::vector<int> envAS;
    // lagged friction state (Friction::initialize, Friction.cpp:81-156; SimulationMesh.h:86-93)
    std::vector<I4> selfLast; std::vector<double> selfLam;
    std::vector<std::array<double, 2>> distCoord; std::vector<std::array<V3, 2>> tanBasis;
    std::vector<int> envLast; std::vector<double> envLam;
    // close constraints recorded by postLineSearch (SimulationMesh.h; IPCSolver.cpp:844-866)
    std::vector<int> closeEnv; std::vector<double> closeEnvVal;
    std::vector<I4> closeSelf; std::vector<double> closeSelfVal;
    std::vector<V3> grad;
    std::vector<V3> gabs;      // per entry, the sum of |addends| of grad (its rounding scale)
    std::vector<int> outer, inner; std::vector<double> val, rhs;
    std::vector<double> vscale;    // per CSR entry, the summed magnitude of its contributing blocks
    LBVH faceTree, edgeTree;
    std::vector<AABB> sweptPointBoxes, faceBoxes, edgeBoxes;
    int threads = 0;
};

// ----------------------------------------------------------------------------------
// Stable neo-Hookean element (Elasticity.cpp)
// ----------------------------------------------------------------------------------
M3 deformation(const Mesh& m, int t) {                       // Elasticity.cpp:570-598 x Dm^-1
    const I4& id = m.tets[t];
    M3 Ds;
    for (int c = 0; c < 3; ++c) {
        const V3 e = m.x[id[c + 1]] - m.x[id[0]];
        Ds(0, c) = e.x; Ds(1, c) = e.y; Ds(2, c) = e.z;
    }
    return matmul(Ds, m.dminv[t]);
}
M3 cofactor(const M3& F) {                                   // pI3pF, Elasticity.cpp:1155-1165
    M3 c;
    c(0, 0) = F(1, 1) * F(2, 2) - F(1, 2) * F(2, 1);
    c(0, 1) = F(1, 2) * F(2, 0) - F(1, 0) * F(2, 2);
    c(0, 2) = F(1, 0) * F(2, 1) - F(1, 1) * F(2, 0);
    c(1, 0) = F(2, 1) * F(0, 2) - F(2, 2) * F(0, 1);
    c(1, 1) = F(2, 2) * F(0, 0) - F(2, 0) * F(0, 2);
    c(1, 2) = F(2, 0) * F(0, 1) - F(2, 1) * F(0, 0);
    c(2, 0) = F(0, 1) * F(1, 2) - F(1, 1) * F(0, 2);
    c(2, 1) = F(0, 2) * F(1, 0) - F(0, 0) * F(1, 2);
    c(2, 2) = F(0, 0) * F(1, 1) - F(0, 1) * F(1, 0);
    return c;
}
// computePEPF_StableNHK3D_2_double (Elasticity.cpp:1144-1170)
M3 snh_pk1(const M3& F, double u, double r) {
    M3 U, V; double s[3];
    svd3(F, U, s, V);
    const double I3 = s[0] * s[1] * s[2];
    const double I2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    const M3 c = cofactor(F);
    M3 P;
    const double a = u * (1 - 1 / (I2 + 1)), b = r * (I3 - 1 - u * 3 / (r * 4));
    for (int i = 0; i < 9; ++i) P.m[i] = a * F.m[i] + b * c.m[i];
    return P;
}
// Rounding scale of each entry of snh_pk1: the magnitudes of the terms a F and b cof(F)
// and of the cancellations inside b (r (I3 - 1) - 3u/4) and inside each cofactor.
void snh_pk1_scale(const M3& F, double u, double r, double ps[9]) {
    M3 U, V; double s[3];
    svd3(F, U, s, V);
    const double I3 = s[0] * s[1] * s[2];
    const double I2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    const M3 c = cofactor(F);
    const double a = u * (1 - 1 / (I2 + 1)), b = r * (I3 - 1 - u * 3 / (r * 4));
    const double bs = std::fabs(r) * (std::fabs(I3) + 1) + 0.75 * std::fabs(u);
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) {
            const int r1 = (row + 1) % 3, r2 = (row + 2) % 3, c1 = (col + 1) % 3, c2 = (col + 2) % 3;
            const double cm = std::fabs(F(r1, c1) * F(r2, c2)) + std::fabs(F(r1, c2) * F(r2, c1));
            const int k = row + 3 * col;
            ps[k] = std::fabs(a * F.m[k]) + bs * std::fabs(c.m[k]) + std::fabs(b) * cm;
        }
}
// project_StabbleNHK_2_H_3D active branch (Elasticity.cpp:1337-1345,1436-1477), then makePD<9>.
// H9 row-major 9x9 over column-major vec(F).
void snh_hessian(const M3& F, double u, double r, double* H) {
    M3 U, V; double s[3];
    svd3(F, U, s, V);
    const double I3 = s[0] * s[1] * s[2];
    const double I2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    const double mu = u, lambda = r, Ic = I2, J = I3;
    double g1[9], gJ[9];
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 3; ++i) g1[3 * c + i] = 2 * F(i, c);
    const V3 f0(F(0, 0), F(1, 0), F(2, 0)), f1(F(0, 1), F(1, 1), F(2, 1)), f2(F(0, 2), F(1, 2), F(2, 2));
    const V3 c0 = cross(f1, f2), c1 = cross(f2, f0), c2 = cross(f0, f1);
    for (int i = 0; i < 3; ++i) { gJ[i] = c0[i]; gJ[3 + i] = c1[i]; gJ[6 + i] = c2[i]; }
    auto hat = [](const V3& f, double h[3][3]) {
        h[0][0] = 0; h[0][1] = -f.z; h[0][2] = f.y;
        h[1][0] = f.z; h[1][1] = 0; h[1][2] = -f.x;
        h[2][0] = -f.y; h[2][1] = f.x; h[2][2] = 0;
    };
    double fh[3][3][3];
    hat(f0, fh[0]); hat(f1, fh[1]); hat(f2, fh[2]);
    double HJ[9][9] = {};
    // blocks (0,1)=-f2hat (0,2)=f1hat (1,0)=f2hat (1,2)=-f0hat (2,0)=-f1hat (2,1)=f0hat
    const int bk[6][2] = {{0, 1}, {0, 2}, {1, 0}, {1, 2}, {2, 0}, {2, 1}};
    const int bf[6] = {2, 1, 2, 0, 1, 0};
    const double bs[6] = {-1, 1, 1, -1, -1, 1};
    for (int q = 0; q < 6; ++q)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) HJ[3 * bk[q][0] + i][3 * bk[q][1] + j] = bs[q] * fh[bf[q]][i][j];
    const double a1 = (Ic * mu) / (2 * (Ic + 1));
    const double a2 = lambda * (J - 1 - (3 * mu) / (4.0 * lambda));
    const double a3 = mu / (2 * (Ic + 1) * (Ic + 1));
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j)
            H[9 * i + j] = a1 * (i == j ? 2.0 : 0.0) + a2 * HJ[i][j] + a3 * g1[i] * g1[j] + lambda * gJ[i] * gJ[j];
    make_pd<9>(H);
}
// getObjEnergy_StableNHK2_3D per-element term (Elasticity.cpp:695-727), without the volume.
double snh_energy_density(const M3& F, double u, double r) {
    M3 U, V; double s[3];
    svd3(F, U, s, V);
    M3 S;   // V diag(s) V^T
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) S(i, j) = V(i, 0) * s[0] * V(j, 0) + V(i, 1) * s[1] * V(j, 1) + V(i, 2) * s[2] * V(j, 2);
    const M3 S2 = matmul(S, S);
    const double I2 = S2(0, 0) + S2(1, 1) + S2(2, 2);
    const double I3 = S(0, 0) * (S(1, 1) * S(2, 2) - S(1, 2) * S(2, 1)) - S(0, 1) * (S(1, 0) * S(2, 2) - S(1, 2) * S(2, 0))
                    + S(0, 2) * (S(1, 0) * S(2, 1) - S(1, 1) * S(2, 0));
    const double k = I3 - 1 - 3 * u / 4 / r;
    return 0.5 * u * (I2 - 3) + 0.5 * r * k * k - 0.5 * u * std::log(I2 + 1);
}
// computePFPX3D_double (Elasticity.cpp:803-848): row 3c+i of vec(F) against 3a+i of x.
void pfpx(const M3& Di, double P[9][12]) {
    for (int i = 0; i < 9; ++i) for (int j = 0; j < 12; ++j) P[i][j] = 0;
    for (int c = 0; c < 3; ++c) {
        const double t = -(Di(0, c) + Di(1, c) + Di(2, c));
        for (int i = 0; i < 3; ++i) {
            P[3 * c + i][i] = t;
            for (int a = 1; a < 4; ++a) P[3 * c + i][3 * a + i] = Di(a - 1, c);
        }
    }
}

// ----------------------------------------------------------------------------------
// Cloth: Baraff-Witkin triangles (Elasticity.cpp) and bending (Simulator.cpp:48-109,
// HingeBending.cpp, Elasticity.cpp:12-150).  A 3x2 deformation gradient is stored
// column-major: f[0..2] = F a (a = (1,0)), f[3..5] = F b (b = (0,1)).
// ----------------------------------------------------------------------------------
V3 normalized(const V3& v) {   // Eigen 3.4 normalized(): v / |v|, unchanged when |v| = 0
    const double n2 = sqn(v);
    if (!(n2 > 0.0)) return v;
    const double n = std::sqrt(n2);
    return V3(v.x / n, v.y / n, v.z / n);
}
// calculateDs32D_double x Dm^-1 (Elasticity.cpp:601-607; IPCSolver.cpp:169-170)
void tri_deformation(const Mesh& m, int t, double f[6]) {
    const auto& id = m.tris[t];
    const V3 e1 = m.x[id[1]] - m.x[id[0]], e2 = m.x[id[2]] - m.x[id[0]];
    const auto& D = m.dmt[t];   // column-major 2x2
    for (int j = 0; j < 2; ++j) {
        const V3 c = D[2 * j] * e1 + D[2 * j + 1] * e2;
        f[3 * j] = c.x; f[3 * j + 1] = c.y; f[3 * j + 2] = c.z;
    }
}
// computePFPX32D_double (Elasticity.cpp:851-897): row 3j+i of vec(F) against 3c+i of x.
void tri_pfpx(const std::array<double, 4>& D, double P[6][9]) {
    for (int i = 0; i < 6; ++i) for (int j = 0; j < 9; ++j) P[i][j] = 0;
    const double s0 = D[0] + D[1], s1 = D[2] + D[3];
    for (int i = 0; i < 3; ++i) {
        P[i][i] = -s0; P[3 + i][i] = -s1;
        P[i][3 + i] = D[0]; P[3 + i][3 + i] = D[2];
        P[i][6 + i] = D[1]; P[3 + i][6 + i] = D[3];
    }
}
// computePEPF_baraffwitkin_double (Elasticity.cpp:1173-1205)
void bw_pk1(const double f[6], double stretchS, double shearS, double k, double P[6]) {
    const V3 fu(f[0], f[1], f[2]), fv(f[3], f[4], f[5]);
    const double I6 = dot(fu, fv);
    const double I5u = dot(fu, fu), I5v = dot(fv, fv);
    double uc = 1.0 - 1 / std::sqrt(I5u), vc = 1.0 - 1 / std::sqrt(I5v);
    if (I5u > 1) uc += 1.5 * k * (std::sqrt(I5u) + 1 / std::sqrt(I5u) - 2);
    if (I5v > 1) vc += 1.5 * k * (std::sqrt(I5v) + 1 / std::sqrt(I5v) - 2);
    // shear: 2 (I6 - a.b)(F a b^T + F b a^T) = 2 I6 [F b, F a]; stretch: 2 c_u F a a^T + 2 c_v F b b^T
    for (int i = 0; i < 3; ++i) {
        P[i] = stretchS * (uc * 2. * fu[i]) + shearS * (2 * I6 * fv[i]);
        P[3 + i] = stretchS * (vc * 2. * fv[i]) + shearS * (2 * I6 * fu[i]);
    }
}
// Rounding scale of each entry of bw_pk1: the stretch coefficients' terms (1 and 1/|f|,
// and the strain-rate terms) and the shear term's products.
void bw_pk1_scale(const double f[6], double stretchS, double shearS, double k, double ps[6]) {
    const V3 fu(f[0], f[1], f[2]), fv(f[3], f[4], f[5]);
    const double su = std::sqrt(dot(fu, fu)), sv = std::sqrt(dot(fv, fv));
    const double cu = 1 + 1 / su + (su > 1 ? 1.5 * std::fabs(k) * (su + 1 / su + 2) : 0.0);
    const double cv = 1 + 1 / sv + (sv > 1 ? 1.5 * std::fabs(k) * (sv + 1 / sv + 2) : 0.0);
    const double i6 = std::fabs(fu.x * fv.x) + std::fabs(fu.y * fv.y) + std::fabs(fu.z * fv.z);
    for (int i = 0; i < 3; ++i) {
        ps[i] = std::fabs(stretchS) * 2 * cu * std::fabs(fu[i]) + std::fabs(shearS) * 2 * i6 * std::fabs(fv[i]);
        ps[3 + i] = std::fabs(stretchS) * 2 * cv * std::fabs(fv[i]) + std::fabs(shearS) * 2 * i6 * std::fabs(fu[i]);
    }
}
// project_baraffwitkint_H_3D (Elasticity.cpp:1481-1544): 6x6 row-major over vec(F).
void bw_hessian(const double f[6], double stretchS, double shearS, double k, double Hout[36]) {
    const V3 fu(f[0], f[1], f[2]), fv(f[3], f[4], f[5]);
    double H[36] = {};
    const double I5u = dot(fu, fu), I5v = dot(fv, fv);
    const double su = std::sqrt(I5u), sv = std::sqrt(I5v);
    if (su > 1) { const double h = 2 * (((su - 1) * (3 * su * k - 3 * k + 2)) / (2 * su)); for (int i = 0; i < 3; ++i) H[7 * i] = h; }
    if (sv > 1) { const double h = 2 * (((sv - 1) * (3 * sv * k - 3 * k + 2)) / (2 * sv)); for (int i = 3; i < 6; ++i) H[7 * i] = h; }
    const V3 nu = normalized(fu), nv = normalized(fv);
    const double uC = (su > 1.0) ? (3 * I5u * k - 3 * k + 2) / su : 2.0;
    const double vC = (sv > 1.0) ? (3 * I5v * k - 3 * k + 2) / sv : 2.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { H[6 * i + j] += uC * (nu[i] * nu[j]); H[6 * (3 + i) + 3 + j] += vC * (nv[i] * nv[j]); }
    for (int i = 0; i < 36; ++i) Hout[i] = stretchS * H[i];
    // shear: analytic PSD form from the eigensystem of the I6^2 Hessian
    const double I6 = dot(fu, fv);
    const double sgn = (I6 >= 0) ? 1.0 : -1.0;
    double g[6] = {fv.x, fv.y, fv.z, fu.x, fu.y, fu.z};           // vec(F (a b^T + b a^T))
    const double I2 = I5u + I5v;
    const double l0 = 0.5 * (I2 + std::sqrt(I2 * I2 + 12.0 * I6 * I6));
    double q[6], qn = 0;
    for (int i = 0; i < 6; ++i) { q[i] = I6 * g[(i + 3) % 6] + l0 * g[i]; qn += q[i] * q[i]; }
    if (qn > 0) { qn = std::sqrt(qn); for (int i = 0; i < 6; ++i) q[i] /= qn; }
    double T[36], Tq[6], tn = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) T[6 * i + j] = 0.5 * ((i == j ? 1.0 : 0.0) + sgn * ((j == (i + 3) % 6) ? 1.0 : 0.0));
    for (int i = 0; i < 6; ++i) { double a = 0; for (int j = 0; j < 6; ++j) a += T[6 * i + j] * q[j]; Tq[i] = a; tn += a * a; }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            const double hs = std::fabs(I6) * (T[6 * i + j] - (Tq[i] * Tq[j]) / tn) + l0 * (q[i] * q[j]);
            Hout[6 * i + j] += shearS * (hs * 2);
        }
}
// getObjEnergy_baraffwitkin_3D per element (Elasticity.cpp:729-771), without the area.
double bw_energy_density(const double f[6], double stretchS, double shearS, double k) {
    const V3 fu(f[0], f[1], f[2]), fv(f[3], f[4], f[5]);
    const double I6 = dot(fu, fv);
    const double sh = I6 * I6;
    const double nu = std::sqrt(sqn(fu)), nv = std::sqrt(sqn(fv));
    const double cu = nu <= 1 ? 0.0 : 1.0, cv = nv <= 1 ? 0.0 : 1.0;
    const double st = std::pow(nu - 1, 2) + cu * k * std::pow(nu - 1, 3) + std::pow(nv - 1, 2) + cv * k * std::pow(nv - 1, 3);
    return stretchS * st + shearS * sh;
}

// LibShell::angle (Elasticity.cpp:28-58): angle of v to w about axis, derivative over
// (v, w, axis) and the second-derivative blocks the reference fills ((0,0), (3,3),
// (0,6), (3,6); the (6,0) and (6,3) blocks stay zero there).
double shell_angle(const V3& v, const V3& w, const V3& axis, double* der, double (*hes)[9]) {
    const double an = std::sqrt(sqn(axis));
    const double theta = 2.0 * std::atan2(dot(cross(v, w), axis) / an, dot(v, w) + std::sqrt(sqn(v)) * std::sqrt(sqn(w)));
    if (der) {
        const V3 a = cross(axis, v), b = cross(axis, w);
        const double vv = sqn(v), ww = sqn(w);
        for (int i = 0; i < 3; ++i) { der[i] = -a[i] / vv / an; der[3 + i] = b[i] / ww / an; der[6 + i] = 0.0; }
    }
    if (hes) {
        for (int i = 0; i < 9; ++i) for (int j = 0; j < 9; ++j) hes[i][j] = 0.0;
        const double vv = sqn(v), ww = sqn(w);
        const V3 av = cross(axis, v), aw = cross(axis, w);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                hes[i][j] += 2.0 * av[i] * v[j] / vv / vv / an;
                hes[3 + i][3 + j] += -2.0 * aw[i] * w[j] / ww / ww / an;
            }
        const double cx[3][3] = {{0, -axis.z, axis.y}, {axis.z, 0, -axis.x}, {-axis.y, axis.x, 0}};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) { hes[i][j] += -cx[i][j] / vv / an; hes[3 + i][3 + j] += cx[i][j] / ww / an; }
        double dah[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) dah[i][j] = (i == j ? 1.0 : 0.0) / an - axis[i] * axis[j] / an / an / an;
        const double cv[3][3] = {{0, -v.z, v.y}, {v.z, 0, -v.x}, {-v.y, v.x, 0}};
        const double cw[3][3] = {{0, -w.z, w.y}, {w.z, 0, -w.x}, {-w.y, w.x, 0}};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double a = 0, b = 0;
                for (int k = 0; k < 3; ++k) { a += cv[i][k] * dah[k][j]; b += cw[i][k] * dah[k][j]; }
                hes[i][6 + j] += a / vv;
                hes[3 + i][6 + j] += -b / ww;
            }
    }
    return theta;
}
// edgeTheta (Elasticity.cpp:59-150): dihedral angle of the hinge (q0, q1; q2, q3), its
// 12-gradient and its 12x12 second-derivative matrix (row-major) as the reference
// assembles them from shell_angle's blocks.
double edge_theta(const V3& q0, const V3& q1, const V3& q2, const V3& q3, double* der, double* hess) {
    const V3 n0 = cross(q0 - q2, q1 - q2), n1 = cross(q1 - q3, q0 - q3), axis = q1 - q0;
    double ad[9], ah[9][9];
    const double theta = shell_angle(n0, n1, axis, (der || hess) ? ad : nullptr, hess ? ah : nullptr);
    auto cm = [](const V3& v, double o[3][3]) {
        o[0][0] = 0; o[0][1] = -v.z; o[0][2] = v.y; o[1][0] = v.z; o[1][1] = 0; o[1][2] = -v.x; o[2][0] = -v.y; o[2][1] = v.x; o[2][2] = 0;
    };
    // row vector r times cross matrix of e: (r [e]x)_j = sum_i r_i [e]x(i, j)
    auto rowcm = [&](const double* r, const V3& e, double* out) {
        double c[3][3]; cm(e, c);
        for (int j = 0; j < 3; ++j) out[j] += r[0] * c[0][j] + r[1] * c[1][j] + r[2] * c[2][j];
    };
    if (der) {
        for (int i = 0; i < 12; ++i) der[i] = 0.0;
        rowcm(ad, q2 - q1, der + 0); rowcm(ad, q0 - q2, der + 3); rowcm(ad, q1 - q0, der + 6);
        rowcm(ad + 3, q1 - q3, der + 0); rowcm(ad + 3, q3 - q0, der + 3); rowcm(ad + 3, q0 - q1, der + 9);
    }
    if (hess) {
        for (int i = 0; i < 144; ++i) hess[i] = 0.0;
        double vq[3][3][3], wq[3][3][3];
        cm(q0 - q2, vq[0]); cm(q1 - q0, vq[1]); cm(q2 - q1, vq[2]);
        cm(q0 - q1, wq[0]); cm(q1 - q3, wq[1]); cm(q3 - q0, wq[2]);
        const int vi[3] = {3, 6, 0}, wi[3] = {9, 0, 3};
        // blk(r0, c0) += A^T B C  /  A B  /  B C patterns over 3x3 blocks
        auto blk = [&](int r0, int c0, const double (*L)[3], bool LT, int br, int bc, const double (*R)[3], double sg) {
            // L^T (or L) * ah[br.., bc..] * R (R may be null = identity)
            double t[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double a = 0;
                    for (int k = 0; k < 3; ++k) a += (L ? (LT ? L[k][i] : L[i][k]) : (i == k ? 1.0 : 0.0)) * ah[br + k][bc + j];
                    t[i][j] = a;
                }
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double a = 0;
                    for (int k = 0; k < 3; ++k) a += t[i][k] * (R ? R[k][j] : (k == j ? 1.0 : 0.0));
                    hess[12 * (r0 + i) + c0 + j] += sg * a;
                }
        };
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) {
                blk(vi[i], vi[j], vq[i], true, 0, 0, vq[j], 1.0);
                blk(vi[i], wi[j], vq[i], true, 0, 3, wq[j], 1.0);
                blk(wi[i], vi[j], wq[i], true, 3, 0, vq[j], 1.0);
                blk(wi[i], wi[j], wq[i], true, 3, 3, wq[j], 1.0);
            }
            blk(vi[i], 3, vq[i], true, 0, 6, nullptr, 1.0);
            blk(3, vi[i], nullptr, false, 6, 0, vq[i], 1.0);
            blk(vi[i], 0, vq[i], true, 0, 6, nullptr, -1.0);
            blk(0, vi[i], nullptr, false, 6, 0, vq[i], -1.0);
            blk(wi[i], 3, wq[i], true, 3, 6, nullptr, 1.0);
            blk(3, wi[i], nullptr, false, 6, 3, wq[i], 1.0);
            blk(wi[i], 0, wq[i], true, 3, 6, nullptr, -1.0);
            blk(0, wi[i], nullptr, false, 6, 3, wq[i], -1.0);
        }
        double d1[3][3], d2[3][3];
        cm(V3(ad[0], ad[1], ad[2]), d1); cm(V3(ad[3], ad[4], ad[5]), d2);
        auto add = [&](int r0, int c0, const double (*M)[3], double sg) {
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) hess[12 * (r0 + i) + c0 + j] += sg * M[i][j];
        };
        add(6, 3, d1, 1); add(0, 3, d1, -1); add(0, 6, d1, 1); add(3, 0, d1, 1); add(3, 6, d1, -1); add(6, 0, d1, -1);
        add(9, 0, d2, 1); add(3, 0, d2, -1); add(3, 9, d2, 1); add(0, 3, d2, 1); add(0, 9, d2, -1); add(9, 3, d2, -1);
    }
    return theta;
}
// remainder(angle - rest, 2 pi) (HingeBending.cpp:38-42)
double hinge_difference(double angle, double rest) { return std::remainder(angle - rest, 6.28318530717958647692); }

// ----------------------------------------------------------------------------------
// Barrier, distances and their classification (ContactMechanics.cpp)
// ----------------------------------------------------------------------------------
void barrier_b(double d, double dh, double& b) { b = -(d - dh) * (d - dh) * std::log(d / dh); }  // :775-778
void barrier_g(double d, double dh, double& g) {                                                 // :780-784
    const double t = d - dh;
    g = t * std::log(d / dh) * -2.0 - (t * t) / d;
}
void barrier_h(double d, double dh, double& h) {                                                 // :786-790
    const double t = d - dh;
    h = (std::log(d / dh) * -2.0 - t * 4.0 / d) + 1.0 / (d * d) * (t * t);
}

double d_PP(const V3& a, const V3& b) { return sqn(a - b); }                                      // :282-287
double d_PE(const V3& p, const V3& e0, const V3& e1) { return sqn(cross(e0 - p, e1 - p)) / sqn(e1 - e0); } // :807-813
double d_PT(const V3& p, const V3& t0, const V3& t1, const V3& t2) {                             // :298-306
    const V3 b = cross(t1 - t0, t2 - t0);
    const double aTb = dot(p - t0, b);
    return aTb * aTb / sqn(b);
}
double d_EE(const V3& a0, const V3& a1, const V3& b0, const V3& b1) {                            // :764-773
    const V3 b = cross(a1 - a0, b1 - b0);
    const double aTb = dot(b0 - a0, b);
    return aTb * aTb / sqn(b);
}
double ee_cross_sqn(const V3& a0, const V3& a1, const V3& b0, const V3& b1) { return sqn(cross(a1 - a0, b1 - b0)); } // :308-316

// Eigen::Matrix2d::ldlt().solve(b) for a symmetric 2x2 (Eigen 3.4 LDLT: diagonal pivoting
// on the larger |diagonal|, unit-lower L, pseudo-inverse of D below DBL_MIN).
void ldlt2_solve(double a00, double a10, double a11, double b0, double b1, double& x0, double& x1) {
    bool swapped = std::fabs(a11) > std::fabs(a00);
    double d0 = swapped ? a11 : a00, d1 = swapped ? a00 : a11, l = a10;
    bool valid0 = std::fabs(d0) > 0.0;
    if (!valid0) {
        swapped = false;  // all-zero diagonal: transpositions reset to identity, L untouched
        d0 = a00; d1 = a11;
    } else {
        l = l / d0;
        const double tmp = d0 * l;
        d1 = d1 - l * tmp;
    }
    double y0 = swapped ? b1 : b0, y1 = swapped ? b0 : b1;
    y1 = y1 - l * y0;
    const double tol = std::numeric_limits<double>::min();
    y0 = std::fabs(d0) > tol ? y0 / d0 : 0.0;
    y1 = std::fabs(d1) > tol ? y1 / d1 : 0.0;
    y0 = y0 - l * y1;
    if (swapped) { x0 = y1; x1 = y0; } else { x0 = y0; x1 = y1; }
}
// One projection of dType_PT: basis rows r0 = e, r1 = e x n; returns the 2 parameters.
void pt_param(const V3& r0, const V3& nvec, const V3& rhs, double& p0, double& p1) {
    const V3 r1 = cross(r0, nvec);
    const double a00 = dot(r0, r0), a10 = dot(r1, r0), a11 = dot(r1, r1);
    ldlt2_solve(a00, a10, a11, dot(r0, rhs), dot(r1, rhs), p0, p1);
}
// dType_PT (ContactMechanics.cpp:143-193): v0 the point, v1 v2 v3 the triangle.
int dType_PT(const V3& v0, const V3& v1, const V3& v2, const V3& v3) {
    const V3 n = cross(v2 - v1, v3 - v1);
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
// dType_EE (ContactMechanics.cpp:195-280) and edge_edge_distance_type (AdditiveCCD.cpp:64-139).
int dType_EE(const V3& v0, const V3& v1, const V3& v2, const V3& v3) {
    const V3 u = v1 - v0, v = v3 - v2, w = v0 - v2;
    const double a = sqn(u), b = dot(u, v), c = sqn(v), d = dot(u, w), e = dot(v, w);
    const double D = a * c - b * b;
    double tD = D, sN, tN;
    int def = 8;
    sN = (b * e - c * d);
    if (sN <= 0.0) { tN = e; tD = c; def = 2; }
    else if (sN >= D) { tN = e + b; tD = c; def = 5; }
    else {
        tN = (a * e - b * d);
        const V3 uv = cross(u, v);
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

// Exact distance derivatives through second-order AD (replaces the generated
// g_PE/H_PE/g_PT/H_PT/g_EE/H_EE, ContactMechanics.cpp:815-2467).
template <int N>
void dist_derivs(int type, const V3* p, double& d, double* g, double* H) {
    HV3<N> q[4];
    for (int k = 0; k < N / 3; ++k) q[k] = hvar<N>(p[k], k);
    HD<N> r;
    switch (type) {
    case 0: { HV3<N> e = hsub(q[0], q[1]); r = hdot(e, e); break; }
    case 1: { HV3<N> c = hcross(hsub(q[1], q[0]), hsub(q[2], q[0])); HV3<N> e = hsub(q[2], q[1]); r = hdot(c, c) / hdot(e, e); break; }
    case 2: { HV3<N> b = hcross(hsub(q[2], q[1]), hsub(q[3], q[1])); HD<N> t = hdot(hsub(q[0], q[1]), b); r = (t * t) / hdot(b, b); break; }
    case 3: { HV3<N> b = hcross(hsub(q[1], q[0]), hsub(q[3], q[2])); HD<N> t = hdot(hsub(q[2], q[0]), b); r = (t * t) / hdot(b, b); break; }
    default: { HV3<N> b = hcross(hsub(q[1], q[0]), hsub(q[3], q[2])); r = hdot(b, b); break; }
    }
    d = r.v;
    for (int i = 0; i < N; ++i) { if (g) g[i] = r.g[i]; if (H) for (int j = 0; j < N; ++j) H[i * N + j] = r.h[i][j]; }
}
void deriv_any(int type, const V3* p, double& d, double* g, double* H) {
    if (type == 0) dist_derivs<6>(0, p, d, g, H);
    else if (type == 1) dist_derivs<9>(1, p, d, g, H);
    else dist_derivs<12>(type, p, d, g, H);
}

double eps_x(const Mesh& m, int a0, int a1, int b0, int b1) {                                  // :758-762
    return 1e-3 * sqn(m.vrest[a0] - m.vrest[a1]) * sqn(m.vrest[b0] - m.vrest[b1]);
}
// mollifier q(c) and derivatives (ContactMechanics.cpp:678-693)
double mol_e(double c, double ex) { if (!(c < ex)) return 1.0; const double t = c / ex; return (-t + 2.0) * t; }
void mol_eg(const V3* p, double ex, double* g) {                                                // :711-730
    double c, gc[12];
    dist_derivs<12>(4, p, c, gc, nullptr);
    if (c < ex) {
        const double oe = 1.0 / ex, qg = 2.0 * oe * (-oe * c + 1.0);
        for (int i = 0; i < 12; ++i) g[i] = gc[i] * qg;
    } else for (int i = 0; i < 12; ++i) g[i] = 0;
}
void mol_eH(const V3* p, double ex, double* H) {                                                // :732-756
    double c, gc[12], Hc[144];
    dist_derivs<12>(4, p, c, gc, Hc);
    if (c < ex) {
        const double oe = 1.0 / ex, qg = 2.0 * oe * (-oe * c + 1.0), qH = -2.0 / (ex * ex);
        for (int i = 0; i < 12; ++i) for (int j = 0; j < 12; ++j) H[12 * i + j] = Hc[12 * i + j] * qg + (qH * gc[i]) * gc[j];
    } else for (int i = 0; i < 144; ++i) H[i] = 0;
}

// SelfConstraintVal (ContactMechanics.cpp:2478-2503)
double self_val(const Mesh& m, const I4& c) {
    if (c[0] >= 0) return d_EE(m.x[c[0]], m.x[c[1]], m.x[c[2]], m.x[c[3]]);
    const int v0 = -c[0] - 1;
    if (c[2] < 0) return d_PP(m.x[v0], m.x[c[1]]);
    if (c[3] < 0) return d_PE(m.x[v0], m.x[c[1]], m.x[c[2]]);
    return d_PT(m.x[v0], m.x[c[1]], m.x[c[2]], m.x[c[3]]);
}
// vertices and distance type of an encoded contact: type 0 PP 1 PE 2 PT 3 EE
int decode(const I4& c, int vid[4]) {
    if (c[0] >= 0) { for (int k = 0; k < 4; ++k) vid[k] = c[k]; return 3; }
    vid[0] = -c[0] - 1; vid[1] = c[1];
    if (c[2] < 0) return 0;
    vid[2] = c[2];
    if (c[3] < 0) return 1;
    vid[3] = c[3];
    return 2;
}

// ----------------------------------------------------------------------------------
// Lagged friction (Friction.cpp, FrictionKinematics.h with SFCLAMPING_ORDER 1)
// ----------------------------------------------------------------------------------
void f0_SF(double x2, double e, double& f0) { f0 = x2 * (-std::sqrt(x2) / 3.0 + e) / (e * e) + e / 3.0; }    // :262-265
void f1_SF_div(double x2, double e, double& r) { r = (-std::sqrt(x2) + 2.0 * e) / (e * e); }             // :267-270
void f2_SF(double x2, double e, double& f2) { f2 = 2.0 * (e - std::sqrt(x2)) / (e * e); }                 // :272-275

// Tangent bases (FrictionKinematics.h:18-28,81-91,155-164,211-227) and closest points
// (:30-41,93-114,166-172).  Columns of the 3x2 basis are b[0], b[1].
void tangent_PT(const V3& v0, const V3& v1, const V3& v2, const V3& v3, std::array<V3, 2>& b) {
    const V3 v12 = v2 - v1;
    b[0] = normalized(v12);
    b[1] = normalized(cross(cross(v12, v3 - v1), v12));
    (void)v0;
}
void closest_PT(const V3& v0, const V3& v1, const V3& v2, const V3& v3, double beta[2]) {
    const V3 r0 = v2 - v1, r1 = v3 - v1, q = v0 - v1;
    ldlt2_solve(dot(r0, r0), dot(r1, r0), dot(r1, r1), dot(r0, q), dot(r1, q), beta[0], beta[1]);
}
void tangent_EE(const V3& v0, const V3& v1, const V3& v2, const V3& v3, std::array<V3, 2>& b) {
    const V3 v01 = v1 - v0;
    b[0] = normalized(v01);
    b[1] = normalized(cross(cross(v01, v3 - v2), v01));
}
void closest_EE(const V3& v0, const V3& v1, const V3& v2, const V3& v3, double g[2]) {
    const V3 e20 = v0 - v2, e01 = v1 - v0, e23 = v3 - v2;
    const double c00 = sqn(e01), c01 = -dot(e23, e01), c11 = sqn(e23);
    ldlt2_solve(c00, c01, c11, -dot(e20, e01), dot(e20, e23), g[0], g[1]);
}
void tangent_PE(const V3& v0, const V3& v1, const V3& v2, std::array<V3, 2>& b) {
    const V3 v12 = v2 - v1;
    b[0] = normalized(v12);
    b[1] = normalized(cross(v12, v0 - v1));
}
double closest_PE(const V3& v0, const V3& v1, const V3& v2) { const V3 e12 = v2 - v1; return dot(v0 - v1, e12) / sqn(e12); }
void tangent_PP(const V3& v0, const V3& v1, std::array<V3, 2>& b) {
    const V3 v01 = v1 - v0;
    const V3 xc = cross(V3(1, 0, 0), v01), yc = cross(V3(0, 1, 0), v01);
    if (sqn(xc) > sqn(yc)) { b[0] = normalized(xc); b[1] = normalized(cross(v01, xc)); }
    else { b[0] = normalized(yc); b[1] = normalized(cross(v01, yc)); }
}

// Relative tangential displacement of a lagged contact (computeRelDX_*), in 3D.
V3 rel_dx(const Mesh& m, size_t c) {
    const I4& id = m.selfLast[c];
    auto dx = [&](int v) { return m.x[v] - m.vprev[v]; };
    const double b1 = m.distCoord[c][0], b2 = m.distCoord[c][1];
    if (id[0] >= 0) return dx(id[0]) + b1 * (dx(id[1]) - dx(id[0])) - (dx(id[2]) + b2 * (dx(id[3]) - dx(id[2])));   // :117-126
    const int p = -id[0] - 1;
    if (id[2] < 0) return dx(p) - dx(id[1]);                                                             // :229-235
    if (id[3] < 0) return dx(p) - (dx(id[1]) + b1 * (dx(id[2]) - dx(id[1])));                               // :174-181
    return dx(p) - (dx(id[1]) + b1 * (dx(id[2]) - dx(id[1])) + b2 * (dx(id[3]) - dx(id[1])));              // :43-51
}
// liftRelDXTanToMesh_*: the 3M-vector of a tangent direction t (2D), and its vertex ids.
int lift(const Mesh& m, size_t c, const double t[2], double* out, int vid[4]) {
    const I4& id = m.selfLast[c];
    const V3 T = t[0] * m.tanBasis[c][0] + t[1] * m.tanBasis[c][1];
    const double b1 = m.distCoord[c][0], b2 = m.distCoord[c][1];
    auto put = [&](int k, const V3& v) { out[3 * k] = v.x; out[3 * k + 1] = v.y; out[3 * k + 2] = v.z; };
    if (id[0] >= 0) {
        put(0, (1.0 - b1) * T); put(1, b1 * T); put(2, (b2 - 1.0) * T); put(3, -b2 * T);                  // :128-139
        for (int k = 0; k < 4; ++k) vid[k] = id[k];
        return 4;
    }
    const int p = -id[0] - 1;
    vid[0] = p; vid[1] = id[1];
    put(0, T);
    if (id[2] < 0) { put(1, -1.0 * T); return 2; }                                                          // :237-244
    vid[2] = id[2];
    if (id[3] < 0) { put(1, (b1 - 1.0) * T); put(2, -b1 * T); return 3; }                                 // :183-192
    vid[3] = id[3];
    put(1, (-1 + b1 + b2) * T); put(2, -b1 * T); put(3, -b2 * T);                                          // :53-63
    return 4;
}

void friction_initialize(Mesh& m, double kappa) {
    const size_t n = m.selfAS.size();
    m.selfLam.resize(n); m.distCoord.assign(n, {0.0, 0.0}); m.tanBasis.resize(n);
    for (size_t i = 0; i < n; ++i) {                                                                     // :88-96
        const I4& c = m.selfAS[i];
        const double d = self_val(m, c);
        double lam;
        barrier_g(d, m.hhat, lam);
        lam *= -kappa * 2.0 * std::sqrt(d);
        if (c[3] < -1) lam *= -c[3];
        m.selfLam[i] = lam;
        if (c[0] >= 0) {
            closest_EE(m.x[c[0]], m.x[c[1]], m.x[c[2]], m.x[c[3]], m.distCoord[i].data());
            tangent_EE(m.x[c[0]], m.x[c[1]], m.x[c[2]], m.x[c[3]], m.tanBasis[i]);
        } else {
            const int p = -c[0] - 1;
            if (c[2] < 0) tangent_PP(m.x[p], m.x[c[1]], m.tanBasis[i]);
            else if (c[3] < 0) {
                m.distCoord[i][0] = closest_PE(m.x[p], m.x[c[1]], m.x[c[2]]);
                tangent_PE(m.x[p], m.x[c[1]], m.x[c[2]], m.tanBasis[i]);
            } else {
                closest_PT(m.x[p], m.x[c[1]], m.x[c[2]], m.x[c[3]], m.distCoord[i].data());
                tangent_PT(m.x[p], m.x[c[1]], m.x[c[2]], m.x[c[3]], m.tanBasis[i]);
            }
        }
    }
    m.selfLast = m.selfAS;
    m.envLam.resize(m.envAS.size());                                                                     // :141-152
    for (size_t i = 0; i < m.envAS.size(); ++i) {
        const double dist = dot(m.gn, m.x[m.envAS[i]]) - m.gd;
        const double d = dist * dist;
        double lam;
        barrier_g(d, m.hhat, lam);
        lam *= -kappa * 2.0 * std::sqrt(d);
        m.envLam[i] = lam;
    }
    m.envLast = m.envAS;
}

// Tangent-plane basis of the ground normal (Friction.cpp:66-76).
std::array<V3, 2> ground_basis(const V3& nrm) {
    const V3 helper = std::fabs(nrm.x) < 0.9 ? V3(1, 0, 0) : V3(0, 1, 0);
    std::array<V3, 2> b;
    b[0] = normalized(helper - dot(helper, nrm) * nrm);
    b[1] = normalized(cross(nrm, b[0]));
    return b;
}
// grad[v] += c; gabs[v] += the rounding scale of c: its Euclidean length (the terms are
// vectors formed through rotations, projections and distance gradients, whose components
// carry errors relative to the vector's length), or an explicit per-component scale `sc`
// (element forces, whose components are sums of products that cancel)
inline void gadd(Mesh& m, int v, const V3& c) {
    m.grad[v] = m.grad[v] + c;
    const double n = std::sqrt(sqn(c));
    m.gabs[v] = m.gabs[v] + V3(n, n, n);
}
inline void gadd(Mesh& m, int v, const V3& c, const V3& sc) {
    m.grad[v] = m.grad[v] + c;
    m.gabs[v] = m.gabs[v] + sc;
}


// Friction::addGradient (Friction.cpp:158-302).
void friction_gradient(Mesh& m, double eps2, double coef) {
    const double eps = std::sqrt(eps2);
    for (size_t c = 0; c < m.selfLast.size(); ++c) {
        const V3 r3 = rel_dx(m, c);
        double t[2] = {dot(r3, m.tanBasis[c][0]), dot(r3, m.tanBasis[c][1])};
        const double sq = t[0] * t[0] + t[1] * t[1];
        if (sq > eps2) { const double nr = std::sqrt(sq); t[0] /= nr; t[1] /= nr; }
        else { double f1; f1_SF_div(sq, eps, f1); t[0] *= f1; t[1] *= f1; }
        double L[12]; int vid[4];
        const int nv = lift(m, c, t, L, vid);
        const double s = coef * m.selfLam[c];
        for (int k = 0; k < nv; ++k) gadd(m, vid[k], V3(L[3 * k] * s, L[3 * k + 1] * s, L[3 * k + 2] * s));
    }
    for (size_t c = 0; c < m.envLast.size(); ++c) {
        const int v = m.envLast[c];
        const V3 d = m.x[v] - m.vprev[v];
        const V3 pj = d - dot(d, m.gn) * m.gn;
        const double mag2 = sqn(pj);
        if (mag2 > eps2) gadd(m, v, (coef * m.envLam[c] / std::sqrt(mag2)) * pj);
        else gadd(m, v, (coef * m.envLam[c] / eps) * pj);
    }
}

// Friction::addHessian (Friction.cpp:304-457) with analyticFrictionHessian (:11-64).
void friction_hessian(Mesh& m, Blocks& BH, double eps2, double coef) {
    const double eps = std::sqrt(eps2);
    const std::array<V3, 2> gb = ground_basis(m.gn);
    for (size_t c = 0; c < m.envLast.size(); ++c) {
        const int v = m.envLast[c];
        const double mult = std::max(0.0, coef * m.envLam[c]);
        const V3 d = m.x[v] - m.vprev[v];
        const double r0 = dot(gb[0], d), r1 = dot(gb[1], d);
        const double nr = std::sqrt(r0 * r0 + r1 * r1);
        std::array<double, 9> H{};
        if (mult > 0.0) {
            if (nr > eps) {
                const double p0 = r0 / nr, p1 = r1 / nr;
                const V3 l = (-p1) * gb[0] + p0 * gb[1];
                const double s = mult / nr;
                for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) H[3 * i + j] = (s * l[i]) * l[j];
            } else {
                const double s = mult / eps;
                for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) H[3 * i + j] = s * (gb[0][i] * gb[0][j] + gb[1][i] * gb[1][j]);
            }
        }
        BH.h3.push_back(H); BH.d1.push_back(v);
    }
    for (size_t c = 0; c < m.selfLast.size(); ++c) {
        const V3 r3 = rel_dx(m, c);
        const double u[2] = {dot(r3, m.tanBasis[c][0]), dot(r3, m.tanBasis[c][1])};
        const double w = std::max(0.0, coef * m.selfLam[c]);
        double H[144] = {};
        int vid[4] = {-1, -1, -1, -1};
        int nv = 0;
        {
            double L[12];
            const double t0[2] = {1.0, 0.0};
            nv = lift(m, c, t0, L, vid);
        }
        const int N = 3 * nv;
        auto add = [&](const double t[2], double ev) {
            const double e = std::max(0.0, ev);
            if (e == 0.0) return;
            double L[12]; int vv[4];
            lift(m, c, t, L, vv);
            const double s = w * e;
            for (int i = 0; i < N; ++i) for (int j = 0; j < N; ++j) H[N * i + j] += (s * L[i]) * L[j];
        };
        if (w != 0.0) {
            const double sq = u[0] * u[0] + u[1] * u[1];
            const double nr = std::sqrt(sq);
            if (nr > 0.0) {
                const double par[2] = {u[0] / nr, u[1] / nr};
                const double perp[2] = {-par[1], par[0]};
                if (nr > eps) add(perp, 1.0 / nr);
                else {
                    double pe, pa;
                    f1_SF_div(sq, eps, pe); f2_SF(sq, eps, pa);
                    add(par, pa); add(perp, pe);
                }
            } else {
                double iso;
                f1_SF_div(0.0, eps, iso);
                const double ex[2] = {1.0, 0.0}, ey[2] = {0.0, 1.0};
                add(ex, iso); add(ey, iso);
            }
        }
        if (nv == 2) { std::array<double, 36> h; std::copy(H, H + 36, h.begin()); BH.h6.push_back(h); BH.d2.push_back({vid[0], vid[1]}); }
        else if (nv == 3) { std::array<double, 81> h; std::copy(H, H + 81, h.begin()); BH.h9.push_back(h); BH.d3.push_back({vid[0], vid[1], vid[2]}); }
        else { std::array<double, 144> h; std::copy(H, H + 144, h.begin()); BH.h12.push_back(h); BH.d4.push_back({vid[0], vid[1], vid[2], vid[3]}); }
    }
}

// Friction::energy (Friction.cpp:459-545), without the friction coefficient's factor.
double friction_energy(const Mesh& m) {
    const double eps2 = m.fhat * m.dt * m.dt, eps = std::sqrt(eps2);
    double sum = 0.0;
    for (size_t c = 0; c < m.selfLast.size(); ++c) {
        const V3 r3 = rel_dx(m, c);
        const double a = dot(r3, m.tanBasis[c][0]), b = dot(r3, m.tanBasis[c][1]);
        const double sq = a * a + b * b;
        if (sq > eps2) sum += m.selfLam[c] * std::sqrt(sq);
        else { double f0; f0_SF(sq, eps, f0); sum += m.selfLam[c] * f0; }
    }
    double result = m.fric * sum;
    for (size_t c = 0; c < m.envLast.size(); ++c) {
        const int v = m.envLast[c];
        const V3 d = m.x[v] - m.vprev[v];
        const V3 pj = d - dot(d, m.gn) * m.gn;
        const double sq = sqn(pj);
        const double cl = sq > eps2 ? std::sqrt(sq) - 0.5 * eps : 0.5 * sq / eps;
        result += m.fric * m.envLam[c] * cl;
    }
    return result;
}

// ----------------------------------------------------------------------------------
// computeGradientAndHessian (IPCSolver.cpp:108-319), tet + ground + self contact subset
// ----------------------------------------------------------------------------------
int gradient_and_hessian(Mesh& m, double kappa, Blocks& BH) {
    m.grad.assign(m.V, V3());
    m.gabs.assign(m.V, V3());
    for (int v = 0; v < m.V; ++v) gadd(m, v, m.mass[v] * (m.x[v] - m.xt[v]));                     // :118-123
    for (int v = 0; v < m.V; ++v) gadd(m, v, (m.drag * m.mass[v]) * (m.x[v] - m.vprev[v])); // :124-128
    // soft targets (addSoftDerivatives, BoundaryConditions.cpp:175-208; IPCSolver.cpp:130)
    for (size_t i = 0; i < m.softIdx.size(); ++i) {
        const int v = m.softIdx[i];
        gadd(m, v, m.softW * (m.x[v] - m.softT[i]));
        std::array<double, 9> h{};
        h[0] = h[4] = h[8] = m.softW;
        BH.h3.push_back(h); BH.d1.push_back(v);
    }
    // tets (IPCSolver.cpp:139-164)
    std::vector<std::array<double, 12>> fe(m.T), fs(m.T);
    BH.h12.assign(m.T, {}); BH.d4.assign(m.T, {});
#pragma omp parallel for schedule(dynamic, 256)
    for (int t = 0; t < m.T; ++t) {
        double P[9][12];
        pfpx(m.dminv[t], P);
        const M3 F = deformation(m, t);
        const M3 PE = snh_pk1(F, m.u, m.r);
        double ps[9];
        snh_pk1_scale(F, m.u, m.r, ps);
        for (int j = 0; j < 12; ++j) {
            double acc = 0, sc = 0;
            for (int k = 0; k < 9; ++k) {
                acc += (m.vol[t] * P[k][j]) * PE.m[k];
                sc += std::fabs(m.vol[t] * P[k][j]) * ps[k];
            }
            fe[t][j] = acc;
            fs[t][j] = sc;
        }
        double Hq[81];
        snh_hessian(F, m.u, m.r, Hq);
        double T1[9][12];
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 12; ++j) {
                double acc = 0;
                for (int k = 0; k < 9; ++k) acc += Hq[9 * i + k] * P[k][j];
                T1[i][j] = acc;
            }
        const double s = m.vol[t] * m.dt * m.dt;
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) {
                double acc = 0;
                for (int k = 0; k < 9; ++k) acc += (s * P[k][i]) * T1[k][j];
                BH.h12[t][12 * i + j] = acc;
            }
        BH.d4[t] = m.tets[t];
    }
    const double dt2 = m.dt * m.dt;
    for (int t = 0; t < m.T; ++t)
        for (int a = 0; a < 4; ++a) {
            V3 f(fe[t][3 * a], fe[t][3 * a + 1], fe[t][3 * a + 2]);
            gadd(m, m.tets[t][a], dt2 * f, dt2 * V3(fs[t][3 * a], fs[t][3 * a + 1], fs[t][3 * a + 2]));
        }
    // cloth triangles (IPCSolver.cpp:166-194): f = dt^2 area PFPX^T vec(P), H9 = area dt^2 PFPX^T Hq PFPX
    std::vector<std::array<double, 9>> ft(m.Tr), fts(m.Tr);
    BH.h9.assign(m.Tr, {}); BH.d3.assign(m.Tr, {});
#pragma omp parallel for schedule(dynamic, 256)
    for (int t = 0; t < m.Tr; ++t) {
        double P[6][9], f[6], pe[6], Hq[36], ps[6];
        tri_pfpx(m.dmt[t], P);
        tri_deformation(m, t, f);
        bw_pk1(f, m.stretchS, m.shearS, m.srate, pe);
        bw_pk1_scale(f, m.stretchS, m.shearS, m.srate, ps);
        for (int j = 0; j < 9; ++j) {
            double acc = 0, sc = 0;
            for (int k = 0; k < 6; ++k) { acc += P[k][j] * pe[k]; sc += std::fabs(P[k][j]) * ps[k]; }
            ft[t][j] = m.dt * m.dt * m.area[t] * acc;
            fts[t][j] = m.dt * m.dt * m.area[t] * sc;
        }
        bw_hessian(f, m.stretchS, m.shearS, m.srate, Hq);
        double T1[6][9];
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 9; ++j) {
                double acc = 0;
                for (int k = 0; k < 6; ++k) acc += Hq[6 * i + k] * P[k][j];
                T1[i][j] = acc;
            }
        const double sc = m.area[t] * m.dt * m.dt;
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 9; ++j) {
                double acc = 0;
                for (int k = 0; k < 6; ++k) acc += P[k][i] * T1[k][j];
                BH.h9[t][9 * i + j] = sc * acc;
            }
        BH.d3[t] = m.tris[t];
    }
    for (int t = 0; t < m.Tr; ++t)
        for (int a = 0; a < 3; ++a)
            gadd(m, m.tris[t][a], V3(ft[t][3 * a], ft[t][3 * a + 1], ft[t][3 * a + 2]), V3(fts[t][3 * a], fts[t][3 * a + 1], fts[t][3 * a + 2]));
    // bending (IPCSolver.cpp:197-267): quadratic form D dt^2 (Q x I3) or the hinge angle
    std::vector<std::array<double, 12>> fb(m.Nb), fbs(m.Nb);
    std::vector<std::array<double, 144>> hb(m.Nb);
#pragma omp parallel for schedule(dynamic, 256)
    for (int h = 0; h < m.Nb; ++h) {
        const I4& bv = m.bverts[h];
        if (m.bmode == 0) {
            const double sc = m.plate * m.dt * m.dt;
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d) {
                    double acc = 0, as = 0;
                    for (int b = 0; b < 4; ++b) {
                        acc += (sc * m.bq[h][4 * a + b]) * m.x[bv[b]][d];
                        as += std::fabs((sc * m.bq[h][4 * a + b]) * m.x[bv[b]][d]);
                    }
                    fb[h][3 * a + d] = acc;
                    fbs[h][3 * a + d] = as;
                }
            for (int i = 0; i < 144; ++i) hb[h][i] = 0.0;
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    for (int d = 0; d < 3; ++d) hb[h][12 * (3 * a + d) + 3 * b + d] = sc * m.bq[h][4 * a + b];
        } else {
            double ag[12], aH[144];
            const double th = edge_theta(m.x[bv[0]], m.x[bv[1]], m.x[bv[2]], m.x[bv[3]], ag, aH);
            const double df = hinge_difference(th, m.brest[h]);
            const double sc = m.plate * m.bweight[h];
            for (int i = 0; i < 12; ++i) fb[h][i] = sc * df * ag[i];
            for (int i = 0; i < 12; ++i)
                for (int j = 0; j < 12; ++j) hb[h][12 * i + j] = sc * (ag[i] * ag[j] + df * aH[12 * i + j]);
            make_pd<12>(hb[h].data());
            for (int i = 0; i < 12; ++i) fb[h][i] *= m.dt * m.dt;
            for (int a = 0; a < 4; ++a) {
                const double gn = std::sqrt(ag[3 * a] * ag[3 * a] + ag[3 * a + 1] * ag[3 * a + 1] + ag[3 * a + 2] * ag[3 * a + 2]);
                for (int d = 0; d < 3; ++d)
                    fbs[h][3 * a + d] = std::fabs(sc) * (std::fabs(df) + std::fabs(th) + std::fabs(m.brest[h])) * gn * m.dt * m.dt;
            }
            for (int i = 0; i < 144; ++i) hb[h][i] *= m.dt * m.dt;
        }
    }
    for (int h = 0; h < m.Nb; ++h) {
        for (int a = 0; a < 4; ++a)
            gadd(m, m.bverts[h][a], V3(fb[h][3 * a], fb[h][3 * a + 1], fb[h][3 * a + 2]), V3(fbs[h][3 * a], fbs[h][3 * a + 1], fbs[h][3 * a + 2]));
        BH.h12.push_back(hb[h]); BH.d4.push_back(m.bverts[h]);
    }
    // ground (IPCSolver.cpp:270-293)
    for (int v : m.envAS) {
        const double dist = dot(m.gn, m.x[v]) - m.gd;
        const double d = dist * dist;
        double gb, hb;
        barrier_g(d, m.hhat, gb); barrier_h(d, m.hhat, hb);
        gadd(m, v, (((kappa * gb) * 2.0) * dist) * m.gn);
        const double param = 4.0 * hb * d + 2.0 * gb;
        if (param > 0) {
            std::array<double, 9> h;
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) h[3 * i + j] = (kappa * param) * (m.gn[i] * m.gn[j]);
            BH.h3.push_back(h); BH.d1.push_back(v);
        }
    }
    // self contact gradient, compute_g_dpt (ContactMechanics.cpp:2970-3032), serial
    int collisions = 0;
    auto add_g = [&](int type, const int* vid, double scale) {
        V3 p[4]; const int nv = type == 0 ? 2 : (type == 1 ? 3 : 4);
        for (int k = 0; k < nv; ++k) p[k] = m.x[vid[k]];
        double d, g[12];
        deriv_any(type, p, d, g, nullptr);
        for (int k = 0; k < nv; ++k) gadd(m, vid[k], V3(scale * g[3 * k], scale * g[3 * k + 1], scale * g[3 * k + 2]));
    };
    auto g_dpt = [&](const std::vector<I4>& set, const std::vector<double>& in) {
        int n = 0;
        for (size_t i = 0; i < set.size(); ++i) {
            int vid[4]; const int type = decode(set[i], vid);
            double scale;
            if (type == 0 || type == 1) { scale = kappa * -set[i][3] * in[i]; n += -set[i][3]; }
            else { scale = kappa * in[i]; n += 1; }
            add_g(type, vid, scale);
        }
        return n;
    };
    std::vector<double> gb(m.selfAS.size());
    for (size_t i = 0; i < m.selfAS.size(); ++i) barrier_g(self_val(m, m.selfAS[i]), m.hhat, gb[i]);
    collisions += g_dpt(m.selfAS, gb);
    // compute_g_dee (ContactMechanics.cpp:3034-3080)
    std::vector<double> edb(m.eeAS.size());
    for (size_t i = 0; i < m.eeAS.size(); ++i) {
        const I4& c = m.eeAS[i];
        const double d = self_val(m, c);
        double b, g;
        barrier_b(d, m.hhat, b); barrier_g(d, m.hhat, g);
        int e4[4];
        if (c[3] >= 0) { for (int k = 0; k < 4; ++k) e4[k] = c[k]; }
        else {
            const auto& eI = m.se[m.eeIe[i].first]; const auto& eJ = m.se[m.eeIe[i].second];
            e4[0] = eI.first; e4[1] = eI.second; e4[2] = eJ.first; e4[3] = eJ.second;
        }
        const double ex = eps_x(m, e4[0], e4[1], e4[2], e4[3]);
        V3 p[4] = {m.x[e4[0]], m.x[e4[1]], m.x[e4[2]], m.x[e4[3]]};
        const double e = mol_e(ee_cross_sqn(p[0], p[1], p[2], p[3]), ex);
        double eg[12];
        mol_eg(p, ex, eg);
        const double cb = kappa * b;
        for (int k = 0; k < 4; ++k) gadd(m, e4[k], cb * V3(eg[3 * k], eg[3 * k + 1], eg[3 * k + 2]));
        edb[i] = g * e;
    }
    g_dpt(m.eeAS, edb);

    // compute_H_dpt (ContactMechanics.cpp:3466-3573)
    const size_t ns = m.selfAS.size();
    std::vector<std::array<double, 144>> Hs(ns);
    std::vector<int> types(ns);
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < (long)ns; ++i) {
        int vid[4]; const int type = decode(m.selfAS[i], vid);
        types[i] = type;
        const int nv = type == 0 ? 2 : (type == 1 ? 3 : 4), n = 3 * nv;
        V3 p[4];
        for (int k = 0; k < nv; ++k) p[k] = m.x[vid[k]];
        double d, g[12], H[144];
        deriv_any(type, p, d, g, H);
        double g_b, H_b;
        barrier_g(d, m.hhat, g_b); barrier_h(d, m.hhat, H_b);
        const double coef = (type <= 1) ? kappa * -m.selfAS[i][3] : kappa;
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b) Hs[i][n * a + b] = ((coef * H_b) * g[a]) * g[b] + (coef * g_b) * H[n * a + b];
        if (n == 6) make_pd<6>(Hs[i].data());
        else if (n == 9) make_pd<9>(Hs[i].data());
        else make_pd<12>(Hs[i].data());
    }
    for (size_t i = 0; i < ns; ++i) {
        int vid[4]; decode(m.selfAS[i], vid);
        if (types[i] == 0) { std::array<double, 36> h; std::copy(Hs[i].begin(), Hs[i].begin() + 36, h.begin()); BH.h6.push_back(h); BH.d2.push_back({vid[0], vid[1]}); }
        else if (types[i] == 1) { std::array<double, 81> h; std::copy(Hs[i].begin(), Hs[i].begin() + 81, h.begin()); BH.h9.push_back(h); BH.d3.push_back({vid[0], vid[1], vid[2]}); }
        else { BH.h12.push_back(Hs[i]); BH.d4.push_back({vid[0], vid[1], vid[2], vid[3]}); }
    }
    // compute_H_dee (ContactMechanics.cpp:3575-3700)
    const size_t ne = m.eeAS.size();
    std::vector<std::array<double, 144>> He(ne);
    std::vector<I4> De(ne);
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < (long)ne; ++i) {
        const I4& c = m.eeAS[i];
        const double d = self_val(m, c);
        double b, g_b, H_b;
        barrier_b(d, m.hhat, b); barrier_g(d, m.hhat, g_b); barrier_h(d, m.hhat, H_b);
        int e4[4];
        double gd[12] = {}, Hd[144] = {};
        if (c[3] >= 0) {
            for (int k = 0; k < 4; ++k) e4[k] = c[k];
            V3 p[4] = {m.x[c[0]], m.x[c[1]], m.x[c[2]], m.x[c[3]]};
            double dd;
            dist_derivs<12>(3, p, dd, gd, Hd);
        } else {
            const auto& eI = m.se[m.eeIe[i].first]; const auto& eJ = m.se[m.eeIe[i].second];
            e4[0] = eI.first; e4[1] = eI.second; e4[2] = eJ.first; e4[3] = eJ.second;
            int vid[4]; const int type = decode(c, vid);
            const int nv = type == 0 ? 2 : 3;
            int map[3] = {0, 0, 0};
            for (int k = 0; k < 4; ++k) {
                if (vid[0] == e4[k]) map[0] = k;
                else if (vid[1] == e4[k]) map[1] = k;
                else if (nv == 3 && vid[2] == e4[k]) map[2] = k;
            }
            V3 p[3];
            for (int k = 0; k < nv; ++k) p[k] = m.x[vid[k]];
            double dd, gl[9], Hl[81];
            deriv_any(type, p, dd, gl, Hl);
            const int n = 3 * nv;
            for (int a = 0; a < nv; ++a) {
                for (int r = 0; r < 3; ++r) gd[3 * map[a] + r] = gl[3 * a + r];
                for (int bb = 0; bb < nv; ++bb)
                    for (int r = 0; r < 3; ++r)
                        for (int s = 0; s < 3; ++s) Hd[12 * (3 * map[a] + r) + 3 * map[bb] + s] = Hl[n * (3 * a + r) + 3 * bb + s];
            }
        }
        const double ex = eps_x(m, e4[0], e4[1], e4[2], e4[3]);
        V3 p[4] = {m.x[e4[0]], m.x[e4[1]], m.x[e4[2]], m.x[e4[3]]};
        const double e = mol_e(ee_cross_sqn(p[0], p[1], p[2], p[3]), ex);
        double eg[12], eH[144];
        mol_eg(p, ex, eg); mol_eH(p, ex, eH);
        for (int a = 0; a < 12; ++a)
            for (int bb = 0; bb < 12; ++bb) {
                const double kgg = ((kappa * g_b) * gd[a]) * eg[bb];
                const double kggT = ((kappa * g_b) * gd[bb]) * eg[a];
                He[i][12 * a + bb] = kgg + kggT + (kappa * b) * eH[12 * a + bb] + ((kappa * e * H_b) * gd[a]) * gd[bb]
                                   + (kappa * e * g_b) * Hd[12 * a + bb];
            }
        make_pd<12>(He[i].data());
        De[i] = {e4[0], e4[1], e4[2], e4[3]};
    }
    for (size_t i = 0; i < ne; ++i) { BH.h12.push_back(He[i]); BH.d4.push_back(De[i]); }
    // lagged friction (IPCSolver.cpp:309-317)
    friction_gradient(m, m.fhat * m.dt * m.dt, m.fric);
    friction_hessian(m, BH, m.fhat * m.dt * m.dt, m.fric);
    return collisions;
}

// BHessian::toTriplets + NewtonLinearSystem::assemble (ContactMechanics.cpp:13-134;
// NewtonLinearSystem.cpp:126-154).  setFromTriplets sums duplicates in insertion order.
void assemble_csr(Mesh& m, const Blocks& BH) {
    struct Tri { int r, c; double v, s; };
    std::vector<Tri> trip;
    auto append = [&](int nv, const double* H, const int* vid) {
        const int n = 3 * nv;
        // rounding scale of every entry of this block: n max|H|, a bound of its 2-norm, to which
        // the error of its eigendecomposition (the PSD projection) is proportional
        double hmax = 0.0;
        for (int i = 0; i < n * n; ++i) hmax = std::max(hmax, std::fabs(H[i]));
        hmax *= n;
        for (int a = 0; a < nv; ++a) {
            const int rv = vid[a];
            if (rv < 0 || rv >= m.V || !isFree(m.bt[rv])) continue;
            for (int b = 0; b < nv; ++b) {
                const int cv = vid[b];
                if (cv < 0 || cv >= m.V || !isFree(m.bt[cv])) continue;
                for (int rd = 0; rd < 3; ++rd)
                    for (int cd = 0; cd < 3; ++cd) {
                        const int gr = 3 * rv + rd, gc = 3 * cv + cd;
                        const double v = H[n * (3 * a + rd) + 3 * b + cd];
                        if (gr >= gc && v != 0.0) trip.push_back({gr, gc, v, hmax});
                    }
            }
        }
    };
    for (size_t i = 0; i < BH.h3.size(); ++i) append(1, BH.h3[i].data(), &BH.d1[i]);
    for (size_t i = 0; i < BH.h6.size(); ++i) append(2, BH.h6[i].data(), BH.d2[i].data());
    for (size_t i = 0; i < BH.h9.size(); ++i) append(3, BH.h9[i].data(), BH.d3[i].data());
    for (size_t i = 0; i < BH.h12.size(); ++i) append(4, BH.h12[i].data(), BH.d4[i].data());
    for (int dof = 0; dof < 3 * m.V; ++dof) {
        const double v = m.mass[dof / 3] * (1.0 + m.drag);
        trip.push_back({dof, dof, v, std::fabs(v)});
    }
    const int n = 3 * m.V;
    std::vector<int> cnt(n + 1, 0);
    for (const auto& t : trip) cnt[t.c + 1]++;
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    std::vector<int> order(trip.size());
    for (size_t i = 0; i < trip.size(); ++i) order[pos[trip[i].c]++] = (int)i;
    m.outer.assign(n + 1, 0); m.inner.clear(); m.val.clear(); m.vscale.clear();
    for (int c = 0; c < n; ++c) {
        std::vector<int> ids(order.begin() + cnt[c], order.begin() + cnt[c + 1]);
        std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) { return trip[a].r < trip[b].r; });
        for (size_t k = 0; k < ids.size(); ++k) {
            const Tri& t = trip[ids[k]];
            if (k > 0 && trip[ids[k - 1]].r == t.r) { m.val.back() += t.v; m.vscale.back() += t.s; }
            else { m.inner.push_back(t.r); m.val.push_back(t.v); m.vscale.push_back(t.s); }
        }
        m.outer[c + 1] = (int)m.inner.size();
    }
    m.rhs.assign(n, 0.0);
    for (int v = 0; v < m.V; ++v)
        if (isFree(m.bt[v])) { m.rhs[3 * v] = m.grad[v].x; m.rhs[3 * v + 1] = m.grad[v].y; m.rhs[3 * v + 2] = m.grad[v].z; }
}

// ----------------------------------------------------------------------------------
// computeEnergy (IPCSolver.cpp:322-456), tet + inertia + barrier subset
// ----------------------------------------------------------------------------------
void energy_terms(Mesh& m, double kappa, double out[9]) {
    std::vector<double> et(m.T);
#pragma omp parallel for schedule(dynamic, 256)
    for (int t = 0; t < m.T; ++t) et[t] = snh_energy_density(deformation(m, t), m.u, m.r) * m.vol[t];
    double esnh = 0;
    for (int t = 0; t < m.T; ++t) esnh += et[t];
    // Baraff-Witkin triangles and bending, before the dt^2 factor (IPCSolver.cpp:330-368)
    std::vector<double> etr(m.Tr), ebd(m.Nb);
#pragma omp parallel for schedule(dynamic, 256)
    for (int t = 0; t < m.Tr; ++t) {
        double f[6];
        tri_deformation(m, t, f);
        etr[t] = bw_energy_density(f, m.stretchS, m.shearS, m.srate) * m.area[t];
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int h = 0; h < m.Nb; ++h) {
        const I4& bv = m.bverts[h];
        if (m.bmode == 0) {
            double acc = 0;
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) acc += m.bq[h][4 * a + b] * dot(m.x[bv[a]], m.x[bv[b]]);
            ebd[h] = acc;
        } else {
            const double df = hinge_difference(edge_theta(m.x[bv[0]], m.x[bv[1]], m.x[bv[2]], m.x[bv[3]], nullptr, nullptr), m.brest[h]);
            ebd[h] = 0.5 * m.plate * m.bweight[h] * df * df;
        }
    }
    double ebw = 0, ebend = 0;
    for (int t = 0; t < m.Tr; ++t) ebw += etr[t];
    for (int h = 0; h < m.Nb; ++h) ebend += ebd[h];
    if (m.bmode == 0) ebend *= 0.5 * m.plate;
    double de = 0;
    for (int v = 0; v < m.V; ++v) {
        de += sqn(m.x[v] - m.xt[v]) * m.mass[v] / 2.0;
        de += (sqn(m.x[v] - m.vprev[v]) * (m.drag) * m.mass[v] / 2.0);
    }
    // softEnergy (BoundaryConditions.cpp:210-229)
    double esoft = 0;
    for (size_t i = 0; i < m.softIdx.size(); ++i) esoft += 0.5 * m.softW * sqn(m.x[m.softIdx[i]] - m.softT[i]);
    double eg = 0, es = 0, ee = 0;
    for (int v : m.envAS) {
        const double dist = dot(m.gn, m.x[v]) - m.gd;
        double b; barrier_b(dist * dist, m.hhat, b); eg += b;
    }
    for (const I4& c : m.selfAS) {
        double b; barrier_b(self_val(m, c), m.hhat, b);
        if (c[3] < -1) b *= -c[3];
        es += b;
    }
    for (size_t i = 0; i < m.eeAS.size(); ++i) {
        const I4& c = m.eeAS[i];
        int e4[4];
        if (c[3] >= 0) { for (int k = 0; k < 4; ++k) e4[k] = c[k]; }
        else {
            const auto& eI = m.se[m.eeIe[i].first]; const auto& eJ = m.se[m.eeIe[i].second];
            e4[0] = eI.first; e4[1] = eI.second; e4[2] = eJ.first; e4[3] = eJ.second;
        }
        const double ex = eps_x(m, e4[0], e4[1], e4[2], e4[3]);
        const double e = mol_e(ee_cross_sqn(m.x[e4[0]], m.x[e4[1]], m.x[e4[2]], m.x[e4[3]]), ex);
        double b; barrier_b(self_val(m, c), m.hhat, b);
        ee += b * e;
    }
    out[0] = esnh; out[1] = de; out[2] = eg; out[3] = es; out[4] = ee;
    out[5] = friction_energy(m);
    out[6] = ebw; out[7] = ebend; out[8] = esoft;
    (void)kappa;
}
// initKappa's reductions (IPCSolver.cpp:766-799): g_E from computeEGradient (inertia
// without drag, tet elastic forces times dt^2; IPCSolver.cpp:74-106), g_c the ground and
// self barrier gradients with kappa = 1 (compute_g_dpt with multiplicity).
// out = {sum_v g_c . g_E, sum_v |g_c|^2}; returns the number of ground + self constraints.
int kappa_gradients(Mesh& m, double out[2]) {
    const int nc = (int)(m.envAS.size() + m.selfAS.size());
    out[0] = out[1] = 0.0;
    if (!nc) return 0;
    std::vector<V3> gE(m.V, V3()), gc(m.V, V3());
    for (int v = 0; v < m.V; ++v) gE[v] = gE[v] + m.mass[v] * (m.x[v] - m.xt[v]);
    const double dt2 = m.dt * m.dt;
    std::vector<std::array<double, 12>> fe(m.T);
#pragma omp parallel for schedule(dynamic, 256)
    for (int t = 0; t < m.T; ++t) {
        double P[9][12];
        pfpx(m.dminv[t], P);
        const M3 PE = snh_pk1(deformation(m, t), m.u, m.r);
        for (int j = 0; j < 12; ++j) {
            double acc = 0;
            for (int k = 0; k < 9; ++k) acc += (m.vol[t] * P[k][j]) * PE.m[k];
            fe[t][j] = acc;
        }
    }
    for (int t = 0; t < m.T; ++t)
        for (int a = 0; a < 4; ++a) gE[m.tets[t][a]] = gE[m.tets[t][a]] + dt2 * V3(fe[t][3 * a], fe[t][3 * a + 1], fe[t][3 * a + 2]);
    for (int v : m.envAS) {
        const double dist = dot(m.gn, m.x[v]) - m.gd;
        double gb;
        barrier_g(dist * dist, m.hhat, gb);
        gc[v] = gc[v] + ((gb * 2.0) * dist) * m.gn;
    }
    for (const I4& c : m.selfAS) {
        int vid[4];
        const int type = decode(c, vid);
        double gb;
        barrier_g(self_val(m, c), m.hhat, gb);
        const double scale = (type == 0 || type == 1) ? -c[3] * gb : gb;
        V3 p[4];
        const int nv = type == 0 ? 2 : (type == 1 ? 3 : 4);
        for (int k = 0; k < nv; ++k) p[k] = m.x[vid[k]];
        double d, g[12];
        deriv_any(type, p, d, g, nullptr);
        for (int k = 0; k < nv; ++k) gc[vid[k]] = gc[vid[k]] + V3(scale * g[3 * k], scale * g[3 * k + 1], scale * g[3 * k + 2]);
    }
    for (int v = 0; v < m.V; ++v) { out[0] += dot(gc[v], gE[v]); out[1] += sqn(gc[v]); }
    return nc;
}

// postLineSearch's constraint bookkeeping (IPCSolver.cpp:815-866): whether a recorded
// close constraint is now at or below its recorded value, then the new records (ground
// and self constraints with value < dTol, in set order).
int close_constraints(Mesh& m, double dtol) {
    int closer = 0;
    for (size_t i = 0; i < m.closeEnv.size() && !closer; ++i) {
        const double dist = dot(m.gn, m.x[m.closeEnv[i]]) - m.gd;
        if (dist * dist <= m.closeEnvVal[i]) closer = 1;
    }
    for (size_t i = 0; i < m.closeSelf.size() && !closer; ++i)
        if (self_val(m, m.closeSelf[i]) <= m.closeSelfVal[i]) closer = 1;
    m.closeEnv.clear(); m.closeEnvVal.clear(); m.closeSelf.clear(); m.closeSelfVal.clear();
    for (int v : m.envAS) {
        const double dist = dot(m.gn, m.x[v]) - m.gd;
        const double d = dist * dist;
        if (d < dtol) { m.closeEnv.push_back(v); m.closeEnvVal.push_back(d); }
    }
    for (const I4& c : m.selfAS) {
        const double d = self_val(m, c);
        if (d < dtol) { m.closeSelf.push_back(c); m.closeSelfVal.push_back(d); }
    }
    return closer;
}

// The kappa loop's exit test of solveIPCStep (IPCSolver.cpp:1171-1188): the smallest and
// largest value over the ground constraints (Evaluate_GroundConstraintVals) followed by the
// self constraints (Evaluate_SelfPTConstraintVals); returns their number.
int constraint_extrema(Mesh& m, double out[2]) {
    int n = 0;
    double lo = 0.0, hi = 0.0;
    auto take = [&](double d) {
        if (n == 0 || d < lo) lo = d;
        if (n == 0 || d > hi) hi = d;
        ++n;
    };
    for (int v : m.envAS) {
        const double dist = dot(m.gn, m.x[v]) - m.gd;
        take(dist * dist);
    }
    for (const I4& c : m.selfAS) take(self_val(m, c));
    out[0] = lo; out[1] = hi;
    return n;
}

double energy(Mesh& m, double kappa) {
    double t[9];
    energy_terms(m, kappa, t);
    double E = t[0];
    E += t[6];
    E += t[7];
    E *= m.dt * m.dt;
    E += t[1];
    E += t[8];
    E += kappa * (t[2] + t[3] + t[4]);
    E += t[5];
    return E;
}

// ----------------------------------------------------------------------------------
// Broad phase and active set (CollisionBroadPhase.cpp:765-1092; LBVH.cpp:284-339)
// ----------------------------------------------------------------------------------
void build_static(Mesh& m) {
    m.faceBoxes.resize(m.F);
    for (int f = 0; f < m.F; ++f) m.faceBoxes[f] = triBox(m.x[m.faces[f][0]], m.x[m.faces[f][1]], m.x[m.faces[f][2]]);
    m.edgeBoxes.resize(m.Es);
    for (int e = 0; e < m.Es; ++e) m.edgeBoxes[e] = edgeBox(m.x[m.se[e].first], m.x[m.se[e].second]);
    m.sweptPointBoxes.clear();
    m.faceTree.build(m.faceBoxes);
    m.edgeTree.build(m.edgeBoxes);
}
void build_swept(Mesh& m, const std::vector<V3>& p, double a) {
    m.sweptPointBoxes.resize(m.Vs);
    for (int s = 0; s < m.Vs; ++s) m.sweptPointBoxes[s] = sweptPoint(m.x[m.sv[s]], p[m.sv[s]], a);
    m.faceBoxes.resize(m.F);
    for (int f = 0; f < m.F; ++f) {
        const auto& fv = m.faces[f];
        m.faceBoxes[f] = sweptTri(m.x[fv[0]], p[fv[0]], m.x[fv[1]], p[fv[1]], m.x[fv[2]], p[fv[2]], a);
    }
    m.edgeBoxes.resize(m.Es);
    for (int e = 0; e < m.Es; ++e) m.edgeBoxes[e] = sweptEdge(m.x[m.se[e].first], p[m.se[e].first], m.x[m.se[e].second], p[m.se[e].second], a);
    m.faceTree.build(m.faceBoxes);
    m.edgeTree.build(m.edgeBoxes);
}

void active_set(Mesh& m) {
    build_static(m);
    const double sq = std::sqrt(m.hhat);
    std::vector<std::vector<I4>> csPT(m.Vs), csEE(m.Es);
    std::vector<std::vector<int>> ccPT(m.Vs), ccEE(m.Es);
#pragma omp parallel
    {
        std::vector<int> cand;
#pragma omp for schedule(dynamic, 64)
        for (int s = 0; s < m.Vs; ++s) {
            const int vI = m.sv[s];
            m.faceTree.query(pointBox(m.x[vI]), sq, cand);
            const AABB pb = pointBox(m.x[vI]);
            for (int f : cand) {
                const auto& fv = m.faces[f];
                if (isCollider(m.bt[vI]) && isCollider(m.bt[fv[0]]) && isCollider(m.bt[fv[1]]) && isCollider(m.bt[fv[2]])) continue;
                if (vI == fv[0] || vI == fv[1] || vI == fv[2]) continue;
                if (!pb.overlaps(triBox(m.x[fv[0]], m.x[fv[1]], m.x[fv[2]]), sq)) continue;
                ccPT[s].push_back(f);
                const int dt = dType_PT(m.x[vI], m.x[fv[0]], m.x[fv[1]], m.x[fv[2]]);
                const V3 &P = m.x[vI], &A = m.x[fv[0]], &B = m.x[fv[1]], &C = m.x[fv[2]];
                double d;
                switch (dt) {
                case 0: d = d_PP(P, A); if (d < m.hhat) csPT[s].push_back({-vI - 1, fv[0], -1, -1}); break;
                case 1: d = d_PP(P, B); if (d < m.hhat) csPT[s].push_back({-vI - 1, fv[1], -1, -1}); break;
                case 2: d = d_PP(P, C); if (d < m.hhat) csPT[s].push_back({-vI - 1, fv[2], -1, -1}); break;
                case 3: d = d_PE(P, A, B); if (d < m.hhat) csPT[s].push_back({-vI - 1, fv[0], fv[1], -1}); break;
                case 4: d = d_PE(P, B, C); if (d < m.hhat) csPT[s].push_back({-vI - 1, fv[1], fv[2], -1}); break;
                case 5: d = d_PE(P, C, A); if (d < m.hhat) csPT[s].push_back({-vI - 1, fv[2], fv[0], -1}); break;
                case 6: d = d_PT(P, A, B, C); if (d < m.hhat) csPT[s].push_back({-vI - 1, fv[0], fv[1], fv[2]}); break;
                default: break;
                }
            }
        }
#pragma omp for schedule(dynamic, 64)
        for (int eI = 0; eI < m.Es; ++eI) {
            const auto& EI = m.se[eI];
            m.edgeTree.query(edgeBox(m.x[EI.first], m.x[EI.second]), sq, cand);
            cand.erase(std::remove_if(cand.begin(), cand.end(), [eI](int c) { return c <= eI; }), cand.end());
            const AABB b1 = edgeBox(m.x[EI.first], m.x[EI.second]);
            for (int eJ : cand) {
                const auto& EJ = m.se[eJ];
                if (isCollider(m.bt[EI.first]) && isCollider(m.bt[EI.second]) && isCollider(m.bt[EJ.first]) && isCollider(m.bt[EJ.second])) continue;
                if (EI.first == EJ.first || EI.first == EJ.second || EI.second == EJ.first || EI.second == EJ.second || eI > eJ) continue;
                if (!b1.overlaps(edgeBox(m.x[EJ.first], m.x[EJ.second]), sq)) continue;
                ccEE[eI].push_back(eJ);
                const V3 &a0 = m.x[EI.first], &a1 = m.x[EI.second], &b0 = m.x[EJ.first], &b1v = m.x[EJ.second];
                const int dt = dType_EE(a0, a1, b0, b1v);
                const double cr = ee_cross_sqn(a0, a1, b0, b1v);
                const double ex = eps_x(m, EI.first, EI.second, EJ.first, EJ.second);
                const int add_e = (cr < ex) ? -eJ - 2 : -1;
                double d;
                switch (dt) {
                case 0: d = d_PP(a0, b0); if (d < m.hhat) csEE[eI].push_back({-EI.first - 1, EJ.first, -1, add_e}); break;
                case 1: d = d_PP(a0, b1v); if (d < m.hhat) csEE[eI].push_back({-EI.first - 1, EJ.second, -1, add_e}); break;
                case 2: d = d_PE(a0, b0, b1v); if (d < m.hhat) csEE[eI].push_back({-EI.first - 1, EJ.first, EJ.second, add_e}); break;
                case 3: d = d_PP(a1, b0); if (d < m.hhat) csEE[eI].push_back({-EI.second - 1, EJ.first, -1, add_e}); break;
                case 4: d = d_PP(a1, b1v); if (d < m.hhat) csEE[eI].push_back({-EI.second - 1, EJ.second, -1, add_e}); break;
                case 5: d = d_PE(a1, b0, b1v); if (d < m.hhat) csEE[eI].push_back({-EI.second - 1, EJ.first, EJ.second, add_e}); break;
                case 6: d = d_PE(b0, a0, a1); if (d < m.hhat) csEE[eI].push_back({-EJ.first - 1, EI.first, EI.second, add_e}); break;
                case 7: d = d_PE(b1v, a0, a1); if (d < m.hhat) csEE[eI].push_back({-EJ.second - 1, EI.first, EI.second, add_e}); break;
                case 8:
                    d = d_EE(a0, a1, b0, b1v);
                    if (d < m.hhat) {
                        if (add_e <= -2) csEE[eI].push_back({EI.first, EI.second, EJ.first, -EJ.second - m.Es - 2});
                        else csEE[eI].push_back({EI.first, EI.second, EJ.first, EJ.second});
                    }
                    break;
                default: break;
                }
            }
        }
    }
    // serial compaction (CollisionBroadPhase.cpp:990-1058)
    m.ccdAS.clear();
    for (int s = 0; s < m.Vs; ++s) for (int f : ccPT[s]) m.ccdAS.emplace_back(-s - 1, f);
    for (int e = 0; e < m.Es; ++e) for (int f : ccEE[e]) m.ccdAS.emplace_back(e, f);
    m.selfAS.clear(); m.eeAS.clear(); m.eeIe.clear();
    std::map<I4, int> counter;
    for (const auto& cs : csPT) for (const auto& c : cs) { if (c[3] < 0) ++counter[c]; else m.selfAS.push_back(c); }
    for (int eI = 0; eI < m.Es; ++eI)
        for (const auto& c : csEE[eI]) {
            if (c[3] >= 0) m.selfAS.push_back(c);
            else if (c[3] == -1) ++counter[c];
            else if (c[3] >= -m.Es - 1) { m.eeAS.push_back({c[0], c[1], c[2], -1}); m.eeIe.emplace_back(eI, -c[3] - 2); }
            else { m.eeAS.push_back({c[0], c[1], c[2], -c[3] - m.Es - 2}); m.eeIe.emplace_back(-1, -1); }
        }
    for (const auto& kv : counter) m.selfAS.push_back({kv.first[0], kv.first[1], kv.first[2], -kv.second});
    // Ground::calculateActivateSet (CollisionBroadPhase.cpp:1076-1092), ascending surface order
    m.envAS.clear();
    for (int s = 0; s < m.Vs; ++s) {
        const double dist = dot(m.gn, m.x[m.sv[s]]) - m.gd;
        if (dist * dist < m.hhat) m.envAS.push_back(m.sv[s]);
    }
}

// ----------------------------------------------------------------------------------
// Additive CCD (AdditiveCCD.cpp:8-329)
// ----------------------------------------------------------------------------------
int pt_type_ccd(const V3& p, const V3& t0, const V3& t1, const V3& t2) {   // :8-62
    const V3 n = cross(t1 - t0, t2 - t0);
    double p00, p10, p01, p11, p02, p12;
    pt_param(t1 - t0, n, p - t0, p00, p10);
    if (p00 > 0.0 && p00 < 1.0 && p10 >= 0.0) return 3;
    pt_param(t2 - t1, n, p - t1, p01, p11);
    if (p01 > 0.0 && p01 < 1.0 && p11 >= 0.0) return 4;
    pt_param(t0 - t2, n, p - t2, p02, p12);
    if (p02 > 0.0 && p02 < 1.0 && p12 >= 0.0) return 5;
    if (p00 <= 0.0 && p02 >= 1.0) return 0;
    if (p01 <= 0.0 && p00 >= 1.0) return 1;
    if (p02 <= 0.0 && p01 >= 1.0) return 2;
    return 6;
}
double ccd_pe(const V3& p, const V3& e0, const V3& e1) { return sqn(cross(e0 - p, e1 - p)) / sqn(e1 - e0); }  // :146-149
double ccd_pt(const V3& p, const V3& t0, const V3& t1, const V3& t2) {                                     // :151-156
    const V3 b = cross(t1 - t0, t2 - t0);
    const double aTb = dot(p - t0, b);
    return aTb * aTb / sqn(b);
}
double ccd_ee(const V3& a0, const V3& a1, const V3& b0, const V3& b1) {                                     // :158-163
    const V3 b = cross(a1 - a0, b1 - b0);
    const double aTb = dot(b0 - a0, b);
    return aTb * aTb / sqn(b);
}
double ee_unclassified(const V3& a0, const V3& a1, const V3& b0, const V3& b1) {                            // :165-191
    switch (dType_EE(a0, a1, b0, b1)) {
    case 0: return sqn(a0 - b0);
    case 1: return sqn(a0 - b1);
    case 2: return ccd_pe(a0, b0, b1);
    case 3: return sqn(a1 - b0);
    case 4: return sqn(a1 - b1);
    case 5: return ccd_pe(a1, b0, b1);
    case 6: return ccd_pe(b0, a0, a1);
    case 7: return ccd_pe(b1, a0, a1);
    case 8: return ccd_ee(a0, a1, b0, b1);
    default: return std::numeric_limits<double>::max();
    }
}
double pt_unclassified(const V3& p, const V3& t0, const V3& t1, const V3& t2) {                             // :193-219
    switch (pt_type_ccd(p, t0, t1, t2)) {
    case 0: return sqn(p - t0);
    case 1: return sqn(p - t1);
    case 2: return sqn(p - t2);
    case 3: return ccd_pe(p, t0, t1);
    case 4: return ccd_pe(p, t1, t2);
    case 5: return ccd_pe(p, t2, t0);
    case 6: return ccd_pt(p, t0, t1, t2);
    default: return std::numeric_limits<double>::max();
    }
}
double accd_ee(V3 ea0, V3 ea1, V3 eb0, V3 eb1, V3 dea0, V3 dea1, V3 deb0, V3 deb1, double eta, double th) {  // :221-279
    const V3 mov = (dea0 + dea1 + deb0 + deb1) / 4;
    dea0 = dea0 - mov; dea1 = dea1 - mov; deb0 = deb0 - mov; deb1 = deb1 - mov;
    const double maxd = std::sqrt(std::max(sqn(dea0), sqn(dea1))) + std::sqrt(std::max(sqn(deb0), sqn(deb1)));
    if (maxd == 0) return 1.0;
    auto minpp = [&]() {
        double a = sqn(ea0 - eb0), b = sqn(ea0 - eb1), c = sqn(ea1 - eb0), d = sqn(ea1 - eb1);
        return std::min(std::min(a, b), std::min(c, d));
    };
    double d2 = ee_unclassified(ea0, ea1, eb0, eb1);
    double dF = d2 - th * th;
    if (dF <= 0) { d2 = minpp(); dF = d2 - th * th; }
    double dc = std::sqrt(d2);
    const double gap = eta * dF / (dc + th);
    double toc = 0.0;
    for (int it = 0; it < 10000; ++it) {
        const double tl = (1 - eta) * dF / ((dc + th) * maxd);
        if (!std::isfinite(tl) || tl <= 0.0) return std::max(0.0, std::min(toc, 1.0));
        ea0 = ea0 + tl * dea0; ea1 = ea1 + tl * dea1; eb0 = eb0 + tl * deb0; eb1 = eb1 + tl * deb1;
        d2 = ee_unclassified(ea0, ea1, eb0, eb1);
        dF = d2 - th * th;
        if (dF <= 0) { d2 = minpp(); dF = d2 - th * th; }
        dc = std::sqrt(d2);
        if (toc != 0.0 && (dF / (dc + th) < gap)) break;
        toc += tl;
        if (toc > 1.0) return 1.0;
    }
    return std::max(0.0, std::min(toc, 1.0));
}
double accd_pt(V3 p, V3 t0, V3 t1, V3 t2, V3 dp, V3 dt0, V3 dt1, V3 dt2, double eta, double th) {          // :281-329
    const V3 mov = (dt0 + dt1 + dt2 + dp) / 4;
    dt0 = dt0 - mov; dt1 = dt1 - mov; dt2 = dt2 - mov; dp = dp - mov;
    const double mx = std::max(std::max(sqn(dt0), sqn(dt1)), sqn(dt2));
    const double maxd = std::sqrt(sqn(dp)) + std::sqrt(mx);
    if (maxd == 0) return 1.0;
    double d2 = pt_unclassified(p, t0, t1, t2);
    double dc = std::sqrt(d2);
    const double gap = eta * (d2 - th * th) / (dc + th);
    double toc = 0.0;
    for (int it = 0; it < 10000; ++it) {
        const double tl = (1 - eta) * (d2 - th * th) / ((dc + th) * maxd);
        if (!std::isfinite(tl) || tl <= 0.0) return std::max(0.0, std::min(toc, 1.0));
        p = p + tl * dp; t0 = t0 + tl * dt0; t1 = t1 + tl * dt1; t2 = t2 + tl * dt2;
        d2 = pt_unclassified(p, t0, t1, t2);
        dc = std::sqrt(d2);
        if (toc != 0.0 && ((d2 - th * th) / (dc + th) < gap)) break;
        toc += tl;
        if (toc > 1.0) return 1.0;
    }
    return std::max(0.0, std::min(toc, 1.0));
}

// Newton step bound: limitStepByGround (FeasibleStep.cpp:8-40), partial CCD
// (ContactMechanics.cpp:3704-3752), CFL (IPCSolver.cpp:960-974), swept rebuild and full
// CCD (ContactMechanics.cpp:3754-3867), merge (IPCSolver.cpp:985-989).
// Self_largestFeasibleStepSize_CCD (ContactMechanics.cpp:3754-3867) after
// SpatialHash::build(mesh, p, alpha0): the swept trees at step alpha0, then the
// point-triangle and edge-edge phases; returns the full-CCD step.
double full_ccd(Mesh& m, const std::vector<V3>& p, double alpha0) {
    const double eta = 1.0 - 0.8;
    double full = alpha0;
    build_swept(m, p, full);
    {   // PT phase
        const double a = full;
        std::vector<double> lt(m.Vs, 1.0);
#pragma omp parallel
        {
            std::vector<int> cand;
#pragma omp for schedule(dynamic, 64)
            for (int s = 0; s < m.Vs; ++s) {
                const int vI = m.sv[s];
                m.faceTree.query(m.sweptPointBoxes[s], 0.0, cand);
                const AABB pb = sweptPoint(m.x[vI], p[vI], a);
                for (int f : cand) {
                    const auto& fv = m.faces[f];
                    if (isCollider(m.bt[vI]) && isCollider(m.bt[fv[0]]) && isCollider(m.bt[fv[1]]) && isCollider(m.bt[fv[2]])) continue;
                    if (vI == fv[0] || vI == fv[1] || vI == fv[2]) continue;
                    const AABB fb = sweptTri(m.x[fv[0]], p[fv[0]], m.x[fv[1]], p[fv[1]], m.x[fv[2]], p[fv[2]], a);
                    if (!pb.overlaps(fb, 0.0)) continue;
                    const double la = accd_pt(m.x[vI], m.x[fv[0]], m.x[fv[1]], m.x[fv[2]], -p[vI], -p[fv[0]], -p[fv[1]], -p[fv[2]], eta, 0);
                    lt[s] = std::min(la, lt[s]);
                }
            }
        }
        double mn = 1.0;
        if (m.Vs > 0) { mn = lt[0]; for (double v : lt) mn = std::min(mn, v); }
        full = std::min(full, mn);
    }
    {   // EE phase
        const double a = full;
        std::vector<double> le(m.Es, 1.0);
#pragma omp parallel
        {
            std::vector<int> cand;
#pragma omp for schedule(dynamic, 64)
            for (int eJ = 0; eJ < m.Es; ++eJ) {
                const auto& EJ = m.se[eJ];
                m.edgeTree.query(m.edgeBoxes[eJ], 0.0, cand);
                cand.erase(std::remove_if(cand.begin(), cand.end(), [eJ](int c) { return c <= eJ; }), cand.end());
                const AABB b1 = sweptEdge(m.x[EJ.first], p[EJ.first], m.x[EJ.second], p[EJ.second], a);
                for (int eI : cand) {
                    const auto& EI = m.se[eI];
                    if (isCollider(m.bt[EI.first]) && isCollider(m.bt[EI.second]) && isCollider(m.bt[EJ.first]) && isCollider(m.bt[EJ.second])) continue;
                    if (EI.first == EJ.first || EI.first == EJ.second || EI.second == EJ.first || EI.second == EJ.second || eI < eJ) continue;
                    const AABB b2 = sweptEdge(m.x[EI.first], p[EI.first], m.x[EI.second], p[EI.second], a);
                    if (!b1.overlaps(b2, 0.0)) continue;
                    const double la = accd_ee(m.x[EI.first], m.x[EI.second], m.x[EJ.first], m.x[EJ.second],
                                              -p[EI.first], -p[EI.second], -p[EJ.first], -p[EJ.second], eta, 0);
                    le[eJ] = std::min(la, le[eJ]);
                }
            }
        }
        double mn = 1.0;
        if (m.Es > 0) { mn = le[0]; for (double v : le) mn = std::min(mn, v); }
        full = std::min(full, mn);
    }
    return full;
}

void step_bound(Mesh& m, const std::vector<V3>& p, double out[5]) {
    const double slack = 0.8;
    double alpha = 1.0;
    {   // ground
        double mn = 1.0;
        for (int s = 0; s < m.Vs; ++s) {
            double best = 1.0;
            const int v = m.sv[s];
            const double nm = dot(m.gn, -p[v]);
            if (!(nm >= 0.0)) {
                const double dist = dot(m.gn, m.x[v]) - m.gd;
                const double cand = -dist / nm * slack;
                if (cand >= 0.0) best = cand;
            }
            mn = std::min(mn, best);
        }
        if (m.Vs > 0) alpha = std::min(alpha, mn);
    }
    const double aground = alpha;
    if (alpha <= 0) alpha = 1;
    const double eta = 1.0 - slack;
    double partial = alpha;
    if (!m.ccdAS.empty()) {
        std::vector<double> la(m.ccdAS.size());
#pragma omp parallel for schedule(dynamic, 64)
        for (long i = 0; i < (long)m.ccdAS.size(); ++i) {
            const auto& pr = m.ccdAS[i];
            if (pr.first < 0) {
                const int pt = m.sv[-pr.first - 1];
                const auto& f = m.faces[pr.second];
                la[i] = accd_pt(m.x[pt], m.x[f[0]], m.x[f[1]], m.x[f[2]], -p[pt], -p[f[0]], -p[f[1]], -p[f[2]], eta, 0.0);
            } else {
                const auto& e1 = m.se[pr.first]; const auto& e2 = m.se[pr.second];
                la[i] = accd_ee(m.x[e1.first], m.x[e1.second], m.x[e2.first], m.x[e2.second],
                                -p[e1.first], -p[e1.second], -p[e2.first], -p[e2.second], eta, 0.0);
            }
        }
        double mn = la[0];
        for (double v : la) mn = std::min(mn, v);
        partial = std::min(partial, mn);
    }
    double maxmove = 0.0;
    for (int s = 0; s < m.Vs; ++s) maxmove = std::max(maxmove, std::sqrt(sqn(p[m.sv[s]])));
    const double acfl = maxmove > 0.0 ? std::sqrt(m.hhat) / (maxmove * 2.0) : 1.0;
    const double full = full_ccd(m, p, alpha);
    alpha = std::min(alpha, acfl);
    if (partial > 2.0 * acfl) {
        alpha = std::min(partial, full);
        alpha = std::max(alpha, acfl);
    }
    out[0] = alpha; out[1] = aground; out[2] = partial; out[3] = acfl; out[4] = full;
}

// IglUtils::segTriIntersect (ContactMechanics.h:47-71) with Eigen 3.4 FullPivLU on the
// 3x3 system: complete pivoting on the first largest |a_ij| in column-major order,
// rank threshold 3 eps |max pivot|.
bool seg_tri(const V3& e0, const V3& e1, const V3& t0, const V3& t1, const V3& t2) {
    M3 A;
    const V3 c0 = t1 - t0, c1 = t2 - t0, c2 = e0 - e1;
    for (int i = 0; i < 3; ++i) { A(i, 0) = c0[i]; A(i, 1) = c1[i]; A(i, 2) = c2[i]; }
    const V3 n = cross(c0, c1);
    if (dot(n, e0 - t0) * dot(n, e1 - t0) > 0.0) return false;
    const double det = A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) - A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0))
                     + A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
    if (det == 0.0) return false;
    M3 lu = A;
    int rt[3], ct[3], nz = 3;
    double maxpiv = 0;
    for (int k = 0; k < 3; ++k) {
        int br = k, bc = k; double best = std::fabs(lu(k, k));
        for (int c = k; c < 3; ++c)
            for (int r = k; r < 3; ++r)
                if (std::fabs(lu(r, c)) > best) { best = std::fabs(lu(r, c)); br = r; bc = c; }
        if (best == 0.0) { nz = k; for (int i = k; i < 3; ++i) { rt[i] = i; ct[i] = i; } break; }
        if (best > maxpiv) maxpiv = best;
        rt[k] = br; ct[k] = bc;
        if (br != k) for (int c = 0; c < 3; ++c) std::swap(lu(k, c), lu(br, c));
        if (bc != k) for (int r = 0; r < 3; ++r) std::swap(lu(r, k), lu(r, bc));
        for (int r = k + 1; r < 3; ++r) lu(r, k) /= lu(k, k);
        for (int c = k + 1; c < 3; ++c)
            for (int r = k + 1; r < 3; ++r) lu(r, c) -= lu(r, k) * lu(k, c);
    }
    int rank = 0;
    const double thr = maxpiv * (std::numeric_limits<double>::epsilon() * 3.0);
    for (int i = 0; i < nz; ++i) rank += std::fabs(lu(i, i)) > thr;
    V3 rhs = e0 - t0;
    double c[3] = {rhs.x, rhs.y, rhs.z}, x[3] = {0, 0, 0};
    if (rank > 0) {
        for (int k = 0; k < 3; ++k) std::swap(c[k], c[rt[k]]);
        // Eigen's triangular_solve_vector skips a column whose right-hand side is exactly 0
        for (int i = 0; i < 3; ++i) if (c[i] != 0.0) for (int r = i + 1; r < 3; ++r) c[r] -= c[i] * lu(r, i);
        for (int i = rank - 1; i >= 0; --i) {
            if (c[i] == 0.0) continue;
            c[i] /= lu(i, i);
            for (int r = 0; r < i; ++r) c[r] -= c[i] * lu(r, i);
        }
        int q[3] = {0, 1, 2};
        for (int k = 0; k < 3; ++k) std::swap(q[k], q[ct[k]]);
        for (int i = 0; i < rank; ++i) x[q[i]] = c[i];
        for (int i = rank; i < 3; ++i) x[q[i]] = 0.0;
    }
    return x[0] >= 0.0 && x[1] >= 0.0 && x[0] + x[1] <= 1.0 && x[2] >= 0.0 && x[2] <= 1.0;
}

}  // namespace

struct oracle_mesh { Mesh m; };

extern "C" {

oracle_mesh* oracle_create(const oracle_scene_desc* d) {
    auto* om = new oracle_mesh;
    Mesh& m = om->m;
    m.V = d->V; m.T = d->T; m.F = d->F; m.Vs = d->Vs; m.Es = d->Es;
    auto v3s = [&](const double* p, std::vector<V3>& out) {
        out.resize(m.V);
        for (int i = 0; i < m.V; ++i) out[i] = V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    };
    v3s(d->x, m.x); v3s(d->v_rest, m.vrest); v3s(d->V_prev, m.vprev); v3s(d->xtilde, m.xt);
    m.mass.assign(d->masses, d->masses + m.V);
    m.bt.assign(d->boundary, d->boundary + m.V);
    m.tets.resize(m.T); m.dminv.resize(m.T); m.vol.assign(d->volumes, d->volumes + m.T);
    for (int t = 0; t < m.T; ++t) {
        for (int k = 0; k < 4; ++k) m.tets[t][k] = d->tets[4 * t + k];
        for (int k = 0; k < 9; ++k) m.dminv[t].m[k] = d->dm_inv[9 * t + k];
    }
    m.faces.resize(m.F);
    for (int f = 0; f < m.F; ++f) for (int k = 0; k < 3; ++k) m.faces[f][k] = d->faces[4 * f + k];
    m.sv.resize(m.Vs);
    for (int s = 0; s < m.Vs; ++s) m.sv[s] = (int)d->surf_verts[s];
    m.se.resize(m.Es);
    for (int e = 0; e < m.Es; ++e) m.se[e] = {(int)d->surf_edges[2 * e], (int)d->surf_edges[2 * e + 1]};
    m.Tr = d->Tr; m.Nb = d->Nb; m.bmode = d->bending;
    m.tris.resize(m.Tr); m.dmt.resize(m.Tr); m.area.assign(d->areas, d->areas + m.Tr);
    for (int t = 0; t < m.Tr; ++t) {
        for (int k = 0; k < 3; ++k) m.tris[t][k] = d->triangles[3 * t + k];
        for (int k = 0; k < 4; ++k) m.dmt[t][k] = d->dm_tri_inv[4 * t + k];
    }
    m.stretchS = d->stretch; m.shearS = d->shear; m.srate = d->strain_rate; m.plate = d->plate_rigidity;
    m.bverts.resize(m.Nb);
    for (int h = 0; h < m.Nb; ++h) for (int k = 0; k < 4; ++k) m.bverts[h][k] = d->bend_verts[4 * h + k];
    if (m.bmode == 0) {
        m.bq.resize(m.Nb);
        for (int h = 0; h < m.Nb; ++h) for (int k = 0; k < 16; ++k) m.bq[h][k] = d->bend_q[16 * h + k];
    } else {
        m.brest.assign(d->bend_rest, d->bend_rest + m.Nb);
        m.bweight.assign(d->bend_weight, d->bend_weight + m.Nb);
    }
    m.dt = d->dt; m.u = d->length_rate; m.r = d->volume_rate; m.drag = d->drag;
    m.hhat = d->hhat; m.fhat = d->fhat; m.fric = d->friction;
    m.gn = V3(d->ground_normal[0], d->ground_normal[1], d->ground_normal[2]); m.gd = d->ground_d;
    return om;
}
void oracle_destroy(oracle_mesh* om) { delete om; }
void oracle_set_threads(oracle_mesh* om, int threads) { om->m.threads = threads; if (threads > 0) omp_set_num_threads(threads); }
void oracle_set_state(oracle_mesh* om, const double* x, const double* vp, const double* xt) {
    Mesh& m = om->m;
    for (int i = 0; i < m.V; ++i) {
        if (x) m.x[i] = V3(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        if (vp) m.vprev[i] = V3(vp[3 * i], vp[3 * i + 1], vp[3 * i + 2]);
        if (xt) m.xt[i] = V3(xt[3 * i], xt[3 * i + 1], xt[3 * i + 2]);
    }
}
void oracle_active_set(oracle_mesh* om, int64_t counts[4]) {
    active_set(om->m);
    counts[0] = om->m.selfAS.size(); counts[1] = om->m.eeAS.size(); counts[2] = om->m.ccdAS.size(); counts[3] = om->m.envAS.size();
}
void oracle_get_active_set(const oracle_mesh* om, int* s4, int* e4, int* eie, int* ccd, int* env) {
    const Mesh& m = om->m;
    for (size_t i = 0; i < m.selfAS.size(); ++i) for (int k = 0; k < 4; ++k) s4[4 * i + k] = m.selfAS[i][k];
    for (size_t i = 0; i < m.eeAS.size(); ++i) { for (int k = 0; k < 4; ++k) e4[4 * i + k] = m.eeAS[i][k]; eie[2 * i] = m.eeIe[i].first; eie[2 * i + 1] = m.eeIe[i].second; }
    for (size_t i = 0; i < m.ccdAS.size(); ++i) { ccd[2 * i] = m.ccdAS[i].first; ccd[2 * i + 1] = m.ccdAS[i].second; }
    for (size_t i = 0; i < m.envAS.size(); ++i) env[i] = m.envAS[i];
}
void oracle_set_active_set(oracle_mesh* om, const int64_t c[4], const int* s4, const int* e4, const int* eie, const int* ccd, const int* env) {
    Mesh& m = om->m;
    m.selfAS.resize(c[0]); m.eeAS.resize(c[1]); m.eeIe.resize(c[1]); m.ccdAS.resize(c[2]); m.envAS.assign(env, env + c[3]);
    for (int64_t i = 0; i < c[0]; ++i) for (int k = 0; k < 4; ++k) m.selfAS[i][k] = s4[4 * i + k];
    for (int64_t i = 0; i < c[1]; ++i) { for (int k = 0; k < 4; ++k) m.eeAS[i][k] = e4[4 * i + k]; m.eeIe[i] = {eie[2 * i], eie[2 * i + 1]}; }
    for (int64_t i = 0; i < c[2]; ++i) m.ccdAS[i] = {ccd[2 * i], ccd[2 * i + 1]};
}
int oracle_assemble(oracle_mesh* om, double kappa, int64_t* nnz) {
    Blocks BH;
    const int n = gradient_and_hessian(om->m, kappa, BH);
    assemble_csr(om->m, BH);
    *nnz = (int64_t)om->m.inner.size();
    return n;
}
void oracle_get_gradient(const oracle_mesh* om, double* g) {
    for (int v = 0; v < om->m.V; ++v) { g[3 * v] = om->m.grad[v].x; g[3 * v + 1] = om->m.grad[v].y; g[3 * v + 2] = om->m.grad[v].z; }
}
void oracle_get_csr_scale(const oracle_mesh* om, double* s) {
    for (size_t i = 0; i < om->m.vscale.size(); ++i) s[i] = om->m.vscale[i];
}
void oracle_get_gradient_scale(const oracle_mesh* om, double* g) {
    for (int v = 0; v < om->m.V; ++v) { g[3 * v] = om->m.gabs[v].x; g[3 * v + 1] = om->m.gabs[v].y; g[3 * v + 2] = om->m.gabs[v].z; }
}
void oracle_get_csr(const oracle_mesh* om, int* outer, int* inner, double* val, double* rhs) {
    const Mesh& m = om->m;
    std::copy(m.outer.begin(), m.outer.end(), outer);
    std::copy(m.inner.begin(), m.inner.end(), inner);
    std::copy(m.val.begin(), m.val.end(), val);
    if (rhs) std::copy(m.rhs.begin(), m.rhs.end(), rhs);
}
double oracle_energy(oracle_mesh* om, double kappa) { return energy(om->m, kappa); }
void oracle_energy_terms(oracle_mesh* om, double kappa, double t[9]) { energy_terms(om->m, kappa, t); }
int oracle_kappa_gradients(oracle_mesh* om, double out[2]) { return kappa_gradients(om->m, out); }
int oracle_close_constraints(oracle_mesh* om, double dtol, int64_t counts[2]) {
    const int r = close_constraints(om->m, dtol);
    counts[0] = (int64_t)om->m.closeEnv.size(); counts[1] = (int64_t)om->m.closeSelf.size();
    return r;
}
int oracle_constraint_extrema(oracle_mesh* om, double out[2]) { return constraint_extrema(om->m, out); }
void oracle_clear_close_constraints(oracle_mesh* om) {
    Mesh& m = om->m;
    m.closeEnv.clear(); m.closeEnvVal.clear(); m.closeSelf.clear(); m.closeSelfVal.clear();
}
void oracle_set_soft(oracle_mesh* om, int n, const int* idx, const double* targets, double weight) {
    Mesh& m = om->m;
    m.softIdx.assign(idx, idx + n);
    m.softT.resize(n);
    for (int i = 0; i < n; ++i) m.softT[i] = V3(targets[3 * i], targets[3 * i + 1], targets[3 * i + 2]);
    m.softW = weight;
}
void oracle_friction_update(oracle_mesh* om, double kappa) { friction_initialize(om->m, kappa); }
void oracle_get_friction(const oracle_mesh* om, int64_t counts[2], double* lam, double* coord, double* basis, double* env_lam) {
    const Mesh& m = om->m;
    counts[0] = (int64_t)m.selfLast.size(); counts[1] = (int64_t)m.envLast.size();
    if (lam) std::copy(m.selfLam.begin(), m.selfLam.end(), lam);
    if (coord) for (size_t i = 0; i < m.distCoord.size(); ++i) { coord[2 * i] = m.distCoord[i][0]; coord[2 * i + 1] = m.distCoord[i][1]; }
    if (basis)
        for (size_t i = 0; i < m.tanBasis.size(); ++i)
            for (int k = 0; k < 2; ++k) { basis[6 * i + 3 * k] = m.tanBasis[i][k].x; basis[6 * i + 3 * k + 1] = m.tanBasis[i][k].y; basis[6 * i + 3 * k + 2] = m.tanBasis[i][k].z; }
    if (env_lam) std::copy(m.envLam.begin(), m.envLam.end(), env_lam);
}
void oracle_step_bound(oracle_mesh* om, const double* p, double out[5]) {
    std::vector<V3> pv(om->m.V);
    for (int i = 0; i < om->m.V; ++i) pv[i] = V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    step_bound(om->m, pv, out);
}
double oracle_full_ccd(oracle_mesh* om, const double* p, double alpha0) {
    std::vector<V3> pv(om->m.V);
    for (int i = 0; i < om->m.V; ++i) pv[i] = V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    return full_ccd(om->m, pv, alpha0);
}
void oracle_step_forward(oracle_mesh* om, const double* x0, const double* p, double a) {
    Mesh& m = om->m;
    for (int v = 0; v < m.V; ++v)
        if (isFree(m.bt[v])) m.x[v] = V3(x0[3 * v], x0[3 * v + 1], x0[3 * v + 2]) - a * V3(p[3 * v], p[3 * v + 1], p[3 * v + 2]);
}
void oracle_get_positions(const oracle_mesh* om, double* x) {
    for (int v = 0; v < om->m.V; ++v) { x[3 * v] = om->m.x[v].x; x[3 * v + 1] = om->m.x[v].y; x[3 * v + 2] = om->m.x[v].z; }
}
int oracle_is_intersected(oracle_mesh* om) {
    Mesh& m = om->m;
    for (int v : m.envAS) { const double d = dot(m.gn, m.x[v]) - m.gd; if (d * d <= 0.0) return 1; }
    build_static(m);
    int hit = 0;
#pragma omp parallel
    {
        std::vector<int> cand;
#pragma omp for schedule(dynamic, 64) reduction(| : hit)
        for (int f = 0; f < m.F; ++f) {
            const auto& fv = m.faces[f];
            m.edgeTree.query(triBox(m.x[fv[0]], m.x[fv[1]], m.x[fv[2]]), 0.0, cand);
            for (int e : cand) {
                const auto& E = m.se[e];
                if (E.first == fv[0] || E.first == fv[1] || E.first == fv[2] || E.second == fv[0] || E.second == fv[1] || E.second == fv[2]) continue;
                if (isCollider(m.bt[E.first]) && isCollider(m.bt[E.second]) && isCollider(m.bt[fv[0]]) && isCollider(m.bt[fv[1]]) && isCollider(m.bt[fv[2]])) continue;
                if (seg_tri(m.x[E.first], m.x[E.second], m.x[fv[0]], m.x[fv[1]], m.x[fv[2]])) { hit = 1; break; }
            }
        }
    }
    return hit;
}

void oracle_svd3(const double* F, double* U, double* s, double* V) {
    M3 A, Um, Vm;
    for (int i = 0; i < 9; ++i) A.m[i] = F[i];
    svd3(A, Um, s, Vm);
    for (int i = 0; i < 9; ++i) { U[i] = Um.m[i]; V[i] = Vm.m[i]; }
}
void oracle_snh_element(const double* Fp, double u, double r, double* P, double* H9, double* e) {
    M3 F;
    for (int i = 0; i < 9; ++i) F.m[i] = Fp[i];
    const M3 Pm = snh_pk1(F, u, r);
    for (int i = 0; i < 9; ++i) P[i] = Pm.m[i];
    snh_hessian(F, u, r, H9);
    *e = snh_energy_density(F, u, r);
}
void oracle_bw_element(const double* F6, double stretch, double shear, double k, double* P6, double* H36, double* e) {
    bw_pk1(F6, stretch, shear, k, P6);
    bw_hessian(F6, stretch, shear, k, H36);
    *e = bw_energy_density(F6, stretch, shear, k);
}
double oracle_edge_theta(const double* p, double* g, double* H) {
    return edge_theta(V3(p[0], p[1], p[2]), V3(p[3], p[4], p[5]), V3(p[6], p[7], p[8]), V3(p[9], p[10], p[11]), g, H);
}
void oracle_hinge_element(const double* p, double rest, double weight, double plate, double* g, double* H, double* e) {
    double ag[12], aH[144];
    const double th = edge_theta(V3(p[0], p[1], p[2]), V3(p[3], p[4], p[5]), V3(p[6], p[7], p[8]), V3(p[9], p[10], p[11]), ag, aH);
    const double df = hinge_difference(th, rest);
    const double sc = plate * weight;
    for (int i = 0; i < 12; ++i) g[i] = sc * df * ag[i];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) H[12 * i + j] = sc * (ag[i] * ag[j] + df * aH[12 * i + j]);
    make_pd<12>(H);
    *e = 0.5 * plate * weight * df * df;
}
int oracle_make_pd(int n, double* H) {
    switch (n) {
    case 3: return make_pd<3>(H);
    case 6: return make_pd<6>(H);
    case 9: return make_pd<9>(H);
    case 12: return make_pd<12>(H);
    default: return -1;
    }
}
void oracle_barrier(double d, double dh, double* b, double* g, double* h) { barrier_b(d, dh, *b); barrier_g(d, dh, *g); barrier_h(d, dh, *h); }
void oracle_distance(int type, const double* pts, double* d, double* g, double* H) {
    V3 p[4];
    for (int k = 0; k < 4; ++k) p[k] = V3(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]);
    deriv_any(type, p, *d, g, H);
}
int oracle_dtype_pt(const double* q) {
    return dType_PT(V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), V3(q[6], q[7], q[8]), V3(q[9], q[10], q[11]));
}
int oracle_dtype_ee(const double* q) {
    return dType_EE(V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), V3(q[6], q[7], q[8]), V3(q[9], q[10], q[11]));
}
double oracle_accd_pt(const double* q, const double* d, double eta, double th) {
    auto v = [](const double* a, int k) { return V3(a[3 * k], a[3 * k + 1], a[3 * k + 2]); };
    return accd_pt(v(q, 0), v(q, 1), v(q, 2), v(q, 3), v(d, 0), v(d, 1), v(d, 2), v(d, 3), eta, th);
}
double oracle_accd_ee(const double* q, const double* d, double eta, double th) {
    auto v = [](const double* a, int k) { return V3(a[3 * k], a[3 * k + 1], a[3 * k + 2]); };
    return accd_ee(v(q, 0), v(q, 1), v(q, 2), v(q, 3), v(d, 0), v(d, 1), v(d, 2), v(d, 3), eta, th);
}
int oracle_seg_tri_intersect(const double* q) {
    auto v = [&](int k) { return V3(q[3 * k], q[3 * k + 1], q[3 * k + 2]); };
    return seg_tri(v(0), v(1), v(2), v(3), v(4)) ? 1 : 0;
}
int64_t oracle_lbvh_query(int n, const double* bx, const double* qb, double gap, int* out, int64_t cap) {
    std::vector<AABB> boxes(n);
    for (int i = 0; i < n; ++i) { boxes[i].lo = V3(bx[6 * i], bx[6 * i + 1], bx[6 * i + 2]); boxes[i].hi = V3(bx[6 * i + 3], bx[6 * i + 4], bx[6 * i + 5]); }
    LBVH t; t.build(boxes);
    AABB q; q.lo = V3(qb[0], qb[1], qb[2]); q.hi = V3(qb[3], qb[4], qb[5]);
    std::vector<int> ids;
    t.query(q, gap, ids);
    for (int64_t i = 0; i < (int64_t)ids.size() && i < cap; ++i) out[i] = ids[i];
    return (int64_t)ids.size();
}

}  // extern "C"
