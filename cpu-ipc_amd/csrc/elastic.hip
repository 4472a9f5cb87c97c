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
    sp.h_blk_of_slot.assign(gsrc.size(), 0);
    for (int k = 0; k < sp.nblocks; ++k)
        for (int j = gptr[k]; j < gptr[k + 1]; ++j) sp.h_blk_of_slot[j] = k;
    sp.h_slot_h.swap(slot_h);
    sp.chunked = false;     // build_tet_chunks decides, once the scene ranges are known
    c.stagG.resize((size_t)3 * std::max<size_t>(vsrc.size(), 1));
    c.sbv.resize((size_t)9 * std::max(sp.nblocks, 1));
    c.sbf.resize(std::max(sp.nblocks, 1));
    c.grad.resize((size_t)3 * std::max(V, 1));
}

// ---------------------------------------------------------------------------------
// Chunk plan of the tet assembly.  The summation order of a block is fixed by the plan:
// within a chunk, phase by phase (local block 0..9), and within a phase by rounds that
// take the chunk's lanes in order; across chunks, in chunk order.  The plan depends only
// on each scene's own tets, so a scene packed into a batch is summed exactly as alone.
// ---------------------------------------------------------------------------------
namespace {
constexpr int kAccBytes = 9 * 8 + 2;   // one accumulator: 9 doubles and a 16-bit mask
constexpr int kChunkAcc = 216;         // block accumulators per chunk: 216 x 74 B < 16 KiB of LDS

uint32_t morton30(double px, double py, double pz, const double lo[3], double ext) {
    auto q = [&](double v, int k) {
        const double t = ext > 0 ? (v - lo[k]) / ext : 0.0;
        return (uint32_t)std::max(0.0, std::min(t * 1024.0, 1023.0));
    };
    auto ex = [](uint32_t v) {
        v = (v * 0x00010001u) & 0xFF0000FFu; v = (v * 0x00000101u) & 0x0F00F00Fu;
        v = (v * 0x00000011u) & 0xC30C30C3u; v = (v * 0x00000005u) & 0x49249249u;
        return v;
    };
    return (ex(q(px, 0)) << 2) | (ex(q(py, 1)) << 1) | ex(q(pz, 2));
}
}  // namespace

__global__ void chunk_static_gather(int NP, int T, const int* __restrict__ ch_tet, const int4* __restrict__ tets,
                                    const double* __restrict__ dminv, const double* __restrict__ vol, int4* __restrict__ ctet,
                                    double* __restrict__ cdm, double* __restrict__ cvol) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NP) return;
    int t = ch_tet[p];
    if (t < 0) t = ch_tet[p - p % kChunk];
    ctet[p] = tets[t];
    cvol[p] = vol[t];
#pragma unroll
    for (int k = 0; k < 9; ++k) cdm[(size_t)k * NP + p] = dminv[(size_t)k * T + t];
}

void build_tet_chunks(Context& c, const double* hx) {
    Scene& s = c.sc;
    StaticPattern& sp = c.sp;
    const int T = s.T;
    sp.chunked = false;
    auto staged = [&] { c.stagH.resize((size_t)9 * std::max<size_t>(sp.h_blk_of_slot.size(), 1)); };
    if (T == 0 || s.Tr > 0 || s.Nb > 0 || !hx) { staged(); return; }
    // tets of each scene in Morton order of their centroids (the scene's own box), ties by index
    std::vector<int> order(T);
    const std::vector<int>& toff = s.h_toff;
    for (size_t k = 0; k + 1 < toff.size(); ++k) {
        const int t0 = toff[k], t1 = toff[k + 1];
        if (t1 <= t0) continue;
        std::vector<double> cen(3 * (size_t)(t1 - t0));
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int t = t0; t < t1; ++t)
            for (int q = 0; q < 3; ++q) {
                double m = 0;
                for (int a = 0; a < 4; ++a) m += hx[3 * (size_t)s.h_tets[4 * (size_t)t + a] + q];
                m *= 0.25;
                cen[3 * (size_t)(t - t0) + q] = m;
                lo[q] = std::min(lo[q], m); hi[q] = std::max(hi[q], m);
            }
        const double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        std::vector<std::pair<uint32_t, int>> o((size_t)(t1 - t0));
        for (int t = t0; t < t1; ++t) {
            const double* m = &cen[3 * (size_t)(t - t0)];
            o[t - t0] = {morton30(m[0], m[1], m[2], lo, ext), t};
        }
        std::sort(o.begin(), o.end());
        for (int t = t0; t < t1; ++t) order[t] = o[t - t0].second;
    }
    auto target = [&](int t, int lb) {
        const int j = sp.h_slot_h[10 * (size_t)t + lb];
        return j < 0 ? -1 : sp.h_blk_of_slot[j];
    };
    // one wavefront (up to kChunk tets) per chunk: the rounds need no work-group barrier.  A
    // chunk also ends before its block accumulators would exceed kChunkAcc, which bounds the
    // LDS of a work-group (10 work-groups per CU) so that every chunk is resident at once.
    constexpr int C = kChunk;
    std::vector<int> ch_first, ch_end;   // positions in `order` of every chunk
    {
        std::vector<int> mark(sp.nblocks, -1);
        for (size_t k = 0; k + 1 < toff.size(); ++k) {
            int p = toff[k];
            while (p < toff[k + 1]) {
                const int ch = (int)ch_first.size();
                ch_first.push_back(p);
                int na = 0;
                while (p < toff[k + 1] && p - ch_first.back() < C) {
                    int fresh = 0;
                    for (int lb = 0; lb < 10; ++lb) {
                        const int b = target(order[p], lb);
                        if (b >= 0 && mark[b] != ch) {
                            bool dup = false;   // a block twice in one tet counts once
                            for (int l2 = 0; l2 < lb; ++l2) dup |= target(order[p], l2) == b;
                            fresh += dup ? 0 : 1;
                        }
                    }
                    if (na + fresh > kChunkAcc) break;
                    for (int lb = 0; lb < 10; ++lb) {
                        const int b = target(order[p], lb);
                        if (b >= 0) mark[b] = ch;
                    }
                    na += fresh;
                    ++p;
                }
                ch_end.push_back(p);
            }
        }
    }
    const int nch = (int)ch_first.size();
    auto chunk_end = [&](int ch) { return ch_end[ch]; };
    // per chunk: block accumulators (distinct blocks of its tets) and vertex accumulators
    // (distinct vertices); per lane 10 block codes then 4 corner codes, accumulator << 6 |
    // round; per chunk 10 + 4 round counts
    std::vector<int> ch_tet((size_t)nch * C, -1), acc_ptr(nch + 1, 0), acc_blk, vacc_ptr(nch + 1, 0), vacc_vtx;
    std::vector<unsigned short> code((size_t)nch * 14 * C, 0xFFFF);
    std::vector<unsigned char> rounds((size_t)nch * 16, 0);   // 14 used, padded to 16 bytes
    std::vector<int> tg, vg, cnt;
    int maxlds = 0;
    for (int ch = 0; ch < nch; ++ch) {
        const int p0 = ch_first[ch], p1 = chunk_end(ch), n = p1 - p0;
        tg.clear();
        vg.clear();
        for (int l = 0; l < n; ++l) {
            const int t = order[p0 + l];
            ch_tet[(size_t)ch * C + l] = t;
            for (int lb = 0; lb < 10; ++lb) {
                const int b = target(t, lb);
                if (b >= 0) tg.push_back(b);
            }
            for (int a = 0; a < 4; ++a) vg.push_back(s.h_tets[4 * (size_t)t + a]);
        }
        std::sort(tg.begin(), tg.end());
        tg.erase(std::unique(tg.begin(), tg.end()), tg.end());
        std::sort(vg.begin(), vg.end());
        vg.erase(std::unique(vg.begin(), vg.end()), vg.end());
        const int na = (int)tg.size(), nv = (int)vg.size();   // na <= 640, nv <= 256: 10-bit fields
        acc_ptr[ch + 1] = acc_ptr[ch] + na;
        acc_blk.insert(acc_blk.end(), tg.begin(), tg.end());
        vacc_ptr[ch + 1] = vacc_ptr[ch] + nv;
        vacc_vtx.insert(vacc_vtx.end(), vg.begin(), vg.end());
        maxlds = std::max(maxlds, std::max(na * kAccBytes, nv * 24));   // vertex accumulators alias the block ones
        for (int ph = 0; ph < 14; ++ph) {
            const bool blk = ph < 10;
            const std::vector<int>& keys = blk ? tg : vg;
            cnt.assign(keys.size(), 0);
            int rmax = 0;
            for (int l = 0; l < n; ++l) {
                const int t = order[p0 + l];
                const int key = blk ? target(t, ph) : s.h_tets[4 * (size_t)t + (ph - 10)];
                if (key < 0) continue;
                const int a = (int)(std::lower_bound(keys.begin(), keys.end(), key) - keys.begin());
                const int r = cnt[a]++;
                if (r > 63) { staged(); return; }   // the 6-bit round field
                code[((size_t)ch * 14 + ph) * C + l] = (unsigned short)((a << 6) | r);
                rmax = std::max(rmax, r + 1);
            }
            rounds[(size_t)ch * 16 + ph] = (unsigned char)rmax;
        }
    }
    // blocks and vertices touched by one chunk are written by it; the others (and those no
    // tet touches) are completed from their partials, in chunk order
    auto runs = [](const std::vector<int>& keys, int n, std::vector<int>& list, std::vector<int>& ptr, std::vector<int>& dst) {
        std::vector<int> nof(n, 0), cursor(n, -1);
        for (int k : keys) nof[k]++;
        ptr.assign(1, 0);
        list.clear();
        auto completed = [&](int k) { return nof[k] != 1; };
        for (int k = 0; k < n; ++k)
            if (completed(k)) {
                cursor[k] = ptr.back();
                list.push_back(k);
                ptr.push_back(ptr.back() + nof[k]);
            }
        dst.resize(keys.size());
        for (size_t i = 0; i < keys.size(); ++i)   // chunks in order: partial runs in chunk order
            dst[i] = completed(keys[i]) ? -(cursor[keys[i]]++ + 1) : keys[i];
    };
    std::vector<int> cp_blk, cp_ptr, acc_dst, cpv_vtx, cpv_ptr, vacc_dst;
    runs(acc_blk, sp.nblocks, cp_blk, cp_ptr, acc_dst);
    runs(vacc_vtx, s.V, cpv_vtx, cpv_ptr, vacc_dst);
    auto up = [&](auto& d, const auto& h) {
        d.resize(std::max<size_t>(h.size(), 1));
        if (!h.empty()) CIPC_HIP(hipMemcpy(d.p, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice));
    };
    up(sp.ch_tet, ch_tet); up(sp.ch_code, code); up(sp.ch_rounds, rounds);
    up(sp.ch_acc_ptr, acc_ptr); up(sp.ch_acc_dst, acc_dst);
    up(sp.ch_vacc_ptr, vacc_ptr); up(sp.ch_vacc_dst, vacc_dst);
    up(sp.cp_blk, cp_blk); up(sp.cp_ptr, cp_ptr);
    up(sp.cpv_vtx, cpv_vtx); up(sp.cpv_ptr, cpv_ptr);
    sp.part.resize((size_t)9 * std::max(cp_ptr.back(), 1));
    sp.partf.resize((size_t)std::max(cp_ptr.back(), 1));
    sp.vpart.resize((size_t)3 * std::max(cpv_ptr.back(), 1));
    sp.chunk = C;
    sp.nchunks = nch;
    sp.ncomplete = (int)cp_blk.size();
    sp.ncomplete_v = (int)cpv_vtx.size();
    sp.lds_bytes = ((maxlds + 15) / 16) * 16;
    // the tets' static data in chunk order (coalesced loads in the chunk kernel); a padding
    // lane holds its chunk's first tet
    const size_t NP = (size_t)nch * C;
    sp.ctet.resize(NP);
    sp.cdm.resize(9 * NP);
    sp.cvol.resize(NP);
    hipLaunchKernelGGL(chunk_static_gather, dim3(blocks_for(NP, 256)), dim3(256), 0, c.stream, (int)NP, T, sp.ch_tet.p,
                       s.tets.p, s.dminv.p, s.vol.p, sp.ctet.p, sp.cdm.p, sp.cvol.p);
    CIPC_CHECK_LAUNCH();
    CIPC_HIP(hipStreamSynchronize(c.stream));
    sp.chunked = true;
    c.stagH.release();
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

// Per-tet state of the block formula, in the frames of F = U diag(s) V^T (U, V rotations):
// with beta_a = V^T B_a, F B_a = U (s o beta_a), cof F B_a = U (s' o beta_a) (s' = s1 s2,
// s0 s2, s0 s1), B_a . B_b = beta_a . beta_b and F (B_a x B_b) = U (s o (beta_a x beta_b)),
// so every block is K_ab = U X_ab U^T with
//   X_ab = A (beta_a . beta_b) I - a2 hat(s o (beta_a x beta_b)) + 4 a3 (s o beta_a)(s o beta_b)^T
//          + r (s' o beta_a)(s' o beta_b)^T + P_ab,
// P_ab the projection of the negative modes (SNH::Sw).  Identical in real arithmetic to the
// world-frame form of the element stiffness; 44 doubles of state per tet instead of 90.
struct TetSetup {
    double beta[4][3];
    M33 U;
    double s[3], sc3[3];   // singular values (s2 signed) and their cofactor products s'
    double A, a2, c1, r;   // Hessian coefficients, c1 = 4 a3
    double Sw[3][3], pp[3], pm[3];
    bool project;
    int vid[4];
    double sc;             // vol dt^2
};

// Per-tet set-up shared by the staged and the chunked kernels, in two halves so that the
// corner forces can be used (and dropped) before the SVD.  First half: F, the invariants and
// the corner forces f_a = vol P B_a (dt^2 is applied where they are summed, as
// IPCSolver.cpp:149-151).  The tet's static data (indices, Dm^-1 with stride NP, volume) is
// at position p; its material (a batch's per-tet Lame parameters) at tet index t.
__device__ inline void tet_setup_forces(TetSetup& q, SNH& e, d3 B[4], d3 f[4], double& u, double& r, int t, int p, int NP,
                                        const int4* __restrict__ tets, const double* __restrict__ x,
                                        const double* __restrict__ dminv, const double* __restrict__ vol, const Lame& lame,
                                        double dt2) {
    load_tet(p, NP, tets, x, dminv, e.F, B);
    const int4 tv = tets[p];
    q.vid[0] = tv.x; q.vid[1] = tv.y; q.vid[2] = tv.z; q.vid[3] = tv.w;
    lame.at(t, u, r);
    snh_invariants(e, u, r);
    const double vt = vol[p];
#pragma unroll
    for (int a = 0; a < 4; ++a) f[a] = vt * (e.pa * mvec(e.F, B[a]) + e.pb * mvec(e.C, B[a]));
    q.sc = vt * dt2;
}
// Second half: SVD frames, eigenvalues and the block formula's state.
__device__ inline void tet_setup_modes(TetSetup& q, SNH& e, const d3 B[4], double u, double r) {
    snh_modes(e, u, r);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) q.beta[a][k] = dot(mcol(e.V, k), B[a]);
    q.U = e.U;
#pragma unroll
    for (int k = 0; k < 3; ++k) q.s[k] = e.s[k];
    q.sc3[0] = e.s[1] * e.s[2]; q.sc3[1] = e.s[0] * e.s[2]; q.sc3[2] = e.s[0] * e.s[1];
    q.A = e.A; q.a2 = e.a2; q.c1 = 4.0 * e.a3; q.r = e.r;
    q.project = e.project;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        q.pp[i] = e.pp[i]; q.pm[i] = e.pm[i];
#pragma unroll
        for (int j = 0; j < 3; ++j) q.Sw[i][j] = e.Sw[i][j];
    }
}

// Block K_ab (3x3) of the projected element stiffness before the vol dt^2 factor.
__device__ inline void snh_block(const TetSetup& q, int a, int b, double K[3][3]) {
    const double* ba = q.beta[a];
    const double* bb = q.beta[b];
    const double dab = ba[0] * bb[0] + ba[1] * bb[1] + ba[2] * bb[2];
    const double w[3] = {q.s[0] * (ba[1] * bb[2] - ba[2] * bb[1]), q.s[1] * (ba[2] * bb[0] - ba[0] * bb[2]),
                         q.s[2] * (ba[0] * bb[1] - ba[1] * bb[0])};
    const double h[3][3] = {{0, -w[2], w[1]}, {w[2], 0, -w[0]}, {-w[1], w[0], 0}};
    double X[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            X[i][j] = (i == j ? q.A * dab : 0.0) - q.a2 * h[i][j] + q.c1 * (q.s[i] * ba[i]) * (q.s[j] * bb[j])
                    + q.r * (q.sc3[i] * ba[i]) * (q.sc3[j] * bb[j]);
    if (q.project) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) X[i][j] += q.Sw[i][j] * ba[i] * bb[j];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int i = (k == 0) ? 1 : 0, j = (k == 2) ? 1 : 2;
            X[i][i] += q.pp[k] * ba[j] * bb[j];
            X[j][j] += q.pp[k] * ba[i] * bb[i];
            X[i][j] += q.pm[k] * ba[j] * bb[i];
            X[j][i] += q.pm[k] * ba[i] * bb[j];
        }
    }
    double M[3][3];   // U X
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int j = 0; j < 3; ++j) M[r][j] = q.U.a[r][0] * X[0][j] + q.U.a[r][1] * X[1][j] + q.U.a[r][2] * X[2][j];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) K[r][c] = M[r][0] * q.U.a[c][0] + M[r][1] * q.U.a[c][1] + M[r][2] * q.U.a[c][2];
}

// Local block lb as stored (row vertex max(va, vb), column vertex min): vol dt^2 K_ab, or
// its transpose when va < vb (a diagonal block is symmetric and stored whole), rounded to
// the staging type H.
template <typename H>
__device__ inline void tet_block(const TetSetup& q, int lb, double o[9]) {
    const int a = lb_a(lb), b = lb_b(lb);
    double K[3][3];
    snh_block(q, a, b, K);
    const bool tr = q.vid[a] < q.vid[b];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) o[3 * i + k] = (double)(H)(q.sc * (tr ? K[k][i] : K[i][k]));
}

// Staged form (scenes with cloth elements): every (tet, local block) and (tet, corner) has
// a slot fixed when the pattern is built; the block gather sums the slots of a block.
// H: staging type of the Hessian blocks (double; float for the mixed form)
template <typename H>
__global__ __launch_bounds__(256) void snh_element_kernel(int T, const int4* __restrict__ tets, const double* __restrict__ x,
                                                          const double* __restrict__ dminv, const double* __restrict__ vol,
                                                          Lame lame, double dt2, const int* __restrict__ slot_h,
                                                          const int* __restrict__ slot_g, H* __restrict__ stagH,
                                                          double* __restrict__ stagG) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    TetSetup q;
    SNH e;
    d3 B[4], f[4];
    double u, r;
    tet_setup_forces(q, e, B, f, u, r, t, t, T, tets, x, dminv, vol, lame, dt2);
#pragma unroll
    for (int a = 0; a < 4; ++a) st3(stagG, slot_g[4 * t + a], f[a]);
    tet_setup_modes(q, e, B, u, r);
#pragma unroll
    for (int lb = 0; lb < 10; ++lb) {
        const int j = slot_h[10 * t + lb];
        if (j < 0) continue;   // a vertex of the block is not free
        double o[9];
        tet_block<H>(q, lb, o);
        H* d = stagH + (size_t)9 * j;
#pragma unroll
        for (int i = 0; i < 9; ++i) d[i] = (H)o[i];
    }
}

// Orders the LDS accesses of one wavefront: a compiler fence only.  A wavefront's LDS
// instructions execute in issue order, so what a lane wrote before this point is what any
// lane of the same wavefront reads after it.
__device__ inline void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// g = m (x - x~) + drag m (x - V_prev) (IPCSolver.cpp:118-128), to which the elastic forces
// are added.
__device__ inline d3 inertia_gradient(int v, const double* __restrict__ x, const double* __restrict__ xt,
                                      const double* __restrict__ vprev, const double* __restrict__ mass, double drag) {
    const d3 xv = ld3(x, v);
    const double m = mass[v];
    return m * (xv - ld3(xt, v)) + (drag * m) * (xv - ld3(vprev, v));
}

// Code and round count of phase PH (0..9 local blocks, 10..13 corners) from the packed words.
template <int PH>
__device__ __forceinline__ unsigned phase_code(const unsigned cdp[7]) { return (cdp[PH / 2] >> (16 * (PH % 2))) & 0xFFFFu; }
template <int PH>
__device__ __forceinline__ int phase_rounds(const unsigned rdp[4]) { return (int)((rdp[PH / 4] >> (8 * (PH % 4))) & 0xFFu); }

// One block phase of the chunked form: local block LB of every lane's tet into the chunk's
// accumulators, round by round (LB is a template argument so that every index into the
// tet's state is static).
template <typename H, int LB>
__device__ __forceinline__ void chunk_phase(TetSetup& q, const unsigned cdp[7], const unsigned rdp[4], double* acc,
                                            unsigned short* msk) {
    // the frame coordinates pass through an empty asm statement: products that several blocks
    // share are then recomputed in each phase rather than kept live across all ten (which
    // would double the register footprint and halve the occupancy)
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) __asm__ volatile("" : "+v"(q.beta[a][k]));
    const unsigned cd = phase_code<LB>(cdp);
    const int R = phase_rounds<LB>(rdp);
    double o[9];
    // (a block outside the pattern, or a padding lane, computes nothing; the branch also
    // keeps the phases' arithmetic apart, which bounds the register pressure)
    if (cd != 0xFFFFu) tet_block<H>(q, LB, o);
    const int mine = cd == 0xFFFFu ? -1 : (int)(cd & 63u);
    const unsigned a = cd >> 6;
    // zero entries add nothing (an accumulator starts at +0 and never holds -0, so adding +0
    // in their place is exact) and set no mask bit; the update is branch-free so that its
    // nine LDS reads go out together
    double v[9];
    unsigned m = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const bool nz = o[i] != 0.0;
        v[i] = nz ? o[i] : 0.0;
        m |= nz ? (1u << i) : 0u;
    }
    for (int r = 0; r < R; ++r) {
        if (mine == r) {
            double cur[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) cur[i] = acc[9 * a + i];
            const unsigned short mo = msk[a];
#pragma unroll
            for (int i = 0; i < 9; ++i) acc[9 * a + i] = cur[i] + v[i];
            msk[a] = (unsigned short)(mo | m);
        }
        wave_lds_order();
    }
    // keep the next phase's arithmetic out of this one (it would only raise register pressure)
    __builtin_amdgcn_sched_barrier(0);
}

// One corner phase: dt^2 f_A of every lane's tet into the chunk's vertex accumulators.
template <int A>
__device__ __forceinline__ void corner_phase(const d3 f[4], double dt2, const unsigned cdp[7], const unsigned rdp[4],
                                             double* vacc) {
    const unsigned cd = phase_code<10 + A>(cdp);
    const int R = phase_rounds<10 + A>(rdp);
    const int mine = cd == 0xFFFFu ? -1 : (int)(cd & 63u);
    const unsigned a = cd >> 6;
    const d3 g = dt2 * f[A];
    for (int r = 0; r < R; ++r) {
        if (mine == r) {
            vacc[3 * a] += g.x;
            vacc[3 * a + 1] += g.y;
            vacc[3 * a + 2] += g.z;
        }
        wave_lds_order();
    }
}

// Chunked form: one wavefront per chunk of 64 tets, one lane per tet.  The corner forces go
// to the chunk's vertex accumulators and then, for each local block lb, the lanes add their
// block to the chunk's block accumulators, in the planned rounds (no two lanes of a round
// share an accumulator).  Block entries: zero entries are skipped and the entries that
// received a nonzero value recorded (the reference drops zero entries before summing
// duplicates).  Then each accumulator goes to its block or vertex (a vertex with its inertia
// term), or to its partial when other chunks touch the block or vertex too.
template <typename H>
__global__ __launch_bounds__(kChunk) void snh_chunk_kernel(int NP, const int4* __restrict__ ctet, const double* __restrict__ x,
                                                           const double* __restrict__ cdm, const double* __restrict__ cvol,
                                                           Lame lame, double dt2, const int* __restrict__ ch_tet,
                                                           const unsigned short* __restrict__ ch_code,
                                                           const unsigned char* __restrict__ ch_rounds,
                                                           const int* __restrict__ acc_ptr, const int* __restrict__ acc_dst,
                                                           const int* __restrict__ vacc_ptr, const int* __restrict__ vacc_dst,
                                                           const double* __restrict__ xt, const double* __restrict__ vprev,
                                                           const double* __restrict__ mass, double drag, double* __restrict__ sbv,
                                                           unsigned short* __restrict__ sbf, double* __restrict__ part,
                                                           unsigned short* __restrict__ partf, double* __restrict__ grad,
                                                           double* __restrict__ vpart) {
    extern __shared__ double lds_acc[];
    const int ch = blockIdx.x, lane = threadIdx.x;
    const int a0 = acc_ptr[ch], na = acc_ptr[ch + 1] - a0;
    const int v0 = vacc_ptr[ch], nv = vacc_ptr[ch + 1] - v0;
    double* acc = lds_acc;
    double* vacc = lds_acc;   // the corner phases finish before the block phases start
    unsigned short* msk = reinterpret_cast<unsigned short*>(acc + 9 * (size_t)na);
    const int p = ch * kChunk + lane;
    // the lane's 14 codes and the chunk's 14 round counts, loaded before any LDS ordering point
    // so that their latency overlaps the tet's evaluation
    unsigned cdp[7], rdp[4];
#pragma unroll
    for (int i = 0; i < 7; ++i)
        cdp[i] = (unsigned)ch_code[((size_t)ch * 14 + 2 * i) * kChunk + lane] |
                 ((unsigned)ch_code[((size_t)ch * 14 + 2 * i + 1) * kChunk + lane] << 16);
    const uint4 rd = reinterpret_cast<const uint4*>(ch_rounds)[ch];
    rdp[0] = rd.x; rdp[1] = rd.y; rdp[2] = rd.z; rdp[3] = rd.w;
    for (int k = lane; k < 3 * nv; k += kChunk) vacc[k] = 0.0;
    // padding lanes (a scene's last chunk) evaluate the chunk's first tet (their static data
    // is that tet's) and add nothing: their codes are all 0xFFFF
    const int t = ch_tet[p];
    const int tt = t >= 0 ? t : ch_tet[ch * kChunk];
    TetSetup q;
    SNH e;
    d3 B[4], f[4];
    double u, r;
    tet_setup_forces(q, e, B, f, u, r, tt, p, NP, ctet, x, cdm, cvol, lame, dt2);
    wave_lds_order();
    corner_phase<0>(f, dt2, cdp, rdp, vacc);
    corner_phase<1>(f, dt2, cdp, rdp, vacc);
    corner_phase<2>(f, dt2, cdp, rdp, vacc);
    corner_phase<3>(f, dt2, cdp, rdp, vacc);
    for (int k = lane; k < nv; k += kChunk) {
        const int d = vacc_dst[v0 + k];
        const d3 g = {vacc[3 * k], vacc[3 * k + 1], vacc[3 * k + 2]};
        if (d >= 0) st3(grad, d, inertia_gradient(d, x, xt, vprev, mass, drag) + g);
        else st3(vpart, -d - 1, g);
    }
    wave_lds_order();
    for (int k = lane; k < 9 * na; k += kChunk) acc[k] = 0.0;
    for (int k = lane; k < na; k += kChunk) msk[k] = 0;
    wave_lds_order();
    __builtin_amdgcn_sched_barrier(0);
    tet_setup_modes(q, e, B, u, r);
    chunk_phase<H, 0>(q, cdp, rdp, acc, msk);
    chunk_phase<H, 1>(q, cdp, rdp, acc, msk);
    chunk_phase<H, 2>(q, cdp, rdp, acc, msk);
    chunk_phase<H, 3>(q, cdp, rdp, acc, msk);
    chunk_phase<H, 4>(q, cdp, rdp, acc, msk);
    chunk_phase<H, 5>(q, cdp, rdp, acc, msk);
    chunk_phase<H, 6>(q, cdp, rdp, acc, msk);
    chunk_phase<H, 7>(q, cdp, rdp, acc, msk);
    chunk_phase<H, 8>(q, cdp, rdp, acc, msk);
    chunk_phase<H, 9>(q, cdp, rdp, acc, msk);
    for (int k = lane; k < na; k += kChunk) {
        const int d = acc_dst[a0 + k];
        double* o = d >= 0 ? sbv + (size_t)9 * d : part + (size_t)9 * (-d - 1);
#pragma unroll
        for (int i = 0; i < 9; ++i) o[i] = acc[9 * (size_t)k + i];
        if (d >= 0) sbf[d] = msk[k];
        else partf[-d - 1] = msk[k];
    }
}

// Blocks (i < nb) and vertices (i >= nb) touched by several chunks or by none: the sum of
// their partials in chunk order; a vertex with its inertia term.
__global__ __launch_bounds__(256) void chunk_complete_kernel(int nb, int nv, const int* __restrict__ blk, const int* __restrict__ ptr,
                                                             const double* __restrict__ part,
                                                             const unsigned short* __restrict__ partf, double* __restrict__ sbv,
                                                             unsigned short* __restrict__ sbf, const int* __restrict__ vtx,
                                                             const int* __restrict__ vptr, const double* __restrict__ vpart,
                                                             const double* __restrict__ x, const double* __restrict__ xt,
                                                             const double* __restrict__ vprev, const double* __restrict__ mass,
                                                             double drag, double* __restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) {
        double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        unsigned m = 0;
        for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
#pragma unroll
            for (int k = 0; k < 9; ++k) acc[k] += part[(size_t)9 * q + k];
            m |= partf[q];
        }
        const int b = blk[i];
#pragma unroll
        for (int k = 0; k < 9; ++k) sbv[(size_t)9 * b + k] = acc[k];
        sbf[b] = (unsigned short)m;
    } else if (i < nb + nv) {
        const int j = i - nb;
        d3 g = {0, 0, 0};
        for (int q = vptr[j]; q < vptr[j + 1]; ++q) g = g + ld3(vpart, q);
        const int v = vtx[j];
        st3(grad, v, inertia_gradient(v, x, xt, vprev, mass, drag) + g);
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

static void elastic_staged(Context& c, hipStream_t st, double dt2);

void elastic_assemble(Context& c, hipStream_t st) {
    Scene& s = c.sc;
    const double dt2 = s.dt * s.dt;
    StaticPattern& sp = c.sp;
    if (sp.chunked) {
        auto k = c.hessian_fp32 ? snh_chunk_kernel<float> : snh_chunk_kernel<double>;
        hipLaunchKernelGGL(k, dim3(sp.nchunks), dim3(kChunk), sp.lds_bytes, st, sp.nchunks * kChunk, sp.ctet.p, s.x.p, sp.cdm.p,
                           sp.cvol.p, scene_lame(s), dt2, sp.ch_tet.p, sp.ch_code.p, sp.ch_rounds.p, sp.ch_acc_ptr.p,
                           sp.ch_acc_dst.p, sp.ch_vacc_ptr.p, sp.ch_vacc_dst.p, s.xt.p, s.vprev.p, s.mass.p, s.drag, c.sbv.p,
                           c.sbf.p, sp.part.p, sp.partf.p, c.grad.p, sp.vpart.p);
        CIPC_CHECK_LAUNCH();
        const int nc = sp.ncomplete + sp.ncomplete_v;
        if (nc > 0)
            hipLaunchKernelGGL(chunk_complete_kernel, dim3(blocks_for(nc, 256)), dim3(256), 0, st, sp.ncomplete, sp.ncomplete_v,
                               sp.cp_blk.p, sp.cp_ptr.p, sp.part.p, sp.partf.p, c.sbv.p, c.sbf.p, sp.cpv_vtx.p, sp.cpv_ptr.p,
                               sp.vpart.p, s.x.p, s.xt.p, s.vprev.p, s.mass.p, s.drag, c.grad.p);
        CIPC_CHECK_LAUNCH();
        hipLaunchKernelGGL(mass_diag_kernel, dim3(blocks_for(s.V, 256)), dim3(256), 0, st, s.V, sp.diag_block.p, s.mass.p,
                           s.drag, c.sbv.p, c.sbf.p);
        CIPC_CHECK_LAUNCH();
    } else {
        elastic_staged(c, st, dt2);
        hipLaunchKernelGGL(vertex_gather_kernel, dim3(blocks_for(s.V, 256)), dim3(256), 0, st, s.V, s.T, s.x.p, s.xt.p,
                           s.vprev.p, s.mass.p, s.drag, dt2, c.sp.vptr.p, c.stagG.p, c.grad.p);
        CIPC_CHECK_LAUNCH();
    }
    if (c.n_soft > 0) {
        hipLaunchKernelGGL(soft_kernel, dim3(blocks_for(c.n_soft, 256)), dim3(256), 0, st, c.n_soft, c.soft_idx.p, c.soft_t.p,
                           c.soft_w, s.x.p, c.sp.diag_block.p, c.grad.p, c.sbv.p, c.sbf.p);
        CIPC_CHECK_LAUNCH();
    }
}

// Staged form: element kernels write every contribution to its slot, the gathers sum them.
static void elastic_staged(Context& c, hipStream_t st, double dt2) {
    Scene& s = c.sc;
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
