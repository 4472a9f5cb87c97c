// Shared device/host definitions of the MI355X (gfx950) Newton-iteration library.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cipc_gpu.h"

namespace cipc {

// ------------------------------------------------------------------------------------
// Errors: every ABI entry point catches cipc::Error and maps it to a status code plus a
// message retrievable with cipc_gpu_last_error().
// ------------------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CIPC_HIP(call)                                                                         \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::cipc::Error(CIPC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CIPC_CHECK_LAUNCH() CIPC_HIP(hipGetLastError())

// Grown-on-demand device buffer; never shrinks during a context's life.
// Bumped by every device or pinned allocation and free of the buffers below: a captured HIP
// graph holds buffer addresses, so it is replayed only while this count is unchanged.
inline std::atomic<unsigned long long>& alloc_generation() {
    static std::atomic<unsigned long long> g{0};
    return g;
}

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void reserve(size_t want) {
        if (want <= cap) return;
        ++alloc_generation();
        if (p) CIPC_HIP(hipFree(p));
        p = nullptr;
        cap = 0;
        size_t c = want + want / 4 + 16;
        CIPC_HIP(hipMalloc(&p, c * sizeof(T)));
        cap = c;
    }
    void resize(size_t want) { reserve(want); n = want; }
    void release() {
        if (p) {
            ++alloc_generation();
            (void)hipFree(p);
        }
        p = nullptr; cap = n = 0;
    }
    T* data() { return p; }
    const T* data() const { return p; }
};

// Grown-on-demand pinned host buffer: the destination of asynchronous device-to-host copies
// that the host reads after a later synchronisation of the stream (so that one
// synchronisation serves several results).  Grown only while no copy into it is pending.
template <typename T>
struct PinnedBuf {
    T* p = nullptr;
    size_t cap = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    void reserve(size_t want) {
        if (want <= cap) return;
        ++alloc_generation();
        if (p) CIPC_HIP(hipHostFree(p));
        p = nullptr;
        cap = 0;
        CIPC_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), want * sizeof(T), hipHostMallocDefault));
        cap = want;
    }
};

inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// ------------------------------------------------------------------------------------
// Device vector helpers.  Contact classification, distances, boxes and CCD use these
// with floating-point contraction disabled in their translation units, so that each
// expression rounds exactly as written (and as the CPU restatement rounds it).
// ------------------------------------------------------------------------------------
struct d3 {
    double x, y, z;
};
__host__ __device__ inline d3 mk(double a, double b, double c) { return d3{a, b, c}; }
__host__ __device__ inline d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline d3 operator-(d3 a) { return {-a.x, -a.y, -a.z}; }
__host__ __device__ inline d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline d3 operator/(d3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline double sqn(d3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__host__ __device__ inline d3 cross(d3 a, d3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline d3 vmin(d3 a, d3 b) { return {fmin(a.x, b.x), fmin(a.y, b.y), fmin(a.z, b.z)}; }
__host__ __device__ inline d3 vmax(d3 a, d3 b) { return {fmax(a.x, b.x), fmax(a.y, b.y), fmax(a.z, b.z)}; }
__device__ inline d3 ld3(const double* p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ inline void st3(double* p, int i, d3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }

// Local 3x3 blocks of a 4-vertex element: 0..3 diagonal (a,a); 4..9 off-diagonal (a,b),
// a > b: (1,0) (2,0) (3,0) (2,1) (3,1) (3,2).  A triangle uses those with a, b < 3.
__host__ __device__ constexpr int lb_a(int lb) { return lb < 4 ? lb : (lb < 7 ? (lb - 3) : (lb < 9 ? lb - 5 : 3)); }
__host__ __device__ constexpr int lb_b(int lb) { return lb < 4 ? lb : (lb < 7 ? 0 : (lb < 9 ? 1 : 2)); }

__host__ __device__ inline bool is_free(int t) { return t == 0; }
__host__ __device__ inline bool is_collider(int t) { return t >= 2; }

// Axis-aligned box (lower, upper) in 48 bytes.
struct Box {
    d3 lo, hi;
};
// Packed LBVH node, one 64-byte record: the box, the first child (internal node) or the
// primitive id (leaf), the node that follows this node's subtree in depth-first order (-1
// after the last) and the second child (internal node).  One record is read per traversal
// step.
struct alignas(64) Node {
    d3 lo;
    int left, skip;
    d3 hi;
    int right, pad;
};
// Four-wide record of an internal node: the boxes of its grandchildren (its children where
// those are leaves), rounded outward to float, and what they are: the binary node index of
// an internal node, -(primitive + 1) for a leaf, kNoChild for an empty slot.  One 128-byte
// line per step of a wide walk; boxes that only touch after rounding are kept, so a wide
// walk returns a superset that the caller confirms with the exact boxes.
constexpr int kNoChild = 0x7fffffff;
struct alignas(128) Node4 {
    float lo[4][3];
    float hi[4][3];
    int child[4];
    int pad[4];
};
// Touching counts as overlapping (CPU IPC/LBVH.cpp:157-161); `q` is the querying box.
__host__ __device__ inline bool overlaps(const Box& q, const Box& b, double gap) {
    return !((b.lo.x - q.hi.x > gap) || (b.lo.y - q.hi.y > gap) || (b.lo.z - q.hi.z > gap) ||
             (q.lo.x - b.hi.x > gap) || (q.lo.y - b.hi.y > gap) || (q.lo.z - b.hi.z > gap));
}

// ------------------------------------------------------------------------------------
// Device-resident scene and per-iteration state.
// ------------------------------------------------------------------------------------
struct Scene {
    int V = 0, T = 0, F = 0, Vs = 0, Es = 0;
    int kb = 1;                // block-record key: column << kb | row, 2^kb > V
    double dt = 0, u = 0, r = 0, drag = 0, hhat = 0, fhat = 0, fric = 0;
    d3 gn{0, 1, 0};
    double gd = -1;
    DBuf<double> x, x0, vrest, vprev, xt, mass, dir;
    DBuf<double> vel;          // mesh3D::velocities (3V), advanced by a time step
    DBuf<int> bt;
    DBuf<int4> tets;           // v0..v3
    DBuf<double> dminv;        // SoA: 9 x T (entry k of tet t at k*T + t)
    DBuf<double> vol;
    DBuf<int4> faces;          // v0, v1, v2, tet
    DBuf<int> sv;              // surface vertices
    DBuf<int2> se;             // surface edges
    DBuf<int> sv_perm, se_perm; // query orders: surface vertices / edges by Morton code of the rest shape
    // cloth: Baraff-Witkin triangles and bending elements
    int Tr = 0, Nb = 0, bmode = 0;  // bmode 0: quadratic bending form, 1: hinge angle
    double stretch = 0, shear = 0, srate = 0, plate = 0;
    DBuf<int4> tris;           // v0, v1, v2, 0
    DBuf<double> dmt;          // SoA: 4 x Tr, column-major 2x2 Dm^-1
    DBuf<double> area;         // Tr, area x thickness
    DBuf<int4> bverts;         // Nb
    DBuf<double> bq;           // SoA: 16 x Nb (quadratic form Q, row-major 4x4)
    DBuf<double> brest, bweight;  // Nb (hinge form)
    std::vector<int> h_bt;     // host copies for pattern construction
    std::vector<int> h_tets, h_tris, h_bverts;
    // Batch of independent scenes packed into this one (S > 1): vertices, tets, surface
    // vertices, edges and faces of scene k follow those of scene k - 1.  Contacts never join
    // two scenes; stiffness, energies, step bounds and solves are per scene.
    int S = 1;
    DBuf<int> vscene;          // V: scene of each vertex (batch only)
    DBuf<double> tmu, tlam;    // T: per-tet Lame parameters u, r (batch only)
    DBuf<double> kap;          // S: barrier stiffness of the current call (batch only)
    std::vector<int> h_voff, h_toff;   // S + 1 vertex / tet offsets
};

// Barrier stiffness of the scene a vertex belongs to: one value, or the per-scene value of
// a batch.
struct Kap {
    double k = 0.0;
    const double* ks = nullptr;
    const int* vs = nullptr;
    __device__ double at(int v) const { return ks ? ks[vs[v]] : k; }
};
// Lame parameters of a tet: one pair, or per tet in a batch.
struct Lame {
    double u = 0.0, r = 0.0;
    const double* tu = nullptr;
    const double* tr = nullptr;
    __device__ void at(int t, double& uu, double& rr) const {
        if (tu) { uu = tu[t]; rr = tr[t]; }
        else { uu = u; rr = r; }
    }
};
inline Kap scene_kappa(const Scene& s, double k) { return Kap{k, s.S > 1 ? s.kap.p : nullptr, s.S > 1 ? s.vscene.p : nullptr}; }
inline Lame scene_lame(const Scene& s) { return Lame{s.u, s.r, s.S > 1 ? s.tmu.p : nullptr, s.S > 1 ? s.tlam.p : nullptr}; }
inline const int* scene_of_vertex(const Scene& s) { return s.S > 1 ? s.vscene.p : nullptr; }

constexpr int kChunk = 64;   // tets per chunk of the chunked tet assembly: one wavefront

// Static elastic/mass block pattern and gather lists (built once per scene).
struct StaticPattern {
    int nblocks = 0;
    DBuf<int> col_ptr;      // V+1: blocks of block-column b in [col_ptr[b], col_ptr[b+1])
    DBuf<int> row;          // nblocks: block row a (a >= b), ascending within a column
    DBuf<int> gptr;         // nblocks+1: contribution slots of block k in [gptr[k], gptr[k+1])
    DBuf<int> vptr;         // V+1: element-corner slots of vertex v in [vptr[v], vptr[v+1])
    // Elements e: tets [0, T), triangles [T, T+Tr), bending elements [T+Tr, T+Tr+Nb).
    DBuf<int> slot_h;       // 10 per element: slot of (element, local block), -1 when not stored
    DBuf<int> slot_g;       // 4 per element: slot of (element, corner)
    DBuf<int> vtet;         // V: tet corners of vertex v (the first slots of its run)
    DBuf<int> diag_block;   // V: block index of (v, v) in the static pattern
    // host copies for the chunk plan: the block of every contribution slot, and slot_h
    std::vector<int> h_blk_of_slot, h_slot_h;

    // Chunked tet assembly (scenes whose elements are all tets): the tets of each scene in
    // Morton order of their centroids, cut into chunks of kChunk tets, one wavefront per
    // chunk and one lane per tet.  A chunk sums its tets' blocks in LDS accumulators, one per
    // distinct block it touches; a block touched by one chunk only is written straight into
    // the block values, the others leave one partial per chunk that the completion pass adds
    // in chunk order.
    // Vertex gradients go the same way: the tets' corner forces into per-chunk vertex
    // accumulators, a vertex interior to one chunk written with its inertia term by that
    // chunk, the others completed from their partials.
    bool chunked = false;
    int chunk = 0, nchunks = 0, lds_bytes = 0, ncomplete = 0, ncomplete_v = 0;
    DBuf<int> ch_tet;               // nchunks*chunk: tet of each lane, -1 for padding
    DBuf<unsigned short> ch_code;   // ((ch*14 + phase)*chunk + lane): accumulator << 6 | round; 0xFFFF: none
                                    // (phases 0..9 the local blocks, 10..13 the corners)
    DBuf<unsigned char> ch_rounds;  // ch*16 + phase: rounds of that phase (16 bytes per chunk)
    DBuf<int> ch_acc_ptr;           // nchunks+1: block accumulators of chunk ch in [ptr[ch], ptr[ch+1])
    DBuf<int> ch_acc_dst;           // per block accumulator: block index (>= 0), or -(partial + 1)
    DBuf<int> ch_vacc_ptr;          // nchunks+1: vertex accumulators of chunk ch
    DBuf<int> ch_vacc_dst;          // per vertex accumulator: vertex (>= 0), or -(vertex partial + 1)
    DBuf<int> cp_blk, cp_ptr;       // completion: block, and its partials in [cp_ptr[i], cp_ptr[i+1])
    DBuf<int> cpv_vtx, cpv_ptr;     // completion: vertex, and its partials
    DBuf<double> part;              // 9 per block partial
    DBuf<unsigned short> partf;     // nonzero masks of the block partials
    DBuf<double> vpart;             // 3 per vertex partial
    DBuf<int4> ctet;                // chunk order copies of the tets' static data (padding: the
    DBuf<double> cdm, cvol;         //   chunk's first tet); cdm SoA, entry k of lane p at k*NP + p
};

}  // namespace cipc
