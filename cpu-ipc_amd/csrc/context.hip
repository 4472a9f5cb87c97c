// C ABI of the device path (declared in include/cipc_gpu.h).  Every entry point catches
// cipc::Error / std::exception and returns a status code; the message is kept per thread.
#include "context.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>

namespace {
thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

template <typename F>
int guard(F&& f) {
    try {
        f();
        return CIPC_OK;
    } catch (const cipc::Error& e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc&) {
        return fail(CIPC_ERR_RUNTIME, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(CIPC_ERR_RUNTIME, e.what());
    }
}

void require(bool ok, const char* msg) {
    if (!ok) throw cipc::Error(CIPC_ERR_INVALID_ARGUMENT, msg);
}

template <typename T>
void upload(cipc::DBuf<T>& d, const T* h, size_t n, hipStream_t s) {
    d.resize(std::max<size_t>(n, 1));
    if (n) CIPC_HIP(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
}

void bind(cipc::Context& c) { CIPC_HIP(hipSetDevice(c.device)); }
}  // namespace

struct cipc_gpu_context : cipc::Context {};

extern "C" int cipc_gpu_eval_snh_impl(int n, const double* F, double u, double r, double* P, double* H, double* e);
extern "C" int cipc_gpu_eval_cloth_impl(int n, const double* F, double st, double sh, double k, double* P, double* H, double* e);
extern "C" int cipc_gpu_eval_hinge_impl(int n, const double* pts, const double* rest, const double* weight, double plate, double* g,
                                        double* H, double* e);
extern "C" int cipc_gpu_eval_distance_impl(int n, int kind, const double* pts, double* d, double* g, double* H);
extern "C" int cipc_gpu_eval_make_pd_impl(int n, int dim, double* H);
extern "C" int cipc_gpu_eval_contact_pd_impl(int n, int nverts, double* H, const double* G);
extern "C" int cipc_gpu_eval_accd_impl(int n, int kind, const double* pts, const double* dirs, double eta, double th, double* toc);
extern "C" int cipc_gpu_eval_classify_impl(int n, int kind, const double* pts, int* type);

using namespace cipc;

extern "C" {

const char* cipc_gpu_last_error(void) { return g_last_error.c_str(); }
// the error channel of the host-side entry points (include/cipc_host.h)
int cipc_internal_fail(int code, const char* msg) { return fail(code, msg ? msg : "unknown error"); }

int cipc_gpu_context_create(int device, cipc_gpu_context** out) {
    return guard([&] {
        require(out != nullptr, "output pointer is null");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw Error(CIPC_ERR_NO_DEVICE, "no HIP device is available");
        require(device >= 0 && device < n, "device ordinal out of range");
        CIPC_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        CIPC_HIP(hipGetDeviceProperties(&prop, device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            throw Error(CIPC_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
        auto* c = new cipc_gpu_context();
        c->device = device;
        CIPC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        CIPC_HIP(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
        CIPC_HIP(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
        CIPC_HIP(hipEventCreateWithFlags(&c->join, hipEventDisableTiming));
        CIPC_HIP(hipEventCreateWithFlags(&c->fr_fork, hipEventDisableTiming));
        CIPC_HIP(hipEventCreateWithFlags(&c->fr_join, hipEventDisableTiming));
        CIPC_HIP(hipEventCreateWithFlags(&c->cl_done, hipEventDisableTiming));
        CIPC_HIP(hipStreamCreateWithFlags(&c->sbs, hipStreamNonBlocking));
        CIPC_HIP(hipEventCreateWithFlags(&c->as_done, hipEventDisableTiming));
        CIPC_HIP(hipEventCreateWithFlags(&c->sb_done, hipEventDisableTiming));
        CIPC_HIP(hipEventCreateWithFlags(&c->it_start, hipEventDisableTiming));
        CIPC_HIP(hipEventCreateWithFlags(&c->en_done, hipEventDisableTiming));
        *out = c;
    });
}

int cipc_gpu_context_destroy(cipc_gpu_context* c) {
    return guard([&] {
        if (!c) return;
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->aux);
        (void)hipStreamSynchronize(c->sbs);
        (void)hipEventDestroy(c->fork);
        (void)hipEventDestroy(c->join);
        (void)hipEventDestroy(c->fr_fork);
        (void)hipEventDestroy(c->fr_join);
        (void)hipEventDestroy(c->cl_done);
        (void)hipEventDestroy(c->as_done);
        (void)hipEventDestroy(c->sb_done);
        (void)hipEventDestroy(c->it_start);
        (void)hipEventDestroy(c->en_done);
        if (c->graph.exec) (void)hipGraphExecDestroy(c->graph.exec);
        (void)hipStreamDestroy(c->sbs);
        (void)hipStreamDestroy(c->aux);
        (void)hipStreamDestroy(c->stream);
        destroy_learned(*c);
        destroy_direct(*c);
        delete c;  // DBuf members free on destruction of their owner only when released:
    });
}

}  // extern "C"

namespace {
void upload_scene_impl(cipc_gpu_context* c, const cipc_scene_desc* d, bool plan = true) {
    {
        require(c && d, "null context or descriptor");
        require(d->V > 0 && d->T >= 0 && d->F >= 0 && d->Vs >= 0 && d->Es >= 0, "invalid scene counts");
        require(d->x && d->v_rest && d->V_prev && d->xtilde && d->masses && d->boundary, "missing vertex arrays");
        require(d->T == 0 || (d->tets && d->dm_inv && d->volumes), "missing tet arrays");
        require(d->F == 0 || d->faces, "missing surface faces");
        require(d->Vs == 0 || d->surf_verts, "missing surface vertices");
        require(d->Es == 0 || d->surf_edges, "missing surface edges");
        require(d->hhat > 0 && d->dt > 0, "dHat and dt must be positive");
        require(d->Tr >= 0 && d->Nb >= 0 && (d->bending == 0 || d->bending == 1), "invalid cloth counts or bending mode");
        require(d->Tr == 0 || (d->triangles && d->dm_tri_inv && d->areas), "missing triangle arrays");
        require(d->Nb == 0 || (d->bend_verts && (d->bending == 0 ? d->bend_q != nullptr : (d->bend_rest && d->bend_weight))),
                "missing bending arrays");
        bind(*c);
        Scene& s = c->sc;
        const int V = d->V, T = d->T;
        for (int i = 0; i < 4 * T; ++i)
            if (d->tets[i] < 0 || d->tets[i] >= V) throw Error(CIPC_ERR_INVALID_ARGUMENT, "tet vertex index out of range");
        for (int i = 0; i < d->F; ++i)
            for (int k = 0; k < 3; ++k)
                if (d->faces[4 * i + k] < 0 || d->faces[4 * i + k] >= V)
                    throw Error(CIPC_ERR_INVALID_ARGUMENT, "surface face index out of range");
        for (int i = 0; i < 3 * d->Tr; ++i)
            if (d->triangles[i] < 0 || d->triangles[i] >= V) throw Error(CIPC_ERR_INVALID_ARGUMENT, "triangle vertex index out of range");
        for (int i = 0; i < 4 * d->Nb; ++i)
            if (d->bend_verts[i] < 0 || d->bend_verts[i] >= V) throw Error(CIPC_ERR_INVALID_ARGUMENT, "bending vertex index out of range");
        for (int i = 0; i < d->Vs; ++i)
            if (d->surf_verts[i] >= (uint64_t)V) throw Error(CIPC_ERR_INVALID_ARGUMENT, "surface vertex index out of range");
        for (int i = 0; i < 2 * d->Es; ++i)
            if (d->surf_edges[i] >= (uint64_t)V) throw Error(CIPC_ERR_INVALID_ARGUMENT, "surface edge index out of range");
        s.V = V; s.T = T; s.F = d->F; s.Vs = d->Vs; s.Es = d->Es;
        s.kb = 1;
        while ((1ll << s.kb) <= (long long)V) ++s.kb;
        s.dt = d->dt; s.u = d->length_rate; s.r = d->volume_rate; s.drag = d->drag; s.hhat = d->hhat; s.fhat = d->fhat;
        s.fric = d->friction;
        s.gn = d3{d->ground_normal[0], d->ground_normal[1], d->ground_normal[2]};
        s.gd = d->ground_d;
        const hipStream_t st = c->stream;
        upload(s.x, d->x, 3 * (size_t)V, st);
        upload(s.vrest, d->v_rest, 3 * (size_t)V, st);
        upload(s.vprev, d->V_prev, 3 * (size_t)V, st);
        upload(s.xt, d->xtilde, 3 * (size_t)V, st);
        s.vel.resize(3 * (size_t)V);
        CIPC_HIP(hipMemsetAsync(s.vel.p, 0, 3 * sizeof(double) * (size_t)V, st));
        upload(s.mass, d->masses, (size_t)V, st);
        upload(s.bt, d->boundary, (size_t)V, st);
        s.x0.resize(3 * (size_t)V);
        s.dir.resize(3 * (size_t)V);
        CIPC_HIP(hipMemsetAsync(s.dir.p, 0, 3 * sizeof(double) * V, st));
        s.h_tets.assign(d->tets, d->tets + 4 * (size_t)T);
        s.h_bt.assign(d->boundary, d->boundary + V);
        upload(s.tets, reinterpret_cast<const int4*>(d->tets), (size_t)T, st);
        // Dm^-1 to SoA: entry k of tet t at k*T + t
        std::vector<double> soa(9 * (size_t)std::max(T, 1));
        for (int t = 0; t < T; ++t)
            for (int k = 0; k < 9; ++k) soa[(size_t)k * T + t] = d->dm_inv[9 * (size_t)t + k];
        upload(s.dminv, soa.data(), 9 * (size_t)T, st);
        upload(s.vol, d->volumes, (size_t)T, st);
        upload(s.faces, reinterpret_cast<const int4*>(d->faces), (size_t)d->F, st);
        // cloth: triangles padded to int4, Dm^-1 and the quadratic forms to SoA
        {
            const int Tr = d->Tr, Nb = d->Nb;
            s.Tr = Tr; s.Nb = Nb; s.bmode = d->bending;
            s.stretch = d->stretch; s.shear = d->shear; s.srate = d->strain_rate; s.plate = d->plate_rigidity;
            s.h_tris.assign(d->triangles, d->triangles + 3 * (size_t)Tr);
            s.h_bverts.assign(d->bend_verts, d->bend_verts + 4 * (size_t)Nb);
            std::vector<int4> tr4(std::max(Tr, 1));
            for (int t = 0; t < Tr; ++t) tr4[t] = make_int4(d->triangles[3 * t], d->triangles[3 * t + 1], d->triangles[3 * t + 2], 0);
            upload(s.tris, tr4.data(), (size_t)Tr, st);
            std::vector<double> dm(4 * (size_t)std::max(Tr, 1));
            for (int t = 0; t < Tr; ++t)
                for (int k = 0; k < 4; ++k) dm[(size_t)k * Tr + t] = d->dm_tri_inv[4 * (size_t)t + k];
            upload(s.dmt, dm.data(), 4 * (size_t)Tr, st);
            upload(s.area, d->areas, (size_t)Tr, st);
            upload(s.bverts, reinterpret_cast<const int4*>(d->bend_verts), (size_t)Nb, st);
            std::vector<double> q(16 * (size_t)std::max(Nb, 1), 0.0);
            if (d->bending == 0)
                for (int h = 0; h < Nb; ++h)
                    for (int k = 0; k < 16; ++k) q[(size_t)k * Nb + h] = d->bend_q[16 * (size_t)h + k];
            upload(s.bq, q.data(), d->bending == 0 ? 16 * (size_t)Nb : 0, st);
            upload(s.brest, d->bending == 1 ? d->bend_rest : nullptr, d->bending == 1 ? (size_t)Nb : 0, st);
            upload(s.bweight, d->bending == 1 ? d->bend_weight : nullptr, d->bending == 1 ? (size_t)Nb : 0, st);
            CIPC_HIP(hipStreamSynchronize(st));   // the host staging vectors go out of scope
        }
        std::vector<int> sv(d->Vs), se(2 * (size_t)d->Es);
        for (int i = 0; i < d->Vs; ++i) sv[i] = (int)d->surf_verts[i];
        for (int i = 0; i < 2 * d->Es; ++i) se[i] = (int)d->surf_edges[i];
        upload(s.sv, sv.data(), sv.size(), st);
        upload(s.se, reinterpret_cast<const int2*>(se.data()), (size_t)d->Es, st);
        // spatially coherent query orders for the tree traversals
        {
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            for (int v = 0; v < V; ++v)
                for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], d->x[3 * v + k]); hi[k] = std::max(hi[k], d->x[3 * v + k]); }
            auto code = [&](double px, double py, double pz) {
                auto q = [&](double v, int k) {
                    const double e = hi[k] - lo[k];
                    const double t = e > 0 ? (v - lo[k]) / e : 0.0;
                    return (uint32_t)std::max(0.0, std::min(t * 1024.0, 1023.0));
                };
                auto ex = [](uint32_t v) {
                    v = (v * 0x00010001u) & 0xFF0000FFu; v = (v * 0x00000101u) & 0x0F00F00Fu;
                    v = (v * 0x00000011u) & 0xC30C30C3u; v = (v * 0x00000005u) & 0x49249249u;
                    return v;
                };
                return (ex(q(px, 0)) << 2) | (ex(q(py, 1)) << 1) | ex(q(pz, 2));
            };
            std::vector<std::pair<uint32_t, int>> ord(d->Vs);
            for (int i = 0; i < d->Vs; ++i) ord[i] = {code(d->x[3 * sv[i]], d->x[3 * sv[i] + 1], d->x[3 * sv[i] + 2]), i};
            std::sort(ord.begin(), ord.end());
            std::vector<int> perm(d->Vs);
            for (int i = 0; i < d->Vs; ++i) perm[i] = ord[i].second;
            upload(s.sv_perm, perm.data(), perm.size(), st);
            std::vector<std::pair<uint32_t, int>> oe(d->Es);
            for (int i = 0; i < d->Es; ++i) {
                const double* a = d->x + 3 * se[2 * i];
                const double* b = d->x + 3 * se[2 * i + 1];
                oe[i] = {code(0.5 * (a[0] + b[0]), 0.5 * (a[1] + b[1]), 0.5 * (a[2] + b[2])), i};
            }
            std::sort(oe.begin(), oe.end());
            std::vector<int> pe(d->Es);
            for (int i = 0; i < d->Es; ++i) pe[i] = oe[i].second;
            upload(s.se_perm, pe.data(), pe.size(), st);
        }
        CIPC_HIP(hipStreamSynchronize(st));
        build_static_pattern(*c);
        c->n_self = c->n_ee = c->n_ccd = c->n_env = 0;
        c->self_set.resize(1); c->ee_set.resize(1); c->ee_ie.resize(1); c->ccd_set.resize(1); c->env_set.resize(1);
        c->have_scene = true;
        c->face_tree.age = c->edge_tree.age = c->sface_tree.age = c->sedge_tree.age = kMaxRefits;
        c->n_self = c->n_ee = c->n_ccd = c->n_env = 0;   // sets and lagged state of an earlier scene
        upload_counts(*c);
        c->as.capPT = c->as.capEE = c->as.capM = 0;   // capacities of an earlier scene
        c->capS = 0;
        // Lanes per query of the four walks until the autotuner chooses: about 2^17 lanes
        // in flight, floor(log2(2^17 / queries)) bits, 1 to 16 lanes (what the autotuner
        // picks for the full CCD on the four-cube scene and for every walk of a 256-scene
        // batch).
        auto bits_for = [](int queries) {
            int b = 0;
            while (b < 4 && ((long long)queries << (b + 1)) <= (1ll << 17)) ++b;
            return b;
        };
        c->tune.ccd_pt_bits = c->tune.as_pt_bits = bits_for(d->Vs);
        c->tune.ccd_ee_bits = c->tune.as_ee_bits = bits_for(d->Es);
        if (c->graph.exec) (void)hipGraphExecDestroy(c->graph.exec);   // launches of an earlier scene
        c->graph.exec = nullptr;
        c->graph.key.clear();
        c->n_fself = c->n_fenv = 0;
        c->n_soft = 0; c->soft_w = 0.0;
        c->n_cenv = c->n_cself = 0;
        c->assembled = false;
        s.S = 1;
        s.h_voff = {0, V};
        s.h_toff = {0, T};
        if (plan) build_tet_chunks(*c, d->x);
    }
}

// One descriptor holding the scenes of a batch one after another: vertex, tet and face
// indices shifted by the scene's offsets.  The storage lives in `keep`.
struct BatchPack {
    std::vector<double> x, vr, vp, xt, m, dm, vol;
    std::vector<int> bt, tets, faces;
    std::vector<uint64_t> sv, se;
    std::vector<double> tmu, tlam;
    std::vector<int> vscene, voff, toff;
};
void pack_batch(const cipc_scene_desc* const* ds, int S, BatchPack& b, cipc_scene_desc& u) {
    const cipc_scene_desc& f = *ds[0];
    u = cipc_scene_desc{};
    b.voff.assign(1, 0);
    b.toff.assign(1, 0);
    for (int k = 0; k < S; ++k) {
        const cipc_scene_desc* d = ds[k];
        require(d != nullptr, "null scene descriptor in the batch");
        require(d->V > 0 && d->T > 0 && d->F >= 0 && d->Vs >= 0 && d->Es >= 0, "invalid scene counts in the batch");
        require(d->x && d->v_rest && d->V_prev && d->xtilde && d->masses && d->boundary && d->tets && d->dm_inv && d->volumes,
                "missing arrays in a batch scene");
        require(d->F == 0 || d->faces, "missing surface faces");
        require(d->Vs == 0 || d->surf_verts, "missing surface vertices");
        require(d->Es == 0 || d->surf_edges, "missing surface edges");
        require(d->Tr == 0 && d->Nb == 0, "batched scenes hold tets only");
        require(d->dt == f.dt && d->hhat == f.hhat && d->fhat == f.fhat && d->drag == f.drag && d->friction == f.friction &&
                    d->ground_d == f.ground_d && d->ground_normal[0] == f.ground_normal[0] &&
                    d->ground_normal[1] == f.ground_normal[1] && d->ground_normal[2] == f.ground_normal[2],
                "the scenes of a batch share dt, dHat, the friction parameters and the ground");
        const int V = d->V, T = d->T, vo = b.voff.back(), to = b.toff.back();
        if ((long long)vo + V > (1ll << 30) || (long long)to + T > (1ll << 30)) throw Error(CIPC_ERR_OVERFLOW, "batch too large");
        for (int i = 0; i < 4 * T; ++i)
            if (d->tets[i] < 0 || d->tets[i] >= V) throw Error(CIPC_ERR_INVALID_ARGUMENT, "tet vertex index out of range");
        for (int i = 0; i < d->F; ++i) {
            for (int q = 0; q < 3; ++q)
                if (d->faces[4 * i + q] < 0 || d->faces[4 * i + q] >= V)
                    throw Error(CIPC_ERR_INVALID_ARGUMENT, "surface face index out of range");
        }
        for (int i = 0; i < d->Vs; ++i)
            if (d->surf_verts[i] >= (uint64_t)V) throw Error(CIPC_ERR_INVALID_ARGUMENT, "surface vertex index out of range");
        for (int i = 0; i < 2 * d->Es; ++i)
            if (d->surf_edges[i] >= (uint64_t)V) throw Error(CIPC_ERR_INVALID_ARGUMENT, "surface edge index out of range");
        b.x.insert(b.x.end(), d->x, d->x + 3 * (size_t)V);
        b.vr.insert(b.vr.end(), d->v_rest, d->v_rest + 3 * (size_t)V);
        b.vp.insert(b.vp.end(), d->V_prev, d->V_prev + 3 * (size_t)V);
        b.xt.insert(b.xt.end(), d->xtilde, d->xtilde + 3 * (size_t)V);
        b.m.insert(b.m.end(), d->masses, d->masses + V);
        b.bt.insert(b.bt.end(), d->boundary, d->boundary + V);
        for (int i = 0; i < 4 * T; ++i) b.tets.push_back(d->tets[i] + vo);
        b.dm.insert(b.dm.end(), d->dm_inv, d->dm_inv + 9 * (size_t)T);
        b.vol.insert(b.vol.end(), d->volumes, d->volumes + T);
        for (int i = 0; i < d->F; ++i) {
            for (int q = 0; q < 3; ++q) b.faces.push_back(d->faces[4 * i + q] + vo);
            b.faces.push_back(d->faces[4 * i + 3] + to);
        }
        for (int i = 0; i < d->Vs; ++i) b.sv.push_back(d->surf_verts[i] + vo);
        for (int i = 0; i < 2 * d->Es; ++i) b.se.push_back(d->surf_edges[i] + vo);
        b.tmu.insert(b.tmu.end(), (size_t)T, d->length_rate);
        b.tlam.insert(b.tlam.end(), (size_t)T, d->volume_rate);
        b.vscene.insert(b.vscene.end(), (size_t)V, k);
        b.voff.push_back(vo + V);
        b.toff.push_back(to + T);
    }
    u.V = b.voff.back(); u.T = b.toff.back(); u.F = (int)(b.faces.size() / 4); u.Vs = (int)b.sv.size(); u.Es = (int)(b.se.size() / 2);
    u.x = b.x.data(); u.v_rest = b.vr.data(); u.V_prev = b.vp.data(); u.xtilde = b.xt.data(); u.masses = b.m.data();
    u.boundary = b.bt.data(); u.tets = b.tets.data(); u.dm_inv = b.dm.data(); u.volumes = b.vol.data();
    u.faces = b.faces.data(); u.surf_verts = b.sv.data(); u.surf_edges = b.se.data();
    u.dt = f.dt; u.length_rate = f.length_rate; u.volume_rate = f.volume_rate; u.drag = f.drag; u.hhat = f.hhat; u.fhat = f.fhat;
    u.friction = f.friction;
    for (int q = 0; q < 3; ++q) u.ground_normal[q] = f.ground_normal[q];
    u.ground_d = f.ground_d;
}

void require_batch(cipc_gpu_context* c) { require(c && c->have_scene, "no scene uploaded"); }
}  // namespace

extern "C" {

int cipc_gpu_upload_scene(cipc_gpu_context* c, const cipc_scene_desc* d) {
    return guard([&] { upload_scene_impl(c, d); });
}

int cipc_gpu_upload_batch(cipc_gpu_context* c, const cipc_scene_desc* const* scenes, int n_scenes) {
    return guard([&] {
        require(c && scenes && n_scenes > 0, "null context or descriptors, or no scenes");
        BatchPack b;
        cipc_scene_desc u;
        pack_batch(scenes, n_scenes, b, u);
        upload_scene_impl(c, &u, false);
        Scene& s = c->sc;
        const hipStream_t st = c->stream;
        s.S = n_scenes;
        s.h_voff = b.voff;
        s.h_toff = b.toff;
        build_tet_chunks(*c, b.x.data());
        upload(s.vscene, b.vscene.data(), b.vscene.size(), st);
        upload(s.tmu, b.tmu.data(), b.tmu.size(), st);
        upload(s.tlam, b.tlam.data(), b.tlam.size(), st);
        s.kap.resize(n_scenes);
        std::vector<int> off(b.toff);
        off.insert(off.end(), b.voff.begin(), b.voff.end());
        upload(c->batch_off, off.data(), off.size(), st);
        CIPC_HIP(hipStreamSynchronize(st));
    });
}

int cipc_gpu_batch_size(cipc_gpu_context* c, int* n_scenes, int* vertex_offsets, int* tet_offsets) {
    return guard([&] {
        require_batch(c);
        const Scene& s = c->sc;
        if (n_scenes) *n_scenes = s.S;
        if (vertex_offsets) std::copy(s.h_voff.begin(), s.h_voff.end(), vertex_offsets);
        if (tet_offsets) std::copy(s.h_toff.begin(), s.h_toff.end(), tet_offsets);
    });
}

int cipc_gpu_set_state(cipc_gpu_context* c, const double* x, const double* vprev, const double* xt) {
    return guard([&] {
        require(c && c->have_scene, "no scene uploaded");
        bind(*c);
        const size_t n = 3 * (size_t)c->sc.V * sizeof(double);
        if (x) {
            CIPC_HIP(hipMemcpyAsync(c->sc.x.p, x, n, hipMemcpyHostToDevice, c->stream));
            // new positions: fresh topologies on the next builds
            c->face_tree.age = c->edge_tree.age = c->sface_tree.age = c->sedge_tree.age = kMaxRefits;
        }
        if (vprev) CIPC_HIP(hipMemcpyAsync(c->sc.vprev.p, vprev, n, hipMemcpyHostToDevice, c->stream));
        if (xt) CIPC_HIP(hipMemcpyAsync(c->sc.xt.p, xt, n, hipMemcpyHostToDevice, c->stream));
        CIPC_HIP(hipStreamSynchronize(c->stream));
    });
}

int cipc_gpu_get_positions(cipc_gpu_context* c, double* x) {
    return guard([&] {
        require(c && c->have_scene && x, "no scene uploaded or null output");
        bind(*c);
        CIPC_HIP(hipMemcpyAsync(x, c->sc.x.p, 3 * sizeof(double) * c->sc.V, hipMemcpyDeviceToHost, c->stream));
        CIPC_HIP(hipStreamSynchronize(c->stream));
    });
}

int cipc_gpu_set_direction(cipc_gpu_context* c, const double* p) {
    return guard([&] {
        require(c && c->have_scene && p, "no scene uploaded or null direction");
        bind(*c);
        CIPC_HIP(hipMemcpyAsync(c->sc.dir.p, p, 3 * sizeof(double) * c->sc.V, hipMemcpyHostToDevice, c->stream));
        CIPC_HIP(hipStreamSynchronize(c->stream));
    });
}

int cipc_gpu_active_set(cipc_gpu_context* c, int64_t counts[4]) {
    return guard([&] {
        require(c && c->have_scene, "no scene uploaded");
        bind(*c);
        active_set_now(*c);   // the call returns with the sets built
        if (counts) { counts[0] = c->n_self; counts[1] = c->n_ee; counts[2] = c->n_ccd; counts[3] = c->n_env; }
    });
}

int cipc_gpu_get_active_set(cipc_gpu_context* c, int* self4, int* ee4, int* eeie2, int* ccd2, int* env) {
    return guard([&] {
        require(c && c->have_scene, "no scene uploaded");
        bind(*c);
        const hipStream_t st = c->stream;
        if (self4 && c->n_self) CIPC_HIP(hipMemcpyAsync(self4, c->self_set.p, 16 * c->n_self, hipMemcpyDeviceToHost, st));
        if (ee4 && c->n_ee) CIPC_HIP(hipMemcpyAsync(ee4, c->ee_set.p, 16 * c->n_ee, hipMemcpyDeviceToHost, st));
        if (eeie2 && c->n_ee) CIPC_HIP(hipMemcpyAsync(eeie2, c->ee_ie.p, 8 * c->n_ee, hipMemcpyDeviceToHost, st));
        if (ccd2 && c->n_ccd) CIPC_HIP(hipMemcpyAsync(ccd2, c->ccd_set.p, 8 * c->n_ccd, hipMemcpyDeviceToHost, st));
        if (env && c->n_env) CIPC_HIP(hipMemcpyAsync(env, c->env_set.p, 4 * c->n_env, hipMemcpyDeviceToHost, st));
        CIPC_HIP(hipStreamSynchronize(st));
    });
}

int cipc_gpu_set_active_set(cipc_gpu_context* c, const int64_t n[4], const int* self4, const int* ee4, const int* eeie2,
                            const int* ccd2, const int* env) {
    return guard([&] {
        require(c && c->have_scene && n, "no scene uploaded or null counts");
        require(n[0] >= 0 && n[1] >= 0 && n[2] >= 0 && n[3] >= 0, "negative active-set count");
        require((n[0] == 0 || self4) && (n[1] == 0 || (ee4 && eeie2)) && (n[2] == 0 || ccd2) && (n[3] == 0 || env),
                "missing active-set array");
        const int V = c->sc.V, Es = c->sc.Es, F = c->sc.F, Vs = c->sc.Vs;
        for (int64_t i = 0; i < n[0]; ++i) {
            const int* e = self4 + 4 * i;
            const int v0 = e[0] >= 0 ? e[0] : -e[0] - 1;
            bool ok = v0 < V && e[1] >= 0 && e[1] < V && e[2] < V && e[3] < V;
            if (e[0] >= 0) ok = ok && e[2] >= 0 && e[3] >= 0;
            if (!ok) throw Error(CIPC_ERR_INVALID_ARGUMENT, "Self_ActiveSet entry out of range");
        }
        for (int64_t i = 0; i < n[1]; ++i) {
            const int* e = ee4 + 4 * i;
            const int v0 = e[0] >= 0 ? e[0] : -e[0] - 1;
            bool ok = v0 < V && e[1] >= 0 && e[1] < V && e[2] < V && e[3] < V;
            if (e[3] < 0) ok = ok && eeie2[2 * i] >= 0 && eeie2[2 * i] < Es && eeie2[2 * i + 1] >= 0 && eeie2[2 * i + 1] < Es;
            if (!ok) throw Error(CIPC_ERR_INVALID_ARGUMENT, "Self_EE_ActiveSet entry out of range");
        }
        for (int64_t i = 0; i < n[2]; ++i) {
            const int a = ccd2[2 * i], b = ccd2[2 * i + 1];
            const bool ok = a < 0 ? (-a - 1 < Vs && b >= 0 && b < F) : (a < Es && b >= 0 && b < Es);
            if (!ok) throw Error(CIPC_ERR_INVALID_ARGUMENT, "Self_CCD_ActiveSet entry out of range");
        }
        for (int64_t i = 0; i < n[3]; ++i)
            if (env[i] < 0 || env[i] >= V) throw Error(CIPC_ERR_INVALID_ARGUMENT, "Environment_ActiveSet entry out of range");
        bind(*c);
        const hipStream_t st = c->stream;
        c->n_self = n[0]; c->n_ee = n[1]; c->n_ccd = n[2]; c->n_env = n[3];
        upload(c->self_set, reinterpret_cast<const int4*>(self4), (size_t)n[0], st);
        upload(c->ee_set, reinterpret_cast<const int4*>(ee4), (size_t)n[1], st);
        upload(c->ee_ie, reinterpret_cast<const int2*>(eeie2), (size_t)n[1], st);
        upload(c->ccd_set, reinterpret_cast<const int2*>(ccd2), (size_t)n[2], st);
        upload(c->env_set, env, (size_t)n[3], st);
        CIPC_HIP(hipStreamSynchronize(st));
        upload_counts(*c);
    });
}

// The elastic part depends only on the state: it runs on the auxiliary stream from the
// current point of the main stream, and the main stream waits for it before the merge.
static void elastic_fork(cipc_gpu_context* c, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr) {
    CIPC_HIP(hipEventRecord(c->fork, c->stream));
    CIPC_HIP(hipStreamWaitEvent(c->aux, c->fork, 0));
    if (t0) CIPC_HIP(hipEventRecord(t0, c->aux));
    elastic_assemble(*c, c->aux);
    if (t1) CIPC_HIP(hipEventRecord(t1, c->aux));
    CIPC_HIP(hipEventRecord(c->join, c->aux));
}
static void elastic_join(cipc_gpu_context* c) { CIPC_HIP(hipStreamWaitEvent(c->stream, c->join, 0)); }

// Barrier stiffness per scene of a batch (ks, S values; NULL: k for every scene).
static void set_scene_kappa(cipc_gpu_context* c, const double* ks, double k) {
    Scene& s = c->sc;
    if (s.S == 1) return;
    std::vector<double> h(s.S, k);
    if (ks)
        for (int i = 0; i < s.S; ++i) h[i] = ks[i];
    for (double v : h) require(std::isfinite(v) && v >= 0.0, "barrier stiffness must be finite and non-negative");
    CIPC_HIP(hipMemcpyAsync(s.kap.p, h.data(), s.S * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CIPC_HIP(hipStreamSynchronize(c->stream));
}
static void require_single(cipc_gpu_context* c, const char* what) {
    if (c->sc.S != 1) throw Error(CIPC_ERR_LOGIC, std::string(what) + ": a batch context takes the per-scene entry points");
}

static void assemble_impl(cipc_gpu_context* c, double kappa, const double* ks = nullptr) {
    set_scene_kappa(c, ks, kappa);
    elastic_fork(c);
    contact_assemble(*c, kappa);
    elastic_join(c);
    finalize_csr(*c);
    c->assembled = true;
    CIPC_HIP(hipStreamSynchronize(c->stream));
    settle_nnz(*c);   // the nonzero and collision counts
}

int cipc_gpu_set_hessian_precision(cipc_gpu_context* c, int precision) {
    return guard([&] {
        require(c != nullptr, "null context");
        require(precision == 0 || precision == 1, "precision must be 0 (FP64) or 1 (FP32 elastic Hessian blocks)");
        c->hessian_fp32 = precision == 1;
        c->assembled = false;
    });
}

int cipc_gpu_assemble(cipc_gpu_context* c, double kappa, int64_t* nnz, int* collisions) {
    return guard([&] {
        require(c && c->have_scene, "no scene uploaded");
        bind(*c);
        assemble_impl(c, kappa);
        if (nnz) *nnz = c->nnz;
        if (collisions) *collisions = c->collisions;
    });
}

int cipc_gpu_get_gradient(cipc_gpu_context* c, double* g) {
    return guard([&] {
        require(c && c->have_scene && g, "no scene uploaded or null output");
        bind(*c);
        CIPC_HIP(hipMemcpyAsync(g, c->grad.p, 3 * sizeof(double) * c->sc.V, hipMemcpyDeviceToHost, c->stream));
        CIPC_HIP(hipStreamSynchronize(c->stream));
    });
}

int cipc_gpu_get_csr(cipc_gpu_context* c, int* outer, int* inner, double* values, double* rhs) {
    return guard([&] {
        require(c && c->have_scene, "no scene uploaded");
        bind(*c);
        const hipStream_t st = c->stream;
        const size_t n = 3 * (size_t)c->sc.V;
        CIPC_HIP(hipStreamSynchronize(st));
        settle_nnz(*c);
        if (outer) CIPC_HIP(hipMemcpyAsync(outer, c->outer.p, (n + 1) * sizeof(int), hipMemcpyDeviceToHost, st));
        if (inner && c->nnz) CIPC_HIP(hipMemcpyAsync(inner, c->inner.p, c->nnz * sizeof(int), hipMemcpyDeviceToHost, st));
        if (values && c->nnz) CIPC_HIP(hipMemcpyAsync(values, c->val.p, c->nnz * sizeof(double), hipMemcpyDeviceToHost, st));
        if (rhs) CIPC_HIP(hipMemcpyAsync(rhs, c->rhs.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
        CIPC_HIP(hipStreamSynchronize(st));
    });
}

int cipc_gpu_set_soft_targets(cipc_gpu_context* c, int n, const int* vertices, const double* targets, double weight) {
    return guard([&] {
        require(c && c->have_scene, "no scene uploaded");
        require(n >= 0 && (n == 0 || (vertices && targets)), "invalid soft boundary arguments");
        bind(*c);
        if (n > 0) {
            // validateSoftState / initialize (BoundaryConditions.cpp:36-58,64-110)
            require(std::isfinite(weight) && weight > 0.0, "soft boundary weight must be finite and positive");
            std::vector<char> seen(c->sc.V, 0);
            for (int i = 0; i < n; ++i) {
                const int v = vertices[i];
                require(v >= 0 && v < c->sc.V, "soft boundary vertex index out of range");
                require(is_free(c->sc.h_bt[v]), "soft boundary vertices must be free");
                require(!seen[v], "soft boundary vertices must be distinct");
                seen[v] = 1;
                for (int k = 0; k < 3; ++k) require(std::isfinite(targets[3 * i + k]), "soft boundary target must be finite");
            }
        }
        upload(c->soft_idx, vertices, (size_t)n, c->stream);
        upload(c->soft_t, targets, 3 * (size_t)n, c->stream);
        CIPC_HIP(hipStreamSynchronize(c->stream));
        c->n_soft = n;
        c->soft_w = n > 0 ? weight : 0.0;
    });
}

int64_t cipc_gpu_preconditioner_param_count(void) { return learned_param_count(); }

int cipc_gpu_set_preconditioner_weights(cipc_gpu_context* c, const float* weights, int64_t n) {
    return guard([&] {
        require(c && weights, "null context or weights");
        bind(*c);
        learned_set_weights(*c, weights, n);
    });
}

int cipc_gpu_preconditioner_gradient(cipc_gpu_context* c, int objective, const float* params, const float* probes, int n_probes,
                                     double* loss, float* grad) {
    return guard([&] {
        require(c && c->have_scene && params && loss, "invalid arguments");
        require(objective == 0 || objective == 1, "objective must be 0 (probe loss) or 1 (block FSAI regression)");
        require(objective == 1 || (probes && n_probes > 0), "the probe loss needs probes");
        require(c->assembled, "no assembled system: call cipc_gpu_assemble first");
        bind(*c);
        *loss = learned_loss_grad(*c, params, probes, n_probes, 0, grad, objective);
    });
}

int cipc_gpu_preconditioner_train_step(cipc_gpu_context* c, int objective, float* params, float* adam_m, float* adam_v, int step,
                                       double lr, int n_probes, uint64_t seed, double* loss) {
    return guard([&] {
        require(c && c->have_scene && params && adam_m && adam_v, "invalid arguments");
        require(objective == 0 || objective == 1, "objective must be 0 (probe loss) or 1 (block FSAI regression)");
        require(step >= 1 && (objective == 1 || n_probes > 0) && lr > 0, "step >= 1, probes > 0 and lr > 0");
        require(c->assembled, "no assembled system: call cipc_gpu_assemble first");
        bind(*c);
        const double l = learned_train_step(*c, params, adam_m, adam_v, step, lr, n_probes, seed, objective);
        if (loss) *loss = l;
    });
}

int cipc_gpu_kappa_gradients(cipc_gpu_context* c, double out[2], int64_t* n_constraints) {
    return guard([&] {
        require(c && c->have_scene && out, "no scene uploaded or null output");
        require_single(c, "cipc_gpu_kappa_gradients");
        bind(*c);
        const int64_t n = kappa_gradients(*c, out);
        if (n_constraints) *n_constraints = n;
    });
}

int cipc_gpu_close_constraints(cipc_gpu_context* c, double dtol, int* closer, int64_t counts[2]) {
    return guard([&] {
        require(c && c->have_scene && closer, "no scene uploaded or null output");
        require_single(c, "cipc_gpu_close_constraints");
        bind(*c);
        int64_t cnt[2];
        *closer = close_constraints(*c, dtol, cnt);
        if (counts) { counts[0] = cnt[0]; counts[1] = cnt[1]; }
    });
}

int cipc_gpu_clear_close_constraints(cipc_gpu_context* c) {
    return guard([&] {
        require(c && c->have_scene, "no scene uploaded");
        c->n_cenv = c->n_cself = 0;
    });
}

int cipc_gpu_energy_terms_noalloc(cipc_gpu_context* c, double kappa, double* terms_device) {
    return guard([&] {
        (void)kappa;
        require(c && c->have_scene && terms_device, "no scene uploaded or null output");
        bind(*c);
        hipPointerAttribute_t at;
        const hipError_t pe = hipPointerGetAttributes(&at, terms_device);
        if (pe != hipSuccess) (void)hipGetLastError();   // an unknown (host) pointer: clear the error it leaves
        if (pe != hipSuccess || at.type != hipMemoryTypeDevice)
            throw Error(CIPC_ERR_INVALID_ARGUMENT, "terms_device must be device memory (9 doubles)");
        energy_terms_noalloc(*c, terms_device);
    });
}

int cipc_gpu_energy_terms(cipc_gpu_context* c, double kappa, double terms[9]) {
    return guard([&] {
        require(c && c->have_scene && terms, "no scene uploaded or null output");
        bind(*c);
        energy_terms(*c, kappa, terms);
    });
}

int cipc_gpu_energy(cipc_gpu_context* c, double kappa, double* e) {
    return guard([&] {
        require(c && c->have_scene && e, "no scene uploaded or null output");
        bind(*c);
        double t[9];
        energy_terms(*c, kappa, t);
        *e = total_energy(*c, kappa, t);
    });
}

int cipc_gpu_solve(cipc_gpu_context* c, int precision, double rel_tol, int max_iter, double* direction, cipc_solve_stats* stats) {
    return cipc_gpu_solve_ex(c, precision, 0, rel_tol, max_iter, direction, stats);
}

int cipc_gpu_solve_ex(cipc_gpu_context* c, int precision, int preconditioner, double rel_tol, int max_iter, double* direction,
                      cipc_solve_stats* stats) {
    return guard([&] {
        require(c && c->have_scene, "no scene uploaded");
        require(precision == 0 || precision == 1, "precision must be 0 (FP64) or 1 (FP32 with FP64 refinement)");
        require(preconditioner >= 0 && preconditioner <= 3,
                "preconditioner must be 0 (block-Jacobi), 1 (block FSAI), 2 (learned) or 3 (direct)");
        require(std::isfinite(rel_tol) && rel_tol > 0.0 && max_iter > 0, "tolerance and iteration cap must be positive");
        require(c->assembled, "no assembled system: call cipc_gpu_assemble first");
        bind(*c);
        hipEvent_t e0, e1;
        CIPC_HIP(hipEventCreate(&e0));
        CIPC_HIP(hipEventCreate(&e1));
        CIPC_HIP(hipEventRecord(e0, c->stream));
        SolveResult r;
        linear_solve(*c, precision, preconditioner, rel_tol, max_iter, r);
        CIPC_HIP(hipEventRecord(e1, c->stream));
        CIPC_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        CIPC_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (direction) {
            CIPC_HIP(hipMemcpyAsync(direction, c->sc.dir.p, 3 * sizeof(double) * c->sc.V, hipMemcpyDeviceToHost, c->stream));
            CIPC_HIP(hipStreamSynchronize(c->stream));
        }
        if (stats) {
            stats->iterations = r.iterations;
            stats->refinements = r.refinements;
            stats->rel_residual = r.rel_residual;
            stats->solve_ms = ms;
        }
        if (!(r.rel_residual <= rel_tol))
            throw Error(CIPC_ERR_RUNTIME, "conjugate gradients stopped after " + std::to_string(r.iterations) +
                                              " iterations at relative residual " + std::to_string(r.rel_residual));
    });
}

int cipc_gpu_get_solve_factor(cipc_gpu_context* c, int64_t* nblocks, int* rowptr, int* col, double* values) {
    return guard([&] {
        require(c && c->have_scene && nblocks, "no scene uploaded or null output");
        const SolveBuffers& w = c->solve;
        if (w.factor == 0) throw Error(CIPC_ERR_LOGIC, "no factor: call cipc_gpu_solve_ex with preconditioner 1 or 2 first");
        bind(*c);
        const int V = c->sc.V;
        int nb = 0;
        read_back(*c, &nb, w.rowptr.p + V, 1);
        *nblocks = nb;
        if (rowptr) CIPC_HIP(hipMemcpyAsync(rowptr, w.rowptr.p, sizeof(int) * ((size_t)V + 1), hipMemcpyDeviceToHost, c->stream));
        if (col && nb) CIPC_HIP(hipMemcpyAsync(col, w.col.p, sizeof(int) * (size_t)nb, hipMemcpyDeviceToHost, c->stream));
        if (values && nb) CIPC_HIP(hipMemcpyAsync(values, w.G.p, sizeof(double) * 9 * (size_t)nb, hipMemcpyDeviceToHost, c->stream));
        CIPC_HIP(hipStreamSynchronize(c->stream));
    });
}

// ---- batched scenes -------------------------------------------------------------------
int cipc_gpu_batch_assemble(cipc_gpu_context* c, const double* kappa, int64_t* nnz, int* collisions) {
    return guard([&] {
        require_batch(c);
        require(kappa != nullptr, "null stiffness array");
        bind(*c);
        assemble_impl(c, kappa[0], kappa);
        if (nnz) *nnz = c->nnz;
        if (collisions) {
            // per scene, counted as compute_g_dpt returns it (scene of the entry's point / first edge vertex)
            const Scene& s = c->sc;
            std::fill(collisions, collisions + s.S, 0);
            if (c->n_self) {
                std::vector<int4> h(c->n_self);
                CIPC_HIP(hipMemcpyAsync(h.data(), c->self_set.p, 16 * c->n_self, hipMemcpyDeviceToHost, c->stream));
                CIPC_HIP(hipStreamSynchronize(c->stream));
                for (const int4& e : h) {
                    const int v = e.x >= 0 ? e.x : -e.x - 1;
                    const int k = (int)(std::upper_bound(s.h_voff.begin(), s.h_voff.end(), v) - s.h_voff.begin()) - 1;
                    collisions[k] += (e.x < 0 && (e.z < 0 || e.w < 0)) ? -e.w : 1;
                }
            }
        }
    });
}

int cipc_gpu_batch_friction_update(cipc_gpu_context* c, const double* kappa) {
    return guard([&] {
        require_batch(c);
        require(kappa != nullptr, "null stiffness array");
        bind(*c);
        set_scene_kappa(c, kappa, kappa[0]);
        friction_update(*c, kappa[0]);
        CIPC_HIP(hipStreamSynchronize(c->stream));
    });
}

int cipc_gpu_batch_energy(cipc_gpu_context* c, const double* kappa, double* energy, double* terms) {
    return guard([&] {
        require_batch(c);
        require(kappa && (energy || terms), "null stiffness array or outputs");
        bind(*c);
        const int S = c->sc.S;
        std::vector<double> t(9 * (size_t)S);
        energy_terms_scenes(*c, t.data());
        for (int k = 0; k < S; ++k) {
            if (energy) energy[k] = total_energy(*c, kappa[k], t.data() + 9 * (size_t)k);
            if (terms) std::copy(t.begin() + 9 * (size_t)k, t.begin() + 9 * (size_t)k + 9, terms + 9 * (size_t)k);
        }
    });
}

int cipc_gpu_batch_step_bound(cipc_gpu_context* c, double* out) {
    return guard([&] {
        require_batch(c);
        require(out != nullptr, "null output");
        bind(*c);
        step_bound_scenes(*c, out);
    });
}

int cipc_gpu_batch_step_forward(cipc_gpu_context* c, const double* alpha) {
    return guard([&] {
        require_batch(c);
        require(alpha != nullptr, "null step array");
        for (int k = 0; k < c->sc.S; ++k) require(std::isfinite(alpha[k]), "steps must be finite");
        bind(*c);
        step_forward_scenes(*c, alpha);
        CIPC_HIP(hipStreamSynchronize(c->stream));
    });
}

int cipc_gpu_batch_is_intersected(cipc_gpu_context* c, int* r) {
    return guard([&] {
        require_batch(c);
        require(r != nullptr, "null output");
        bind(*c);
        scene_intersections(*c, r);
    });
}

int cipc_gpu_batch_solve(cipc_gpu_context* c, int precision, int preconditioner, double rel_tol, int max_iter, const int* active,
                         double* direction, cipc_solve_stats* stats) {
    return guard([&] {
        require_batch(c);
        require(precision == 0 || precision == 1, "precision must be 0 (FP64) or 1 (FP32 with FP64 refinement)");
        require(preconditioner >= 0 && preconditioner <= 2, "preconditioner must be 0 (block-Jacobi), 1 (block FSAI) or 2 (learned)");
        require(std::isfinite(rel_tol) && rel_tol > 0.0 && max_iter > 0, "tolerance and iteration cap must be positive");
        require(c->assembled, "no assembled system: call cipc_gpu_batch_assemble first");
        bind(*c);
        const int S = c->sc.S;
        hipEvent_t e0, e1;
        CIPC_HIP(hipEventCreate(&e0));
        CIPC_HIP(hipEventCreate(&e1));
        CIPC_HIP(hipEventRecord(e0, c->stream));
        std::vector<SolveResult> r(S);
        linear_solve_scenes(*c, precision, preconditioner, rel_tol, max_iter, active, r.data());
        CIPC_HIP(hipEventRecord(e1, c->stream));
        CIPC_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        CIPC_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (direction) {
            CIPC_HIP(hipMemcpyAsync(direction, c->sc.dir.p, 3 * sizeof(double) * c->sc.V, hipMemcpyDeviceToHost, c->stream));
            CIPC_HIP(hipStreamSynchronize(c->stream));
        }
        int missed = -1;
        for (int k = 0; k < S; ++k) {
            if (stats) {
                stats[k].iterations = r[k].iterations;
                stats[k].refinements = r[k].refinements;
                stats[k].rel_residual = r[k].rel_residual;
                stats[k].solve_ms = ms;
            }
            if ((!active || active[k]) && !(r[k].rel_residual <= rel_tol) && missed < 0) missed = k;
        }
        if (missed >= 0)
            throw Error(CIPC_ERR_RUNTIME, "conjugate gradients of scene " + std::to_string(missed) + " stopped after " +
                                              std::to_string(r[missed].iterations) + " iterations at relative residual " +
                                              std::to_string(r[missed].rel_residual));
    });
}

}  // extern "C"

namespace {
// Timing events of an iteration enqueued eagerly (a captured one is timed as a whole).
struct IterEvents {
    hipEvent_t ev[4] = {}, es = nullptr, ee = nullptr, e0 = nullptr, en = nullptr, s0 = nullptr, s1 = nullptr;
    IterEvents() {
        for (auto& e : ev) CIPC_HIP(hipEventCreate(&e));
        for (hipEvent_t* e : {&es, &ee, &e0, &en, &s0, &s1}) CIPC_HIP(hipEventCreate(e));
    }
    ~IterEvents() {
        for (auto& e : ev) (void)hipEventDestroy(e);
        for (hipEvent_t e : {es, ee, e0, en, s0, s1}) (void)hipEventDestroy(e);
    }
    IterEvents(const IterEvents&) = delete;
    IterEvents& operator=(const IterEvents&) = delete;
};

// One Newton iteration's launches on the three streams, without a host synchronisation.
// Main: active set, barrier blocks, CSR.  Second (`aux`): the elastic stage beside the active
// set, the friction records and small projections beside the barrier blocks, then (with
// `energy`) the energy.  Step-bound stream (`sbs`): the ground bound, swept trees and full CCD
// (positions and direction only) from the start, the partial CCD once the active set exists.
// The streams fork from and join into the main stream through the context's events, so the
// same launches can be captured into a graph.  t (eager only): per-stage timing events.
void enqueue_iteration(cipc_gpu_context* c, double kappa, bool energy, const IterEvents* t) {
    // The host enqueues the critical path first (the active set on the main stream), the
    // elastic stage and the lagged friction records on `aux` (they
    // read the state and the lagged friction only), the barrier records on the main stream;
    // then what has slack: the ground bound, swept trees, full and partial CCD on `sbs`, the
    // energy on `aux`; then the CSR on the main stream.  Host enqueue time is close to half
    // the iteration's GPU time, so its order decides how long the main stream waits.
    if (t) CIPC_HIP(hipEventRecord(t->ev[0], c->stream));
    CIPC_HIP(hipEventRecord(c->it_start, c->stream));
    contact_records_prepare(*c);
    compute_active_set(*c, false);   // contact_assemble publishes the counts
    if (t) CIPC_HIP(hipEventRecord(t->ev[1], c->stream));
    CIPC_HIP(hipEventRecord(c->as_done, c->stream));
    CIPC_HIP(hipStreamWaitEvent(c->aux, c->it_start, 0));
    if (t) CIPC_HIP(hipEventRecord(t->es, c->aux));
    elastic_assemble(*c, c->aux);
    if (t) CIPC_HIP(hipEventRecord(t->ee, c->aux));
    friction_records_launch(*c, c->aux);
    CIPC_HIP(hipEventRecord(c->join, c->aux));
    contact_assemble(*c, kappa, true);
    // the step bound after the barrier records: it has slack, the main stream has none
    CIPC_HIP(hipStreamWaitEvent(c->sbs, c->it_start, 0));
    if (t) CIPC_HIP(hipEventRecord(t->s0, c->sbs));
    step_bound_early(*c, c->sbs);
    CIPC_HIP(hipStreamWaitEvent(c->sbs, c->as_done, 0));
    step_bound_partial(*c, c->sbs);
    if (t) CIPC_HIP(hipEventRecord(t->s1, c->sbs));
    CIPC_HIP(hipEventRecord(c->sb_done, c->sbs));
    if (energy) {
        if (t) CIPC_HIP(hipEventRecord(t->e0, c->aux));
        energy_terms_launch(*c, c->aux);
        if (t) CIPC_HIP(hipEventRecord(t->en, c->aux));
    }
    collision_count_launch(*c, c->aux);   // it reads the active set only
    CIPC_HIP(hipEventRecord(c->en_done, c->aux));
    elastic_join(c);
    finalize_csr(*c, false);
    if (t) CIPC_HIP(hipEventRecord(t->ev[2], c->stream));
    CIPC_HIP(hipStreamWaitEvent(c->stream, c->sb_done, 0));
    CIPC_HIP(hipStreamWaitEvent(c->stream, c->en_done, 0));
    publish_nnz(*c, c->stream);
    if (t) CIPC_HIP(hipEventRecord(t->ev[3], c->stream));
}

// What the captured launches depend on besides device memory.
struct IterKey {
    double kappa, soft_w;
    long long nfs, nfe, gen, topo;
    int fp32, capPT, capEE, capM, capS, nsoft, tune_pt, tune_ee, tune_apt, tune_aee, S, energy;
};
std::vector<unsigned char> iter_key(const cipc_gpu_context* c, double kappa, bool energy) {
    IterKey k;
    std::memset(&k, 0, sizeof k);
    k.kappa = kappa; k.soft_w = c->soft_w;
    k.nfs = c->n_fself; k.nfe = c->n_fenv; k.gen = (long long)alloc_generation().load(); k.topo = (long long)c->topo_gen;
    k.fp32 = c->hessian_fp32 ? 1 : 0; k.capPT = c->as.capPT; k.capEE = c->as.capEE; k.capM = c->as.capM; k.capS = c->capS;
    k.nsoft = c->n_soft; k.tune_pt = c->tune.ccd_pt_bits; k.tune_ee = c->tune.ccd_ee_bits;
    k.tune_apt = c->tune.as_pt_bits; k.tune_aee = c->tune.as_ee_bits; k.S = c->sc.S; k.energy = energy ? 1 : 0;
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&k);
    return std::vector<unsigned char>(b, b + sizeof k);
}

// The four trees are refitted, not rebuilt, by the next iteration (a captured iteration holds
// the refit launches of the topologies it was captured with).
bool trees_refit_only(const cipc_gpu_context* c) {
    const Scene& s = c->sc;
    auto ok = [](const LBVH& t, int n) { return n <= 0 || (t.n == n && t.age < kMaxRefits); };
    return ok(c->face_tree, s.F) && ok(c->edge_tree, s.Es) && ok(c->sface_tree, s.F) && ok(c->sedge_tree, s.Es);
}

void drop_graph(cipc_gpu_context* c) {
    if (c->graph.exec) (void)hipGraphExecDestroy(c->graph.exec);
    c->graph.exec = nullptr;
    c->graph.key.clear();
}

// Runs the iteration as one graph launch on the main stream when it can: a replay when the
// key matches, else a new capture (discarded, and false returned, when an allocation happened
// while capturing, or the capture failed).  t0/t1 bracket the launch.
bool iteration_graph(cipc_gpu_context* c, double kappa, bool energy, hipEvent_t t0, hipEvent_t t1) {
    // by default a batch goes as one graph launch and a single scene eagerly (measured no
    // faster as a graph: launching the graph's ~120 nodes costs the host what the launches do)
    const bool on = c->graph.mode < 0 ? c->sc.S > 1 : c->graph.mode == 1;
    if (!on || !trees_refit_only(c) || c->capS == 0 || c->as.capPT == 0) return false;
    std::vector<unsigned char> key = iter_key(c, kappa, energy);
    if (!c->graph.exec || key != c->graph.key) {
        drop_graph(c);
        const unsigned long long gen0 = alloc_generation().load();
        hipGraph_t g = nullptr;
        CIPC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        bool ok = true;
        try {
            enqueue_iteration(c, kappa, energy, nullptr);
        } catch (const std::exception&) {
            ok = false;
        }
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (e != hipSuccess || !g) ok = false;
        (void)hipGetLastError();
        if (!ok) {
            // a launch that cannot be captured: no graph for this context from now on; every
            // stream leaves capture and drains before the launches go eagerly
            c->graph.mode = 0;
            for (hipStream_t st : {c->stream, c->aux, c->sbs}) {
                hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
                if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
                    hipGraph_t dropped = nullptr;
                    (void)hipStreamEndCapture(st, &dropped);
                    if (dropped) (void)hipGraphDestroy(dropped);
                }
                (void)hipStreamSynchronize(st);
            }
            (void)hipGetLastError();
        }
        if (ok && alloc_generation().load() == gen0) {
            hipGraphExec_t exec = nullptr;
            if (hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) == hipSuccess && exec) {
                c->graph.exec = exec;
                c->graph.key = iter_key(c, kappa, energy);
                ++c->graph.captures;
            }
            (void)hipGetLastError();
        }
        if (g) (void)hipGraphDestroy(g);
        if (!c->graph.exec) {
            c->bucket_cnt_clear = 0;   // the captured launches did not run
            return false;              // this call runs eagerly
        }
    } else {
        // the host side of a replayed iteration: what enqueue_iteration records on the host
        for (LBVH* t : {&c->face_tree, &c->edge_tree, &c->sface_tree, &c->sedge_tree})
            if (t->n > 0) ++t->age;
        c->dc_pending = true;
        c->nnz_pending = true;
        ++c->graph.replays;
    }
    CIPC_HIP(hipEventRecord(t0, c->stream));
    CIPC_HIP(hipGraphLaunch(c->graph.exec, c->stream));
    CIPC_HIP(hipEventRecord(t1, c->stream));
    return true;
}
}  // namespace

extern "C" {

// One Newton iteration of every scene: active set, assembly with each scene's stiffness, step
// bounds and energies (per scene), timed as cipc_gpu_newton_iteration; total->alpha and
// total->energy hold the minimum step and the summed energy.
int cipc_gpu_batch_newton_iteration(cipc_gpu_context* c, const double* kappa, cipc_iteration_timing* t, double* alpha, double* energy) {
    return guard([&] {
        require_batch(c);
        require(kappa && t, "null stiffness array or output");
        bind(*c);
        const int S = c->sc.S;
        set_scene_kappa(c, kappa, kappa[0]);
        // the launches of cipc_gpu_newton_iteration for the whole batch (one graph when they
        // can go as one), then the per-scene energies, whose element arrays are sized by the
        // settled counts
        IterEvents E;
        hipEvent_t e_en0, e_en1;
        CIPC_HIP(hipEventCreate(&e_en0));
        CIPC_HIP(hipEventCreate(&e_en1));
        bool graphed = false;
        double enq_ms = 0;
        for (int round = 0;; ++round) {
            const auto h0 = std::chrono::steady_clock::now();
            graphed = round == 0 && iteration_graph(c, kappa[0], false, E.ev[0], E.ev[3]);
            if (!graphed) enqueue_iteration(c, kappa[0], false, &E);
            if (round == 0) enq_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
            CIPC_HIP(hipEventSynchronize(E.ev[3]));
            if (!grow_capacities(*c)) break;
            if (round >= 3) throw Error(CIPC_ERR_RUNTIME, "active-set capacities kept overflowing");
        }
        settle_active_set(*c);
        settle_nnz(*c);
        c->assembled = true;
        std::vector<double> sb(5 * (size_t)S), et(9 * (size_t)S);
        CIPC_HIP(hipEventRecord(e_en0, c->stream));
        energy_terms_scenes(*c, et.data());          // synchronises the main stream
        CIPC_HIP(hipEventRecord(e_en1, c->stream));
        CIPC_HIP(hipEventSynchronize(e_en1));
        step_bound_finish(*c, sb.data());
        *t = cipc_iteration_timing{};
        float tot = 0, en = 0;
        CIPC_HIP(hipEventElapsedTime(&tot, E.ev[0], E.ev[3]));
        CIPC_HIP(hipEventElapsedTime(&en, e_en0, e_en1));
        if (!graphed) {
            float ms[3], el = 0;
            for (int i = 0; i < 2; ++i) CIPC_HIP(hipEventElapsedTime(&ms[i], E.ev[i], E.ev[i + 1]));
            CIPC_HIP(hipEventElapsedTime(&ms[2], E.s0, E.s1));
            CIPC_HIP(hipEventElapsedTime(&el, E.es, E.ee));
            t->broadphase_ms = ms[0]; t->assembly_ms = ms[1]; t->elastic_ms = el; t->ccd_ms = ms[2];
        }
        (void)hipEventDestroy(e_en0);
        (void)hipEventDestroy(e_en1);
        t->energy_ms = en;
        t->total_ms = tot + en;
        t->graph = graphed ? 1 : 0;
        t->enqueue_ms = enq_ms;
        double amin = 1.0, esum = 0.0, bmin[4] = {1.0, 1.0, 1.0, 1.0};
        for (int k = 0; k < S; ++k) {
            const double a = sb[5 * (size_t)k], e = total_energy(*c, kappa[k], et.data() + 9 * (size_t)k);
            if (alpha) alpha[k] = a;
            if (energy) energy[k] = e;
            amin = std::min(amin, a);
            esum += e;
            for (int j = 0; j < 4; ++j) bmin[j] = std::min(bmin[j], sb[5 * (size_t)k + 1 + j]);
        }
        t->alpha = amin;
        t->energy = esum;
        t->nnz = c->nnz;
        t->contacts = c->n_self + c->n_ee + c->n_env;
        t->collisions = c->collisions;
        t->bound_ground = bmin[0]; t->bound_partial = bmin[1]; t->bound_cfl = bmin[2]; t->bound_full = bmin[3];
    });
}

int cipc_gpu_friction_update(cipc_gpu_context* c, double kappa) {
    return guard([&] {
        require(c && c->have_scene, "no scene uploaded");
        bind(*c);
        set_scene_kappa(c, nullptr, kappa);
        friction_update(*c, kappa);
        CIPC_HIP(hipStreamSynchronize(c->stream));
    });
}

int cipc_gpu_get_friction(cipc_gpu_context* c, int64_t counts[2], double* lam, double* coord, double* basis, double* env_lam) {
    return guard([&] {
        require(c && c->have_scene && counts, "no scene uploaded or null output");
        bind(*c);
        counts[0] = c->n_fself;
        counts[1] = c->n_fenv;
        const hipStream_t st = c->stream;
        const size_t n = (size_t)c->n_fself;
        if (lam && n) CIPC_HIP(hipMemcpyAsync(lam, c->fr_lam.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
        if (coord && n) CIPC_HIP(hipMemcpyAsync(coord, c->fr_coord.p, 2 * n * sizeof(double), hipMemcpyDeviceToHost, st));
        if (basis && n) CIPC_HIP(hipMemcpyAsync(basis, c->fr_basis.p, 6 * n * sizeof(double), hipMemcpyDeviceToHost, st));
        if (env_lam && c->n_fenv)
            CIPC_HIP(hipMemcpyAsync(env_lam, c->fr_env_lam.p, (size_t)c->n_fenv * sizeof(double), hipMemcpyDeviceToHost, st));
        CIPC_HIP(hipStreamSynchronize(st));
    });
}

int cipc_gpu_step_bound(cipc_gpu_context* c, double out[5]) {
    return guard([&] {
        require(c && c->have_scene && out, "no scene uploaded or null output");
        bind(*c);
        step_bound(*c, out);
    });
}

int cipc_gpu_autotune(cipc_gpu_context* c, cipc_tune_entry* entries, int max_entries, int* n_entries) {
    return guard([&] {
        require(c && c->have_scene, "no scene uploaded");
        require(max_entries >= 0 && (max_entries == 0 || entries), "null entries");
        bind(*c);
        TuneEntry e[4];
        const int n = autotune(*c, e, 4);
        for (int k = 0; k < n && k < max_entries; ++k) {
            entries[k] = cipc_tune_entry{};
            entries[k].kernel = e[k].kernel;
            entries[k].chosen = e[k].chosen;
            entries[k].n_candidates = e[k].n;
            for (int j = 0; j < 8; ++j) entries[k].ms[j] = e[k].ms[j];
        }
        if (n_entries) *n_entries = n;
    });
}

int cipc_gpu_set_tune(cipc_gpu_context* c, int kernel, int choice) {
    return guard([&] {
        require(c != nullptr, "null context");
        require(kernel >= 0 && kernel <= 4,
                "kernel is 0 / 1 (point-triangle / edge-edge full CCD), 2 (iteration graph), 3 / 4 (point-triangle / edge-edge "
                "active-set walk)");
        if (kernel == 2) {
            require(choice == 0 || choice == 1, "the iteration graph is 0 (off) or 1 (on)");
            c->graph.mode = choice;
            if (choice == 0) drop_graph(c);
            return;
        }
        require(choice >= 0 && choice < kTuneSplits, "choice is 0..4 (1 to 16 lanes per query)");
        int* slot[5] = {&c->tune.ccd_pt_bits, &c->tune.ccd_ee_bits, nullptr, &c->tune.as_pt_bits, &c->tune.as_ee_bits};
        *slot[kernel] = choice;
    });
}

int cipc_gpu_save_positions(cipc_gpu_context* c) {
    return guard([&] {
        require(c && c->have_scene, "no scene uploaded");
        bind(*c);
        CIPC_HIP(hipMemcpyAsync(c->sc.x0.p, c->sc.x.p, 3 * sizeof(double) * c->sc.V, hipMemcpyDeviceToDevice, c->stream));
        CIPC_HIP(hipStreamSynchronize(c->stream));
    });
}

int cipc_gpu_step_forward(cipc_gpu_context* c, double alpha) {
    return guard([&] {
        require(c && c->have_scene, "no scene uploaded");
        bind(*c);
        step_forward(*c, alpha);
        CIPC_HIP(hipStreamSynchronize(c->stream));
    });
}

int cipc_gpu_is_intersected(cipc_gpu_context* c, int* r) {
    return guard([&] {
        require(c && c->have_scene && r, "no scene uploaded or null output");
        bind(*c);
        *r = any_intersection(*c) ? 1 : 0;
    });
}

int cipc_gpu_newton_iteration(cipc_gpu_context* c, double kappa, cipc_iteration_timing* t) {
    return guard([&] {
        require(c && c->have_scene && t, "no scene uploaded or null output");
        require_single(c, "cipc_gpu_newton_iteration");
        bind(*c);
        // No stage waits on the host: the active set's counts stay on the device, and the
        // one synchronisation at the end also reports whether a capacity was exceeded (then
        // the capacity is grown and the iteration enqueued again, eagerly).  The launches
        // go as one graph when they can (stage times are then not taken: zero).
        IterEvents E;
        bool graphed = false;
        double enq_ms = 0;
        for (int round = 0;; ++round) {
            const auto h0 = std::chrono::steady_clock::now();
            graphed = round == 0 && iteration_graph(c, kappa, true, E.ev[0], E.ev[3]);
            if (!graphed) enqueue_iteration(c, kappa, true, &E);
            if (round == 0) enq_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
            CIPC_HIP(hipEventSynchronize(E.ev[3]));
            if (!grow_capacities(*c)) break;
            if (round >= 3) throw Error(CIPC_ERR_RUNTIME, "active-set capacities kept overflowing");
        }
        settle_active_set(*c);
        settle_nnz(*c);
        c->assembled = true;
        double sb[5], et[9];
        step_bound_finish(*c, sb);
        energy_terms_finish(*c, et);
        float tot = 0;
        CIPC_HIP(hipEventElapsedTime(&tot, E.ev[0], E.ev[3]));
        *t = cipc_iteration_timing{};
        if (!graphed) {
            float ms[4], el = 0;
            for (int i = 0; i < 2; ++i) CIPC_HIP(hipEventElapsedTime(&ms[i], E.ev[i], E.ev[i + 1]));
            CIPC_HIP(hipEventElapsedTime(&ms[2], E.s0, E.s1));  // step bound, beside the assembly
            CIPC_HIP(hipEventElapsedTime(&ms[3], E.e0, E.en));  // energy, beside the assembly's tail
            CIPC_HIP(hipEventElapsedTime(&el, E.es, E.ee));
            t->broadphase_ms = ms[0]; t->assembly_ms = ms[1]; t->elastic_ms = el; t->ccd_ms = ms[2]; t->energy_ms = ms[3];
        }
        t->total_ms = tot;
        t->graph = graphed ? 1 : 0;
        t->enqueue_ms = enq_ms;
        t->alpha = sb[0];
        t->energy = total_energy(*c, kappa, et);
        t->nnz = c->nnz;
        t->contacts = c->n_self + c->n_ee + c->n_env;
        t->collisions = c->collisions;
        t->bound_ground = sb[1]; t->bound_partial = sb[2]; t->bound_cfl = sb[3]; t->bound_full = sb[4];
    });
}

int cipc_gpu_set_velocities(cipc_gpu_context* c, const double* v) {
    return guard([&] {
        require(c && c->have_scene && v, "no scene uploaded or null velocities");
        bind(*c);
        CIPC_HIP(hipMemcpyAsync(c->sc.vel.p, v, 3 * sizeof(double) * c->sc.V, hipMemcpyHostToDevice, c->stream));
        CIPC_HIP(hipStreamSynchronize(c->stream));
    });
}

int cipc_gpu_get_state(cipc_gpu_context* c, double* x, double* vprev, double* xt, double* v) {
    return guard([&] {
        require(c && c->have_scene, "no scene uploaded");
        bind(*c);
        const size_t n = 3 * sizeof(double) * c->sc.V;
        if (x) CIPC_HIP(hipMemcpyAsync(x, c->sc.x.p, n, hipMemcpyDeviceToHost, c->stream));
        if (vprev) CIPC_HIP(hipMemcpyAsync(vprev, c->sc.vprev.p, n, hipMemcpyDeviceToHost, c->stream));
        if (xt) CIPC_HIP(hipMemcpyAsync(xt, c->sc.xt.p, n, hipMemcpyDeviceToHost, c->stream));
        if (v) CIPC_HIP(hipMemcpyAsync(v, c->sc.vel.p, n, hipMemcpyDeviceToHost, c->stream));
        CIPC_HIP(hipStreamSynchronize(c->stream));
    });
}

int cipc_gpu_time_step(cipc_gpu_context* c, const cipc_step_params* p, double* kappa, const cipc_dirichlet* d, cipc_step_stats* out) {
    return guard([&] {
        require(c && c->have_scene && p && kappa, "no scene uploaded or null parameters");
        require_single(c, "cipc_gpu_time_step");
        require(p->bbox_diag2 > 0 && p->mean_mass > 0 && p->newton_threshold > 0 && std::isfinite(*kappa) && *kappa >= 0,
                "invalid step parameters");
        require(p->precision == 0 || p->precision == 1, "precision must be 0 (FP64) or 1 (FP32 with FP64 refinement)");
        require(p->preconditioner >= 0 && p->preconditioner <= 3, "preconditioner must be 0, 1, 2 or 3 (direct)");
        require(p->solve_rel_tol > 0 && p->solve_max_iter > 0 && p->max_newton > 0, "solve tolerance and limits must be positive");
        DirichletHook hook;
        if (d && d->n > 0) {
            require(d->vertices && d->fn, "animated Dirichlet update without vertices or callback");
            for (int i = 0; i < d->n; ++i)
                require(d->vertices[i] >= 0 && d->vertices[i] < c->sc.V, "Dirichlet vertex index out of range");
            hook.n = d->n; hook.vertices = d->vertices; hook.step = d->step; hook.fn = d->fn; hook.user = d->user;
        }
        bind(*c);
        StepParams P;
        P.bbox_diag2 = p->bbox_diag2; P.mean_mass = p->mean_mass; P.dtol = p->dtol; P.newton_threshold = p->newton_threshold;
        P.apply_gravity = p->apply_gravity; P.quasi_static = p->quasi_static;
        P.precision = p->precision; P.preconditioner = p->preconditioner;
        P.solve_tol = p->solve_rel_tol; P.solve_max_iter = p->solve_max_iter; P.max_newton = p->max_newton;
        StepStats st;
        time_step(*c, P, *kappa, st, hook.n ? &hook : nullptr);
        if (out) {
            out->newton_iterations = st.newton_iterations; out->kappa_iterations = st.kappa_iterations;
            out->energy_backtracks = st.energy_backtracks; out->intersection_backtracks = st.intersection_backtracks;
            out->max_energy_backtracks_per_newton = st.max_energy_backtracks_per_newton;
            out->newton_steps_over_two_backtracks = st.newton_over_two_backtracks;
            out->pcg_iterations = st.pcg_iterations; out->collisions = st.collisions; out->matrix_nnz = st.matrix_nnz;
            out->direct_solves = st.direct_solves;
            out->ground_contacts = st.ground_contacts; out->self_contacts = st.self_contacts;
            out->mollified_contacts = st.mollified_contacts;
            out->kappa = st.kappa; out->min_distance2 = st.min_distance2; out->dirichlet_alpha = st.dirichlet_alpha;
            out->alpha_sum = st.alpha_sum; out->alpha_min = st.alpha_min; out->alpha_max = st.alpha_max;
            out->step_ms = st.step_ms; out->assembly_ms = st.assembly_ms; out->linear_ms = st.linear_ms; out->ccd_ms = st.ccd_ms;
            out->line_search_ms = st.line_search_ms; out->post_ms = st.post_ms;
        }
    });
}

int cipc_gpu_eval_cloth(int n, const double* F, double st, double sh, double k, double* P, double* H, double* e) {
    return guard([&] {
        require(n >= 0 && (n == 0 || (F && P && H && e)), "invalid arguments");
        if (n) cipc_gpu_eval_cloth_impl(n, F, st, sh, k, P, H, e);
    });
}
int cipc_gpu_eval_hinge(int n, const double* pts, const double* rest, const double* weight, double plate, double* g, double* H,
                        double* e) {
    return guard([&] {
        require(n >= 0 && (n == 0 || (pts && rest && weight && g && H && e)), "invalid arguments");
        if (n) cipc_gpu_eval_hinge_impl(n, pts, rest, weight, plate, g, H, e);
    });
}
int cipc_gpu_eval_snh(int n, const double* F, double u, double r, double* P, double* H, double* e) {
    return guard([&] { require(n >= 0 && (n == 0 || (F && P && H && e)), "invalid arguments"); if (n) cipc_gpu_eval_snh_impl(n, F, u, r, P, H, e); });
}
int cipc_gpu_eval_distance(int n, int kind, const double* pts, double* d, double* g, double* H) {
    return guard([&] {
        require(n >= 0 && kind >= 0 && kind <= 4 && (n == 0 || (pts && d && g && H)), "invalid arguments");
        if (n) cipc_gpu_eval_distance_impl(n, kind, pts, d, g, H);
    });
}
int cipc_gpu_eval_make_pd(int n, int dim, double* H) {
    return guard([&] {
        require(n >= 0 && (dim == 6 || dim == 9 || dim == 12) && (n == 0 || H), "invalid arguments");
        if (n) cipc_gpu_eval_make_pd_impl(n, dim, H);
    });
}
int cipc_gpu_eval_contact_pd(int n, int nverts, double* H, const double* G) {
    return guard([&] {
        require(n >= 0 && nverts >= 2 && nverts <= 4 && (n == 0 || H), "invalid arguments");
        require(G == nullptr || nverts >= 3, "the structured projection needs 3 or 4 vertices");
        if (n) cipc_gpu_eval_contact_pd_impl(n, nverts, H, G);
    });
}
int cipc_gpu_eval_accd(int n, int kind, const double* pts, const double* dirs, double eta, double th, double* toc) {
    return guard([&] {
        require(n >= 0 && (kind == 0 || kind == 1) && (n == 0 || (pts && dirs && toc)), "invalid arguments");
        if (n) cipc_gpu_eval_accd_impl(n, kind, pts, dirs, eta, th, toc);
    });
}
int cipc_gpu_eval_classify(int n, int kind, const double* pts, int* type) {
    return guard([&] {
        require(n >= 0 && (kind == 0 || kind == 1) && (n == 0 || (pts && type)), "invalid arguments");
        if (n) cipc_gpu_eval_classify_impl(n, kind, pts, type);
    });
}

}  // extern "C"

namespace cipc {
void assemble(Context& c, double kappa) { assemble_impl(static_cast<cipc_gpu_context*>(&c), kappa); }
}  // namespace cipc
