// C ABI of the host-side scene set-up (declared in include/cipc_host.h).
#include "../../include/cipc_host.h"

#include <cmath>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>

#include "mesh.hpp"

using cipc::host::Mesh;
using cipc::host::Vec3;

struct cipc_host_mesh {
    Mesh m;
};

// the library's shared error channel (context.hip)
extern "C" int cipc_internal_fail(int code, const char* msg);

namespace {
template <typename F>
int guard(F&& f) {
    try {
        f();
        return CIPC_OK;
    } catch (const std::invalid_argument& e) {
        return cipc_internal_fail(CIPC_ERR_INVALID_ARGUMENT, e.what());
    } catch (const std::bad_alloc&) {
        return cipc_internal_fail(CIPC_ERR_RUNTIME, "host allocation failed");
    } catch (const std::exception& e) {
        return cipc_internal_fail(CIPC_ERR_RUNTIME, e.what());
    }
}
void require(bool ok, const char* msg) {
    if (!ok) throw std::invalid_argument(msg);
}
Vec3 v3(const double* p) { return {{p[0], p[1], p[2]}}; }
}  // namespace

extern "C" {

int cipc_host_mesh_create(const char* settings, cipc_host_mesh** out) {
    return guard([&] {
        require(out != nullptr, "null output");
        auto* h = new cipc_host_mesh();
        try {
            if (settings) cipc::host::load_settings(h->m, settings);
            else cipc::host::update_material(h->m);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int cipc_host_mesh_destroy(cipc_host_mesh* mesh) {
    delete mesh;
    return CIPC_OK;
}

int cipc_host_load_tetrahedra(cipc_host_mesh* mesh, const char* path, double scale, const double offset[3]) {
    return guard([&] {
        require(mesh && path && offset, "null mesh, path or offset");
        std::string err;
        if (!cipc::host::load_tetrahedra(mesh->m, path, scale, v3(offset), err)) throw std::runtime_error(err);
    });
}

int cipc_host_load_triangles(cipc_host_mesh* mesh, const char* path, double scale, const double offset[3], int type) {
    return guard([&] {
        require(mesh && path && offset, "null mesh, path or offset");
        std::string err;
        if (!cipc::host::load_triangles(mesh->m, path, scale, v3(offset), type, err)) throw std::runtime_error(err);
    });
}

int cipc_host_finalize(cipc_host_mesh* mesh) {
    return guard([&] {
        require(mesh != nullptr, "null mesh");
        require(!mesh->m.x.empty(), "the mesh has no vertices");
        cipc::host::finalize_scene(mesh->m);
    });
}

int cipc_host_scene_build(const char* scene, const char* asset_dir, const char* settings, cipc_host_mesh** out) {
    return guard([&] {
        require(scene && asset_dir && settings && out, "null scene, asset directory, settings or output");
        auto* h = new cipc_host_mesh();
        try {
            std::string dir(asset_dir);
            if (!dir.empty() && dir.back() != '/') dir += '/';
            cipc::host::build_scene(h->m, scene, dir, settings);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int cipc_host_counts(const cipc_host_mesh* mesh, int64_t counts[9]) {
    return guard([&] {
        require(mesh && counts, "null mesh or output");
        const Mesh& m = mesh->m;
        const int64_t c[9] = {(int64_t)m.x.size(), (int64_t)m.tets.size(), (int64_t)m.surface.size(), (int64_t)m.surf_verts.size(),
                              (int64_t)m.surf_edges.size(), (int64_t)m.triangles.size(), (int64_t)m.bend_verts.size(),
                              (int64_t)m.dirichlet.size(), (int64_t)m.soft.size()};
        for (int i = 0; i < 9; ++i) counts[i] = c[i];
    });
}

int cipc_host_scene_desc(const cipc_host_mesh* mesh, cipc_scene_desc* d, cipc_step_params* p) {
    return guard([&] {
        require(mesh && d, "null mesh or descriptor");
        const Mesh& m = mesh->m;
        require(m.xtilde.size() == m.x.size() && m.dm_inv.size() == m.tets.size(), "the mesh is not finalised");
        *d = cipc_scene_desc{};
        d->V = (int)m.x.size();
        d->T = (int)m.tets.size();
        d->F = (int)m.surface.size();
        d->Vs = (int)m.surf_verts.size();
        d->Es = (int)m.surf_edges.size();
        d->x = m.x.data()->data();
        d->v_rest = m.v_rest.data()->data();
        d->V_prev = m.V_prev.data()->data();
        d->xtilde = m.xtilde.data()->data();
        d->masses = m.masses.data();
        d->boundary = m.boundary.data();
        d->tets = d->T ? m.tets.data()->data() : nullptr;
        d->dm_inv = d->T ? m.dm_inv.data()->data() : nullptr;
        d->volumes = d->T ? m.volumes.data() : nullptr;
        d->faces = d->F ? m.surface.data()->data() : nullptr;
        d->surf_verts = d->Vs ? m.surf_verts.data() : nullptr;
        static_assert(sizeof(std::pair<uint64_t, uint64_t>) == 2 * sizeof(uint64_t), "edge pairs are two packed indices");
        d->surf_edges = d->Es ? reinterpret_cast<const uint64_t*>(m.surf_edges.data()) : nullptr;
        d->dt = m.dt;
        d->length_rate = m.length_rate;
        d->volume_rate = m.volume_rate;
        d->drag = m.drag;
        d->hhat = m.hhat;
        d->fhat = m.fhat;
        d->friction = m.friction;
        // the fixed half-space ground y >= -1 (CollisionBroadPhase.cpp:1062-1074)
        d->ground_normal[0] = 0; d->ground_normal[1] = 1; d->ground_normal[2] = 0;
        d->ground_d = -1;
        d->Tr = (int)m.triangles.size();
        d->triangles = d->Tr ? m.triangles.data()->data() : nullptr;
        d->dm_tri_inv = d->Tr ? m.dm_tri_inv.data()->data() : nullptr;
        d->areas = d->Tr ? m.areas.data() : nullptr;
        d->stretch = m.stretch;
        d->shear = m.shear;
        d->strain_rate = m.strain_rate;
        d->plate_rigidity = m.plate;
        d->Nb = (int)m.bend_verts.size();
        d->bending = 0;
        d->bend_verts = d->Nb ? m.bend_verts.data()->data() : nullptr;
        d->bend_q = d->Nb ? m.bend_q.data()->data() : nullptr;
        if (p) {
            *p = cipc_step_params{};
            p->bbox_diag2 = m.bbox_diag2;
            p->mean_mass = m.mean_mass;
            p->dtol = m.dtol;
            p->newton_threshold = m.newton_threshold;
            p->apply_gravity = m.apply_gravity ? 1 : 0;
            p->quasi_static = m.quasi_static ? 1 : 0;
            p->precision = 0;
            p->preconditioner = 1;
            p->solve_rel_tol = 1e-6;
            p->solve_max_iter = 10000;
            p->max_newton = 10000;
        }
    });
}

int cipc_host_boundary(const cipc_host_mesh* mesh, int* motion, const int** dirichlet, const int** soft, double* soft_weight) {
    return guard([&] {
        require(mesh != nullptr, "null mesh");
        if (motion) *motion = mesh->m.motion;
        if (dirichlet) *dirichlet = mesh->m.dirichlet.empty() ? nullptr : mesh->m.dirichlet.data();
        if (soft) *soft = mesh->m.soft.empty() ? nullptr : mesh->m.soft.data();
        if (soft_weight) *soft_weight = mesh->m.soft_weight;
    });
}

int cipc_host_twisting_directions(void*, int n, const int*, const double* x, const double* v_rest, int, double fraction, double dt,
                                  double* out) {
    return guard([&] {
        require(n >= 0 && (n == 0 || (x && v_rest && out)), "invalid arguments");
        require(std::isfinite(fraction) && fraction >= 0.0 && fraction <= 1.0 && std::isfinite(dt) && dt > 0.0,
                "Dirichlet update received invalid time data");
        for (int i = 0; i < n; ++i) {
            const Vec3 d = cipc::host::twisting_direction(v3(x + 3 * i), v3(v_rest + 3 * i), fraction, dt);
            for (int k = 0; k < 3; ++k) out[3 * i + k] = d[k];
        }
    });
}

int cipc_host_twisting_targets(int n, const double* x, const double* v_rest, double dt, double* targets) {
    return guard([&] {
        require(n >= 0 && (n == 0 || (x && v_rest && targets)), "invalid arguments");
        require(std::isfinite(dt) && dt > 0.0, "soft boundary update received invalid time data");
        for (int i = 0; i < n; ++i) {
            const Vec3 d = cipc::host::twisting_direction(v3(x + 3 * i), v3(v_rest + 3 * i), 1.0, dt);
            for (int k = 0; k < 3; ++k) targets[3 * i + k] = x[3 * i + k] - d[k];
        }
    });
}

int cipc_host_set_state(cipc_host_mesh* mesh, const double* x, const double* xtilde) {
    return guard([&] {
        require(mesh != nullptr, "null mesh");
        Mesh& m = mesh->m;
        for (size_t i = 0; i < m.x.size(); ++i)
            for (int k = 0; k < 3; ++k) {
                if (x) m.x[i][k] = x[3 * i + k];
                if (xtilde && i < m.xtilde.size()) m.xtilde[i][k] = xtilde[3 * i + k];
            }
    });
}

int cipc_host_write_msh(const cipc_host_mesh* mesh, const char* path) {
    return guard([&] {
        require(mesh && path, "null mesh or path");
        if (!cipc::host::write_msh(mesh->m, path)) throw std::runtime_error(std::string("unable to write ") + path);
    });
}

int cipc_host_save_checkpoint(const cipc_host_mesh* mesh, const char* dir) {
    return guard([&] {
        require(mesh && dir, "null mesh or directory");
        if (!cipc::host::save_checkpoint(mesh->m, dir)) throw std::runtime_error(std::string("unable to write a checkpoint under ") + dir);
    });
}

int cipc_host_load_checkpoint(cipc_host_mesh* mesh, const char* dir) {
    return guard([&] {
        require(mesh && dir, "null mesh or directory");
        if (!cipc::host::load_checkpoint(mesh->m, dir)) throw std::runtime_error(std::string("no usable checkpoint under ") + dir);
    });
}

}  // extern "C"
