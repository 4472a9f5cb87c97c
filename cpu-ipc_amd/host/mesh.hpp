// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate logic of CPU-IPC (MPL-2.0);
// each such part cites the source it follows.
//
// Host side of the scene set-up: the mesh3D arrays (CPU IPC/SimulationMesh.h:37-129) as the
// reference's loaders, getSurface and initMesh3D produce them, the parameter file, the
// named scenes of FEMSimulator::buildModels, and the mesh / checkpoint files.  Everything
// here runs once per scene on the host; the device path consumes the resulting arrays.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace cipc {
namespace host {

using Vec3 = std::array<double, 3>;

struct Mesh {
    // state and static arrays (mesh3D)
    std::vector<Vec3> x, v_rest, V_prev, xtilde, velocities;
    std::vector<double> masses;
    std::vector<int> boundary;                       // 0 free, 1 Dirichlet, 2 external collider
    std::vector<std::array<int, 4>> tets;
    std::vector<std::array<int, 3>> triangles;
    std::vector<std::array<double, 9>> dm_inv;       // column-major Dm^-1 per tet
    std::vector<double> volumes;
    std::vector<std::array<double, 4>> dm_tri_inv;   // column-major 2x2 per triangle
    std::vector<double> areas;                       // area x thickness
    std::vector<std::array<int, 4>> surface;         // v0, v1, v2, tet
    std::vector<uint64_t> surf_verts;
    std::vector<std::pair<uint64_t, uint64_t>> surf_edges;
    std::vector<std::array<int, 2>> tri_edges, tri_edge_adj;   // interior cloth edges and their opposite vertices
    std::vector<std::array<int, 4>> bend_verts;      // QuadBendingInfo::verts
    std::vector<std::array<double, 16>> bend_q;      // QuadBendingInfo::Q, row-major
    Vec3 max_corner{{-1e32, -1e32, -1e32}}, min_corner{{1e32, 1e32, 1e32}};
    Vec3 obj_max{{0, 0, 0}}, obj_min{{0, 0, 0}};
    // material and solver scalars (parameterSetting.txt and the scene's overrides)
    double density = 1e3, youngs = 1e4, poisson = 0.49, friction = 0.5;
    double cloth_thickness = 1e-3, cloth_youngs = 1e4, shear_youngs = 1e1, bend_youngs = 1e4, cloth_density = 1e2;
    double strain_rate = 1e2, drag = 0.0, dt = 1.0 / 240.0, newton_threshold = 1e-2;
    double hhat = 9e-8, fhat = 1e-6, kappa = 0.0, dtol = 1e-18;
    bool use_barrier = true, quasi_static = false, apply_gravity = true;
    double length_rate = 0, volume_rate = 0, stretch = 0, shear = 0, plate = 0;
    double bbox_diag2 = 0, mean_mass = 0, avg_edge_len = 0;
    // boundary conditions of the named scenes
    std::vector<int> dirichlet, soft;
    double soft_weight = 0.0;
    int motion = 0;                                  // 0 none, 1 twisting (Dirichlet or soft targets)
};

// load_tetrahedraMesh (SimulationMesh.cpp:174-361); false and a message on a malformed file
bool load_tetrahedra(Mesh& m, const std::string& path, double scale, const Vec3& offset, std::string& err);
// load_triangleMesh (SimulationMesh.cpp:545-716): type 2 marks external colliders, type 3
// puts the faces straight into the surface
bool load_triangles(Mesh& m, const std::string& path, double scale, const Vec3& offset, int type, std::string& err);
// getSurface (SimulationMesh.cpp:363-446) with the reference's face order
void get_surface(Mesh& m);
// initMesh3D (Elasticity.cpp:1050-1116): masses, Dm^-1, volumes, triangle rest data, meanMass
void init_mesh(Mesh& m);
// updateMaterial (Simulator.cpp:128-141)
void update_material(Mesh& m);
// loadSettings (Simulator.cpp:174-263); throws std::runtime_error
void load_settings(Mesh& m, const std::string& path);
// FEMSimulator::buildModels (Simulator.cpp:385-540) for "twisting-mat", "twisting-mat-soft",
// "bunny2" and "cloth-bunny"; assets under asset_dir (the reference's Assets/ layout)
void build_scene(Mesh& m, const std::string& scene, const std::string& asset_dir, const std::string& settings_path);
// the tail of buildModels (Simulator.cpp:468-506): initMesh3D, v_rest, V_prev, the inertial
// target, the scene-scaled hhat / fhat / dTol, getSurface, the average surface edge length
void finalize_scene(Mesh& m);
// prepareQuadraticBending (Simulator.cpp:48-109) on the current positions
void prepare_quadratic_bending(Mesh& m);
// twistingBoundaryDirection (Simulator.cpp:24-46)
Vec3 twisting_direction(const Vec3& x, const Vec3& rest, double fraction, double dt);
// output_tetrahedraMesh (SimulationMesh.cpp:448-468)
bool write_msh(const Mesh& m, const std::string& path);
// output_tetTempData / load_tetTempData (SimulationMesh.cpp:470-543): vertex.txt and
// vertexXtile.txt in dir
bool save_checkpoint(const Mesh& m, const std::string& dir);
bool load_checkpoint(Mesh& m, const std::string& dir);

}  // namespace host
}  // namespace cipc
