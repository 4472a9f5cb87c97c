/* Host-side scene set-up of the MI355X Newton-iteration library: the reference's mesh
 * loaders, surface extraction, rest-shape set-up, parameter file, named scenes and mesh /
 * checkpoint files, producing the mesh3D arrays that cipc_gpu_upload_scene consumes.
 *
 * Same conventions as cipc_gpu.h: every entry point returns CIPC_OK or a CIPC_ERR_* code,
 * with the message in cipc_gpu_last_error().  Nothing here touches the GPU. */
#ifndef CIPC_HOST_H
#define CIPC_HOST_H

#include "cipc_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cipc_host_mesh cipc_host_mesh;

/* An empty mesh3D with the parameter file's settings (loadSettings, Simulator.cpp:174-263;
 * settings may be NULL for the reference's built-in defaults, applyDefaultSettings
 * Simulator.cpp:143-172). */
int cipc_host_mesh_create(const char* settings, cipc_host_mesh** out);
int cipc_host_mesh_destroy(cipc_host_mesh* mesh);
/* mesh3D::load_tetrahedraMesh (SimulationMesh.cpp:174-361): Gmsh nodes and tetrahedra
 * appended to the mesh, scaled and offset. */
int cipc_host_load_tetrahedra(cipc_host_mesh* mesh, const char* path, double scale, const double offset[3]);
/* mesh3D::load_triangleMesh (SimulationMesh.cpp:545-716): OBJ vertices and faces (fan
 * triangulated); type 2 marks the vertices external colliders, type 3 puts the faces into
 * the surface instead of the cloth triangles. */
int cipc_host_load_triangles(cipc_host_mesh* mesh, const char* path, double scale, const double offset[3], int type);
/* The tail of FEMSimulator::buildModels (Simulator.cpp:468-506): initMesh3D (masses, Dm^-1,
 * volumes, triangle rest data), v_rest = V_prev = x, the inertial target, the scene-scaled
 * hhat / fhat / dTol, getSurface (SimulationMesh.cpp:363-446, the reference's face order)
 * and the average surface edge length. */
int cipc_host_finalize(cipc_host_mesh* mesh);
/* FEMSimulator::buildModels for a named scene: "twisting-mat", "twisting-mat-soft",
 * "bunny2" or "cloth-bunny" (Simulator.cpp:265-396,442-466), assets under asset_dir in the
 * reference's Assets/ layout, settings as cipc_host_mesh_create. */
int cipc_host_scene_build(const char* scene, const char* asset_dir, const char* settings, cipc_host_mesh** out);

/* counts = {V, T, F (surface faces), Vs, Es, Tr (cloth triangles), Nb (bending elements),
 * Dirichlet vertices, soft-target vertices} */
int cipc_host_counts(const cipc_host_mesh* mesh, int64_t counts[9]);
/* The scene descriptor over the mesh's own arrays (valid while the mesh lives and is not
 * changed) and the time step's scalars (precision, preconditioner and solve limits at
 * FP64, block FSAI, 1e-6, 10000; max_newton 10000). */
int cipc_host_scene_desc(const cipc_host_mesh* mesh, cipc_scene_desc* desc, cipc_step_params* params);
/* Boundary conditions of the scene: motion 0 none, 1 twisting; the Dirichlet vertices and
 * the soft-target vertices with their weight (any output may be NULL). */
int cipc_host_boundary(const cipc_host_mesh* mesh, int* motion, const int** dirichlet, const int** soft, double* soft_weight);
/* twistingBoundaryDirection (Simulator.cpp:24-46) for n vertices, in the form of
 * cipc_dirichlet_fn (user is ignored). */
int cipc_host_twisting_directions(void* user, int n, const int* vertices, const double* x, const double* v_rest, int step,
                                  double fraction, double dt, double* directions);
/* The twisting-mat-soft targets of updateSoftTargets (BoundaryConditions.cpp:112-135):
 * x - twistingBoundaryDirection(x, rest, 1, dt) for n vertices. */
int cipc_host_twisting_targets(int n, const double* x, const double* v_rest, double dt, double* targets);
/* Positions and inertial target of the mesh (3V each; either may be NULL), as a time step
 * leaves them, for the files below. */
int cipc_host_set_state(cipc_host_mesh* mesh, const double* x, const double* xtilde);
/* output_tetrahedraMesh (SimulationMesh.cpp:448-468). */
int cipc_host_write_msh(const cipc_host_mesh* mesh, const char* path);
/* output_tetTempData / load_tetTempData (SimulationMesh.cpp:470-543): vertex.txt and
 * vertexXtile.txt under dir; loading sets x, V_prev and the inertial target, and fails
 * (CIPC_ERR_RUNTIME) on missing files or a different vertex count. */
int cipc_host_save_checkpoint(const cipc_host_mesh* mesh, const char* dir);
int cipc_host_load_checkpoint(cipc_host_mesh* mesh, const char* dir);

#ifdef __cplusplus
}
#endif
#endif
