// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate expressions of CPU-IPC
// (MPL-2.0) in the reference's operation order; each such part cites the source it follows.
/* ORACLE — test infrastructure only (see oracle_math.hpp for the full notice).
 *
 * C interface of the CPU restatement of CPU-IPC's per-Newton-iteration path.  The
 * scene descriptor mirrors the reference's mesh3D arrays (CPU IPC/SimulationMesh.h:37-129)
 * with their memory layouts: positions are std::vector<Eigen::Vector3d> (3 doubles per
 * vertex), tets and surface faces Eigen::Vector4i, Dm^-1 a column-major Matrix3d,
 * surfVerts uint64 and surfEdges pairs of uint64.
 */
#ifndef CIPC_ORACLE_H
#define CIPC_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int V, T, F, Vs, Es;
    const double* x;          /* 3V */
    const double* v_rest;     /* 3V */
    const double* V_prev;     /* 3V */
    const double* xtilde;     /* 3V inertial target */
    const double* masses;     /* V */
    const int* boundary;      /* V */
    const int* tets;          /* 4T */
    const double* dm_inv;     /* 9T column-major */
    const double* volumes;    /* T */
    const int* faces;         /* 4F (v0, v1, v2, tetId) */
    const uint64_t* surf_verts; /* Vs */
    const uint64_t* surf_edges; /* 2Es */
    double dt, length_rate, volume_rate, drag, hhat, fhat, friction;
    double ground_normal[3], ground_d;
    /* cloth triangles and bending elements (0 / NULL for tet-only scenes) */
    int Tr;                      /* mesh3D::triangleNum */
    const int* triangles;        /* 3Tr */
    const double* dm_tri_inv;    /* 4Tr column-major 2x2 */
    const double* areas;         /* Tr (area x thickness) */
    double stretch, shear, strain_rate, plate_rigidity;
    int Nb, bending;             /* bending: 0 quadratic, 1 hinge */
    const int* bend_verts;       /* 4Nb */
    const double* bend_q;        /* 16Nb (bending 0) */
    const double* bend_rest;     /* Nb (bending 1) */
    const double* bend_weight;   /* Nb (bending 1) */
} oracle_scene_desc;

typedef struct oracle_mesh oracle_mesh;

oracle_mesh* oracle_create(const oracle_scene_desc* d);
void oracle_destroy(oracle_mesh* m);
void oracle_set_threads(oracle_mesh* m, int threads);
void oracle_set_state(oracle_mesh* m, const double* x, const double* V_prev, const double* xtilde);

/* Broad phase + active set (SpatialHash::build + calculateActivateSet with the LBVH
 * backend, Ground::calculateActivateSet).  counts = {Self, SelfEE, CCD, Environment}. */
void oracle_active_set(oracle_mesh* m, int64_t counts[4]);
void oracle_get_active_set(const oracle_mesh* m, int* self4, int* ee4, int* eeIe2, int* ccd2, int* env);
void oracle_set_active_set(oracle_mesh* m, const int64_t counts[4], const int* self4, const int* ee4,
                           const int* eeIe2, const int* ccd2, const int* env);

/* computeGradientAndHessian + NewtonLinearSystem::assemble.  Returns collision count;
 * nnz via out pointer; the CSR is the Eigen lower CSC (== upper CSR). */
int oracle_assemble(oracle_mesh* m, double kappa, int64_t* nnz);
void oracle_get_gradient(const oracle_mesh* m, double* grad);
/* per gradient entry, the sum of the magnitudes of the terms added into it (3V): the scale of
 * its rounding error where the terms cancel */
void oracle_get_gradient_scale(const oracle_mesh* m, double* scale);
/* per CSR entry (nnz), the sum over the element / contact / mass blocks added into it of
 * n times the block's largest magnitude (n its order, a bound of its 2-norm): the scale of
 * its rounding error where the blocks cancel */
void oracle_get_csr_scale(const oracle_mesh* m, double* scale);
void oracle_get_csr(const oracle_mesh* m, int* outer, int* inner, double* val, double* rhs);

/* computeEnergy. */
double oracle_energy(oracle_mesh* m, double kappa);
/* per-term energies: {SNH (before dt^2), inertia+drag, ground barrier, self barrier, mollified EE,
 * friction, Baraff-Witkin triangles (before dt^2), bending (before dt^2), soft targets} */
void oracle_energy_terms(oracle_mesh* m, double kappa, double terms[9]);
/* initKappa's reductions: out = {sum g_c . g_E, sum |g_c|^2}; returns #ground + #self constraints */
int oracle_kappa_gradients(oracle_mesh* m, double out[2]);
/* postLineSearch's bookkeeping: returns 1 when a recorded close constraint got closer, then
 * records the constraints below dtol; counts = {ground, self} recorded */
int oracle_close_constraints(oracle_mesh* m, double dtol, int64_t counts[2]);
void oracle_clear_close_constraints(oracle_mesh* m);
/* the kappa loop's exit test (IPCSolver.cpp:1171-1188): out = {min, max} over the ground then
 * self constraint values; returns their number (0: out undefined) */
int oracle_constraint_extrema(oracle_mesh* m, double out[2]);
/* soft target-position penalty 0.5 w |x_v - t|^2 on n vertices (SoftBoundaryCondition) */
void oracle_set_soft(oracle_mesh* m, int n, const int* idx, const double* targets, double weight);
/* Friction::initialize: lagged friction state from the current active set and positions. */
void oracle_friction_update(oracle_mesh* m, double kappa);
/* counts[0] self contacts, counts[1] ground contacts; lam (n), coord (2n), basis (6n: two
 * columns of 3), env_lam (counts[1]); any output may be NULL. */
void oracle_get_friction(const oracle_mesh* m, int64_t counts[2], double* lam, double* coord, double* basis, double* env_lam);

/* Newton step bound (IPCSolver.cpp:943-990): out = {alpha, ground, partial, cfl, full} */
void oracle_step_bound(oracle_mesh* m, const double* p, double out[5]);
/* Self_largestFeasibleStepSize_CCD alone with slackness 0.8 from step alpha0 (the animated
 * Dirichlet update, IPCSolver.cpp:1131-1134) */
double oracle_full_ccd(oracle_mesh* m, const double* p, double alpha0);

/* x <- x0 - alpha p on free vertices (stepForward, IPCSolver.cpp:458-473) */
void oracle_step_forward(oracle_mesh* m, const double* x0, const double* p, double alpha);
void oracle_get_positions(const oracle_mesh* m, double* x);
/* isIntersected (IPCSolver.cpp:523-542) with a static LBVH built on the current x. */
int oracle_is_intersected(oracle_mesh* m);

/* ---- element-level kernels (for unit tests and golden vectors) ---- */
void oracle_svd3(const double* F, double* U, double* s, double* V);
void oracle_snh_element(const double* F, double u, double r, double* P, double* H9, double* energy);
int oracle_make_pd(int n, double* H);  /* n in {3,6,9,12}; row-major */
/* Baraff-Witkin triangle: F 3x2 column-major -> PK1 (6), projected 6x6 Hessian, energy density */
void oracle_bw_element(const double* F6, double stretch, double shear, double k, double* P6, double* H36, double* e);
/* edgeTheta: hinge points q0..q3 (12) -> angle, gradient (12), second-derivative matrix (144 row-major) */
double oracle_edge_theta(const double* pts, double* g, double* H);
/* hinge bending element before dt^2: gradient (12), makePD'd Hessian (144), energy */
void oracle_hinge_element(const double* pts, double rest, double weight, double plate, double* g, double* H, double* e);
void oracle_barrier(double d, double dhat, double* b, double* g, double* h);
/* type: 0 PP(2 pts) 1 PE(3) 2 PT(4) 3 EE(4) 4 EE cross norm(4); pts 12 doubles */
void oracle_distance(int type, const double* pts, double* d, double* g, double* H);
int oracle_dtype_pt(const double* pts);
int oracle_dtype_ee(const double* pts);
double oracle_accd_pt(const double* pts, const double* dirs, double eta, double thickness);
double oracle_accd_ee(const double* pts, const double* dirs, double eta, double thickness);
int oracle_seg_tri_intersect(const double* pts);
/* LBVH over boxes (6 doubles each: lower xyz, upper xyz); query returns sorted ids. */
int64_t oracle_lbvh_query(int n, const double* boxes, const double* query_box, double gap, int* out, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
