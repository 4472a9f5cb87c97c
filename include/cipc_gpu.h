/*
 * cipc_gpu.h — C ABI of the MI355X (gfx950) per-Newton-iteration path of CPU-IPC.
 *
 * The reference's hot path lives behind three C++ interfaces (paths relative to the
 * reference root, "CPU IPC/" omitted):
 *   - the Newton step API    solveIPCStep / solveBarrierSubproblem   (IPCSolver.h:63-68, IPCSolver.cpp:890-1064)
 *   - the broad-phase facade  SpatialHash::build / calculateActivateSet with BroadPhaseBackend
 *                                                                      (CollisionBroadPhase.h:11-80)
 *   - the linear system       NewtonLinearSystem::solve / assemble    (NewtonLinearSystem.h:22-28)
 * Each entry point below replaces one whole batch function of those interfaces (never a
 * per-primitive query, which the reference calls from inside TBB loops).  All arguments
 * are plain pointers and sizes in the reference's own memory layouts:
 *   std::vector<Eigen::Vector3d>  -> const double*  (3 doubles per vertex, AoS)
 *   std::vector<Eigen::Vector4i>  -> const int*     (4 ints per tet / per surface face)
 *   std::vector<Eigen::Matrix3d>  -> const double*  (9 doubles, column-major)
 *   std::vector<uint64_t>, std::vector<std::pair<uint64_t,uint64_t>> -> const uint64_t*
 *   std::vector<EncodedContact>   -> int*           (4 ints per contact)
 *   std::vector<std::pair<int,int>> -> int*         (2 ints per pair)
 *   Eigen::SparseMatrix<double> (lower, column-major) -> outer/inner/values arrays
 *
 * Every call is synchronous with respect to the host and returns CIPC_OK or an error
 * code; cipc_gpu_last_error() returns the message of the most recent failure in the
 * calling thread.  Error codes map to the exception types the reference throws at the
 * same places (std::invalid_argument, std::runtime_error, std::logic_error,
 * std::overflow_error).  One context owns one device, its streams (all at the default
 * priority) and all device memory of one scene; every call returns with its work done;
 * calls on one context must not overlap.
 */
#ifndef CIPC_GPU_H
#define CIPC_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CIPC_OK 0
#define CIPC_ERR_INVALID_ARGUMENT 1 /* std::invalid_argument in the reference */
#define CIPC_ERR_RUNTIME 2          /* std::runtime_error */
#define CIPC_ERR_LOGIC 3            /* std::logic_error */
#define CIPC_ERR_OVERFLOW 4         /* std::overflow_error */
#define CIPC_ERR_HIP 5              /* HIP runtime failure */
#define CIPC_ERR_NO_DEVICE 6        /* no usable gfx950 device */

typedef struct cipc_gpu_context cipc_gpu_context;

/* Scene arrays of mesh3D (SimulationMesh.h:37-129) and the scalars of the barrier
 * subproblem.  hhat is the squared, scene-scaled dHat (Simulator.cpp:486-489); the ground
 * is the half-space n.x >= D (CollisionBroadPhase.cpp:1062-1074). */
typedef struct {
    int V, T, F, Vs, Es;
    const double* x;            /* mesh3D::vertexes                    3V  */
    const double* v_rest;       /* mesh3D::v_rest                      3V  */
    const double* V_prev;       /* mesh3D::V_prev                      3V  */
    const double* xtilde;       /* mesh3D::inertialTarget              3V  */
    const double* masses;       /* mesh3D::masses                      V   */
    const int* boundary;        /* mesh3D::boundaryTypes               V   */
    const int* tets;            /* mesh3D::tetrahedras                 4T  */
    const double* dm_inv;       /* mesh3D::DM_tetrahedra_inverse       9T  */
    const double* volumes;      /* mesh3D::volum                       T   */
    const int* faces;           /* mesh3D::surface (v0,v1,v2,tet)      4F  */
    const uint64_t* surf_verts; /* mesh3D::surfVerts                   Vs  */
    const uint64_t* surf_edges; /* mesh3D::surfEdges                   2Es */
    double dt, length_rate, volume_rate, drag, hhat, fhat, friction;
    double ground_normal[3], ground_d;
    /* Cloth (Baraff-Witkin triangles, IPCSolver.cpp:166-194) and bending (IPCSolver.cpp:
     * 197-267); Tr = Nb = 0 and NULL pointers for tet-only scenes. */
    int Tr;                     /* mesh3D::triangleNum                 */
    const int* triangles;       /* mesh3D::triangles                   3Tr */
    const double* dm_tri_inv;   /* mesh3D::DM_triangle_inverse         4Tr (2x2 column-major) */
    const double* areas;        /* mesh3D::areas (area x thickness)    Tr  */
    double stretch, shear, strain_rate, plate_rigidity; /* stretchStiffness, shearStiffness, strainRate, plateRigidity */
    int Nb;                     /* bending elements                    */
    int bending;                /* 0: quadratic form (USE_QUADRATIC_BENDING), 1: hinge angle */
    const int* bend_verts;      /* QuadBendingInfo::verts / HingeBendingInfo::vertices   4Nb */
    const double* bend_q;       /* QuadBendingInfo::Q, row-major 4x4 (bending 0)         16Nb */
    const double* bend_rest;    /* HingeBendingInfo::restAngle (bending 1)              Nb  */
    const double* bend_weight;  /* HingeBendingInfo::geometricWeight (bending 1)        Nb  */
} cipc_scene_desc;

/* Stage times of one Newton iteration, milliseconds, measured with HIP events (the
 * reference's StageTimer, StageTimer.h:5-27).  The broad phase and the assembly are timed on
 * the context's stream; the elastic part of the assembly, the step bound and the energy run
 * on streams of their own beside it, and are timed there, so the stages overlap and
 * total_ms is less than their sum. */
typedef struct {
    double broadphase_ms; /* LBVH build + active set (buildCollisionSets, IPCSolver.cpp:544-556) */
    double assembly_ms;   /* gradient + PSD Hessian -> lower CSR (IPCSolver.cpp:918-930 less the solve) */
    double elastic_ms;    /* inertia + tet part of the assembly (element kernel and its gathers) */
    double ccd_ms;        /* ground + partial CCD + CFL + swept LBVH + full CCD (IPCSolver.cpp:943-990), beside the assembly */
    double energy_ms;     /* one line-search energy evaluation (IPCSolver.cpp:322-456), beside the assembly's tail */
    double total_ms;
    double alpha;         /* merged step bound */
    double energy;        /* energy at the current positions */
    int64_t nnz;          /* CSR nonzeros */
    int64_t contacts;     /* |Self_ActiveSet| + |Self_EE_ActiveSet| + |Environment_ActiveSet| */
    int64_t collisions;   /* compute_g_dpt's return value (ContactMechanics.cpp:2970-3032), as cipc_gpu_assemble */
    /* the parts of alpha, as cipc_gpu_step_bound's out[1..4]: ground bound, partial CCD, CFL, full CCD */
    double bound_ground, bound_partial, bound_cfl, bound_full;
    /* 1 when the stages ran as one HIP graph launch (then only total_ms and, for a batch,
     * energy_ms are timed; the other stage times are zero) */
    int graph;
    /* host time of the first round's enqueue (the launches one by one, or the graph launch
     * with its capture when one is made): the iteration waits on the host when this is
     * close to total_ms */
    double enqueue_ms;
} cipc_iteration_timing;

/* Result of one Newton linear solve. */
typedef struct {
    int iterations;       /* conjugate-gradient iterations (inner iterations summed for precision 1) */
    int refinements;      /* FP64 refinement steps (precision 1), 0 for precision 0 */
    double rel_residual;  /* ||rhs - H p|| / ||rhs|| evaluated in FP64 at exit */
    double solve_ms;      /* device time of the solve, HIP events on the context's stream */
} cipc_solve_stats;

const char* cipc_gpu_last_error(void);

/* Context lifetime; device is a HIP device ordinal. */
int cipc_gpu_context_create(int device, cipc_gpu_context** out);
int cipc_gpu_context_destroy(cipc_gpu_context* ctx);

/* Upload topology, material and initial state once per scene (FEMSimulator::buildModels,
 * Simulator.cpp:426-529, after initMesh3D and getSurface).  Builds the static elastic
 * sparsity pattern. */
int cipc_gpu_upload_scene(cipc_gpu_context* ctx, const cipc_scene_desc* desc);

/* State: any pointer may be NULL to keep the device copy.  Positions of the current
 * iterate, the previous step, and the inertial target (IPCSolver.cpp:55-68,1200-1204). */
int cipc_gpu_set_state(cipc_gpu_context* ctx, const double* x, const double* V_prev, const double* xtilde);
int cipc_gpu_get_positions(cipc_gpu_context* ctx, double* x);
/* Newton direction p (x_new = x - alpha p), 3V doubles (moveDir, IPCSolver.cpp:904). */
int cipc_gpu_set_direction(cipc_gpu_context* ctx, const double* p);

/* Static broad phase + active set: SpatialHash::build(mesh, h) with the LBVH backend
 * (CollisionBroadPhase.cpp:58-66; LBVH.cpp:284-305) followed by
 * SpatialHash::calculateActivateSet (CollisionBroadPhase.cpp:765-1059) and
 * Ground::calculateActivateSet (CollisionBroadPhase.cpp:1076-1092).
 * counts = {|Self_ActiveSet|, |Self_EE_ActiveSet|, |Self_CCD_ActiveSet|, |Environment_ActiveSet|}. */
int cipc_gpu_active_set(cipc_gpu_context* ctx, int64_t counts[4]);
/* Copy the sets out in the reference's order and encodings (EncodedContact.h:8-40).
 * Environment_ActiveSet is returned in ascending surface-vertex order (the reference's is
 * unordered).  Any pointer may be NULL. */
int cipc_gpu_get_active_set(cipc_gpu_context* ctx, int* self4, int* ee4, int* ee_ie2, int* ccd2, int* env);
/* Replace the device active set (for callers that build it elsewhere). */
int cipc_gpu_set_active_set(cipc_gpu_context* ctx, const int64_t counts[4], const int* self4, const int* ee4,
                            const int* ee_ie2, const int* ccd2, const int* env);

/* computeGradientAndHessian (IPCSolver.cpp:108-319) fused with
 * NewtonLinearSystem::assemble (NewtonLinearSystem.cpp:126-154): gradient, PSD-projected
 * Hessian blocks and the mass diagonal straight into the Eigen lower-triangular CSC
 * pattern (== upper CSR), exact zeros dropped before summation, rows/columns of
 * non-free vertices dropped, mass diagonal on every DOF.  Returns the nonzero count and
 * the collision count the reference accumulates (compute_g_dpt's return value). */
int cipc_gpu_assemble(cipc_gpu_context* ctx, double kappa, int64_t* nnz, int* collisions);
/* Precision of the elastic Hessian blocks of the next assemblies.  No reference counterpart
 * (the reference assembles in FP64).  0: FP64 (default); 1: the mixed form, tet, triangle
 * and bending blocks rounded to FP32 where they are staged and summed in FP64, with the
 * barrier, friction and mass parts, the gradient and the right-hand side in FP64. */
int cipc_gpu_set_hessian_precision(cipc_gpu_context* ctx, int precision);
/* Gradient of the last assemble (unmasked, 3V) and the CSR + masked right-hand side. */
int cipc_gpu_get_gradient(cipc_gpu_context* ctx, double* grad);
int cipc_gpu_get_csr(cipc_gpu_context* ctx, int* outer, int* inner, double* values, double* rhs);

/* computeEnergy (IPCSolver.cpp:322-456) at the current positions with the current active
 * set.  terms = {SNH elastic (before the dt^2 factor), inertia + drag, ground barrier sum,
 * self barrier sum (with multiplicity), mollified edge-edge barrier sum, lagged friction
 * energy (Friction::energy, Friction.cpp:459-545; 0 before cipc_gpu_friction_update),
 * Baraff-Witkin triangles (before dt^2), bending (before dt^2), soft targets (softEnergy,
 * BoundaryConditions.cpp:210-229)}. */
int cipc_gpu_energy(cipc_gpu_context* ctx, double kappa, double* energy);
int cipc_gpu_energy_terms(cipc_gpu_context* ctx, double kappa, double terms[9]);
/* No counterpart in the reference: cipc_gpu_energy_terms for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same nine terms, written to terms_device (device memory, 9 doubles); the call allocates
 * nothing: its reduction partials live in a fixed device array of the library (11 x 256
 * doubles, 22.5 KiB per GPU, held until the process ends), whose size does not depend on
 * the scene.  One scene only (a batch context is refused, CIPC_ERR_LOGIC); every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU race.  Synchronises the
 * context's stream. */
int cipc_gpu_energy_terms_noalloc(cipc_gpu_context* ctx, double kappa, double* terms_device);

/* Learned block preconditioner (solve preconditioner 2).  No counterpart in the reference.
 * A message-passing network (hidden width 64, 3 layers, bfloat16 weights with float32
 * accumulation) over the vertex-block graph of the Newton matrix predicts the block
 * lower-triangular factor G of M^-1 = G^T G as a correction of the first-order block FSAI
 * of the symmetrically scaled matrix.  Weights are a flat float32 vector of
 * cipc_gpu_preconditioner_param_count() values (layout of oracle/gnn_oracle.py).
 *  set_preconditioner_weights: weights for later solves with preconditioner 2;
 *  preconditioner_gradient: a training loss of the last assembled system and its gradient
 *    (param_count): objective 0, sum_z |G^ A^ G^T z - z|^2 / (n_z 3V) for the given probes
 *    (n_probes x 3V); objective 1, the mean squared difference of G^ from the block FSAI factor
 *    of the same matrix in the scaled basis (probes unused);
 *  preconditioner_train_step: one Adam step (beta 0.9 / 0.999) on the last assembled system
 *    (objective 0 with n_probes Gaussian probes drawn from `seed`); params and moments updated
 *    in place. */
int64_t cipc_gpu_preconditioner_param_count(void);
int cipc_gpu_set_preconditioner_weights(cipc_gpu_context* ctx, const float* weights, int64_t n);
int cipc_gpu_preconditioner_gradient(cipc_gpu_context* ctx, int objective, const float* params, const float* probes, int n_probes,
                                     double* loss, float* grad);
int cipc_gpu_preconditioner_train_step(cipc_gpu_context* ctx, int objective, float* params, float* adam_m, float* adam_v, int step,
                                       double lr, int n_probes, uint64_t seed, double* loss);

/* The data-parallel parts of the barrier-stiffness control; the scalar logic of
 * suggestKappa / upperBoundKappa / initKappa / postLineSearch (IPCSolver.cpp:751-869) stays
 * with the caller.
 *  kappa_gradients: initKappa's sums (IPCSolver.cpp:766-799) over the current positions and
 *    sets: out = {sum_v g_c.g_E, sum_v |g_c|^2} with g_E = computeEGradient (inertia without
 *    drag, tet elastic forces times dt^2; IPCSolver.cpp:74-106) and g_c the ground and
 *    Self_ActiveSet barrier gradients at kappa = 1; n_constraints = |Environment| + |Self|.
 *  close_constraints: postLineSearch's else-branch (IPCSolver.cpp:815-866): *closer = 1 when
 *    a recorded ground or self constraint is now at or below its recorded value, then the
 *    records are replaced by the current constraints with value < dtol (counts = {ground,
 *    self}).  clear_close_constraints empties the records (IPCSolver.cpp:1163-1166). */
int cipc_gpu_kappa_gradients(cipc_gpu_context* ctx, double out[2], int64_t* n_constraints);
int cipc_gpu_close_constraints(cipc_gpu_context* ctx, double dtol, int* closer, int64_t counts[2]);
int cipc_gpu_clear_close_constraints(cipc_gpu_context* ctx);

/* SoftBoundaryCondition (BoundaryConditions.h:54-67) after BoundaryConditionOps::
 * initialize / updateSoftTargets (BoundaryConditions.cpp:64-140): the penalty
 * 0.5 weight |x_v - target|^2 on n distinct free vertices, added to every later assemble
 * (addSoftDerivatives, BoundaryConditions.cpp:175-208) and energy.  n = 0 removes it.
 * CIPC_ERR_INVALID_ARGUMENT for a non-positive weight, non-finite targets, or a vertex
 * out of range, repeated or not free. */
int cipc_gpu_set_soft_targets(cipc_gpu_context* ctx, int n, const int* vertices, const double* targets, double weight);

/* NewtonLinearSystem::solve (NewtonLinearSystem.cpp:21-56) for the system of the last
 * cipc_gpu_assemble: H p = rhs with the relative-residual target and iteration cap of
 * LinearSolverOptions (LinearSolverOptions.h:9-19: 1e-6, 10000).  The solve runs on the
 * device with the 3x3 block-Jacobi preconditioner, in FP64 (precision 0) or in FP32 inside
 * FP64 iterative refinement (precision 1).  The solution becomes the direction of
 * cipc_gpu_step_bound / cipc_gpu_step_forward and is copied to `direction` (3V, may be
 * NULL).  CIPC_ERR_RUNTIME when the target is not met within max_iter, as the reference's
 * solveWithEigenConjugateGradient throws (NewtonLinearSystem.cpp:196-203). */
int cipc_gpu_solve(cipc_gpu_context* ctx, int precision, double rel_tol, int max_iter, double* direction, cipc_solve_stats* stats);
/* The same solve with a choice of preconditioner: 0 block-Jacobi (as cipc_gpu_solve), 1 the
 * 3x3-block factorized sparse approximate inverse M^-1 = G^T G with G on the matrix's lower
 * block pattern (at most 24 blocks per row, the largest kept), 2 the learned factor of the
 * weights set with cipc_gpu_set_preconditioner_weights, 3 no preconditioner but the direct
 * tier: the dense Cholesky factorisation of the system on the GPU (the host direct solvers
 * PARDISO / CHOLMOD, NewtonLinearSystem.cpp:150-174; at most 60,000 unknowns, precision,
 * rel_tol and max_iter ignored, CIPC_ERR_RUNTIME when the system is not positive definite).
 * The iterative choices have no counterpart in the reference, whose iterative solver is Eigen
 * CG with incomplete Cholesky (NewtonLinearSystem.cpp:176-207). */
int cipc_gpu_solve_ex(cipc_gpu_context* ctx, int precision, int preconditioner, double rel_tol, int max_iter, double* direction,
                      cipc_solve_stats* stats);

/* The factor G (M^-1 = G^T G) of the last solve with preconditioner 1 or 2, on the block
 * pattern of the full symmetric Newton matrix: nblocks 3x3 blocks, rowptr (V + 1), col
 * (nblocks, ascending within a row) and values (9 nblocks, row-major blocks; blocks above
 * the diagonal and blocks outside the factor's pattern are zero).  Call with NULL arrays to
 * get nblocks first.  CIPC_ERR_LOGIC when the last solve built no factor.  No counterpart
 * in the reference; it exposes the preconditioner for checking. */
int cipc_gpu_get_solve_factor(cipc_gpu_context* ctx, int64_t* nblocks, int* rowptr, int* col, double* values);

/* Friction::initialize (Friction.cpp:81-156), called at the start of a time step: the
 * lagged friction state (per Self_ActiveSet entry lambda, closest-point coordinates and
 * tangent basis; per Environment_ActiveSet vertex lambda) from the current active set and
 * positions.  Every later assemble and energy adds the friction terms
 * (Friction::addGradient / addHessian / energy, IPCSolver.cpp:309-317,453-455) with the
 * scene's friction coefficient and epsilon^2 = fhat dt^2. */
int cipc_gpu_friction_update(cipc_gpu_context* ctx, double kappa);
/* The lagged state: counts = {self contacts, ground contacts}; lam (n), coord (2n),
 * basis (6n: column 0 xyz, column 1 xyz), env_lam (counts[1]); outputs may be NULL. */
int cipc_gpu_get_friction(cipc_gpu_context* ctx, int64_t counts[2], double* lam, double* coord, double* basis, double* env_lam);

/* Step bound of one Newton iteration for the direction set with cipc_gpu_set_direction:
 * limitStepByGround (FeasibleStep.cpp:8-40), Self_largestFeasibleStepSize
 * (ContactMechanics.cpp:3704-3752), CFL (IPCSolver.cpp:960-974), the swept rebuild
 * SpatialHash::build(mesh, p, alpha) and Self_largestFeasibleStepSize_CCD
 * (ContactMechanics.cpp:3754-3867), merged as IPCSolver.cpp:985-989.
 * out = {alpha, ground bound, partial CCD, CFL, full CCD}. */
int cipc_gpu_step_bound(cipc_gpu_context* ctx, double out[5]);

/* The launch-shape autotuner.  No reference counterpart: the reference's CCD loops are CPU
 * loops (ContactMechanics.cpp:3754-3867) with no launch shape to choose.  It times each
 * candidate split of the two full-CCD walks of the step bound (kernel 0 point-triangle,
 * 1 edge-edge) and of the active set's two walks (kernel 3 point-triangle, 4 edge-edge;
 * candidate c runs 2^c lanes per query, c = 0..4) on the context's current positions and
 * direction (cipc_gpu_set_direction; a batch context times its whole batch), and keeps the
 * fastest for every later call of the context.  Every split returns the same results, bit
 * for bit: a full-CCD walk's result is a minimum over pairs, and the active set ranks its
 * hits by primitive.  entries (may be NULL) receives up to max_entries records (4 at most),
 * n_entries the number of tuned kernels.
 * cipc_gpu_set_tune sets a choice directly (an uploaded scene starts with the largest c,
 * at most 4, that keeps 2^c times the walk's query count (surface vertices or surface
 * edges) within 2^17 lanes); its
 * kernel 2 switches the Newton iteration's single graph launch (cipc_gpu_newton_iteration,
 * cipc_gpu_batch_newton_iteration) off (choice 0: every stage launched on its own and timed)
 * or on (1); by default a batch context uses it and a single-scene context does not. */
typedef struct {
    int kernel, chosen, n_candidates;
    double ms[8]; /* mean ms of each candidate over three timed runs */
} cipc_tune_entry;
int cipc_gpu_autotune(cipc_gpu_context* ctx, cipc_tune_entry* entries, int max_entries, int* n_entries);
int cipc_gpu_set_tune(cipc_gpu_context* ctx, int kernel, int choice);

/* Line-search support: remember the current positions as x0, x = x0 - alpha p on free
 * vertices (stepForward, IPCSolver.cpp:458-473), and isIntersected (IPCSolver.cpp:523-542)
 * with a static LBVH rebuilt on the current positions. */
int cipc_gpu_save_positions(cipc_gpu_context* ctx);
int cipc_gpu_step_forward(cipc_gpu_context* ctx, double alpha);
int cipc_gpu_is_intersected(cipc_gpu_context* ctx, int* intersected);

/* One Newton iteration's non-linear stages on the device-resident state, timed per stage:
 * active set, assembly, step bound, energy.  The assembled system is then the one that
 * cipc_gpu_solve / cipc_gpu_get_csr see, as after cipc_gpu_assemble.  No stage waits on the
 * host (the active set's counts stay in device memory; one synchronisation ends the call),
 * so the stages can be captured once into a HIP graph and replayed by later calls while the
 * stiffness, the buffers' capacities, the lagged friction counts and the tree topologies stay
 * the same (cipc_gpu_set_tune kernel 2). */
int cipc_gpu_newton_iteration(cipc_gpu_context* ctx, double kappa, cipc_iteration_timing* timing);

/* Velocities of the scene's vertices (mesh3D::velocities, 3V; zero after an upload) and the
 * full state: positions, V_prev, inertial target, velocities (3V each, any may be NULL). */
int cipc_gpu_set_velocities(cipc_gpu_context* ctx, const double* velocities);
int cipc_gpu_get_state(cipc_gpu_context* ctx, double* x, double* V_prev, double* xtilde, double* velocities);

/* Scene scalars that a time step reads besides the descriptor (mesh3D::bboxDiagSize2,
 * meanMass, dTol, Newton_Solver_Threshold, apply_gravity, is_quasi_static;
 * SimulationMesh.h:37-129) and the Newton linear-solve settings: the device PCG replaces
 * NewtonLinearSystem::solve (NewtonLinearSystem.cpp:21-56) with precision 0 (FP64) or 1
 * (FP32 inside FP64 refinement), preconditioner 0 (block-Jacobi), 1 (block FSAI) or 2
 * (learned), to solve_rel_tol relative residual.  max_newton is solveBarrierSubproblem's
 * iterationLimit (IPCSolver.cpp:901, 10000). */
typedef struct {
    double bbox_diag2, mean_mass, dtol, newton_threshold;
    int apply_gravity, quasi_static;
    int precision, preconditioner;
    double solve_rel_tol;
    int solve_max_iter, max_newton;
} cipc_step_params;

/* IPCStepStats (IPCSolver.h) of one time step.  Stage times are host wall-clock, as the
 * reference's StageTimer; pcg_iterations sums the Newton solves' iterations.  A solve whose
 * PCG stops above solve_rel_tol is repeated by the direct tier (preconditioner 3) when the
 * system has at most 60,000 unknowns. */
typedef struct {
    int newton_iterations, kappa_iterations;
    int energy_backtracks, intersection_backtracks, max_energy_backtracks_per_newton, newton_steps_over_two_backtracks;
    int64_t pcg_iterations, collisions, matrix_nnz;
    int64_t direct_solves;   /* Newton solves that fell back to the direct tier (PCG above tolerance) */
    int64_t ground_contacts, self_contacts, mollified_contacts;
    double kappa, min_distance2, dirichlet_alpha;
    double alpha_sum, alpha_min, alpha_max;
    double step_ms, assembly_ms, linear_ms, ccd_ms, line_search_ms, post_ms;
} cipc_step_stats;

/* DirichletBoundaryCondition::updateDirection (BoundaryConditions.h) for n vertices: from
 * their current positions x (3n) and rest positions v_rest (3n), the step index, the
 * fraction of the step's prescribed motion and dt, write the directions (3n; the update
 * moves x to x - direction).  Returns 0, or non-zero to abort the step. */
typedef int (*cipc_dirichlet_fn)(void* user, int n, const int* vertices, const double* x, const double* v_rest, int step,
                                 double fraction, double dt, double* directions);
typedef struct {
    int n;                  /* animated Dirichlet vertices (DirichletBoundaryCondition::vertexIndices) */
    const int* vertices;
    int step;               /* the step index passed to the callback */
    cipc_dirichlet_fn fn;
    void* user;
} cipc_dirichlet;

/* One implicit time step of the uploaded scene on the device: solveIPCStep
 * (IPCSolver.cpp:1079-1336) with initKappa and Friction::initialize at the step start, the
 * animated Dirichlet update with its CCD bound and intersection backoff when dirichlet is
 * not NULL (IPCSolver.cpp:1127-1155), the kappa loop of barrier subproblems
 * (solveBarrierSubproblem, IPCSolver.cpp:890-1064: assembly, the device solve, the
 * |p|_inf convergence test, the merged step bound, lineSearch IPCSolver.cpp:560-749 and
 * postLineSearch :811-869), then updateVelocity, V_prev = x and updateInertialTarget
 * (:1196-1204).  kappa is mesh.Kappa, read and updated.  The reference's exceptions
 * (non-finite energy, no non-intersecting step, no energy decrease, the kappa-loop limit)
 * return CIPC_ERR_RUNTIME with its message. */
int cipc_gpu_time_step(cipc_gpu_context* ctx, const cipc_step_params* params, double* kappa, const cipc_dirichlet* dirichlet,
                       cipc_step_stats* stats);

/* Batched scenes.  No single reference interface: the reference runs independent scenes
 * as separate FEMSimulator / IPCSolverContext instances (IPCSolver.h:46-61), each with
 * its own mesh3D, contact sets, barrier stiffness, step and solve.  A batch context holds
 * n_scenes such scenes in one set of device arrays (vertices, tets and surface primitives
 * of scene k after those of scene k-1, indices shifted) and evaluates each one exactly as
 * a context holding that scene alone would: contacts never join two scenes, and step
 * bounds, energies and linear solves are per scene.  The scenes are tet meshes and share
 * dt, hhat, fhat, drag, friction and the ground; the Lame parameters (length_rate,
 * volume_rate) and the masses are per scene.  Scene-agnostic entry points (set_state,
 * get_positions, set_direction, active_set, get_active_set, get_gradient, get_csr,
 * save_positions, step_forward, is_intersected, assemble and friction_update with one
 * stiffness for every scene) act on the whole batch; the per-scene ones below replace
 * energy, step_bound, solve and newton_iteration, which refuse a batch
 * (CIPC_ERR_LOGIC), as do kappa_gradients and close_constraints. */
int cipc_gpu_upload_batch(cipc_gpu_context* ctx, const cipc_scene_desc* const* scenes, int n_scenes);
/* n_scenes; vertex_offsets and tet_offsets (n_scenes + 1 each) may be NULL. */
int cipc_gpu_batch_size(cipc_gpu_context* ctx, int* n_scenes, int* vertex_offsets, int* tet_offsets);
/* As cipc_gpu_assemble with kappa[k] for scene k; collisions (n_scenes, may be NULL) per scene. */
int cipc_gpu_batch_assemble(cipc_gpu_context* ctx, const double* kappa, int64_t* nnz, int* collisions);
/* As cipc_gpu_friction_update with kappa[k] for scene k. */
int cipc_gpu_batch_friction_update(cipc_gpu_context* ctx, const double* kappa);
/* As cipc_gpu_energy / cipc_gpu_energy_terms per scene: energy (n_scenes) and terms
 * (9 n_scenes), either may be NULL. */
int cipc_gpu_batch_energy(cipc_gpu_context* ctx, const double* kappa, double* energy, double* terms);
/* As cipc_gpu_step_bound per scene: out[5k .. 5k+4]. */
int cipc_gpu_batch_step_bound(cipc_gpu_context* ctx, double* out);
/* x = x0 - alpha[k] p on scene k's vertices (a converged scene takes alpha 0). */
int cipc_gpu_batch_step_forward(cipc_gpu_context* ctx, const double* alpha);
/* r[k] = 1 when scene k is intersected (cipc_gpu_is_intersected per scene). */
int cipc_gpu_batch_is_intersected(cipc_gpu_context* ctx, int* r);
/* As cipc_gpu_solve_ex per scene, all scenes in lock step: active (n_scenes, NULL = all)
 * masks converged scenes, whose direction is left zero; stats (n_scenes, may be NULL).
 * CIPC_ERR_RUNTIME names the first active scene that misses the tolerance. */
int cipc_gpu_batch_solve(cipc_gpu_context* ctx, int precision, int preconditioner, double rel_tol, int max_iter, const int* active,
                         double* direction, cipc_solve_stats* stats);
/* As cipc_gpu_newton_iteration with kappa[k] for scene k; alpha and energy (n_scenes
 * each, may be NULL) per scene; timing->alpha and the four bound parts are minima over the
 * scenes, timing->energy and timing->collisions sums. */
int cipc_gpu_batch_newton_iteration(cipc_gpu_context* ctx, const double* kappa, cipc_iteration_timing* timing, double* alpha,
                                    double* energy);

/* Element-level evaluation of n independent inputs (host arrays in, host arrays out);
 * used to check the device kernels element by element.
 *  snh:      F (9n, column-major) -> PK1 (9n), projected Hessian (81n row-major), energy density (n)
 *            (Elasticity.cpp:1144-1170,1337-1479,695-727)
 *  distance: kind 0 PP, 1 PE, 2 PT, 3 EE, 4 EE cross norm; pts 12n -> d (n), grad (12n), Hessian (144n)
 *            (ContactMechanics.cpp:282-316,764-2467)
 *  make_pd:  dim in {6,9,12}, H (dim*dim*n row-major) in place (ContactMechanics.h:27-45)
 *  contact_pd: the projection as the assembly applies it to contact Hessians (functions of
 *            vertex differences only), nverts in {2,3,4}, H (9*nverts^2*n row-major) in place;
 *            with G (3*nverts*n, the blocks' gradients, point-edge / point-triangle /
 *            edge-edge distances only) the structured range-basis path, else the general one
 *            (ContactMechanics.cpp:3489,3516,3539,3561)
 *  accd:     kind 0 point-triangle, 1 edge-edge; pts 12n, dirs 12n -> toc (n) (AdditiveCCD.cpp:221-329)
 *  classify: kind 0 dType_PT, 1 dType_EE; pts 12n -> type (n) (ContactMechanics.cpp:143-280) */
int cipc_gpu_eval_snh(int n, const double* F, double length_rate, double volume_rate, double* P, double* H, double* e);
/*  cloth:    F (6n, 3x2 column-major) -> PK1 (6n), projected 6x6 Hessian (36n row-major), energy
 *            density (n) (Elasticity.cpp:729-771,1173-1205,1481-1544)
 *  hinge:    hinge points (12n: edge start, edge end, opposite vertices), rest angle, weight (n)
 *            -> gradient (12n), makePD'd Hessian (144n row-major), energy (n), before dt^2
 *            (HingeBending.cpp:79-106; Elasticity.cpp:12-150; IPCSolver.cpp:238-267) */
int cipc_gpu_eval_cloth(int n, const double* F, double stretch, double shear, double strain_rate, double* P, double* H, double* e);
int cipc_gpu_eval_hinge(int n, const double* pts, const double* rest, const double* weight, double plate_rigidity,
                        double* g, double* H, double* e);
int cipc_gpu_eval_distance(int n, int kind, const double* pts, double* d, double* grad, double* hess);
int cipc_gpu_eval_make_pd(int n, int dim, double* H);
int cipc_gpu_eval_contact_pd(int n, int nverts, double* H, const double* G);
int cipc_gpu_eval_accd(int n, int kind, const double* pts, const double* dirs, double eta, double thickness, double* toc);
int cipc_gpu_eval_classify(int n, int kind, const double* pts, int* type);

#ifdef __cplusplus
}
#endif
#endif /* CIPC_GPU_H */
