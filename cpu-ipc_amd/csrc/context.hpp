// Device context: one scene, its streams (main, auxiliary, step bound), all device memory.  Each translation unit
// implements one stage of the Newton iteration on it.
#pragma once
#include "common.hpp"

#include <cstring>
#include <limits>

namespace cipc {

// Linear BVH in the Karras layout: internal nodes [0, n-2], leaves [n-1, 2n-2] in Morton
// order.  Traversal is stackless: `left` is the first child of an internal node (or the
// primitive id of a leaf) and `skip` the next node in depth-first order once a subtree
// is done (-1 at the end).
// Power-of-two reduction levels over the sorted leaf boxes (level l node j = leaves
// [j 2^l, (j+1) 2^l)), used to form internal boxes without inter-work-group communication.
struct SegLevels {
    int count[40];
    long long offset[40];
};

// Device LBVH.  The topology (Morton order, Karras splits, skip pointers) and the boxes are
// built separately: a query's result set depends only on the leaf boxes, never on the
// topology, so a topology built from earlier positions is refitted with the current boxes
// and rebuilt only every kMaxRefits fits (or when the state is replaced).
struct LBVH {
    int n = 0;
    int age = 1 << 30;         // fits since the topology was built
    DBuf<Box> prim;            // primitive boxes (input)
    DBuf<Node> nodes;          // 2n-1: internal nodes 0..n-2 (root 0), leaves n-1..2n-2
    DBuf<Node4> wide;          // n-1: four-wide records of the internal nodes
    DBuf<int> range_first;     // n-1
    DBuf<int> range_last;      // n-1
    DBuf<int> split_right;     // n-1: right child of the internal node whose split is s
    SegLevels L{};
    int levels = 0;
    DBuf<Box> seg;             // the reduction levels, level 0 = leaf boxes in leaf order
    DBuf<unsigned> codes, codes_sorted;   // 30-bit Morton codes, by primitive / sorted
    DBuf<int> order;                      // primitive ids in leaf order (+ n scratch)
    DBuf<int> leafpos;                    // leaf position of each primitive
    DBuf<unsigned long long> keys_sorted; // (code << 32 | primitive), sorted and unique
    DBuf<Box> scene_partial;
    DBuf<int2> wslot;                     // 2n-1: the four-wide record slots of each node's box, (record << 2 | k) or -1
    bool swept = false;
};
constexpr int kMaxRefits = 32;

// Scratch of the active-set construction (broadphase.hip).
// Device counts of the active set and of the contact records (Context::dcnt): written by the
// kernels that find them, read by every later kernel from device memory, and by the host only
// at its next synchronisation (settle_active_set), so that no stage waits on the host.
enum : int {
    kDcPT = 0,        // point-triangle candidate hits
    kDcEE,            // edge-edge candidate hits
    kDcSelf,          // Self_ActiveSet entries
    kDcMol,           // Self_EE_ActiveSet entries
    kDcCCD,           // partial-CCD pairs
    kDcEnv,           // Environment_ActiveSet entries
    kDcFlags,         // overflow bits of the active set: 1 PT hits, 2 EE hits, 4 map keys
    kDcContactOver,   // 1 when n_self + n_ee exceeds the contact records' capacity
    kDcMapKeys,       // point-point / point-edge map keys before merging
    kDcA, kDcA2, kDcAK, kDcB, kDcB2, kDcBK,   // per-category totals: the scatters' bases
    kDcWords = 16
};

struct ActiveSetScratch {
    // capacities of the candidate hits (per type) and of the map keys; grown by
    // settle_active_set when the device reports more
    int capPT = 0, capEE = 0, capM = 0;
    // per point query, per edge query and per surface vertex (ground flag), in one array
    DBuf<int> cnt_all, off_all;
    DBuf<int> cntpPT, cntpEE, bufPT, bufEE;
    DBuf<int> candUPT, candUEE, candPT, candEE, ownPT, ownEE, catPT, catEE;
    DBuf<int4> recPT, recEE;
    DBuf<int2> auxEE;
    DBuf<int4> pos_all;   // per-hit exclusive counts of four categories: point hits, then edge hits
    DBuf<unsigned long long> mkeys, mkeys_sorted, ukeys;
    DBuf<int> ucounts, nruns;
};

// Buffers of the Newton linear solve (pcg.hip).
struct SolveBuffers {
    DBuf<unsigned long long> key, key_sorted;
    DBuf<int> src, src_sorted, col, rowptr;   // full symmetric 3x3-block matrix (BSR)
    DBuf<double> val, minv;                   // FP64 blocks and block-Jacobi inverses
    DBuf<float> valf, minvf, xf;              // FP32 copies for the mixed form
    DBuf<double> r, z, p, q, x, res, part, scal, t;
    DBuf<double> G;                           // factor preconditioner at the lower BSR entries
    DBuf<float> Gf;
    DBuf<int> mirror;                         // index of the transposed entry
    DBuf<double> vil, Gil;                    // interleaved copies for the row products
    DBuf<float> vilf, Gilf;
    size_t ne = 0;                            // BSR entry capacity (sentinels included)
    int factor = 0;                           // factor in G: 0 none, 1 block FSAI, 2 learned
    // batch: work-group tables of the row kernels (per vertex partition) and per-scene controls
    std::vector<int> layout_key;
    int layout_ng = 0;
    DBuf<int4> bl;
    DBuf<int> gb, ctl_i;
    DBuf<double> ctl_d;
    DBuf<double> dense;                       // direct tier: the dense lower factor (n x n, column-major)
    DBuf<int> info;
};
struct SolveResult {
    int iterations = 0, refinements = 0;
    double rel_residual = 0.0;
};

struct LearnedState;

struct Context {
    int device = 0;
    LearnedState* learned = nullptr;    // learned preconditioner state (learned.hip)
    void* blas = nullptr;               // rocBLAS handle of the direct tier (created on first use)
    hipStream_t stream = nullptr;
    // Elastic assembly depends on the state only; it runs on `aux` (default priority) while
    // `stream` builds the active set and the contact blocks (fork / join events).
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // friction records and the small contact projections on `aux` beside the barrier blocks
    hipEvent_t fr_fork = nullptr, fr_join = nullptr, cl_done = nullptr;
    // In a Newton iteration the step bound runs on `sbs` (default priority) beside the
    // assembly: it reads the positions, the direction and the active set only.  It refits
    // the trees (the assembly does not read them) and sorts with its own temporary storage.
    hipStream_t sbs = nullptr;
    hipEvent_t as_done = nullptr, sb_done = nullptr;
    // start of an iteration (the step-bound stream forks there) and end of its energy (joins
    // the main stream): a captured iteration forks and joins through these
    hipEvent_t it_start = nullptr, en_done = nullptr;
    DBuf<unsigned char> tmp_sb;
    // The Newton iteration's launches as one HIP graph (context.hip), replayed while nothing
    // they depend on changes (the key: stiffness, precision, capacities, lagged and soft
    // counts, launch shapes, allocations, tree topologies), else captured again.
    struct IterGraph {
        hipGraphExec_t exec = nullptr;
        std::vector<unsigned char> key;
        int mode = -1;   // 1 on, 0 off, -1 the default: on for a batch, off for one scene
        long long captures = 0, replays = 0;
    } graph;
    unsigned long long topo_gen = 0;   // tree topologies built (broadphase.hip)
    // tile choices of the autotuner (ccd.hip): lanes per query of the full-CCD walks (2^bits)
    // and of the active set's walks
    struct Tune {
        int ccd_pt_bits = 3, ccd_ee_bits = 3;
        int as_pt_bits = 3, as_ee_bits = 3;
    } tune;
    Scene sc;
    StaticPattern sp;
    bool have_scene = false;

    // elastic staging: 3x3 blocks and corner forces in their pattern slots; with
    // hessian_fp32 (the mixed form: FP32 elastic Hessian, FP64 barriers, friction and
    // gradient) the blocks are staged as float in the same buffer
    DBuf<double> stagH, stagG;
    bool hessian_fp32 = false;
    DBuf<double> sbv;          // static block values, 9 per block
    DBuf<unsigned short> sbf;  // static block nonzero-contribution masks
    DBuf<double> grad;         // 3V

    // active set (host copies of the device counts, valid after settle_active_set)
    int64_t n_self = 0, n_ee = 0, n_ccd = 0, n_env = 0;
    DBuf<int> dcnt;            // kDcWords device counts
    PinnedBuf<int> pin_dc;
    bool dc_pending = false;   // the device counts are newer than the host copies
    int capS = 0;              // contact (self + mollified) capacity of the contact records
    int records_capS = 0;      // the capacity the current record slots were laid out with
    DBuf<int4> self_set, ee_set;
    DBuf<int2> ee_ie, ccd_set;
    DBuf<int> env_set;

    // lagged friction state (Friction::initialize): per self contact the encoded contact,
    // lambda, closest-point coordinates (2) and tangent basis (6: column 0 xyz, column 1 xyz);
    // per ground contact the vertex and lambda
    int64_t n_fself = 0, n_fenv = 0;
    // soft target-position penalty (SoftBoundaryCondition, BoundaryConditions.h:54-67)
    int n_soft = 0;
    double soft_w = 0.0;
    DBuf<int> soft_idx;
    DBuf<double> soft_t;       // 3 n_soft
    // barrier-stiffness control: initKappa vectors, postLineSearch close-constraint records
    DBuf<double> kgE, kgc, kvals;
    int64_t n_cenv = 0, n_cself = 0;
    DBuf<int> cenv;
    DBuf<double> cenv_val, cself_val;
    DBuf<int4> cself;
    DBuf<int4> fr_self;
    DBuf<double> fr_lam, fr_coord, fr_basis;
    DBuf<int> fr_env;
    DBuf<double> fr_env_lam;
    DBuf<double> fhess;        // per friction contact: 78 upper-triangle Hessian + 12 gradient entries (SoA)
    DBuf<int> fdims;
    DBuf<int4> fverts;

    // contact contributions (fixed slots, invalid key = ~0ull)
    DBuf<unsigned long long> ckey, ckey_sorted;
    DBuf<int> cidx, cidx_sorted;
    DBuf<double> cval;         // 9 per record
    DBuf<int> cdims;           // per contact: local Hessian dimension
    DBuf<int4> cverts;         // per contact: vertices (-1 past the last)
    DBuf<double> chess;        // per contact: 78 upper-triangle Hessian + 12 gradient entries (SoA)
    DBuf<int> gkey, gkey_sorted, gidx, gidx_sorted;
    // vertex buckets that order the records (sparse.hip): counts and cursors, starts, entries
    DBuf<int> bucket_cnt, bucket_start;
    size_t bucket_cnt_clear = 0;        // leading entries of bucket_cnt known to be zero (left so by the last assembly)
    const int* bucket_cnt_ptr = nullptr;  // the buffer they are known zero in
    DBuf<unsigned long long> bucket_entries;
    DBuf<double> gval;         // 3 per record
    DBuf<int> run_start;       // start of each run of equal sorted contact block keys
    DBuf<unsigned long long> ukey;  // unique contact block keys
    DBuf<double> uval;
    DBuf<unsigned short> uflag;
    int64_t n_ublocks = 0;
    DBuf<int> counters;        // small scratch integers

    // merged block pattern and output CSR
    DBuf<int> mptr, mrow, mcol;
    DBuf<int> pcnt, ppos;      // scalar runs of (merged block, column component) pairs
    DBuf<double> mval;
    DBuf<unsigned short> mflag;
    DBuf<int> outer, inner;
    DBuf<double> val, rhs;
    int64_t nnz = 0;
    DBuf<int> coll;            // collision count of the last finalize_csr (device)
    bool nnz_pending = false;  // nnz and the collision count are in pin_nnz once the stream has been synchronised
    // pinned destinations of the asynchronous reads that one final synchronisation collects
    PinnedBuf<int> pin_nnz;
    PinnedBuf<int> pin_as;     // counts of the active-set construction
    PinnedBuf<unsigned char> pin_rb;   // staging of read_back
    PinnedBuf<unsigned long long> pin_slots;
    PinnedBuf<double> pin_energy;
    bool assembled = false;    // a system exists for the linear solve
    SolveBuffers solve;
    int collisions = 0;

    // broad phase
    LBVH face_tree, edge_tree;     // static boxes: the active set, isIntersected
    LBVH sface_tree, sedge_tree;   // swept boxes: the full CCD (their own, so that the step
                                   // bound can run beside the active set)
    ActiveSetScratch as;

    // generic scratch
    DBuf<unsigned char> tmp;
    DBuf<int> iscratch, iscratch2;
    DBuf<double> dscratch, dscratch2;
    // batch: device tet offsets (S + 1) then vertex offsets (S + 1); per-scene steps
    DBuf<int> batch_off;
    DBuf<double> batch_alpha;
    DBuf<unsigned long long> u64scratch, u64scratch2;
    DBuf<int4> i4scratch;
};

// n values from the device to the host through pinned memory (a copy into pageable memory
// is staged by the runtime and costs several times the copy), after the stream's pending work
template <typename T>
void read_back(Context& c, T* host, const T* dev, size_t n) {
    if (n == 0) return;
    c.pin_rb.reserve(n * sizeof(T));
    CIPC_HIP(hipMemcpyAsync(c.pin_rb.p, dev, n * sizeof(T), hipMemcpyDeviceToHost, c.stream));
    CIPC_HIP(hipStreamSynchronize(c.stream));
    std::memcpy(static_cast<void*>(host), c.pin_rb.p, n * sizeof(T));
}

// elastic.hip
void build_static_pattern(Context& c);
// chunk plan of the tet assembly from the host positions (after the scene ranges are set)
void build_tet_chunks(Context& c, const double* hx);
void elastic_assemble(Context& c, hipStream_t st);
void cloth_assemble(Context& c, hipStream_t st);
// contact.hip
// contact_records_prepare sizes the records (host only); friction_records_launch builds the
// lagged friction records on a stream; contact_assemble the barrier records (and, unless
// friction_launched, prepares and launches the friction records itself on `aux`).
void contact_records_prepare(Context& c);
void friction_records_launch(Context& c, hipStream_t st);
void contact_assemble(Context& c, double kappa, bool friction_launched = false);
// sparse.hip
// count_collisions false: the caller launches collision_count_launch and publish_nnz itself
void finalize_csr(Context& c, bool count_collisions = true);
void collision_count_launch(Context& c, hipStream_t st);
void publish_nnz(Context& c, hipStream_t st);
void exclusive_scan_int(Context& c, const int* in, int* out, int n);
// broadphase.hip
void build_tree(Context& c, LBVH& t, int n);
void build_static_trees(Context& c);
// on stream st (the step bound may run beside the assembly); temporary storage of its own
void build_swept_trees(Context& c, const unsigned long long* abits, hipStream_t st);
// Enqueues the active set on the main stream; its counts stay on the device (publish: also
// copied to pinned memory for the host's next synchronisation; a caller that enqueues the
// contact assembly next leaves that to it).
void compute_active_set(Context& c, bool publish = true);
// Reads the device counts (one synchronisation of the main stream) into the host copies.
// When a capacity was exceeded it is grown from the count the device found and the active
// set is built again.  Returns true when it was rebuilt (work enqueued after the first build
// then has to be enqueued again).
bool settle_active_set(Context& c);
// compute_active_set + settle_active_set
void active_set_now(Context& c);
// After a synchronisation that covers the last copy of the device counts (the active set's or
// the contact records'): grows every capacity that the device reports exceeded; true when one
// grew (the work that used it has to be enqueued again).
bool grow_capacities(Context& c);
// Device counts from the host copies (after the sets were uploaded or reset).
void upload_counts(Context& c);
// Contact-record capacity for the next assembly (at least the settled contact count).
void size_contact_records(Context& c);
void sort_u64(Context& c, const unsigned long long* in, unsigned long long* out, int n, int bits);
void sort_u64_pairs(Context& c, const unsigned long long* kin, unsigned long long* kout,
                    const int* vin, int* vout, int n, int bits);
void sort_i32_pairs(Context& c, const int* kin, int* kout, const int* vin, int* vout, int n, int bits);
void sort_u32_pairs(Context& c, const unsigned* kin, unsigned* kout, const int* vin, int* vout, int n, int bits);
void sort_u32_pairs(const unsigned* kin, unsigned* kout, const int* vin, int* vout, int n, int bits, hipStream_t st,
                    DBuf<unsigned char>& tmp);
// sums vertex-keyed 3-vector records (keys sorted, record order within a key) into grad
void add_gradient_records_into(Context& c, int n, const int* k, const int* idx, const double* val, double* grad);
// initKappa's reductions {sum g_c.g_E, sum |g_c|^2}; returns the ground + self constraint count
int64_t kappa_gradients(Context& c, double out[2]);
// postLineSearch's close-constraint check and records; returns 1 when a record got closer
int close_constraints(Context& c, double dtol, int64_t counts[2]);
// the kappa loop's exit test: out = {min, max} over the ground then self constraint values;
// returns their number (0: out is {0, 0})
int64_t constraint_extrema(Context& c, double out[2]);
bool any_intersection(Context& c);
// per scene: 1 when a ground gap is zero or an edge crosses a face
void scene_intersections(Context& c, int* out);
// ccd.hip: step-bound slots per scene (ground, CFL, partial CCD, full CCD, point-triangle
// result, swept step)
constexpr int kSlots = 8;
void step_bound(Context& c, double out[5]);
// step_bound in two halves: the kernels and an asynchronous read of the slots, then (after
// the caller has synchronised the stream) the merge
void step_bound_launch(Context& c, hipStream_t st);
// step_bound_launch in two parts: what needs only the positions and the direction (ground
// bound, CFL, swept trees, full CCD), then the partial CCD over the active set's CCD pairs and
// the read-back of the bounds
// (ground false: the ground bound and CFL slots are left at 1 and 0, so the full CCD starts
// from step 1)
void step_bound_early(Context& c, hipStream_t st, bool ground = true);
void step_bound_partial(Context& c, hipStream_t st);
void step_bound_finish(Context& c, double* out);
void size_active_walks(Context& c, int pt, int ee);
void launch_active_walk(Context& c, bool pt, bool compact, int pb, hipStream_t st = nullptr);
void launch_full_ccd_pt(Context& c, hipStream_t st, int pb);
void launch_full_ccd_ee(Context& c, hipStream_t st, int pb);
// the autotuner: one entry per tuned kernel (0 full CCD point-triangle, 1 edge-edge, 3 the
// active set's point-triangle walk, 4 its edge-edge walk), the mean ms of each candidate
// split and the chosen one; returns the number of entries
constexpr int kTuneSplits = 5;     // 1, 2, 4, 8 or 16 lanes per query
struct TuneEntry {
    int kernel = 0, chosen = 0, n = 0;
    double ms[8] = {};
};
int autotune(Context& c, TuneEntry* out, int max_out);
// Self_largestFeasibleStepSize_CCD alone (FeasibleStep.cpp), seeded with step 1 and
// slackness 0.8, over the current direction: the animated Dirichlet update's bound
double full_ccd_bound(Context& c);
// per scene: out[5 k .. 5 k + 4] as step_bound's five values
void step_bound_scenes(Context& c, double* out);
// energy.hip
void energy_terms(Context& c, double kappa, double out[9]);
// energy_terms in two halves, as step_bound_launch / step_bound_finish
void energy_terms_launch(Context& c, hipStream_t st);
void energy_terms_finish(Context& c, double out[9]);
// the partials and sums of energy_terms_launch into P: kEnergyScratch doubles
constexpr int kEnergyScratch = 11 * 256;
void energy_partials_launch(Context& c, hipStream_t st, double* P);
// energy_terms without allocating: partials in a file-scope device array; the nine terms
// written to out_device (device memory, 9 doubles); synchronises the context's stream
void energy_terms_noalloc(Context& c, double* out_device);
// after a synchronisation of the stream: the CSR nonzero count and the collision count of the
// last finalize_csr
inline void settle_nnz(Context& c) {
    if (c.nnz_pending) { c.nnz = c.pin_nnz.p[0]; c.collisions = c.pin_nnz.p[1]; c.nnz_pending = false; }
}
void cloth_energy_partials(Context& c, double* part_bw, double* part_bend, hipStream_t st);
double total_energy(const Context& c, double kappa, const double t[9]);
// per scene: out[9 k .. 9 k + 8] as energy_terms' nine values
void energy_terms_scenes(Context& c, double* out);

// Deterministic energy reductions: a fixed grid of kEnergyGrid work-groups of kEnergyBlock
// threads, each summing a grid-strided subset in index order, then a shared-memory tree;
// one partial per work-group.
constexpr int kEnergyBlock = 256;
constexpr int kEnergyGrid = 256;
__device__ inline void energy_block_sum(double v, double* part) {
    __shared__ double sh[kEnergyBlock];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kEnergyBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
void friction_update(Context& c, double kappa);
void build_solve_matrix(Context& c, bool fp32);
// preconditioner: 0 block-Jacobi, 1 block FSAI, 2 learned factor
// preconditioner kDirectSolve: the dense Cholesky direct solve (precision ignored)
constexpr int kDirectSolve = 3;
constexpr int kDirectMaxDofs = 60000;   // 28.8 GB of dense factor at the limit
void linear_solve(Context& c, int precision, int preconditioner, double rel_tol, int max_iter, SolveResult& res);
void direct_solve(Context& c, SolveResult& res);
void destroy_direct(Context& c);
void build_fsai(Context& c, bool fp32);
void build_learned(Context& c, bool fp32);
void learned_set_weights(Context& c, const float* w, int64_t n);
int64_t learned_param_count();
void destroy_learned(Context& c);
double learned_loss_grad(Context& c, const float* params, const float* host_probes, int nz, unsigned long long seed, float* grad_out,
                         int objective);
double learned_train_step(Context& c, float* params, float* m1, float* m2, int step, double lr, int nz, unsigned long long seed,
                          int objective);
// Friction records after the contact records; returns the friction energy parts' launch
void friction_records(Context& c, size_t slot_b, size_t slot_g, hipStream_t st);
void friction_energy_partials(Context& c, double* part, double* elem_self, double* elem_ground, hipStream_t st);
void step_forward(Context& c, double alpha);
// per scene: x = x0 - alpha[scene] p
void step_forward_scenes(Context& c, const double* alpha);
// computeGradientAndHessian + NewtonLinearSystem::assemble on the context's stream (synchronised)
void assemble(Context& c, double kappa);

// One implicit time step (solver.hip): scene scalars that the step reads besides the scene
// descriptor, and the linear-solve settings of its Newton iterations.
struct StepParams {
    double bbox_diag2 = 0, mean_mass = 0, dtol = 0, newton_threshold = 1e-2;
    int apply_gravity = 1, quasi_static = 0;
    int precision = 0, preconditioner = 1;
    double solve_tol = 1e-6;
    int solve_max_iter = 10000;
    int max_newton = 10000;
};
// IPCStepStats (IPCSolver.h) fields that the device step produces
struct StepStats {
    int newton_iterations = 0, kappa_iterations = 0;
    int energy_backtracks = 0, intersection_backtracks = 0, max_energy_backtracks_per_newton = 0, newton_over_two_backtracks = 0;
    int64_t pcg_iterations = 0, collisions = 0, matrix_nnz = 0, direct_solves = 0;
    int64_t ground_contacts = 0, self_contacts = 0, mollified_contacts = 0;
    double kappa = 0, min_distance2 = 0, dirichlet_alpha = 1.0;
    double alpha_sum = 0, alpha_min = std::numeric_limits<double>::infinity(), alpha_max = 0;
    double step_ms = 0, assembly_ms = 0, linear_ms = 0, ccd_ms = 0, line_search_ms = 0, post_ms = 0;
};
// The prescribed motion of animated Dirichlet vertices (DirichletBoundaryCondition::
// updateDirection, BoundaryConditions.h), evaluated by the caller on the host
struct DirichletHook {
    int n = 0;
    const int* vertices = nullptr;
    int step = 0;
    cipc_dirichlet_fn fn = nullptr;
    void* user = nullptr;
};
void time_step(Context& c, const StepParams& P, double& kappa, StepStats& st, const DirichletHook* D);

// per scene lock-step solves; active[k] == 0 leaves scene k's direction zero
void linear_solve_scenes(Context& c, int precision, int preconditioner, double rel_tol, int max_iter, const int* active,
                         SolveResult* res);

}  // namespace cipc
