// This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
// If a copy of the MPL was not distributed with this file, You can obtain one at
// https://mozilla.org/MPL/2.0/.  Parts of this file restate logic of CPU-IPC (MPL-2.0);
// each such part cites the source it follows.
//
// One implicit time step on the device context: solveIPCStep (CPU IPC/IPCSolver.cpp:
// 1079-1336) with its barrier subproblems (solveBarrierSubproblem, :890-1064), the line
// search (lineSearch, :560-749), the stiffness control (initKappa / postLineSearch,
// :766-869) and the end-of-step updates (updateVelocity :1066-1077, updateInertialTarget
// :55-70).  Compiled with -ffp-contract=off.
//
// The state never leaves the device: positions, directions, active sets, the assembled
// system and the solve stay in HBM, and the host reads back only the scalars that steer the
// loop (step bounds, energies, the direction's norm, counts), each through pinned memory.
// The control flow, its limits and its exceptions are the reference's, in the same order,
// so that a caller of solveIPCStep finds the same iteration counts, stiffness values and
// failure modes.
#include "context.hpp"

#include <chrono>
#include <cmath>
#include <limits>

namespace cipc {

namespace {

constexpr int kRB = 256;   // reduction work-group size
constexpr int kRG = 256;   // reduction grid (one partial per work-group, summed by one work-group)

// g.p in a fixed order (grid-strided per thread, then a fixed tree) and |p|_inf
__global__ __launch_bounds__(kRB) void search_partials(int n, const double* __restrict__ g, const double* __restrict__ p,
                                                       double* __restrict__ part) {
    double s = 0.0, m = 0.0;
    for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += kRG * kRB) {
        s += g[i] * p[i];
        m = fmax(m, fabs(p[i]));
    }
    __shared__ double ss[kRB], sm[kRB];
    ss[threadIdx.x] = s;
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int k = kRB / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            ss[threadIdx.x] += ss[threadIdx.x + k];
            sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + k]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[blockIdx.x] = ss[0]; part[kRG + blockIdx.x] = sm[0]; }
}
__global__ __launch_bounds__(kRB) void search_final(const double* __restrict__ part, double* __restrict__ out) {
    static_assert(kRG == kRB, "one partial per thread");
    __shared__ double ss[kRB], sm[kRB];
    ss[threadIdx.x] = part[threadIdx.x];
    sm[threadIdx.x] = part[kRG + threadIdx.x];
    __syncthreads();
    for (int k = kRB / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            ss[threadIdx.x] += ss[threadIdx.x + k];
            sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + k]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = ss[0]; out[1] = sm[0]; }
}

// updateVelocity (IPCSolver.cpp:1066-1077), V_prev = x (:1203) and updateInertialTarget
// (:55-70): v = (x - V_prev) / dt (0 when quasi-static), V_prev = x,
// x~ = V_prev + v dt + (0, g dt dt, 0).
__global__ void end_of_step_kernel(int V, const double* __restrict__ x, double* __restrict__ vprev, double* __restrict__ vel,
                                   double* __restrict__ xt, double dt, double gdt2, int quasi_static) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * V) return;
    const double xi = x[i];
    const double v = quasi_static ? 0.0 : (xi - vprev[i]) / dt;
    vel[i] = v;
    vprev[i] = xi;
    const double g = (i % 3 == 1) ? gdt2 : 0.0;
    xt[i] = (xi + v * dt) + g;
}

// x = x0 - p on every vertex (stepForward with updateConstrainedVertices, IPCSolver.cpp:
// 458-473, as the animated Dirichlet update calls it with step 1)
__global__ void step_all_kernel(int n, const double* __restrict__ x0, const double* __restrict__ p, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = x0[i] - 1.0 * p[i];
}

__global__ void zero_kernel(double* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}
__global__ void scatter_dirs(int n, const int* __restrict__ idx, const double* __restrict__ d, double* __restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 3; ++k) p[3 * idx[i] + k] = d[3 * i + k];
}
__global__ void gather_rows(int n, const int* __restrict__ idx, const double* __restrict__ a, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 3; ++k) out[3 * i + k] = a[3 * idx[i] + k];
}

// compute_H_b (ContactMechanics.cpp:786-790), suggestKappa / upperBoundKappa
// (IPCSolver.cpp:751-764)
double barrier_h(double d, double dhat) {
    const double t = d - dhat;
    return (std::log(d / dhat) * -2.0 - t * 4.0 / d) + 1.0 / (d * d) * (t * t);
}
double suggest_kappa(const StepParams& P, double hhat) {
    const double hb = barrier_h(1.0e-16 * P.bbox_diag2, hhat);
    return 1e13 * P.mean_mass / (4.0e-16 * P.bbox_diag2 * hb);
}
void upper_bound_kappa(double& kappa, const StepParams& P, double hhat) {
    const double hb = barrier_h(1.0e-16 * P.bbox_diag2, hhat);
    const double kmax = 100 * 1e13 * P.mean_mass / (4.0e-16 * P.bbox_diag2 * hb);
    if (kappa > kmax) kappa = kmax;
}

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

// One time step's driver; `mesh_kappa` is mesh.Kappa.
struct Stepper {
    Context& c;
    const StepParams& P;
    StepStats& st;
    double& mesh_kappa;
    double thr = 0.0;          // Newton_Solver_Threshold sqrt(bboxDiagSize2) dt

    // {g.p, |p|_inf} of the gradient of the last assembly and the current direction
    void search_scalars(double out[2]) {
        Scene& s = c.sc;
        DBuf<double>& part = c.dscratch2;
        part.resize(2 * (size_t)kRG + 2);
        hipLaunchKernelGGL(search_partials, dim3(kRG), dim3(kRB), 0, c.stream, 3 * s.V, c.grad.p, s.dir.p, part.p);
        hipLaunchKernelGGL(search_final, dim3(1), dim3(kRB), 0, c.stream, part.p, part.p + 2 * kRG);
        CIPC_CHECK_LAUNCH();
        read_back(c, out, part.p + 2 * kRG, 2);
    }

    double energy(double kappa) {
        double t[9];
        energy_terms(c, kappa, t);
        return total_energy(c, kappa, t);
    }

    void active_set() { active_set_now(c); }

    void step(double alpha) {
        step_forward(c, alpha);
        CIPC_HIP(hipStreamSynchronize(c.stream));
    }

    // initKappa (IPCSolver.cpp:766-808) on `kappa`; its final upperBoundKappa applies to
    // mesh.Kappa, which is `kappa` itself when `aliased`
    void init_kappa(double& kappa, bool aliased) {
        double g[2];
        if (!kappa_gradients(c, g)) return;
        double minKappa = -g[0] / g[1];
        if (minKappa > 0.0) kappa = minKappa;
        minKappa = suggest_kappa(P, c.sc.hhat);
        if (kappa < minKappa) kappa = minKappa;
        if (aliased) upper_bound_kappa(kappa, P, c.sc.hhat);
        else upper_bound_kappa(mesh_kappa, P, c.sc.hhat);
    }

    // postLineSearch (IPCSolver.cpp:811-869)
    void post_line_search(double& kappa) {
        if (kappa == 0.0) {
            init_kappa(kappa, false);
            return;
        }
        int64_t counts[2];
        if (close_constraints(c, P.dtol, counts)) {
            kappa *= 2.0;
            upper_bound_kappa(mesh_kappa, P, c.sc.hhat);
        }
    }

    // lineSearch (IPCSolver.cpp:560-749) without the diagnostic branch; returns whether the
    // step was accepted; alpha in/out
    bool line_search(double& stepSize, double kappa, double dnorm) {
        Scene& s = c.sc;
        const double lastEnergyVal = energy(kappa);
        double sc[2];
        search_scalars(sc);
        const double gradientDotDirection = sc[0];
        if (!std::isfinite(lastEnergyVal) || !std::isfinite(gradientDotDirection))
            throw Error(CIPC_ERR_RUNTIME, "line search received a non-finite energy or direction");
        const double predictedDecrease = std::max(0.0, gradientDotDirection);
        constexpr double armijoCoefficient = 0.0;
        const double eps = std::numeric_limits<double>::epsilon();
        CIPC_HIP(hipMemcpyAsync(s.x0.p, s.x.p, 3 * sizeof(double) * s.V, hipMemcpyDeviceToDevice, c.stream));
        step(stepSize);
        int numOfIntersect = 0;
        constexpr int maxIntersectionBacktracks = 64;
        while (any_intersection(c)) {
            if (numOfIntersect >= maxIntersectionBacktracks || stepSize <= eps)
                throw Error(CIPC_ERR_RUNTIME, "line search could not find a non-intersecting step");
            ++numOfIntersect;
            stepSize /= 2.0;
            step(stepSize);
        }
        active_set();
        double testingE = energy(kappa);
        int numOfLineSearch = 0;
        const double LFStepSize = stepSize;
        auto acceptable = [&](double e, double a) {
            if (!std::isfinite(e)) return false;
            return e < lastEnergyVal - armijoCoefficient * a * predictedDecrease;
        };
        constexpr int maxEnergyBacktracks = 64;
        while (!acceptable(testingE, stepSize)) {
            if (numOfLineSearch >= maxEnergyBacktracks || stepSize <= eps * std::max(1.0, LFStepSize)) {
                if (dnorm < thr) {
                    CIPC_HIP(hipMemcpyAsync(s.x.p, s.x0.p, 3 * sizeof(double) * s.V, hipMemcpyDeviceToDevice, c.stream));
                    active_set();
                    tally(numOfLineSearch, numOfIntersect);
                    return false;
                }
                throw Error(CIPC_ERR_RUNTIME, "line search failed to find a strict energy decrease");
            }
            stepSize *= 0.5;
            ++numOfLineSearch;
            step(stepSize);
            active_set();
            testingE = energy(kappa);
        }
        if (stepSize < LFStepSize) {
            bool needRecomputeCS = false;
            int safeguardIterations = 0;
            while (any_intersection(c)) {
                if (safeguardIterations++ >= maxIntersectionBacktracks || stepSize <= eps)
                    throw Error(CIPC_ERR_RUNTIME, "backtracked step remains intersecting");
                stepSize /= 2.0;
                ++numOfIntersect;
                step(stepSize);
                needRecomputeCS = true;
            }
            if (needRecomputeCS) active_set();
        }
        tally(numOfLineSearch, numOfIntersect);
        st.alpha_sum += stepSize;
        st.alpha_min = std::min(st.alpha_min, stepSize);
        st.alpha_max = std::max(st.alpha_max, stepSize);
        return true;
    }
    void tally(int nE, int nI) {
        st.energy_backtracks += nE;
        st.intersection_backtracks += nI;
        st.max_energy_backtracks_per_newton = std::max(st.max_energy_backtracks_per_newton, nE);
        if (nE > 2) ++st.newton_over_two_backtracks;
    }

    // solveBarrierSubproblem (IPCSolver.cpp:890-1064) with a by-value kappa
    int barrier_subproblem(double kappa) {
        Scene& s = c.sc;
        int k = 0;
        double beta = 1.0;
        const int semi_implicit_kmin = 6;
        for (; k < P.max_newton; ++k) {
            auto t0 = Clock::now();
            assemble(c, kappa);
            st.collisions += c.collisions;
            st.matrix_nnz = c.nnz;
            const double a_ms = ms_since(t0);
            t0 = Clock::now();
            SolveResult r;
            linear_solve(c, P.precision, P.preconditioner, P.solve_tol, P.solve_max_iter, r);
            st.pcg_iterations += r.iterations;
            if (!(r.rel_residual <= P.solve_tol) && P.preconditioner != kDirectSolve && 3 * s.V <= kDirectMaxDofs) {
                // the direct tier takes over where conjugate gradients stall
                const int it = r.iterations;
                direct_solve(c, r);
                r.iterations = it;
                ++st.direct_solves;
            }
            if (!(r.rel_residual <= P.solve_tol))
                throw Error(CIPC_ERR_RUNTIME, "the Newton solve stopped after " + std::to_string(r.iterations) +
                                                  " iterations at relative residual " + std::to_string(r.rel_residual));
            double sc[2];
            search_scalars(sc);
            const double l_ms = ms_since(t0);
            const double directionNorm = sc[1];
            if (directionNorm < thr) {
                st.assembly_ms += a_ms;
                st.linear_ms += l_ms;
                break;
            }
            t0 = Clock::now();
            double b[5];
            step_bound(c, b);
            double alpha = b[0];
            const double c_ms = ms_since(t0);
            t0 = Clock::now();
            const bool accepted = line_search(alpha, kappa, directionNorm);
            const double ls_ms = ms_since(t0);
            if (!accepted) {
                st.assembly_ms += a_ms; st.linear_ms += l_ms; st.ccd_ms += c_ms; st.line_search_ms += ls_ms;
                break;
            }
            t0 = Clock::now();
            post_line_search(kappa);
            st.assembly_ms += a_ms; st.linear_ms += l_ms; st.ccd_ms += c_ms; st.line_search_ms += ls_ms;
            st.post_ms += ms_since(t0);
            if (k + 1 >= semi_implicit_kmin) beta = (1 - alpha) * beta;
            if (beta <= thr) break;
        }
        (void)s;
        st.newton_iterations += k;
        return k;
    }

    // The animated Dirichlet update of solveIPCStep (IPCSolver.cpp:1127-1155): the
    // prescribed directions, cut by the full CCD with slackness 0.8, then halved until no
    // edge crosses a face
    void dirichlet_update(const DirichletHook& D) {
        Scene& s = c.sc;
        const int n = D.n;
        const int V = s.V;
        DBuf<int>& idx = c.iscratch;
        DBuf<double>& dd = c.dscratch;
        idx.resize(std::max(n, 1));
        dd.resize(6 * (size_t)std::max(n, 1));
        CIPC_HIP(hipMemcpyAsync(idx.p, D.vertices, sizeof(int) * n, hipMemcpyHostToDevice, c.stream));
        std::vector<double> hx(3 * (size_t)n), hr(3 * (size_t)n), hd(3 * (size_t)n);
        hipLaunchKernelGGL(gather_rows, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, idx.p, s.vrest.p, dd.p + 3 * (size_t)n);
        CIPC_CHECK_LAUNCH();
        read_back(c, hr.data(), dd.p + 3 * (size_t)n, 3 * (size_t)n);
        auto directions = [&](double fraction) {
            hipLaunchKernelGGL(gather_rows, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, idx.p, s.x.p, dd.p);
            CIPC_CHECK_LAUNCH();
            read_back(c, hx.data(), dd.p, 3 * (size_t)n);
            if (D.fn(D.user, n, D.vertices, hx.data(), hr.data(), D.step, fraction, s.dt, hd.data()) != 0)
                throw Error(CIPC_ERR_RUNTIME, "Dirichlet boundary update callback failed");
            for (double v : hd)
                if (!std::isfinite(v)) throw Error(CIPC_ERR_RUNTIME, "Dirichlet boundary update produced a non-finite direction");
            CIPC_HIP(hipMemcpyAsync(dd.p, hd.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, c.stream));
            hipLaunchKernelGGL(zero_kernel, dim3(blocks_for(3 * V, 256)), dim3(256), 0, c.stream, s.dir.p, 3 * V);
            hipLaunchKernelGGL(scatter_dirs, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, idx.p, dd.p, s.dir.p);
            CIPC_CHECK_LAUNCH();
            CIPC_HIP(hipStreamSynchronize(c.stream));
        };
        double new_alpha = 1.0;
        directions(new_alpha);
        new_alpha = full_ccd_bound(c);
        directions(new_alpha);
        CIPC_HIP(hipMemcpyAsync(s.x0.p, s.x.p, 3 * sizeof(double) * V, hipMemcpyDeviceToDevice, c.stream));
        auto move = [&] {
            hipLaunchKernelGGL(step_all_kernel, dim3(blocks_for(3 * V, 256)), dim3(256), 0, c.stream, 3 * V, s.x0.p, s.dir.p, s.x.p);
            CIPC_CHECK_LAUNCH();
            CIPC_HIP(hipStreamSynchronize(c.stream));
        };
        move();
        int numOfIntersect = 0;
        while (any_intersection(c)) {
            if (numOfIntersect++ >= 64 || new_alpha <= std::numeric_limits<double>::epsilon())
                throw Error(CIPC_ERR_RUNTIME, "animated boundary update remains intersecting");
            new_alpha /= 2.0;
            CIPC_HIP(hipMemcpyAsync(s.x.p, s.x0.p, 3 * sizeof(double) * V, hipMemcpyDeviceToDevice, c.stream));
            directions(new_alpha);
            move();
        }
        st.dirichlet_alpha = new_alpha;
        active_set();
    }

    void run(const DirichletHook* D) {
        Scene& s = c.sc;
        const auto begin = Clock::now();
        thr = P.newton_threshold * std::sqrt(P.bbox_diag2) * s.dt;
        upper_bound_kappa(mesh_kappa, P, s.hhat);
        if (mesh_kappa < 1e-16) mesh_kappa = suggest_kappa(P, s.hhat);
        // the step starts from the active sets of its positions (the reference keeps those of
        // the previous step's last line search, built at the same positions)
        active_set();
        init_kappa(mesh_kappa, true);
        friction_update(c, mesh_kappa);
        CIPC_HIP(hipStreamSynchronize(c.stream));
        if (D && D->n > 0) dirichlet_update(*D);
        constexpr int maxKappaIterations = 64;
        for (; st.kappa_iterations < maxKappaIterations; ++st.kappa_iterations) {
            c.n_cenv = c.n_cself = 0;    // close-constraint records cleared
            barrier_subproblem(mesh_kappa);
            double ext[2];
            if (constraint_extrema(c, ext)) {
                st.min_distance2 = ext[0];
                if (ext[0] < P.dtol) break;
                else if (ext[1] < s.hhat) break;
            } else {
                break;
            }
        }
        if (st.kappa_iterations == maxKappaIterations) throw Error(CIPC_ERR_RUNTIME, "IPC kappa loop exceeded its iteration limit");
        ++st.kappa_iterations;
        const double g = P.apply_gravity ? -9.8 : 0.0;
        hipLaunchKernelGGL(end_of_step_kernel, dim3(blocks_for(3 * s.V, 256)), dim3(256), 0, c.stream, s.V, s.x.p, s.vprev.p, s.vel.p,
                           s.xt.p, s.dt, g * s.dt * s.dt, P.quasi_static);
        CIPC_CHECK_LAUNCH();
        CIPC_HIP(hipStreamSynchronize(c.stream));
        st.kappa = mesh_kappa;
        st.ground_contacts = c.n_env;
        st.self_contacts = c.n_self;
        st.mollified_contacts = c.n_ee;
        st.step_ms = ms_since(begin);
    }
};

}  // namespace

void time_step(Context& c, const StepParams& P, double& kappa, StepStats& st, const DirichletHook* D) {
    if (c.sc.S != 1) throw Error(CIPC_ERR_LOGIC, "a batch steps its scenes through the per-scene entry points");
    st = StepStats{};
    Stepper s{c, P, st, kappa};
    s.run(D);
}

}  // namespace cipc
