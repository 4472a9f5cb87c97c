// Headless driver of the device path: the reference's apps/cipc_headless.cpp interface
// (scene, steps, output directory, checkpoints; metrics.csv and the RESULT line) over the
// C ABI, with every time step run on the GPU by cipc_gpu_time_step.
//
// Differences from the reference's flags: --linear-solver selects the device solve
// (pcg-fp64 or pcg-mixed) instead of CHOLMOD / PARDISO / Eigen CG; --preconditioner
// jacobi|fsai; --assets / --settings locate the reference's Assets/ tree (the reference
// compiles its path in); --broad-phase accepts lbvh only (the device broad phase);
// --disable-barrier and --diagnose-line-search are not provided.
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cipc_gpu.h"
#include "../../include/cipc_host.h"

namespace {

struct Options {
    std::string scene = "twisting-mat-soft";
    int steps = 1;
    bool resume = false, write_output = true, write_checkpoints = false, verbose = false;
    std::string output, assets, settings;
    int precision = 0, preconditioner = 1, device = 0;
    double solve_tol = 1e-6;
};

void usage(const char* exe) {
    std::cout << "Usage: " << exe << " [options]\n"
              << "  --scene cloth-bunny|twisting-mat|twisting-mat-soft|bunny2   (default twisting-mat-soft)\n"
              << "  --steps N\n"
              << "  --output DIRECTORY\n"
              << "  --assets DIRECTORY   the reference's Assets/ tree (default $CIPC_ASSETS_DIR)\n"
              << "  --settings FILE      (default ASSETS/scene/parameterSetting.txt)\n"
              << "  --broad-phase lbvh\n"
              << "  --linear-solver pcg-fp64|pcg-mixed   (default pcg-fp64)\n"
              << "  --preconditioner jacobi|fsai         (default fsai)\n"
              << "  --solve-tol X        relative residual of each Newton solve (default 1e-6)\n"
              << "  --device N\n"
              << "  --resume\n"
              << "  --write-checkpoints\n"
              << "  --no-output\n"
              << "  --verbose\n"
              << "  --help\n";
}

int non_negative(const std::string& v, const char* opt) {
    size_t used = 0;
    const int r = std::stoi(v, &used);
    if (used != v.size() || r < 0) throw std::invalid_argument(std::string(opt) + " expects a non-negative integer");
    return r;
}

Options parse(int argc, char** argv) {
    Options o;
    const char* env = std::getenv("CIPC_ASSETS_DIR");
    if (env) o.assets = env;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&](const char* opt) -> std::string {
            if (++i >= argc) throw std::invalid_argument(std::string(opt) + " requires a value");
            return argv[i];
        };
        if (a == "--scene") o.scene = value("--scene");
        else if (a == "--steps") o.steps = non_negative(value("--steps"), "--steps");
        else if (a == "--output") o.output = value("--output");
        else if (a == "--assets") o.assets = value("--assets");
        else if (a == "--settings") o.settings = value("--settings");
        else if (a == "--broad-phase") {
            const std::string v = value("--broad-phase");
            if (v != "lbvh") throw std::invalid_argument("unknown or unavailable broad phase: " + v + " (the device broad phase is lbvh)");
        } else if (a == "--linear-solver") {
            const std::string v = value("--linear-solver");
            if (v == "pcg-fp64") o.precision = 0;
            else if (v == "pcg-mixed") o.precision = 1;
            else throw std::invalid_argument("unknown linear solver: " + v);
        } else if (a == "--preconditioner") {
            const std::string v = value("--preconditioner");
            if (v == "jacobi") o.preconditioner = 0;
            else if (v == "fsai") o.preconditioner = 1;
            else throw std::invalid_argument("unknown preconditioner: " + v);
        } else if (a == "--solve-tol") o.solve_tol = std::stod(value("--solve-tol"));
        else if (a == "--device") o.device = non_negative(value("--device"), "--device");
        else if (a == "--resume") o.resume = true;
        else if (a == "--write-checkpoints") o.write_checkpoints = true;
        else if (a == "--no-output") o.write_output = false;
        else if (a == "--verbose") o.verbose = true;
        else if (a == "--help" || a == "-h") { usage(argv[0]); std::exit(0); }
        else throw std::invalid_argument("unknown option: " + a);
    }
    if (o.assets.empty()) throw std::invalid_argument("no asset directory: pass --assets or set CIPC_ASSETS_DIR");
    if (o.assets.back() != '/') o.assets += '/';
    if (o.settings.empty()) o.settings = o.assets + "scene/parameterSetting.txt";
    if (o.output.empty()) o.output = "cipc_output/headless";
    return o;
}

void check(int rc) {
    if (rc != CIPC_OK) throw std::runtime_error(cipc_gpu_last_error());
}

void make_dirs(const std::string& path) {
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty()) mkdir(cur.c_str(), 0755);
        }
        if (i < path.size()) cur += path[i];
    }
}

// Cumulative counters of IPCSolverContext (IPCSolver.h) kept across steps and checkpoints.
struct Totals {
    double stage_ms[5] = {0, 0, 0, 0, 0};
    double step_ms = 0;
    long long newton = 0, collisions = 0;
    int step_index = 0;
};

}  // namespace

int main(int argc, char** argv) {
    try {
        const Options opt = parse(argc, argv);
        const std::string temp = opt.output + "/temp";
        if (opt.write_output || opt.write_checkpoints) make_dirs(temp);

        cipc_host_mesh* mesh = nullptr;
        check(cipc_host_scene_build(opt.scene.c_str(), opt.assets.c_str(), opt.settings.c_str(), &mesh));
        int64_t cnt[9];
        check(cipc_host_counts(mesh, cnt));
        if (opt.verbose) std::printf("tets num: %lld\nverts num: %lld\ntriangles num: %lld\n", (long long)cnt[1], (long long)cnt[0], (long long)cnt[2]);

        Totals tot;
        double kappa = 0.0;
        if (opt.resume) {
            // IPCSolver.cpp:1085-1100 and SimulationMesh.cpp:470-505: counters and kappa, then positions
            std::ifstream t(temp + "/timeCost.txt");
            const bool state = cipc_host_load_checkpoint(mesh, temp.c_str()) == CIPC_OK;
            if (t && state) {
                t >> tot.stage_ms[0] >> tot.stage_ms[1] >> tot.stage_ms[2] >> tot.stage_ms[3] >> tot.stage_ms[4] >> tot.step_ms >>
                    tot.newton >> tot.step_index >> kappa;
            }
            if (opt.verbose) std::printf(state ? "load temp data\n" : "no temp data\n");
        }
        cipc_scene_desc desc;
        cipc_step_params params;
        check(cipc_host_scene_desc(mesh, &desc, &params));
        params.precision = opt.precision;
        params.preconditioner = opt.preconditioner;
        params.solve_rel_tol = opt.solve_tol;
        int motion = 0;
        const int* dirichlet = nullptr;
        const int* soft = nullptr;
        double soft_weight = 0.0;
        check(cipc_host_boundary(mesh, &motion, &dirichlet, &soft, &soft_weight));

        cipc_gpu_context* ctx = nullptr;
        check(cipc_gpu_context_create(opt.device, &ctx));
        check(cipc_gpu_upload_scene(ctx, &desc));
        const int V = desc.V;
        std::vector<double> x(3 * (size_t)V), xt(3 * (size_t)V);
        const int nsoft = (int)cnt[8];
        std::vector<double> sx(3 * (size_t)nsoft), sr(3 * (size_t)nsoft), st(3 * (size_t)nsoft);
        for (int i = 0; i < nsoft; ++i)
            for (int k = 0; k < 3; ++k) sr[3 * i + k] = desc.v_rest[3 * (size_t)soft[i] + k];

        std::ofstream metrics;
        if (opt.write_output) {
            const bool append = opt.resume && tot.step_index > 0;
            metrics.open(opt.output + "/metrics.csv", append ? std::ios::app : std::ios::trunc);
            if (!append)
                metrics << "frame,step_ms,assembly_ms,linear_ms,ccd_ms,line_search_ms,post_ms,newton,kappa_loops,backtracks,"
                           "energy_backtracks,intersection_backtracks,max_energy_backtracks_per_newton,"
                           "newton_steps_over_two_backtracks,mean_alpha,min_alpha,max_alpha,collisions,kappa,min_distance2,"
                           "ground_contacts,self_contacts,mollified_contacts,matrix_nnz,symbolic_analyses,numeric_factorizations,"
                           "pardiso_analysis_ms,pardiso_factorization_ms,pardiso_solve_ms,linear_solver_threads,factor_nnz,"
                           "pcg_iterations\n";
        }

        cipc_step_stats last{};
        int done = 0;
        for (int s = 0; s < opt.steps; ++s) {
            const int step_id = tot.step_index;
            if (motion == 1 && nsoft > 0) {
                // updateSoftTargets (BoundaryConditions.cpp:112-135) at the step start
                check(cipc_gpu_get_positions(ctx, x.data()));
                for (int i = 0; i < nsoft; ++i)
                    for (int k = 0; k < 3; ++k) sx[3 * i + k] = x[3 * (size_t)soft[i] + k];
                check(cipc_host_twisting_targets(nsoft, sx.data(), sr.data(), desc.dt, st.data()));
                check(cipc_gpu_set_soft_targets(ctx, nsoft, soft, st.data(), soft_weight));
            }
            cipc_dirichlet dir{};
            const bool animated = motion == 1 && cnt[7] > 0;
            if (animated) {
                dir.n = (int)cnt[7];
                dir.vertices = dirichlet;
                dir.step = step_id;
                dir.fn = cipc_host_twisting_directions;
            }
            check(cipc_gpu_time_step(ctx, &params, &kappa, animated ? &dir : nullptr, &last));
            ++done;
            ++tot.step_index;
            tot.stage_ms[0] += last.assembly_ms; tot.stage_ms[1] += last.linear_ms; tot.stage_ms[2] += last.ccd_ms;
            tot.stage_ms[3] += last.line_search_ms; tot.stage_ms[4] += last.post_ms;
            tot.step_ms += last.step_ms;
            tot.newton += last.newton_iterations;
            tot.collisions += last.collisions;
            if (opt.verbose)
                std::cout << "finished step " << step_id << " in " << last.step_ms << " ms, Newton=" << last.newton_iterations
                          << ", kappa loops=" << last.kappa_iterations << ", nnz=" << last.matrix_nnz << std::endl;
            if (opt.write_output) {
                std::ofstream tc(opt.output + "/timeCost.txt");
                const double avg = tot.newton > 0 ? (double)tot.collisions / tot.newton : 0.0;
                for (int k = 0; k < 5; ++k) tc << "time" << k << ": " << tot.stage_ms[k] / 1000.0 << "\n";
                tc << "timeAll: " << tot.step_ms / 1000.0 << "\ntotal iter: " << tot.newton << "\nframes: " << tot.step_index
                   << "\ntotalCollisionNum: " << tot.collisions << "\naverageCollision: " << avg << "\n";
                const double mean_alpha = last.newton_iterations > 0 ? last.alpha_sum / last.newton_iterations : 0.0;
                metrics << std::setprecision(17) << step_id << ',' << last.step_ms << ',' << last.assembly_ms << ',' << last.linear_ms << ','
                        << last.ccd_ms << ',' << last.line_search_ms << ',' << last.post_ms << ',' << last.newton_iterations << ','
                        << last.kappa_iterations << ',' << last.energy_backtracks + last.intersection_backtracks << ','
                        << last.energy_backtracks << ',' << last.intersection_backtracks << ',' << last.max_energy_backtracks_per_newton
                        << ',' << last.newton_steps_over_two_backtracks << ',' << mean_alpha << ',' << last.alpha_min << ','
                        << last.alpha_max << ',' << last.collisions << ',' << last.kappa << ',' << last.min_distance2 << ','
                        << last.ground_contacts << ',' << last.self_contacts << ',' << last.mollified_contacts << ','
                        << last.matrix_nnz << ",0,0,0,0,0,0,0," << last.pcg_iterations << '\n';
                metrics.flush();
            }
            if (opt.write_checkpoints && tot.step_index % 10 == 0) {
                check(cipc_gpu_get_state(ctx, x.data(), nullptr, xt.data(), nullptr));
                check(cipc_host_set_state(mesh, x.data(), xt.data()));
                std::ofstream tc(temp + "/timeCost.txt");
                tc << std::setprecision(17);
                for (int k = 0; k < 5; ++k) tc << tot.stage_ms[k] << "\n";
                tc << tot.step_ms << "\n" << tot.newton << "\n" << tot.step_index << "\n" << kappa << "\n";
                if (cipc_host_save_checkpoint(mesh, temp.c_str()) != CIPC_OK)
                    std::cerr << "Warning: failed to save simulation checkpoint" << std::endl;
            }
        }

        // printFinalSummary (apps/cipc_headless.cpp:150-178)
        check(cipc_gpu_get_positions(ctx, x.data()));
        double sum[3] = {0, 0, 0}, sq = 0.0;
        for (int v = 0; v < V; ++v) {
            for (int k = 0; k < 3; ++k) sum[k] += x[3 * (size_t)v + k];
            sq += x[3 * (size_t)v] * x[3 * (size_t)v] + x[3 * (size_t)v + 1] * x[3 * (size_t)v + 1] + x[3 * (size_t)v + 2] * x[3 * (size_t)v + 2];
        }
        std::cout << std::setprecision(17) << "RESULT steps=" << done << " vertices=" << V << " sum_x=" << sum[0] << " sum_y=" << sum[1]
                  << " sum_z=" << sum[2] << " squared_norm_sum=" << sq;
        if (done > 0)
            std::cout << " last_step_ms=" << last.step_ms << " last_newton=" << last.newton_iterations
                      << " last_min_distance2=" << last.min_distance2 << " last_matrix_nnz=" << last.matrix_nnz;
        std::cout << std::endl;
        cipc_gpu_context_destroy(ctx);
        cipc_host_mesh_destroy(mesh);
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "cipc_headless: " << e.what() << std::endl;
        return 1;
    }
}
