"""Benchmark: one Newton iteration's non-linear stages on one MI355X.

Workload: four 20^3-cell lattice cubes (37,044 vertices / 160,000 tets, cell 0.1 m),
seeded yaw rotations, stacked with 1e-3 m gaps on the ground (every interface in
contact), E=1e5 Pa, nu=0.4, rho=1000, dt=0.01 s, relative dHat 1e-3, friction mu=0.3,
FP64.  A step is one Newton iteration of the reference's barrier subproblem without the
linear solve (CPU IPC/IPCSolver.cpp:912-998), on a frozen state with inputs resident in HBM:
  1. LBVH build + active set      (buildCollisionSets)
  2. gradient + projected Hessian with lagged friction -> the reference's lower CSR
     (computeGradientAndHessian + assemble)
  3. step bound: ground + partial CCD + CFL + swept LBVH + full CCD
  4. one line-search energy evaluation (with friction)
The frozen state: x as generated, each cube given a seeded velocity in [-0.5, 0.5]^3 m/s,
positions at the step start V_prev = x - v dt, x~ from velocity and gravity, direction
p = x - x~ perturbed per vertex by +-5%; the lagged friction state (Friction::initialize)
is taken once at that state, as at the start of a time step.

Prints one JSON line.  --gpus above 1 exits: every run here uses one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cpu-ipc_amd"))

METRIC = "ms per Newton iter + assembly HBM GB/s; learned-PCG vs plain at 1e-6 rel residual"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def elastic_algorithmic_bytes(scene):
    """Minimal HBM traffic of the inertia + tet assembly per launch: read x (24 B/vertex),
    per tet Dm^-1 (72 B) + volume (8 B) + indices (16 B), write the gradient (24 B/vertex)
    and the elastic lower-CSR values (8 B per nonzero: 6 per vertex + 9 per coupled
    vertex pair)."""
    import numpy as np
    t = scene.tets.astype(np.int64)
    pairs = np.concatenate([t[:, [a, b]] for a in range(4) for b in range(a + 1, 4)])
    pairs.sort(axis=1)
    E = len(np.unique(pairs[:, 0] * scene.V + pairs[:, 1]))
    nnz = 6 * scene.V + 9 * E
    return 24 * scene.V + scene.T * (72 + 8 + 16) + 24 * scene.V + 8 * nnz, nnz


def stage_rooflines(scene, a, nnz, nnz_el, mean):
    """HBM rate of each stage of the iteration against its minimal traffic (each input read
    once, each output written once), from the stage times of the HIP events.  a: the active
    set's arrays (their bytes are the active set's output and the contact stages' input).
    Surface positions 24 B per surface vertex, faces 12 B, edges 8 B; the lagged friction
    state 80 B per contact (basis, coordinates, multiplier); the CSR 4 B per row start,
    12 B per nonzero; gradient and right-hand side 24 B per vertex each."""
    Vs, F, Es, V, T = len(scene.surf_verts), len(scene.faces), len(scene.surf_edges), scene.V, scene.T
    topo = 24 * Vs + 12 * F + 8 * Es
    sets = sum(int(v.nbytes) for v in a.values())
    contact_sets = sets - int(a["ccd_set"].nbytes)
    contacts = len(a["self_set"]) + len(a["ee_set"]) + len(a["env_set"])
    b = {
        "broadphase_ms": ("active set: tree refit and walks, candidate filters, sets", topo + sets),
        "assembly_ms": ("contact blocks, friction, merge into the CSR with the elastic values",
                        24 * Vs + contact_sets + 80 * contacts + 8 * nnz_el + 24 * V
                        + 4 * (3 * V + 1) + 12 * nnz + 48 * V),
        "ccd_ms": ("step bound: ground, CFL, swept trees, full and partial CCD",
                   topo + 24 * Vs + int(a["ccd_set"].nbytes) + 8),
        "energy_ms": ("line-search energy: inertia, elasticity, contacts, friction",
                      56 * V + 96 * T + 24 * Vs + contact_sets + 80 * contacts + 8),
    }
    out = {}
    for k, (what, nbytes) in b.items():
        gbs = nbytes / (mean[k] * 1e-3) / 1e9
        out[k[:-3]] = {"what": what, "algorithmic_bytes": int(nbytes), "ms": round(mean[k], 4),
                       "achieved_GBs": round(gbs, 2), "frac": round(gbs / HBM_PEAK_GBS, 5)}
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, xt, vprev, p, kappa, runs=5, thread_counts=(16, 1)):
    """The four stages of the timed step on the oracle (the CPU restatement of the reference
    path, OpenMP where the reference runs TBB loops), on the host cores of this machine:
    the median of `runs` warm runs per stage, at 16 threads (the reference's default thread
    count, LinearSolverOptions.h:18, and the CPU share of one GPU on this machine) and at 1."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    out = {}
    for threads in thread_counts:
        m = O.OracleMesh(scene, xtilde=xt, threads=threads)
        m.set_state(V_prev=vprev)
        m.active_set()
        m.friction_update(kappa)
        t = {"active_set": [], "assembly": [], "step_bound": [], "energy": []}
        for _ in range(runs):
            for name, fn in (("active_set", m.active_set), ("assembly", lambda: m.assemble(kappa)),
                             ("step_bound", lambda: m.step_bound(p)), ("energy", lambda: m.energy(kappa))):
                t0 = time.perf_counter()
                fn()
                t[name].append((time.perf_counter() - t0) * 1e3)
        m.close()
        med = {k: round(float(np.median(v)), 3) for k, v in t.items()}
        out[threads] = {"ms_per_iter": round(sum(med.values()), 3), "stages_ms": med}
    main = out[thread_counts[0]]
    return {"value": main["ms_per_iter"], "unit": "ms/iter", "cores": thread_counts[0], "kind": "port",
            "sample": f"the timed step's four stages (active set, assembly, step bound, energy) on the same frozen state, "
                      f"median of {runs} warm runs per stage, on the oracle (CPU restatement of the reference path, OpenMP)",
            "stages_ms": main["stages_ms"], "cpu_model": cpu_model(),
            "other_thread_counts": {str(k): v for k, v in out.items() if k != thread_counts[0]}}


def time_step_leg(scene, steps=3, precision=1, preconditioner=1):
    """Whole implicit time steps of the same scene on the device (cipc_gpu_time_step:
    solveIPCStep with its Newton loop, the linear solve, step bounds, line search and
    stiffness control, IPCSolver.cpp:890-1336), from the stacked state with the cubes'
    seeded velocities: wall time per step and per Newton iteration including the solve."""
    from cipc import gpu
    ctx = gpu.GpuContext(scene, xtilde=scene.inertial_target())
    ctx.set_velocities(scene.velocities)
    params = gpu.step_params(scene, precision=precision, preconditioner=preconditioner, rel_tol=1e-6)
    kappa = 0.0
    rows = []
    for _ in range(steps):
        t0 = time.perf_counter()
        kappa, st = ctx.time_step(kappa, params)
        rows.append(((time.perf_counter() - t0) * 1e3, st))
    ctx.close()
    newton = sum(st["newton_iterations"] for _, st in rows)
    wall = sum(ms for ms, _ in rows)
    return {"steps": steps, "solve": {"precision": "fp32 PCG in fp64 refinement" if precision else "fp64 PCG",
                                      "preconditioner": ["block-Jacobi", "block FSAI", "learned"][preconditioner]},
            "ms_per_step": round(wall / steps, 3), "newton_iterations": newton,
            "ms_per_newton_iter": round(wall / max(newton, 1), 3),
            "pcg_iterations": int(sum(st["pcg_iterations"] for _, st in rows)),
            "stage_ms_per_step": {k: round(sum(st[k] for _, st in rows) / steps, 3)
                                  for k in ("assembly_ms", "linear_ms", "ccd_ms", "line_search_ms", "post_ms")},
            "energy_backtracks": int(sum(st["energy_backtracks"] for _, st in rows)),
            "kappa": kappa}


def pcg_leg(ctx, kappa, tol=1e-6, max_iter=10000, repeats=3):
    """Newton linear solve of the same frozen state's system (NewtonLinearSystem::solve,
    NewtonLinearSystem.cpp:21-56) to the reference's 1e-6 relative residual: FP64 PCG and
    FP32 PCG inside FP64 refinement, each with block-Jacobi, block FSAI and (when trained
    weights are present) the learned factor; device time (factor set-up and, for the learned
    factor, network inference included) from HIP events, median of `repeats` solves."""
    import numpy as np
    ctx.assemble(kappa)
    out = {"tol": tol, "system": "the frozen state's assembled Newton system (lower CSR, n = 3V)"}
    legs = [("plain_fp64_jacobi", 0, 0), ("mixed_fp32_fp64_jacobi", 1, 0), ("plain_fp64_fsai", 0, 1), ("mixed_fp32_fp64_fsai", 1, 1)]
    from cipc import train_precond
    weights = train_precond.DEFAULT_WEIGHTS
    if os.path.exists(weights):
        ctx.set_preconditioner_weights(np.load(weights))
        legs += [("plain_fp64_learned", 0, 2), ("mixed_fp32_fp64_learned", 1, 2)]
    for name, prec, pc in legs:
        runs = []
        for _ in range(repeats):
            try:
                _, st = ctx.solve(precision=prec, rel_tol=tol, max_iter=max_iter, preconditioner=pc)
                runs.append(st)
            except RuntimeError as e:
                runs.append({"error": str(e)})
                break
        ok = [r for r in runs if "error" not in r]
        if ok:
            ms = float(np.median([r["solve_ms"] for r in ok]))
            out[name] = {"ms": round(ms, 3), "iterations": int(ok[-1]["iterations"]),
                         "refinements": int(ok[-1]["refinements"]), "rel_residual": float(ok[-1]["rel_residual"])}
        else:
            out[name] = {"error": runs[-1]["error"]}
    base = out.get("plain_fp64_jacobi", {}).get("ms")
    if base:
        out["speedup_vs_plain_fp64_jacobi"] = {name: round(base / out[name]["ms"], 3) for name, _, _ in legs[1:]
                                               if "ms" in out.get(name, {})}
    out["learned_weights"] = os.path.relpath(weights, ROOT) if os.path.exists(weights) else None
    return out


def batch_leg(n_scenes=256, steps=5, seed=0, solve=True):
    """The batched workload: n independent lattice cubes (8^3..24^3 cells, per-scene E, nu,
    drop height) in one batch context, each at its first Newton iteration (frozen state,
    probe direction).  Times the batched Newton iteration (HIP events) with the default
    full-CCD split and again after the autotuner has picked one, and one per-scene
    lock-step solve (FP32 PCG with FP64 refinement, block FSAI, 1e-6)."""
    import numpy as np
    from cipc import gpu, scenes
    t0 = time.perf_counter()
    sc = scenes.batched_cubes(n_scenes, seed=seed)
    probes = [scenes.newton_probe_state(s) for s in sc]
    kappa = np.array([scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass) for s in sc])
    b = gpu.BatchContext(sc, xtildes=[xt for xt, _ in probes])
    b.set_state(V_prev=np.concatenate([s.x - s.velocities * s.material.dt for s in sc]))
    b.active_set()
    b.friction_update(kappa)
    b.set_direction(np.concatenate([p for _, p in probes]))
    setup_s = time.perf_counter() - t0
    b.newton_iteration(kappa)
    it = [b.newton_iteration(kappa) for _ in range(steps)]
    ms = float(np.mean([t["total_ms"] for t in it]))
    tune = b.autotune()                               # full-CCD lanes per query, on this state
    itt = [b.newton_iteration(kappa) for _ in range(steps)]
    b.set_hessian_precision(1)                        # the mixed form
    b.newton_iteration(kappa)
    itm = [b.newton_iteration(kappa) for _ in range(steps)]
    b.set_hessian_precision(0)
    solve_ms, iters = None, None
    if solve:
        b.assemble(kappa)
        _, st = b.solve(precision=1, rel_tol=1e-6, preconditioner=1)
        solve_ms = round(st[0]["solve_ms"], 3)
        iters = [int(min(s["iterations"] for s in st)), int(max(s["iterations"] for s in st))]
    out = {"scenes": n_scenes, "V": int(b.V), "T": int(sum(s.T for s in sc)), "newton_iteration_ms": round(ms, 3),
           "scene_iterations_per_s": round(n_scenes / (ms * 1e-3), 1),
           "stages_ms": {k: round(float(np.mean([t[k] for t in it])), 3) for k in
                         ("broadphase_ms", "assembly_ms", "elastic_ms", "ccd_ms", "energy_ms")},
           "autotuned": {"newton_iteration_ms": round(float(np.mean([t["total_ms"] for t in itt])), 3),
                         "ccd_ms": round(float(np.mean([t["ccd_ms"] for t in itt])), 3),
                         "lanes_per_query": {k: 1 << c for k, (c, _) in tune.items()},
                         "candidate_ms": {k: [round(m, 3) for m in ms_] for k, (_, ms_) in tune.items()}},
           "mixed_fp32_hessian": {"newton_iteration_ms": round(float(np.mean([t["total_ms"] for t in itm])), 3),
                                  "elastic_ms": round(float(np.mean([t["elastic_ms"] for t in itm])), 3)},
           "contacts": int(it[-1]["contacts"]), "min_alpha": float(it[-1]["alpha"]),
           "solve_mixed_fsai_ms": solve_ms, "solve_iterations_min_max": iters,
           "setup_s": round(setup_s, 1)}
    b.close()
    return out


def pcg_large_leg(tol=1e-6, max_iter=20000):
    """The large linear-solve workload: eight 32^3-cell lattice cubes (287,496 vertices /
    1,310,720 tets) stacked in contact with seeded velocities; the assembled Newton system of
    a frozen state solved to 1e-6 relative residual with FP64 block-Jacobi PCG (the plain
    baseline), FP64 and mixed block FSAI, and the learned factor when weights exist."""
    import numpy as np
    from cipc import gpu, scenes, train_precond
    sc = scenes.stacked_cubes(8, 32, 0.1, 1e-3, seed=4, rotate="yaw", velocity_scale=0.5, name="8x32cubes")
    xt, p = scenes.newton_probe_state(sc)
    kappa = scenes.suggest_kappa(sc.hhat, sc.bbox_diag2, sc.mean_mass)
    ctx = gpu.GpuContext(sc, xtilde=xt)
    ctx.set_state(V_prev=sc.x - sc.velocities * sc.material.dt)
    a = ctx.active_set()
    ctx.friction_update(kappa)
    r = ctx.assemble(kappa)
    out = {"V": sc.V, "T": sc.T, "n": 3 * sc.V, "nnz": int(len(r["val"])),
           "contacts": int(len(a["self_set"]) + len(a["ee_set"]) + len(a["env_set"]))}
    legs = [("plain_fp64_jacobi", 0, 0), ("mixed_fp32_fp64_jacobi", 1, 0), ("plain_fp64_fsai", 0, 1), ("mixed_fp32_fp64_fsai", 1, 1)]
    if os.path.exists(train_precond.DEFAULT_WEIGHTS):
        ctx.set_preconditioner_weights(np.load(train_precond.DEFAULT_WEIGHTS))
        legs += [("plain_fp64_learned", 0, 2), ("mixed_fp32_fp64_learned", 1, 2)]
    for name, prec, pc in legs:
        try:
            ctx.solve(precision=prec, rel_tol=tol, max_iter=max_iter, preconditioner=pc)   # warm-up
            _, st = ctx.solve(precision=prec, rel_tol=tol, max_iter=max_iter, preconditioner=pc)
            out[name] = {"ms": round(st["solve_ms"], 3), "iterations": st["iterations"], "refinements": st["refinements"],
                         "rel_residual": st["rel_residual"]}
        except RuntimeError as e:
            out[name] = {"error": str(e)}
    base = out["plain_fp64_jacobi"].get("ms")
    if base:
        out["speedup_vs_plain_fp64_jacobi"] = {n: round(base / out[n]["ms"], 3) for n, _, _ in legs[1:] if "ms" in out[n]}
    ctx.close()
    return out


def traffic_probe(n=5):
    """Workload for the HBM-traffic counters of the elastic stage: n Newton iterations of the
    frozen state (each launches the elastic kernels once), then one 256 MiB device-to-device
    copy through torch, a plain copy kernel of known traffic (256 MiB read, 256 MiB written)
    that scales the counters.  Profiled once with --pmc FETCH_SIZE and once with WRITE_SIZE."""
    import torch
    from cipc import gpu, scenes
    scene = scenes.four_cubes20(velocity_scale=0.5)
    xt, p = scenes.newton_probe_state(scene)
    kappa = scenes.suggest_kappa(scene.hhat, scene.bbox_diag2, scene.mean_mass)
    ctx = gpu.GpuContext(scene, xtilde=xt)
    ctx.set_state(V_prev=scene.x - scene.velocities * scene.material.dt)
    ctx.active_set()
    ctx.friction_update(kappa)
    ctx.set_direction(p)
    for _ in range(n):
        ctx.newton_iteration(kappa)
    ctx.close()
    a = torch.empty(256 * 1024 * 1024 // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print("traffic probe done", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-pcg", action="store_true", help="skip the linear solves (the single-scene leg and the batch's)")
    ap.add_argument("--no-time-step", action="store_true", help="skip the whole-time-step leg")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "elastic_traffic_4x20cubes.json"),
                    help="JSON file with the measured HBM bytes per launch of the elastic stage (rocprofv3 FETCH_SIZE + "
                         "WRITE_SIZE of --traffic-probe, scaled by a copy of known size)")
    ap.add_argument("--traffic-probe", action="store_true", help="run the HBM-counter workload instead of the benchmark")
    ap.add_argument("--batch", type=int, default=0, help="also time a batch of this many independent scenes (256: the batched workload)")
    ap.add_argument("--pcg-large", action="store_true", help="also time the linear solves of the eight-32^3-cube system")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (every run uses a single MI355X)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from cipc import gpu, scenes

    scene = scenes.four_cubes20(velocity_scale=0.5)
    xt, p = scenes.newton_probe_state(scene)
    vprev = scene.x - scene.velocities * scene.material.dt
    kappa = scenes.suggest_kappa(scene.hhat, scene.bbox_diag2, scene.mean_mass)
    ctx = gpu.GpuContext(scene, xtilde=xt)
    ctx.set_state(V_prev=vprev)
    ctx.active_set()
    ctx.friction_update(kappa)
    ctx.set_direction(p)
    stage_keys = ("broadphase_ms", "assembly_ms", "elastic_ms", "ccd_ms", "energy_ms", "total_ms")

    def timed(graph, steps, warmup):
        """Iterations timed by the host clock around them (each call ends with its one
        synchronisation); graph: the iteration as one HIP graph launch, or its stages launched
        one by one (the single-scene default) and timed per stage with HIP events."""
        ctx.set_tune(2, 1 if graph else 0)
        for _ in range(warmup):
            ctx.newton_iteration(kappa)
        t0 = time.perf_counter()
        its = [ctx.newton_iteration(kappa) for _ in range(steps)]
        return (time.perf_counter() - t0) * 1e3 / steps, its

    wall_ms, stages = timed(False, args.steps, args.warmup)
    last = stages[-1]
    mean = {k: float(np.mean([s[k] for s in stages])) for k in stage_keys}
    graph_ms, graph_its = timed(True, args.steps, 2)
    graphed = int(sum(s["graph"] for s in graph_its))
    # the mixed form: FP32 elastic Hessian blocks, FP64 barriers / friction / gradient
    ctx.set_hessian_precision(1)
    mixed_wall, mixed = timed(False, args.steps, args.warmup)
    mixed_graph_ms, _ = timed(True, args.steps, 2)
    ctx.set_hessian_precision(0)
    ctx.set_tune(2, 0)
    mixed_out = {"ms_per_iter": round(mixed_wall, 4), "graph_ms_per_iter": round(mixed_graph_ms, 4),
                 "stages_ms": {k: round(float(np.mean([s[k] for s in mixed])), 4) for k in stage_keys}}
    aset = ctx.active_set()
    pcg = pcg_leg(ctx, kappa) if not args.no_pcg else None
    ctx.close()

    bytes_el, nnz_el = elastic_algorithmic_bytes(scene)
    achieved = bytes_el / (mean["elastic_ms"] * 1e-3) / 1e9
    traffic = flops = None
    if args.traffic and os.path.exists(args.traffic):
        counters = json.load(open(args.traffic))
        traffic = counters.get("bytes_per_launch")
        flops = counters.get("fp64_flops_per_launch")
    out = {
        "metric": METRIC,
        "value": round(wall_ms, 4),
        "unit": "ms/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "4x20^3 lattice cubes, 37044 verts / 160000 tets, stacked in contact, friction 0.3, frozen Newton state",
                   "V": scene.V, "T": scene.T, "nnz": int(last["nnz"]), "contacts": int(last["contacts"]),
                   "step": "one Newton iteration's non-linear stages (active set, assembly into the CSR, step bound, "
                           "line-search energy); the linear solve is timed on its own in pcg"},
        "launch": {"graph_ms_per_iter": round(graph_ms, 4), "graph_launches": graphed, "of": args.steps,
                   "note": "value and stages_ms: the stages launched one by one (the single-scene default), each "
                           "timed with HIP events; graph_ms_per_iter: the same stages as one HIP graph launch, "
                           "replayed (recaptured after each tree-topology rebuild)"},
        "stages_ms": {k: round(v, 4) for k, v in mean.items()},
        # the assembly's HBM rate, the metric's second figure
        "assembly_hbm": {"kernel": "elastic assembly (snh_chunk_kernel + chunk_complete_kernel + mass_diag_kernel)",
                         "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes": int(bytes_el),
                         "elastic_nnz": int(nnz_el), "traffic": traffic, "stage_ms": round(mean["elastic_ms"], 4)},
    }
    if flops:
        # The elastic stage sits on the compute side of the ridge (18 FP64 flop per algorithmic
        # byte against 9.8): its roof is the FP64 peak.  FLOPs per launch counted by rocprofv3
        # (FP64 add, mul, FMA x2, transcendental; 64 lanes per wavefront instruction) on the
        # same frozen state; time from the HIP events around the stage in this run.
        tf = flops / (mean["elastic_ms"] * 1e-3) / 1e12
        out["roofline"] = {"kernel": "elastic assembly (snh_chunk_kernel + chunk_complete_kernel + mass_diag_kernel)",
                           "bound": "mfma", "achieved": round(tf, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": round(tf / FP64_PEAK_TFLOPS, 4), "traffic": traffic, "flops_per_launch": int(flops),
                           "flops": "executed FP64 FLOPs (rocprofv3 SQ_INSTS_VALU_*_F64 x 64 lanes)",
                           "limiter": "FP64 issue latency of the per-tet SVD and projection chains (divisions, square "
                                      "roots) at about 2.4 wavefronts per SIMD for 160,000 tets; the peak is the FP64 "
                                      "vector/matrix rate"}
    else:
        out["roofline"] = dict(out["assembly_hbm"], bound="hbm")
    # every other stage against HBM: all far below it (launch- and dependency-bound)
    out["stages_roofline"] = stage_rooflines(scene, aset, int(last["nnz"]), nnz_el, mean)
    out["mixed_fp32_hessian"] = mixed_out
    if pcg is not None:
        out["pcg"] = pcg
        best = min((v["ms"] for k, v in pcg.items() if isinstance(v, dict) and "ms" in v), default=None)
        if best is not None:
            # a whole Newton iteration: the non-linear stages plus the fastest solve to 1e-6
            out["ms_per_newton_iter_with_solve"] = round(wall_ms + best, 4)
    if not args.no_time_step:
        out["time_step"] = time_step_leg(scene)
    if args.batch > 0:
        out["batch"] = batch_leg(args.batch, solve=not args.no_pcg)
    if args.pcg_large:
        out["pcg_large"] = pcg_large_leg()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, xt, vprev, p, kappa)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
