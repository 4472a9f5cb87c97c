"""Training of the learned block preconditioner on Newton systems the simulator generates.

Every training system is the assembled Newton matrix of a frozen state (lagged friction
included) of one of the benchmark scene families: stacked lattice cubes of several sizes
and seeds (the four-cube and batched-cube workloads) and folded cloth on a ball.  Each step
is one Adam step of the device's own forward and backward kernels on one system
(GpuContext.preconditioner_train_step), cycling through the systems.  The network starts
at the analytic first-order block FSAI (zero decoders) and is fitted to the block FSAI
factor of each system (objective 1): it then predicts an FSAI-quality factor without the
per-row dense solves.

    python -m cipc.train_precond --steps 3000 --out cpu-ipc_amd/cipc/data/precond_weights.npy
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "precond_weights.npy")


def init_params(n: int, seed: int = 0) -> np.ndarray:
    """He-style initialization of the hidden layers, zero decoders (layout of the device
    library: encoders, three layers, node decoder, edge decoder)."""
    from cipc import gnn_layout
    rng = np.random.default_rng(seed)
    out = []
    for name, shape in gnn_layout.param_shapes():
        if name.startswith("b") or name in ("Wd", "Wo"):
            out.append(np.zeros(shape))
        else:
            scale = math.sqrt(2.0 / shape[0]) * (0.5 if name.startswith("Wu") else 1.0)
            out.append(rng.standard_normal(shape) * scale)
    flat = np.concatenate([a.reshape(-1) for a in out]).astype(np.float32)
    assert len(flat) == n
    return flat


def training_systems(gpu, count_scale: int = 1):
    """(context, label) pairs with an assembled system each."""
    from cipc import scenes
    out = []

    def add(s, label, seed):
        xt, _ = scenes.newton_probe_state(s, seed=seed)
        k = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
        g = gpu.GpuContext(s, xtilde=xt)
        g.set_state(V_prev=s.x - s.velocities * s.material.dt)
        g.active_set()
        g.friction_update(k)
        g.assemble(k)
        out.append((g, label))

    for sd in range(3 * count_scale):
        cells = (8, 12, 16)[sd % 3]
        add(scenes.stacked_cubes(2, cells, 0.1, 1e-3, seed=sd, rotate="yaw", velocity_scale=0.5, name="cubes"),
            f"2x{cells}^3 cubes seed {sd}", sd)
    add(scenes.four_cubes20(seed=11, velocity_scale=0.5), "4x20^3 cubes seed 11", 11)
    add(scenes.cloth_on_sphere(64, sphere_cells=6, radius=0.25, fold=40, seed=4), "cloth 64 on ball", 4)
    return out


def train(steps: int = 3000, lr: float = 1e-3, probes: int = 4, seed: int = 0, log_every: int = 100, objective: int = 1):
    from cipc import gpu
    gpu.load()
    n = gpu.preconditioner_param_count()
    systems = training_systems(gpu)
    p = init_params(n, seed)
    m1 = np.zeros_like(p)
    m2 = np.zeros_like(p)
    t0 = time.time()
    losses = []
    for step in range(1, steps + 1):
        g, _ = systems[step % len(systems)]
        lr_t = lr * min(1.0, step / 100) * 0.5 * (1 + math.cos(math.pi * step / steps))
        loss = g.preconditioner_train_step(p, m1, m2, step, max(lr_t, 1e-7), probes, seed * 1000003 + step, objective)
        losses.append(loss)
        if step % log_every == 0:
            print(f"step {step} loss {np.mean(losses[-log_every:]):.5f} {time.time() - t0:.1f}s", flush=True)
    for g, _ in systems:
        g.close()
    return p, {"steps": steps, "lr": lr, "probes": probes, "seed": seed, "objective": objective, "seconds": time.time() - t0,
               "systems": [label for _, label in systems], "final_loss": float(np.mean(losses[-log_every:]))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--probes", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    a = ap.parse_args()
    p, meta = train(a.steps, a.lr, a.probes, a.seed)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.save(a.out, p)
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
