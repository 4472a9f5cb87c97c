"""The mixed form of the Newton assembly: elastic Hessian blocks (tets, triangles, bending)
rounded to FP32 where they are staged, everything else (barriers, friction, mass, gradient,
right-hand side) in FP64.  No reference counterpart: the reference assembles in FP64.

Bars: CSR pattern, gradient and right-hand side identical to the FP64 assembly; CSR values
within 1e-6 of the largest magnitude (FP32 rounding of each contribution, at most a few
dozen per entry, summed in FP64); the mixed system solved to 1e-6 relative residual.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from cipc import gpu as g
    g.load()
    return g


def _scene(name):
    from cipc import scenes
    if name == "4x20cubes":
        return scenes.four_cubes20(velocity_scale=0.5)
    return scenes.cloth_on_sphere(24, sphere_cells=4, radius=0.25, fold=16, bending=1, seed=2, name=name)


@pytest.mark.parametrize("name", ["4x20cubes", "cloth24_hinge"])
def test_mixed_hessian_matches_fp64(gpu, name):
    from cipc import scenes
    s = _scene(name)
    xt, p = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    g = gpu.GpuContext(s, xtilde=xt)
    try:
        g.set_state(V_prev=s.x - s.velocities * s.material.dt)
        g.active_set()
        g.friction_update(kappa)
        r64 = g.assemble(kappa)
        g.set_hessian_precision(1)
        r32 = g.assemble(kappa)
        assert np.array_equal(r64["outer"], r32["outer"]) and np.array_equal(r64["inner"], r32["inner"]), name
        assert np.array_equal(r64["grad"], r32["grad"]) and np.array_equal(r64["rhs"], r32["rhs"]), name
        scale = np.abs(r64["val"]).max()
        err = np.abs(r32["val"] - r64["val"]).max() / scale
        assert 0 < err <= 1e-6, (name, err)
        _, st = g.solve(precision=1, rel_tol=1e-6, preconditioner=1)
        assert st["rel_residual"] <= 1e-6
        g.set_hessian_precision(0)
        r = g.assemble(kappa)
        assert np.array_equal(r["val"], r64["val"])      # back to FP64
    finally:
        g.close()


def test_hessian_precision_rejects_other_values(gpu):
    from cipc import scenes
    g = gpu.GpuContext(scenes.cube10())
    try:
        with pytest.raises(ValueError):
            g.set_hessian_precision(2)
    finally:
        g.close()


def test_mixed_assembly_against_oracle(gpu, oracle_lib):
    """The mixed form against the FP64 oracle, per result: gradient, right-hand side and
    energy at the FP64 bars (they never touch FP32); CSR values per entry within 1e-7 of the
    magnitude summed into the entry (the FP32 rounding of each elastic block, 6e-8)."""
    from cipc import scenes
    from parity_util import csr_entry_err, vec3_entry_err
    s = scenes.four_cubes20(velocity_scale=0.5)
    xt, p = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    g = gpu.GpuContext(s, xtilde=xt)
    o = oracle_lib.OracleMesh(s, xtilde=xt)
    try:
        for m in (g, o):
            m.set_state(V_prev=s.x - s.velocities * s.material.dt)
            m.active_set()
            m.friction_update(kappa)
        g.set_hessian_precision(1)
        rg, ro = g.assemble(kappa), o.assemble(kappa)
        assert np.array_equal(rg["outer"], ro["outer"]) and np.array_equal(rg["inner"], ro["inner"])
        ev = csr_entry_err(ro["outer"], ro["inner"], rg["val"], ro["val"], scale=ro["val_scale"], scale_floor=1.0)
        eg = vec3_entry_err(rg["grad"], ro["grad"], scale=ro["grad_scale"])
        ee = abs(g.energy(kappa) - o.energy(kappa)) / abs(o.energy(kappa))
        print(f"mixed form vs oracle: CSR per entry {ev:.2e} of the summed magnitude, gradient {eg:.2e}, energy {ee:.2e}")
        assert ev <= 1e-7 and eg <= 1e-10 and ee <= 1e-10, (ev, eg, ee)
    finally:
        g.close()
        o.close()


def test_mixed_time_steps_against_oracle(gpu, oracle_lib):
    """Whole time steps with the mixed form (FP32 elastic Hessian, FP32 PCG inside FP64
    refinement) against the oracle's FP64 steps: the Hessian only shapes the Newton
    direction, so both loops converge to the same positions within the Newton tolerance
    (Newton_Solver_Threshold sqrt(bboxDiagSize2) dt, IPCSolver.cpp:933-935)."""
    from cipc import scenes
    from step_oracle import OracleStep
    s = scenes.stacked_cubes(2, 3, 0.1, 4e-4, seed=3, rotate="yaw", velocity_scale=0.2, name="mixed_steps")
    xt = s.inertial_target()
    g = gpu.GpuContext(s, xtilde=xt)
    g.set_velocities(s.velocities)
    g.set_hessian_precision(1)
    o = OracleStep(oracle_lib.OracleMesh(s, xtilde=xt), s)
    thr = s.material.newton_threshold * np.sqrt(s.bbox_diag2) * s.material.dt
    params = gpu.step_params(s, precision=1, preconditioner=1, rel_tol=1e-6)
    try:
        kg = 0.0
        for step in range(3):
            kg, sg = g.time_step(kg, params)
            so = o.step()
            dx = np.abs(g.state()[0] - o.x).max()
            print(f"mixed step {step}: newton {sg['newton_iterations']}/{so['newton_iterations']} |dx| {dx:.2e} "
                  f"(Newton tolerance {thr:.1e})")
            assert dx <= thr, (step, dx, thr)
    finally:
        g.close()
