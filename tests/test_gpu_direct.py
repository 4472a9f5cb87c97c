"""The direct tier of the Newton solve (cipc_gpu_solve_ex preconditioner 3: the dense
Cholesky factorisation of the assembled system on the GPU, standing in for PARDISO /
CHOLMOD, NewtonLinearSystem.cpp:150-174).

Bars: on config[0]'s cube (3,993 unknowns, ground contact and friction) the direction
equals scipy's sparse direct solve of the oracle's CSR to 1e-9 relative and the FP64
residual is below 1e-12; time steps with the direct tier follow the oracle's steps as the
PCG steps do; a system above the size limit is refused.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_direct_solve_matches_sparse_direct(oracle_lib):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from cipc import gpu, scenes
    s = scenes.cube10()
    xt, p = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    g = gpu.GpuContext(s, xtilde=xt)
    o = oracle_lib.OracleMesh(s, xtilde=xt)
    try:
        for m in (g, o):
            m.set_state(V_prev=s.x - s.velocities * s.material.dt)
            m.active_set()
            m.friction_update(kappa)
        g.assemble(kappa)
        ro = o.assemble(kappa)
        d, st = g.solve(precision=0, rel_tol=1e-6, preconditioner=3)
        n = 3 * s.V
        L = sp.csc_matrix((ro["val"], ro["inner"], ro["outer"]), shape=(n, n))
        A = (L + L.T - sp.diags(L.diagonal())).tocsc()
        want = spla.spsolve(A, ro["rhs"])
        err = np.abs(d.reshape(-1) - want).max() / np.abs(want).max()
        print(f"direct tier: n {n}, residual {st['rel_residual']:.2e}, vs sparse direct {err:.2e}, {st['solve_ms']:.1f} ms")
        assert st["rel_residual"] <= 1e-12 and err <= 1e-9, (st, err)
        assert st["iterations"] == 0
    finally:
        g.close()
        o.close()


def test_direct_time_steps_match_oracle(oracle_lib):
    from cipc import gpu, scenes
    from step_oracle import OracleStep
    s = scenes.stacked_cubes(2, 3, 0.1, 4e-4, seed=3, rotate="yaw", velocity_scale=0.2, name="direct_steps")
    xt = s.inertial_target()
    g = gpu.GpuContext(s, xtilde=xt)
    g.set_velocities(s.velocities)
    o = OracleStep(oracle_lib.OracleMesh(s, xtilde=xt), s)
    params = gpu.step_params(s, precision=0, preconditioner=3)
    try:
        kg = 0.0
        for step in range(3):
            kg, sg = g.time_step(kg, params)
            so = o.step()
            dx = np.abs(g.state()[0] - o.x).max()
            print(f"direct step {step}: newton {sg['newton_iterations']}/{so['newton_iterations']} |dx| {dx:.2e}")
            assert sg["newton_iterations"] == so["newton_iterations"] and sg["pcg_iterations"] == 0
            assert dx <= 1e-9, (step, dx)
    finally:
        g.close()


def test_direct_solve_size_limit():
    from cipc import gpu, scenes
    s = scenes.four_cubes20(velocity_scale=0.5)     # 111,132 unknowns
    xt, p = scenes.newton_probe_state(s)
    g = gpu.GpuContext(s, xtilde=xt)
    try:
        g.active_set()
        g.assemble(scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass))
        with pytest.raises(ValueError, match="at most"):
            g.solve(precision=0, preconditioner=3)
    finally:
        g.close()
