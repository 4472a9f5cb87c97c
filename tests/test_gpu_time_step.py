"""Whole time steps inside the device library (cipc_gpu_time_step: solveIPCStep with its
barrier subproblems, line search, stiffness control and end-of-step updates,
IPCSolver.cpp:1079-1336), against the oracle's restatement of the same step
(oracle/step_oracle.py, sparse direct solves standing in for PARDISO).

Bars: the step's counters (Newton iterations, kappa loops, energy and intersection
backtracks) equal the oracle's; positions within 1e-8 m and velocities within 1e-6 m/s
after every step (the device solves each Newton system to 1e-10 relative residual, the
oracle exactly); the stiffness within 1e-9 relative.  Free fall lands on the inertial target
to 1e-12 m, as the oracle's own pin requires.  The animated Dirichlet update (a prescribed
press of the top layer, IPCSolver.cpp:1127-1155) is checked the same way, its CCD-limited
fraction to 1e-12.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(oracle_lib, s, quasi_static=False):
    from cipc import gpu
    from step_oracle import OracleStep
    xt = s.inertial_target()
    g = gpu.GpuContext(s, xtilde=xt)
    g.set_velocities(s.velocities)
    o = OracleStep(oracle_lib.OracleMesh(s, xtilde=xt), s, quasi_static=quasi_static)
    return g, o


def _params(s):
    from cipc import gpu
    return gpu.step_params(s, precision=0, preconditioner=1, rel_tol=1e-10, max_iter=20000)


def _compare(g, o, sg, so, kg, step):
    x, vprev, xt, v = g.state()
    dx = np.abs(x - o.x).max()
    dv = np.abs(v - o.v).max()
    print(f"step {step}: newton {sg['newton_iterations']}/{so['newton_iterations']} kappa loops "
          f"{sg['kappa_iterations']}/{so['kappa_iterations']} backtracks {sg['energy_backtracks']}+"
          f"{sg['intersection_backtracks']} kappa {kg:.6e}/{so['kappa']:.6e} |dx| {dx:.2e} |dv| {dv:.2e} "
          f"{sg['step_ms']:.1f} ms", flush=True)
    for k in ("newton_iterations", "kappa_iterations", "energy_backtracks", "intersection_backtracks"):
        assert sg[k] == so[k], (step, k, sg[k], so[k])
    assert abs(kg - so["kappa"]) <= 1e-9 * abs(so["kappa"]), (step, kg, so["kappa"])
    assert dx <= 1e-8, (step, dx)
    assert dv <= 1e-6, (step, dv)
    assert np.array_equal(vprev, x)
    dt = o.s.material.dt
    assert np.abs(xt - (x + v * dt + np.array([0.0, -9.8 * dt * dt, 0.0]))).max() <= 1e-15


def test_free_fall_lands_on_inertial_target(oracle_lib):
    from cipc import scenes
    s = scenes.stacked_cubes(1, 3, 0.1, 0.5, seed=2, rotate="yaw", name="fall")
    s.velocities[:] = np.array([0.1, -0.2, 0.05])
    g, o = _pair(oracle_lib, s)
    try:
        xt = s.inertial_target()
        kg, sg = g.time_step(0.0, _params(s))
        so = o.step()
        assert sg["newton_iterations"] == 1
        x = g.state()[0]
        assert np.abs(x - xt).max() <= 1e-12
        _compare(g, o, sg, so, kg, 0)
    finally:
        g.close()


@pytest.mark.parametrize("seed", [3])
def test_time_steps_match_oracle(oracle_lib, seed):
    from cipc import scenes
    s = scenes.stacked_cubes(2, 3, 0.1, 4e-4, seed=seed, rotate="yaw", velocity_scale=0.2, name="steps")
    g, o = _pair(oracle_lib, s)
    try:
        kg = 0.0
        contacts = 0
        for step in range(3):
            kg, sg = g.time_step(kg, _params(s))
            so = o.step()
            contacts = max(contacts, sg["self_contacts"] + sg["ground_contacts"])
            _compare(g, o, sg, so, kg, step)
        assert contacts > 0
        assert g.state()[0][:, 1].min() > scenes.GROUND_Y
    finally:
        g.close()


def test_animated_dirichlet_press_matches_oracle(oracle_lib):
    from cipc import scenes
    s = scenes.stacked_cubes(1, 3, 0.1, 4e-4, seed=5, rotate="yaw", name="press")
    top = np.where(s.x[:, 1] > s.x[:, 1].max() - 0.05)[0].astype(np.int32)
    s.boundary[top] = 1

    def press(x, v_rest, step, fraction, dt):
        d = np.zeros_like(x)
        d[:, 1] = 2e-3 * fraction          # x - d: the top layer moves 2 mm down per step
        return d

    g, o = _pair(oracle_lib, s)
    try:
        kg = 0.0
        for step in range(2):
            kg, sg = g.time_step(kg, _params(s), dirichlet=(top, step, press))
            so = o.step(dirichlet=(top, step, press))
            assert abs(sg["dirichlet_alpha"] - so["dirichlet_alpha"]) <= 1e-12
            _compare(g, o, sg, so, kg, step)
        x = g.state()[0]
        assert np.abs(x[top, 1] - (s.x[top, 1] - 4e-3)).max() <= 1e-12   # both presses taken in full
    finally:
        g.close()
