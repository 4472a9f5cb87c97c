"""The oracle's time step (oracle/step_oracle.py, solveIPCStep restated on the oracle's
stages) pinned by what the mechanics fixes:

* free fall: a body at its rest shape, away from the ground, with a uniform velocity, lands
  exactly on its inertial target x + v dt + g dt^2 (the elastic Hessian annihilates the
  uniform translation that the first Newton direction is, so one Newton step solves the
  step), and its velocity becomes v + g dt;
* quasi-static: the velocity is zeroed;
* no gravity and no velocity: nothing moves and no Newton iteration is taken;
* a body resting just above the ground keeps a positive gap under gravity (the barrier
  holds it), the ground contacts are active and the stiffness is positive.
"""
import numpy as np
import pytest


def _cube(height, velocity=(0.0, 0.0, 0.0), cells=3, gravity=True):
    from cipc import scenes
    s = scenes.stacked_cubes(1, cells, 0.1, height, seed=2, rotate="yaw", name="cube")
    s.velocities[:] = np.asarray(velocity)
    s.apply_gravity = gravity
    return s


def test_free_fall_lands_on_inertial_target(oracle_lib):
    from step_oracle import OracleStep
    s = _cube(0.5, velocity=(0.1, -0.2, 0.05))
    xt = s.inertial_target()
    m = oracle_lib.OracleMesh(s, xtilde=xt)
    st = OracleStep(m, s)
    stats = st.step()
    assert stats["newton_iterations"] == 1
    assert np.abs(st.x - xt).max() <= 1e-12
    dt = s.material.dt
    v_want = s.velocities + np.array([0.0, -9.8 * dt, 0.0])
    assert np.abs(st.v - v_want).max() <= 1e-9
    assert np.array_equal(st.vprev, st.x)
    assert np.abs(st.xt - (st.x + st.v * dt + np.array([0.0, -9.8 * dt * dt, 0.0]))).max() <= 1e-15
    m.close()


def test_quasi_static_zeroes_velocity(oracle_lib):
    from step_oracle import OracleStep
    s = _cube(0.5, velocity=(0.1, 0.0, 0.0))
    m = oracle_lib.OracleMesh(s, xtilde=s.inertial_target())
    st = OracleStep(m, s, quasi_static=True)
    st.step()
    assert np.array_equal(st.v, np.zeros_like(st.v))
    m.close()


def test_at_rest_without_gravity_takes_no_newton_iteration(oracle_lib):
    from step_oracle import OracleStep
    s = _cube(0.5, gravity=False)
    m = oracle_lib.OracleMesh(s, xtilde=s.inertial_target())
    st = OracleStep(m, s)
    stats = st.step()
    assert stats["newton_iterations"] == 0
    assert np.array_equal(st.x, s.x)
    m.close()


def test_resting_on_ground_keeps_a_gap(oracle_lib):
    from cipc import scenes
    from step_oracle import OracleStep
    s = _cube(3e-4)
    m = oracle_lib.OracleMesh(s, xtilde=s.inertial_target())
    st = OracleStep(m, s)
    for _ in range(2):
        stats = st.step()
        assert stats["kappa"] > 0
        assert stats["newton_iterations"] >= 1
    assert st.x[:, 1].min() > scenes.GROUND_Y
    assert len(m.active_set()["env_set"]) > 0
    m.close()


def test_gradient_scale_bounds_the_gradient(oracle_lib):
    """The per-entry magnitude of the terms summed into the gradient is at least the
    gradient's own magnitude, and exceeds it where terms cancel."""
    from cipc import scenes
    s = scenes.stacked_cubes(2, 2, 0.1, 3e-4, seed=5, rotate="yaw", velocity_scale=0.3, name="scale")
    m = oracle_lib.OracleMesh(s, xtilde=s.inertial_target())
    m.active_set()
    r = m.assemble(1e5)
    assert (r["grad_scale"] >= np.abs(r["grad"]) * (1 - 1e-15)).all()
    assert (r["grad_scale"] > 1.5 * np.abs(r["grad"])).any()
    m.close()
