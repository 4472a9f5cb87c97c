"""Whole time steps driven through the C ABI: the host Newton loop of solveBarrierSubproblem
(IPCSolver.cpp:890-1064) with the line search of IPCSolver.cpp:560-749, run once on the
device context (linear solve: the device PCG) and once on the CPU oracle (linear solve: a
sparse direct solve of the oracle's CSR).  The loop here is test scaffolding, a plain
restatement of the reference's control flow: lagged friction per step, active set +
assembly + solve + step bound per iteration, backtracking on the energy with the
intersection check, convergence on |p|_inf / dt.

Bar: after every time step the two trajectories agree to 1e-6 m (both Newton loops run to
|p|_inf / dt <= 1e-7, far below it).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRAV = np.array([0.0, -9.8, 0.0])


class _Gpu:
    def __init__(self, gpu, s, xt):
        self.c = gpu.GpuContext(s, xtilde=xt)

    def set_state(self, **kw):
        self.c.set_state(**kw)

    def friction_update(self, k):
        self.c.friction_update(k)

    def active_set(self):
        self.c.active_set()

    def newton_direction(self, k):
        self.c.assemble(k)
        p, st = self.c.solve(precision=0, rel_tol=1e-10, max_iter=20000, preconditioner=1)
        return p

    def step_bound(self, p):
        return self.c.step_bound(p)["alpha"]

    def energy(self, k):
        return self.c.energy(k)

    def begin_search(self, p):
        self.c.set_direction(p)
        self.c.save_positions()

    def try_step(self, p, alpha):
        self.c.step_forward(alpha)

    def intersected(self):
        return self.c.is_intersected()

    def positions(self):
        return self.c.positions()

    def close(self):
        self.c.close()


class _Oracle:
    def __init__(self, O, s, xt):
        self.o = O.OracleMesh(s, xtilde=xt)
        self.V = s.V

    def set_state(self, **kw):
        self.o.set_state(**kw)

    def friction_update(self, k):
        self.o.friction_update(k)

    def active_set(self):
        self.o.active_set()

    def newton_direction(self, k):
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        r = self.o.assemble(k)
        n = 3 * self.V
        L = sp.csc_matrix((r["val"], r["inner"], r["outer"]), shape=(n, n))
        A = (L + L.T - sp.diags(L.diagonal())).tocsc()
        return spla.spsolve(A, r["rhs"]).reshape(-1, 3)

    def step_bound(self, p):
        return self.o.step_bound(p)["alpha"]

    def energy(self, k):
        return self.o.energy(k)

    def begin_search(self, p):
        self.x0 = self.o.positions()

    def try_step(self, p, alpha):
        self.o.step_forward(self.x0, p, alpha)

    def intersected(self):
        return self.o.is_intersected()

    def positions(self):
        return self.o.positions()

    def close(self):
        self.o.close()


def _time_step(m, s, x, v, kappa, tol=1e-7, max_newton=60):
    dt = s.material.dt
    free = (s.boundary == 0)[:, None]
    xt = np.where(free, x + dt * v + dt * dt * GRAV, x)
    m.set_state(x=x, V_prev=x, xtilde=xt)
    m.active_set()
    m.friction_update(kappa)
    for _ in range(max_newton):
        m.active_set()
        p = m.newton_direction(kappa)
        if np.abs(p).max() / dt <= tol:
            break
        alpha = m.step_bound(p)
        e0 = m.energy(kappa)
        m.begin_search(p)
        for _ in range(60):
            m.try_step(p, alpha)
            m.active_set()
            if not m.intersected() and m.energy(kappa) <= e0:
                break
            alpha *= 0.5
    xn = m.positions()
    return xn, (xn - x) / dt


@pytest.mark.parametrize("seed", [3])
def test_time_steps_match_oracle(oracle_lib, seed):
    from cipc import gpu as G
    from cipc import scenes
    G.load()
    s = scenes.stacked_cubes(2, 3, 0.1, 4e-4, seed=seed, rotate="yaw", velocity_scale=0.2, name="steps")
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    xt0 = s.inertial_target()
    mg, mo = _Gpu(G, s, xt0), _Oracle(oracle_lib, s, xt0)
    try:
        xg, vg = s.x.copy(), s.velocities.copy()
        xo, vo = s.x.copy(), s.velocities.copy()
        moved = 0.0
        for step in range(3):
            xg, vg = _time_step(mg, s, xg, vg, kappa)
            xo, vo = _time_step(mo, s, xo, vo, kappa)
            moved = max(moved, np.abs(xo - s.x).max())
            assert np.abs(xg - xo).max() <= 1e-6, (step, np.abs(xg - xo).max())
        assert moved > 1e-4                      # the scene did move
        assert xo[:, 1].min() > scenes.GROUND_Y   # and stayed above the ground
    finally:
        mg.close()
        mo.close()
