# This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
# If a copy of the MPL was not distributed with this file, You can obtain one at
# https://mozilla.org/MPL/2.0/.  Parts of this file restate logic of CPU-IPC (MPL-2.0);
# each such part cites the source it follows.
"""ORACLE — test infrastructure only; never part of the product path.

CPU restatement of one implicit time step, solveIPCStep (CPU IPC/IPCSolver.cpp:1079-1336),
on the oracle's stages (oracle.OracleMesh): initKappa and Friction::initialize at the step
start, the animated Dirichlet update (:1127-1155), the kappa loop of barrier subproblems
(solveBarrierSubproblem :890-1064 with lineSearch :560-749 and postLineSearch :811-869),
then updateVelocity (:1066-1077), V_prev = x and updateInertialTarget (:55-70).

The linear solve is a sparse direct solve of the oracle's CSR (scipy), standing in for
PARDISO / CHOLMOD (NewtonLinearSystem.cpp:21-56).  The control flow, limits and exceptions
follow the reference line by line; the step's positions, velocities, stiffness and counters
are what a caller of solveIPCStep would see.
"""
from __future__ import annotations

import math

import numpy as np

EPS = np.finfo(np.float64).eps


def _barrier_h(d, dhat):
    """compute_H_b (ContactMechanics.cpp:786-790)."""
    t = d - dhat
    return (math.log(d / dhat) * -2.0 - t * 4.0 / d) + 1.0 / (d * d) * (t * t)


def suggest_kappa(hhat, diag2, mean_mass):
    """suggestKappa (IPCSolver.cpp:751-755)."""
    return 1e13 * mean_mass / (4.0e-16 * diag2 * _barrier_h(1.0e-16 * diag2, hhat))


def upper_bound_kappa(kappa, hhat, diag2, mean_mass):
    """upperBoundKappa (IPCSolver.cpp:757-764)."""
    kmax = 100 * 1e13 * mean_mass / (4.0e-16 * diag2 * _barrier_h(1.0e-16 * diag2, hhat))
    return kmax if kappa > kmax else kappa


class LineSearchFailure(RuntimeError):
    pass


class OracleStep:
    """Time steps of one scene on an OracleMesh.  Host copies of x, V_prev, x~ and the
    velocities are kept here and pushed to the mesh as the reference's mesh3D would hold
    them."""

    def __init__(self, mesh, scene, quasi_static=False):
        self.m = mesh
        self.s = scene
        self.x = scene.x.copy()
        self.vprev = scene.V_prev.copy()
        self.v = scene.velocities.copy()
        self.xt = scene.inertial_target()
        self.free = (scene.boundary == 0)[:, None]
        self.quasi_static = quasi_static
        self.mesh_kappa = scene.kappa
        self.m.set_state(x=self.x, V_prev=self.vprev, xtilde=self.xt)

    # ---- stages ----------------------------------------------------------------------
    def _positions(self):
        self.x = self.m.positions()
        return self.x

    def _solve(self, kappa):
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        r = self.m.assemble(kappa)
        n = 3 * self.s.V
        L = sp.csc_matrix((r["val"], r["inner"], r["outer"]), shape=(n, n))
        A = (L + L.T - sp.diags(L.diagonal())).tocsc()
        p = spla.spsolve(A, r["rhs"]).reshape(-1, 3)
        return p, r

    def _step_forward(self, x0, p, alpha):
        self.m.step_forward(x0, p, alpha)

    # ---- kappa -----------------------------------------------------------------------
    def _init_kappa(self, kappa, aliased):
        """initKappa (IPCSolver.cpp:766-808)."""
        s = self.s
        gsum, gsnorm, n = self.m.kappa_gradients()
        if not n:
            return kappa
        min_kappa = -gsum / gsnorm
        if min_kappa > 0.0:
            kappa = min_kappa
        min_kappa = suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
        if kappa < min_kappa:
            kappa = min_kappa
        if aliased:
            kappa = upper_bound_kappa(kappa, s.hhat, s.bbox_diag2, s.mean_mass)
        else:
            self.mesh_kappa = upper_bound_kappa(self.mesh_kappa, s.hhat, s.bbox_diag2, s.mean_mass)
        return kappa

    def _post_line_search(self, kappa):
        """postLineSearch (IPCSolver.cpp:811-869)."""
        s = self.s
        if kappa == 0.0:
            return self._init_kappa(kappa, aliased=False)
        closer, _ = self.m.close_constraints(s.dtol)
        if closer:
            kappa *= 2.0
            self.mesh_kappa = upper_bound_kappa(self.mesh_kappa, s.hhat, s.bbox_diag2, s.mean_mass)
        return kappa

    # ---- line search ----------------------------------------------------------------------
    def _line_search(self, p, grad, alpha, kappa, dnorm, thr, st):
        """lineSearch (IPCSolver.cpp:560-749) without the diagnostic branch."""
        e0 = self.m.energy(kappa)
        gp = float(np.sum(grad * p))
        if not (math.isfinite(e0) and math.isfinite(gp)):
            raise LineSearchFailure("line search received a non-finite energy or direction")
        x0 = self.m.positions()
        self._step_forward(x0, p, alpha)
        n_int = 0
        while self.m.is_intersected():
            if n_int >= 64 or alpha <= EPS:
                raise LineSearchFailure("line search could not find a non-intersecting step")
            n_int += 1
            alpha /= 2.0
            self._step_forward(x0, p, alpha)
        self.m.active_set()
        e = self.m.energy(kappa)
        n_e = 0
        lf = alpha
        while not (math.isfinite(e) and e < e0):
            if n_e >= 64 or alpha <= EPS * max(1.0, lf):
                if dnorm < thr:
                    self.m.set_state(x=x0)
                    self.m.active_set()
                    self._tally(st, n_e, n_int)
                    return False, alpha
                raise LineSearchFailure("line search failed to find a strict energy decrease")
            alpha *= 0.5
            n_e += 1
            self._step_forward(x0, p, alpha)
            self.m.active_set()
            e = self.m.energy(kappa)
        if alpha < lf:
            recompute = False
            guard = 0
            while self.m.is_intersected():
                if guard >= 64 or alpha <= EPS:
                    raise LineSearchFailure("backtracked step remains intersecting")
                guard += 1
                alpha /= 2.0
                n_int += 1
                self._step_forward(x0, p, alpha)
                recompute = True
            if recompute:
                self.m.active_set()
        self._tally(st, n_e, n_int)
        st["alpha_sum"] += alpha
        st["alpha_min"] = min(st["alpha_min"], alpha)
        st["alpha_max"] = max(st["alpha_max"], alpha)
        return True, alpha

    @staticmethod
    def _tally(st, n_e, n_i):
        st["energy_backtracks"] += n_e
        st["intersection_backtracks"] += n_i
        st["max_energy_backtracks_per_newton"] = max(st["max_energy_backtracks_per_newton"], n_e)
        if n_e > 2:
            st["newton_steps_over_two_backtracks"] += 1

    # ---- Newton ----------------------------------------------------------------------------
    def _barrier_subproblem(self, kappa, st, max_newton=10000):
        """solveBarrierSubproblem (IPCSolver.cpp:890-1064) with a by-value kappa."""
        s = self.s
        thr = s.material.newton_threshold * math.sqrt(s.bbox_diag2) * s.material.dt
        beta = 1.0
        k = 0
        while k < max_newton:
            p, r = self._solve(kappa)
            st["collisions"] += r["collisions"]
            grad = r["grad"].reshape(-1, 3)
            dnorm = float(np.abs(p).max())
            if dnorm < thr:
                break
            alpha = self.m.step_bound(p)["alpha"]
            ok, alpha = self._line_search(p, grad, alpha, kappa, dnorm, thr, st)
            if not ok:
                break
            kappa = self._post_line_search(kappa)
            if k + 1 >= 6:
                beta = (1 - alpha) * beta
            if beta <= thr:
                break
            k += 1
        st["newton_iterations"] += k
        return k

    def _dirichlet(self, verts, step, fn):
        """The animated Dirichlet update (IPCSolver.cpp:1127-1155)."""
        s = self.s
        dt = s.material.dt

        def dirs(frac):
            x = self.m.positions()
            p = np.zeros((s.V, 3))
            p[verts] = np.asarray(fn(x[verts].copy(), s.v_rest[verts].copy(), step, frac, dt), dtype=np.float64)
            return p

        new_alpha = 1.0
        p = dirs(new_alpha)
        new_alpha = self.m.full_ccd(p, new_alpha)
        p = dirs(new_alpha)
        x0 = self.m.positions()
        self.m.set_state(x=x0 - 1.0 * p)
        n = 0
        while self.m.is_intersected():
            if n >= 64 or new_alpha <= EPS:
                raise LineSearchFailure("animated boundary update remains intersecting")
            n += 1
            new_alpha /= 2.0
            self.m.set_state(x=x0)
            p = dirs(new_alpha)
            self.m.set_state(x=x0 - 1.0 * p)
        self.m.active_set()
        return new_alpha

    def step(self, dirichlet=None, max_newton=10000):
        """One solveIPCStep; returns the stats dict."""
        s = self.s
        st = dict(newton_iterations=0, kappa_iterations=0, energy_backtracks=0, intersection_backtracks=0,
                  max_energy_backtracks_per_newton=0, newton_steps_over_two_backtracks=0, collisions=0,
                  alpha_sum=0.0, alpha_min=math.inf, alpha_max=0.0, min_distance2=0.0, dirichlet_alpha=1.0)
        self.mesh_kappa = upper_bound_kappa(self.mesh_kappa, s.hhat, s.bbox_diag2, s.mean_mass)
        if self.mesh_kappa < 1e-16:
            self.mesh_kappa = suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
        self.m.active_set()
        self.mesh_kappa = self._init_kappa(self.mesh_kappa, aliased=True)
        self.m.friction_update(self.mesh_kappa)
        if dirichlet is not None:
            st["dirichlet_alpha"] = self._dirichlet(*dirichlet)
        while st["kappa_iterations"] < 64:
            self.m.clear_close_constraints()
            self._barrier_subproblem(self.mesh_kappa, st, max_newton)
            lo, hi, n = self.m.constraint_extrema()
            if n:
                st["min_distance2"] = lo
                if lo < s.dtol or hi < s.hhat:
                    break
            else:
                break
            st["kappa_iterations"] += 1
        if st["kappa_iterations"] == 64:
            raise LineSearchFailure("IPC kappa loop exceeded its iteration limit")
        st["kappa_iterations"] += 1
        # updateVelocity, V_prev = x, updateInertialTarget (IPCSolver.cpp:1196-1204, 55-70)
        x = self.m.positions()
        self.v = np.zeros_like(x) if self.quasi_static else (x - self.vprev) / s.material.dt
        self.vprev = x.copy()
        g = np.zeros(3)
        if s.apply_gravity:
            g[1] = (-9.8 * s.material.dt) * s.material.dt
        self.xt = (self.vprev + self.v * s.material.dt) + g
        self.x = x
        self.m.set_state(V_prev=self.vprev, xtilde=self.xt)
        st["kappa"] = self.mesh_kappa
        return st
