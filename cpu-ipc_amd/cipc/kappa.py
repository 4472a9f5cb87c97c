# This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
# If a copy of the MPL was not distributed with this file, You can obtain one at
# https://mozilla.org/MPL/2.0/.  Parts of this file restate logic of CPU-IPC (MPL-2.0);
# each such part cites the source it follows.
"""Barrier stiffness (kappa) control of the IPC solver, host side.

Mirrors suggestKappa / upperBoundKappa / initKappa / postLineSearch
(``CPU IPC/IPCSolver.cpp:751-869``) and their call sites in solveIPCStep
(``IPCSolver.cpp:1118-1123,1161-1170``) and solveBarrierSubproblem (``IPCSolver.cpp:1007``).
The data-parallel parts run on the device: the two reductions of initKappa
(``GpuContext.kappa_gradients``) and the close-constraint check and records of
postLineSearch (``GpuContext.close_constraints``).  What remains here is scalar logic.

The reference keeps two kappas: ``mesh.Kappa`` and the by-value copy that one barrier
subproblem works with.  upperBoundKappa is always applied to ``mesh.Kappa``; inside a
subproblem it therefore leaves the working copy unbounded.  ``KappaControl`` keeps both,
so that a caller sees the reference's exact sequence of values.
"""
from __future__ import annotations

import math


def barrier_hessian(d: float, dhat: float) -> float:
    """compute_H_b (ContactMechanics.cpp:786-790)."""
    t = d - dhat
    return (math.log(d / dhat) * -2.0 - t * 4.0 / d) + 1.0 / (d * d) * (t * t)


def suggest_kappa(hhat: float, bbox_diag2: float, mean_mass: float) -> float:
    """suggestKappa (IPCSolver.cpp:751-755)."""
    h_b = barrier_hessian(1.0e-16 * bbox_diag2, hhat)
    return 1e13 * mean_mass / (4.0e-16 * bbox_diag2 * h_b)


def upper_bound_kappa(kappa: float, hhat: float, bbox_diag2: float, mean_mass: float) -> float:
    """upperBoundKappa (IPCSolver.cpp:757-764)."""
    h_b = barrier_hessian(1.0e-16 * bbox_diag2, hhat)
    kmax = 100 * 1e13 * mean_mass / (4.0e-16 * bbox_diag2 * h_b)
    return kmax if kappa > kmax else kappa


class KappaControl:
    """kappa bookkeeping of one mesh over time steps and barrier subproblems.

    ``backend`` is a device context (or the oracle) with ``kappa_gradients()`` ->
    (sum g_c.g_E, sum |g_c|^2, n_constraints), ``close_constraints(dtol)`` -> (closer, counts)
    and ``clear_close_constraints()``, all on its current positions and active sets."""

    def __init__(self, backend, hhat: float, bbox_diag2: float, mean_mass: float, dtol: float,
                 kappa: float = 0.0):
        self.backend = backend
        self.hhat, self.bbox_diag2, self.mean_mass, self.dtol = hhat, bbox_diag2, mean_mass, dtol
        self.mesh_kappa = kappa     # mesh.Kappa
        self.kappa = kappa          # the working copy of the current barrier subproblem

    def _bound_mesh(self):
        self.mesh_kappa = upper_bound_kappa(self.mesh_kappa, self.hhat, self.bbox_diag2, self.mean_mass)

    def _init(self, kappa: float, aliased: bool) -> float:
        """initKappa (IPCSolver.cpp:766-808) on `kappa`; the final bound goes to mesh.Kappa,
        which is `kappa` itself when ``aliased``."""
        gsum, gsnorm, n = self.backend.kappa_gradients()
        if not n:
            return kappa
        min_kappa = -gsum / gsnorm
        if min_kappa > 0.0:
            kappa = min_kappa
        min_kappa = suggest_kappa(self.hhat, self.bbox_diag2, self.mean_mass)
        if kappa < min_kappa:
            kappa = min_kappa
        if aliased:
            kappa = upper_bound_kappa(kappa, self.hhat, self.bbox_diag2, self.mean_mass)
            self.mesh_kappa = kappa
        else:
            self._bound_mesh()
        return kappa

    def step_start(self) -> float:
        """solveIPCStep (IPCSolver.cpp:1118-1123): bound, suggest when ~0, initKappa on
        mesh.Kappa.  Call with the step's active sets built."""
        self._bound_mesh()
        if self.mesh_kappa < 1e-16:
            self.mesh_kappa = suggest_kappa(self.hhat, self.bbox_diag2, self.mean_mass)
        self.mesh_kappa = self._init(self.mesh_kappa, aliased=True)
        return self.mesh_kappa

    def subproblem_start(self) -> float:
        """IPCSolver.cpp:1163-1170: close records cleared, working copy = mesh.Kappa."""
        self.backend.clear_close_constraints()
        self.kappa = self.mesh_kappa
        return self.kappa

    def post_line_search(self) -> float:
        """postLineSearch (IPCSolver.cpp:811-869) on the working copy, after an accepted
        line search (positions and active sets of the new iterate)."""
        if self.kappa == 0.0:
            self.kappa = self._init(self.kappa, aliased=False)
            return self.kappa
        closer, _ = self.backend.close_constraints(self.dtol)
        if closer:
            self.kappa *= 2.0
            self._bound_mesh()
        return self.kappa
