"""Pins of the oracle's soft target-position penalty (SoftBoundaryCondition,
BoundaryConditions.h:54-67; addSoftDerivatives / softEnergy, BoundaryConditions.cpp:
175-229): the energy term is 0.5 w |x - t|^2 summed, the gradient is its derivative, and the
assembled Hessian gains exactly w on the soft vertices' diagonal DOFs."""
import numpy as np
import pytest

from cipc import scenes


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def _dense_diag(r, n):
    d = np.zeros(n)
    for c in range(n):
        for k in range(r["outer"][c], r["outer"][c + 1]):
            if r["inner"][k] == c:
                d[c] = r["val"][k]
    return d


def test_soft_targets(O):
    s = scenes.stacked_cubes(2, 4, 0.1, 3e-4, seed=4, rotate="yaw", name="soft")
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    rng = np.random.default_rng(0)
    idx = rng.choice(s.V, 12, replace=False)
    tgt = s.x[idx] + 0.01 * rng.standard_normal((12, 3))
    w = 3.7e2
    m0, m1 = O.OracleMesh(s), O.OracleMesh(s)
    m1.set_soft_targets(idx, tgt, w)
    for m in (m0, m1):
        m.active_set()
    r0, r1 = m0.assemble(kappa), m1.assemble(kappa)
    t1 = m1.energy_terms(kappa)
    assert t1[8] == pytest.approx(0.5 * w * np.sum((s.x[idx] - tgt) ** 2), rel=1e-13)
    dg = (r1["grad"] - r0["grad"])
    expect = np.zeros_like(dg)
    expect[idx] = w * (s.x[idx] - tgt)
    assert np.allclose(dg, expect, rtol=1e-9, atol=1e-12 * np.abs(r0["grad"]).max())
    n = 3 * s.V
    dd = _dense_diag(r1, n) - _dense_diag(r0, n)
    ex = np.zeros(n)
    ex[(3 * idx[:, None] + np.arange(3)).reshape(-1)] = w
    assert np.allclose(dd, ex, rtol=1e-12, atol=1e-9 * w)
    assert np.array_equal(r0["inner"], r1["inner"])        # the diagonal already exists
    # energy derivative
    v = rng.standard_normal(s.x.shape)
    h = 1e-7
    m1.set_state(x=s.x + h * v)
    Ep = m1.energy(kappa)
    m1.set_state(x=s.x - h * v)
    Em = m1.energy(kappa)
    assert (Ep - Em) / (2 * h) == pytest.approx(np.sum(r1["grad"] * v), rel=1e-5)
