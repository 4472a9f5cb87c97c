"""Pins of the oracle's lagged friction (Friction.cpp, FrictionKinematics.h): the lagged
state against its defining conditions, and the friction energy, gradient and Hessian
against each other by central differences.  The friction part is isolated as the
difference between two meshes that differ only in the friction coefficient, so the
elastic and barrier terms cancel exactly."""
import dataclasses

import numpy as np
import pytest

from cipc import scenes


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def _stack(mu, seed=5):
    mat = dataclasses.replace(scenes.Material(), friction=mu)
    return scenes.stacked_cubes(2, 4, 0.1, 3e-4, seed=seed, rotate="yaw", velocity_scale=0.0, jitter=0.002,
                                mat=mat, name="fric")


def _slid_state(s, seed=0):
    """V_prev = x minus a per-object tangential slide plus small per-vertex noise, so that
    relative tangential displacements fall on both sides of the static threshold."""
    rng = np.random.default_rng(seed)
    eps = np.sqrt(s.fhat) * s.material.dt
    vprev = s.x.copy()
    upper = s.x[:, 1] > np.median(s.x[:, 1])
    vprev[upper] -= np.array([1.5 * eps, 0.0, -0.7 * eps])
    vprev += 0.6 * eps * rng.standard_normal(s.x.shape)
    return vprev, eps


def _dense(r, n):
    H = np.zeros((n, n))
    for c in range(n):
        for k in range(r["outer"][c], r["outer"][c + 1]):
            H[r["inner"][k], c] += r["val"][k]
            if r["inner"][k] != c:
                H[c, r["inner"][k]] += r["val"][k]
    return H


def test_lagged_state_definitions(O):
    s = _stack(0.3)
    vprev, _ = _slid_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    m = O.OracleMesh(s)
    m.set_state(V_prev=vprev)
    a = m.active_set()
    m.friction_update(kappa)
    f = m.friction_state()
    S, x = a["self_set"], s.x
    assert len(S) > 0 and len(f["lam"]) == len(S) and len(f["env_lam"]) == len(a["env_set"])
    for i, c in enumerate(S):
        B = f["basis"][i]
        assert np.allclose(B @ B.T, np.eye(2), atol=1e-12)
        b1, b2 = f["coord"][i]
        if c[0] >= 0:       # edge-edge: closest points of the two lines
            kind, pts = O.DIST_EE, x[c]
            r = (x[c[0]] + b1 * (x[c[1]] - x[c[0]])) - (x[c[2]] + b2 * (x[c[3]] - x[c[2]]))
            for e in (x[c[1]] - x[c[0]], x[c[3]] - x[c[2]]):
                assert abs(r @ e) <= 1e-9 * np.linalg.norm(r) * np.linalg.norm(e) + 1e-15
            mult = 1
        else:
            p = -c[0] - 1
            if c[2] < 0:
                kind, pts, r = O.DIST_PP, np.stack([x[p], x[c[1]]]), None
            elif c[3] < 0:
                kind, pts = O.DIST_PE, np.stack([x[p], x[c[1]], x[c[2]]])
                r = x[p] - (x[c[1]] + b1 * (x[c[2]] - x[c[1]]))
                assert abs(r @ (x[c[2]] - x[c[1]])) <= 1e-9 * np.linalg.norm(r) * np.linalg.norm(x[c[2]] - x[c[1]]) + 1e-15
            else:
                kind, pts = O.DIST_PT, np.stack([x[p], x[c[1]], x[c[2]], x[c[3]]])
                r = x[p] - (x[c[1]] + b1 * (x[c[2]] - x[c[1]]) + b2 * (x[c[3]] - x[c[1]]))
                for e in (x[c[2]] - x[c[1]], x[c[3]] - x[c[1]]):
                    assert abs(r @ e) <= 1e-9 * np.linalg.norm(r) * np.linalg.norm(e) + 1e-15
            mult = -c[3] if c[3] < -1 else 1
        if r is not None:   # the tangent basis spans the plane orthogonal to the normal direction
            assert np.abs(B @ r).max() <= 1e-8 * np.linalg.norm(r)
        d = O.distance(kind, pts.reshape(-1))[0]
        _, bg, _ = O.barrier(d, s.hhat)
        assert f["lam"][i] == pytest.approx(-kappa * 2.0 * np.sqrt(d) * bg * mult, rel=1e-12)
        assert f["lam"][i] > 0


@pytest.mark.parametrize("seed", [0, 1])
def test_friction_energy_gradient_hessian_consistent(O, seed):
    """With the lagged state fixed, d/dh E_f(x + h v) = g_f . v and
    d/dh g_f(x + h v) = H_f v, where (E_f, g_f, H_f) is the friction part."""
    s1, s0 = _stack(0.3), _stack(0.0)
    vprev, eps = _slid_state(s1, seed)
    kappa = scenes.suggest_kappa(s1.hhat, s1.bbox_diag2, s1.mean_mass)
    ms = []
    for s in (s1, s0):
        m = O.OracleMesh(s)
        m.set_state(V_prev=vprev)
        m.active_set()
        m.friction_update(kappa)
        ms.append(m)
    n = 3 * s1.V
    x = s1.x

    def parts(xx):
        out = []
        for m in ms:
            m.set_state(x=xx)
            r = m.assemble(kappa)
            out.append((m.energy(kappa), r["grad"].reshape(-1).copy(), r))
        return out[0][0] - out[1][0], out[0][1] - out[1][1], out

    E, g, res = parts(x)
    H = _dense(res[0][2], n) - _dense(res[1][2], n)
    assert np.abs(g).max() > 0 and np.abs(H).max() > 0
    ef = ms[0].energy_terms(kappa)[5]
    assert ef > 0 and E == pytest.approx(ef, rel=1e-9)
    rng = np.random.default_rng(10 + seed)
    v = rng.standard_normal(n)
    v /= np.abs(v).max()
    h = 1e-3 * eps
    Ep, gp, _ = parts(x + h * v.reshape(-1, 3))
    Em, gm, _ = parts(x - h * v.reshape(-1, 3))
    fd = (Ep - Em) / (2 * h)
    assert abs(fd - g @ v) <= 1e-6 * np.abs(g).sum()
    fdg = (gp - gm) / (2 * h)
    assert np.allclose(H @ v, fdg, rtol=1e-5, atol=1e-5 * np.abs(fdg).max())
