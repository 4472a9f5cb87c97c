"""Pins of the oracle's barrier-stiffness reductions and of the host kappa logic
(initKappa / postLineSearch, IPCSolver.cpp:751-869).

initKappa's vectors are fixed independently of their code: with the mollified edge-edge set
removed, no friction and no drag, the assembled gradient at kappa = 0 is g_E (inertia +
elastic) and its change from kappa = 0 to kappa = 1 is g_c, so both sums follow from two
assemblies.  The close-constraint records are checked against a brute-force evaluation of
the constraint values."""
import numpy as np
import pytest

from cipc import kappa as K
from cipc import scenes


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def _scene():
    s = scenes.stacked_cubes(2, 5, 0.1, 3e-4, seed=8, rotate="yaw", velocity_scale=0.3, name="kappa")
    return s


def test_kappa_gradients_from_assemblies(O):
    s = _scene()
    xt, _ = scenes.newton_probe_state(s)
    m = O.OracleMesh(s, xtilde=xt)
    a = m.active_set()
    a["ee_set"] = np.zeros((0, 4), np.int32)
    a["ee_ie"] = np.zeros((0, 2), np.int32)
    m.set_active_set(a)
    g0 = m.assemble(0.0)["grad"].reshape(-1)
    g1 = m.assemble(1.0)["grad"].reshape(-1)
    gc = g1 - g0
    gsum, gsnorm, n = m.kappa_gradients()
    assert n == len(a["env_set"]) + len(a["self_set"]) > 0
    assert gsum == pytest.approx(gc @ g0, rel=1e-9)
    assert gsnorm == pytest.approx(gc @ gc, rel=1e-9)


def _values(O, s, x, a):
    gnd = (x[a["env_set"], 1] + 1.0) ** 2
    sv = []
    for c in a["self_set"]:
        if c[0] >= 0:
            sv.append(O.distance(O.DIST_EE, x[c].reshape(-1))[0])
            continue
        p = -c[0] - 1
        if c[2] < 0:
            sv.append(O.distance(O.DIST_PP, np.concatenate([x[p], x[c[1]]]))[0])
        elif c[3] < 0:
            sv.append(O.distance(O.DIST_PE, np.concatenate([x[p], x[c[1]], x[c[2]]]))[0])
        else:
            sv.append(O.distance(O.DIST_PT, np.concatenate([x[p], x[c[1]], x[c[2]], x[c[3]]]))[0])
    return gnd, np.array(sv)


def test_close_constraints_brute_force(O):
    s = _scene()
    m = O.OracleMesh(s)
    a = m.active_set()
    gnd, sv = _values(O, s, s.x, a)
    dtol = float(np.quantile(np.concatenate([gnd, sv]), 0.3))
    closer, cnt = m.close_constraints(dtol)
    assert not closer and cnt == (int((gnd < dtol).sum()), int((sv < dtol).sum())) and sum(cnt) > 0
    closer, _ = m.close_constraints(dtol)
    assert closer          # unchanged positions: every record's value <= itself
    m.clear_close_constraints()
    closer, _ = m.close_constraints(dtol)
    assert not closer


def test_kappa_control_sequence(O):
    """KappaControl's values against the reference's formulas, step start and doubling."""
    s = _scene()
    xt, _ = scenes.newton_probe_state(s)
    m = O.OracleMesh(s, xtilde=xt)
    m.active_set()
    kc = K.KappaControl(m, s.hhat, s.bbox_diag2, s.mean_mass, s.dtol)
    k0 = kc.step_start()
    gsum, gsnorm, _ = m.kappa_gradients()
    sug = K.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    expect = max(-gsum / gsnorm if -gsum / gsnorm > 0 else sug, sug)
    expect = K.upper_bound_kappa(expect, s.hhat, s.bbox_diag2, s.mean_mass)
    assert k0 == expect and kc.mesh_kappa == k0
    kc.subproblem_start()
    kc.dtol = 1e300                       # every constraint is "close": the next check doubles
    assert kc.post_line_search() == k0    # first call only records
    assert kc.post_line_search() == 2 * k0
    assert kc.mesh_kappa == k0            # the bound and the doubling act on different copies
