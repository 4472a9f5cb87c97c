"""Batched scenes on the device: one context holding several independent scenes (the
reference's equivalent is one FEMSimulator / IPCSolverContext per scene, IPCSolver.h:46-61)
must evaluate every scene exactly as a context holding that scene alone.

The batch mixes cube stacks of different sizes, Lame parameters, masses and barrier
stiffness, and holds one scene twice at the same place: any contact across scenes would
show up there.  Bars: active sets, CSR patterns and values, gradients, step bounds, solve
iteration counts and directions bit-exact against the single-scene contexts (the batch keeps
each scene's order of every sum); energies within 1e-12 relative (their per-scene reduction
order differs), and against the CPU oracle within 1e-10 as the single-scene tests.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [(3, 1e5, 0.40, 0, 1.0), (4, 3e4, 0.33, 1, 2.0), (5, 6e5, 0.45, 2, 0.5), (3, 1e5, 0.40, 0, 1.0), (2, 2e5, 0.30, 4, 3.0)]


def _scenes():
    from cipc import scenes
    out = []
    for n, E, nu, seed, _ in SPECS:
        mat = scenes.Material(youngs=E, poisson=nu, friction=0.2)
        out.append(scenes.stacked_cubes(2, n, 0.1, 1e-3, seed=seed, rotate="yaw", velocity_scale=0.5, mat=mat, name=f"b{n}"))
    hh, fh, dt = max(s.hhat for s in out), max(s.fhat for s in out), max(s.dtol for s in out)
    for s in out:
        s.hhat, s.fhat, s.dtol = hh, fh, dt
    return out


@pytest.fixture(scope="module")
def gpu():
    from cipc import gpu as g
    g.load()
    return g


@pytest.fixture(scope="module")
def setup(gpu):
    from cipc import scenes
    sc = _scenes()
    probes = [scenes.newton_probe_state(s, seed=3 + k) for k, s in enumerate(sc)]
    kappa = np.array([scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass) * f for s, (*_, f) in zip(sc, SPECS)])
    b = gpu.BatchContext(sc, xtildes=[xt for xt, _ in probes])
    b.set_state(V_prev=np.concatenate([s.x - s.velocities * s.material.dt for s in sc]))
    singles = []
    for s, (xt, _) in zip(sc, probes):
        g = gpu.GpuContext(s, xtilde=xt)
        g.set_state(V_prev=s.x - s.velocities * s.material.dt)
        singles.append(g)
    yield dict(b=b, singles=singles, scenes=sc, probes=probes, kappa=kappa)
    b.close()
    for g in singles:
        g.close()


def _shift4(a, off):
    a = a.copy()
    pos = a >= 0
    a[pos] -= off
    neg_x = a[:, 0] < 0
    a[neg_x, 0] += off
    return a


def _first_vertex(a):
    return np.where(a[:, 0] >= 0, a[:, 0], -a[:, 0] - 1)


def _scene_of(v, voff):
    return np.searchsorted(voff, v, side="right") - 1


def _sub(A, lo, hi):
    return A[lo:hi, lo:hi]


def _csr(r, n):
    import scipy.sparse as sp
    return sp.csc_matrix((r["val"], r["inner"], r["outer"]), shape=(n, n)).tocsr()


def test_batch_layout(setup):
    b, sc = setup["b"], setup["scenes"]
    assert b.S == len(sc)
    assert list(np.diff(b.voff)) == [s.V for s in sc]
    assert list(np.diff(b.toff)) == [s.T for s in sc]


def test_batch_active_set_matches_scenes(setup):
    b, singles = setup["b"], setup["singles"]
    ab = b.active_set()
    voff = b.voff
    total = 0
    for k, g in enumerate(singles):
        a = g.active_set()
        for key in ("self_set", "ee_set"):
            rec = ab[key]
            mine = rec[_scene_of(_first_vertex(rec), voff) == k] if len(rec) else rec
            assert np.array_equal(_shift4(mine, voff[k]), a[key]), (k, key)
            total += len(a[key])
        env = ab["env_set"]
        assert np.array_equal(env[_scene_of(env, voff) == k] - voff[k], a["env_set"]), k
    assert total > 0
    assert sum(len(g.active_set()["self_set"]) for g in singles) == len(ab["self_set"])   # nothing across scenes


def test_batch_assembly_matches_scenes(setup):
    b, singles, kappa = setup["b"], setup["singles"], setup["kappa"]
    b.active_set()
    b.friction_update(kappa)
    rb = b.assemble(kappa)
    n = 3 * b.V
    A = _csr(rb, n)
    inside = 0
    for k, g in enumerate(singles):
        g.active_set()
        g.friction_update(kappa[k])
        r = g.assemble(kappa[k])
        lo, hi = 3 * b.voff[k], 3 * b.voff[k + 1]
        B = _sub(A, lo, hi).tocsc()
        B.sort_indices()
        S = _csr(r, hi - lo).tocsc()
        S.sort_indices()
        assert np.array_equal(B.indptr, S.indptr) and np.array_equal(B.indices, S.indices), k
        assert np.array_equal(B.data, S.data), k
        assert np.array_equal(rb["grad"][b.voff[k]:b.voff[k + 1]], r["grad"]), k
        assert np.array_equal(rb["rhs"][lo:hi], r["rhs"]), k
        assert rb["collisions"][k] == r["collisions"], k
        inside += B.nnz
    assert inside == A.nnz          # no entry couples two scenes


def test_batch_energy_matches_scenes_and_oracle(setup, oracle_lib):
    b, singles, kappa, sc, probes = setup["b"], setup["singles"], setup["kappa"], setup["scenes"], setup["probes"]
    b.active_set()
    b.friction_update(kappa)
    tb = b.energy_terms(kappa)
    eb = b.energy(kappa)
    for k, g in enumerate(singles):
        g.active_set()
        g.friction_update(kappa[k])
        t = g.energy_terms(kappa[k])
        for q in range(9):
            assert abs(tb[k, q] - t[q]) <= 1e-12 * max(abs(t[q]), 1e-300) or tb[k, q] == t[q], (k, q, tb[k, q], t[q])
        e = g.energy(kappa[k])
        assert abs(eb[k] - e) <= 1e-12 * abs(e), k
        o = oracle_lib.OracleMesh(sc[k], xtilde=probes[k][0])
        try:
            o.set_state(V_prev=sc[k].x - sc[k].velocities * sc[k].material.dt)
            o.active_set()
            o.friction_update(kappa[k])
            eo = o.energy(kappa[k])
        finally:
            o.close()
        assert abs(eb[k] - eo) <= 1e-10 * abs(eo), (k, eb[k], eo)


def test_batch_step_bound_matches_scenes(setup):
    b, singles, probes = setup["b"], setup["singles"], setup["probes"]
    b.active_set()
    p = np.concatenate([pp for _, pp in probes])
    bb = b.step_bound(p)
    assert np.all(bb["alpha"] > 0)
    for k, g in enumerate(singles):
        g.active_set()
        s1 = g.step_bound(probes[k][1])
        for key in ("alpha", "ground", "partial", "cfl", "full"):
            assert bb[key][k] == s1[key], (k, key, bb[key][k], s1[key])


@pytest.mark.parametrize("preconditioner", [0, 1])
@pytest.mark.parametrize("precision", [0, 1])
def test_batch_solve_matches_scenes(setup, precision, preconditioner):
    b, singles, kappa = setup["b"], setup["singles"], setup["kappa"]
    b.active_set()
    b.friction_update(kappa)
    b.assemble(kappa)
    pb, stb = b.solve(precision=precision, rel_tol=1e-6, preconditioner=preconditioner)
    parts = b.split(pb)
    for k, g in enumerate(singles):
        g.active_set()
        g.friction_update(kappa[k])
        g.assemble(kappa[k])
        p, st = g.solve(precision=precision, rel_tol=1e-6, preconditioner=preconditioner)
        assert stb[k]["iterations"] == st["iterations"] and stb[k]["refinements"] == st["refinements"], (k, stb[k], st)
        assert stb[k]["rel_residual"] == st["rel_residual"], k
        assert np.array_equal(parts[k], p), (k, np.abs(parts[k] - p).max())
    # convergence mask: a masked scene keeps a zero direction, the others are unchanged
    act = np.ones(b.S, np.int32)
    act[1] = 0
    pm, stm = b.solve(precision=precision, rel_tol=1e-6, preconditioner=preconditioner, active=act)
    pmp = b.split(pm)
    assert not np.any(pmp[1]) and stm[1]["iterations"] == 0
    for k in range(b.S):
        if k != 1:
            assert np.array_equal(pmp[k], parts[k]), k


def test_batch_step_forward_and_intersection(setup):
    b, singles, probes, sc = setup["b"], setup["singles"], setup["probes"], setup["scenes"]
    x0 = np.concatenate([s.x for s in sc])
    b.set_state(x=x0)
    p = np.concatenate([pp for _, pp in probes])
    b.set_direction(p)
    b.save_positions()
    alpha = np.linspace(0.1, 0.9, b.S)
    b.step_forward(alpha)
    xb = b.split(b.positions())
    hit = b.is_intersected()
    for k, g in enumerate(singles):
        g.set_state(x=sc[k].x)
        g.set_direction(probes[k][1])
        g.save_positions()
        g.step_forward(alpha[k])
        assert np.array_equal(xb[k], g.positions()), k
        assert bool(hit[k]) == g.is_intersected(), k
        g.set_state(x=sc[k].x)
    b.set_state(x=x0)


def test_batch_newton_iteration_matches_scenes(setup):
    b, singles, kappa, probes = setup["b"], setup["singles"], setup["kappa"], setup["probes"]
    b.set_direction(np.concatenate([pp for _, pp in probes]))
    b.active_set()
    b.friction_update(kappa)
    tb = b.newton_iteration(kappa)
    for k, g in enumerate(singles):
        g.set_direction(probes[k][1])
        g.active_set()
        g.friction_update(kappa[k])
        t = g.newton_iteration(kappa[k])
        assert tb["alpha_scenes"][k] == t["alpha"], k
        assert abs(tb["energy_scenes"][k] - t["energy"]) <= 1e-12 * abs(t["energy"]), k
    assert tb["alpha"] == min(tb["alpha_scenes"])


def test_single_scene_entry_points_refuse_a_batch(setup, gpu):
    b = setup["b"]
    with pytest.raises(gpu.CipcLogicError):
        gpu.GpuContext.energy(b, 1.0)
    with pytest.raises(gpu.CipcLogicError):
        gpu.GpuContext.step_bound(b)


def test_batch_rejects_mismatched_scenes(gpu):
    from cipc import scenes
    a = scenes.stacked_cubes(1, 2, 0.1, 0.05, seed=0, rotate="none")
    c = scenes.stacked_cubes(1, 3, 0.1, 0.05, seed=1, rotate="none")
    c.hhat = a.hhat * 2
    with pytest.raises(ValueError):
        gpu.BatchContext([a, c])
