"""Device path vs the CPU oracle, through the C ABI, on identical inputs.

Bars (stated per test): contact pair sets, classifications, CSR patterns and CCD steps
bit-exact; FP64 energies within 1e-10 relative; CSR values and gradients within 1e-10 per
entry, each entry relative to itself with a floor of 1e-6 of its 3x3 block's (or its
vertex's) largest entry (parity_util); element-level results within 1e-10 relative to the
largest magnitude of the compared element.
"""
import numpy as np
import pytest

from parity_util import csr_entry_err, vec3_entry_err

pytestmark = pytest.mark.gpu

REL = 1e-10


def check_system(rg, ro, name):
    """CSR pattern bit-exact; values, gradient and right-hand side per entry."""
    assert np.array_equal(rg["outer"], ro["outer"]) and np.array_equal(rg["inner"], ro["inner"]), name
    ev = csr_entry_err(ro["outer"], ro["inner"], rg["val"], ro["val"], scale=ro["val_scale"])
    eg = vec3_entry_err(rg["grad"], ro["grad"], scale=ro["grad_scale"])
    er = vec3_entry_err(rg["rhs"], ro["rhs"], scale=ro["grad_scale"])
    print(f"{name}: per-entry CSR {ev:.2e} gradient {eg:.2e} rhs {er:.2e} (nnz {len(ro['val'])})")
    assert ev <= REL, (name, "CSR", ev)
    assert eg <= REL, (name, "gradient", eg)
    assert er <= REL, (name, "rhs", er)


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(initial=0.0), 1e-300)
    return np.abs(a - b).max(initial=0.0) / scale


@pytest.fixture(scope="module")
def gpu():
    from cipc import gpu as g
    g.load()
    return g


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def random_F(rng, n):
    """Deformation gradients: near rest, compressed, stretched, sheared and inverted."""
    F = np.eye(3) + 0.05 * rng.standard_normal((n, 3, 3))
    k = n // 5
    F[:k] *= rng.uniform(0.5, 0.95, (k, 1, 1))            # compression
    F[k:2 * k] *= rng.uniform(1.05, 1.6, (k, 1, 1))       # stretch
    F[2 * k:3 * k] += 0.6 * rng.standard_normal((k, 3, 3))
    F[3 * k:4 * k, :, 0] *= -1.0                          # inverted
    return F


def test_snh_element_matches_oracle(gpu, O):
    rng = np.random.default_rng(11)
    F = random_F(rng, 500)
    u, r = 4 * 35714.2857 / 3, 142857.142 + 5 * 35714.2857 / 6
    P, H, e = gpu.eval_snh(F, u, r)
    for i in range(len(F)):
        Po, Ho, eo = O.snh_element(F[i], u, r)
        assert rel_err(P[i], Po) <= REL, i
        assert rel_err(H[i], Ho) <= REL, (i, rel_err(H[i], Ho))
        assert abs(e[i] - eo) <= REL * max(abs(eo), u), i


@pytest.mark.parametrize("kind", [1, 2, 3, 4])
def test_distance_derivatives_match_oracle(gpu, O, kind):
    rng = np.random.default_rng(20 + kind)
    pts = rng.standard_normal((300, 12))
    d, g, H = gpu.eval_distance(kind, pts)
    for i in range(len(pts)):
        do, go, Ho = O.distance(kind, pts[i])
        n = go.shape[0]
        assert abs(d[i] - do) <= 1e-12 * max(abs(do), 1e-300) or abs(d[i] - do) <= 1e-14
        assert rel_err(g[i][:n], go) <= REL, (kind, i)
        assert rel_err(H[i], Ho) <= REL, (kind, i, rel_err(H[i], Ho))


@pytest.mark.parametrize("dim", [6, 9, 12])
def test_make_pd_matches_oracle(gpu, O, dim):
    rng = np.random.default_rng(dim)
    A = rng.standard_normal((200, dim, dim))
    H = A + A.transpose(0, 2, 1)
    H[:20] = np.einsum("nij,nkj->nik", A[:20], A[:20])   # already PSD: left unchanged
    Hg = gpu.eval_make_pd(H)
    for i in range(len(H)):
        Ho = O.make_pd(H[i])
        assert rel_err(Hg[i], Ho) <= REL, (dim, i)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_contact_projection_matches_make_pd(gpu, O, kind):
    """The projections the assembly applies to contact Hessians against the oracle's
    makePD of the full block: the general one on the translation-free subspace (all kinds)
    and the structured range-basis one (point-edge, point-triangle, edge-edge)."""
    from contact_cases import barrier_blocks
    H, G = barrier_blocks(O, kind, 300, seed=80 + kind)
    Ho = [O.make_pd(H[i]) for i in range(len(H))]
    Hg = gpu.eval_contact_pd(H)
    for i in range(len(H)):
        assert rel_err(Hg[i], Ho[i]) <= REL, (kind, i, rel_err(Hg[i], Ho[i]))
    if kind >= 1:
        Hs = gpu.eval_contact_pd(H, G)
        for i in range(len(H)):
            assert rel_err(Hs[i], Ho[i]) <= REL, ("structured", kind, i, rel_err(Hs[i], Ho[i]))


def _ccd_inputs(rng, n):
    pts = rng.uniform(-1, 1, (n, 12))
    dirs = rng.uniform(-1, 1, (n, 12)) * rng.uniform(0.01, 3.0, (n, 1))
    dirs[:10] = 0.0                          # no motion -> 1
    dirs[10:20] = np.tile(rng.uniform(-1, 1, 3), 4)  # pure translation -> 1
    return pts, dirs


@pytest.mark.parametrize("kind", [0, 1])
def test_accd_bitexact(gpu, O, kind):
    rng = np.random.default_rng(40 + kind)
    pts, dirs = _ccd_inputs(rng, 400)
    eta = 1.0 - 0.8
    tg = gpu.eval_accd(kind, pts, dirs, eta)
    f = O.accd_pt if kind == 0 else O.accd_ee
    to = np.array([f(pts[i], dirs[i], eta) for i in range(len(pts))])
    assert np.array_equal(tg, to)


@pytest.mark.parametrize("kind", [0, 1])
def test_classification_bitexact(gpu, O, kind):
    rng = np.random.default_rng(60 + kind)
    pts = rng.uniform(-1, 1, (2000, 12))
    pts[:200, 9:] = pts[:200, 6:9] + 1e-9 * rng.standard_normal((200, 3))  # near-degenerate
    pts[200:300, 3:6] = pts[200:300, 0:3] + (pts[200:300, 6:9] - pts[200:300, 0:3]) * 0.5  # collinear-ish
    tg = gpu.eval_classify(kind, pts)
    f = O.dtype_pt if kind == 0 else O.dtype_ee
    to = np.array([f(p) for p in pts])
    assert np.array_equal(tg, to)


def _scenes():
    from cipc import scenes
    return {
        "cube10": scenes.cube10(),
        "two_cubes": scenes.stacked_cubes(2, 6, 0.1, 1e-3, seed=3, rotate="yaw", velocity_scale=0.5, name="two_cubes"),
        "four_20cubes": scenes.four_cubes20(velocity_scale=0.5),
    }


@pytest.fixture(scope="module", params=["cube10", "two_cubes", "four_20cubes"])
def pair(request, gpu, O):
    from cipc import scenes
    s = _scenes()[request.param]
    xt, p = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    g = gpu.GpuContext(s, xtilde=xt)
    o = O.OracleMesh(s, xtilde=xt)
    yield dict(scene=s, g=g, o=o, p=p, kappa=kappa, name=request.param)
    g.close()
    o.close()


def test_active_set_bitexact(pair):
    ag, ao = pair["g"].active_set(), pair["o"].active_set()
    for k in ("self_set", "ee_set", "ee_ie", "ccd_set", "env_set"):
        assert ag[k].shape == ao[k].shape, (pair["name"], k, ag[k].shape, ao[k].shape)
        assert np.array_equal(ag[k], ao[k]), (pair["name"], k)


def test_assembly_matches_oracle(pair):
    g, o, kappa = pair["g"], pair["o"], pair["kappa"]
    g.active_set()
    o.active_set()
    rg, ro = g.assemble(kappa), o.assemble(kappa)
    check_system(rg, ro, pair["name"])
    assert rg["collisions"] == ro["collisions"]


def test_energy_matches_oracle(pair):
    g, o, kappa = pair["g"], pair["o"], pair["kappa"]
    g.active_set()
    o.active_set()
    tg, to = g.energy_terms(kappa), o.energy_terms(kappa)
    for k in range(6):
        assert abs(tg[k] - to[k]) <= REL * max(abs(to[k]), 1e-300) or (tg[k] == to[k]), (pair["name"], k, tg[k], to[k])
    eg, eo = g.energy(kappa), o.energy(kappa)
    assert abs(eg - eo) <= REL * abs(eo)


def test_step_bound_matches_oracle(pair):
    g, o, p = pair["g"], pair["o"], pair["p"]
    g.active_set()
    o.active_set()
    bg, bo = g.step_bound(p), o.step_bound(p)
    for k in ("alpha", "ground", "partial", "cfl", "full"):
        assert bg[k] <= bo[k] and abs(bg[k] - bo[k]) <= 1e-6, (pair["name"], k, bg[k], bo[k])
        assert bg[k] == bo[k], (pair["name"], k, bg[k], bo[k])


def test_step_forward_and_intersection(pair):
    g, o, p = pair["g"], pair["o"], pair["p"]
    s = pair["scene"]
    g.set_direction(p)
    g.save_positions()
    g.step_forward(0.5)
    o.step_forward(s.x, p, 0.5)
    assert np.array_equal(g.positions(), o.positions())
    g.active_set()
    o.active_set()
    assert g.is_intersected() == o.is_intersected()
    g.set_state(x=s.x)
    o.set_state(x=s.x)


def test_is_intersected_when_a_cube_passes_into_another(gpu, O):
    """The positive path of isIntersected (IPCSolver.cpp:476-542, segTriIntersect): the upper
    cube pushed 5 cm into the lower one intersects, on the device and in the oracle; pushed
    by less than the gap it does not."""
    s = _scenes()["two_cubes"]
    g = gpu.GpuContext(s)
    o = O.OracleMesh(s)
    try:
        top = s.x[:, 1] > np.median(s.x[:, 1])
        for depth, expect in ((5e-4, False), (0.05, True)):
            p = np.zeros_like(s.x)
            p[top, 1] = depth
            g.set_state(x=s.x)
            g.set_direction(p)
            g.save_positions()
            g.step_forward(1.0)
            o.step_forward(s.x, p, 1.0)
            assert np.array_equal(g.positions(), o.positions())
            g.active_set()
            o.active_set()
            assert o.is_intersected() == expect
            assert g.is_intersected() == expect, depth
            o.set_state(x=s.x)
    finally:
        g.close()
        o.close()


def test_refitted_trees_after_a_step(pair):
    """After step_forward the device refits trees whose topology was built at the old
    positions; queries still return exactly the reference's sets and step bounds."""
    g, o, p = pair["g"], pair["o"], pair["p"]
    s = pair["scene"]
    g.active_set()
    g.set_direction(p)
    g.save_positions()
    g.step_forward(0.7)
    o.step_forward(s.x, p, 0.7)
    ag, ao = g.active_set(), o.active_set()
    for k in ("self_set", "ee_set", "ee_ie", "ccd_set", "env_set"):
        assert np.array_equal(ag[k], ao[k]), (pair["name"], k)
    bg, bo = g.step_bound(p), o.step_bound(p)
    for k in ("alpha", "ground", "partial", "cfl", "full"):
        assert bg[k] == bo[k], (pair["name"], k, bg[k], bo[k])
    g.set_state(x=s.x)
    o.set_state(x=s.x)


@pytest.mark.parametrize("name", ["two_cubes", "four_20cubes"])
def test_friction_matches_oracle(gpu, O, name):
    """Lagged friction (Friction::initialize, addGradient, addHessian, energy): lagged state,
    assembled CSR and gradient, and energy, from a state whose previous positions are slid
    tangentially so that contacts are both sticking and sliding."""
    from cipc import scenes
    s = _scenes()[name]
    rng = np.random.default_rng(3)
    eps = np.sqrt(s.fhat) * s.material.dt
    vprev = s.x.copy()
    upper = s.x[:, 1] > np.median(s.x[:, 1])
    vprev[upper] -= np.array([1.5 * eps, 0.0, -0.7 * eps])
    vprev += 0.6 * eps * rng.standard_normal(s.x.shape)
    xt, _ = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    g = gpu.GpuContext(s, xtilde=xt)
    o = O.OracleMesh(s, xtilde=xt)
    try:
        g.set_state(V_prev=vprev)
        o.set_state(V_prev=vprev)
        g.active_set()
        o.active_set()
        g.friction_update(kappa)
        o.friction_update(kappa)
        fg, fo = g.friction_state(), o.friction_state()
        assert len(fg["lam"]) == len(fo["lam"]) > 0 and len(fg["env_lam"]) == len(fo["env_lam"])
        for k in ("lam", "coord", "basis", "env_lam"):
            assert rel_err(fg[k], fo[k]) <= REL, (name, k, rel_err(fg[k], fo[k]))
        rg, ro = g.assemble(kappa), o.assemble(kappa)
        check_system(rg, ro, name + " friction")
        tg, to = g.energy_terms(kappa), o.energy_terms(kappa)
        assert to[5] > 0 and abs(tg[5] - to[5]) <= REL * abs(to[5]), (name, tg[5], to[5])
        assert abs(g.energy(kappa) - o.energy(kappa)) <= REL * abs(o.energy(kappa))
    finally:
        g.close()
        o.close()


def test_newton_iteration_matches_oracle(pair):
    """The fused iteration (active set, assembly, step bound on its own stream beside the
    assembly, energy on a second stream): its system, step bound and energy against the
    oracle's stage-by-stage results on the same state."""
    g, o, kappa = pair["g"], pair["o"], pair["kappa"]
    g.set_direction(pair["p"])
    for it in range(2):   # the second call refits the trees the first one built
        t = g.newton_iteration(kappa)
        assert t["total_ms"] > 0 and t["nnz"] > 0
        rg = g.system(t["nnz"])
        o.active_set()
        ro = o.assemble(kappa)
        check_system(rg, ro, f"{pair['name']} iteration {it}")
        assert t["collisions"] == ro["collisions"], (pair["name"], t["collisions"], ro["collisions"])
        bo = o.step_bound(pair["p"])
        # the merged bound and each of its parts bit-exact, as the stand-alone step bound
        assert t["alpha"] == bo["alpha"], (pair["name"], it, t["alpha"], bo["alpha"])
        for k in ("ground", "partial", "cfl", "full"):
            assert t["bound_" + k] == bo[k], (pair["name"], it, k, t["bound_" + k], bo[k])
        assert abs(t["energy"] - o.energy(kappa)) <= REL * abs(t["energy"])


def test_solve_after_newton_iteration(gpu, O):
    """The system a Newton iteration assembles is the one the solve sees, on a fresh
    context with no separate assemble call."""
    from cipc import scenes
    s = _scenes()["two_cubes"]
    xt, p = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    g = gpu.GpuContext(s, xtilde=xt)
    o = O.OracleMesh(s, xtilde=xt)
    try:
        g.set_direction(p)
        t = g.newton_iteration(kappa)
        d, st = g.solve(precision=0, rel_tol=1e-6, preconditioner=1)
        o.active_set()
        ro = o.assemble(kappa)
        assert t["nnz"] == len(ro["val"])
        n = 3 * s.V
        col = np.repeat(np.arange(n), np.diff(ro["outer"]))
        import scipy.sparse as sp
        L = sp.csc_matrix((ro["val"], (ro["inner"], col)), shape=(n, n))
        A = L + L.T - sp.diags(L.diagonal())
        res = np.linalg.norm(ro["rhs"] - A @ d.reshape(-1)) / np.linalg.norm(ro["rhs"])
        assert res <= 1.5e-6, res
    finally:
        g.close()
        o.close()


def test_soft_targets_match_oracle(gpu, O):
    """SoftBoundaryCondition on the device: assembled system and energy against the
    oracle, and the argument checks of validateSoftState."""
    from cipc import scenes
    s = _scenes()["two_cubes"]
    xt, _ = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    rng = np.random.default_rng(9)
    idx = rng.choice(s.V, 40, replace=False)
    tgt = s.x[idx] + 0.02 * rng.standard_normal((40, 3))
    g = gpu.GpuContext(s, xtilde=xt)
    o = O.OracleMesh(s, xtilde=xt)
    try:
        g.set_soft_targets(idx, tgt, 5e2)
        o.set_soft_targets(idx, tgt, 5e2)
        g.active_set()
        o.active_set()
        rg, ro = g.assemble(kappa), o.assemble(kappa)
        check_system(rg, ro, "soft targets")
        tg, to = g.energy_terms(kappa), o.energy_terms(kappa)
        assert to[8] > 0 and abs(tg[8] - to[8]) <= REL * to[8]
        assert abs(g.energy(kappa) - o.energy(kappa)) <= REL * abs(o.energy(kappa))
        with pytest.raises(ValueError):
            g.set_soft_targets([idx[0], idx[0]], tgt[:2], 5e2)      # repeated vertex
        with pytest.raises(ValueError):
            g.set_soft_targets(idx[:2], tgt[:2], 0.0)               # weight must be positive
        g.set_soft_targets([], np.zeros((0, 3)), 0.0)               # removal
        assert g.energy_terms(kappa)[8] == 0.0
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("name", ["two_cubes", "four_20cubes"])
def test_kappa_control_matches_oracle(gpu, O, name):
    """initKappa's sums and postLineSearch's close-constraint records on the device."""
    from cipc import kappa as K
    from cipc import scenes
    s = _scenes()[name]
    xt, p = scenes.newton_probe_state(s)
    g = gpu.GpuContext(s, xtilde=xt)
    o = O.OracleMesh(s, xtilde=xt)
    try:
        g.active_set()
        o.active_set()
        sg, so = g.kappa_gradients(), o.kappa_gradients()
        assert sg[2] == so[2] > 0
        assert abs(sg[0] - so[0]) <= REL * abs(so[0]) and abs(sg[1] - so[1]) <= REL * abs(so[1]), (sg, so)
        kg = K.KappaControl(g, s.hhat, s.bbox_diag2, s.mean_mass, s.dtol)
        ko = K.KappaControl(o, s.hhat, s.bbox_diag2, s.mean_mass, s.dtol)
        assert abs(kg.step_start() - ko.step_start()) <= REL * ko.mesh_kappa
        # records at a threshold that catches part of the constraints, then a step closer
        dtol = s.hhat * 0.5
        assert g.close_constraints(dtol) == o.close_constraints(dtol)
        g.set_direction(p)
        g.save_positions()
        g.step_forward(0.3)
        o.step_forward(s.x, p, 0.3)
        g.active_set()
        o.active_set()
        assert g.close_constraints(dtol) == o.close_constraints(dtol)
        g.clear_close_constraints()
        o.clear_close_constraints()
        assert g.close_constraints(dtol) == o.close_constraints(dtol)
    finally:
        g.close()
        o.close()
