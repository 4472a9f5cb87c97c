"""Scene-level pins of the oracle: broad phase against brute force, the reference's own
finite-difference Taylor test of energy against gradient (IPCSolver.cpp:596-631), a
Hessian-vector Taylor test where no projection is active, the CSR layout rules of
NewtonLinearSystem::assemble, and the step bound's guarantees."""
import numpy as np
import pytest

from cipc import scenes


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def small_stack(seed=3, n=4, vel=0.5):
    """Two yawed cubes 3e-4 m apart (below dHat = 1e-3 of the scene diagonal)."""
    return scenes.stacked_cubes(2, n, 0.1, 3e-4, seed=seed, rotate="yaw", velocity_scale=vel, jitter=0.002,
                                name="small")


def test_lattice_counts_match_closed_forms():
    for n in (3, 5, 10):
        s = scenes.stacked_cubes(1, n, 0.1, 0.05, seed=0, rotate="none", jitter=0.0)
        assert s.V == (n + 1) ** 3 and s.T == 5 * n ** 3
        assert len(s.faces) == 12 * n * n and len(s.surf_verts) == 6 * n * n + 2 and len(s.surf_edges) == 18 * n * n
        assert abs(s.volumes.sum() - (0.1 * n) ** 3) < 1e-9 * (0.1 * n) ** 3 * 10


def _tri_boxes(x, faces):
    p = x[faces[:, :3]]
    return np.concatenate([p.min(axis=1), p.max(axis=1)], axis=1)


def _overlap(a, b, gap):
    return ~(np.any(b[..., :3] - a[..., 3:] > gap, axis=-1) | np.any(a[..., :3] - b[..., 3:] > gap, axis=-1))


def test_partial_ccd_pairs_equal_brute_force(O):
    """Self_CCD_ActiveSet = every (surface vertex, face) and (edge, later edge) pair that
    passes the collider/incidence filters and the exact box test with gap sqrt(Hhat),
    in the reference's (query, sorted candidate) order (CollisionBroadPhase.cpp:765-1008)."""
    s = small_stack()
    m = O.OracleMesh(s)
    a = m.active_set()
    sq = np.sqrt(s.hhat)
    fb = _tri_boxes(s.x, s.faces)
    want = []
    for si, v in enumerate(s.surf_verts):
        pb = np.concatenate([s.x[v], s.x[v]])
        ok = _overlap(pb[None], fb, sq)
        for f in np.nonzero(ok)[0]:
            if v in s.faces[f, :3]:
                continue
            want.append((-si - 1, f))
    eb = np.concatenate([np.minimum(s.x[s.surf_edges[:, 0]], s.x[s.surf_edges[:, 1]]),
                         np.maximum(s.x[s.surf_edges[:, 0]], s.x[s.surf_edges[:, 1]])], axis=1)
    for i, (a0, a1) in enumerate(s.surf_edges):
        ok = _overlap(eb[i][None], eb, sq)
        for j in np.nonzero(ok)[0]:
            if j <= i or len({a0, a1} & set(s.surf_edges[j])):
                continue
            want.append((i, j))
    got = [tuple(r) for r in a["ccd_set"]]
    assert got == want


def test_active_constraints_are_closer_than_dhat(O):
    s = small_stack()
    m = O.OracleMesh(s)
    a = m.active_set()
    assert len(a["self_set"]) > 0
    x = s.x
    for c in a["self_set"]:
        if c[0] >= 0:
            d = O.distance(O.DIST_EE, x[c].reshape(-1))[0]
        else:
            v = -c[0] - 1
            if c[2] < 0:
                d = O.distance(O.DIST_PP, np.concatenate([x[v], x[c[1]]]))[0]
            elif c[3] < 0:
                d = O.distance(O.DIST_PE, np.concatenate([x[v], x[c[1]], x[c[2]]]))[0]
            else:
                d = O.distance(O.DIST_PT, np.concatenate([x[v], x[c[1]], x[c[2]], x[c[3]]]))[0]
        assert 0 < d < s.hhat
    # multiplicities are negative counts on the deduplicated point-point/point-edge tail
    tail = a["self_set"][(a["self_set"][:, 0] < 0) & ((a["self_set"][:, 2] < 0) | (a["self_set"][:, 3] < 0))]
    assert np.all(tail[:, 3] <= -1)
    keys = [tuple(r[:3]) for r in tail]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    ground = a["env_set"]
    assert np.all((s.x[ground, 1] + 1.0) ** 2 < s.hhat)


def test_energy_gradient_taylor(O):
    """Central-difference directional derivative of the energy equals g . p with the
    active set held fixed (the reference's --diagnose-line-search check)."""
    s = small_stack(n=5)
    xt, p = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    m = O.OracleMesh(s, xtilde=xt)
    m.active_set()
    r = m.assemble(kappa)
    g = r["grad"]
    assert len(m.active_set()["self_set"]) > 0
    gp = float(np.sum(g * p))
    for h in (1e-7, 1e-8):
        m.step_forward(s.x, p, h)
        ep = m.energy(kappa)
        m.step_forward(s.x, p, -h)
        em = m.energy(kappa)
        m.set_state(x=s.x)
        fd = (em - ep) / (2 * h)
        assert abs(fd - gp) <= 1e-5 * abs(gp), (h, fd, gp)


def test_hessian_vector_taylor_without_projection(O):
    """A uniformly stretched, contact-free cube keeps every SNH eigenvalue positive, so
    the assembled CSR is the exact Hessian: H v = (g(x + h v) - g(x - h v)) / 2h."""
    s = scenes.stacked_cubes(1, 4, 0.1, 0.5, seed=2, rotate="none")
    c = s.x.mean(axis=0)
    x = c + 1.1 * (s.x - c)
    m = O.OracleMesh(s)
    m.set_state(x=x)
    m.active_set()
    r = m.assemble(0.0)
    n = 3 * s.V
    H = np.zeros((n, n))
    for col in range(n):
        for k in range(r["outer"][col], r["outer"][col + 1]):
            H[r["inner"][k], col] = r["val"][k]
            H[col, r["inner"][k]] = r["val"][k]
    v = np.random.default_rng(0).standard_normal(n)
    h = 1e-6
    m.set_state(x=x + h * v.reshape(-1, 3))
    gp = m.assemble(0.0)["grad"].reshape(-1)
    m.set_state(x=x - h * v.reshape(-1, 3))
    gm = m.assemble(0.0)["grad"].reshape(-1)
    fd = (gp - gm) / (2 * h)
    assert np.allclose(H @ v, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())


def test_csr_layout_rules(O):
    """Lower triangle, ascending rows per column, the mass diagonal on every DOF, rows and
    columns of non-free vertices carry only the diagonal, RHS zero there."""
    s = small_stack(n=3)
    s.boundary[:5] = 1
    xt, _ = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    m = O.OracleMesh(s, xtilde=xt)
    m.active_set()
    r = m.assemble(kappa)
    outer, inner, val = r["outer"], r["inner"], r["val"]
    n = 3 * s.V
    assert outer[0] == 0 and outer[-1] == len(inner)
    for col in range(n):
        rows = inner[outer[col]:outer[col + 1]]
        assert rows[0] == col and np.all(np.diff(rows) > 0)
        assert val[outer[col]] > 0
        if s.boundary[col // 3] != 0:
            assert len(rows) == 1
    fixed = np.repeat(s.boundary != 0, 3)
    assert np.all(r["rhs"][fixed] == 0) and np.array_equal(r["rhs"][~fixed], r["grad"].reshape(-1)[~fixed])
    nonfree_rows = fixed[inner] & (inner != np.repeat(np.arange(n), np.diff(outer)))
    assert not np.any(nonfree_rows)


def test_step_bound_leaves_a_feasible_step(O):
    """The merged step never exceeds 1 and x - alpha p is intersection-free with every
    ground gap positive (IPCSolver.cpp:943-990)."""
    s = small_stack(n=4, vel=2.0)
    xt, p = scenes.newton_probe_state(s, scale=20.0)
    m = O.OracleMesh(s, xtilde=xt)
    m.active_set()
    b = m.step_bound(p)
    assert 0 < b["alpha"] <= 1.0
    assert b["ground"] <= 1.0 and b["full"] <= b["ground"] + 0.0 or b["ground"] == 1.0
    m.step_forward(s.x, p, b["alpha"])
    assert not m.is_intersected()
    x = m.positions()
    assert np.all(x[s.surf_verts, 1] > -1.0)


def test_batched_cubes_share_the_batch_scalars():
    """The batched workload: cell counts in [8, 24] per axis, per-scene materials, drops of
    0.02-0.2 m, and the hhat / fhat / dtol that a batch requires to be shared."""
    from cipc import scenes
    b = scenes.batched_cubes(6, seed=2)
    assert len(b) == 6
    for s in b:
        assert 5 * 8 ** 3 <= s.T <= 5 * 24 ** 3
        assert s.hhat == b[0].hhat and s.fhat == b[0].fhat and s.dtol == b[0].dtol
        gap = s.x[:, 1].min() - scenes.GROUND_Y
        assert 0.02 - 1e-12 <= gap <= 0.2 + 1e-12
        assert 1e4 <= s.material.youngs <= 1e6 and 0.3 <= s.material.poisson <= 0.45
    assert len({s.material.youngs for s in b}) == 6


def test_intersection_check_finds_interpenetration(O):
    """isIntersected (IPCSolver.cpp:476-542): false for the stacked cubes as built and after a
    step smaller than their gap; true once the upper cube is pushed 5 cm into the lower one,
    where its vertical side edges cross the lower cube's top faces."""
    s = small_stack()
    o = O.OracleMesh(s)
    try:
        o.active_set()
        assert not o.is_intersected()
        top = s.x[:, 1] > np.median(s.x[:, 1])
        for depth, expect in ((1e-4, False), (0.05, True)):
            p = np.zeros_like(s.x)
            p[top, 1] = depth
            o.step_forward(s.x, p, 1.0)
            o.active_set()
            assert o.is_intersected() == expect, depth
            o.set_state(x=s.x)
    finally:
        o.close()


def test_step_bound_closed_forms(O):
    """The step bound's parts and both branches of its merge (IPCSolver.cpp:945-989) on hand-
    built motions of three axis-aligned, exactly planar cubes stacked 3e-4 m apart (below
    dHat), the lowest 3e-4 m above the ground:
      * all cubes moving down at v: no relative motion, so the CCD bounds equal the ground
        bound, limitStepByGround's closed form 0.8 h / v (h the lowest height above the
        ground; FeasibleStep.cpp:8-40);
      * the middle cube moving down onto the lowest one at v1 and the top one moving up at
        v2 = v1: the face pairs close head-on, so the partial and full CCD are
        (1 - eta) gap / v1 = 0.8 gap / v1 (the ACCD closed form), which is not above twice
        the CFL bound sqrt(hhat) / (2 max|p|): alpha = min(ground, CFL) = CFL;
      * the same with v2 = 1000 v1: the CFL bound shrinks below half the partial CCD, and
        alpha = max(min(partial, full), CFL) = 0.8 gap / v1."""
    s = scenes.stacked_cubes(3, 3, 0.1, 3e-4, seed=2, rotate="none", jitter=0.0, name="three")
    n = s.V // 3
    ids = [np.arange(k * n, (k + 1) * n) for k in range(3)]
    gap = s.x[ids[1], 1].min() - s.x[ids[0], 1].max()
    sq = np.sqrt(s.hhat)
    o = O.OracleMesh(s)
    try:
        o.active_set()
        v = 1.0
        p = np.zeros_like(s.x)
        p[:, 1] = v
        b = o.step_bound(p)
        h = s.x[:, 1].min() + 1.0
        assert b["ground"] == h / v * 0.8 and b["alpha"] == b["ground"]
        assert b["partial"] == b["ground"] and b["full"] == b["ground"]
        p[:, 1] = -v                                     # moving up: the ground does not limit
        assert o.step_bound(p)["ground"] == 1.0
        v1 = 1e-3
        for v2, branch in ((v1, "cfl"), (1000 * v1, "ccd")):
            p = np.zeros_like(s.x)
            p[ids[1], 1] = v1
            p[ids[2], 1] = -v2
            b = o.step_bound(p)
            head_on = 0.8 * gap / v1
            cfl = sq / (max(v1, v2) * 2.0)
            assert b["ground"] == 1.0
            assert abs(b["partial"] - head_on) <= 1e-12 * head_on and abs(b["full"] - head_on) <= 1e-12 * head_on
            assert b["cfl"] == cfl
            if branch == "cfl":
                assert b["partial"] <= 2 * cfl and b["alpha"] == cfl
            else:
                assert b["partial"] > 2 * cfl and b["alpha"] == min(b["partial"], b["full"])
    finally:
        o.close()
