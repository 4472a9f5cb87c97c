"""Pins of the oracle's cloth terms: the Baraff-Witkin triangle (Elasticity.cpp:729-771,
1173-1205,1481-1544), the quadratic bending form (Simulator.cpp:48-109; IPCSolver.cpp:
197-237,333-353) and hinge bending (HingeBending.cpp; Elasticity.cpp:12-150).

What fixes each one independently of its formula:
* PK1 is the derivative of the energy density (central differences);
* each projected Hessian equals the PSD projection (eigenvalues clamped at zero) of the
  exact Hessian, which is the finite-difference Jacobian of the PK1 / gradient;
* edgeTheta's angle equals the textbook dihedral angle atan2(((n0 x n1).e)/|e|, n0.n1),
  and its derivatives match central differences;
* the quadratic form vanishes on flat configurations and on translations, and an
  isometric bend of one hinge by theta costs 3 |e|^2 / (2 (A0 + A1)) * 4 sin^2(theta/2)
  (Bergou et al. 2006, the model the reference implements);
* on a whole cloth scene, the assembled gradient is the derivative of the energy.
"""
import numpy as np
import pytest

from cipc import scenes


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def _psd(H):
    H = 0.5 * (H + H.T)
    w, Z = np.linalg.eigh(H)
    return (Z * np.maximum(w, 0.0)) @ Z.T


def _bw_fd(O, F, st, sh, k, h=1e-6):
    """central differences of the energy (-> PK1) and of the PK1 (-> Hessian)"""
    f = F.T.reshape(-1)      # column-major vec(F)
    P = np.zeros(6)
    H = np.zeros((6, 6))
    for i in range(6):
        e = np.zeros(6)
        e[i] = h
        Fp, Fm = (f + e).reshape(2, 3).T, (f - e).reshape(2, 3).T
        P[i] = (O.bw_element(Fp, st, sh, k)[2] - O.bw_element(Fm, st, sh, k)[2]) / (2 * h)
        H[:, i] = (O.bw_element(Fp, st, sh, k)[0].T.reshape(-1) - O.bw_element(Fm, st, sh, k)[0].T.reshape(-1)) / (2 * h)
    return P, H


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("regime", ["tension", "compression", "mixed"])
def test_bw_pk1_and_projected_hessian(O, seed, regime):
    rng = np.random.default_rng(seed)
    # a rotation of a 2D stretch: columns scaled by (su, sv), sheared a little
    su, sv = {"tension": (1.2, 1.1), "compression": (0.8, 0.9), "mixed": (1.15, 0.85)}[regime]
    A = np.array([[su, 0.1 * rng.standard_normal()], [0.1 * rng.standard_normal(), sv]])
    R, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    F = R[:, :2] @ A
    st, sh, k = 1.1e4, 3.6e3, 1e2
    P, H, e = O.bw_element(F, st, sh, k)
    Pfd, _ = _bw_fd(O, F, st, sh, k)
    assert np.allclose(P.T.reshape(-1), Pfd, rtol=1e-6, atol=1e-6 * np.abs(Pfd).max())
    # each part is the PSD projection of its exact Hessian
    for s1, s2 in ((st, 0.0), (0.0, sh)):
        Hp = O.bw_element(F, s1, s2, k)[1]
        _, Hfd = _bw_fd(O, F, s1, s2, k)
        ref = _psd(Hfd)
        assert np.allclose(Hp, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max()), (s1, s2)
    assert np.allclose(H, O.bw_element(F, st, 0.0, k)[1] + O.bw_element(F, 0.0, sh, k)[1], rtol=1e-12, atol=1e-9)


def test_bw_rest_is_stress_free(O):
    R, _ = np.linalg.qr(np.random.default_rng(3).standard_normal((3, 3)))
    P, H, e = O.bw_element(R[:, :2], 1e4, 1e3, 1e2)
    assert e == pytest.approx(0.0, abs=1e-20) and np.abs(P).max() < 1e-10
    assert np.all(np.linalg.eigvalsh(0.5 * (H + H.T)) > -1e-9)


def _dihedral(q):
    q0, q1, q2, q3 = q
    n0, n1, ax = np.cross(q0 - q2, q1 - q2), np.cross(q1 - q3, q0 - q3), q1 - q0
    return np.arctan2(np.dot(np.cross(n0, n1), ax) / np.linalg.norm(ax), np.dot(n0, n1))


@pytest.mark.parametrize("seed", range(5))
def test_edge_theta_angle_and_derivatives(O, seed):
    q = np.random.default_rng(seed).standard_normal((4, 3))
    th, g, H = O.edge_theta(q)
    assert th == pytest.approx(_dihedral(q), abs=1e-12)
    h = 1e-6
    gfd, Hfd = np.zeros(12), np.zeros((12, 12))
    for i in range(12):
        e = np.zeros(12)
        e[i] = h
        tp, gp, _ = O.edge_theta((q.reshape(-1) + e).reshape(4, 3))
        tm, gm, _ = O.edge_theta((q.reshape(-1) - e).reshape(4, 3))
        gfd[i] = (tp - tm) / (2 * h)
        Hfd[:, i] = (gp - gm) / (2 * h)
    assert np.allclose(g, gfd, rtol=1e-6, atol=1e-7 * np.abs(g).max())
    assert np.allclose(H, Hfd, rtol=1e-5, atol=1e-6 * np.abs(H).max())


@pytest.mark.parametrize("seed", range(4))
def test_hinge_element(O, seed):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal((4, 3))
    rest, w, D = rng.uniform(-2.5, 2.5), rng.uniform(0.3, 2.0), 1e-4
    g, H, e = O.hinge_element(q, rest, w, D)
    th = _dihedral(q)
    diff = np.remainder(th - rest + np.pi, 2 * np.pi) - np.pi
    assert e == pytest.approx(0.5 * D * w * diff * diff, rel=1e-10)
    h = 1e-6
    gfd, Hfd = np.zeros(12), np.zeros((12, 12))
    for i in range(12):
        d = np.zeros(12)
        d[i] = h
        gp, _, ep = O.hinge_element((q.reshape(-1) + d).reshape(4, 3), rest, w, D)
        gm, _, em = O.hinge_element((q.reshape(-1) - d).reshape(4, 3), rest, w, D)
        gfd[i] = (ep - em) / (2 * h)
        Hfd[:, i] = (gp - gm) / (2 * h)
    assert np.allclose(g, gfd, rtol=1e-6, atol=1e-7 * np.abs(g).max())
    ref = _psd(Hfd)
    assert np.allclose(H, ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())


def test_quadratic_bending_form(O):
    s = scenes.cloth_on_sphere(8, sphere_cells=2)
    c = s.cloth
    xr = s.v_rest
    for h in range(c.Nb):
        Q = c.bend_q[h].reshape(4, 4)
        assert np.allclose(Q, Q.T)
        assert np.abs(Q.sum(axis=1)).max() <= 1e-9 * np.abs(Q).max()           # translations
        xl = xr[c.bend_verts[h]]
        assert np.abs(Q @ xl).max() <= 1e-9 * np.abs(Q).max()                   # flat rest
        assert np.linalg.matrix_rank(Q, tol=1e-9 * np.abs(Q).max()) == 1


@pytest.mark.parametrize("edge_len", [1.0, 1.7])
def test_quadratic_bending_isometric_hinge(O, edge_len):
    e0, e1 = np.zeros(3), np.array([edge_len, 0.0, 0.0])
    o1, o2 = np.array([0.3, 0.0, 0.7]), np.array([0.6, 0.0, -0.5])
    x = np.array([o1, e0, e1, o2])
    bv, bq = scenes.prepare_quadratic_bending(x, np.array([[1, 2, 0], [2, 1, 3]]))
    assert bv.tolist() == [[1, 2, 0, 3]]
    Q = bq[0].reshape(4, 4)
    A = 0.5 * np.linalg.norm(np.cross(e1, o1)) + 0.5 * np.linalg.norm(np.cross(e1, o2))
    for th in (1e-3, 0.1, 1.0):
        R = np.array([[1, 0, 0], [0, np.cos(th), -np.sin(th)], [0, np.sin(th), np.cos(th)]])
        xx = x.copy()
        xx[3] = R @ o2
        xl = xx[bv[0]]
        E = 0.5 * np.einsum("ab,ad,bd->", Q, xl, xl)
        assert E == pytest.approx(3 * edge_len ** 2 / (2 * A) * 4 * np.sin(th / 2) ** 2, rel=1e-9)


@pytest.mark.parametrize("bending", [0, 1])
def test_cloth_scene_gradient_is_energy_derivative(O, bending):
    """Whole assembly on a folded cloth over the ball (triangles, bending, self contact
    including mollified edge-edge pairs): d/dh E(x + h v) = g . v."""
    s = scenes.cloth_on_sphere(12, sphere_cells=3, radius=0.2, fold=8, bending=bending)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    m = O.OracleMesh(s)
    a = m.active_set()
    assert len(a["self_set"]) > 0 and len(a["ee_set"]) > 0
    assert not m.is_intersected()
    r = m.assemble(kappa)
    t = m.energy_terms(kappa)
    assert t[6] > 0 and t[7] > 0
    g = r["grad"].reshape(-1)
    rng = np.random.default_rng(2)
    v = rng.standard_normal(g.shape)
    v[np.repeat(s.boundary != 0, 3)] = 0.0
    v /= np.abs(v).max()
    h = 1e-7
    x = s.x
    m.set_state(x=x + h * v.reshape(-1, 3))
    Ep = m.energy(kappa)
    m.set_state(x=x - h * v.reshape(-1, 3))
    Em = m.energy(kappa)
    fd = (Ep - Em) / (2 * h)
    assert fd == pytest.approx(g @ v, rel=1e-5, abs=1e-6 * np.abs(g).sum())


def _loop_quadratic(x, tris):
    """prepareQuadraticBending as the reference loops it (Simulator.cpp:48-109)."""
    first, out = {}, []
    for t in tris:
        for i in range(3):
            e = (int(t[i]), int(t[(i + 1) % 3]))
            other = int(t[(i + 2) % 3])
            e = (min(e), max(e))
            if e not in first:
                first[e] = other
            else:
                out.append((e[0], e[1], first[e], other))
    return np.array(out, np.int64).reshape(-1, 4)


def _loop_hinges(tris):
    """load_triangleMesh's edge map (SimulationMesh.cpp:659-713) as the reference loops it."""
    adj = {}
    for a, b, c in tris:
        for (p, q), o in (((a, b), c), ((b, c), a), ((a, c), b)):
            adj.setdefault((min(p, q), max(p, q)), []).append(int(o))
    return np.array([(k[0], k[1], v[0], v[1]) for k, v in sorted(adj.items()) if len(v) == 2], np.int64).reshape(-1, 4)


@pytest.mark.parametrize("seed", [0, 1])
def test_bending_element_lists_follow_reference_order(seed):
    rng = np.random.default_rng(seed)
    x, tris = scenes.cloth_grid(7, 1.0)
    tris = tris[rng.permutation(len(tris))]
    # rotate each triangle's corners (same orientation, different starting corner)
    r = rng.integers(0, 3, len(tris))
    tris = np.stack([np.roll(t, -k) for t, k in zip(tris, r)])
    x = x + 0.01 * rng.standard_normal(x.shape)
    bv, bq = scenes.prepare_quadratic_bending(x, tris)
    assert np.array_equal(bv, _loop_quadratic(x, tris))
    hv, rest, w = scenes.prepare_hinge_bending(x, tris)
    assert np.array_equal(hv, _loop_hinges(tris))
    for h in range(len(hv)):
        q = x[hv[h]]
        assert rest[h] == pytest.approx(_dihedral(q), abs=1e-12)
        e = q[1] - q[0]
        s2 = np.linalg.norm(np.cross(e, q[2] - q[0])) + np.linalg.norm(np.cross(e, q[3] - q[0]))
        assert w[h] == pytest.approx(e @ e / s2, rel=1e-12)
