"""Pins the oracle's element-level functions against what mathematics fixes: numpy's
SVD and symmetric eigensolver, central finite differences, closed forms and invariants.
Each check is chosen so a dropped term, a sign or index slip, or a transposed operand
fails it."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def random_F(rng, n):
    F = np.eye(3) + 0.05 * rng.standard_normal((n, 3, 3))
    k = n // 5
    F[:k] *= rng.uniform(0.5, 0.95, (k, 1, 1))
    F[k:2 * k] *= rng.uniform(1.05, 1.6, (k, 1, 1))
    F[2 * k:3 * k] += 0.6 * rng.standard_normal((k, 3, 3))
    F[3 * k:4 * k, :, 0] *= -1.0
    return F


# ---------------------------------------------------------------------------------------
def test_svd_conventions_against_numpy(O):
    """RotationAwareSVD.h:11-13: U, V rotations, |s| descending, only s2 negative,
    sign(s2) = sign(det F); singular values equal numpy's."""
    rng = np.random.default_rng(1)
    for F in random_F(rng, 300):
        U, s, V = O.svd3(F)
        assert np.allclose(U @ U.T, np.eye(3), atol=1e-13) and np.allclose(V @ V.T, np.eye(3), atol=1e-13)
        assert np.linalg.det(U) > 0 and np.linalg.det(V) > 0
        assert np.allclose(U @ np.diag(s) @ V.T, F, atol=1e-13 * np.abs(F).max())
        assert s[0] >= s[1] >= abs(s[2]) and s[1] >= 0
        assert np.sign(s[2]) == np.sign(np.linalg.det(F))
        assert np.allclose(np.abs(s), np.linalg.svd(F, compute_uv=False), rtol=1e-13, atol=1e-15)


def test_svd_degenerate_inputs(O):
    for F in (np.zeros((3, 3)), np.diag([2.0, 0.0, 0.0]), np.diag([1.0, 1.0, 0.0]), -np.eye(3),
              np.outer([1, 2, 3], [0.5, -1, 2.0])):
        U, s, V = O.svd3(F)
        assert np.allclose(U @ np.diag(s) @ V.T, F, atol=1e-13)
        assert abs(np.linalg.det(U) - 1) < 1e-12 and abs(np.linalg.det(V) - 1) < 1e-12


def _numpy_make_pd(H):
    w, Z = np.linalg.eigh(H)
    if w[0] >= 0:
        return H
    return Z @ np.diag(np.maximum(w, 0)) @ Z.T


@pytest.mark.parametrize("n", [3, 6, 9, 12])
def test_make_pd_against_numpy_eigh(O, n):
    """ContactMechanics.h:27-45: unchanged when PSD, else negative eigenvalues zeroed."""
    rng = np.random.default_rng(n)
    for k in range(100):
        A = rng.standard_normal((n, n))
        H = A + A.T if k % 4 else A @ A.T
        R = O.make_pd(H)
        assert np.allclose(R, _numpy_make_pd(H), atol=1e-12 * np.abs(H).max())
        assert np.linalg.eigvalsh(R)[0] >= -1e-12 * np.abs(H).max()
        if k % 4 == 0:
            assert np.array_equal(R, H)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_contact_make_pd_on_translation_free_subspace(O, kind):
    """Contact Hessians annihilate rigid translations, so makePD(H) equals
    H - B^T (sum over negative eigenpairs of R = B H B^T of lambda z z^T) B with B the
    Helmert basis (x) I3 of the translation-free subspace; the device projects contacts
    this way (the 12x12 solve becomes 9x9)."""
    from contact_cases import NVERTS, barrier_hessians
    m = NVERTS[kind]
    hm = np.zeros((m - 1, m))
    for i in range(m - 1):
        hm[i, :i + 1] = 1.0 / np.sqrt((i + 1) * (i + 2))
        hm[i, i + 1] = -(i + 1) / np.sqrt((i + 1) * (i + 2))
    B = np.kron(hm, np.eye(3))
    assert np.allclose(B @ B.T, np.eye(3 * m - 3)) and np.allclose(B @ np.tile(np.eye(3), (m, 1)), 0)
    npd = 0
    for H in barrier_hessians(O, kind, 60, seed=70 + kind):
        assert np.abs(H @ np.tile(np.eye(3), (m, 1))).max() <= 1e-12 * np.abs(H).max()
        w, Z = np.linalg.eigh(B @ H @ B.T)
        P = H - B.T @ (Z @ np.diag(np.minimum(w, 0)) @ Z.T) @ B if w[0] < 0 else H
        npd += w[0] >= 0
        assert np.allclose(P, O.make_pd(H), atol=1e-12 * np.abs(H).max())
    assert npd < 60   # the projection path is exercised


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_structured_contact_blocks_live_on_a_small_range(O, kind):
    """Point-edge, point-triangle and edge-edge barrier blocks map into a 4 / 5 / 5
    dimensional space built from the rank-one factors c (x) u of their own gradient:
    {w (x) u : sum w = 0} + {c (x) t : t perpendicular to u} (point-triangle, edge-edge) and
    {c (x) t : t in R^3} + {(0, 1, -1) (x) u} (point-edge).  makePD of the block equals the
    projection computed from the eigenpairs of the block restricted to that space."""
    from contact_cases import NVERTS, barrier_blocks
    m = NVERTS[kind]
    H, G = barrier_blocks(O, kind, 200, seed=90 + kind)
    ones = np.ones(m) / np.sqrt(m)
    for i in range(len(H)):
        Gm = G[i].reshape(m, 3)
        u = Gm[np.argmax(np.sum(Gm * Gm, axis=1))]
        u = u / np.linalg.norm(u)
        c = Gm @ u
        c = c / np.linalg.norm(c)
        assert abs(c @ ones) < 1e-12
        t = np.linalg.svd(u[None])[2][1:]           # two unit vectors perpendicular to u
        if m == 4:
            w = np.linalg.svd(ones[None])[2][1:]    # orthonormal basis of sum-zero weights
            cols = [np.kron(wi, u) for wi in w] + [np.kron(c, ti) for ti in t]
        else:
            dv = np.array([0.0, 1.0, -1.0]) / np.sqrt(2)
            dv = dv - (dv @ c) * c
            cols = [np.kron(c, e) for e in np.eye(3)] + [np.kron(dv / np.linalg.norm(dv), u)]
        Q = np.stack(cols, axis=1)
        assert np.allclose(Q.T @ Q, np.eye(Q.shape[1]), atol=1e-12)
        P = Q @ Q.T
        scale = np.abs(H[i]).max()
        assert np.abs(H[i] - P @ H[i] @ P).max() <= 1e-12 * scale
        lam, Y = np.linalg.eigh(Q.T @ H[i] @ Q)
        Z = Q @ Y
        proj = H[i] - Z @ np.diag(np.minimum(lam, 0)) @ Z.T if lam[0] < 0 else H[i]
        assert np.allclose(proj, O.make_pd(H[i]), atol=1e-11 * scale)


# ---------------------------------------------------------------------------------------
U_RATE, R_RATE = 4 * 35714.2857 / 3, 142857.142 + 5 * 35714.2857 / 6


def test_snh_rest_state_is_stress_free(O):
    P, H, e = O.snh_element(np.eye(3), U_RATE, R_RATE)
    assert np.abs(P).max() < 1e-9 * U_RATE


def test_snh_pk1_is_energy_gradient(O):
    """Elasticity.cpp:1144-1170 against central differences of Elasticity.cpp:695-727."""
    rng = np.random.default_rng(3)
    h = 1e-6
    for F in random_F(rng, 60):
        P, _, _ = O.snh_element(F, U_RATE, R_RATE)
        fd = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                Fp, Fm = F.copy(), F.copy()
                Fp[i, j] += h
                Fm[i, j] -= h
                fd[i, j] = (O.snh_element(Fp, U_RATE, R_RATE)[2] - O.snh_element(Fm, U_RATE, R_RATE)[2]) / (2 * h)
        assert np.allclose(P, fd, rtol=1e-6, atol=1e-6 * np.abs(P).max())


def test_snh_projected_hessian_is_projected_pk1_jacobian(O):
    """Hessian of Elasticity.cpp:1436-1477 (finite differences of the PK1 over column-major
    vec F) followed by numpy's eigen-projection equals the oracle's projected Hessian."""
    rng = np.random.default_rng(4)
    h = 1e-6
    for F in random_F(rng, 40):
        _, H, _ = O.snh_element(F, U_RATE, R_RATE)
        J = np.zeros((9, 9))
        for q in range(9):
            i, j = q % 3, q // 3
            Fp, Fm = F.copy(), F.copy()
            Fp[i, j] += h
            Fm[i, j] -= h
            dP = (O.snh_element(Fp, U_RATE, R_RATE)[0] - O.snh_element(Fm, U_RATE, R_RATE)[0]) / (2 * h)
            J[:, q] = dP.T.reshape(9)
        J = 0.5 * (J + J.T)
        assert np.allclose(H, _numpy_make_pd(J), atol=2e-5 * np.abs(J).max())


def test_snh_energy_is_rotation_invariant(O):
    rng = np.random.default_rng(5)
    for F in random_F(rng, 30):
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        if np.linalg.det(Q) < 0:
            Q[:, 0] *= -1
        e1 = O.snh_element(F, U_RATE, R_RATE)[2]
        e2 = O.snh_element(Q @ F, U_RATE, R_RATE)[2]
        assert abs(e1 - e2) <= 1e-12 * max(abs(e1), U_RATE)


# ---------------------------------------------------------------------------------------
def test_barrier_closed_forms(O):
    """ContactMechanics.cpp:775-790: b(dHat) = b'(dHat) = 0, b' and b'' are derivatives."""
    dh = 2.3e-4
    b, g, H = O.barrier(dh, dh)
    assert b == 0.0 and g == 0.0
    for d in np.geomspace(1e-9, 0.99 * dh, 25):
        b, g, H = O.barrier(d, dh)
        assert b > 0 and g < 0 and H > 0
        e = 1e-7 * d
        assert abs((O.barrier(d + e, dh)[0] - O.barrier(d - e, dh)[0]) / (2 * e) - g) <= 1e-5 * abs(g)
        assert abs((O.barrier(d + e, dh)[1] - O.barrier(d - e, dh)[1]) / (2 * e) - H) <= 1e-5 * abs(H)


def _closest_pt_brute(p, t0, t1, t2, n=400):
    u, v = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n))
    m = (u + v) <= 1
    q = t0[None] + u[m][:, None] * (t1 - t0) + v[m][:, None] * (t2 - t0)
    return np.min(np.sum((q - p) ** 2, axis=1))


def _closest_ee_brute(a0, a1, b0, b1, n=600):
    s = np.linspace(0, 1, n)
    A = a0[None] + s[:, None] * (a1 - a0)
    B = b0[None] + s[:, None] * (b1 - b0)
    return np.min(np.sum((A[:, None] - B[None]) ** 2, axis=2))


def test_point_triangle_distance_and_type_against_brute_force(O):
    """dType_PT + d_PP/d_PE/d_PT (ContactMechanics.cpp:143-306) give the true squared
    distance of the point to the triangle (dense sampling bound)."""
    rng = np.random.default_rng(6)
    for _ in range(60):
        pts = rng.uniform(-1, 1, (4, 3))
        t = O.dtype_pt(pts)
        p, a, b, c = pts
        if t in (0, 1, 2):
            d = O.distance(O.DIST_PP, np.concatenate([p, (a, b, c)[t]]))[0]
        elif t in (3, 4, 5):
            e = ((a, b), (b, c), (c, a))[t - 3]
            d = O.distance(O.DIST_PE, np.concatenate([p, e[0], e[1]]))[0]
        else:
            d = O.distance(O.DIST_PT, pts.reshape(-1))[0]
        brute = _closest_pt_brute(p, a, b, c)
        assert d <= brute + 1e-12 and brute - d <= 2e-2 * max(brute, 1e-3), (t, d, brute)


def test_edge_edge_distance_and_type_against_brute_force(O):
    rng = np.random.default_rng(7)
    for _ in range(60):
        pts = rng.uniform(-1, 1, (4, 3))
        t = O.dtype_ee(pts)
        a0, a1, b0, b1 = pts
        pp = {0: (a0, b0), 1: (a0, b1), 3: (a1, b0), 4: (a1, b1)}
        pe = {2: (a0, b0, b1), 5: (a1, b0, b1), 6: (b0, a0, a1), 7: (b1, a0, a1)}
        if t in pp:
            d = O.distance(O.DIST_PP, np.concatenate(pp[t]))[0]
        elif t in pe:
            d = O.distance(O.DIST_PE, np.concatenate(pe[t]))[0]
        else:
            d = O.distance(O.DIST_EE, pts.reshape(-1))[0]
        brute = _closest_ee_brute(a0, a1, b0, b1)
        assert d <= brute + 1e-12 and brute - d <= 2e-2 * max(brute, 1e-3), (t, d, brute)


@pytest.mark.parametrize("kind,n", [(0, 6), (1, 9), (2, 12), (3, 12), (4, 12)])
def test_distance_derivatives_finite_differences(O, kind, n):
    rng = np.random.default_rng(8 + kind)
    h = 1e-6
    for _ in range(20):
        x = rng.standard_normal(12)
        d, g, H = O.distance(kind, x)
        for j in range(n):
            xp, xm = x.copy(), x.copy()
            xp[j] += h
            xm[j] -= h
            dp, gp, _ = O.distance(kind, xp)
            dm, gm, _ = O.distance(kind, xm)
            assert abs((dp - dm) / (2 * h) - g[j]) <= 1e-6 * max(1.0, np.abs(g).max())
            assert np.allclose((gp - gm) / (2 * h), H[:, j], atol=1e-5 * max(1.0, np.abs(H).max()))


# ---------------------------------------------------------------------------------------
def _ccd_config(rng):
    pts = rng.uniform(-1, 1, (4, 3))
    dirs = rng.uniform(-1, 1, (4, 3)) * rng.uniform(0.05, 2.0)
    return pts, dirs


def test_accd_trivial_motions_return_one(O):
    rng = np.random.default_rng(9)
    for _ in range(20):
        pts, dirs = _ccd_config(rng)
        assert O.accd_pt(pts, np.zeros((4, 3)), 0.2) == 1.0
        assert O.accd_ee(pts, np.tile(dirs[0], (4, 1)), 0.2) == 1.0


@pytest.mark.parametrize("kind", ["pt", "ee"])
def test_accd_is_conservative(O, kind):
    """AdditiveCCD.cpp:221-329: moving by toc * dir leaves a positive gap, at least about
    eta times the starting one, never crossing a contact a dense time sampling finds."""
    rng = np.random.default_rng(10)
    f = O.accd_pt if kind == "pt" else O.accd_ee
    dist = (lambda q: _closest_pt_brute(q[0], q[1], q[2], q[3], 120)) if kind == "pt" else \
           (lambda q: _closest_ee_brute(q[0], q[1], q[2], q[3], 200))
    for _ in range(25):
        pts, dirs = _ccd_config(rng)
        toc = f(pts, dirs, 0.2)
        assert 0.0 <= toc <= 1.0
        for s in np.linspace(0, toc, 8):
            assert dist(pts + s * dirs) > 0.0


def _random_frame(rng):
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    return q * np.sign(np.linalg.det(q))


@pytest.mark.parametrize("eta", [0.2, 0.5])
def test_accd_point_triangle_head_on_closed_form(O, eta):
    """Additive CCD (AdditiveCCD.cpp:281-329; Li et al. 2021, Codimensional IPC, Alg. 1) for
    a point moving along the triangle's normal onto an interior point: the motion bound is
    the relative speed, the first advance covers (1 - eta) of the gap and leaves eta of it,
    and the next one ends the loop, so toc = (1 - eta) t*, t* the time of impact, or 1 once
    that exceeds 1.  Moving the triangle instead of, or as well as, the point gives the same
    toc for the same relative motion.  A step of half, or twice, the right size fails."""
    rng = np.random.default_rng(int(eta * 100))
    for trial in range(40):
        R = _random_frame(rng)
        tri = np.array([[-0.6, 0.0, -0.5], [0.7, 0.0, -0.4], [0.0, 0.0, 0.8]])
        foot = np.array([rng.uniform(-0.1, 0.1), 0.0, rng.uniform(-0.1, 0.1)])
        d0 = rng.uniform(0.01, 0.5)
        speed = rng.uniform(0.05, 2.0)
        tstar = d0 / speed
        share = [1.0, 0.0, 0.5][trial % 3]           # part of the approach done by the point
        pts = np.vstack([foot + [0.0, d0, 0.0], tri]) @ R.T + rng.uniform(-2, 2, 3)
        dp = np.array([0.0, -speed * share, 0.0]) @ R.T
        dt = np.array([0.0, speed * (1.0 - share), 0.0]) @ R.T
        dirs = np.vstack([dp, dt, dt, dt])
        toc = O.accd_pt(pts, dirs, eta)
        want = min(1.0, (1.0 - eta) * tstar)
        assert toc <= tstar
        assert abs(toc - want) <= 1e-12 * max(want, 1e-300) + 1e-15, (trial, toc, want, tstar)


@pytest.mark.parametrize("eta", [0.2, 0.5])
def test_accd_edge_edge_crossing_closed_form(O, eta):
    """Additive CCD for two edges crossing perpendicularly (AdditiveCCD.cpp:221-279): one
    edge approaches the other along their common normal, closest points interior to both;
    toc = (1 - eta) t*, or 1 once that exceeds 1, with either or both edges moving."""
    rng = np.random.default_rng(7 + int(eta * 100))
    for trial in range(40):
        R = _random_frame(rng)
        d0 = rng.uniform(0.01, 0.5)
        speed = rng.uniform(0.05, 2.0)
        tstar = d0 / speed
        a = rng.uniform(-0.2, 0.2, 2)
        ea = np.array([[-1.0, d0, a[0]], [1.0, d0, a[0]]])     # along x, above
        eb = np.array([[a[1], 0.0, -1.0], [a[1], 0.0, 1.0]])   # along z, below
        share = [1.0, 0.0, 0.5][trial % 3]
        pts = np.vstack([ea, eb]) @ R.T + rng.uniform(-2, 2, 3)
        da = np.array([0.0, -speed * share, 0.0]) @ R.T
        db = np.array([0.0, speed * (1.0 - share), 0.0]) @ R.T
        dirs = np.vstack([da, da, db, db])
        toc = O.accd_ee(pts, dirs, eta)
        want = min(1.0, (1.0 - eta) * tstar)
        assert toc <= tstar
        assert abs(toc - want) <= 1e-12 * max(want, 1e-300) + 1e-15, (trial, toc, want, tstar)


def test_accd_sliding_motion_returns_one(O):
    """A point sliding parallel to a triangle's plane, and an edge sliding along its own
    direction past another, never approach: the loop runs to toc > 1 and returns 1."""
    tri = np.array([[-5.0, 0.0, -5.0], [5.0, 0.0, -5.0], [0.0, 0.0, 5.0]])
    pts = np.vstack([[0.0, 0.1, 0.0], tri])
    dirs = np.vstack([[0.3, 0.0, 0.2], np.zeros((3, 3))])
    assert O.accd_pt(pts, dirs, 0.2) == 1.0
    pts = np.array([[-1.0, 0.1, 0.0], [1.0, 0.1, 0.0], [0.0, 0.0, -1.0], [0.0, 0.0, 1.0]])
    dirs = np.array([[0.5, 0.0, 0.0], [0.5, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    assert O.accd_ee(pts, dirs, 0.2) == 1.0


def test_seg_tri_intersection_closed_form(O):
    """IglUtils::segTriIntersect (ContactMechanics.h:47-71): a segment through the
    interior crosses, one beside the triangle or short of the plane does not."""
    t = [np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])]
    cross = np.concatenate([[0.2, 0.2, -1], [0.2, 0.2, 1], *t])
    miss = np.concatenate([[0.8, 0.8, -1], [0.8, 0.8, 1], *t])
    short = np.concatenate([[0.2, 0.2, 0.5], [0.2, 0.2, 1], *t])
    assert O.seg_tri_intersect(cross) and not O.seg_tri_intersect(miss) and not O.seg_tri_intersect(short)


# ---------------------------------------------------------------------------------------
def _brute_overlap(boxes, q, gap):
    lo, hi = boxes[:, :3], boxes[:, 3:]
    ok = ~(np.any(lo - q[3:] > gap, axis=1) | np.any(q[:3] - hi > gap, axis=1))
    return np.nonzero(ok)[0]


@pytest.mark.parametrize("case", ["random", "duplicates", "flat", "single", "empty"])
def test_lbvh_query_equals_brute_force(O, case):
    """The reference's deleted LBVH tests (agent_docs/09_optimization_report.md:94):
    random boxes, duplicate centres, a degenerate axis, one leaf and an empty tree."""
    rng = np.random.default_rng(11)
    n = {"random": 257, "duplicates": 200, "flat": 300, "single": 1, "empty": 0}[case]
    c = rng.uniform(-1, 1, (n, 3))
    if case == "duplicates":
        c[::2] = c[1::2][: len(c[::2])]
    if case == "flat":
        c[:, 2] = 0.25
    ext = rng.uniform(0, 0.1, (n, 3))
    boxes = np.concatenate([c - ext, c + ext], axis=1)
    for _ in range(50):
        q0 = rng.uniform(-1, 1, 3)
        q = np.concatenate([q0 - 0.05, q0 + 0.05])
        gap = rng.choice([0.0, 0.02])
        got = O.lbvh_query(boxes, q, gap) if n else np.zeros(0, np.int32)
        assert np.array_equal(got, _brute_overlap(boxes, q, gap) if n else np.zeros(0, np.int64))
