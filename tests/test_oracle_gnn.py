"""Pins of the learned-preconditioner oracle (oracle/gnn_oracle.py): its loss gradient
against central differences of its own loss (in exact float64, without the bfloat16
storage roundings), the preconditioner's algebra (G^ block lower-triangular with positive
diagonal, so M^-1 = S G^T G^ S is symmetric positive definite), and the analytic base it
starts from: B_ii A^_ii B_ii^T = I (block Jacobi) and B_ij = -w B_ii A^_ij A^_jj^-1."""
import numpy as np
import pytest
import scipy.sparse as sp

from cipc import scenes


@pytest.fixture(scope="module")
def G():
    import gnn_oracle
    return gnn_oracle


def newton_bsr(O, scene, kappa=None):
    """Full symmetric 3x3-block BSR (rowptr, col, blocks) of the oracle's Newton system."""
    xt, _ = scenes.newton_probe_state(scene)
    m = O.OracleMesh(scene, xtilde=xt)
    m.active_set()
    k = kappa if kappa is not None else scenes.suggest_kappa(scene.hhat, scene.bbox_diag2, scene.mean_mass)
    r = m.assemble(k)
    n = 3 * scene.V
    L = sp.csc_matrix((r["val"], r["inner"], r["outer"]), shape=(n, n))
    A = (L + L.T - sp.diags(L.diagonal())).tocsr()
    B = sp.bsr_matrix(A, blocksize=(3, 3))
    B.sort_indices()
    m.close()
    return B.indptr.astype(np.int64), B.indices.astype(np.int64), np.asarray(B.data, np.float64), r["rhs"]


@pytest.fixture(scope="module")
def system(oracle_lib):
    s = scenes.stacked_cubes(2, 3, 0.1, 3e-4, seed=3, rotate="yaw", velocity_scale=0.3, name="gnn")
    return newton_bsr(oracle_lib, s)


def test_gradient_matches_finite_differences(G, system):
    rowptr, col, val, _ = system
    g = G.Graph(rowptr, col)
    rng = np.random.default_rng(0)
    flat = G.init_params(1)
    flat += 0.05 * rng.standard_normal(flat.shape)      # non-zero decoders: every path active
    Z = rng.standard_normal((2, g.n, 3))
    L, grad, _, _ = G.loss_and_grad(flat, g, val, Z, use_bf16=False)
    idx = rng.choice(len(flat), 60, replace=False)
    h = 1e-6
    for k in idx:
        e = np.zeros_like(flat)
        e[k] = h
        Lp = G.loss_and_grad(flat + e, g, val, Z, use_bf16=False)[0]
        Lm = G.loss_and_grad(flat - e, g, val, Z, use_bf16=False)[0]
        fd = (Lp - Lm) / (2 * h)
        assert fd == pytest.approx(grad[k], rel=1e-5, abs=1e-8 * np.abs(grad).max()), k


def test_preconditioner_algebra(G, system):
    rowptr, col, val, _ = system
    g = G.Graph(rowptr, col)
    rng = np.random.default_rng(1)
    flat = G.init_params(2) + 0.002 * rng.standard_normal(G.n_params())
    p = G.unpack(flat)
    X, E, d = G.features(g, val)
    An = G.normalized_blocks(g, val, d)
    Bd, Bo = G.base_factor(g, An)
    Gd, Go, _ = G.forward(p, g, X, E, Bd, Bo)
    n = g.n
    M = np.zeros((3 * n, 3 * n))
    for i in range(n):
        M[3 * i:3 * i + 3, 3 * i:3 * i + 3] = Gd[i]
        assert Gd[i][0, 1] == Gd[i][0, 2] == Gd[i][1, 2] == 0.0 and np.all(np.diag(Gd[i]) > 0)
    for k, e in enumerate(g.low):
        i, j = g.src[e], g.dst[e]
        assert i > j
        M[3 * i:3 * i + 3, 3 * j:3 * j + 3] = Go[k]
    # M^-1 = S G^T G S is symmetric positive definite iff G^ is nonsingular: block
    # lower-triangular with lower-triangular diagonal blocks of positive diagonal
    assert np.all(np.diag(M) > 0) and np.array_equal(np.triu(M, 1), np.zeros_like(M))
    w = np.linalg.eigvalsh(M.T @ M)
    assert w.min() > 1e-8 * w.max()
    # zero decoders: the analytic base
    p0 = G.unpack(G.init_params(3))
    Gd0, Go0, _ = G.forward(p0, g, X, E, Bd, Bo)
    assert np.array_equal(Gd0, Bd) and np.array_equal(Go0, Bo)
    Dii = An[g.diag]
    assert np.allclose(np.einsum("nab,nbc,ndc->nad", Bd, Dii, Bd), np.eye(3)[None], atol=1e-10)
    k = 5
    e, i, j = g.eid[g.low[k]], g.src[g.low[k]], g.dst[g.low[k]]
    assert np.allclose(Bo[k], -G.OMEGA * Bd[i] @ An[e] @ np.linalg.inv(Dii[j]))


def test_bf16_rounding():
    import gnn_oracle as G
    # bfloat16 spacing at 1 is 2^-7
    x = np.array([1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -3.14159, 0.0])
    r = G.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 + 2 ** -7          # representable: unchanged
    assert r[2] == 1.0                                     # tie: to the even neighbour (1.0)
    assert r[3] == 1.0 + 2 ** -6                           # tie: to the even neighbour (up)
    assert r[4] == 1.0                                     # below half a spacing: down
    assert abs(r[5] - x[5]) <= 2 ** -8 * abs(x[5]) and r[6] == 0.0


def test_product_layout_matches_oracle():
    import gnn_oracle as G
    from cipc import gnn_layout
    assert gnn_layout.param_shapes() == G.param_shapes() and gnn_layout.n_params() == G.n_params()


def test_fsai_objective_gradient_matches_finite_differences(G, system):
    rowptr, col, val, _ = system
    g = G.Graph(rowptr, col)
    rng = np.random.default_rng(4)
    flat = G.init_params(1) + 0.05 * rng.standard_normal(G.n_params())
    L, grad, _, _ = G.loss_and_grad_fsai(flat, g, val, use_bf16=False)
    h = 1e-6
    for k in rng.choice(len(flat), 30, replace=False):
        e = np.zeros_like(flat)
        e[k] = h
        fd = (G.loss_and_grad_fsai(flat + e, g, val, use_bf16=False)[0] - G.loss_and_grad_fsai(flat - e, g, val, use_bf16=False)[0]) / (2 * h)
        assert fd == pytest.approx(grad[k], rel=1e-5, abs=1e-8 * np.abs(grad).max()), k
