"""Learned block preconditioner on the device (learned.hip) against the numpy oracle
(oracle/gnn_oracle.py): training loss and parameter gradient of the same system, weights
and probes (the forward and backward BF16 MFMA kernels), and a training run on Newton
systems the simulator generates, after which the learned factor needs fewer PCG
iterations than the analytic factor it starts from.

Bars: loss within 1e-3 relative; gradient within 3e-2 in relative L2 norm (bfloat16
storage: the device accumulates in float32, the oracle in float64, and values that fall
near a bfloat16 rounding boundary can round to neighbouring values)."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from cipc import gpu as g
    g.load()
    return g


def _system(gpu, s, kappa=None):
    from cipc import scenes
    xt, _ = scenes.newton_probe_state(s)
    k = kappa if kappa is not None else scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    g = gpu.GpuContext(s, xtilde=xt)
    g.active_set()
    r = g.assemble(k)
    return g, r


def _bsr(r):
    n = len(r["outer"]) - 1
    L = sp.csc_matrix((r["val"], r["inner"], r["outer"]), shape=(n, n))
    A = (L + L.T - sp.diags(L.diagonal())).tocsr()
    B = sp.bsr_matrix(A, blocksize=(3, 3))
    B.sort_indices()
    return B.indptr.astype(np.int64), B.indices.astype(np.int64), np.asarray(B.data, np.float64)


def test_param_count_matches_oracle(gpu):
    import gnn_oracle as G
    assert gpu.preconditioner_param_count() == G.n_params()


def test_loss_and_gradient_match_oracle(gpu):
    import gnn_oracle as G
    from cipc import scenes
    s = scenes.stacked_cubes(2, 4, 0.1, 3e-4, seed=3, rotate="yaw", velocity_scale=0.3, name="gnn")
    g, r = _system(gpu, s)
    try:
        rowptr, col, val = _bsr(r)
        graph = G.Graph(rowptr, col)
        rng = np.random.default_rng(0)
        flat = G.init_params(1) + 0.02 * rng.standard_normal(G.n_params())
        flat32 = flat.astype(np.float32)
        Z = rng.standard_normal((2, graph.n, 3)).astype(np.float32)
        lo, go, _, _ = G.loss_and_grad(flat32.astype(np.float64), graph, val, Z.astype(np.float64), use_bf16=True)
        lg, gg = g.preconditioner_gradient(flat32, Z, objective=0)
        assert abs(lg - lo) <= 1e-3 * abs(lo), (lg, lo)
        err = np.linalg.norm(gg - go) / np.linalg.norm(go)
        assert err <= 3e-2, err
        # objective 1: regression on the block FSAI factor
        lo1, go1, _, _ = G.loss_and_grad_fsai(flat32.astype(np.float64), graph, val, use_bf16=True)
        lg1, gg1 = g.preconditioner_gradient(flat32, objective=1)
        assert abs(lg1 - lo1) <= 1e-3 * abs(lo1), (lg1, lo1)
        err1 = np.linalg.norm(gg1 - go1) / np.linalg.norm(go1)
        assert err1 <= 3e-2, err1
    finally:
        g.close()
