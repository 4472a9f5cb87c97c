"""Mathematical pins of the block FSAI oracle (oracle/fsai_oracle.py), the factorized sparse
approximate inverse of Kolotilina & Yeremin that the device builds as a PCG preconditioner.

For a symmetric positive definite A and a lower block pattern P_i per block row, the factor
G has G_{i,P_i} = the last block row of L_i^-1 with A[P_i,P_i] = L_i L_i^T.  Hence
  * the diagonal 3x3 blocks of G A G^T are the identity, whatever the pattern;
  * with the full lower pattern, G = L^-1 for A = L L^T (the leading principal submatrices
    share the leading part of the Cholesky factor), so G A G^T = I;
  * G is block lower triangular with its pattern inside the pattern of A, and where a row
    has more than 23 lower blocks it keeps the 23 largest by Frobenius norm (ties to the
    smaller column) and the diagonal.
A factor scaled by any constant other than 1, or with one misplaced row, fails these.
"""
import numpy as np
import scipy.sparse as sp

import fsai_oracle as F


def _bsr(A):
    B = sp.bsr_matrix(A, blocksize=(3, 3))
    B.sort_indices()
    return B.indptr, B.indices, np.asarray(B.data)


def _spd_dense(rng, nb):
    M = rng.standard_normal((3 * nb, 3 * nb))
    return M @ M.T + 3 * nb * np.eye(3 * nb)


def _spd_sparse(rng, nb, density):
    """Block-sparse SPD: random symmetric block pattern, diagonally dominant blocks."""
    mask = np.triu(rng.random((nb, nb)) < density, 1)
    mask = mask | mask.T
    A = np.zeros((3 * nb, 3 * nb))
    for i in range(nb):
        for j in range(nb):
            if mask[i, j] and j < i:
                blk = rng.standard_normal((3, 3))
                A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blk
                A[3 * j:3 * j + 3, 3 * i:3 * i + 3] = blk.T
    A += (np.abs(A).sum(axis=1).max() + 1.0) * np.eye(3 * nb)
    return A


def test_full_pattern_gives_the_inverse_cholesky_factor():
    rng = np.random.default_rng(0)
    for nb in (1, 2, 5, 8):
        A = _spd_dense(rng, nb)
        G = F.block_fsai(*_bsr(A)).toarray()
        Linv = np.linalg.inv(np.linalg.cholesky(A))
        assert np.allclose(G, Linv, rtol=1e-10, atol=1e-12 * np.abs(Linv).max()), nb
        assert np.allclose(G @ A @ G.T, np.eye(3 * nb), atol=1e-10)


def test_diagonal_blocks_of_the_preconditioned_matrix_are_identity():
    rng = np.random.default_rng(1)
    for nb, dens in ((12, 0.2), (20, 0.35), (40, 0.1)):
        A = _spd_sparse(rng, nb, dens)
        G = F.block_fsai(*_bsr(A)).toarray()
        S = G @ A @ G.T
        for i in range(nb):
            assert np.allclose(S[3 * i:3 * i + 3, 3 * i:3 * i + 3], np.eye(3), atol=1e-11), (nb, i)
        # block lower triangular, inside the pattern of A
        for i in range(nb):
            for j in range(nb):
                blk = G[3 * i:3 * i + 3, 3 * j:3 * j + 3]
                if j > i or not np.any(A[3 * i:3 * i + 3, 3 * j:3 * j + 3]):
                    assert not np.any(blk), (i, j)
        # a wrongly scaled factor breaks the identity
        assert not np.allclose((1.01 * G) @ A @ (1.01 * G).T, S, atol=1e-6)


def test_rows_with_many_blocks_keep_the_largest():
    rng = np.random.default_rng(2)
    nb = 32
    A = _spd_dense(rng, nb)
    scale = np.exp(rng.uniform(-3, 3, (nb, nb)))
    scale = np.sqrt(scale * scale.T)
    np.fill_diagonal(scale, 1.0)
    A = A * np.kron(scale, np.ones((3, 3)))
    A = 0.5 * (A + A.T) + 4 * np.abs(A).sum(axis=1).max() * np.eye(3 * nb)
    G = F.block_fsai(*_bsr(A)).toarray()
    S = G @ A @ G.T
    for i in range(nb):
        low = list(range(i))
        if len(low) > F.K_MAX - 1:
            nrm = {j: np.sum(A[3 * i:3 * i + 3, 3 * j:3 * j + 3] ** 2) for j in low}
            low = sorted(low, key=lambda j: (-nrm[j], j))[:F.K_MAX - 1]
        keep = set(low) | {i}
        for j in range(nb):
            assert np.any(G[3 * i:3 * i + 3, 3 * j:3 * j + 3]) == (j in keep), (i, j)
        assert np.allclose(S[3 * i:3 * i + 3, 3 * i:3 * i + 3], np.eye(3), atol=1e-10), i
