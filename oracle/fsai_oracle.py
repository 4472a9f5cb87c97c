"""ORACLE - test infrastructure only.  numpy restatement of the 3x3-block factorized sparse
approximate inverse (Kolotilina & Yeremin) that the device builds as a PCG preconditioner:
for block row i with lower pattern P (its blocks j <= i; at most 24, the diagonal and the
largest off-diagonal blocks by Frobenius norm, ties to the smaller column), G_{i,P} is the
last block row of L^-1 where A[P,P] = L L^T.  M^-1 = G^T G."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

K_MAX = 24


def block_fsai(rowptr, col, val):
    """rowptr/col/val: full symmetric 3x3-block BSR (val (nnzb,3,3)).  Returns G as a scipy
    CSR matrix (3n x 3n)."""
    n = len(rowptr) - 1
    rows, cols, vals = [], [], []
    A = sp.bsr_matrix((val, col, rowptr), shape=(3 * n, 3 * n)).tocsr()
    for i in range(n):
        e = np.arange(rowptr[i], rowptr[i + 1])
        low = e[col[e] < i]
        if len(low) > K_MAX - 1:
            nrm = (val[low] ** 2).sum(axis=(1, 2))
            order = sorted(range(len(low)), key=lambda t: (-nrm[t], t))
            low = np.sort(low[order[:K_MAX - 1]])
        P = list(col[low]) + [i]
        idx = np.concatenate([[3 * p, 3 * p + 1, 3 * p + 2] for p in P])
        M = A[idx][:, idx].toarray()
        L = np.linalg.cholesky(M)
        Linv = np.linalg.inv(L)
        Gr = Linv[-3:, :]
        for u in range(3):
            rows += [3 * i + u] * len(idx)
            cols += list(idx)
            vals += list(Gr[u])
    return sp.csr_matrix((vals, (rows, cols)), shape=(3 * n, 3 * n))


def pcg_iterations(A, b, apply_minv, tol=1e-6, maxit=20000):
    x = np.zeros_like(b)
    r = b.copy()
    z = apply_minv(r)
    p = z.copy()
    rz = r @ z
    bn = np.linalg.norm(b)
    for it in range(maxit):
        q = A @ p
        al = rz / (p @ q)
        x += al * p
        r -= al * q
        if np.linalg.norm(r) <= tol * bn:
            return it + 1, x
        z = apply_minv(r)
        rz1 = r @ z
        p = z + (rz1 / rz) * p
        rz = rz1
    return maxit, x
