"""ORACLE - test infrastructure only.  numpy restatement of the learned block preconditioner
(forward, loss and backward), used by tests/ as the checker of the device kernels; never
imported by the product path.

The preconditioner has no counterpart in the reference (it replaces the plain block-Jacobi
PCG that stands in for NewtonLinearSystem::solve, NewtonLinearSystem.cpp:21-56).  The
model, shared with the device:

* graph: one node per vertex block of the Newton matrix A (full symmetric 3x3-block BSR),
  one directed edge per off-diagonal block;
* scaling: d_i = trace(A_ii) / 3, S = diag(1/sqrt(d_i)) per block, A^ = S A S;
* node features (16, zero padded): the six upper entries of A_ii / d_i, log(d_i / mean d),
  degree / 16, log(1 + sum_j |A_ij|_F / d_i); edge features (16): the nine entries of
  A_ij / sqrt(d_i d_j) and their Frobenius norm;
* encoders h0 = relu(X Wn + bn), e = relu(E We + be) (hidden width 64);
* three layers: m = h W_m; a_i = sum_j e_ij * m_j; h <- h + relu([h, a] W_u + b_u);
* base (analytic, no parameters): the first-order block FSAI of A^,
  B_ii = chol(A^_ii)^-1 (lower triangular) and B_ij = -w B_ii A^_ij A^_jj^-1 for j < i,
  w = 0.7;
* decoders: per node o_i = h_i W_d + b_d gives the lower-triangular T_i (diagonal exp(o),
  clamped to [-8, 8]; strict lower part o); per lower edge (i > j)
  o_ij = [h_i * h_j / 8, e_ij] W_o + b_o gives D_ij (nine entries);
* G^_ii = T_i B_ii, G^_ij = B_ij + D_ij: block lower-triangular on A's pattern, equal to
  the base when the decoders are zero;
* preconditioner M^-1 = S G^T G S;
* losses: L = sum_z |G^ A^ G^T z - z|^2 / (n_z 3n) over Gaussian probes z (objective 0), or
  the mean squared difference of G^ from the block FSAI factor of A in the scaled basis
  (objective 1, oracle/fsai_oracle.py).

Storage roundings of the device are reproduced with ``bf16=True``: features, embeddings,
hidden states, messages, aggregates and every GEMM input are rounded to bfloat16 (round to
nearest even); GEMMs accumulate in float64 here (float32 on the device).
"""
from __future__ import annotations

import numpy as np

H = 64
F_IN = 16
LAYERS = 3


def param_shapes():
    """Flat parameter layout (name, shape), in storage order."""
    s = [("Wn", (F_IN, H)), ("bn", (H,)), ("We", (F_IN, H)), ("be", (H,))]
    for l in range(LAYERS):
        s += [(f"Wm{l}", (H, H)), (f"Wu{l}", (2 * H, H)), (f"bu{l}", (H,))]
    s += [("Wd", (H, 16)), ("bd", (16,)), ("Wo", (2 * H, 16)), ("bo", (16,))]
    return s


def n_params():
    return sum(int(np.prod(sh)) for _, sh in param_shapes())


def unpack(flat):
    out, o = {}, 0
    for name, sh in param_shapes():
        k = int(np.prod(sh))
        out[name] = np.asarray(flat[o:o + k], np.float64).reshape(sh)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in param_shapes()])


def bf16(x):
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


class Graph:
    """Block graph of a full symmetric BSR (rowptr, col over V block rows)."""

    def __init__(self, rowptr, col):
        self.rowptr = np.asarray(rowptr, np.int64)
        self.col = np.asarray(col, np.int64)
        self.n = len(self.rowptr) - 1
        row = np.repeat(np.arange(self.n), np.diff(self.rowptr))
        self.row = row
        off = self.col != row
        self.eid = np.nonzero(off)[0]                     # BSR entries that are edges
        self.src = row[self.eid]                          # i (receiver)
        self.dst = self.col[self.eid]                     # j (sender)
        self.diag = np.full(self.n, -1, np.int64)
        self.diag[row[~off]] = np.nonzero(~off)[0]
        lower = self.src > self.dst
        self.low = np.nonzero(lower)[0]                   # edge ids with i > j


def features(g: Graph, val):
    """val: (nnzb, 3, 3) FP64 blocks in BSR order.  Returns X (n,16), E (m,16), d (n,)."""
    n = g.n
    D = np.zeros((n, 3, 3))
    has = g.diag >= 0
    D[has] = val[g.diag[has]]
    d = np.trace(D, axis1=1, axis2=2) / 3.0
    d = np.where(d > 0, d, 1.0)
    X = np.zeros((n, F_IN))
    Dn = D / d[:, None, None]
    iu = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for k, (a, b) in enumerate(iu):
        X[:, k] = Dn[:, a, b]
    X[:, 6] = np.log(d / d.mean())
    deg = np.bincount(g.src, minlength=n)
    X[:, 7] = deg / 16.0
    Ab = val[g.eid]
    fro = np.sqrt((Ab ** 2).sum(axis=(1, 2)))
    X[:, 8] = np.log1p(np.bincount(g.src, weights=fro, minlength=n) / d)
    sc = 1.0 / np.sqrt(d[g.src] * d[g.dst])
    An = Ab * sc[:, None, None]
    E = np.zeros((len(g.eid), F_IN))
    E[:, :9] = An.reshape(-1, 9)
    E[:, 9] = np.sqrt((An ** 2).sum(axis=(1, 2)))
    return X, E, d


OMEGA = 0.7


def base_factor(g: Graph, An):
    """First-order block FSAI of the normalized matrix: (B_ii (n,3,3), B_ij (n_low,3,3))."""
    Dii = An[g.diag]
    L = np.linalg.cholesky(Dii)
    a, b, c, d, e, f = L[:, 0, 0], L[:, 1, 0], L[:, 1, 1], L[:, 2, 0], L[:, 2, 1], L[:, 2, 2]
    Bd = np.zeros_like(L)                                  # inverse of a lower-triangular 3x3
    Bd[:, 0, 0], Bd[:, 1, 1], Bd[:, 2, 2] = 1 / a, 1 / c, 1 / f
    Bd[:, 1, 0] = -b / (a * c)
    Bd[:, 2, 1] = -e / (c * f)
    Bd[:, 2, 0] = (b * e - c * d) / (a * c * f)
    Dinv = np.linalg.inv(Dii)
    i, j = g.src[g.low], g.dst[g.low]
    Bo = -OMEGA * np.einsum("eab,ebc,ecd->ead", Bd[i], An[g.eid[g.low]], Dinv[j])
    return Bd, Bo


def _relu(x):
    return np.maximum(x, 0.0)


def forward(p, g: Graph, X, E, Bd, Bo, use_bf16=True):
    """Returns (Gd (n,3,3) lower-triangular, Go (n_low,3,3), cache)."""
    R = bf16 if use_bf16 else (lambda v: np.asarray(v, np.float64))
    c = {}
    Xb, Eb = R(X), R(E)
    c["Xb"], c["Eb"] = Xb, Eb
    pre = Xb @ R(p["Wn"]) + p["bn"]
    c["pre_n"] = pre
    h = R(_relu(pre))
    pre = Eb @ R(p["We"]) + p["be"]
    c["pre_e"] = pre
    e = R(_relu(pre))
    c["e"] = e
    c["h"] = [h]
    for l in range(LAYERS):
        m = R(h @ R(p[f"Wm{l}"]))
        a = np.zeros_like(h)
        np.add.at(a, g.src, e * m[g.dst])
        a = R(a)
        cat = np.concatenate([h, a], axis=1)
        pre = cat @ R(p[f"Wu{l}"]) + p[f"bu{l}"]
        h = R(h + _relu(pre))
        c[f"m{l}"], c[f"a{l}"], c[f"pre_u{l}"] = m, a, pre
        c["h"].append(h)
    od = h @ R(p["Wd"]) + p["bd"]
    i, j = g.src[g.low], g.dst[g.low]
    cat = R(np.concatenate([h[i] * h[j] * 0.125, e[g.low]], axis=1))
    c["cat_o"] = cat
    oo = cat @ R(p["Wo"]) + p["bo"]
    c["od"], c["oo"] = od, oo
    n = g.n
    T = np.zeros((n, 3, 3))
    dg = np.exp(np.clip(od[:, [0, 2, 5]], -8.0, 8.0))
    c["dg"] = dg
    T[:, 0, 0], T[:, 1, 1], T[:, 2, 2] = dg[:, 0], dg[:, 1], dg[:, 2]
    T[:, 1, 0], T[:, 2, 0], T[:, 2, 1] = od[:, 1], od[:, 3], od[:, 4]
    c["Bd"] = Bd
    Gd = np.einsum("nab,nbc->nac", T, Bd)
    Go = Bo + oo[:, :9].reshape(-1, 3, 3)
    return Gd, Go, c


def _gmul(g: Graph, Gd, Go, v):
    """y = G^ v with v (n,3): y_i = Gd_i v_i + sum_{j<i} Go_ij v_j."""
    y = np.einsum("nij,nj->ni", Gd, v)
    i, j = g.src[g.low], g.dst[g.low]
    np.add.at(y, i, np.einsum("eij,ej->ei", Go, v[j]))
    return y


def _gtmul(g: Graph, Gd, Go, v):
    """y = G^T v."""
    y = np.einsum("nji,nj->ni", Gd, v)
    i, j = g.src[g.low], g.dst[g.low]
    np.add.at(y, j, np.einsum("eji,ej->ei", Go, v[i]))
    return y


def _amul(g: Graph, An, v):
    """y = A^ v with the full normalized BSR An (nnzb,3,3)."""
    y = np.zeros_like(v)
    np.add.at(y, g.row, np.einsum("eij,ej->ei", An, v[g.col]))
    return y


def normalized_blocks(g: Graph, val, d):
    s = 1.0 / np.sqrt(d)
    return val * (s[g.row] * s[g.col])[:, None, None]


def loss_and_grad(flat, g: Graph, val, Z, use_bf16=True):
    """Z: (n_z, n, 3) probes.  Returns (loss, flat gradient, Gd, Go)."""
    p = unpack(flat)
    X, E, d = features(g, val)
    An = normalized_blocks(g, val, d)
    Bd, Bo = base_factor(g, An)
    Gd, Go, c = forward(p, g, X, E, Bd, Bo, use_bf16)
    nz = Z.shape[0]
    cst = 1.0 / (nz * 3 * g.n)
    L = 0.0
    dGd = np.zeros_like(Gd)
    dGo = np.zeros_like(Go)
    i, j = g.src[g.low], g.dst[g.low]
    for z in Z:
        w = _gtmul(g, Gd, Go, z)
        y = _amul(g, An, w)
        u = _gmul(g, Gd, Go, y)
        r = u - z
        L += cst * np.sum(r * r)
        gu = 2.0 * cst * r
        # u = G y
        dGd += np.einsum("ni,nj->nij", gu, y)
        dGo += np.einsum("ei,ej->eij", gu[i], y[j])
        gy = _gtmul(g, Gd, Go, gu)
        gw = _amul(g, An, gy)
        # w = G^T z: w_j = sum_i G_ij^T z_i -> dG_ij += z_i gw_j^T
        dGd += np.einsum("ni,nj->nij", z, gw)
        dGo += np.einsum("ei,ej->eij", z[i], gw[j])
    grads = backward(p, g, c, dGd, dGo, use_bf16)
    return L, pack(grads), Gd, Go


def fsai_targets(g: Graph, val, d):
    """The block FSAI factor (oracle/fsai_oracle.py) in the scaled basis, F^_ij = F_ij sqrt(d_j),
    as (diagonal blocks (n,3,3), lower off-diagonal blocks (n_low,3,3))."""
    import fsai_oracle
    F = fsai_oracle.block_fsai(g.rowptr, g.col, val).toarray()
    n = g.n
    Fd = np.stack([F[3 * i:3 * i + 3, 3 * i:3 * i + 3] * np.sqrt(d[i]) for i in range(n)])
    i, j = g.src[g.low], g.dst[g.low]
    Fo = np.stack([F[3 * a:3 * a + 3, 3 * b:3 * b + 3] * np.sqrt(d[b]) for a, b in zip(i, j)]) if len(i) else np.zeros((0, 3, 3))
    return Fd, Fo


def loss_and_grad_fsai(flat, g: Graph, val, use_bf16=True):
    """Objective 1: mean squared difference of G^ from the scaled block FSAI factor over the
    lower blocks (diagonal included), per entry."""
    p = unpack(flat)
    X, E, d = features(g, val)
    An = normalized_blocks(g, val, d)
    Bd, Bo = base_factor(g, An)
    Gd, Go, c = forward(p, g, X, E, Bd, Bo, use_bf16)
    Fd, Fo = fsai_targets(g, val, d)
    cst = 1.0 / (9.0 * (g.n + len(g.low)))
    L = cst * (np.sum((Gd - Fd) ** 2) + np.sum((Go - Fo) ** 2))
    grads = backward(p, g, c, 2 * cst * (Gd - Fd), 2 * cst * (Go - Fo), use_bf16)
    return L, pack(grads), Gd, Go


def backward(p, g: Graph, c, dGd, dGo, use_bf16=True):
    """Parameter gradients.  With ``use_bf16`` the upstream gradient of every GEMM is rounded
    to bfloat16 as the GEMM operand (weight gradient X^T dY and input gradient dY W^T), as
    on the device; bias gradients use the unrounded values."""
    R = bf16 if use_bf16 else (lambda v: np.asarray(v, np.float64))
    gr = {k: np.zeros_like(v) for k, v in p.items()}
    n = g.n
    od = c["od"]
    dT = np.einsum("nac,nbc->nab", dGd, c["Bd"])          # G_ii = T_i B_ii
    dod = np.zeros((n, 16))
    clip = (od[:, [0, 2, 5]] > -8.0) & (od[:, [0, 2, 5]] < 8.0)
    dod[:, 0] = dT[:, 0, 0] * c["dg"][:, 0] * clip[:, 0]
    dod[:, 2] = dT[:, 1, 1] * c["dg"][:, 1] * clip[:, 1]
    dod[:, 5] = dT[:, 2, 2] * c["dg"][:, 2] * clip[:, 2]
    dod[:, 1], dod[:, 3], dod[:, 4] = dT[:, 1, 0], dT[:, 2, 0], dT[:, 2, 1]
    doo = np.zeros((len(g.low), 16))
    doo[:, :9] = dGo.reshape(-1, 9)
    h = c["h"][-1]
    e = c["e"]
    gr["Wd"] = h.T @ R(dod)
    gr["bd"] = dod.sum(axis=0)
    gr["Wo"] = c["cat_o"].T @ R(doo)
    gr["bo"] = doo.sum(axis=0)
    dh = R(dod) @ R(p["Wd"]).T
    dcat = R(doo) @ R(p["Wo"]).T
    i, j = g.src[g.low], g.dst[g.low]
    de = np.zeros_like(e)
    de[g.low] += dcat[:, H:]
    np.add.at(dh, i, dcat[:, :H] * h[j] * 0.125)
    np.add.at(dh, j, dcat[:, :H] * h[i] * 0.125)
    for l in reversed(range(LAYERS)):
        hprev, m, a, pre = c["h"][l], c[f"m{l}"], c[f"a{l}"], c[f"pre_u{l}"]
        dpre = dh * (pre > 0)
        cat = np.concatenate([hprev, a], axis=1)
        gr[f"Wu{l}"] = cat.T @ R(dpre)
        gr[f"bu{l}"] = dpre.sum(axis=0)
        dcat_u = R(dpre) @ R(p[f"Wu{l}"]).T
        dhp = dh + dcat_u[:, :H]
        da = dcat_u[:, H:]
        # a_i = sum_j e_ij m_j
        de += da[g.src] * m[g.dst]
        dm = np.zeros_like(m)
        np.add.at(dm, g.dst, da[g.src] * e)
        gr[f"Wm{l}"] = hprev.T @ R(dm)
        dhp += R(dm) @ R(p[f"Wm{l}"]).T
        dh = dhp
    dpre_e = de * (c["pre_e"] > 0)
    gr["We"] = c["Eb"].T @ R(dpre_e)
    gr["be"] = dpre_e.sum(axis=0)
    dpre_n = dh * (c["pre_n"] > 0)
    gr["Wn"] = c["Xb"].T @ R(dpre_n)
    gr["bn"] = dpre_n.sum(axis=0)
    return gr


def init_params(seed=0):
    """He-style initialization of the hidden layers; decoders zero so that G^ starts at the
    analytic base."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, sh in param_shapes():
        if name.startswith("b") or name in ("Wd", "Wo"):
            p[name] = np.zeros(sh)
        else:
            p[name] = rng.standard_normal(sh) * np.sqrt(2.0 / sh[0]) * (0.5 if name.startswith("Wu") else 1.0)
    return pack(p)
