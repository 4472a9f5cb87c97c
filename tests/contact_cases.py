"""Contact barrier Hessians kappa (b'' g g^T + b' H_d) of random configurations, built from
the oracle's distance derivatives and barrier (ContactMechanics.cpp:3470-3561)."""
import numpy as np

NVERTS = {0: 2, 1: 3, 2: 4, 3: 4}


def barrier_blocks(O, kind, count, seed):
    """`count` barrier Hessians (count, n, n) and gradients (count, n), n = 3 * NVERTS[kind],
    for distance kind 0 PP, 1 PE, 2 PT, 3 EE.  Points are scaled over four decades, dHat is
    drawn between 1.0001 d and 40 d and kappa over six decades, so b' < 0 < b'' with a wide
    range of b''|g|^2 / |b'| (some blocks are already PSD, most are not)."""
    rng = np.random.default_rng(seed)
    n = 3 * NVERTS[kind]
    Hs, Gs = np.zeros((count, n, n)), np.zeros((count, n))
    for i in range(count):
        pts = rng.standard_normal(12) * 10.0 ** rng.uniform(-2, 1)
        d, g, H = O.distance(kind, pts)
        dhat = d * rng.uniform(1.0001, 40.0)
        _, bg, bh = O.barrier(d, dhat)
        kappa = 10.0 ** rng.uniform(-2, 4)
        Hs[i] = (kappa * bh) * np.outer(g, g) + (kappa * bg) * H
        Gs[i] = (kappa * bg) * g
    return Hs, Gs


def barrier_hessians(O, kind, count, seed):
    return barrier_blocks(O, kind, count, seed)[0]
