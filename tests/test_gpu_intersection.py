"""isIntersected (IPCSolver.cpp:476-542) where the answer is True as well as False, on the
device against the oracle: an edge of one cube crossing a face of another (two cubes
overlapping by 5 cm), a surface vertex exactly on the ground (a zero ground gap in the
Environment_ActiveSet), and separated cubes; per scene in a batch holding one intersected
scene among separated ones.  The oracle's answers are pinned on the CPU by a brute-force
enumeration of every surface edge against every surface face (test_oracle_intersection.py).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scenes():
    from cipc import scenes
    overlap = scenes.stacked_cubes(2, 3, 0.1, -0.05, seed=1, rotate="yaw", name="overlap")
    touch = scenes.stacked_cubes(1, 3, 0.1, 1e-4, seed=1, rotate="yaw", name="touch")
    touch.x[np.argmin(touch.x[:, 1]), 1] = scenes.GROUND_Y
    apart = scenes.stacked_cubes(2, 3, 0.1, 4e-4, seed=1, rotate="yaw", name="apart")
    return [(overlap, True), (touch, True), (apart, False)]


def test_is_intersected_matches_oracle(oracle_lib):
    from cipc import gpu
    for s, want in _scenes():
        xt = s.inertial_target()
        g = gpu.GpuContext(s, xtilde=xt)
        o = oracle_lib.OracleMesh(s, xtilde=xt)
        try:
            g.active_set()
            o.active_set()
            rg, ro = g.is_intersected(), o.is_intersected()
            print(s.name, rg, ro)
            assert ro == want, s.name
            assert rg == ro, s.name
        finally:
            g.close()
            o.close()


def test_batch_is_intersected_per_scene():
    from cipc import gpu, scenes
    sc = scenes.batched_cubes(6, seed=4, landed=True)
    bad = scenes.stacked_cubes(2, 3, 0.1, -0.05, seed=1, rotate="yaw", name="overlap")
    # a batch holds tet scenes with common scalars: take the batch's material for the bad one
    bad = scenes.finalize("overlap", bad.x, bad.tets, sc[0].material, velocities=bad.velocities)
    bad.hhat, bad.fhat, bad.dtol = sc[0].hhat, sc[0].fhat, sc[0].dtol
    allsc = sc[:3] + [bad] + sc[3:]
    b = gpu.BatchContext(allsc, xtildes=[s.inertial_target() for s in allsc])
    try:
        b.active_set()
        r = b.is_intersected()
        assert list(r) == [False, False, False, True, False, False, False], r
    finally:
        b.close()
