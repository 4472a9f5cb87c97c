"""The oracle's isIntersected (IPCSolver.cpp:476-542) pinned by brute force: for scenes
where the answer is True (overlapping cubes; a vertex on the ground) and False (separated
cubes), the LBVH-filtered answer equals the enumeration of every surface edge against every
surface face with the element routine segTriIntersect (oracle_seg_tri_intersect, pinned
in test_oracle_math.py) under the reference's skip rules (shared vertex, all-collider), or
a zero ground gap among the Environment_ActiveSet vertices."""
import numpy as np


def _brute(O, s):
    x = s.x
    for f in s.faces:
        fv = f[:3]
        for e in s.surf_edges:
            if e[0] in fv or e[1] in fv:
                continue
            pts = np.concatenate([x[e[0]], x[e[1]], x[fv[0]], x[fv[1]], x[fv[2]]])
            if O.seg_tri_intersect(pts):
                return True
    return False


def test_is_intersected_matches_brute_force(oracle_lib):
    from cipc import scenes
    O = oracle_lib
    overlap = scenes.stacked_cubes(2, 2, 0.1, -0.05, seed=1, rotate="yaw", name="overlap")
    apart = scenes.stacked_cubes(2, 2, 0.1, 4e-4, seed=1, rotate="yaw", name="apart")
    for s, want in ((overlap, True), (apart, False)):
        m = O.OracleMesh(s, xtilde=s.inertial_target())
        m.active_set()
        assert _brute(O, s) == want, s.name
        assert m.is_intersected() == want, s.name
        m.close()


def test_ground_touch_is_an_intersection(oracle_lib):
    from cipc import scenes
    O = oracle_lib
    s = scenes.stacked_cubes(1, 2, 0.1, 1e-4, seed=1, rotate="yaw", name="touch")
    m = O.OracleMesh(s, xtilde=s.inertial_target())
    m.active_set()
    assert not m.is_intersected()
    x = s.x.copy()
    x[np.argmin(x[:, 1]), 1] = scenes.GROUND_Y
    m.set_state(x=x)
    m.active_set()
    env = m.active_set()["env_set"]
    assert ((x[env, 1] - scenes.GROUND_Y) ** 2 <= 0.0).any()
    assert m.is_intersected()
    m.close()
