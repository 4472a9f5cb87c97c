"""The launch-shape autotuner (cipc_gpu_autotune / cipc_gpu_set_tune): every split of the
full-CCD walks and of the active set's walks (1 to 16 lanes per query) returns the oracle's
step bound and active set bit for bit, and the tuner keeps the fastest candidate and leaves
the results unchanged.

The full-CCD minimum over pairs does not depend on which lane meets which pair, and the
active set ranks its hits by primitive whatever lane found them, so the bar is bit-exact
against the oracle (Self_largestFeasibleStepSize_CCD, ContactMechanics.cpp:3754-3867;
buildCollisionSets, CollisionBroadPhase.cpp:765-1059) for every choice, on single scenes
and on a batch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("alpha", "ground", "partial", "cfl", "full")


@pytest.fixture(scope="module")
def gpu():
    from cipc import gpu as g
    g.load()
    return g


@pytest.mark.parametrize("name", ["two_cubes", "four_20cubes"])
def test_every_split_matches_oracle(gpu, oracle_lib, name):
    from cipc import scenes
    s = (scenes.stacked_cubes(2, 6, 0.1, 1e-3, seed=3, rotate="yaw", velocity_scale=0.5, name="two_cubes")
         if name == "two_cubes" else scenes.four_cubes20(velocity_scale=0.5))
    xt, p = scenes.newton_probe_state(s)
    g = gpu.GpuContext(s, xtilde=xt)
    o = oracle_lib.OracleMesh(s, xtilde=xt)
    try:
        ao = o.active_set()
        bo = o.step_bound(p)
        for choice in range(5):
            g.set_tune(0, choice)
            g.set_tune(1, 4 - choice)
            g.set_tune(3, choice)
            g.set_tune(4, 4 - choice)
            ag = g.active_set()
            for k in ("self_set", "ee_set", "ee_ie", "ccd_set", "env_set"):
                assert np.array_equal(ag[k], ao[k]), (name, choice, k)
            bg = g.step_bound(p)
            for k in KEYS:
                assert bg[k] == bo[k], (name, choice, k, bg[k], bo[k])
        t = g.autotune(p)
        print(name, {k: (c, [f"{m:.3f}" for m in ms]) for k, (c, ms) in t.items()})
        assert set(t) == {"full_ccd_pt", "full_ccd_ee", "active_pt", "active_ee"}
        for chosen, ms in t.values():
            assert len(ms) == 5 and all(m > 0 for m in ms)
            assert ms[chosen] == min(ms)
        ag = g.active_set()
        for k in ("self_set", "ee_set", "ee_ie", "ccd_set", "env_set"):
            assert np.array_equal(ag[k], ao[k]), (name, "tuned", k)
        bg = g.step_bound(p)
        for k in KEYS:
            assert bg[k] == bo[k], (name, "tuned", k, bg[k], bo[k])
    finally:
        g.close()
        o.close()


def test_batch_tuned_step_bound_unchanged(gpu):
    from cipc import scenes
    sc = []
    for n, seed in ((3, 0), (5, 2), (2, 4)):
        sc.append(scenes.stacked_cubes(2, n, 0.1, 1e-3, seed=seed, rotate="yaw", velocity_scale=0.5, name=f"b{n}"))
    hh, fh, dt = max(s.hhat for s in sc), max(s.fhat for s in sc), max(s.dtol for s in sc)
    for s in sc:
        s.hhat, s.fhat, s.dtol = hh, fh, dt
    probes = [scenes.newton_probe_state(s, seed=3 + k) for k, s in enumerate(sc)]
    b = gpu.BatchContext(sc, xtildes=[xt for xt, _ in probes])
    try:
        p = np.concatenate([pp for _, pp in probes])
        b.active_set()
        before = b.step_bound(p)
        t = b.autotune(p)
        assert len(t) == 4
        after = b.step_bound(p)
        for k in KEYS:
            assert np.array_equal(before[k], after[k]), k
    finally:
        b.close()


def test_set_tune_rejects_bad_choices(gpu):
    from cipc import scenes
    g = gpu.GpuContext(scenes.cube10())
    try:
        with pytest.raises(ValueError):
            g.set_tune(5, 0)
        with pytest.raises(ValueError):
            g.set_tune(0, 5)
        with pytest.raises(ValueError):
            g.set_tune(2, 2)      # the iteration graph is off (0) or on (1)
    finally:
        g.close()
