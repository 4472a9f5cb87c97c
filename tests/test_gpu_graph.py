"""The Newton iteration as one HIP graph launch.

A context replays the iteration's captured launches while the stiffness, capacities, lagged
friction counts and tree topologies stay the same.  Every replay must give exactly what the
same launches give when enqueued one by one (a second context with the graph switched off):
the CSR, gradient and right-hand side, the step bound and its parts, the energy, bit for
bit.  The states change between calls as in a line search (x = x0 - alpha p on the device,
so new contact sets on the same tree topologies), new positions from the host make new
topologies and so a new capture, other entry points run eagerly between replays, and a
change of stiffness makes a new capture.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("alpha", "bound_ground", "bound_partial", "bound_cfl", "bound_full", "energy", "nnz", "contacts", "collisions")


@pytest.fixture(scope="module")
def gpu():
    from cipc import gpu as g
    g.load()
    return g


def _same(ta, tb, ra, rb, what):
    for k in KEYS:
        assert ta[k] == tb[k], (what, k, ta[k], tb[k])
    for k in ("outer", "inner", "val", "grad", "rhs"):
        assert np.array_equal(ra[k], rb[k]), (what, k)


@pytest.mark.parametrize("name", ["two_cubes", "four_20cubes"])
def test_replays_equal_eager_launches(gpu, name):
    from cipc import scenes
    s = (scenes.stacked_cubes(2, 6, 0.1, 1e-3, seed=3, rotate="yaw", velocity_scale=0.5, name="two_cubes")
         if name == "two_cubes" else scenes.four_cubes20(velocity_scale=0.5))
    xt, p = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    a = gpu.GpuContext(s, xtilde=xt)
    b = gpu.GpuContext(s, xtilde=xt)
    a.set_tune(2, 1)      # a single scene launches eagerly unless asked
    b.set_tune(2, 0)
    try:
        graphs = 0
        for g in (a, b):
            g.set_direction(p)
            g.save_positions()
        b.active_set()
        amax = b.step_bound(p)["alpha"]       # steps below the feasible one: no penetration
        for it in range(8):
            k = kappa * (2.0 if it == 4 else 1.0)      # a new stiffness: a new capture
            for g in (a, b):
                if it == 6:                            # new positions from the host
                    g.set_state(x=s.x)
                    g.set_direction(p)
                    g.save_positions()
                else:
                    g.step_forward(0.1 * it * amax)
            ta, tb = a.newton_iteration(k), b.newton_iteration(k)
            assert tb["graph"] == 0
            graphs += ta["graph"]
            _same(ta, tb, a.system(ta["nnz"]), b.system(tb["nnz"]), (name, it))
            # eager work between replays on the graph's own buffers
            ea, eb = a.energy(k), b.energy(k)
            assert ea == eb
            sa, sb = a.step_bound(p), b.step_bound(p)
            assert sa == sb
        print(name, "graph launches", graphs)
        assert graphs >= 5   # the calls after the first, but for the one after new positions
    finally:
        a.close()
        b.close()


def test_batch_replays_equal_eager_launches(gpu):
    from cipc import scenes
    sc = []
    for n, seed in ((3, 0), (5, 2), (4, 4), (2, 6)):
        sc.append(scenes.stacked_cubes(2, n, 0.1, 1e-3, seed=seed, rotate="yaw", velocity_scale=0.5, name=f"b{n}"))
    hh, fh, dt = max(s.hhat for s in sc), max(s.fhat for s in sc), max(s.dtol for s in sc)
    for s in sc:
        s.hhat, s.fhat, s.dtol = hh, fh, dt
    probes = [scenes.newton_probe_state(s, seed=3 + k) for k, s in enumerate(sc)]
    kappa = np.array([scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass) for s in sc])
    a = gpu.BatchContext(sc, xtildes=[xt for xt, _ in probes])   # a batch: one graph launch by default
    b = gpu.BatchContext(sc, xtildes=[xt for xt, _ in probes])
    b.set_tune(2, 0)
    try:
        p = np.concatenate([pp for _, pp in probes])
        graphs = 0
        for g in (a, b):
            g.set_direction(p)
            g.save_positions()
        b.active_set()
        amax = float(np.min(b.step_bound(p)["alpha"]))
        for it in range(4):
            for g in (a, b):
                g.step_forward(0.1 * it * amax)
            ta, tb = a.newton_iteration(kappa), b.newton_iteration(kappa)
            graphs += ta["graph"]
            assert np.array_equal(ta["alpha_scenes"], tb["alpha_scenes"]), it
            assert np.array_equal(ta["energy_scenes"], tb["energy_scenes"]), it
            for k in KEYS:
                assert ta[k] == tb[k], (it, k)
            ra, rb = a.system(ta["nnz"]), b.system(tb["nnz"])
            for k in ("outer", "inner", "val", "grad", "rhs"):
                assert np.array_equal(ra[k], rb[k]), (it, k)
        assert graphs >= 2
    finally:
        a.close()
        b.close()
