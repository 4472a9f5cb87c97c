"""Device-resident counts of the active set: no stage of a Newton iteration waits on the host
for a count.  The hit, map-key and contact buffers are sized by capacities; a fresh context
starts them small (1,024 contacts), so the first Newton iteration on the four-cube scene
(about 13,000 contacts) exceeds them, grows them from the counts the device found and runs
again.  Its results must be exactly those of a context whose capacities already fit, and the
oracle's (sets bit-exact, system and energy per entry within 1e-10, step bound bit-exact).
"""
import numpy as np
import pytest

from parity_util import csr_entry_err, vec3_entry_err

pytestmark = pytest.mark.gpu

REL = 1e-10


@pytest.fixture(scope="module")
def gpu():
    from cipc import gpu as g
    g.load()
    return g


def test_first_iteration_on_a_fresh_context(gpu, oracle_lib):
    from cipc import scenes
    s = scenes.four_cubes20(velocity_scale=0.5)
    xt, p = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    g = gpu.GpuContext(s, xtilde=xt)
    o = oracle_lib.OracleMesh(s, xtilde=xt)
    try:
        g.set_direction(p)
        t = g.newton_iteration(kappa)        # capacities grown inside the call
        print(f"contacts {t['contacts']} nnz {t['nnz']} ms {t['total_ms']:.3f}")
        assert t["contacts"] > 4096
        ao = o.active_set()
        assert t["contacts"] == len(ao["self_set"]) + len(ao["ee_set"]) + len(ao["env_set"])
        rg, ro = g.system(t["nnz"]), o.assemble(kappa)
        assert np.array_equal(rg["outer"], ro["outer"]) and np.array_equal(rg["inner"], ro["inner"])
        assert csr_entry_err(ro["outer"], ro["inner"], rg["val"], ro["val"], scale=ro["val_scale"]) <= REL
        assert vec3_entry_err(rg["grad"], ro["grad"], scale=ro["grad_scale"]) <= REL
        bo = o.step_bound(p)
        assert t["alpha"] == bo["alpha"]
        for k in ("ground", "partial", "cfl", "full"):
            assert t["bound_" + k] == bo[k], k
        assert abs(t["energy"] - o.energy(kappa)) <= REL * abs(t["energy"])
        # the same iteration with the capacities now in place: identical results
        t2 = g.newton_iteration(kappa)
        r2 = g.system(t2["nnz"])
        assert t2["nnz"] == t["nnz"] and t2["contacts"] == t["contacts"] and t2["alpha"] == t["alpha"]
        assert np.array_equal(r2["val"], rg["val"]) and np.array_equal(r2["grad"], rg["grad"])
        assert t2["energy"] == t["energy"]
    finally:
        g.close()
        o.close()


def test_active_set_after_a_larger_one(gpu, oracle_lib):
    """Capacities grown for a state with many contacts serve a state with few: the counts come
    from the device, not from the buffers' sizes."""
    from cipc import scenes
    s = scenes.four_cubes20(velocity_scale=0.5)
    g = gpu.GpuContext(s)
    o = oracle_lib.OracleMesh(s)
    try:
        a1 = g.active_set()
        lifted = s.x + np.array([0.0, 0.5, 0.0]) * (s.x[:, 1:2] > np.median(s.x[:, 1]))
        g.set_state(x=lifted)
        o.set_state(x=lifted)
        a2, ao = g.active_set(), o.active_set()
        assert len(a2["self_set"]) < len(a1["self_set"])
        for k in ("self_set", "ee_set", "ee_ie", "ccd_set", "env_set"):
            assert np.array_equal(a2[k], ao[k]), k
    finally:
        g.close()
        o.close()
