"""The batched workload at full size: 256 independent lattice cubes (8^3..24^3 cells, 2,560 to
69,120 tets each, 4.99M tets in all; seeded E, nu and drop height) in one batch context, at
the frame where each drop reaches the ground, so that every scene has ground contacts,
lagged friction and a ground-limited step.  The reference runs such scenes as independent
FEMSimulator / IPCSolverContext instances (IPCSolver.h:46-61).

Checked on every scene: the ground contacts (Ground::calculateActivateSet,
CollisionBroadPhase.cpp:1076-1092) against a numpy enumeration, no contact and no matrix
entry coupling two scenes, the ground and CFL parts of the step bound against their closed
forms (FeasibleStep.cpp:8-40; IPCSolver.cpp:960-974), and finite positive energies and steps.
Checked against the CPU oracle on the smallest, the median and the largest scene, extracted
from the batch: active sets, CSR pattern and the five step-bound values bit-exact; CSR
values and gradients within 1e-10 per entry (parity_util); energy within 1e-10 relative.
"""
import numpy as np
import pytest

from parity_util import csr_entry_err, vec3_entry_err

pytestmark = pytest.mark.gpu

N_SCENES = 256


@pytest.fixture(scope="module")
def batch():
    import time
    from cipc import gpu, scenes
    t0 = time.perf_counter()
    sc = scenes.batched_cubes(N_SCENES, seed=0, landed=True)
    probes = [scenes.newton_probe_state(s, seed=11 + k) for k, s in enumerate(sc)]
    kappa = np.array([scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass) for s in sc])
    b = gpu.BatchContext(sc, xtildes=[xt for xt, _ in probes])
    vprev = [s.x - s.velocities * s.material.dt for s in sc]
    b.set_state(V_prev=np.concatenate(vprev))
    b.active_set()
    b.friction_update(kappa)
    b.set_direction(np.concatenate([p for _, p in probes]))
    t = b.newton_iteration(kappa)
    print(f"config[3] batch: {N_SCENES} scenes, V={b.V}, T={sum(s.T for s in sc)}, set-up "
          f"{time.perf_counter() - t0:.1f} s, iteration {t['total_ms']:.2f} ms, contacts {t['contacts']}", flush=True)
    yield dict(b=b, scenes=sc, probes=probes, kappa=kappa, vprev=vprev, it=t)
    b.close()


def test_every_scene_has_its_ground_contacts(batch):
    b, sc = batch["b"], batch["scenes"]
    a = b.active_set()
    assert len(a["self_set"]) == 0 and len(a["ee_set"]) == 0     # single cubes at rest shape
    env = a["env_set"]
    voff = b.voff
    for k, s in enumerate(sc):
        want = s.surf_verts[(s.x[s.surf_verts, 1] + 1.0) ** 2 < s.hhat] + voff[k]
        mine = env[(env >= voff[k]) & (env < voff[k + 1])]
        assert len(want) > 0, k
        assert np.array_equal(mine, want), k
    assert len(env) == sum(((s.x[s.surf_verts, 1] + 1.0) ** 2 < s.hhat).sum() for s in sc)


def test_every_scene_step_bound_and_energy(batch):
    b, sc, probes, kappa = batch["b"], batch["scenes"], batch["probes"], batch["kappa"]
    bb = b.step_bound(np.concatenate([p for _, p in probes]))
    e = b.energy(kappa)
    for k, s in enumerate(sc):
        p = probes[k][1]
        sv = s.surf_verts
        nm = -p[sv, 1]
        dist = s.x[sv, 1] + 1.0
        cand = np.where(nm < 0.0, (-dist / np.where(nm < 0.0, nm, -1.0)) * 0.8, 1.0)
        cand = np.where(cand >= 0.0, cand, 1.0)
        ground = min(1.0, float(cand.min()))
        move = np.sqrt(p[sv, 0] * p[sv, 0] + p[sv, 1] * p[sv, 1] + p[sv, 2] * p[sv, 2]).max()
        cfl = np.sqrt(s.hhat) / (move * 2.0) if move > 0 else 1.0
        assert bb["ground"][k] == ground, (k, bb["ground"][k], ground)
        assert bb["cfl"][k] == cfl, (k, bb["cfl"][k], cfl)
        assert 0.0 < bb["alpha"][k] <= 1.0, k
        assert np.isfinite(e[k]) and e[k] != 0.0, k
    it = batch["it"]
    assert np.array_equal(it["alpha_scenes"], bb["alpha"])


def test_no_matrix_entry_couples_two_scenes(batch):
    b, kappa = batch["b"], batch["kappa"]
    r = b.assemble(kappa)
    col = np.repeat(np.arange(3 * b.V), np.diff(r["outer"]))
    vs = np.searchsorted(b.voff, np.arange(b.V), side="right") - 1
    assert np.array_equal(vs[r["inner"] // 3], vs[col // 3])
    assert np.all(np.diff(r["outer"]) > 0)           # every DOF keeps its diagonal


def _extract(r, lo, hi):
    import scipy.sparse as sp
    n = len(r["outer"]) - 1
    A = sp.csc_matrix((r["val"], r["inner"], r["outer"]), shape=(n, n))[lo:hi, lo:hi].tocsc()
    A.sort_indices()
    return A.indptr, A.indices, A.data


@pytest.mark.parametrize("which", ["smallest", "median", "largest"])
def test_scene_matches_oracle(batch, oracle_lib, which):
    b, sc, probes, kappa, vprev = batch["b"], batch["scenes"], batch["probes"], batch["kappa"], batch["vprev"]
    T = np.array([s.T for s in sc])
    k = {"smallest": int(np.argmin(T)), "largest": int(np.argmax(T)),
         "median": int(np.argsort(T)[len(T) // 2])}[which]
    s = sc[k]
    lo, hi = b.voff[k], b.voff[k + 1]
    o = oracle_lib.OracleMesh(s, xtilde=probes[k][0])
    try:
        o.set_state(V_prev=vprev[k])
        ao = o.active_set()
        ab = b.active_set()
        env = ab["env_set"]
        assert np.array_equal(env[(env >= lo) & (env < hi)] - lo, ao["env_set"]), k
        assert len(ao["self_set"]) == 0     # the surface's own neighbours give CCD candidates only
        b.friction_update(kappa)
        o.friction_update(kappa[k])
        rb = b.assemble(kappa)
        ro = o.assemble(kappa[k])
        outer, inner, val = _extract(rb, 3 * lo, 3 * hi)
        assert np.array_equal(outer, ro["outer"]) and np.array_equal(inner, ro["inner"]), k
        ev = csr_entry_err(ro["outer"], ro["inner"], val, ro["val"], scale=ro["val_scale"])
        eg = vec3_entry_err(rb["grad"][lo:hi], ro["grad"], scale=ro["grad_scale"])
        print(f"scene {k} ({which}, T={s.T}): per-entry CSR {ev:.2e} gradient {eg:.2e}")
        assert ev <= 1e-10 and eg <= 1e-10, (k, ev, eg)
        assert rb["collisions"][k] == ro["collisions"]
        eb = b.energy(kappa)[k]
        eo = o.energy(kappa[k])
        assert abs(eb - eo) <= 1e-10 * abs(eo), (k, eb, eo)
        bb = b.step_bound(np.concatenate([p for _, p in probes]))
        bo = o.step_bound(probes[k][1])
        for key in ("alpha", "ground", "partial", "cfl", "full"):
            assert bb[key][k] == bo[key], (k, key, bb[key][k], bo[key])
    finally:
        o.close()
