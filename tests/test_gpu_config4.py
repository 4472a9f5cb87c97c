"""The large workload at full size: eight 32^3-cell lattice cubes (287,496 vertices, 1,310,720
tets) stacked in contact with seeded velocities, at a frozen Newton state: about 100,000
self contacts, 590,000 partial-CCD pairs and 17 million CSR nonzeros.

Bars against the CPU oracle on the same state: the five active-set lists, the CSR pattern,
the collision count and the five step-bound values bit-exact; CSR values and gradients within
1e-10 per entry (parity_util); energy within 1e-10 relative; the fused Newton iteration the
same as the separate stages.
"""
import numpy as np
import pytest

from parity_util import csr_entry_err, vec3_entry_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big(oracle_lib):
    from cipc import gpu, scenes
    s = scenes.stacked_cubes(8, 32, 0.1, 1e-3, seed=4, rotate="yaw", velocity_scale=0.5, name="8x32cubes")
    xt, p = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    vprev = s.x - s.velocities * s.material.dt
    g = gpu.GpuContext(s, xtilde=xt)
    o = oracle_lib.OracleMesh(s, xtilde=xt)
    g.set_state(V_prev=vprev)
    o.set_state(V_prev=vprev)
    print(f"8x32^3 cubes: V={s.V} T={s.T}", flush=True)
    yield dict(s=s, g=g, o=o, p=p, kappa=kappa)
    g.close()
    o.close()


def test_active_set_bitexact(big):
    ag, ao = big["g"].active_set(), big["o"].active_set()
    for k in ("self_set", "ee_set", "ee_ie", "ccd_set", "env_set"):
        assert ag[k].shape == ao[k].shape, (k, ag[k].shape, ao[k].shape)
        assert np.array_equal(ag[k], ao[k]), k
    print({k: len(v) for k, v in ao.items()})
    assert len(ao["self_set"]) > 50000 and len(ao["ee_set"]) > 0


def test_assembly_with_friction_matches_oracle(big):
    g, o, kappa = big["g"], big["o"], big["kappa"]
    g.active_set()
    o.active_set()
    g.friction_update(kappa)
    o.friction_update(kappa)
    rg, ro = g.assemble(kappa), o.assemble(kappa)
    assert np.array_equal(rg["outer"], ro["outer"]) and np.array_equal(rg["inner"], ro["inner"])
    ev = csr_entry_err(ro["outer"], ro["inner"], rg["val"], ro["val"], scale=ro["val_scale"])
    eg = vec3_entry_err(rg["grad"], ro["grad"], scale=ro["grad_scale"])
    er = vec3_entry_err(rg["rhs"], ro["rhs"], scale=ro["grad_scale"])
    print(f"8x32^3: nnz {len(ro['val'])}, per-entry CSR {ev:.2e} gradient {eg:.2e} rhs {er:.2e}")
    assert ev <= 1e-10 and eg <= 1e-10 and er <= 1e-10, (ev, eg, er)
    assert rg["collisions"] == ro["collisions"]
    eo = o.energy(kappa)
    assert abs(g.energy(kappa) - eo) <= 1e-10 * abs(eo)


def test_step_bound_bitexact(big):
    g, o, p = big["g"], big["o"], big["p"]
    g.active_set()
    o.active_set()
    bg, bo = g.step_bound(p), o.step_bound(p)
    for k in ("alpha", "ground", "partial", "cfl", "full"):
        assert bg[k] == bo[k], (k, bg[k], bo[k])


def test_newton_iteration_matches_stages(big):
    g, o, p, kappa = big["g"], big["o"], big["p"], big["kappa"]
    o.active_set()
    bo = o.step_bound(p)
    g.set_direction(p)
    t = g.newton_iteration(kappa)
    assert t["alpha"] == bo["alpha"]
    for k in ("ground", "partial", "cfl", "full"):
        assert t["bound_" + k] == bo[k], k
    eo = o.energy(kappa)
    assert abs(t["energy"] - eo) <= 1e-10 * abs(eo)
    print(f"8x32^3 Newton iteration: {t['total_ms']:.2f} ms, contacts {t['contacts']}, nnz {t['nnz']}")


@pytest.mark.skip(reason="a profiled run of the 8x32^3 linear solve ended in a host segmentation fault whose cause "
                         "has not been found; the solve at this size stays off the GPU until it is")
def test_linear_solve_residual(big):
    g, kappa = big["g"], big["kappa"]
    g.active_set()
    g.friction_update(kappa)
    r = g.assemble(kappa)
    import scipy.sparse as sp
    n = len(r["outer"]) - 1
    L = sp.csc_matrix((r["val"], r["inner"], r["outer"]), shape=(n, n))
    A = (L + L.T - sp.diags(L.diagonal())).tocsr()
    for precision, pc in ((0, 0), (0, 1), (1, 1)):
        d, st = g.solve(precision=precision, rel_tol=1e-6, max_iter=20000, preconditioner=pc)
        res = np.linalg.norm(r["rhs"] - A @ d.reshape(-1)) / np.linalg.norm(r["rhs"])
        assert res <= 1.5e-6, (precision, pc, res, st)
