"""Cloth on the device vs the CPU oracle, through the C ABI: Baraff-Witkin triangles
(IPCSolver.cpp:166-194), the quadratic bending form (IPCSolver.cpp:197-237) and hinge
bending (IPCSolver.cpp:238-267), element by element and assembled into whole Newton
systems of a folded cloth lying on a fixed ball collider.

Bars: pair sets, CSR patterns and step bounds bit-exact; FP64 element values and energies
within 1e-10 relative to the largest magnitude compared; assembled CSR values and gradients
within 1e-10 per entry, each relative to itself with a floor of 1e-6 of its 3x3 block's (or
vertex's) largest entry (parity_util).
"""
import numpy as np
import pytest

from parity_util import csr_entry_err, vec3_entry_err

pytestmark = pytest.mark.gpu

REL = 1e-10


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(initial=0.0), 1e-300)
    return np.abs(a - b).max(initial=0.0) / scale


@pytest.fixture(scope="module")
def gpu():
    from cipc import gpu as g
    g.load()
    return g


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def test_bw_element_matches_oracle(gpu, O):
    rng = np.random.default_rng(5)
    n = 400
    R = np.linalg.qr(rng.standard_normal((n, 3, 3)))[0]
    A = np.eye(2) + 0.3 * rng.standard_normal((n, 2, 2))
    A[: n // 4] *= 0.7                              # compression
    A[n // 4: n // 2] *= 1.3                        # tension
    F = R[:, :, :2] @ A
    st, sh, k = 1e4 / 0.84, 10 / 2.8 / 1e-3, 1e2
    P, H, e = gpu.eval_cloth(F, st, sh, k)
    for i in range(n):
        Po, Ho, eo = O.bw_element(F[i], st, sh, k)
        assert rel_err(P[i], Po) <= REL, i
        assert rel_err(H[i], Ho) <= REL, (i, rel_err(H[i], Ho))
        assert abs(e[i] - eo) <= REL * max(abs(eo), st), i


def test_hinge_element_matches_oracle(gpu, O):
    rng = np.random.default_rng(6)
    n = 300
    pts = rng.standard_normal((n, 4, 3))
    rest = rng.uniform(-3.0, 3.0, n)
    w = rng.uniform(0.3, 2.0, n)
    D = 1e-4
    g, H, e = gpu.eval_hinge(pts, rest, w, D)
    for i in range(n):
        go, Ho, eo = O.hinge_element(pts[i], rest[i], w[i], D)
        assert rel_err(g[i], go) <= REL, i
        assert rel_err(H[i], Ho) <= REL, (i, rel_err(H[i], Ho))
        assert abs(e[i] - eo) <= REL * max(abs(eo), 1e-300) or e[i] == eo, i


def _cloth(name):
    from cipc import scenes
    if name == "cloth24_quad":
        return scenes.cloth_on_sphere(24, sphere_cells=4, radius=0.25, fold=16, bending=0, seed=1, name=name)
    if name == "cloth24_hinge":
        return scenes.cloth_on_sphere(24, sphere_cells=4, radius=0.25, fold=16, bending=1, seed=2, name=name)
    return scenes.cloth_on_sphere(256, fold=160, bending=0, seed=0, name=name)


@pytest.fixture(scope="module", params=["cloth24_quad", "cloth24_hinge", "cloth256_quad"])
def cloth(request, gpu, O):
    from cipc import scenes
    s = _cloth(request.param)
    xt, p = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    g = gpu.GpuContext(s, xtilde=xt)
    o = O.OracleMesh(s, xtilde=xt)
    yield dict(scene=s, g=g, o=o, p=p, kappa=kappa, name=request.param)
    g.close()
    o.close()


def test_cloth_active_set_bitexact(cloth):
    ag, ao = cloth["g"].active_set(), cloth["o"].active_set()
    assert len(ao["self_set"]) > 0 and len(ao["ee_set"]) > 0, cloth["name"]
    for k in ("self_set", "ee_set", "ee_ie", "ccd_set", "env_set"):
        assert np.array_equal(ag[k], ao[k]), (cloth["name"], k)


def _pattern_report(rg, ro, n):
    """Entries present in one CSR only, with their magnitude relative to the largest value."""
    import scipy.sparse as sp
    A = sp.csc_matrix((rg["val"], rg["inner"], rg["outer"]), shape=(n, n)).tocoo()
    B = sp.csc_matrix((ro["val"], ro["inner"], ro["outer"]), shape=(n, n)).tocoo()
    ka = dict(zip(zip(A.row.tolist(), A.col.tolist()), A.data))
    kb = dict(zip(zip(B.row.tolist(), B.col.tolist()), B.data))
    mx = np.abs(ro["val"]).max()
    only_g = [(k, ka[k] / mx) for k in ka.keys() - kb.keys()]
    only_o = [(k, kb[k] / mx) for k in kb.keys() - ka.keys()]
    return f"device-only {len(only_g)} {only_g[:4]}, oracle-only {len(only_o)} {only_o[:4]}"


def test_cloth_assembly_matches_oracle(cloth):
    g, o, kappa = cloth["g"], cloth["o"], cloth["kappa"]
    g.active_set()
    o.active_set()
    rg, ro = g.assemble(kappa), o.assemble(kappa)
    name = cloth["name"]
    same = np.array_equal(rg["outer"], ro["outer"]) and np.array_equal(rg["inner"], ro["inner"])
    assert same, (name, _pattern_report(rg, ro, 3 * cloth["scene"].V))
    ev = csr_entry_err(ro["outer"], ro["inner"], rg["val"], ro["val"], scale=ro["val_scale"])
    eg, er = vec3_entry_err(rg["grad"], ro["grad"], scale=ro["grad_scale"]), vec3_entry_err(rg["rhs"], ro["rhs"], scale=ro["grad_scale"])
    print(f"{name}: per-entry CSR {ev:.2e} gradient {eg:.2e} rhs {er:.2e}")
    assert ev <= REL and eg <= REL and er <= REL, (name, ev, eg, er)
    assert rg["collisions"] == ro["collisions"]


def test_cloth_energy_matches_oracle(cloth):
    g, o, kappa = cloth["g"], cloth["o"], cloth["kappa"]
    g.active_set()
    o.active_set()
    tg, to = g.energy_terms(kappa), o.energy_terms(kappa)
    assert to[6] > 0 and to[7] > 0
    for k in range(8):
        assert abs(tg[k] - to[k]) <= REL * max(abs(to[k]), 1e-300) or tg[k] == to[k], (cloth["name"], k, tg[k], to[k])
    assert abs(g.energy(kappa) - o.energy(kappa)) <= REL * abs(o.energy(kappa))


def test_cloth_step_bound_matches_oracle(cloth):
    g, o, p = cloth["g"], cloth["o"], cloth["p"]
    g.active_set()
    o.active_set()
    bg, bo = g.step_bound(p), o.step_bound(p)
    for k in ("alpha", "ground", "partial", "cfl", "full"):
        assert bg[k] == bo[k], (cloth["name"], k, bg[k], bo[k])
