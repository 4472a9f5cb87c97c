"""Brute-force pin of the oracle's active set (SpatialHash::calculateActivateSet,
CollisionBroadPhase.cpp:765-1059; Ground::calculateActivateSet, :1076-1092).

The oracle finds candidates with its LBVH and compacts them into the reference's five
output lists.  Here every (surface vertex, face) and (surface edge, surface edge) pair is
enumerated with no tree at all: a numpy box test with the reference's overlap rule
(CPUAABB::overlaps, LBVH.cpp:157-161) selects the candidates, each kept pair is classified
with the distance-type functions (pinned against dense-sampling brute force in
test_oracle_math) and measured with the closed-form squared distances (pinned by finite
differences there), and the lists are built with the reference's encoding rules: the
point-triangle and edge-edge entries in query order, the point-point and point-edge
entries counted in a dictionary and emitted in ascending EncodedContact order with minus
their multiplicity, and the nearly parallel edge-edge split (EE cross norm below
eps_x = 1e-3 |ea|^2 |eb|^2 at rest, ContactMechanics.cpp:758-762) into Self_EE_ActiveSet
and Self_EEeIe_ActiveSet.  The oracle must reproduce every list exactly: completeness, order,
multiplicities and the split.
"""
import numpy as np
import pytest

from cipc import scenes


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def _overlap(alo, ahi, blo, bhi, gap):
    """CPUAABB::overlaps: touching counts, a gap of at most `gap` per axis counts."""
    return ~(((blo - ahi) > gap).any(axis=-1) | ((alo - bhi) > gap).any(axis=-1))


def brute_force_active_set(O, s):
    x, hhat = s.x, s.hhat
    sq = np.sqrt(hhat)
    collider = s.boundary >= 2
    faces = s.faces[:, :3].astype(np.int64)
    sv = s.surf_verts.astype(np.int64)
    se = s.surf_edges.astype(np.int64)
    Es = len(se)
    tri = x[faces]                                       # (F,3,3)
    flo, fhi = tri.min(axis=1), tri.max(axis=1)
    seg = x[se]
    elo, ehi = seg.min(axis=1), seg.max(axis=1)

    pt_entries, ccd_pt = [], []
    counts = {}
    for si, v in enumerate(sv):
        p = x[v]
        ok = _overlap(p[None], p[None], flo, fhi, sq)
        ok &= ~(faces == v).any(axis=1)
        ok &= ~(collider[v] & collider[faces].all(axis=1))
        for f in np.nonzero(ok)[0]:                      # ascending face id, as the LBVH query sorts
            ccd_pt.append((-si - 1, int(f)))
            a, b, c = faces[f]
            t = O.dtype_pt(np.concatenate([p, x[a], x[b], x[c]]))
            if t in (0, 1, 2):
                q = (a, b, c)[t]
                d = O.distance(0, np.concatenate([p, x[q]]))[0]
                e = (-v - 1, q, -1, -1)
            elif t in (3, 4, 5):
                q0, q1 = ((a, b), (b, c), (c, a))[t - 3]
                d = O.distance(1, np.concatenate([p, x[q0], x[q1]]))[0]
                e = (-v - 1, q0, q1, -1)
            else:
                d = O.distance(2, np.concatenate([p, x[a], x[b], x[c]]))[0]
                e = (-v - 1, a, b, c)
            if d < hhat:
                if e[3] < 0:
                    counts[e] = counts.get(e, 0) + 1
                else:
                    pt_entries.append(e)

    ee_entries, ee_set, ee_ie, ccd_ee = [], [], [], []
    for ei in range(Es):
        i0, i1 = se[ei]
        ok = _overlap(elo[ei][None], ehi[ei][None], elo, ehi, sq)
        ok[:ei + 1] = False                              # candidates after the query edge only
        ok &= ~((se == i0) | (se == i1)).any(axis=1)
        ok &= ~(collider[i0] & collider[i1] & collider[se].all(axis=1))
        for ej in np.nonzero(ok)[0]:
            j0, j1 = se[ej]
            ccd_ee.append((ei, int(ej)))
            P = np.concatenate([x[i0], x[i1], x[j0], x[j1]])
            t = O.dtype_ee(P)
            cross = O.distance(4, P)[0]
            eps_x = 1e-3 * np.sum((s.v_rest[i0] - s.v_rest[i1]) ** 2) * np.sum((s.v_rest[j0] - s.v_rest[j1]) ** 2)
            mol = cross < eps_x
            add_e = -int(ej) - 2 if mol else -1
            pp = {0: (i0, j0), 1: (i0, j1), 3: (i1, j0), 4: (i1, j1)}
            pe = {2: (i0, j0, j1), 5: (i1, j0, j1), 6: (j0, i0, i1), 7: (j1, i0, i1)}
            if t in pp:
                a, b = pp[t]
                d = O.distance(0, np.concatenate([x[a], x[b]]))[0]
                e = (-a - 1, b, -1)
            elif t in pe:
                a, b, c = pe[t]
                d = O.distance(1, np.concatenate([x[a], x[b], x[c]]))[0]
                e = (-a - 1, b, c)
            else:
                d = O.distance(3, P)[0]
                e = None
            if not d < hhat:
                continue
            if e is None:                               # edge-edge
                if mol:
                    ee_set.append((i0, i1, j0, j1))
                    ee_ie.append((-1, -1))
                else:
                    ee_entries.append((i0, i1, j0, j1))
            elif mol:                                   # nearly parallel point-point / point-edge
                ee_set.append(e + (-1,))
                ee_ie.append((ei, int(ej)))
            else:
                key = e + (-1,)
                counts[key] = counts.get(key, 0) + 1
    self_set = pt_entries + ee_entries + [k[:3] + (-n,) for k, n in sorted(counts.items())]
    ground = [int(v) for v in sv if (x[v, 1] - (-1.0)) ** 2 < hhat]
    as2 = lambda a, w: np.array(a, np.int32).reshape(-1, w)
    return dict(self_set=as2(self_set, 4), ee_set=as2(ee_set, 4), ee_ie=as2(ee_ie, 2),
                ccd_set=as2(ccd_pt + ccd_ee, 2), env_set=np.array(ground, np.int32)), counts


def _shifted():
    """Two axis-aligned cubes, the upper one shifted by part of a cell and turned by one
    degree about the vertical, so that its lower vertices lie over the lower cube's edges and
    face interiors and its edges cross the lower cube's at small angles (nearly parallel
    edge-edge pairs with interior closest points)."""
    s = scenes.stacked_cubes(2, 3, 0.1, 3e-4, seed=5, rotate="none", jitter=1e-4, name="shifted")
    top = s.x[:, 1] > np.median(s.x[:, 1])
    c, sn = np.cos(np.radians(1.0)), np.sin(np.radians(1.0))
    R = np.array([[c, 0.0, sn], [0.0, 1.0, 0.0], [-sn, 0.0, c]])
    ctr = s.x[top].mean(axis=0)
    s.x[top] = (s.x[top] - ctr) @ R.T + ctr + np.array([0.05, 0.0, 0.0005])
    return s


def _scenes():
    # yawed faces at 3e-4 m (below dHat), axis-aligned faces (every parallel edge pair
    # across the interface is mollified), and aligned faces shifted by part of a cell
    return {
        "yaw": scenes.stacked_cubes(2, 3, 0.1, 3e-4, seed=3, rotate="yaw", jitter=0.002, name="yaw"),
        "aligned": scenes.stacked_cubes(2, 3, 0.1, 3e-4, seed=4, rotate="none", jitter=1e-4, name="aligned"),
        "shifted": _shifted(),
    }


@pytest.mark.parametrize("name", ["yaw", "aligned", "shifted"])
def test_active_set_equals_brute_force(O, name):
    """The oracle's five lists equal the brute-force enumeration exactly."""
    s = _scenes()[name]
    want, counts = brute_force_active_set(O, s)
    o = O.OracleMesh(s)
    try:
        got = o.active_set()
    finally:
        o.close()
    for k in ("self_set", "ee_set", "ee_ie", "ccd_set", "env_set"):
        assert got[k].shape == want[k].shape, (name, k, got[k].shape, want[k].shape)
        assert np.array_equal(got[k], want[k]), (name, k)
    assert len(want["self_set"]) > 0 and len(want["ccd_set"]) > 0, name


def test_brute_force_scenes_cover_the_encodings(O):
    """The scenes above exercise every branch the lists encode: point-triangle and
    edge-edge entries, point-point / point-edge entries with multiplicities above one,
    nearly parallel edge-edge pairs of both kinds, and ground contacts."""
    seen = dict(pt=0, ee=0, mult=0, mol_ee=0, mol_pe=0, ground=0)
    for s in _scenes().values():
        want, counts = brute_force_active_set(O, s)
        ss = want["self_set"]
        seen["pt"] += int(((ss[:, 0] < 0) & (ss[:, 3] >= 0)).sum())
        seen["ee"] += int(((ss[:, 0] >= 0) & (ss[:, 3] >= 0)).sum())
        seen["mult"] += sum(1 for n in counts.values() if n > 1)
        seen["mol_ee"] += int((want["ee_ie"][:, 0] == -1).sum())
        seen["mol_pe"] += int((want["ee_ie"][:, 0] >= 0).sum())
        seen["ground"] += len(want["env_set"])
    assert all(v > 0 for v in seen.values()), seen
