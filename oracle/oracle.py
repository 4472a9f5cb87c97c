# This Source Code Form is subject to the terms of the Mozilla Public License, v. 2.0.
# If a copy of the MPL was not distributed with this file, You can obtain one at
# https://mozilla.org/MPL/2.0/.  Parts of this file restate logic of CPU-IPC (MPL-2.0);
# each such part cites the source it follows.
"""ctypes binding of the oracle (test infrastructure only).

The oracle is the CPU restatement of CPU-IPC's per-Newton-iteration path; only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg use it,
and only as the checker.  The product path never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_u64p = C.POINTER(C.c_uint64)


class SceneDesc(C.Structure):
    _fields_ = [("V", C.c_int), ("T", C.c_int), ("F", C.c_int), ("Vs", C.c_int), ("Es", C.c_int),
                ("x", _dp), ("v_rest", _dp), ("V_prev", _dp), ("xtilde", _dp), ("masses", _dp),
                ("boundary", _ip), ("tets", _ip), ("dm_inv", _dp), ("volumes", _dp), ("faces", _ip),
                ("surf_verts", _u64p), ("surf_edges", _u64p),
                ("dt", C.c_double), ("length_rate", C.c_double), ("volume_rate", C.c_double),
                ("drag", C.c_double), ("hhat", C.c_double), ("fhat", C.c_double), ("friction", C.c_double),
                ("ground_normal", C.c_double * 3), ("ground_d", C.c_double),
                ("Tr", C.c_int), ("triangles", _ip), ("dm_tri_inv", _dp), ("areas", _dp),
                ("stretch", C.c_double), ("shear", C.c_double), ("strain_rate", C.c_double),
                ("plate_rigidity", C.c_double), ("Nb", C.c_int), ("bending", C.c_int), ("bend_verts", _ip),
                ("bend_q", _dp), ("bend_rest", _dp), ("bend_weight", _dp)]


def _fill_cloth(d, keep, scene):
    """Cloth fields of the descriptor; the arrays are appended to ``keep``."""
    from cipc import scenes
    for name, val in scenes.cloth_arrays(scene).items():
        if isinstance(val, np.ndarray):
            keep.append(val)
            setattr(d, name, val.ctypes.data_as(_ip if val.dtype == np.int32 else _dp))
        else:
            setattr(d, name, val)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(SceneDesc)]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.oracle_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.oracle_active_set.argtypes = [C.c_void_p, _i64p]
        L.oracle_get_active_set.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip, _ip]
        L.oracle_set_active_set.argtypes = [C.c_void_p, _i64p, _ip, _ip, _ip, _ip, _ip]
        L.oracle_assemble.restype = C.c_int
        L.oracle_assemble.argtypes = [C.c_void_p, C.c_double, _i64p]
        L.oracle_get_gradient.argtypes = [C.c_void_p, _dp]
        L.oracle_get_gradient_scale.argtypes = [C.c_void_p, _dp]
        L.oracle_get_csr_scale.argtypes = [C.c_void_p, _dp]
        L.oracle_get_csr.argtypes = [C.c_void_p, _ip, _ip, _dp, _dp]
        L.oracle_energy.restype = C.c_double
        L.oracle_energy.argtypes = [C.c_void_p, C.c_double]
        L.oracle_energy_terms.argtypes = [C.c_void_p, C.c_double, _dp]
        L.oracle_friction_update.argtypes = [C.c_void_p, C.c_double]
        L.oracle_set_soft.argtypes = [C.c_void_p, C.c_int, _ip, _dp, C.c_double]
        L.oracle_kappa_gradients.restype = C.c_int
        L.oracle_kappa_gradients.argtypes = [C.c_void_p, _dp]
        L.oracle_close_constraints.restype = C.c_int
        L.oracle_close_constraints.argtypes = [C.c_void_p, C.c_double, _i64p]
        L.oracle_clear_close_constraints.argtypes = [C.c_void_p]
        L.oracle_constraint_extrema.restype = C.c_int
        L.oracle_constraint_extrema.argtypes = [C.c_void_p, _dp]
        L.oracle_get_friction.argtypes = [C.c_void_p, _i64p, _dp, _dp, _dp, _dp]
        L.oracle_step_bound.argtypes = [C.c_void_p, _dp, _dp]
        L.oracle_full_ccd.restype = C.c_double
        L.oracle_full_ccd.argtypes = [C.c_void_p, _dp, C.c_double]
        L.oracle_step_forward.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
        L.oracle_get_positions.argtypes = [C.c_void_p, _dp]
        L.oracle_is_intersected.restype = C.c_int
        L.oracle_is_intersected.argtypes = [C.c_void_p]
        L.oracle_svd3.argtypes = [_dp, _dp, _dp, _dp]
        L.oracle_snh_element.argtypes = [_dp, C.c_double, C.c_double, _dp, _dp, _dp]
        L.oracle_bw_element.argtypes = [_dp, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        L.oracle_edge_theta.restype = C.c_double
        L.oracle_edge_theta.argtypes = [_dp, _dp, _dp]
        L.oracle_hinge_element.argtypes = [_dp, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        L.oracle_make_pd.restype = C.c_int
        L.oracle_make_pd.argtypes = [C.c_int, _dp]
        L.oracle_barrier.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp]
        L.oracle_distance.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.oracle_dtype_pt.restype = C.c_int
        L.oracle_dtype_pt.argtypes = [_dp]
        L.oracle_dtype_ee.restype = C.c_int
        L.oracle_dtype_ee.argtypes = [_dp]
        L.oracle_accd_pt.restype = C.c_double
        L.oracle_accd_pt.argtypes = [_dp, _dp, C.c_double, C.c_double]
        L.oracle_accd_ee.restype = C.c_double
        L.oracle_accd_ee.argtypes = [_dp, _dp, C.c_double, C.c_double]
        L.oracle_seg_tri_intersect.restype = C.c_int
        L.oracle_seg_tri_intersect.argtypes = [_dp]
        L.oracle_lbvh_query.restype = C.c_int64
        L.oracle_lbvh_query.argtypes = [C.c_int, _dp, _dp, C.c_double, _ip, C.c_int64]
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


class OracleMesh:
    """One scene on the CPU oracle; mirrors the device context's call sequence."""

    def __init__(self, scene, xtilde=None, threads: int = 0):
        L = lib()
        self.scene = scene
        xt = scene.inertial_target() if xtilde is None else xtilde
        self._keep = [_d(scene.x), _d(scene.v_rest), _d(scene.V_prev), _d(xt), _d(scene.masses),
                      _i(scene.boundary), _i(scene.tets), _d(scene.dm_inv), _d(scene.volumes),
                      _i(scene.faces), np.ascontiguousarray(scene.surf_verts, dtype=np.uint64),
                      np.ascontiguousarray(scene.surf_edges, dtype=np.uint64)]
        k = self._keep
        d = SceneDesc()
        d.V, d.T, d.F = scene.V, scene.T, len(scene.faces)
        d.Vs, d.Es = len(scene.surf_verts), len(scene.surf_edges)
        d.x, d.v_rest, d.V_prev, d.xtilde, d.masses = dptr(k[0]), dptr(k[1]), dptr(k[2]), dptr(k[3]), dptr(k[4])
        d.boundary, d.tets, d.dm_inv, d.volumes, d.faces = iptr(k[5]), iptr(k[6]), dptr(k[7]), dptr(k[8]), iptr(k[9])
        d.surf_verts, d.surf_edges = k[10].ctypes.data_as(_u64p), k[11].ctypes.data_as(_u64p)
        m = scene.material
        d.dt, d.length_rate, d.volume_rate, d.drag = m.dt, m.length_rate, m.volume_rate, m.drag
        d.hhat, d.fhat, d.friction = scene.hhat, scene.fhat, m.friction
        d.ground_normal[0], d.ground_normal[1], d.ground_normal[2] = 0.0, 1.0, 0.0
        d.ground_d = -1.0
        _fill_cloth(d, self._keep, scene)
        self.h = L.oracle_create(C.byref(d))
        if threads:
            L.oracle_set_threads(self.h, threads)
        self.V = scene.V

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, x=None, V_prev=None, xtilde=None):
        a = [None if v is None else _d(v) for v in (x, V_prev, xtilde)]
        lib().oracle_set_state(self.h, *[None if v is None else dptr(v) for v in a])

    def active_set(self):
        c = np.zeros(4, np.int64)
        lib().oracle_active_set(self.h, c.ctypes.data_as(_i64p))
        return self._get_as(c)

    def _get_as(self, c):
        s4 = np.zeros((max(c[0], 1), 4), np.int32)
        e4 = np.zeros((max(c[1], 1), 4), np.int32)
        eie = np.zeros((max(c[1], 1), 2), np.int32)
        ccd = np.zeros((max(c[2], 1), 2), np.int32)
        env = np.zeros(max(c[3], 1), np.int32)
        lib().oracle_get_active_set(self.h, iptr(s4), iptr(e4), iptr(eie), iptr(ccd), iptr(env))
        return dict(self_set=s4[:c[0]], ee_set=e4[:c[1]], ee_ie=eie[:c[1]], ccd_set=ccd[:c[2]], env_set=env[:c[3]])

    def set_active_set(self, a):
        c = np.array([len(a["self_set"]), len(a["ee_set"]), len(a["ccd_set"]), len(a["env_set"])], np.int64)
        arr = [_i(a[k]).reshape(-1) if len(a[k]) else np.zeros(1, np.int32)
               for k in ("self_set", "ee_set", "ee_ie", "ccd_set", "env_set")]
        lib().oracle_set_active_set(self.h, c.ctypes.data_as(_i64p), *[iptr(x) for x in arr])

    def assemble(self, kappa):
        nnz = np.zeros(1, np.int64)
        ncol = lib().oracle_assemble(self.h, kappa, nnz.ctypes.data_as(_i64p))
        n = 3 * self.V
        g = np.zeros(n)
        lib().oracle_get_gradient(self.h, dptr(g))
        gs = np.zeros(n)
        lib().oracle_get_gradient_scale(self.h, dptr(gs))
        outer = np.zeros(n + 1, np.int32)
        inner = np.zeros(int(nnz[0]), np.int32)
        val = np.zeros(int(nnz[0]))
        rhs = np.zeros(n)
        lib().oracle_get_csr(self.h, iptr(outer), iptr(inner), dptr(val), dptr(rhs))
        vs = np.zeros(int(nnz[0]))
        lib().oracle_get_csr_scale(self.h, dptr(vs))
        return dict(grad=g.reshape(-1, 3), grad_scale=gs.reshape(-1, 3), outer=outer, inner=inner, val=val, val_scale=vs, rhs=rhs,
                    collisions=ncol)

    def energy(self, kappa):
        return lib().oracle_energy(self.h, kappa)

    def energy_terms(self, kappa):
        """[elastic, inertia, ground barrier, self barrier, mollified barrier, friction,
        cloth triangles, bending, soft targets]"""
        t = np.zeros(9)
        lib().oracle_energy_terms(self.h, kappa, dptr(t))
        return t

    def set_soft_targets(self, idx, targets, weight):
        """SoftBoundaryCondition: 0.5 weight |x_v - target|^2 on the listed vertices."""
        idx, t = _i(np.asarray(idx).reshape(-1)), _d(np.asarray(targets).reshape(-1, 3))
        self._soft = (idx, t)
        lib().oracle_set_soft(self.h, len(idx), iptr(idx), dptr(t), weight)

    def kappa_gradients(self):
        """initKappa's reductions: (sum g_c . g_E, sum |g_c|^2, number of constraints)."""
        out = np.zeros(2)
        n = lib().oracle_kappa_gradients(self.h, dptr(out))
        return out[0], out[1], n

    def close_constraints(self, dtol):
        """postLineSearch's bookkeeping: (a recorded constraint got closer, (#ground, #self) recorded)."""
        c = np.zeros(2, np.int64)
        r = lib().oracle_close_constraints(self.h, dtol, c.ctypes.data_as(_i64p))
        return bool(r), (int(c[0]), int(c[1]))

    def clear_close_constraints(self):
        lib().oracle_clear_close_constraints(self.h)

    def constraint_extrema(self):
        """The kappa loop's exit test: (min, max, n) over the ground then self constraint values."""
        out = np.zeros(2)
        n = lib().oracle_constraint_extrema(self.h, dptr(out))
        return out[0], out[1], n

    def friction_update(self, kappa):
        """Friction::initialize on the current active set and positions."""
        lib().oracle_friction_update(self.h, kappa)

    def friction_state(self):
        n = np.zeros(2, np.int64)
        lib().oracle_get_friction(self.h, n.ctypes.data_as(_i64p), None, None, None, None)
        lam, coord, basis, env = np.zeros(n[0]), np.zeros(2 * n[0]), np.zeros(6 * n[0]), np.zeros(n[1])
        lib().oracle_get_friction(self.h, n.ctypes.data_as(_i64p), dptr(lam), dptr(coord), dptr(basis), dptr(env))
        return dict(lam=lam, coord=coord.reshape(-1, 2), basis=basis.reshape(-1, 2, 3), env_lam=env)

    def step_bound(self, p):
        p = _d(p)
        out = np.zeros(5)
        lib().oracle_step_bound(self.h, dptr(p), dptr(out))
        return dict(alpha=out[0], ground=out[1], partial=out[2], cfl=out[3], full=out[4])

    def full_ccd(self, p, alpha0=1.0):
        """Self_largestFeasibleStepSize_CCD alone from step alpha0."""
        return lib().oracle_full_ccd(self.h, dptr(_d(p)), alpha0)

    def step_forward(self, x0, p, alpha):
        x0, p = _d(x0), _d(p)
        lib().oracle_step_forward(self.h, dptr(x0), dptr(p), alpha)

    def positions(self):
        x = np.zeros((self.V, 3))
        lib().oracle_get_positions(self.h, dptr(x))
        return x

    def is_intersected(self):
        return bool(lib().oracle_is_intersected(self.h))


# ---- element-level helpers -------------------------------------------------------
def svd3(F):
    F = _d(np.asarray(F).T.reshape(-1))           # column-major
    U, s, V = np.zeros(9), np.zeros(3), np.zeros(9)
    lib().oracle_svd3(dptr(F), dptr(U), dptr(s), dptr(V))
    return U.reshape(3, 3).T, s, V.reshape(3, 3).T


def snh_element(F, u, r):
    Fc = _d(np.asarray(F).T.reshape(-1))
    P, H, e = np.zeros(9), np.zeros(81), np.zeros(1)
    lib().oracle_snh_element(dptr(Fc), u, r, dptr(P), dptr(H), dptr(e))
    return P.reshape(3, 3).T, H.reshape(9, 9), e[0]


def bw_element(F, stretch, shear, k):
    """Baraff-Witkin triangle: F (3x2) -> PK1 (3x2), projected 6x6 Hessian over
    column-major vec(F), energy density."""
    Fc = _d(np.asarray(F).T.reshape(-1))
    P, H, e = np.zeros(6), np.zeros(36), np.zeros(1)
    lib().oracle_bw_element(dptr(Fc), stretch, shear, k, dptr(P), dptr(H), dptr(e))
    return P.reshape(2, 3).T, H.reshape(6, 6), e[0]


def edge_theta(pts):
    p = _d(np.asarray(pts).reshape(-1))
    g, H = np.zeros(12), np.zeros(144)
    th = lib().oracle_edge_theta(dptr(p), dptr(g), dptr(H))
    return th, g, H.reshape(12, 12)


def hinge_element(pts, rest, weight, plate):
    p = _d(np.asarray(pts).reshape(-1))
    g, H, e = np.zeros(12), np.zeros(144), np.zeros(1)
    lib().oracle_hinge_element(dptr(p), rest, weight, plate, dptr(g), dptr(H), dptr(e))
    return g, H.reshape(12, 12), e[0]


def make_pd(H):
    H = _d(H).copy()
    n = H.shape[0]
    lib().oracle_make_pd(n, dptr(H))
    return H


def barrier(d, dhat):
    b, g, h = np.zeros(1), np.zeros(1), np.zeros(1)
    lib().oracle_barrier(d, dhat, dptr(b), dptr(g), dptr(h))
    return b[0], g[0], h[0]


DIST_PP, DIST_PE, DIST_PT, DIST_EE, DIST_EE_CROSS = 0, 1, 2, 3, 4


def distance(kind, pts):
    n = {0: 6, 1: 9}.get(kind, 12)
    p = np.zeros(12)
    p[:n] = np.asarray(pts, np.float64).reshape(-1)[:n]
    d, g, H = np.zeros(1), np.zeros(12), np.zeros(144)
    lib().oracle_distance(kind, dptr(p), dptr(d), dptr(g), dptr(H))
    return d[0], g[:n].copy(), H[:n * n].reshape(n, n).copy()


def dtype_pt(pts):
    p = _d(np.asarray(pts).reshape(-1))
    return lib().oracle_dtype_pt(dptr(p))


def dtype_ee(pts):
    p = _d(np.asarray(pts).reshape(-1))
    return lib().oracle_dtype_ee(dptr(p))


def accd_pt(pts, dirs, eta, thickness=0.0):
    p, d = _d(np.asarray(pts).reshape(-1)), _d(np.asarray(dirs).reshape(-1))
    return lib().oracle_accd_pt(dptr(p), dptr(d), eta, thickness)


def accd_ee(pts, dirs, eta, thickness=0.0):
    p, d = _d(np.asarray(pts).reshape(-1)), _d(np.asarray(dirs).reshape(-1))
    return lib().oracle_accd_ee(dptr(p), dptr(d), eta, thickness)


def seg_tri_intersect(pts):
    p = _d(np.asarray(pts).reshape(-1))
    return bool(lib().oracle_seg_tri_intersect(dptr(p)))


def lbvh_query(boxes, qbox, gap):
    boxes = _d(boxes)
    q = _d(qbox)
    out = np.zeros(max(len(boxes), 1), np.int32)
    n = lib().oracle_lbvh_query(len(boxes), dptr(boxes), dptr(q), gap, iptr(out), len(out))
    return out[:n].copy()
