"""Python binding of the device library's C ABI (include/cipc_gpu.h).

This mirrors the call sequence the reference drives per Newton iteration
(``CPU IPC/IPCSolver.cpp:890-1064``): ``active_set`` replaces
``SpatialHash::build`` + ``calculateActivateSet`` + ``Ground::calculateActivateSet``,
``assemble`` replaces ``computeGradientAndHessian`` + ``NewtonLinearSystem::assemble``,
``step_bound`` the ground/CCD/CFL block, ``energy`` ``computeEnergy``.  Errors raise the
Python counterpart of the exception the reference throws at the same place.

There is no fallback: if ``libcipc_gpu.so`` is missing or no gfx950 device is present the
constructor raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libcipc_gpu.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_u64p = C.POINTER(C.c_uint64)
_fp = C.POINTER(C.c_float)

OK, ERR_INVALID_ARGUMENT, ERR_RUNTIME, ERR_LOGIC, ERR_OVERFLOW, ERR_HIP, ERR_NO_DEVICE = range(7)

#: every symbol include/cipc_gpu.h declares
EXPORTS = (
    "cipc_gpu_last_error", "cipc_gpu_context_create", "cipc_gpu_context_destroy", "cipc_gpu_upload_scene",
    "cipc_gpu_set_state", "cipc_gpu_get_positions", "cipc_gpu_set_direction", "cipc_gpu_active_set",
    "cipc_gpu_get_active_set", "cipc_gpu_set_active_set", "cipc_gpu_assemble", "cipc_gpu_set_hessian_precision",
    "cipc_gpu_get_gradient",
    "cipc_gpu_get_csr", "cipc_gpu_energy", "cipc_gpu_energy_terms", "cipc_gpu_energy_terms_noalloc", "cipc_gpu_step_bound",
    "cipc_gpu_autotune", "cipc_gpu_set_tune",
    "cipc_gpu_friction_update", "cipc_gpu_get_friction", "cipc_gpu_set_soft_targets", "cipc_gpu_kappa_gradients",
    "cipc_gpu_close_constraints", "cipc_gpu_clear_close_constraints", "cipc_gpu_preconditioner_param_count",
    "cipc_gpu_set_preconditioner_weights", "cipc_gpu_preconditioner_gradient", "cipc_gpu_preconditioner_train_step", "cipc_gpu_solve", "cipc_gpu_solve_ex", "cipc_gpu_get_solve_factor", "cipc_gpu_save_positions", "cipc_gpu_step_forward",
    "cipc_gpu_is_intersected", "cipc_gpu_newton_iteration", "cipc_gpu_set_velocities", "cipc_gpu_get_state",
    "cipc_gpu_time_step",
    "cipc_gpu_upload_batch", "cipc_gpu_batch_size", "cipc_gpu_batch_assemble", "cipc_gpu_batch_friction_update",
    "cipc_gpu_batch_energy", "cipc_gpu_batch_step_bound", "cipc_gpu_batch_step_forward", "cipc_gpu_batch_is_intersected",
    "cipc_gpu_batch_solve", "cipc_gpu_batch_newton_iteration",
    "cipc_gpu_eval_snh", "cipc_gpu_eval_cloth", "cipc_gpu_eval_hinge", "cipc_gpu_eval_distance", "cipc_gpu_eval_make_pd", "cipc_gpu_eval_contact_pd",
    "cipc_gpu_eval_accd",
    "cipc_gpu_eval_classify",
)


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


class IterationTiming(C.Structure):
    _fields_ = [("broadphase_ms", C.c_double), ("assembly_ms", C.c_double), ("elastic_ms", C.c_double),
                ("ccd_ms", C.c_double),
                ("energy_ms", C.c_double), ("total_ms", C.c_double), ("alpha", C.c_double),
                ("energy", C.c_double), ("nnz", C.c_int64), ("contacts", C.c_int64), ("collisions", C.c_int64),
                ("bound_ground", C.c_double), ("bound_partial", C.c_double), ("bound_cfl", C.c_double),
                ("bound_full", C.c_double), ("graph", C.c_int), ("enqueue_ms", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("refinements", C.c_int), ("rel_residual", C.c_double), ("solve_ms", C.c_double)]


class StepParams(C.Structure):
    _fields_ = [("bbox_diag2", C.c_double), ("mean_mass", C.c_double), ("dtol", C.c_double), ("newton_threshold", C.c_double),
                ("apply_gravity", C.c_int), ("quasi_static", C.c_int), ("precision", C.c_int), ("preconditioner", C.c_int),
                ("solve_rel_tol", C.c_double), ("solve_max_iter", C.c_int), ("max_newton", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("newton_iterations", "kappa_iterations", "energy_backtracks", "intersection_backtracks",
                                       "max_energy_backtracks_per_newton", "newton_steps_over_two_backtracks")] + \
               [(n, C.c_int64) for n in ("pcg_iterations", "collisions", "matrix_nnz", "direct_solves", "ground_contacts",
                                         "self_contacts", "mollified_contacts")] + \
               [(n, C.c_double) for n in ("kappa", "min_distance2", "dirichlet_alpha", "alpha_sum", "alpha_min", "alpha_max",
                                          "step_ms", "assembly_ms", "linear_ms", "ccd_ms", "line_search_ms", "post_ms")]


class TuneEntry(C.Structure):
    _fields_ = [("kernel", C.c_int), ("chosen", C.c_int), ("n_candidates", C.c_int), ("ms", C.c_double * 8)]


DirichletFn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double),
                          C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double))


class Dirichlet(C.Structure):
    _fields_ = [("n", C.c_int), ("vertices", C.POINTER(C.c_int)), ("step", C.c_int), ("fn", DirichletFn), ("user", C.c_void_p)]


class CipcError(RuntimeError):
    """HIP or device failure (CIPC_ERR_HIP / CIPC_ERR_NO_DEVICE)."""


class CipcLogicError(RuntimeError):
    """std::logic_error in the reference."""


_lib = None


def load(path: str = LIB_PATH):
    """Load the device library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(path)
    L.cipc_gpu_last_error.restype = C.c_char_p
    sig = {
        "cipc_gpu_context_create": [C.c_int, C.POINTER(C.c_void_p)],
        "cipc_gpu_context_destroy": [C.c_void_p],
        "cipc_gpu_upload_scene": [C.c_void_p, C.POINTER(SceneDesc)],
        "cipc_gpu_set_state": [C.c_void_p, _dp, _dp, _dp],
        "cipc_gpu_get_positions": [C.c_void_p, _dp],
        "cipc_gpu_set_direction": [C.c_void_p, _dp],
        "cipc_gpu_active_set": [C.c_void_p, _i64p],
        "cipc_gpu_get_active_set": [C.c_void_p, _ip, _ip, _ip, _ip, _ip],
        "cipc_gpu_set_active_set": [C.c_void_p, _i64p, _ip, _ip, _ip, _ip, _ip],
        "cipc_gpu_assemble": [C.c_void_p, C.c_double, _i64p, _ip],
        "cipc_gpu_set_hessian_precision": [C.c_void_p, C.c_int],
        "cipc_gpu_get_gradient": [C.c_void_p, _dp],
        "cipc_gpu_get_csr": [C.c_void_p, _ip, _ip, _dp, _dp],
        "cipc_gpu_energy": [C.c_void_p, C.c_double, _dp],
        "cipc_gpu_energy_terms": [C.c_void_p, C.c_double, _dp],
        "cipc_gpu_energy_terms_noalloc": [C.c_void_p, C.c_double, C.c_void_p],
        "cipc_gpu_friction_update": [C.c_void_p, C.c_double],
        "cipc_gpu_set_soft_targets": [C.c_void_p, C.c_int, _ip, _dp, C.c_double],
        "cipc_gpu_kappa_gradients": [C.c_void_p, _dp, _i64p],
        "cipc_gpu_set_preconditioner_weights": [C.c_void_p, _fp, C.c_int64],
        "cipc_gpu_preconditioner_gradient": [C.c_void_p, C.c_int, _fp, _fp, C.c_int, _dp, _fp],
        "cipc_gpu_preconditioner_train_step": [C.c_void_p, C.c_int, _fp, _fp, _fp, C.c_int, C.c_double, C.c_int, C.c_uint64, _dp],
        "cipc_gpu_close_constraints": [C.c_void_p, C.c_double, _ip, _i64p],
        "cipc_gpu_clear_close_constraints": [C.c_void_p],
        "cipc_gpu_solve": [C.c_void_p, C.c_int, C.c_double, C.c_int, _dp, C.POINTER(SolveStats)],
        "cipc_gpu_solve_ex": [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, _dp, C.POINTER(SolveStats)],
        "cipc_gpu_get_solve_factor": [C.c_void_p, _i64p, _ip, _ip, _dp],
        "cipc_gpu_get_friction": [C.c_void_p, C.POINTER(C.c_int64), _dp, _dp, _dp, _dp],
        "cipc_gpu_step_bound": [C.c_void_p, _dp],
        "cipc_gpu_autotune": [C.c_void_p, C.POINTER(TuneEntry), C.c_int, C.POINTER(C.c_int)],
        "cipc_gpu_set_tune": [C.c_void_p, C.c_int, C.c_int],
        "cipc_gpu_save_positions": [C.c_void_p],
        "cipc_gpu_step_forward": [C.c_void_p, C.c_double],
        "cipc_gpu_is_intersected": [C.c_void_p, _ip],
        "cipc_gpu_newton_iteration": [C.c_void_p, C.c_double, C.POINTER(IterationTiming)],
        "cipc_gpu_set_velocities": [C.c_void_p, _dp],
        "cipc_gpu_get_state": [C.c_void_p, _dp, _dp, _dp, _dp],
        "cipc_gpu_time_step": [C.c_void_p, C.POINTER(StepParams), _dp, C.POINTER(Dirichlet), C.POINTER(StepStats)],
        "cipc_gpu_upload_batch": [C.c_void_p, C.POINTER(C.POINTER(SceneDesc)), C.c_int],
        "cipc_gpu_batch_size": [C.c_void_p, _ip, _ip, _ip],
        "cipc_gpu_batch_assemble": [C.c_void_p, _dp, _i64p, _ip],
        "cipc_gpu_batch_friction_update": [C.c_void_p, _dp],
        "cipc_gpu_batch_energy": [C.c_void_p, _dp, _dp, _dp],
        "cipc_gpu_batch_step_bound": [C.c_void_p, _dp],
        "cipc_gpu_batch_step_forward": [C.c_void_p, _dp],
        "cipc_gpu_batch_is_intersected": [C.c_void_p, _ip],
        "cipc_gpu_batch_solve": [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, _ip, _dp, C.POINTER(SolveStats)],
        "cipc_gpu_batch_newton_iteration": [C.c_void_p, _dp, C.POINTER(IterationTiming), _dp, _dp],
        "cipc_gpu_eval_snh": [C.c_int, _dp, C.c_double, C.c_double, _dp, _dp, _dp],
        "cipc_gpu_eval_cloth": [C.c_int, _dp, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp],
        "cipc_gpu_eval_hinge": [C.c_int, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp],
        "cipc_gpu_eval_distance": [C.c_int, C.c_int, _dp, _dp, _dp, _dp],
        "cipc_gpu_eval_make_pd": [C.c_int, C.c_int, _dp],
        "cipc_gpu_eval_contact_pd": [C.c_int, C.c_int, _dp, _dp],
        "cipc_gpu_eval_accd": [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_double, _dp],
        "cipc_gpu_eval_classify": [C.c_int, C.c_int, _dp, _ip],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.restype = C.c_int
        fn.argtypes = args
    L.cipc_gpu_preconditioner_param_count.restype = C.c_int64
    L.cipc_gpu_preconditioner_param_count.argtypes = []
    _lib = L
    return L


def _check(code: int):
    if code == OK:
        return
    msg = load().cipc_gpu_last_error().decode(errors="replace")
    if code == ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if code == ERR_OVERFLOW:
        raise OverflowError(msg)
    if code == ERR_LOGIC:
        raise CipcLogicError(msg)
    if code == ERR_RUNTIME:
        raise RuntimeError(msg)
    raise CipcError(msg)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


def scene_desc(scene, xtilde=None):
    """Descriptor of a host scene; returns (desc, arrays kept alive)."""
    xt = scene.inertial_target() if xtilde is None else xtilde
    keep = [_d(scene.x), _d(scene.v_rest), _d(scene.V_prev), _d(xt), _d(scene.masses), _i(scene.boundary),
            _i(scene.tets), _d(scene.dm_inv), _d(scene.volumes), _i(scene.faces),
            np.ascontiguousarray(scene.surf_verts, dtype=np.uint64),
            np.ascontiguousarray(scene.surf_edges, dtype=np.uint64)]
    d = SceneDesc()
    d.V, d.T, d.F = scene.V, scene.T, len(scene.faces)
    d.Vs, d.Es = len(scene.surf_verts), len(scene.surf_edges)
    d.x, d.v_rest, d.V_prev, d.xtilde, d.masses = (_dptr(keep[0]), _dptr(keep[1]), _dptr(keep[2]),
                                                   _dptr(keep[3]), _dptr(keep[4]))
    d.boundary, d.tets, d.dm_inv, d.volumes, d.faces = (_iptr(keep[5]), _iptr(keep[6]), _dptr(keep[7]),
                                                        _dptr(keep[8]), _iptr(keep[9]))
    d.surf_verts, d.surf_edges = keep[10].ctypes.data_as(_u64p), keep[11].ctypes.data_as(_u64p)
    m = scene.material
    d.dt, d.length_rate, d.volume_rate, d.drag = m.dt, m.length_rate, m.volume_rate, m.drag
    d.hhat, d.fhat, d.friction = scene.hhat, scene.fhat, m.friction
    d.ground_normal[0], d.ground_normal[1], d.ground_normal[2] = 0.0, 1.0, 0.0
    d.ground_d = -1.0
    _fill_cloth(d, keep, scene)
    return d, keep


class GpuContext:
    """One scene resident on one MI355X (one context, one stream)."""

    def __init__(self, scene, xtilde=None, device: int = 0):
        L = load()
        h = C.c_void_p()
        _check(L.cipc_gpu_context_create(device, C.byref(h)))
        self.h = h
        self.scene = scene
        self.V = scene.V
        desc, keep = scene_desc(scene, xtilde)
        _check(L.cipc_gpu_upload_scene(self.h, C.byref(desc)))

    def close(self):
        if getattr(self, "h", None):
            load().cipc_gpu_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # state
    def set_state(self, x=None, V_prev=None, xtilde=None):
        a = [None if v is None else _d(v) for v in (x, V_prev, xtilde)]
        _check(load().cipc_gpu_set_state(self.h, *[None if v is None else _dptr(v) for v in a]))

    def positions(self):
        x = np.zeros((self.V, 3))
        _check(load().cipc_gpu_get_positions(self.h, _dptr(x)))
        return x

    def set_direction(self, p):
        p = _d(p)
        _check(load().cipc_gpu_set_direction(self.h, _dptr(p)))

    # broad phase + active set
    def active_set(self):
        c = np.zeros(4, np.int64)
        _check(load().cipc_gpu_active_set(self.h, c.ctypes.data_as(_i64p)))
        s4 = np.zeros((max(c[0], 1), 4), np.int32)
        e4 = np.zeros((max(c[1], 1), 4), np.int32)
        eie = np.zeros((max(c[1], 1), 2), np.int32)
        ccd = np.zeros((max(c[2], 1), 2), np.int32)
        env = np.zeros(max(c[3], 1), np.int32)
        _check(load().cipc_gpu_get_active_set(self.h, _iptr(s4), _iptr(e4), _iptr(eie), _iptr(ccd), _iptr(env)))
        return dict(self_set=s4[:c[0]], ee_set=e4[:c[1]], ee_ie=eie[:c[1]], ccd_set=ccd[:c[2]], env_set=env[:c[3]])

    def set_active_set(self, a):
        c = np.array([len(a["self_set"]), len(a["ee_set"]), len(a["ccd_set"]), len(a["env_set"])], np.int64)
        arr = [_i(a[k]).reshape(-1) if len(a[k]) else np.zeros(4, np.int32)
               for k in ("self_set", "ee_set", "ee_ie", "ccd_set", "env_set")]
        _check(load().cipc_gpu_set_active_set(self.h, c.ctypes.data_as(_i64p), *[_iptr(x) for x in arr]))

    # assembly
    def set_hessian_precision(self, precision):
        """0: FP64 elastic Hessian blocks; 1: the mixed form (FP32 elastic blocks, FP64
        barriers, friction, gradient)."""
        _check(load().cipc_gpu_set_hessian_precision(self.h, precision))

    def assemble(self, kappa):
        nnz = np.zeros(1, np.int64)
        col = np.zeros(1, np.int32)
        _check(load().cipc_gpu_assemble(self.h, kappa, nnz.ctypes.data_as(_i64p), _iptr(col)))
        n = 3 * self.V
        g = np.zeros(n)
        _check(load().cipc_gpu_get_gradient(self.h, _dptr(g)))
        outer = np.zeros(n + 1, np.int32)
        inner = np.zeros(max(int(nnz[0]), 1), np.int32)
        val = np.zeros(max(int(nnz[0]), 1))
        rhs = np.zeros(n)
        _check(load().cipc_gpu_get_csr(self.h, _iptr(outer), _iptr(inner), _dptr(val), _dptr(rhs)))
        k = int(nnz[0])
        return dict(grad=g.reshape(-1, 3), outer=outer, inner=inner[:k], val=val[:k], rhs=rhs, collisions=int(col[0]))

    def system(self, nnz):
        """The gradient and lower CSR held by the context (from the last assembly or Newton
        iteration), without assembling again; nnz as that call returned it."""
        n = 3 * self.V
        g = np.zeros(n)
        _check(load().cipc_gpu_get_gradient(self.h, _dptr(g)))
        outer = np.zeros(n + 1, np.int32)
        inner = np.zeros(max(int(nnz), 1), np.int32)
        val = np.zeros(max(int(nnz), 1))
        rhs = np.zeros(n)
        _check(load().cipc_gpu_get_csr(self.h, _iptr(outer), _iptr(inner), _dptr(val), _dptr(rhs)))
        k = int(nnz)
        return dict(grad=g.reshape(-1, 3), outer=outer, inner=inner[:k], val=val[:k], rhs=rhs)

    def energy(self, kappa):
        e = np.zeros(1)
        _check(load().cipc_gpu_energy(self.h, kappa, _dptr(e)))
        return float(e[0])

    def energy_terms(self, kappa):
        """[elastic, inertia, ground barrier, self barrier, mollified barrier, friction,
        cloth triangles, bending, soft targets]"""
        t = np.zeros(9)
        _check(load().cipc_gpu_energy_terms(self.h, kappa, _dptr(t)))
        return t

    def set_preconditioner_weights(self, w):
        """Weights of the learned preconditioner (solve(preconditioner=2))."""
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._pw = w
        _check(load().cipc_gpu_set_preconditioner_weights(self.h, w.ctypes.data_as(_fp), len(w)))

    def preconditioner_gradient(self, params, probes=None, objective=0):
        """Training loss of the last assembled system and its gradient: objective 0 for probes
        (n_z, V, 3), objective 1 the regression on the block FSAI factor."""
        params = np.ascontiguousarray(params, dtype=np.float32)
        g = np.zeros_like(params)
        loss = C.c_double(0.0)
        if probes is None:
            zp, nz = None, 0
        else:
            z = np.ascontiguousarray(probes, dtype=np.float32).reshape(len(probes), -1)
            zp, nz = z.ctypes.data_as(_fp), len(z)
        _check(load().cipc_gpu_preconditioner_gradient(self.h, objective, params.ctypes.data_as(_fp), zp, nz, C.byref(loss),
                                                       g.ctypes.data_as(_fp)))
        return loss.value, g

    def preconditioner_train_step(self, params, m1, m2, step, lr, n_probes=4, seed=0, objective=1):
        """One Adam step on the last assembled system; float32 arrays updated in place."""
        loss = C.c_double(0.0)
        _check(load().cipc_gpu_preconditioner_train_step(self.h, objective, params.ctypes.data_as(_fp), m1.ctypes.data_as(_fp),
                                                         m2.ctypes.data_as(_fp), step, lr, n_probes, seed, C.byref(loss)))
        return loss.value

    def kappa_gradients(self):
        """initKappa's reductions: (sum g_c . g_E, sum |g_c|^2, number of constraints)."""
        out = np.zeros(2)
        n = np.zeros(1, np.int64)
        _check(load().cipc_gpu_kappa_gradients(self.h, _dptr(out), n.ctypes.data_as(_i64p)))
        return out[0], out[1], int(n[0])

    def close_constraints(self, dtol):
        """postLineSearch's bookkeeping: (a recorded constraint got closer, (#ground, #self) recorded)."""
        closer = C.c_int(0)
        cnt = np.zeros(2, np.int64)
        _check(load().cipc_gpu_close_constraints(self.h, dtol, C.byref(closer), cnt.ctypes.data_as(_i64p)))
        return bool(closer.value), (int(cnt[0]), int(cnt[1]))

    def clear_close_constraints(self):
        _check(load().cipc_gpu_clear_close_constraints(self.h))

    def set_soft_targets(self, idx, targets, weight):
        """SoftBoundaryCondition: 0.5 weight |x_v - target|^2 on the listed (distinct, free)
        vertices; an empty list removes it."""
        idx = _i(np.asarray(idx).reshape(-1))
        t = _d(np.asarray(targets).reshape(-1, 3))
        n = len(idx)
        _check(load().cipc_gpu_set_soft_targets(self.h, n, _iptr(idx) if n else None, _dptr(t) if n else None, weight))

    def solve(self, precision=0, rel_tol=1e-6, max_iter=10000, preconditioner=0):
        """Newton linear solve of the last assembled system; returns (direction (V,3), stats).
        precision 0: FP64 PCG; 1: FP32 PCG inside FP64 refinement.  preconditioner 0:
        block-Jacobi; 1: block FSAI; 2: the learned factor."""
        p = np.zeros((self.V, 3))
        st = SolveStats()
        _check(load().cipc_gpu_solve_ex(self.h, precision, preconditioner, rel_tol, max_iter, _dptr(p), C.byref(st)))
        return p, dict(iterations=st.iterations, refinements=st.refinements, rel_residual=st.rel_residual,
                       solve_ms=st.solve_ms)

    def solve_factor(self):
        """The factor G of the last solve with preconditioner 1 or 2 as a scipy CSR matrix (3V x 3V)."""
        import scipy.sparse as sp
        nb = np.zeros(1, np.int64)
        _check(load().cipc_gpu_get_solve_factor(self.h, nb.ctypes.data_as(_i64p), None, None, None))
        rowptr, col, val = np.zeros(self.V + 1, np.int32), np.zeros(max(nb[0], 1), np.int32), np.zeros((max(nb[0], 1), 3, 3))
        _check(load().cipc_gpu_get_solve_factor(self.h, nb.ctypes.data_as(_i64p), _iptr(rowptr), _iptr(col), _dptr(val)))
        n = int(nb[0])
        return sp.bsr_matrix((val[:n], col[:n], rowptr), shape=(3 * self.V, 3 * self.V)).tocsr()

    def friction_update(self, kappa):
        """Friction::initialize on the current active set and positions."""
        _check(load().cipc_gpu_friction_update(self.h, kappa))

    def friction_state(self):
        n = np.zeros(2, np.int64)
        ip = n.ctypes.data_as(C.POINTER(C.c_int64))
        _check(load().cipc_gpu_get_friction(self.h, ip, None, None, None, None))
        lam, coord, basis, env = np.zeros(n[0]), np.zeros(2 * n[0]), np.zeros(6 * n[0]), np.zeros(n[1])
        _check(load().cipc_gpu_get_friction(self.h, ip, _dptr(lam), _dptr(coord), _dptr(basis), _dptr(env)))
        return dict(lam=lam, coord=coord.reshape(-1, 2), basis=basis.reshape(-1, 2, 3), env_lam=env)

    def step_bound(self, p=None):
        if p is not None:
            self.set_direction(p)
        out = np.zeros(5)
        _check(load().cipc_gpu_step_bound(self.h, _dptr(out)))
        return dict(alpha=out[0], ground=out[1], partial=out[2], cfl=out[3], full=out[4])

    def autotune(self, p=None):
        """Times the full-CCD launch splits on the current state (direction p if given) and
        keeps the fastest; returns {kernel: (chosen, [mean ms per candidate])}."""
        if p is not None:
            self.set_direction(p)
        e = (TuneEntry * 4)()
        n = C.c_int()
        _check(load().cipc_gpu_autotune(self.h, e, 4, C.byref(n)))
        names = {0: "full_ccd_pt", 1: "full_ccd_ee", 3: "active_pt", 4: "active_ee"}
        return {names[e[k].kernel]: (e[k].chosen, list(e[k].ms[:e[k].n_candidates])) for k in range(n.value)}

    def set_tune(self, kernel, choice):
        """kernel 0 / 1: full-CCD walk splits, 3 / 4: the active set's walk splits (2^choice
        lanes per query); kernel 2: the iteration graph off (0) or on (1)."""
        _check(load().cipc_gpu_set_tune(self.h, kernel, choice))

    def save_positions(self):
        _check(load().cipc_gpu_save_positions(self.h))

    def step_forward(self, alpha):
        _check(load().cipc_gpu_step_forward(self.h, alpha))

    def is_intersected(self):
        r = np.zeros(1, np.int32)
        _check(load().cipc_gpu_is_intersected(self.h, _iptr(r)))
        return bool(r[0])

    def set_velocities(self, v):
        _check(load().cipc_gpu_set_velocities(self.h, _dptr(_d(v))))

    def state(self):
        """(x, V_prev, x~, velocities), each (V, 3)."""
        out = [np.zeros((self.V, 3)) for _ in range(4)]
        _check(load().cipc_gpu_get_state(self.h, *[_dptr(a) for a in out]))
        return tuple(out)

    def time_step(self, kappa, params, dirichlet=None):
        """One implicit time step (solveIPCStep) on the device; returns (mesh.Kappa after the
        step, stats dict).  `dirichlet` = (vertices, step, fn(x (n,3), v_rest (n,3), step,
        fraction, dt) -> directions (n,3)) for animated Dirichlet vertices."""
        k = np.array([kappa], dtype=np.float64)
        st = StepStats()
        dp = None
        if dirichlet is not None:
            verts, step, fn = dirichlet
            verts = np.ascontiguousarray(verts, dtype=np.int32)
            if isinstance(fn, int):            # the address of a C cipc_dirichlet_fn
                cfn = DirichletFn(fn)
            else:
                def cb(user, n, vi, x, vr, stp, frac, dt, out):
                    xs = np.ctypeslib.as_array(x, shape=(n, 3)).copy()
                    rs = np.ctypeslib.as_array(vr, shape=(n, 3)).copy()
                    d = np.asarray(fn(xs, rs, stp, frac, dt), dtype=np.float64).reshape(n, 3)
                    np.ctypeslib.as_array(out, shape=(n, 3))[:] = d
                    return 0

                cfn = DirichletFn(cb)
            dp = Dirichlet(len(verts), verts.ctypes.data_as(C.POINTER(C.c_int)), step, cfn, None)
            self._keep = (verts, cfn)
        _check(load().cipc_gpu_time_step(self.h, C.byref(params), _dptr(k), None if dp is None else C.byref(dp), C.byref(st)))
        return float(k[0]), {f: getattr(st, f) for f, _ in StepStats._fields_}

    def newton_iteration(self, kappa):
        t = IterationTiming()
        _check(load().cipc_gpu_newton_iteration(self.h, kappa, C.byref(t)))
        return {f: getattr(t, f) for f, _ in IterationTiming._fields_}


def step_params(scene, precision=0, preconditioner=1, rel_tol=1e-6, max_iter=20000, max_newton=10000):
    """The time step's scalars of a scene (mesh3D fields the descriptor does not carry)."""
    return StepParams(scene.bbox_diag2, scene.mean_mass, scene.dtol, scene.material.newton_threshold,
                      1 if scene.apply_gravity else 0, 0, precision, preconditioner, rel_tol, max_iter, max_newton)


class BatchContext(GpuContext):
    """Independent tet scenes packed into one context (cipc_gpu_upload_batch): each scene
    is evaluated as a context holding it alone would evaluate it.  Arrays over the whole
    batch (positions, gradient, CSR) list scene k's vertices after scene k-1's; the
    per-scene methods take and return one value per scene."""

    def __init__(self, scenes, xtildes=None, device: int = 0):
        L = load()
        h = C.c_void_p()
        _check(L.cipc_gpu_context_create(device, C.byref(h)))
        self.h = h
        self.scenes = list(scenes)
        self.S = len(self.scenes)
        xts = xtildes if xtildes is not None else [None] * self.S
        descs, self._keep = [], []
        for sc, xt in zip(self.scenes, xts):
            d, keep = scene_desc(sc, xt)
            descs.append(d)
            self._keep.append(keep)
        ptrs = (C.POINTER(SceneDesc) * self.S)(*[C.pointer(d) for d in descs])
        _check(L.cipc_gpu_upload_batch(self.h, ptrs, self.S))
        self._keep = None
        n = np.zeros(1, np.int32)
        self.voff = np.zeros(self.S + 1, np.int32)
        self.toff = np.zeros(self.S + 1, np.int32)
        _check(L.cipc_gpu_batch_size(self.h, _iptr(n), _iptr(self.voff), _iptr(self.toff)))
        assert int(n[0]) == self.S
        self.V = int(self.voff[-1])
        self.scene = None

    def _k(self, kappa):
        k = np.broadcast_to(np.asarray(kappa, np.float64), (self.S,))
        return np.ascontiguousarray(k)

    def split(self, a):
        """Per-scene views of a per-vertex array of the batch."""
        return [a[self.voff[k]:self.voff[k + 1]] for k in range(self.S)]

    def assemble(self, kappa):
        k = self._k(kappa)
        nnz = np.zeros(1, np.int64)
        col = np.zeros(self.S, np.int32)
        _check(load().cipc_gpu_batch_assemble(self.h, _dptr(k), nnz.ctypes.data_as(_i64p), _iptr(col)))
        n = 3 * self.V
        g = np.zeros(n)
        _check(load().cipc_gpu_get_gradient(self.h, _dptr(g)))
        outer = np.zeros(n + 1, np.int32)
        inner = np.zeros(max(int(nnz[0]), 1), np.int32)
        val = np.zeros(max(int(nnz[0]), 1))
        rhs = np.zeros(n)
        _check(load().cipc_gpu_get_csr(self.h, _iptr(outer), _iptr(inner), _dptr(val), _dptr(rhs)))
        m = int(nnz[0])
        return dict(grad=g.reshape(-1, 3), outer=outer, inner=inner[:m], val=val[:m], rhs=rhs, collisions=col)

    def friction_update(self, kappa):
        k = self._k(kappa)
        _check(load().cipc_gpu_batch_friction_update(self.h, _dptr(k)))

    def energy(self, kappa):
        k = self._k(kappa)
        e = np.zeros(self.S)
        _check(load().cipc_gpu_batch_energy(self.h, _dptr(k), _dptr(e), None))
        return e

    def energy_terms(self, kappa=0.0):
        k = self._k(kappa)
        t = np.zeros((self.S, 9))
        _check(load().cipc_gpu_batch_energy(self.h, _dptr(k), None, _dptr(t)))
        return t

    def step_bound(self, p=None):
        if p is not None:
            self.set_direction(p)
        out = np.zeros((self.S, 5))
        _check(load().cipc_gpu_batch_step_bound(self.h, _dptr(out)))
        return dict(alpha=out[:, 0], ground=out[:, 1], partial=out[:, 2], cfl=out[:, 3], full=out[:, 4])

    def step_forward(self, alpha):
        a = self._k(alpha)
        _check(load().cipc_gpu_batch_step_forward(self.h, _dptr(a)))

    def is_intersected(self):
        r = np.zeros(self.S, np.int32)
        _check(load().cipc_gpu_batch_is_intersected(self.h, _iptr(r)))
        return r.astype(bool)

    def solve(self, precision=0, rel_tol=1e-6, max_iter=10000, preconditioner=0, active=None):
        """Per-scene Newton solves in lock step; returns (direction (V,3), list of stats)."""
        p = np.zeros((self.V, 3))
        st = (SolveStats * self.S)()
        act = None if active is None else _i(np.asarray(active).astype(np.int32))
        _check(load().cipc_gpu_batch_solve(self.h, precision, preconditioner, rel_tol, max_iter,
                                           None if act is None else _iptr(act), _dptr(p), st))
        return p, [dict(iterations=s.iterations, refinements=s.refinements, rel_residual=s.rel_residual,
                        solve_ms=s.solve_ms) for s in st]

    def newton_iteration(self, kappa):
        k = self._k(kappa)
        t = IterationTiming()
        a = np.zeros(self.S)
        e = np.zeros(self.S)
        _check(load().cipc_gpu_batch_newton_iteration(self.h, _dptr(k), C.byref(t), _dptr(a), _dptr(e)))
        out = {f: getattr(t, f) for f, _ in IterationTiming._fields_}
        out["alpha_scenes"] = a
        out["energy_scenes"] = e
        return out


# ---- element-level evaluation -----------------------------------------------------------
def eval_snh(F, u, r):
    """F: (n,3,3) -> PK1 (n,3,3), projected Hessian (n,9,9) over column-major vec F, psi (n,)."""
    F = np.asarray(F, np.float64)
    n = F.shape[0]
    Fc = _d(F.transpose(0, 2, 1).reshape(n, 9))
    P, H, e = np.zeros((n, 9)), np.zeros((n, 81)), np.zeros(n)
    _check(load().cipc_gpu_eval_snh(n, _dptr(Fc), u, r, _dptr(P), _dptr(H), _dptr(e)))
    return P.reshape(n, 3, 3).transpose(0, 2, 1), H.reshape(n, 9, 9), e


def preconditioner_param_count():
    return int(load().cipc_gpu_preconditioner_param_count())


def eval_cloth(F, stretch, shear, strain_rate):
    """F: (n,3,2) -> PK1 (n,3,2), projected 6x6 Hessian over column-major vec F, energy density."""
    F = np.asarray(F, np.float64)
    n = F.shape[0]
    Fc = _d(F.transpose(0, 2, 1).reshape(n, 6))
    P, H, e = np.zeros((n, 6)), np.zeros((n, 36)), np.zeros(n)
    _check(load().cipc_gpu_eval_cloth(n, _dptr(Fc), stretch, shear, strain_rate, _dptr(P), _dptr(H), _dptr(e)))
    return P.reshape(n, 2, 3).transpose(0, 2, 1), H.reshape(n, 6, 6), e


def eval_hinge(pts, rest, weight, plate):
    """pts: (n,4,3) hinge points -> gradient (n,12), projected Hessian (n,12,12), energy (n,),
    before the dt^2 factor."""
    pts = _d(np.asarray(pts).reshape(-1, 12))
    n = pts.shape[0]
    r, w = _d(np.broadcast_to(rest, (n,))), _d(np.broadcast_to(weight, (n,)))
    g, H, e = np.zeros((n, 12)), np.zeros((n, 144)), np.zeros(n)
    _check(load().cipc_gpu_eval_hinge(n, _dptr(pts), _dptr(r), _dptr(w), plate, _dptr(g), _dptr(H), _dptr(e)))
    return g, H.reshape(n, 12, 12), e


def eval_distance(kind, pts):
    pts = _d(np.asarray(pts).reshape(-1, 12))
    n = pts.shape[0]
    d, g, H = np.zeros(n), np.zeros((n, 12)), np.zeros((n, 144))
    _check(load().cipc_gpu_eval_distance(n, kind, _dptr(pts), _dptr(d), _dptr(g), _dptr(H)))
    N = {0: 6, 1: 9}.get(kind, 12)
    return d, g[:, :N].copy(), H[:, :N * N].reshape(n, N, N).copy()


def eval_make_pd(H):
    H = _d(H).copy()
    n, dim = H.shape[0], H.shape[1]
    _check(load().cipc_gpu_eval_make_pd(n, dim, _dptr(H)))
    return H


def eval_contact_pd(H, G=None):
    """Projection of translation-invariant contact Hessians (n, 3m, 3m), m vertices; with the
    blocks' gradients G (n, 3m) the structured path used for distance-barrier blocks."""
    H = _d(H).copy()
    n, dim = H.shape[0], H.shape[1]
    g = None
    if G is not None:
        g = _d(np.asarray(G).reshape(n, dim))
    _check(load().cipc_gpu_eval_contact_pd(n, dim // 3, _dptr(H), _dptr(g) if g is not None else None))
    return H


def eval_accd(kind, pts, dirs, eta, thickness=0.0):
    pts, dirs = _d(np.asarray(pts).reshape(-1, 12)), _d(np.asarray(dirs).reshape(-1, 12))
    n = pts.shape[0]
    out = np.zeros(n)
    _check(load().cipc_gpu_eval_accd(n, kind, _dptr(pts), _dptr(dirs), eta, thickness, _dptr(out)))
    return out


def eval_classify(kind, pts):
    pts = _d(np.asarray(pts).reshape(-1, 12))
    n = pts.shape[0]
    out = np.zeros(n, np.int32)
    _check(load().cipc_gpu_eval_classify(n, kind, _dptr(pts), _iptr(out)))
    return out
