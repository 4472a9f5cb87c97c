"""ctypes binding of the host-side scene set-up (include/cipc_host.h): the reference's mesh
loaders, getSurface, initMesh3D, parameter file, named scenes and mesh / checkpoint files,
producing a ``scenes.Scene`` that the device contexts take."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import gpu, scenes

_dp = C.POINTER(C.c_double)


def _lib():
    L = gpu.load()
    if not getattr(L, "_host_sigs", False):
        sig = {
            "cipc_host_mesh_create": [C.c_char_p, C.POINTER(C.c_void_p)],
            "cipc_host_mesh_destroy": [C.c_void_p],
            "cipc_host_load_tetrahedra": [C.c_void_p, C.c_char_p, C.c_double, _dp],
            "cipc_host_load_triangles": [C.c_void_p, C.c_char_p, C.c_double, _dp, C.c_int],
            "cipc_host_finalize": [C.c_void_p],
            "cipc_host_scene_build": [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)],
            "cipc_host_counts": [C.c_void_p, C.POINTER(C.c_int64)],
            "cipc_host_scene_desc": [C.c_void_p, C.POINTER(gpu.SceneDesc), C.POINTER(gpu.StepParams)],
            "cipc_host_boundary": [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_int)),
                                   _dp],
            "cipc_host_twisting_targets": [C.c_int, _dp, _dp, C.c_double, _dp],
            "cipc_host_set_state": [C.c_void_p, _dp, _dp],
            "cipc_host_write_msh": [C.c_void_p, C.c_char_p],
            "cipc_host_save_checkpoint": [C.c_void_p, C.c_char_p],
            "cipc_host_load_checkpoint": [C.c_void_p, C.c_char_p],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.restype = C.c_int
            fn.argtypes = args
        L.cipc_host_twisting_directions.restype = C.c_int
        L._host_sigs = True
    return L


class _Rates(scenes.Material):
    """A material given by its SNH rates (lengthRate, volumeRate) as the mesh holds them."""

    def __init__(self, length_rate, volume_rate):
        super().__init__()
        self._lr, self._vr = length_rate, volume_rate

    @property
    def length_rate(self):
        return self._lr

    @property
    def volume_rate(self):
        return self._vr


def _b(s):
    return None if s is None else str(s).encode()


def twisting_dirichlet_fn():
    """The library's twistingBoundaryDirection as a cipc_dirichlet_fn pointer."""
    return C.cast(_lib().cipc_host_twisting_directions, C.c_void_p).value


class HostMesh:
    """A mesh3D built on the host by the library (loaders, getSurface, initMesh3D)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def create(cls, settings=None):
        h = C.c_void_p()
        gpu._check(_lib().cipc_host_mesh_create(_b(settings), C.byref(h)))
        return cls(h)

    @classmethod
    def scene(cls, name, asset_dir, settings):
        h = C.c_void_p()
        gpu._check(_lib().cipc_host_scene_build(_b(name), _b(asset_dir), _b(settings), C.byref(h)))
        return cls(h)

    def close(self):
        if self.h:
            _lib().cipc_host_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_tetrahedra(self, path, scale=1.0, offset=(0.0, 0.0, 0.0)):
        o = np.ascontiguousarray(offset, dtype=np.float64)
        gpu._check(_lib().cipc_host_load_tetrahedra(self.h, _b(path), scale, o.ctypes.data_as(_dp)))

    def load_triangles(self, path, scale=1.0, offset=(0.0, 0.0, 0.0), type=0):
        o = np.ascontiguousarray(offset, dtype=np.float64)
        gpu._check(_lib().cipc_host_load_triangles(self.h, _b(path), scale, o.ctypes.data_as(_dp), type))

    def finalize(self):
        gpu._check(_lib().cipc_host_finalize(self.h))

    def counts(self):
        c = np.zeros(9, np.int64)
        gpu._check(_lib().cipc_host_counts(self.h, c.ctypes.data_as(C.POINTER(C.c_int64))))
        return dict(zip(("V", "T", "F", "Vs", "Es", "Tr", "Nb", "dirichlet", "soft"), (int(v) for v in c)))

    def desc(self):
        d, p = gpu.SceneDesc(), gpu.StepParams()
        gpu._check(_lib().cipc_host_scene_desc(self.h, C.byref(d), C.byref(p)))
        return d, p

    def boundary(self):
        motion = C.c_int()
        dp, sp = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
        w = C.c_double()
        gpu._check(_lib().cipc_host_boundary(self.h, C.byref(motion), C.byref(dp), C.byref(sp), C.byref(w)))
        c = self.counts()
        dv = np.ctypeslib.as_array(dp, shape=(c["dirichlet"],)).copy() if c["dirichlet"] else np.zeros(0, np.int32)
        sv = np.ctypeslib.as_array(sp, shape=(c["soft"],)).copy() if c["soft"] else np.zeros(0, np.int32)
        return motion.value, dv, sv, w.value

    def arrays(self):
        """Copies of the mesh3D arrays: x, v_rest, masses, boundary, tets, dm_inv, volumes,
        faces, surf_verts, surf_edges."""
        d, _ = self.desc()
        V, T, F, Vs, Es = d.V, d.T, d.F, d.Vs, d.Es

        def arr(p, shape):
            return np.ctypeslib.as_array(p, shape=shape).copy() if int(np.prod(shape)) else np.zeros(shape)

        return dict(x=arr(d.x, (V, 3)), v_rest=arr(d.v_rest, (V, 3)), masses=arr(d.masses, (V,)),
                    boundary=arr(d.boundary, (V,)).astype(np.int32), tets=arr(d.tets, (T, 4)).astype(np.int32),
                    dm_inv=arr(d.dm_inv, (T, 9)), volumes=arr(d.volumes, (T,)), faces=arr(d.faces, (F, 4)).astype(np.int32),
                    surf_verts=arr(d.surf_verts, (Vs,)).astype(np.int64), surf_edges=arr(d.surf_edges, (Es, 2)).astype(np.int64))

    def to_scene(self, name="host", friction=None):
        """A scenes.Scene over copies of the arrays, with the mesh's scalars."""
        d, p = self.desc()
        a = self.arrays()
        mat = _Rates(d.length_rate, d.volume_rate)
        mat.dt = d.dt
        mat.drag = d.drag
        mat.friction = d.friction if friction is None else friction
        mat.newton_threshold = p.newton_threshold
        s = scenes.Scene(name=name, x=a["x"], v_rest=a["v_rest"], V_prev=a["x"].copy(), velocities=np.zeros_like(a["x"]),
                         tets=a["tets"], dm_inv=a["dm_inv"], volumes=a["volumes"], masses=a["masses"], boundary=a["boundary"],
                         faces=a["faces"], surf_verts=a["surf_verts"].astype(np.int32),
                         surf_edges=a["surf_edges"].astype(np.int32), material=mat, bbox_diag2=p.bbox_diag2, hhat=d.hhat,
                         fhat=d.fhat, dtol=p.dtol, mean_mass=p.mean_mass, avg_edge_len=0.0, apply_gravity=bool(p.apply_gravity))
        return s

    def write_msh(self, path):
        gpu._check(_lib().cipc_host_write_msh(self.h, _b(path)))

    def set_state(self, x=None, xtilde=None):
        xs = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        ts = None if xtilde is None else np.ascontiguousarray(xtilde, dtype=np.float64)
        gpu._check(_lib().cipc_host_set_state(self.h, None if xs is None else xs.ctypes.data_as(_dp),
                                              None if ts is None else ts.ctypes.data_as(_dp)))

    def save_checkpoint(self, directory):
        gpu._check(_lib().cipc_host_save_checkpoint(self.h, _b(directory)))

    def load_checkpoint(self, directory):
        gpu._check(_lib().cipc_host_load_checkpoint(self.h, _b(directory)))


def twisting_targets(x, v_rest, dt):
    x = np.ascontiguousarray(x, dtype=np.float64)
    r = np.ascontiguousarray(v_rest, dtype=np.float64)
    out = np.zeros_like(x)
    gpu._check(_lib().cipc_host_twisting_targets(len(x), x.ctypes.data_as(_dp), r.ctypes.data_as(_dp), dt, out.ctypes.data_as(_dp)))
    return out
