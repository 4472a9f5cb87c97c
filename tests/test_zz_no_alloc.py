"""The allocation-free energy (cipc_gpu_energy_terms_noalloc) against cipc_gpu_energy_terms
on the same state, through the C ABI: two calls in a row on one GPU, the second on a smaller
scene, each bit-identical to the allocating entry point (the same kernels in the same order);
a batch context and a host output pointer are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx(s):
    from cipc import gpu, scenes
    xt, _ = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    g = gpu.GpuContext(s, xtilde=xt)
    g.set_state(V_prev=s.x - s.velocities * s.material.dt)
    g.active_set()
    g.friction_update(kappa)
    return g, kappa


def test_noalloc_energy_matches_energy_terms():
    from cipc import gpu, scenes
    from hipmem import DeviceArray
    L = gpu.load()
    out = DeviceArray(9)
    big = scenes.stacked_cubes(2, 8, 0.1, 3e-4, seed=5, rotate="yaw", velocity_scale=0.3, name="big")
    small = scenes.stacked_cubes(1, 3, 0.1, 3e-4, seed=6, rotate="yaw", velocity_scale=0.3, name="small")
    try:
        for s in (big, small):
            g, kappa = _ctx(s)
            try:
                out.fill(np.nan)
                want = np.array(g.energy_terms(kappa))
                gpu._check(L.cipc_gpu_energy_terms_noalloc(g.h, kappa, out.ptr))
                got = out.numpy()
                print(s.name, got)
                assert np.array_equal(got, want), (s.name, got, want)
                assert want[2] > 0                           # the ground barrier is in it
            finally:
                g.close()
    finally:
        out.close()


def test_noalloc_energy_refuses_host_output_and_batches():
    from cipc import gpu, scenes
    L = gpu.load()
    s = scenes.stacked_cubes(1, 3, 0.1, 3e-4, seed=6, rotate="yaw", name="small")
    g, kappa = _ctx(s)
    host = np.zeros(9)
    try:
        with pytest.raises(ValueError, match="device memory"):
            gpu._check(L.cipc_gpu_energy_terms_noalloc(g.h, kappa, host.ctypes.data_as(C.c_void_p)))
    finally:
        g.close()
    b = gpu.BatchContext(scenes.batched_cubes(2, seed=1, landed=True))
    from hipmem import DeviceArray
    out = DeviceArray(9)
    try:
        with pytest.raises(gpu.CipcLogicError):
            gpu._check(L.cipc_gpu_energy_terms_noalloc(b.h, 1.0, out.ptr))
    finally:
        b.close()
        out.close()
