"""The mixed form of the Newton assembly: elastic Hessian blocks (tets, triangles, bending)
rounded to FP32 where they are staged, everything else (barriers, friction, mass, gradient,
right-hand side) in FP64.  No reference counterpart: the reference assembles in FP64.

Bars: CSR pattern, gradient and right-hand side identical to the FP64 assembly; CSR values
within 1e-6 of the largest magnitude (FP32 rounding of each contribution, at most a few
dozen per entry, summed in FP64); the mixed system solved to 1e-6 relative residual.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from cipc import gpu as g
    g.load()
    return g


def _scene(name):
    from cipc import scenes
    if name == "4x20cubes":
        return scenes.four_cubes20(velocity_scale=0.5)
    return scenes.cloth_on_sphere(24, sphere_cells=4, radius=0.25, fold=16, bending=1, seed=2, name=name)


@pytest.mark.parametrize("name", ["4x20cubes", "cloth24_hinge"])
def test_mixed_hessian_matches_fp64(gpu, name):
    from cipc import scenes
    s = _scene(name)
    xt, p = scenes.newton_probe_state(s)
    kappa = scenes.suggest_kappa(s.hhat, s.bbox_diag2, s.mean_mass)
    g = gpu.GpuContext(s, xtilde=xt)
    try:
        g.set_state(V_prev=s.x - s.velocities * s.material.dt)
        g.active_set()
        g.friction_update(kappa)
        r64 = g.assemble(kappa)
        g.set_hessian_precision(1)
        r32 = g.assemble(kappa)
        assert np.array_equal(r64["outer"], r32["outer"]) and np.array_equal(r64["inner"], r32["inner"]), name
        assert np.array_equal(r64["grad"], r32["grad"]) and np.array_equal(r64["rhs"], r32["rhs"]), name
        scale = np.abs(r64["val"]).max()
        err = np.abs(r32["val"] - r64["val"]).max() / scale
        assert 0 < err <= 1e-6, (name, err)
        _, st = g.solve(precision=1, rel_tol=1e-6, preconditioner=1)
        assert st["rel_residual"] <= 1e-6
        g.set_hessian_precision(0)
        r = g.assemble(kappa)
        assert np.array_equal(r["val"], r64["val"])      # back to FP64
    finally:
        g.close()


def test_hessian_precision_rejects_other_values(gpu):
    from cipc import scenes
    g = gpu.GpuContext(scenes.cube10())
    try:
        with pytest.raises(ValueError):
            g.set_hessian_precision(2)
    finally:
        g.close()
