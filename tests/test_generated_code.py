"""The device's generated distance derivatives (cpu-ipc_amd/csrc/distance_gen.inc),
compiled for the host, against central finite differences and against the oracle's
automatic differentiation."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = r"""
#define __device__
#define __restrict__
#include "distance_gen.inc"
extern "C" void eval(int f, const double* x, double* g, double* H) {
    if (f == 1) dist_pe_gh(x, g, H); else if (f == 2) dist_pt_gh(x, g, H);
    else if (f == 3) dist_ee_gh(x, g, H); else ee_cross_gh(x, g, H);
}
"""


@pytest.fixture(scope="module")
def gen(tmp_path_factory):
    d = tmp_path_factory.mktemp("gen")
    src = d / "gen.cpp"
    src.write_text(SRC)
    lib = d / "libgen.so"
    subprocess.run(["g++", "-O1", "-shared", "-fPIC", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "cpu-ipc_amd", "csrc"), str(src), "-o", str(lib)], check=True)
    L = C.CDLL(str(lib))
    dp = C.POINTER(C.c_double)
    L.eval.argtypes = [C.c_int, dp, dp, dp]

    def f(kind, x):
        x = np.ascontiguousarray(x, np.float64)
        g, H = np.zeros(12), np.zeros(144)
        L.eval(kind, x.ctypes.data_as(dp), g.ctypes.data_as(dp), H.ctypes.data_as(dp))
        n = 9 if kind == 1 else 12
        return g[:n], H[:n * n].reshape(n, n)
    return f


@pytest.mark.parametrize("kind", [1, 2, 3, 4])
def test_generated_derivatives(gen, oracle_lib, kind):
    rng = np.random.default_rng(kind)
    h = 1e-6
    for _ in range(30):
        x = rng.standard_normal(12)
        g, H = gen(kind, x)
        _, go, Ho = oracle_lib.distance(kind, x)
        assert np.allclose(g, go, rtol=1e-11, atol=1e-11 * np.abs(go).max())
        assert np.allclose(H, Ho, rtol=1e-10, atol=1e-10 * np.abs(Ho).max())
        n = len(g)
        for j in range(n):
            xp, xm = x.copy(), x.copy()
            xp[j] += h
            xm[j] -= h
            fd = (gen(kind, xp)[0] - gen(kind, xm)[0]) / (2 * h)
            assert np.allclose(fd, H[:, j], atol=1e-5 * max(1.0, np.abs(H).max()))
