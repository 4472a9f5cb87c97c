"""bench.py's contract on the CPU: more than one GPU is refused with a message, and the
algorithmic byte count of the elastic stage follows its definition (checked on a single
tet and on a lattice, where the number of distinct vertex pairs is known in closed form)."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "single" in r.stderr


def test_elastic_algorithmic_bytes():
    sys.path.insert(0, ROOT)
    import bench
    from cipc import scenes
    # one tet: 4 vertices, 6 coupled pairs -> nnz = 6 * 4 + 9 * 6
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    s = scenes.finalize("tet", x, np.array([[0, 1, 2, 3]]), scenes.Material())
    b, nnz = bench.elastic_algorithmic_bytes(s)
    assert nnz == 6 * 4 + 9 * 6
    assert b == 24 * 4 + 96 + 24 * 4 + 8 * nnz
    # n^3 lattice, 5-tet split: the coupled pairs are the lattice edges plus one diagonal per
    # face of every cell (the split's face diagonals are shared between neighbours) plus none
    # inside the cells, i.e. 3 n (n+1)^2 axis edges + 3 n^2 (n+1) face diagonals
    n = 3
    s = scenes.stacked_cubes(1, n, 0.1, 0.05, seed=0, rotate="none", jitter=0.0)
    b, nnz = bench.elastic_algorithmic_bytes(s)
    E = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1)
    assert nnz == 6 * s.V + 9 * E
