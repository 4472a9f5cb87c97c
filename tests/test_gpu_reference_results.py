"""The reference's own published runs, reproduced on the GPU through the headless driver
(cpu-ipc_amd/cipc_headless: the host scene set-up, then every time step on the device by
cipc_gpu_time_step): the twisting mat with animated Dirichlet ends (1 and 5 steps) and with
soft targets (5 steps).  Bars: the last step's Newton iterations and matrix nnz exactly,
squared_norm_sum within 1e-10 relative of the published value (tests/golden/README.md),
with each Newton system solved by FP64 block-FSAI PCG to 1e-12 relative residual; metrics.csv
holds one row per step with the reference's columns.
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cpu-ipc_amd", "cipc_headless")
ASSETS = os.path.join(ROOT, "tests", "golden", "reference_assets")


def _run(scene, steps, out):
    cmd = [EXE, "--scene", scene, "--steps", str(steps), "--assets", ASSETS, "--output", str(out), "--solve-tol", "1e-12"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT ")][-1]
    print(line, flush=True)
    return dict(kv.split("=", 1) for kv in line.split()[1:])


@pytest.mark.parametrize("scene,steps,newton,nnz,sq", [
    ("twisting-mat", 1, 2, 151335, 560.74908434861868),
    ("twisting-mat", 5, 2, 151335, 560.775874160993),
    ("twisting-mat-soft", 5, 1, 158907, 560.77059938845662),
])
def test_headless_reproduces_reference_results(tmp_path, scene, steps, newton, nnz, sq):
    r = _run(scene, steps, tmp_path)
    assert int(r["steps"]) == steps and int(r["vertices"]) == 3200
    assert int(r["last_newton"]) == newton
    assert int(r["last_matrix_nnz"]) == nnz
    got = float(r["squared_norm_sum"])
    assert abs(got - sq) <= 1e-10 * sq, (got, sq)
    rows = open(tmp_path / "metrics.csv").read().splitlines()
    assert rows[0].startswith("frame,step_ms,assembly_ms") and len(rows) == steps + 1
