"""Host-side scene set-up (include/cipc_host.h) and the oracle's whole time step, pinned by
the reference's own published runs.

* The twisting mat built by the library's loader and scene builder, stepped by the oracle
  (oracle/step_oracle.py), reproduces the reference's RESULT lines: Newton iterations and
  matrix nnz exactly, squared_norm_sum to 1e-12 relative (tests/golden/README.md).
* cloth-bunny and bunny2 load with the reference's vertex, tet and surface-face counts
  (when the reference tree is present).
* The Gmsh loader reads compact rows, standard Gmsh 2.2 rows with tags and CRLF line ends
  alike; the surface has the boundary faces and edges of the same mesh that the Python
  scene module derives independently; checkpoint and mesh files round-trip.
"""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASSETS = os.path.join(ROOT, "tests", "golden", "reference_assets") + "/"
SETTINGS = ASSETS + "scene/parameterSetting.txt"
REF_ASSETS = "/root/reference/Assets/"


def twisting(x, rest, step, fraction, dt):
    """twistingBoundaryDirection (CPU IPC/Simulator.cpp:24-46): the rotation about x by
    -a (rest x < 0) or +a (rest x > 0), a = 3.14/5 dt fraction, as the move x - d = R x."""
    a = 3.14 / 5 * dt * fraction
    c, s = math.cos(a), math.sin(a)
    out = np.zeros_like(x)
    for i in range(len(x)):
        if rest[i, 0] < 0:
            sg = 1.0
        elif rest[i, 0] > 0:
            sg = -1.0
        else:
            continue
        R = ((1.0, 0.0, 0.0), (0.0, c, sg * s), (0.0, -(sg * s), c))
        for k in range(3):
            out[i, k] = ((R[k][0] * x[i, 0] + R[k][1] * x[i, 1]) + R[k][2] * x[i, 2]) - x[i, k]
    return out


def _run_oracle_twist(oracle_lib, scene, steps):
    from cipc import host
    from step_oracle import OracleStep
    m = host.HostMesh.scene(scene, ASSETS, SETTINGS)
    s = m.to_scene(scene)
    _, dv, sv, w = m.boundary()
    om = oracle_lib.OracleMesh(s, xtilde=s.x.copy())
    st = OracleStep(om, s, quasi_static=True)
    for step in range(steps):
        if len(sv):
            x = om.positions()
            om.set_soft_targets(sv, x[sv] - twisting(x[sv], s.v_rest[sv], step, 1.0, s.material.dt), w)
            r = st.step()
        else:
            r = st.step(dirichlet=(dv, step, twisting))
    nnz = len(om.assemble(st.mesh_kappa)["val"])
    om.close()
    m.close()
    return r, float(np.sum(st.x * st.x)), nnz


@pytest.mark.parametrize("scene,steps,newton,nnz,sq", [
    ("twisting-mat", 1, 2, 151335, 560.74908434861868),
    ("twisting-mat", 5, 2, 151335, 560.775874160993),
    ("twisting-mat-soft", 5, 1, 158907, 560.77059938845662),
])
def test_oracle_reproduces_reference_results(oracle_lib, scene, steps, newton, nnz, sq):
    r, got, n = _run_oracle_twist(oracle_lib, scene, steps)
    assert r["newton_iterations"] == newton
    assert n == nnz
    assert abs(got - sq) <= 1e-12 * sq, (got, sq)


def test_twisting_mat_counts():
    from cipc import host
    m = host.HostMesh.scene("twisting-mat", ASSETS, SETTINGS)
    c = m.counts()
    assert (c["V"], c["T"], c["dirichlet"], c["soft"]) == (3200, 9126, 160, 0)
    motion, dv, sv, w = m.boundary()
    a = m.arrays()
    assert motion == 1 and np.array_equal(a["boundary"][dv], np.ones(len(dv)))
    assert np.all(np.abs(np.abs(a["x"][dv, 0]) - 0.5) < 1e-4)
    soft = host.HostMesh.scene("twisting-mat-soft", ASSETS, SETTINGS)
    assert soft.counts()["soft"] == 160 and soft.boundary()[3] == 100.0


@pytest.mark.skipif(not os.path.isdir(REF_ASSETS), reason="the reference's asset tree is absent")
def test_reference_scene_counts():
    from cipc import host
    s = REF_ASSETS + "scene/parameterSetting.txt"
    c = host.HostMesh.scene("cloth-bunny", REF_ASSETS, s).counts()
    assert (c["V"], c["T"], c["F"]) == (2893, 7356, 4074)
    assert host.HostMesh.scene("bunny2", REF_ASSETS, s).counts()["F"] == 41664


def test_surface_matches_independent_derivation():
    """The face set (with orientation), the surface vertex set and the surface edges equal
    those the scene module derives with its own (sorted) face order."""
    from cipc import host, scenes
    m = host.HostMesh.scene("twisting-mat", ASSETS, SETTINGS)
    a = m.arrays()
    faces, sv, se = scenes.extract_surface(a["x"], a["tets"].astype(np.int64))
    canon = lambda f: tuple(np.roll(f[:3], -int(np.argmin(f[:3])))) + (int(f[3]),)
    assert sorted(map(canon, a["faces"])) == sorted(map(canon, faces))
    assert set(a["surf_verts"].tolist()) == set(np.asarray(sv).tolist())
    und = lambda e: sorted(map(tuple, np.sort(np.asarray(e, np.int64), axis=1).tolist()))
    assert und(a["surf_edges"]) == und(se)
    # the directed edges and the vertex order follow the face order by the reference's rules
    # (SimulationMesh.cpp:422-445)
    edges, seen, have = set(), [], set()
    for f in a["faces"]:
        for k in range(3):
            e = (int(f[k]), int(f[(k + 1) % 3]))
            if (e[1], e[0]) not in edges:
                edges.add(e)
    assert [tuple(e) for e in a["surf_edges"].tolist()] == sorted(edges)
    for f in a["faces"]:
        for v in f[:3]:
            if int(v) not in have:
                have.add(int(v))
                seen.append(int(v))
    assert a["surf_verts"].tolist() == seen
    # the reference's iteration order, not a sorted one
    assert not np.array_equal(a["faces"], faces)


def _write(path, text, crlf=False):
    with open(path, "wb") as f:
        f.write(text.replace("\n", "\r\n" if crlf else "\n").encode())


def test_gmsh_row_formats(tmp_path):
    from cipc import host
    nodes = [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (1.0, 1.0, 1.0)]
    tets = [(1, 2, 3, 4), (2, 3, 4, 5)]
    compact = "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n5\n" + "".join(f"{x} {y} {z}\n" for x, y, z in nodes) + \
        "$EndNodes\n$Elements\n2\n" + "".join(f"{i + 1} {a} {b} {c} {d}\n" for i, (a, b, c, d) in enumerate(tets)) + "$EndElements\n"
    # standard Gmsh 2.2: ids out of order, a line element to skip, two tags, trailing spaces
    ids = [7, 3, 9, 1, 4]
    std = "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n5\n" + \
        "".join(f"{ids[i]} {x} {y} {z} \n" for i, (x, y, z) in enumerate(nodes)) + "$EndNodes\n$Elements\n3\n" + \
        f"1 1 2 0 1 {ids[0]} {ids[1]}\n" + \
        "".join(f"{i + 2} 4 2 0 1 {ids[a - 1]} {ids[b - 1]} {ids[c - 1]} {ids[d - 1]} \n" for i, (a, b, c, d) in enumerate(tets)) + \
        "$EndElements\n"
    out = []
    for name, text, crlf in (("c.msh", compact, False), ("s.msh", std, False), ("s_crlf.msh", std, True)):
        p = str(tmp_path / name)
        _write(p, text, crlf)
        m = host.HostMesh.create(SETTINGS)
        m.load_tetrahedra(p, 2.0, (0.5, 0.0, -1.0))
        m.finalize()
        out.append(m.arrays())
    for a in out:
        assert np.array_equal(a["tets"], np.array(tets) - 1)
        assert np.array_equal(a["x"], np.array(nodes) * 2.0 + np.array([0.5, 0.0, -1.0]))
        assert np.allclose(a["volumes"], [8.0 / 6.0, 8.0 / 6.0 * 2.0])
    bad = str(tmp_path / "bad.msh")
    _write(bad, "$Nodes\n2\n1 0 0 0\n1 1 1 1\n$EndNodes\n")
    m = host.HostMesh.create(SETTINGS)
    with pytest.raises(RuntimeError, match="invalid Gmsh node"):
        m.load_tetrahedra(bad)


def test_mesh_and_checkpoint_files_round_trip(tmp_path):
    from cipc import host
    m = host.HostMesh.scene("twisting-mat", ASSETS, SETTINGS)
    a = m.arrays()
    p = str(tmp_path / "out.msh")
    m.write_msh(p)
    m2 = host.HostMesh.create(SETTINGS)
    m2.load_tetrahedra(p)
    m2.finalize()
    b = m2.arrays()
    assert np.array_equal(a["tets"], b["tets"]) and np.allclose(a["x"], b["x"], rtol=0, atol=1e-5)
    x = a["x"] + 1e-3 * np.sin(np.arange(a["x"].size)).reshape(-1, 3)
    m.set_state(x=x, xtilde=x * 0.5)
    m.save_checkpoint(str(tmp_path))
    m3 = host.HostMesh.scene("twisting-mat", ASSETS, SETTINGS)
    m3.load_checkpoint(str(tmp_path))
    d, _ = m3.desc()
    V = d.V
    assert np.array_equal(np.ctypeslib.as_array(d.x, shape=(V, 3)), x)
    assert np.array_equal(np.ctypeslib.as_array(d.xtilde, shape=(V, 3)), x * 0.5)
    small = host.HostMesh.create(SETTINGS)
    small.load_tetrahedra(str(tmp_path / "out.msh"))
    small.finalize()
    with open(tmp_path / "vertex.txt", "a") as f:
        f.write("0 0 0\n")
    with pytest.raises(RuntimeError, match="checkpoint"):
        m3.load_checkpoint(str(tmp_path))


def test_headless_arguments():
    exe = os.path.join(ROOT, "cpu-ipc_amd", "cipc_headless")
    r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 0 and "--scene" in r.stdout
    r = subprocess.run([exe, "--broad-phase", "spatial-hash", "--assets", ASSETS], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "broad phase" in r.stderr
    r = subprocess.run([exe, "--scene", "nope", "--assets", ASSETS], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "unknown scene" in r.stderr
