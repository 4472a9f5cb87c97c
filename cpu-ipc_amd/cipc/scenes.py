"""Scene generators and per-scene host set-up for the Newton-iteration path.

The reference builds its scenes from Gmsh/OBJ assets (``CPU IPC/Simulator.cpp:265-399``)
and prepares them with ``initMesh3D`` (``CPU IPC/Elasticity.cpp:1050-1116``),
``mesh3D::getSurface`` (``CPU IPC/SimulationMesh.cpp:364-446``) and the material
conversion ``updateMaterial`` (``CPU IPC/Simulator.cpp:128-141``).  The lattice-cube
workloads measured here do not exist in the reference, so they are generated here,
seeded, and then prepared with the reference's own rules:

* masses: ``vol * density / 4`` per tet corner (``Elasticity.cpp:1066-1071``);
* ``Dm^-1`` of the rest shape, column-major 3x3 (``Elasticity.cpp:1062,1075``);
* surface triangles oriented so the opposite tet vertex lies behind them
  (``SimulationMesh.cpp:396-420``), ``surfVerts`` in first-appearance order
  (``SimulationMesh.cpp:424-433``) and ``surfEdges`` as the ordered set of directed
  edges (``SimulationMesh.cpp:435-445``);
* ``bboxDiagSize2``, ``Hhat = dHat_rel^2 * bboxDiagSize2`` and
  ``averageEdgeLenth`` (``Simulator.cpp:486-506``).

Everything here is host-side, once per scene; it is the data contract that the
device path consumes, not part of the timed path.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

# Unit-cell 5-tet split.  Corners are indexed c = dx + 2*dy + 4*dz.  A cell of
# parity 0 keeps the regular tet on corners {0,3,5,6}; parity 1 on {1,2,4,7}; the
# alternation makes the face diagonals of neighbouring cells agree, so the lattice is
# conforming (the reference's unit split is the same construction,
# SimulationMesh.cpp:15-26).
_SPLIT_EVEN = ((0, 1, 3, 5), (0, 3, 2, 6), (0, 5, 6, 4), (3, 5, 7, 6), (0, 3, 6, 5))
_SPLIT_ODD = ((1, 0, 2, 4), (1, 2, 3, 7), (1, 4, 7, 5), (2, 4, 6, 7), (1, 2, 4, 7))


@dataclasses.dataclass
class Material:
    """Material and IPC scalars with the reference's conversions (Simulator.cpp:128-141)."""

    youngs: float = 1e5
    poisson: float = 0.4
    density: float = 1000.0
    dt: float = 0.01
    rel_dhat: float = 1e-3
    friction: float = 0.2
    drag: float = 0.0
    newton_threshold: float = 1e-2
    fhat_rel: float = 1e-6
    dtol_rel: float = 1e-18

    @property
    def mu_lame(self) -> float:
        return self.youngs / (2 * (1 + self.poisson))

    @property
    def lambda_lame(self) -> float:
        return self.youngs * self.poisson / ((1 + self.poisson) * (1 - 2 * self.poisson))

    @property
    def length_rate(self) -> float:
        return 4 * self.mu_lame / 3

    @property
    def volume_rate(self) -> float:
        return self.lambda_lame + 5 * self.mu_lame / 6


@dataclasses.dataclass
class Scene:
    """Host arrays in the reference's layouts (``mesh3D``, SimulationMesh.h:37-129).

    Positions are AoS float64 ``(V, 3)``; ``dm_inv`` holds each tet's 3x3 ``Dm^-1`` in
    column-major order (Eigen's storage), so ``dm_inv.reshape(T, 3, 3)[t]`` is its
    transpose.
    """

    name: str
    x: np.ndarray            # (V,3) current positions
    v_rest: np.ndarray       # (V,3) rest positions
    V_prev: np.ndarray       # (V,3) positions at the start of the step
    velocities: np.ndarray   # (V,3)
    tets: np.ndarray         # (T,4) int32
    dm_inv: np.ndarray       # (T,9) column-major Dm^-1
    volumes: np.ndarray      # (T,)
    masses: np.ndarray       # (V,)
    boundary: np.ndarray     # (V,) int32, 0 free / 1 Dirichlet / >=2 collider
    faces: np.ndarray        # (F,4) int32 (v0, v1, v2, tetId)
    surf_verts: np.ndarray   # (Vs,) int32
    surf_edges: np.ndarray   # (Es,2) int32
    material: Material
    bbox_diag2: float
    hhat: float
    fhat: float
    dtol: float
    mean_mass: float
    avg_edge_len: float
    kappa: float = 0.0
    apply_gravity: bool = True

    @property
    def V(self) -> int:
        return int(self.x.shape[0])

    @property
    def T(self) -> int:
        return int(self.tets.shape[0])

    def inertial_target(self) -> np.ndarray:
        """x~ = V_prev + v dt + g dt^2 with g = (0,-9.8,0) (IPCSolver.cpp:55-68)."""
        g = -9.8 if self.apply_gravity else 0.0
        dt = self.material.dt
        return self.V_prev + self.velocities * dt + np.array([0.0, g * dt * dt, 0.0])


def lattice(n: int, cell: float):
    """Vertices and 5-tet split of an n^3-cell lattice with cell edge ``cell``.

    V=(n+1)^3, T=5n^3; vertex (i,j,k) has index i + (n+1)(j + (n+1)k).
    """
    m = n + 1
    idx = np.arange(m ** 3)
    i, j, k = idx % m, (idx // m) % m, idx // (m * m)
    verts = np.stack([i, j, k], axis=1).astype(np.float64) * cell
    tets = []
    for ck in range(n):
        for cj in range(n):
            for ci in range(n):
                corner = [(ci + (c & 1)) + m * ((cj + ((c >> 1) & 1)) + m * (ck + ((c >> 2) & 1)))
                          for c in range(8)]
                split = _SPLIT_EVEN if (ci + cj + ck) % 2 == 0 else _SPLIT_ODD
                for t in split:
                    tets.append([corner[a] for a in t])
    tets = np.asarray(tets, dtype=np.int64)
    # orient every tet positively so that det(Dm) > 0
    d = verts[tets[:, 1:]] - verts[tets[:, :1]]
    neg = np.linalg.det(d) < 0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    return verts, tets


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform random rotation from a unit quaternion (Shoemake's method)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
                  math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3),
                  math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def extract_surface(x: np.ndarray, tets: np.ndarray):
    """Boundary faces, surface vertices and surface edges of a tet mesh.

    Follows the reference's rules (SimulationMesh.cpp:364-446): a face seen once is on
    the boundary; it is stored (v0,v1,v2,tet) with the fourth tet vertex on the negative
    side of (v1-v0)x(v2-v0); surfVerts keeps first-appearance order over the face
    list; surfEdges is the lexicographically ordered set of directed edges where a
    directed edge is skipped if its reverse is already present.  The face order itself
    is the ascending order of sorted face keys (the reference's order comes from an
    unordered_map and is not reproducible; the device path consumes these arrays as
    given).
    """
    T = tets.shape[0]
    local = np.array([[0, 1, 2], [1, 2, 3], [2, 3, 0], [3, 0, 1]])
    opp = np.array([3, 0, 1, 2])
    tri = tets[:, local].reshape(-1, 3)
    opp_v = tets[:, opp].reshape(-1)
    tet_id = np.repeat(np.arange(T), 4)
    key = np.sort(tri, axis=1)
    order = np.lexsort((key[:, 2], key[:, 1], key[:, 0]))
    key_s = key[order]
    same_prev = np.zeros(len(order), bool)
    same_prev[1:] = np.all(key_s[1:] == key_s[:-1], axis=1)
    same_next = np.zeros(len(order), bool)
    same_next[:-1] = same_prev[1:]
    boundary = order[~(same_prev | same_next)]
    faces = []
    for f in boundary:
        v0, v1, v2 = tri[f]
        n = np.cross(x[v1] - x[v0], x[v2] - x[v0])
        if np.dot(n, x[opp_v[f]] - x[v0]) < 0:
            faces.append((v0, v1, v2, tet_id[f]))
        else:
            faces.append((v0, v2, v1, tet_id[f]))
    faces = np.asarray(faces, dtype=np.int64).reshape(-1, 4)
    seen = np.zeros(x.shape[0], bool)
    surf_verts = []
    for f in faces:
        for v in f[:3]:
            if not seen[v]:
                seen[v] = True
                surf_verts.append(v)
    edges = set()
    for f in faces:
        for a, b in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            if (b, a) not in edges:
                edges.add((a, b))
    surf_edges = np.asarray(sorted(edges), dtype=np.int64).reshape(-1, 2)
    return faces, np.asarray(surf_verts, dtype=np.int64), surf_edges


def dm_inverse_colmajor(x: np.ndarray, tets: np.ndarray):
    """Rest-shape Dm^-1 (column-major) and volumes (Elasticity.cpp:570-598,1060-1078)."""
    dm = np.stack([x[tets[:, 1]] - x[tets[:, 0]],
                   x[tets[:, 2]] - x[tets[:, 0]],
                   x[tets[:, 3]] - x[tets[:, 0]]], axis=2)       # columns are edges
    inv = np.linalg.inv(dm)
    vol = np.abs(np.linalg.det(dm)) / 6.0
    # column-major storage of inv: element (r,c) at index r + 3c
    return inv.transpose(0, 2, 1).reshape(-1, 9).copy(), vol


def finalize(name: str, x: np.ndarray, tets: np.ndarray, mat: Material,
             boundary: Optional[np.ndarray] = None,
             velocities: Optional[np.ndarray] = None) -> Scene:
    """initMesh3D + getSurface + the scalar set-up of FEMSimulator::buildModels."""
    tets = np.ascontiguousarray(tets, dtype=np.int64)
    V = x.shape[0]
    dm_inv, vol = dm_inverse_colmajor(x, tets)
    masses = np.zeros(V)
    np.add.at(masses, tets.reshape(-1), np.repeat(vol * mat.density / 4.0, 4))
    faces, surf_verts, surf_edges = extract_surface(x, tets)
    lo, hi = x.min(axis=0), x.max(axis=0)
    diag2 = float(np.sum((hi - lo) ** 2))
    elen = np.linalg.norm(x[surf_edges[:, 0]] - x[surf_edges[:, 1]], axis=1).sum()
    if boundary is None:
        boundary = np.zeros(V, np.int64)
    if velocities is None:
        velocities = np.zeros_like(x)
    return Scene(name=name, x=x.copy(), v_rest=x.copy(), V_prev=x.copy(),
                 velocities=np.ascontiguousarray(velocities, dtype=np.float64),
                 tets=tets.astype(np.int32), dm_inv=dm_inv, volumes=vol, masses=masses,
                 boundary=np.asarray(boundary, np.int32),
                 faces=faces.astype(np.int32), surf_verts=surf_verts.astype(np.int32),
                 surf_edges=surf_edges.astype(np.int32), material=mat,
                 bbox_diag2=diag2, hhat=mat.rel_dhat ** 2 * diag2,
                 fhat=mat.fhat_rel * diag2, dtol=mat.dtol_rel * diag2,
                 mean_mass=float(masses.sum() / V), avg_edge_len=float(elen / (len(surf_edges) * 3)))


GROUND_Y = -1.0  # Ground plane y = -1, normal (0,1,0) (CollisionBroadPhase.cpp:1067-1069)


def yaw_rotation(rng: np.random.Generator) -> np.ndarray:
    """Rotation by a uniform random angle about the stacking (y) axis."""
    a = rng.uniform(0.0, 2.0 * math.pi)
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def stacked_cubes(n_cubes: int, cells: int, cell: float, gap: float, seed: int,
                  jitter: float = 0.01, rotate: str = "yaw", mat: Optional[Material] = None,
                  velocity_scale: float = 0.0, name: str = "stacked_cubes") -> Scene:
    """``n_cubes`` lattice cubes of ``cells``^3 with seeded rotations, stacked along +y
    with ``gap`` between consecutive axis-aligned bounds, the lowest ``gap`` above the
    ground.  ``rotate`` is "yaw" (random angle about the stacking axis: faces stay
    parallel, so every interface is a face-to-face contact closer than dHat), "full"
    (uniform random 3D rotation: cubes meet at corners) or "none".  ``jitter`` is the
    relative uniform vertex perturbation (of the cell edge); ``velocity_scale`` gives
    each cube a seeded uniform velocity in [-s, s]^3 m/s (zero for a drop from rest)."""
    mat = mat or Material()
    rng = np.random.default_rng(seed)
    base_x, base_t = lattice(cells, cell)
    xs, ts, vs = [], [], []
    y_floor = GROUND_Y + gap
    offset = 0
    for c in range(n_cubes):
        xc = base_x + rng.uniform(-jitter, jitter, base_x.shape) * cell
        center = 0.5 * cells * cell
        xc = xc - center
        if rotate == "full":
            xc = xc @ random_rotation(rng).T
        elif rotate == "yaw":
            xc = xc @ yaw_rotation(rng).T
        lo = xc.min(axis=0)
        xc = xc + np.array([-0.5 * (xc.max(axis=0)[0] + lo[0]), y_floor - lo[1],
                            -0.5 * (xc.max(axis=0)[2] + lo[2])])
        y_floor = xc[:, 1].max() + gap
        vel = rng.uniform(-velocity_scale, velocity_scale, 3) if velocity_scale > 0 else np.zeros(3)
        xs.append(xc)
        ts.append(base_t + offset)
        vs.append(np.tile(vel, (xc.shape[0], 1)))
        offset += xc.shape[0]
    x = np.concatenate(xs)
    return finalize(name, x, np.concatenate(ts), mat, velocities=np.concatenate(vs))


def cube10(seed: int = 0) -> Scene:
    """One 10^3-cell cube (1331 verts / 5000 tets), edge 0.1 m, 0.05 m above the ground,
    +-1% jitter, E=1e5, nu=0.4, rho=1000, dt=0.01, dHat_rel=1e-3, mu=0.2."""
    mat = Material(friction=0.2)
    s = stacked_cubes(1, 10, 0.1, 0.05, seed, rotate="none", mat=mat, name="cube10")
    return s


def four_cubes20(seed: int = 1, velocity_scale: float = 0.0, rotate: str = "yaw") -> Scene:
    """Four 20^3-cell cubes (9261 verts / 40000 tets each), cell 0.1 m, seeded random
    rotations about the stacking axis, stacked with 1e-3 m gaps on the ground, mu=0.3,
    otherwise as cube10.  The gap is below dHat (1e-3 of the scene diagonal, ~8.9e-3 m),
    so every cube-cube and cube-ground interface is in contact from the first
    iteration."""
    mat = Material(friction=0.3)
    return stacked_cubes(4, 20, 0.1, 1e-3, seed, rotate=rotate, mat=mat,
                         velocity_scale=velocity_scale, name="4x20cubes")


def suggest_kappa(hhat: float, bbox_diag2: float, mean_mass: float) -> float:
    """suggestKappa (IPCSolver.cpp:751-755)."""
    d = 1.0e-16 * bbox_diag2
    t = d - hhat
    h_b = (math.log(d / hhat) * -2.0 - t * 4.0 / d) + 1.0 / (d * d) * (t * t)
    return 1e13 * mean_mass / (4.0e-16 * bbox_diag2 * h_b)


def newton_probe_state(scene: Scene, seed: int = 7, scale: float = 1.0):
    """A frozen Newton-iteration state for timing and parity: positions as stored,
    x~ from the scene's velocities and gravity, and the direction p = x - x~ (the
    inertia-dominated Newton direction for x_new = x - alpha p), perturbed per vertex
    by a seeded +-5% so that contacting primitives move relative to each other.
    Constrained vertices get p = 0."""
    rng = np.random.default_rng(seed)
    xt = scene.inertial_target()
    p = (scene.x - xt) * scale
    p = p * (1.0 + rng.uniform(-0.05, 0.05, p.shape))
    p[scene.boundary != 0] = 0.0
    return xt, np.ascontiguousarray(p)
