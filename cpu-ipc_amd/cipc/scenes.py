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
class ClothMaterial:
    """Cloth knobs and their conversions (updateMaterial, Simulator.cpp:128-141;
    defaults of Assets/scene/parameterSetting.txt).  ``plate_rigidity`` overrides the
    conversion E_b t^3 / (12 (1 - nu^2)) when given."""

    youngs: float = 1e4
    shear_youngs: float = 1e1
    bend_youngs: float = 1e4
    poisson: float = 0.4
    thickness: float = 1e-3
    density: float = 200.0
    strain_rate: float = 1e2
    plate_rigidity_override: Optional[float] = None

    @property
    def stretch(self) -> float:
        return self.youngs / (1 - self.poisson * self.poisson)

    @property
    def shear(self) -> float:
        return self.shear_youngs / (2 * (1 + self.poisson)) / self.thickness

    @property
    def plate_rigidity(self) -> float:
        if self.plate_rigidity_override is not None:
            return self.plate_rigidity_override
        return self.bend_youngs * self.thickness ** 3 / (12 * (1 - self.poisson * self.poisson))


@dataclasses.dataclass
class Cloth:
    """Triangle (cloth) elements and bending elements of a scene (mesh3D::triangles,
    DM_triangle_inverse, areas, quadBendingInfo / hingeBendingInfo; SimulationMesh.h:
    17-35,79-113).  ``bending`` is 0 for the quadratic form (USE_QUADRATIC_BENDING) and 1
    for the hinge-angle form; ``bend_q`` holds Q (row-major 4x4) of the quadratic form,
    ``bend_rest`` / ``bend_weight`` the rest angle and l^2/(2A0+2A1) of the hinge form."""

    triangles: np.ndarray    # (Tr,3) int32
    dm_tri_inv: np.ndarray   # (Tr,4) column-major 2x2 Dm^-1
    areas: np.ndarray        # (Tr,) area x thickness
    material: ClothMaterial
    bending: int
    bend_verts: np.ndarray   # (Nb,4) int32
    bend_q: np.ndarray       # (Nb,16)
    bend_rest: np.ndarray    # (Nb,)
    bend_weight: np.ndarray  # (Nb,)

    @property
    def Nb(self) -> int:
        return int(self.bend_verts.shape[0])


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
    cloth: Optional["Cloth"] = None

    @property
    def Tr(self) -> int:
        return 0 if self.cloth is None else int(self.cloth.triangles.shape[0])

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


def extract_surface(x: np.ndarray, tets: np.ndarray, triangles: Optional[np.ndarray] = None):
    """Boundary faces, surface vertices and surface edges of a tet mesh.

    Follows the reference's rules (SimulationMesh.cpp:364-446): a face seen once is on
    the boundary; it is stored (v0,v1,v2,tet) with the fourth tet vertex on the negative
    side of (v1-v0)x(v2-v0); surfVerts keeps first-appearance order over the face
    list; surfEdges is the lexicographically ordered set of directed edges where a
    directed edge is skipped if its reverse is already present.  The face order itself
    is the ascending order of sorted face keys (the reference's order comes from an
    unordered_map and is not reproducible; the device path consumes these arrays as
    given).  Cloth triangles follow the tet faces as (v0, v1, v2, 0)
    (SimulationMesh.cpp:418-420).
    """
    T = tets.shape[0]
    if T == 0:
        tets = np.zeros((0, 4), np.int64)
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
    if triangles is not None and len(triangles):
        faces.extend((int(a), int(b), int(c), 0) for a, b, c in triangles)
    faces = np.asarray(faces, dtype=np.int64).reshape(-1, 4)
    # surfVerts: first appearance over the face list
    fv = faces[:, :3].reshape(-1)
    uv, first = np.unique(fv, return_index=True)
    surf_verts = uv[np.argsort(first, kind="stable")]
    # surfEdges: an undirected edge keeps the direction of its first appearance (a later
    # reversed copy finds its reverse in the set; a later equal copy is a no-op insert)
    a = faces[:, [0, 1, 2]].reshape(-1)
    b = faces[:, [1, 2, 0]].reshape(-1)
    key = np.stack([np.minimum(a, b), np.maximum(a, b)], axis=1)
    _, fi = np.unique(key, axis=0, return_index=True)
    de = np.stack([a[fi], b[fi]], axis=1)
    surf_edges = de[np.lexsort((de[:, 1], de[:, 0]))].reshape(-1, 2)
    return faces, surf_verts.astype(np.int64), surf_edges.astype(np.int64)


def triangle_rest(x: np.ndarray, tris: np.ndarray, thickness: float):
    """Per-triangle Dm^-1 (column-major 2x2) and area x thickness (initMesh3D triangle
    loop, Elasticity.cpp:1085-1113; calculateDms2D_double, Elasticity.cpp:610-640): the
    triangle is rotated so that its normal maps to +y (Rodrigues' formula
    I + [v]x + [v]x^2 / (1 + cos)) and its (x, z) coordinates are the 2D rest shape."""
    v01 = x[tris[:, 1]] - x[tris[:, 0]]
    v02 = x[tris[:, 2]] - x[tris[:, 0]]
    cr = np.cross(v01, v02)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    nrm = cr / np.linalg.norm(cr, axis=1, keepdims=True)
    target = np.array([0.0, 1.0, 0.0])
    vec = np.cross(nrm, target)
    cos = nrm @ target
    K = np.zeros((len(tris), 3, 3))
    K[:, 0, 1], K[:, 0, 2] = -vec[:, 2], vec[:, 1]
    K[:, 1, 0], K[:, 1, 2] = vec[:, 2], -vec[:, 0]
    K[:, 2, 0], K[:, 2, 1] = -vec[:, 1], vec[:, 0]
    R = np.eye(3)[None] + K + (K @ K) / (1 + cos)[:, None, None]
    uv = [np.einsum("tij,tj->ti", R, x[tris[:, k]])[:, [0, 2]] for k in range(3)]
    dm = np.stack([uv[1] - uv[0], uv[2] - uv[0]], axis=2)       # columns are edges
    inv = np.linalg.inv(dm)
    return inv.transpose(0, 2, 1).reshape(-1, 4).copy(), area * thickness


def prepare_quadratic_bending(x: np.ndarray, tris: np.ndarray):
    """prepareQuadraticBending (Simulator.cpp:48-109): one element per later appearance of
    an edge over the triangle list (edge i of a triangle runs from corner i to corner
    i+1), in the order of those appearances, with vertices (e0, e1, outer vertex of the
    edge's first triangle, outer vertex of this one) and the constant
    Q = 3 / (A0 + A1) K K^T from the cotangent weights."""
    tris = np.asarray(tris, np.int64).reshape(-1, 3)
    a = tris.reshape(-1)
    b = tris[:, [1, 2, 0]].reshape(-1)
    other = tris[:, [2, 0, 1]].reshape(-1)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    order = np.lexsort((np.arange(len(a)), hi, lo))           # by edge, then appearance
    lo_s, hi_s = lo[order], hi[order]
    new = np.ones(len(order), bool)
    new[1:] = (lo_s[1:] != lo_s[:-1]) | (hi_s[1:] != hi_s[:-1])
    first = np.maximum.accumulate(np.where(new, np.arange(len(order)), 0))
    later = np.nonzero(~new)[0]
    pos = order[later]                                        # appearance index of the later occurrence
    o1 = other[order[first[later]]]
    o2 = other[pos]
    e0, e1 = lo[pos], hi[pos]
    srt = np.argsort(pos, kind="stable")
    e0, e1, o1, o2 = e0[srt], e1[srt], o1[srt], o2[srt]
    v0, v1, v2, v3 = x[o1], x[e0], x[e1], x[o2]
    E0, E1, E2, E3, E4 = v2 - v1, v0 - v1, v3 - v1, v0 - v2, v3 - v2
    nrm = lambda u: np.linalg.norm(u, axis=1)
    a0 = 0.5 * nrm(np.cross(E0, E1))
    a1 = 0.5 * nrm(np.cross(E0, E2))
    cot = lambda u, v: np.einsum("ij,ij->i", u, v) / nrm(np.cross(u, v))
    c01, c02, c03, c04 = cot(E0, E1), cot(E0, E2), cot(-E0, E3), cot(-E0, E4)
    K = np.stack([c03 + c04, c01 + c02, -c01 - c03, -c02 - c04], axis=1)
    Q = (3.0 / (a0 + a1))[:, None, None] * (K[:, :, None] * K[:, None, :])
    verts = np.stack([e0, e1, o1, o2], axis=1).astype(np.int32)
    return verts.reshape(-1, 4), Q.reshape(-1, 16)


def hinge_angle(q0, q1, q2, q3) -> float:
    """edgeTheta's angle (Elasticity.cpp:12-40,69-85): signed dihedral angle of the
    triangles (q0, q1, q2) and (q1, q0, q3) about the axis q1 - q0."""
    n0 = np.cross(q0 - q2, q1 - q2)
    n1 = np.cross(q1 - q3, q0 - q3)
    axis = q1 - q0
    return 2.0 * math.atan2(np.dot(np.cross(n0, n1), axis) / np.linalg.norm(axis),
                            np.dot(n0, n1) + np.linalg.norm(n0) * np.linalg.norm(n1))


def prepare_hinge_bending(x: np.ndarray, tris: np.ndarray):
    """The hinge list of load_triangleMesh (SimulationMesh.cpp:659-713: interior edges in
    ascending (min, max) order, opposite vertices in triangle order, the three edges of a
    triangle (a,b,c) taken as (a,b), (b,c), (a,c)) and makeHingeBendingInfo
    (HingeBending.cpp:47-77: rest angle and l^2 / (2A0 + 2A1))."""
    tris = np.asarray(tris, np.int64).reshape(-1, 3)
    p = tris[:, [0, 1, 0]].reshape(-1)
    q = tris[:, [1, 2, 2]].reshape(-1)
    o = tris[:, [2, 0, 1]].reshape(-1)
    lo, hi = np.minimum(p, q), np.maximum(p, q)
    order = np.lexsort((np.arange(len(p)), hi, lo))
    lo_s, hi_s, o_s = lo[order], hi[order], o[order]
    start = np.ones(len(order), bool)
    start[1:] = (lo_s[1:] != lo_s[:-1]) | (hi_s[1:] != hi_s[:-1])
    idx = np.nonzero(start)[0]
    cnt = np.diff(np.append(idx, len(order)))
    two = idx[cnt == 2]
    verts = np.stack([lo_s[two], hi_s[two], o_s[two], o_s[two + 1]], axis=1)
    q0, q1, q2, q3 = (x[verts[:, k]] for k in range(4))
    n0 = np.cross(q0 - q2, q1 - q2)
    n1 = np.cross(q1 - q3, q0 - q3)
    axis = q1 - q0
    ln = np.linalg.norm(axis, axis=1)
    rest = 2.0 * np.arctan2(np.einsum("ij,ij->i", np.cross(n0, n1), axis) / ln,
                            np.einsum("ij,ij->i", n0, n1) + np.linalg.norm(n0, axis=1) * np.linalg.norm(n1, axis=1))
    s2 = np.linalg.norm(np.cross(axis, q2 - q0), axis=1) + np.linalg.norm(np.cross(axis, q3 - q0), axis=1)
    return verts.astype(np.int32), rest, ln * ln / s2


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
             velocities: Optional[np.ndarray] = None,
             triangles: Optional[np.ndarray] = None,
             cloth_mat: Optional[ClothMaterial] = None,
             bending: int = 0) -> Scene:
    """initMesh3D + getSurface + the scalar set-up of FEMSimulator::buildModels; with
    ``triangles`` also the cloth part of initMesh3D (masses area x thickness x density / 3
    per corner) and the bending preparation (quadratic for ``bending`` 0, hinge for 1)."""
    tets = np.ascontiguousarray(tets, dtype=np.int64).reshape(-1, 4)
    V = x.shape[0]
    if len(tets):
        dm_inv, vol = dm_inverse_colmajor(x, tets)
    else:
        dm_inv, vol = np.zeros((0, 9)), np.zeros(0)
    masses = np.zeros(V)
    np.add.at(masses, tets.reshape(-1), np.repeat(vol * mat.density / 4.0, 4))
    cloth = None
    if triangles is not None and len(triangles):
        triangles = np.ascontiguousarray(triangles, dtype=np.int64)
        cm = cloth_mat or ClothMaterial()
        dmt, areas = triangle_rest(x, triangles, cm.thickness)
        np.add.at(masses, triangles.reshape(-1), np.repeat(areas * cm.density / 3.0, 3))
        if bending == 0:
            bv, bq = prepare_quadratic_bending(x, triangles)
            br, bw = np.zeros(len(bv)), np.zeros(len(bv))
        else:
            bv, br, bw = prepare_hinge_bending(x, triangles)
            bq = np.zeros((len(bv), 16))
        cloth = Cloth(triangles=triangles.astype(np.int32), dm_tri_inv=dmt, areas=areas, material=cm,
                      bending=bending, bend_verts=bv, bend_q=bq, bend_rest=br, bend_weight=bw)
    faces, surf_verts, surf_edges = extract_surface(x, tets, triangles)
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
                 mean_mass=float(masses.sum() / V), avg_edge_len=float(elen / (len(surf_edges) * 3)),
                 cloth=cloth)


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


def batched_cubes(n_scenes: int = 256, seed: int = 0, cells_lo: int = 8, cells_hi: int = 24, edge: float = 1.0,
                  spacing: float = 3.0, jitter: float = 0.01, landed: bool = False) -> list:
    """Independent single-cube scenes for a batch: each a lattice cube of edge ``edge`` with
    a cell count per axis log-uniform in [cells_lo, cells_hi] (8^3..24^3 cells: 2,560 to
    69,120 tets), seeded Young's modulus log-uniform in [1e4, 1e6] Pa, Poisson ratio
    uniform in [0.3, 0.45], dropped from rest with its lowest point 0.02-0.2 m above the
    ground; mu = 0.2, +-1% vertex jitter.  dHat is relative to the nominal cube diagonal
    (sqrt(3) edge) in every scene, so the scenes share hhat, fhat and dtol as a batch
    requires.  Scene k sits at a grid position in the ground plane, ``spacing`` apart
    (the scenes never interact in any case).

    ``landed``: each cube instead at the frame where its drop reaches the ground: the lowest
    point a seeded 0.2-0.9 dHat above it (so part of the bottom face is in contact), moving
    down at the speed the drop gives, sqrt(2 g h).  The other draws are unchanged."""
    rng = np.random.default_rng(seed)
    land = np.random.default_rng(seed + 7919)
    cache = {}
    cols = int(math.ceil(math.sqrt(n_scenes)))
    out = []
    diag2 = 3.0 * edge * edge
    for k in range(n_scenes):
        n = int(round(math.exp(rng.uniform(math.log(cells_lo), math.log(cells_hi)))))
        n = min(max(n, cells_lo), cells_hi)
        if n not in cache:
            cache[n] = lattice(n, edge / n)
        base_x, base_t = cache[n]
        E = math.exp(rng.uniform(math.log(1e4), math.log(1e6)))
        nu = rng.uniform(0.3, 0.45)
        drop = rng.uniform(0.02, 0.2)
        mat = Material(youngs=E, poisson=nu, friction=0.2)
        x = base_x + rng.uniform(-jitter, jitter, base_x.shape) * (edge / n)
        lo = x.min(axis=0)
        if landed:
            height = land.uniform(0.2, 0.9) * mat.rel_dhat * math.sqrt(diag2)
            x = x + np.array([(k % cols) * spacing - lo[0], GROUND_Y + height - lo[1], (k // cols) * spacing - lo[2]])
            vel = np.tile([0.0, -math.sqrt(2 * 9.8 * drop), 0.0], (x.shape[0], 1))
            sc = finalize(f"landed{k}_{n}", x, base_t, mat, velocities=vel)
        else:
            x = x + np.array([(k % cols) * spacing - lo[0], GROUND_Y + drop - lo[1], (k // cols) * spacing - lo[2]])
            sc = finalize(f"batch{k}_{n}", x, base_t, mat)
        sc.hhat = mat.rel_dhat ** 2 * diag2
        sc.fhat = mat.fhat_rel * diag2
        sc.dtol = mat.dtol_rel * diag2
        out.append(sc)
    return out


def suggest_kappa(hhat: float, bbox_diag2: float, mean_mass: float) -> float:
    """suggestKappa (IPCSolver.cpp:751-755)."""
    from cipc.kappa import suggest_kappa as sk
    return sk(hhat, bbox_diag2, mean_mass)


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


CLOTH_FIELDS = (("Tr", "int"), ("triangles", "int*"), ("dm_tri_inv", "double*"), ("areas", "double*"),
                ("stretch", "double"), ("shear", "double"), ("strain_rate", "double"), ("plate_rigidity", "double"),
                ("Nb", "int"), ("bending", "int"), ("bend_verts", "int*"), ("bend_q", "double*"),
                ("bend_rest", "double*"), ("bend_weight", "double*"))


def cloth_arrays(scene: Scene) -> dict:
    """The cloth part of the scene descriptor (names of CLOTH_FIELDS), arrays as
    contiguous int32 / float64 (never empty, so that every pointer is valid)."""
    c = scene.cloth
    i32 = lambda a: np.ascontiguousarray(np.asarray(a).reshape(-1) if np.size(a) else np.zeros(1), dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(np.asarray(a).reshape(-1) if np.size(a) else np.zeros(1), dtype=np.float64)
    if c is None:
        z = np.zeros(1)
        return dict(Tr=0, triangles=i32(z), dm_tri_inv=f64(z), areas=f64(z), stretch=0.0, shear=0.0,
                    strain_rate=0.0, plate_rigidity=0.0, Nb=0, bending=0, bend_verts=i32(z), bend_q=f64(z),
                    bend_rest=f64(z), bend_weight=f64(z))
    m = c.material
    return dict(Tr=int(len(c.triangles)), triangles=i32(c.triangles), dm_tri_inv=f64(c.dm_tri_inv),
                areas=f64(c.areas), stretch=m.stretch, shear=m.shear, strain_rate=m.strain_rate,
                plate_rigidity=m.plate_rigidity, Nb=c.Nb, bending=int(c.bending), bend_verts=i32(c.bend_verts),
                bend_q=f64(c.bend_q), bend_rest=f64(c.bend_rest), bend_weight=f64(c.bend_weight))


def cloth_grid(n: int, size: float):
    """(n+1)^2 vertices of a ``size`` square in the xz-plane centred on the origin, and
    2 n^2 triangles, each oriented with its normal along +y (the orientation
    calculateDms2D_double's rotation to +y is defined for)."""
    m = n + 1
    i, k = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    x = np.stack([i.reshape(-1), np.zeros(m * m), k.reshape(-1)], axis=1).astype(np.float64)
    x = (x / n - np.array([0.5, 0.0, 0.5])) * np.array([size, 0.0, size])
    vid = lambda a, b: a * m + b          # vertex (x index a, z index b)
    tris = []
    for a in range(n):
        for b in range(n):
            p00, p01, p10, p11 = vid(a, b), vid(a, b + 1), vid(a + 1, b), vid(a + 1, b + 1)
            if (a + b) % 2 == 0:
                tris += [(p00, p01, p11), (p00, p11, p10)]
            else:
                tris += [(p00, p01, p10), (p01, p11, p10)]
    tris = np.asarray(tris, np.int64)
    nrm = np.cross(x[tris[:, 1]] - x[tris[:, 0]], x[tris[:, 2]] - x[tris[:, 0]])
    assert np.all(nrm[:, 1] > 0)
    return x, tris


def lattice_ball(cells: int, radius: float):
    """A ball of ``cells``^3 lattice cells (5 tets each) obtained by mapping the cube
    [-1, 1]^3 radially onto the ball (x -> x |x|_inf / |x|_2 * radius)."""
    x, tets = lattice(cells, 2.0 / cells)
    x = x - 1.0
    n2 = np.linalg.norm(x, axis=1)
    ninf = np.abs(x).max(axis=1)
    scale = np.where(n2 > 0, ninf / np.where(n2 > 0, n2, 1.0), 0.0)
    x = x * scale[:, None] * radius
    d = x[tets[:, 1:]] - x[tets[:, :1]]
    neg = np.linalg.det(d) < 0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    return x, tets


def cloth_on_sphere(n: int = 256, size: float = 1.0, sphere_cells: int = 10, radius: float = 0.3,
                    bending: int = 0, fold: Optional[int] = None, seed: int = 0,
                    layer_gap: float = 0.6e-3, name: Optional[str] = None) -> Scene:
    """A ``size`` square cloth of ``n`` x ``n`` quads (2 n^2 triangles) above a fixed
    ball collider of ``sphere_cells``^3 cells (5 sphere_cells^3 tets, boundary type 2)
    on the ground.  Cloth: E = 1e4, thickness 1e-3, density 200, plate rigidity 1e-4;
    contact: dHat_rel 1e-3, mu 0.2.

    The rest shape is the flat square.  The returned state (x) lies ``layer_gap`` above
    the ball's top; with ``fold`` = r the cloth rows past z-row r are folded back over
    the rest, ``layer_gap`` above it, shifted by half a cell in x, so that the two layers
    are in contact everywhere they overlap (a frozen, self-contact-heavy state of a
    cloth dropped on the ball).  In-plane positions carry a seeded +-15% (of a cell)
    jitter, heights a +-5% (of ``layer_gap``) one: exactly planar layers make many
    Hessian entries mathematically zero, and whether such an entry comes out as an exact
    zero (and is dropped from the CSR) then depends on rounding order."""
    rng = np.random.default_rng(seed)
    mat = Material(youngs=1e5, friction=0.2)
    cm = ClothMaterial(youngs=1e4, thickness=1e-3, plate_rigidity_override=1e-4)
    xc, tris = cloth_grid(n, size)
    xb, tb = lattice_ball(sphere_cells, radius)
    centre_y = GROUND_Y + radius + 0.2
    xb = xb + np.array([0.0, centre_y, 0.0])
    top = xb[:, 1].max()
    Vb = xb.shape[0]
    x_rest = np.concatenate([xb, xc + np.array([0.0, top + layer_gap, 0.0])])
    tris = tris + Vb
    boundary = np.zeros(x_rest.shape[0], np.int32)
    boundary[:Vb] = 2
    s = finalize(name or f"cloth{n}_sphere", x_rest, tb, mat, boundary=boundary, triangles=tris,
                 cloth_mat=cm, bending=bending)
    # the state: in-plane jitter, then the fold
    cell = size / n
    m = n + 1
    x = s.x.copy()
    ci = np.arange(m * m)
    a, b = ci // m, ci % m            # x index, z index
    jit = rng.uniform(-0.15, 0.15, (m * m, 2)) * cell
    xcl = x[Vb:]
    xcl[:, 0] += jit[:, 0]
    xcl[:, 1] += rng.uniform(-0.05, 0.05, m * m) * layer_gap
    if fold is not None:
        up = b > fold
        keep_z = b == fold + 1        # the crease row stays over row `fold`
        zi = np.where(up, 2 * fold + 1 - b, b)
        zj = np.where(keep_z, fold, zi)
        crease = keep_z | (b == fold)   # the crease rows share a z, up to a +-5% (of layer_gap) jitter
        zjit = np.where(crease, rng.uniform(-0.05, 0.05, m * m) * layer_gap, jit[:, 1])
        xcl[:, 2] = -0.5 * size + zj * cell + zjit
        xcl[up, 0] += 0.5 * cell
        xcl[up, 1] += layer_gap
    else:
        xcl[:, 2] += jit[:, 1]
    x[Vb:] = xcl
    s.x = x
    s.V_prev = x.copy()
    return s
