"""Scenes: body meshes, mass matrices and initial states in the layout the C ABI consumes.

The reference concatenates per-body meshes into one "soup" of rest vertices, edges, triangles
and a vertex->body map (geometry.py:31-73), loading bodies from a JSON scene
(utils/scene.py:65-85) whose meshes are ``.ma`` slab meshes (medial.py:17-51) or ``.obj``
files.  This module builds the same soup from

* the reference's JSON scenes (``load_json_scene``; reference-compat mass model), and
* procedural benchmark scenes: stacked cubes on a ground plane, a 32x32 torus chain net,
  batches of random convex-hull piles, and a pile of random boxes.

Layout (all C-contiguous):
    rest_x      (nv, 3) float64   body-frame rest positions
    vert_body   (nv,)   int32
    edges       (ne, 2) int32     global vertex ids
    tris        (nt, 3) int32     global vertex ids, outward-facing (CCW)
    body_scene  (nb,)   int32     non-decreasing
    body_pinned (nb,)   int32
    mass4       (nb, 4, 4) float64  the 12x12 inertia block is mass4 (x) I3
    q0, qdot0   (nb, 12) float64  [c, a0, a1, a2]; x = c + sum_k xr_k a_k
"""
import json
import os
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Scene:
    rest_x: np.ndarray
    vert_body: np.ndarray
    edges: np.ndarray
    tris: np.ndarray
    body_scene: np.ndarray
    body_pinned: np.ndarray
    mass4: np.ndarray
    q0: np.ndarray
    qdot0: np.ndarray
    name: str = "scene"
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        self.rest_x = np.ascontiguousarray(self.rest_x, dtype=np.float64).reshape(-1, 3)
        self.vert_body = np.ascontiguousarray(self.vert_body, dtype=np.int32)
        self.edges = np.ascontiguousarray(self.edges, dtype=np.int32).reshape(-1, 2)
        self.tris = np.ascontiguousarray(self.tris, dtype=np.int32).reshape(-1, 3)
        self.body_scene = np.ascontiguousarray(self.body_scene, dtype=np.int32)
        self.body_pinned = np.ascontiguousarray(self.body_pinned, dtype=np.int32)
        self.mass4 = np.ascontiguousarray(self.mass4, dtype=np.float64).reshape(-1, 4, 4)
        self.q0 = np.ascontiguousarray(self.q0, dtype=np.float64).reshape(-1, 12)
        self.qdot0 = np.ascontiguousarray(self.qdot0, dtype=np.float64).reshape(-1, 12)
        nb = len(self.body_scene)
        assert self.mass4.shape[0] == nb and self.q0.shape[0] == nb and self.qdot0.shape[0] == nb
        assert len(self.vert_body) == len(self.rest_x)
        assert np.all(np.diff(self.body_scene) >= 0), "body_scene must be non-decreasing"
        assert np.all(np.diff(self.vert_body) >= 0), "vertices must be grouped by body"

    n_bodies = property(lambda s: int(len(s.body_scene)))
    n_verts = property(lambda s: int(len(s.rest_x)))
    n_edges = property(lambda s: int(len(s.edges)))
    n_tris = property(lambda s: int(len(s.tris)))
    n_scenes = property(lambda s: int(s.body_scene.max()) + 1 if len(s.body_scene) else 0)

    def world_vertices(self, q=None):
        """compute_V (abd_simple.py:26-35): x = c + sum_k xr_k a_k."""
        q = self.q0 if q is None else np.asarray(q).reshape(-1, 12)
        qb = q[self.vert_body]
        return qb[:, 0:3] + self.rest_x[:, 0:1] * qb[:, 3:6] + self.rest_x[:, 1:2] * qb[:, 6:9] + \
            self.rest_x[:, 2:3] * qb[:, 9:12]

    def subset_scenes(self, scene_ids):
        """A new Scene holding only the given scenes (renumbered 0..k-1)."""
        return concat_scenes([self._single(s) for s in scene_ids], name=self.name + "_subset")

    def _single(self, s):
        bsel = np.nonzero(self.body_scene == s)[0]
        return _extract_bodies(self, bsel)


# ----------------------------------------------------------------------------------------------
# meshes
# ----------------------------------------------------------------------------------------------
def edges_of_tris(tris):
    """Unique undirected edges of a triangle list (the igl.edges call of medial.py:45)."""
    t = np.asarray(tris, dtype=np.int64).reshape(-1, 3)
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    e = np.sort(e, axis=1)
    return np.unique(e, axis=0).astype(np.int32)


def cube_mesh(size=(1.0, 1.0, 1.0)):
    """Axis-aligned box centred at the origin: 8 verts, 12 outward triangles, 18 edges."""
    sx, sy, sz = (0.5 * np.asarray(size, dtype=np.float64))
    V = np.array([[x, y, z] for x in (-sx, sx) for y in (-sy, sy) for z in (-sz, sz)])
    # vertex index = 4*ix + 2*iy + iz
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    F = []
    for a, b, c, d in quads:
        F += [(a, b, c), (a, c, d)]
    F = np.array(F, dtype=np.int32)
    return V, _orient_outward(V, F)


def torus_mesh(R=1.0, r=0.25, nu=24, nv=12):
    """Torus around the y axis (ring in the xz plane): nu*nv verts, 2*nu*nv tris, 3*nu*nv edges."""
    u = 2 * np.pi * np.arange(nu) / nu
    v = 2 * np.pi * np.arange(nv) / nv
    uu, vv = np.meshgrid(u, v, indexing="ij")
    x = (R + r * np.cos(vv)) * np.cos(uu)
    z = (R + r * np.cos(vv)) * np.sin(uu)
    y = r * np.sin(vv)
    V = np.stack([x, y, z], -1).reshape(-1, 3)
    F = []
    for i in range(nu):
        for j in range(nv):
            a = i * nv + j
            b = ((i + 1) % nu) * nv + j
            c = ((i + 1) % nu) * nv + (j + 1) % nv
            d = i * nv + (j + 1) % nv
            F += [(a, b, c), (a, c, d)]
    F = np.array(F, dtype=np.int32)
    # the (u, v) parametrisation orients all faces consistently; make the volume positive
    if mesh_volume_moments(V, F)[0] < 0:
        F = F[:, [0, 2, 1]].copy()
    return V, F


def hull_mesh(points):
    """Convex hull of a point set (scipy.spatial.ConvexHull), outward triangles."""
    from scipy.spatial import ConvexHull
    h = ConvexHull(points)
    used = np.unique(h.simplices)
    remap = -np.ones(len(points), dtype=np.int64)
    remap[used] = np.arange(len(used))
    V = np.asarray(points, dtype=np.float64)[used]
    F = remap[h.simplices].astype(np.int32)
    return V, _orient_outward(V, F)


def _orient_outward(V, F):
    """Flip each triangle of a convex mesh so its normal points away from the centroid."""
    F = F.copy()
    c = V.mean(0)
    a, b, cc = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    n = np.cross(b - a, cc - a)
    flip = np.einsum("ij,ij->i", n, (a + b + cc) / 3 - c) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return F


def mesh_volume_moments(V, F):
    """Volume integrals of a closed, outward-oriented triangle mesh by a tetrahedral fan from the
    origin: returns (vol, first (3,), second (3,3)) = (int 1, int x, int x x^T)."""
    V = np.asarray(V, dtype=np.float64)
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    vol6 = np.einsum("ij,ij->i", a, np.cross(b, c))
    vol = vol6 / 6.0
    s = a + b + c
    first = (vol[:, None] * s / 4.0).sum(0)
    # int_tet x x^T = vol/20 (sum_i v_i v_i^T + s s^T), with the origin as fourth vertex
    outer = lambda u: u[:, :, None] * u[:, None, :]
    second = (vol[:, None, None] / 20.0 * (outer(a) + outer(b) + outer(c) + outer(s))).sum(0)
    return vol.sum(), first, second


def mass4_from_mesh(V, F, density):
    """ABD mass matrix rho * int [1, x]^T [1, x] dV (4x4) of a closed mesh in its body frame."""
    vol, first, second = mesh_volume_moments(V, F)
    M = np.empty((4, 4))
    M[0, 0] = vol
    M[0, 1:] = first
    M[1:, 0] = first
    M[1:, 1:] = second
    return density * M


def centre_mesh(V, F):
    """Shift a mesh so that its volume centroid is at the origin; returns (V', centroid)."""
    vol, first, _ = mesh_volume_moments(V, F)
    com = first / vol
    return V - com, com


# ----------------------------------------------------------------------------------------------
# scene assembly
# ----------------------------------------------------------------------------------------------
def rot_euler_xyz(deg):
    """Rotation of scipy's Rotation.from_euler('xyz', deg, degrees=True) (utils/scene.py:84):
    extrinsic x, then y, then z."""
    ax, ay, az = np.deg2rad(np.asarray(deg, dtype=np.float64))
    cx, sx, cy, sy, cz, sz = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay), np.cos(az), np.sin(az)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def q_from_pose(c, Rm):
    """DOF vector [c, a0, a1, a2] for x = Rm xr + c (a_k = column k of Rm = row k of q)."""
    Rm = np.asarray(Rm, dtype=np.float64)
    return np.concatenate([np.asarray(c, dtype=np.float64), Rm[:, 0], Rm[:, 1], Rm[:, 2]])


class _Builder:
    def __init__(self):
        self.V, self.E, self.F, self.VB = [], [], [], []
        self.scene, self.pinned, self.M, self.q, self.qd = [], [], [], [], []
        self.nv = 0

    def add(self, V, F, E, mass4, q, qdot=None, pinned=False, scene=0):
        b = len(self.q)
        V = np.asarray(V, dtype=np.float64)
        self.V.append(V)
        self.F.append(np.asarray(F, dtype=np.int64).reshape(-1, 3) + self.nv)
        self.E.append(np.asarray(E, dtype=np.int64).reshape(-1, 2) + self.nv)
        self.VB.append(np.full(len(V), b, dtype=np.int32))
        self.nv += len(V)
        self.scene.append(scene)
        self.pinned.append(int(pinned))
        self.M.append(mass4)
        self.q.append(q)
        self.qd.append(np.zeros(12) if qdot is None else qdot)
        return b

    def build(self, name, meta=None):
        return Scene(rest_x=np.concatenate(self.V), vert_body=np.concatenate(self.VB),
                     edges=np.concatenate(self.E), tris=np.concatenate(self.F),
                     body_scene=np.array(self.scene), body_pinned=np.array(self.pinned),
                     mass4=np.array(self.M), q0=np.array(self.q), qdot0=np.array(self.qd),
                     name=name, meta=meta or {})


def _extract_bodies(sc, bsel):
    bld = _Builder()
    for b in bsel:
        vs = np.nonzero(sc.vert_body == b)[0]
        v0 = vs[0]
        tmask = np.isin(sc.tris[:, 0], vs)
        emask = np.isin(sc.edges[:, 0], vs)
        bld.add(sc.rest_x[vs], sc.tris[tmask] - v0, sc.edges[emask] - v0, sc.mass4[b], sc.q0[b],
                sc.qdot0[b], bool(sc.body_pinned[b]), 0)
    return bld.build(sc.name)


def concat_scenes(scenes, name="batch"):
    """Concatenate independent scenes into one multi-scene soup (scene index per body)."""
    bld = _Builder()
    for s, sc in enumerate(scenes):
        off_b = len(bld.q)
        for b in range(sc.n_bodies):
            vs = np.nonzero(sc.vert_body == b)[0]
            v0 = vs[0]
            tmask = (sc.tris[:, 0] >= v0) & (sc.tris[:, 0] <= vs[-1])
            emask = (sc.edges[:, 0] >= v0) & (sc.edges[:, 0] <= vs[-1])
            bld.add(sc.rest_x[vs], sc.tris[tmask] - v0, sc.edges[emask] - v0, sc.mass4[b],
                    sc.q0[b], sc.qdot0[b], bool(sc.body_pinned[b]), s)
    return bld.build(name, {"scenes": [sc.name for sc in scenes]})


# ----------------------------------------------------------------------------------------------
# benchmark scenes
# ----------------------------------------------------------------------------------------------
def cubes_on_ground(seed=1, n=2, spacing=1.05, jitter=0.02, density=1e3, lift=0.05):
    """8 unit cubes in a 2x2x2 lattice at spacing 1.05 with uniform jitter +-0.02 (seeded),
    above a static ground plane y = 0 (the lowest cube starts `lift` above the ground)."""
    rng = np.random.default_rng(seed)
    V, F = cube_mesh()
    E = edges_of_tris(F)
    M = mass4_from_mesh(V, F, density)
    bld = _Builder()
    for i in range(n):
        for j in range(n):
            for k in range(n):
                c = np.array([i * spacing, 0.5 + lift + j * spacing, k * spacing])
                c = c + rng.uniform(-jitter, jitter, 3)
                bld.add(V, F, E, M, q_from_pose(c, np.eye(3)))
    sc = bld.build("cubes_on_ground", {"seed": seed, "density": density})
    # keep every cube above the plane despite the jitter
    low = sc.world_vertices()[:, 1].min()
    if low < 0.5 * lift:
        sc.q0[:, 1] += 0.5 * lift - low
    return sc


def chain_net(n=32, R=1.0, r=0.25, nu=24, nv=12, density=1e3, pitch=(1.25, 1.26), tilt_deg=35.0,
              pin_corners=True):
    """n x n chain net of interlocked torus links (European 4-in-1 weave): link (i, j) sits at
    (i*a + (j%2)*a/2, 0, j*b), its ring plane tilted by +-tilt about z (alternating by row), so
    each ring passes through two rings of the row before and two of the row after.  The four
    corner links are pinned."""
    V, F = torus_mesh(R, r, nu, nv)
    E = edges_of_tris(F)
    Vc, _ = centre_mesh(V, F)
    M = mass4_from_mesh(Vc, F, density)
    a, b = pitch
    th = np.deg2rad(tilt_deg)
    bld = _Builder()
    corners = {(0, 0), (n - 1, 0), (0, n - 1), (n - 1, n - 1)}
    for j in range(n):
        s = 1.0 if j % 2 == 0 else -1.0
        cz, sz = np.cos(s * th), np.sin(s * th)
        Rm = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
        for i in range(n):
            c = np.array([i * a + (j % 2) * a / 2, 0.0, j * b])
            bld.add(Vc, F, E, M, q_from_pose(c, Rm), pinned=pin_corners and (i, j) in corners)
    return bld.build("chain_net", {"n": n, "pitch": pitch, "tilt_deg": tilt_deg})


def random_rotation(rng):
    """Uniform random rotation (QR of a Gaussian matrix with a sign fix)."""
    Q, Rr = np.linalg.qr(rng.normal(size=(3, 3)))
    Q = Q * np.sign(np.diag(Rr))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


def hull_pile(n_bodies, rng, box=10.0, density=1e3, npts=(20, 200), ground=True):
    """A random pile of convex hulls of U{20..200} points uniform in the unit ball, placed on a
    jittered grid inside a box^3 region (non-overlapping bounding spheres)."""
    bld = _Builder()
    k = int(np.ceil(n_bodies ** (1 / 3)))
    pitch = box / k
    slots = [(i, j, l) for j in range(k) for i in range(k) for l in range(k)][:n_bodies]
    for (i, j, l) in slots:
        m = rng.integers(npts[0], npts[1] + 1)
        p = rng.normal(size=(m, 3))
        p = p / np.linalg.norm(p, axis=1, keepdims=True) * rng.uniform(0, 1, (m, 1)) ** (1 / 3)
        V, F = hull_mesh(p)
        V, _ = centre_mesh(V, F)
        E = edges_of_tris(F)
        scale = min(1.0, 0.45 * pitch)
        V = V * scale
        M = mass4_from_mesh(V, F, density)
        c = np.array([(i + 0.5) * pitch - box / 2, (j + 0.5) * pitch + 0.1, (l + 0.5) * pitch - box / 2])
        c = c + rng.uniform(-0.02, 0.02, 3) * pitch
        bld.add(V, F, E, M, q_from_pose(c, random_rotation(rng)))
    return bld.build("hull_pile")


def batched_hull_piles(n_scenes=256, seed=3, lo=4, hi=512, box=10.0):
    """Many independent hull piles, body counts log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), n_scenes)).astype(int)
    counts = np.clip(counts, lo, hi)
    scenes = [hull_pile(int(c), rng, box=box) for c in counts]
    sc = concat_scenes(scenes, name="batched_hull_piles")
    sc.meta.update({"seed": seed, "counts": counts.tolist()})
    return sc


def box_pile(n=4096, seed=4, density=1e3, edge=(0.2, 1.0), spacing=1.2):
    """n boxes with edge lengths U(0.2, 1.0) and uniform random orientations on a jittered grid
    above the ground plane."""
    rng = np.random.default_rng(seed)
    k = int(np.ceil(np.sqrt(n / 16)))  # a k x 16 x k stack
    bld = _Builder()
    cnt = 0
    for j in range(10**6):
        for i in range(k):
            for l in range(k):
                if cnt == n:
                    break
                size = rng.uniform(edge[0], edge[1], 3)
                V, F = cube_mesh(size)
                E = edges_of_tris(F)
                M = mass4_from_mesh(V, F, density)
                c = np.array([i * spacing, 0.8 + j * spacing, l * spacing]) + rng.uniform(-0.05, 0.05, 3)
                bld.add(V, F, E, M, q_from_pose(c, random_rotation(rng)))
                cnt += 1
        if cnt == n:
            break
    return bld.build("box_pile", {"seed": seed})


# ----------------------------------------------------------------------------------------------
# reference scenes (JSON + .ma/.obj), reference-compat mass model
# ----------------------------------------------------------------------------------------------
def read_ma(path):
    """SlabMesh.load (medial.py:17-51): 'v x y z r', 'e i j', 'f i j k'; edges are the unique
    union of the listed edges and the faces' edges."""
    Vs, Es, Fs = [], [], []
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                Vs.append([float(t) for t in p[1:4]])
            elif p[0] == "e":
                Es.append([int(t) for t in p[1:3]])
            elif p[0] == "f":
                Fs.append([int(t) for t in p[1:4]])
    V = np.array(Vs, dtype=np.float64)
    F = np.array(Fs, dtype=np.int32).reshape(-1, 3)
    E = np.array(Es, dtype=np.int64).reshape(-1, 2)
    if len(F):
        E = np.concatenate([E, edges_of_tris(F)])
    E = np.unique(np.sort(E, axis=1), axis=0).astype(np.int32)
    return V, E, F


def read_obj(path):
    """Minimal Wavefront OBJ reader (positions and triangular faces; 1-based indices)."""
    Vs, Fs = [], []
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                Vs.append([float(t) for t in p[1:4]])
            elif p[0] == "f":
                idx = [int(t.split("/")[0]) - 1 for t in p[1:]]
                for k in range(1, len(idx) - 1):
                    Fs.append([idx[0], idx[k], idx[k + 1]])
    V = np.array(Vs, dtype=np.float64)
    F = np.array(Fs, dtype=np.int32).reshape(-1, 3)
    return V, edges_of_tris(F), F


def load_json_scene(path, base_dir=None, length=0.4, reference_init=True):
    """JSONComplex (utils/scene.py:65-85) + reference inertia (abd_simple.py:18-22,
    rbd_simple.py:42-48): per object v = V R^T * 2 with R = euler 'xyz' degrees; c = p;
    m = mass/1000, J = 0.4 m L^2 (L = 0.4); mass <= 0 means pinned.  The reference ignores the
    JSON velocities and starts every body with v = 0 and qdot = skew((3, 3, 0))
    (abd_simple.py:37-49); ``reference_init=False`` starts from rest instead."""
    base_dir = base_dir or os.path.dirname(os.path.dirname(os.path.abspath(path)))
    with open(path) as f:
        objs = json.load(f)
    cache = {}
    bld = _Builder()
    for o in objs:
        mp = o.get("obj", "")
        fp = mp if os.path.isabs(mp) else os.path.join(base_dir, mp)
        if not os.path.exists(fp):  # mesh stored beside the JSON file
            fp = os.path.join(os.path.dirname(os.path.abspath(path)), os.path.basename(mp))
        if fp not in cache:
            cache[fp] = read_ma(fp) if fp.endswith(".ma") else read_obj(fp)
        V, E, F = cache[fp]
        Rm = rot_euler_xyz(o.get("euler", [0, 0, 0]))
        v = V @ Rm.T * 2.0
        m = float(o.get("mass", 1.0)) / 1000.0
        J = 0.4 * m * length * length
        pinned = m <= 0
        M = np.diag([m, J, J, J]) if not pinned else np.diag([1.0, 1.0, 1.0, 1.0])
        qd = np.zeros(12)
        if reference_init:
            # rows of skew(w), w = (3, 3, 0): [[0, -wz, wy], [wz, 0, -wx], [-wy, wx, 0]]
            qd[3:] = [0.0, 0.0, 3.0, 0.0, 0.0, -3.0, -3.0, 3.0, 0.0]
        bld.add(v, F, E, M, q_from_pose(o.get("p", [0, 0, 0]), np.eye(3)), qd, pinned=pinned)
    return bld.build(os.path.basename(path))
