"""Synthetic scenes of the benchmark configurations.

Every generator is deterministic.  Meshes are returned as ``MeshArrays`` and
scenes as ``SceneSpec`` (params, cloths, obstacles, handles) that can be fed to
the product (``clothsimulator_amd.Simulator``) or to the CPU oracle alike.

Workloads (cloth side n means an n x n node grid, quads split along the main
diagonal, uv = (x, y) in metres as the reference assigns to OBJ faces without
'/', src/Mesh.cu:334-338):

* hang33    1 m x 1 m 33^2 grid at z = 0, the two corners of the y = 1 edge are
            handles, dt = 5 ms, no obstacle.
* drape129  129^2 grid, 1 m x 1 m centred 0.4 m above a fixed icosphere of
            radius 0.3 (642 nodes / 1280 faces), dt = 5 ms.
* crumple513 513^2 grid, z = 0.5 m + N(0, 1e-3 m) i.i.d. (splitmix64 +
            Box-Muller, seed 0), over a ground plane of two triangles, dt = 2 ms.
"""
import math
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

from ._cabi import MeshArrays, default_params, material_struct


def grid_mesh(n, size=1.0, origin=(0.0, 0.0, 0.0)):
    """n x n node grid over [0, size]^2 (+origin); node (i, j) -> index j*n + i.
    Quad (i, j)-(i+1, j+1) is split along its main diagonal into
    (a, b, c) and (a, c, d) with a=(i,j), b=(i+1,j), c=(i+1,j+1), d=(i,j+1)."""
    h = size / (n - 1)
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = np.zeros((n * n, 3), dtype=np.float64)
    x[:, 0] = ii.ravel() * h + origin[0]
    x[:, 1] = jj.ravel() * h + origin[1]
    x[:, 2] = origin[2]
    i = np.arange(n - 1)
    I, J = np.meshgrid(i, i, indexing="xy")
    a = (J * n + I).ravel()
    b = a + 1
    c = a + n + 1
    d = a + n
    faces = np.empty((2 * (n - 1) * (n - 1), 3), dtype=np.int32)
    faces[0::2] = np.stack([a, b, c], axis=1)
    faces[1::2] = np.stack([a, c, d], axis=1)
    return x, faces


def icosphere(level=3, radius=1.0):
    """Subdivided icosahedron: level 3 gives 642 nodes and 1280 faces, the
    counts of the reference's unit sphere asset."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}
        new_faces = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for (a, b, c) in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new_faces
    x = np.array(verts) * radius
    return x, np.array(faces, dtype=np.int32)


def splitmix64_normals(count, seed=0):
    """i.i.d. N(0, 1) from splitmix64 and the Box-Muller transform in FP64
    (a project-defined generator, reproducible across platforms)."""
    mask = (1 << 64) - 1
    state = seed & mask
    out = np.empty(count, dtype=np.float64)
    k = 0
    while k < count:
        us = []
        for _ in range(2):
            state = (state + 0x9E3779B97F4A7C15) & mask
            z = state
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
            z = z ^ (z >> 31)
            us.append(((z >> 11) + 0.5) / float(1 << 53))
        r = math.sqrt(-2.0 * math.log(us[0]))
        out[k] = r * math.cos(2.0 * math.pi * us[1])
        if k + 1 < count:
            out[k + 1] = r * math.sin(2.0 * math.pi * us[1])
        k += 2
    return out


def splitmix64_normals_fast(count, seed=0):
    """Vectorised splitmix64 + Box-Muller; identical output to splitmix64_normals."""
    n_pairs = (count + 1) // 2
    idx = np.arange(1, 2 * n_pairs + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = ((z >> np.uint64(11)).astype(np.float64) + 0.5) / float(1 << 53)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * n_pairs, dtype=np.float64)
    out[0::2] = r * np.cos(2.0 * np.pi * u2)
    out[1::2] = r * np.sin(2.0 * np.pi * u2)
    return out[:count]


@dataclass
class ClothSpec:
    mesh: MeshArrays
    material: object
    handles: List[int] = field(default_factory=list)
    handle_motion: int = -1  # the motion moving the handles (-1: fixed)
    remeshing: object = None  # a Remeshing struct: the cloth is remeshed by Simulator.remesh()


@dataclass
class SceneSpec:
    name: str
    params: object
    cloths: List[ClothSpec]
    obstacles: List[MeshArrays]
    steps: int
    # keyframe motions: (times, keys), a key = 8 floats (scale, translate xyz, rotate s x y z)
    motions: list = field(default_factory=list)
    obstacle_motions: List[int] = field(default_factory=list)  # per obstacle (-1: static)

    @property
    def n_cloth_nodes(self):
        return sum(c.mesh.n_nodes for c in self.cloths)


def hang(n=33, steps=100):
    x, f = grid_mesh(n, 1.0)
    mesh = MeshArrays(x, f)
    handles = [(n - 1) * n + 0, (n - 1) * n + (n - 1)]
    return SceneSpec("hang%d" % n, default_params(dt=0.005), [ClothSpec(mesh, material_struct(), handles)], [], steps)


def drape(n=129, steps=400, sphere_level=3):
    x, f = grid_mesh(n, 1.0, origin=(-0.5, -0.5, 0.4))
    cloth = MeshArrays(x, f)
    sx, sf = icosphere(sphere_level, 0.3)
    return SceneSpec("drape%d" % n, default_params(dt=0.005), [ClothSpec(cloth, material_struct())],
                     [MeshArrays(sx, sf)], steps)


def sphere_sweep(n=129, steps=800, sphere_level=3):
    """A sheet hanging from two corners, pushed back and forth by a sphere: the motion and the
    placement of the reference's conf/sphere.json (sphere scaled by 0.1 in every key, at rest
    until t = 2 s, then swept along y), with an n x n grid in place of the remeshed square and
    an icosphere of radius 1 in place of meshes/sphere.obj."""
    x, f = grid_mesh(n, 1.0, origin=(-0.5, -1.0, 0.5))
    handles = [(n - 1) * n + 0, (n - 1) * n + (n - 1)]  # the two corners at y = 0
    sx, sf = icosphere(sphere_level, 1.0)
    ys = [0.0, 0.0, 0.0, -1.5, 0.0, 1.5, 0.0, -0.4]
    keys = [[0.1, 0.0, y, 0.0, 1.0, 0.0, 0.0, 0.0] for y in ys]
    p = default_params(dt=0.005)
    p.repulsion_thickness = 5e-3
    p.collision_thickness = 5e-4
    p.collision_stiffness = 1e6
    return SceneSpec("sphere_sweep%d" % n, p, [ClothSpec(MeshArrays(x, f), material_struct(), handles)],
                     [MeshArrays(sx, sf)], steps, motions=[(list(range(8)), keys)], obstacle_motions=[0])


def ground_plane(half=2.0, z=0.0):
    x = np.array([[-half, -half, z], [half, -half, z], [half, half, z], [-half, half, z]], dtype=np.float64)
    f = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    return MeshArrays(x, f)


def crumple(n=513, steps=300, seed=0, height=0.5, sigma=1e-3):
    x, f = grid_mesh(n, 1.0, origin=(-0.5, -0.5, height))
    x[:, 2] += sigma * splitmix64_normals_fast(n * n, seed)
    # uv from the unperturbed (x, y): the perturbation is out of plane
    cloth = MeshArrays(x, f)
    return SceneSpec("crumple%d" % n, default_params(dt=0.002), [ClothSpec(cloth, material_struct())],
                     [ground_plane()], steps)


def fold(n=65, gap=1.5e-3, steps=50, seed=0, noise=2e-4):
    """An n x n sheet (1 m, uv = the flat grid) folded over itself: rows j < n // 2 lie flat at
    z = 0; rows j >= n // 2 are laid back over them, reversed, offset by half a cell in x and y,
    at height `gap` (+ U[-noise, noise]), joined to the lower half by one strip of slanted faces.
    Every upper node starts within `gap` of a lower face: a self-contact state from the first
    step, over a ground plane 5 cm below."""
    x, f = grid_mesh(n, 1.0)
    uv = x[:, :2].copy()
    h = 1.0 / (n - 1)
    m = n // 2
    rng = np.random.default_rng(seed)
    X = x.copy()
    for j in range(m, n):
        row = slice(j * n, (j + 1) * n)
        X[row, 0] = x[row, 0] + 0.5 * h
        X[row, 1] = (n - 1 - j) * h + 0.5 * h
        X[row, 2] = gap + rng.uniform(-noise, noise, n)
    mesh = MeshArrays(X, f, uv=uv, face_verts=f)
    return SceneSpec("fold%d" % n, default_params(dt=0.002), [ClothSpec(mesh, material_struct())],
                     [ground_plane(z=-0.05)], steps)


def instances(count=256, seed=0, side_lo=17, side_hi=129, steps=200, spacing=3.0, gap=None):
    """Independent drapes in one context (CS_FLAG_INSTANCES): instance k is a 1 m square
    grid with side drawn uniformly from {side_lo..side_hi} (uniform over the side, not the node
    count), stretch stiffness scaled by U[0.5, 2], 0.5 m above its own icosphere of radius
    U[0.2, 0.4] m whose centre (at height 0) is offset by U[-0.2, 0.2] m in x and y (with
    `gap`, the cloth starts that far above its sphere's top instead); instances sit on a
    square lattice `spacing` m apart.  Cloth k and obstacle k form instance k."""
    from ._cabi import CS_FLAG_INSTANCES
    rng = np.random.default_rng(seed)
    per_row = int(math.ceil(math.sqrt(count)))
    sphere_x, sphere_f = icosphere(3, 1.0)
    cloths, obstacles = [], []
    for k in range(count):
        n = int(rng.integers(side_lo, side_hi + 1))
        stiff = float(rng.uniform(0.5, 2.0))
        radius = float(rng.uniform(0.2, 0.4))
        off = rng.uniform(-0.2, 0.2, 2)
        cx, cy = spacing * (k % per_row), spacing * (k // per_row)
        z = 0.5 if gap is None else radius + gap
        x, f = grid_mesh(n, 1.0, origin=(cx - 0.5, cy - 0.5, z))
        cloths.append(ClothSpec(MeshArrays(x, f), material_struct(stretch=5e3 * stiff)))
        sx = sphere_x * radius + np.array([cx + off[0], cy + off[1], 0.0])
        obstacles.append(MeshArrays(sx, sphere_f))
    p = default_params(dt=0.005)
    p.flags |= CS_FLAG_INSTANCES
    return SceneSpec("batch%d" % count, p, cloths, obstacles, steps)


# the reference's conf/hang.json "remeshing" block (Narain et al. 2012, Figure 2(b))
HANG_REMESHING = dict(refine_angle=0.3, refine_compression=0.005, refine_velocity=0.5, size=(10e-3, 200e-3),
                      aspect_min=0.2)


def drape_remesh(n=65, steps=400):
    """drape(n) with the remeshing block of the reference's conf/hang.json (a sheet falling on a
    sphere, remeshed at the end of every 8-step frame): the remeshing workload."""
    from ._cabi import remeshing_struct
    spec = drape(n, steps)
    spec.name = "remesh_drape%d" % n
    spec.cloths[0].remeshing = remeshing_struct(**HANG_REMESHING)
    return spec


def by_name(name):
    table = {"hang33": lambda: hang(33), "drape129": lambda: drape(129), "crumple513": lambda: crumple(513)}
    if name in table:
        return table[name]()
    # parametrised small variants: hangN, drapeN, crumpleN
    for prefix, fn in (("hang", hang), ("drape", drape), ("crumple", crumple), ("batch", instances),
                       ("sphere_sweep", sphere_sweep), ("remesh_drape", drape_remesh)):
        if name.startswith(prefix) and name[len(prefix):].isdigit():
            return fn(int(name[len(prefix):]))
    raise KeyError(name)
