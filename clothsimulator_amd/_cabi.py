"""ctypes mirror of include/clothsim.h (the C-ABI of libclothsim.so).

The struct layouts here must match the header field for field; the
``test_cabi_layout`` test checks the sizes against the compiled library.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclothsim.so")

CS_OK = 0
CS_E_ARG = 1
CS_E_HIP = 2
CS_E_OOM = 3
CS_E_NOT_CONVERGED = 4
CS_E_COLLISION_FAILED = 5
CS_E_FORMAT = 6
CS_E_STATE = 7
CS_E_CAPACITY = 8
CS_E_NO_DEVICE = 9
STATUS_NAMES = {0: "CS_OK", 1: "CS_E_ARG", 2: "CS_E_HIP", 3: "CS_E_OOM", 4: "CS_E_NOT_CONVERGED",
                5: "CS_E_COLLISION_FAILED", 6: "CS_E_FORMAT", 7: "CS_E_STATE", 8: "CS_E_CAPACITY",
                9: "CS_E_NO_DEVICE"}

CS_PRECOND_BLOCK_JACOBI = 0
CS_PRECOND_LEARNED = 1
CS_PRECOND_AMG = 2
CS_PRECOND_AMG_LEARNED = 3
CS_MESH_CLOTH = 0
CS_MESH_OBSTACLE = 1
CS_FLAG_NO_PROXIMITY = 1
CS_FLAG_NO_COLLISION = 2
CS_FLAG_TIME_KERNELS = 4
CS_FLAG_INSTANCES = 8
CS_FLAG_FIXED_TUNING = 16
CS_FLAG_GRAPH_CYCLES = 32
CS_FLAG_PERSISTENT_BATCH = 64
CS_LBVH_CCD = 1
CS_LBVH_REFIT = 2
CS_KERNEL_PCG_SPMV = 0
CS_KERNEL_ASSEMBLE_ROWS = 1
CS_KERNEL_FACE_RECORDS = 2
CS_KERNEL_EVENT_PAIR = 3


class Params(C.Structure):
    _fields_ = [
        ("dt", C.c_float),
        ("gravity", C.c_float * 3),
        ("wind_density", C.c_float),
        ("wind_drag", C.c_float),
        ("wind_velocity", C.c_float * 3),
        ("handle_stiffness", C.c_float),
        ("collision_stiffness", C.c_float),
        ("repulsion_thickness", C.c_float),
        ("collision_thickness", C.c_float),
        ("friction", C.c_float),
        ("obstacle_friction", C.c_float),
        ("pcg_rel_tol", C.c_float),
        ("pcg_max_iters", C.c_int),
        ("precond", C.c_int),
        ("max_collision_iters", C.c_int),
        ("flags", C.c_int),
    ]


class MeshIn(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int),
        ("x", C.POINTER(C.c_float)),
        ("v", C.POINTER(C.c_float)),
        ("n_verts", C.c_int),
        ("uv", C.POINTER(C.c_float)),
        ("n_faces", C.c_int),
        ("face_nodes", C.POINTER(C.c_int)),
        ("face_verts", C.POINTER(C.c_int)),
    ]


class MaterialIn(C.Structure):
    _fields_ = [
        ("density", C.c_float),
        ("stretching", (C.c_float * 4) * 6),
        ("bending", (C.c_float * 5) * 3),
        ("thicken", C.c_float),
        ("density_mult", C.c_float),
        ("stretching_mult", C.c_float),
        ("bending_mult", C.c_float),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("pcg_iters", C.c_int),
        ("pcg_rel_residual", C.c_float),
        ("pcg_converged", C.c_int),
        ("n_proximities", C.c_int),
        ("n_contact_blocks", C.c_int),
        ("collision_iters", C.c_int),
        ("collision_deform", C.c_int),
        ("n_impacts", C.c_int),
        ("ms_bvh", C.c_float),
        ("ms_proximity", C.c_float),
        ("ms_assemble", C.c_float),
        ("ms_solve", C.c_float),
        ("ms_update", C.c_float),
        ("ms_collision", C.c_float),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ProximityRec(C.Structure):
    _fields_ = [
        ("nodes", C.c_int * 4),
        ("n", C.c_float * 3),
        ("mu", C.c_float),
        ("stiffness", C.c_float),
        ("w", C.c_float * 4),
    ]


class ImpactRec(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("nodes", C.c_int * 4),
        ("t", C.c_double),
        ("w", C.c_double * 4),
        ("n", C.c_double * 3),
    ]


class SceneInfo(C.Structure):
    _fields_ = [
        ("frame_time", C.c_float),
        ("frame_steps", C.c_int),
        ("end_time", C.c_float),
        ("end_frame", C.c_int),
        ("n_cloths", C.c_int),
        ("n_obstacles", C.c_int),
        ("n_handles", C.c_int),
        ("n_cloth_nodes", C.c_int),
        ("n_cloth_faces", C.c_int),
        ("n_motions", C.c_int),
        ("n_moving_obstacles", C.c_int),
        ("remeshing", C.c_int),
    ]


CS_MAX_MOTIONS = 64


class Transform(C.Structure):
    """cs_transform: p -> scale * rotate(p) + translate, rotate = quaternion (s, x, y, z)."""
    _fields_ = [
        ("scale", C.c_float),
        ("translate", C.c_float * 3),
        ("rotate", C.c_float * 4),
    ]

    def as_array(self):
        return np.array([self.scale] + list(self.translate) + list(self.rotate), dtype=np.float32)


def transform_struct(scale=1.0, translate=(0.0, 0.0, 0.0), rotate=(1.0, 0.0, 0.0, 0.0)):
    t = Transform()
    t.scale = scale
    t.translate[:] = [float(v) for v in translate]
    t.rotate[:] = [float(v) for v in rotate]
    return t


def axis_angle(axis, degrees):
    """The quaternion of Transformation(json) "rotate": [degrees, ax, ay, az] (float32 steps of
    src/Transformation.cu:9-14 and src/Quaternion.cu:7-16)."""
    ang = np.float32(np.float32(degrees) * np.pi / np.float32(180.0))
    if ang == 0:
        return (1.0, 0.0, 0.0, 0.0)
    a = np.asarray(axis, dtype=np.float32)
    a = a * (np.float32(1) / np.sqrt(np.float32(a @ a)))
    h = np.float32(0.5) * ang
    s = np.sin(h, dtype=np.float32)
    return (float(np.cos(h, dtype=np.float32)), float(s * a[0]), float(s * a[1]), float(s * a[2]))


class Remeshing(C.Structure):
    """cs_remeshing: a cloth's "remeshing" block."""
    _fields_ = [(n, C.c_float) for n in ("refine_angle", "refine_velocity", "refine_compression", "rib_stiffening",
                                          "size_min", "size_max", "aspect_min", "flip_threshold")]


def remeshing_struct(refine_angle=np.inf, refine_velocity=np.inf, refine_compression=np.inf, rib_stiffening=1.0,
                     size=None, aspect_min=-np.inf, flip_threshold=1e-2):
    """The block's fields with the reference's defaults (src/Remeshing.cu:3-19)."""
    r = Remeshing()
    r.refine_angle, r.refine_velocity, r.refine_compression = refine_angle, refine_velocity, refine_compression
    r.rib_stiffening = rib_stiffening
    r.size_min, r.size_max = (-np.inf, np.inf) if size is None else size
    r.aspect_min, r.flip_threshold = aspect_min, flip_threshold
    return r


PROX_DTYPE = np.dtype([("nodes", "<i4", 4), ("n", "<f4", 3), ("mu", "<f4"), ("stiffness", "<f4"), ("w", "<f4", 4)])
IMPACT_DTYPE = np.dtype({"names": ["type", "nodes", "t", "w", "n"],
                         "formats": ["<i4", ("<i4", 4), "<f8", ("<f8", 4), ("<f8", 3)],
                         "offsets": [0, 4, 24, 32, 64], "itemsize": 88})
assert C.sizeof(ProximityRec) == PROX_DTYPE.itemsize
assert C.sizeof(ImpactRec) == IMPACT_DTYPE.itemsize


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_params(dt=0.005, gravity=(0.0, 0.0, -9.8)):
    """Reference defaults: src/Magic.cu:5-17, src/Wind.cu:3-7, src/Simulator.cu:19-27."""
    p = Params()
    p.dt = dt
    p.gravity[:] = list(gravity)
    p.wind_density = 1.0
    p.wind_drag = 0.0
    p.wind_velocity[:] = [0.0, 0.0, 0.0]
    p.handle_stiffness = 1e3
    p.collision_stiffness = 1e9
    p.repulsion_thickness = 1e-3
    p.collision_thickness = 1e-4
    p.friction = 0.6
    p.obstacle_friction = 0.3
    p.pcg_rel_tol = 1e-5
    p.pcg_max_iters = 4000
    p.precond = CS_PRECOND_BLOCK_JACOBI
    p.max_collision_iters = 100
    p.flags = 0
    return p


class MeshArrays:
    """Host copy of one mesh, kept alive while a MeshIn points into it."""

    def __init__(self, x, faces, uv=None, face_verts=None, v=None):
        self.x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
        self.faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
        if uv is None:  # OBJ faces without '/': uv = raw (x, y) of the node (src/Mesh.cu:334-338)
            uv = self.x[:, :2].copy()
            face_verts = self.faces.copy()
        self.uv = np.ascontiguousarray(uv, dtype=np.float32).reshape(-1, 2)
        self.face_verts = np.ascontiguousarray(face_verts if face_verts is not None else self.faces,
                                               dtype=np.int32).reshape(-1, 3)
        self.v = None if v is None else np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)

    @property
    def n_nodes(self):
        return self.x.shape[0]

    @property
    def n_faces(self):
        return self.faces.shape[0]

    def c_struct(self):
        m = MeshIn()
        m.n_nodes = self.x.shape[0]
        m.x = fptr(self.x)
        m.v = fptr(self.v) if self.v is not None else C.POINTER(C.c_float)()
        m.n_verts = self.uv.shape[0]
        m.uv = fptr(self.uv)
        m.n_faces = self.faces.shape[0]
        m.face_nodes = iptr(self.faces)
        m.face_verts = iptr(self.face_verts)
        return m


def material_struct(density=0.15, stretch=5e3, bend=1e-5, stretching=None, bending=None,
                    thicken=1.0, density_mult=1.0, stretching_mult=1.0, bending_mult=1.0):
    """Material JSON (src/Material.cu:19-40).  The default is the configs' isotropic
    material: every stretching row [k, 0, k, k] (the reference doubles it,
    src/Material.cu:82) and a constant 3x5 bending table."""
    m = MaterialIn()
    m.density = density
    rows = stretching if stretching is not None else [[stretch, 0.0, stretch, stretch]] * 6
    for i in range(6):
        for j in range(4):
            m.stretching[i][j] = rows[i][j]
    btab = bending if bending is not None else [[bend] * 5] * 3
    for i in range(3):
        for j in range(5):
            m.bending[i][j] = btab[i][j]
    m.thicken = thicken
    m.density_mult = density_mult
    m.stretching_mult = stretching_mult
    m.bending_mult = bending_mult
    return m
