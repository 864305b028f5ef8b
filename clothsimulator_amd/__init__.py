"""clothsimulator_amd — MI355X-native per-step cloth physics (ClothSimulator's --gpu step).

The compute lives in libclothsim.so (hand-written gfx950 HIP kernels behind the
C-ABI of include/clothsim.h).  This module is a thin ctypes mirror of that ABI
used by tests, benchmarks and Python callers.  There is no CPU fallback: if
the library is missing, or no gfx950 device is present, every call fails.
"""
import ctypes as C
import os

import numpy as np

from ._cabi import (CS_E_COLLISION_FAILED, CS_E_NOT_CONVERGED, CS_MESH_CLOTH, CS_PRECOND_AMG_LEARNED, CS_PRECOND_LEARNED, CS_MESH_OBSTACLE, CS_OK, IMPACT_DTYPE, LIB_PATH, PROX_DTYPE,
                    STATUS_NAMES, ImpactRec, MaterialIn, MeshArrays, MeshIn, Params, ProximityRec, SceneInfo,
                    Remeshing, Stats, Transform, axis_angle, default_params, fptr, iptr, material_struct,
                    remeshing_struct, transform_struct)

__all__ = ["load_library", "Simulator", "ClothError", "default_params", "material_struct", "MeshArrays"]

_lib = None

EXPORTED = ["cs_version", "cs_default_params", "cs_create", "cs_destroy", "cs_last_error", "cs_set_params",
            "cs_add_cloth", "cs_add_obstacle", "cs_add_handles", "cs_physics_step", "cs_collision_step", "cs_step",
            "cs_run", "cs_get_counts", "cs_get_state", "cs_set_state", "cs_get_geometry", "cs_material_table",
            "cs_lbvh_pairs", "cs_find_proximities", "cs_assemble", "cs_solve", "cs_find_impacts",
            "cs_independent_impacts", "cs_kernel_time", "cs_load_scene", "cs_scene_error", "cs_precondition",
            "cs_lpc_weight_count", "cs_set_learned_weights", "cs_get_learned_weights", "cs_lpc_factor",
            "cs_lpc_train_step", "cs_lpc_smoother_omega", "cs_batch_tuning", "cs_add_motion", "cs_motion_eval",
            "cs_set_obstacle_motion", "cs_add_moving_handles", "cs_set_step_count", "cs_get_step_count",
            "cs_scene_motion", "cs_get_topology", "cs_write_config", "cs_impact_response",
            "cs_solve_bsr_noalloc", "cs_lpc_train_step_smooth", "cs_default_remeshing", "cs_remesh_mesh",
            "cs_remesh_result_counts", "cs_remesh_result_get", "cs_remesh_result_free", "cs_set_remeshing",
            "cs_remesh", "cs_remesh_sizing", "cs_remesh_times", "cs_separation_step", "cs_frame_begin",
            "cs_frame_end"]

# trained weights of the learned preconditioner shipped with the package (float32, .npy)
LPC_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lpc_weights.npy")


class ClothError(RuntimeError):
    def __init__(self, status, what, msg=""):
        self.status = status
        super().__init__("%s failed: %s %s" % (what, STATUS_NAMES.get(status, status), msg))


def load_library(path=LIB_PATH):
    """Load libclothsim.so.  Raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError("%s is not built: run `python -m clothsimulator_amd.build`" % path)
    L = C.CDLL(path)
    vp = C.c_void_p
    L.cs_version.restype = C.c_char_p
    L.cs_default_params.argtypes = [C.POINTER(Params)]
    L.cs_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.cs_destroy.argtypes = [vp]
    L.cs_last_error.restype = C.c_char_p
    L.cs_last_error.argtypes = [vp]
    L.cs_set_params.argtypes = [vp, C.POINTER(Params)]
    L.cs_add_cloth.argtypes = [vp, C.POINTER(MeshIn), C.POINTER(MaterialIn), C.POINTER(C.c_int)]
    L.cs_add_obstacle.argtypes = [vp, C.POINTER(MeshIn), C.POINTER(C.c_int)]
    L.cs_add_handles.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.c_int]
    for name in ("cs_physics_step", "cs_collision_step", "cs_step"):
        getattr(L, name).argtypes = [vp, C.POINTER(Stats)]
    L.cs_run.argtypes = [vp, C.c_int, C.POINTER(Stats)]
    L.cs_get_counts.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    F = C.POINTER(C.c_float)
    L.cs_get_state.argtypes = [vp, C.c_int, C.c_int, F, F, F]
    L.cs_set_state.argtypes = [vp, C.c_int, C.c_int, F, F, F]
    L.cs_frame_begin.argtypes = [vp, C.c_int]
    L.cs_frame_end.argtypes = [vp, C.c_int, C.c_int, F, F, C.POINTER(C.c_int)]
    L.cs_get_geometry.argtypes = [vp, C.c_int, C.c_int, F, F, F]
    L.cs_material_table.argtypes = [C.POINTER(MaterialIn), F, F, F]
    L.cs_lbvh_pairs.argtypes = [vp, C.c_int, C.c_float, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]
    L.cs_find_proximities.argtypes = [vp, C.POINTER(ProximityRec), C.c_int, C.POINTER(C.c_int)]
    L.cs_assemble.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), F, C.c_int, C.POINTER(C.c_int), F]
    L.cs_solve.argtypes = [vp, C.c_int, F, C.POINTER(Stats)]
    L.cs_find_impacts.argtypes = [vp, C.c_float, C.POINTER(ImpactRec), C.c_int, C.POINTER(C.c_int)]
    L.cs_independent_impacts.argtypes = [vp, C.POINTER(ImpactRec), C.c_int, C.c_int, C.POINTER(ImpactRec),
                                         C.POINTER(C.c_int)]
    L.cs_solve_bsr_noalloc.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, C.c_float, C.c_int, C.POINTER(C.c_int),
                                       C.POINTER(C.c_float)]
    L.cs_impact_response.argtypes = [vp, C.POINTER(ImpactRec), C.c_int, C.c_int, C.c_float, C.POINTER(C.c_int)]
    L.cs_kernel_time.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int]
    L.cs_load_scene.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(Params), C.POINTER(vp), C.POINTER(SceneInfo)]
    L.cs_scene_error.restype = C.c_char_p
    L.cs_precondition.argtypes = [vp, C.c_int, F, F, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]
    L.cs_lpc_weight_count.restype = C.c_int
    L.cs_set_learned_weights.argtypes = [vp, F, C.c_int]
    L.cs_get_learned_weights.argtypes = [vp, F, C.c_int]
    L.cs_lpc_factor.argtypes = [vp, C.c_int, F, F, C.c_int, C.POINTER(C.c_int)]
    L.cs_lpc_train_step.argtypes = [vp, C.c_int, C.c_float, C.c_float, F, F]
    L.cs_lpc_train_step_smooth.argtypes = [vp, C.c_int, C.c_float, C.c_float, C.c_float, C.c_uint, F, F]
    L.cs_lpc_smoother_omega.argtypes = [vp, C.c_int, F]
    L.cs_batch_tuning.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    TP = C.POINTER(Transform)
    L.cs_add_motion.argtypes = [vp, F, TP, C.c_int, C.POINTER(C.c_int)]
    L.cs_motion_eval.argtypes = [F, TP, C.c_int, C.c_float, TP]
    L.cs_set_obstacle_motion.argtypes = [vp, C.c_int, C.c_int]
    L.cs_add_moving_handles.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int]
    L.cs_set_step_count.argtypes = [vp, C.c_int]
    L.cs_get_step_count.argtypes = [vp, C.POINTER(C.c_int)]
    L.cs_scene_motion.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_float, TP]
    IP = C.POINTER(C.c_int)
    L.cs_get_topology.argtypes = [vp, C.c_int, C.c_int, IP, IP, F, IP]
    L.cs_write_config.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_char_p), C.c_int]
    RP = C.POINTER(Remeshing)
    UP = C.POINTER(C.c_ubyte)
    L.cs_default_remeshing.argtypes = [RP]
    L.cs_default_remeshing.restype = None
    L.cs_remesh_mesh.argtypes = [RP, C.c_int, F, F, F, UP, C.c_int, F, IP, C.c_int, IP, C.c_int, IP, F, IP,
                                 C.POINTER(vp)]
    L.cs_remesh_result_counts.argtypes = [vp, IP, IP, IP, IP]
    L.cs_remesh_result_get.argtypes = [vp, F, F, F, UP, F, IP, IP, IP, F, IP]
    L.cs_set_remeshing.argtypes = [vp, C.c_int, RP]
    L.cs_remesh.argtypes = [vp]
    L.cs_remesh_sizing.argtypes = [vp, C.c_int, RP, F, F]
    L.cs_remesh_times.argtypes = [vp, F, F, F, F, IP]
    L.cs_separation_step.argtypes = [vp, IP, F, F]
    L.cs_remesh_result_free.argtypes = [vp]
    L.cs_remesh_result_free.restype = None
    _lib = L
    return L


def remesh_mesh(r, x, x0, v, preserve, uv, vert_node, face_verts, edge_nodes, face_sizing, flip_count=0):
    """The remeshing operators on one mesh in host arrays (cs_remesh_mesh; needs no device).
    Returns (dict of x, x0, v, preserve, uv, vnode, fv, en, sizing, node_origin; flip_count)."""
    L = load_library()
    f32 = lambda a, k: np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, k))
    i32 = lambda a, k: np.ascontiguousarray(np.asarray(a, np.int32).reshape(-1, k))
    x, x0, v, uv = f32(x, 3), f32(x0, 3), f32(v, 3), f32(uv, 2)
    pres = np.ascontiguousarray(np.asarray(preserve, np.uint8))
    vnode, fv, en, fs = i32(vert_node, 1).ravel(), i32(face_verts, 3), i32(edge_nodes, 2), f32(face_sizing, 4)
    fc = C.c_int(flip_count)
    res = C.c_void_p()
    up = lambda a: a.ctypes.data_as(C.POINTER(C.c_ubyte))
    st = L.cs_remesh_mesh(C.byref(r), len(x), fptr(x), fptr(x0), fptr(v), up(pres), len(uv), fptr(uv), iptr(vnode),
                          len(fv), iptr(fv), len(en), iptr(en), fptr(fs), C.byref(fc), C.byref(res))
    if st != CS_OK:
        raise ClothError(st, "cs_remesh_mesh")
    try:
        n = [C.c_int() for _ in range(4)]
        L.cs_remesh_result_counts(res, *[C.byref(k) for k in n])
        nn, nv, nf, ne = (k.value for k in n)
        o = dict(x=np.zeros((nn, 3), np.float32), x0=np.zeros((nn, 3), np.float32), v=np.zeros((nn, 3), np.float32),
                 preserve=np.zeros(nn, np.uint8), uv=np.zeros((nv, 2), np.float32), vnode=np.zeros(nv, np.int32),
                 fv=np.zeros((nf, 3), np.int32), en=np.zeros((ne, 2), np.int32), sizing=np.zeros((nv, 4), np.float32),
                 node_origin=np.zeros(nn, np.int32))
        L.cs_remesh_result_get(res, fptr(o["x"]), fptr(o["x0"]), fptr(o["v"]), up(o["preserve"]), fptr(o["uv"]),
                               iptr(o["vnode"]), iptr(o["fv"]), iptr(o["en"]), fptr(o["sizing"]),
                               iptr(o["node_origin"]))
    finally:
        L.cs_remesh_result_free(res)
    o["preserve"] = o["preserve"].astype(bool)
    return o, fc.value


def _keys(times, keys):
    t = np.ascontiguousarray(times, dtype=np.float32)
    k = (Transform * len(keys))(*[x if isinstance(x, Transform) else
                                  transform_struct(x[0], x[1:4], x[4:8]) for x in keys])
    return t, k


def motion_eval(times, keys, time):
    """Motion::computeTransformation of keyframes (times, keys) at `time` (cs_motion_eval; host only).
    keys: Transform structs or 8 floats (scale, translate xyz, rotate s x y z).  Returns a Transform."""
    L = load_library()
    t, k = _keys(times, keys)
    out = Transform()
    st = L.cs_motion_eval(fptr(t), k, len(t), float(time), C.byref(out))
    if st != CS_OK:
        raise ClothError(st, "cs_motion_eval")
    return out


def scene_motion(config_path, motion_id, time, root_dir=None):
    """Motion `motion_id` of a reference configuration file at `time` (cs_scene_motion; host only)."""
    L = load_library()
    out = Transform()
    st = L.cs_scene_motion(config_path.encode(), (root_dir or "").encode(), motion_id, float(time), C.byref(out))
    if st != CS_OK:
        raise ClothError(st, "cs_scene_motion", L.cs_scene_error().decode())
    return out


def write_config(config_path, out_path, cloth_meshes):
    """Simulator::save's configuration (cs_write_config): cloth i's "mesh" = cloth_meshes[i]."""
    L = load_library()
    arr = (C.c_char_p * len(cloth_meshes))(*[m.encode() for m in cloth_meshes])
    st = L.cs_write_config(config_path.encode(), out_path.encode(), arr, len(cloth_meshes))
    if st != CS_OK:
        raise ClothError(st, "cs_write_config", L.cs_scene_error().decode())


def scene_info(config_path, root_dir=None):
    """Parse a reference configuration file (and its meshes and materials) without a device."""
    L = load_library()
    info = SceneInfo()
    st = L.cs_load_scene(config_path.encode(), (root_dir or "").encode(), None, None, C.byref(info))
    if st != CS_OK:
        raise ClothError(st, "cs_load_scene", L.cs_scene_error().decode())
    return info


def default_learned_weights():
    """The packaged trained weights of the learned preconditioner (float32 vector)."""
    w = np.load(LPC_WEIGHTS, allow_pickle=False)
    return np.ascontiguousarray(w, dtype=np.float32)


def material_table(mat):
    L = load_library()
    table = np.zeros(30 * 30 * 30 * 4, dtype=np.float32)
    bend = np.zeros(15, dtype=np.float32)
    dens = C.c_float()
    st = L.cs_material_table(C.byref(mat), fptr(table), C.byref(dens), fptr(bend))
    if st != CS_OK:
        raise ClothError(st, "cs_material_table")
    return table.reshape(30, 30, 30, 4), dens.value, bend.reshape(3, 5)


class Simulator:
    """One cs_ctx: cloths, static obstacles and handles on one gfx950 device."""

    def __init__(self, spec=None, params=None):
        self.L = load_library()
        self.params = params if params is not None else spec.params
        h = C.c_void_p()
        st = self.L.cs_create(C.byref(self.params), C.byref(h))
        if st != CS_OK:
            raise ClothError(st, "cs_create")
        self.h = h
        self._keep = []
        self.n_cloths = 0
        if spec is not None:
            cids = [self.add_cloth(c.mesh, c.material) for c in spec.cloths]
            for times, keys in getattr(spec, "motions", []):
                self.add_motion(times, keys)
            for cid, c in zip(cids, spec.cloths):
                if c.handles:
                    self.add_handles(cid, c.handles, getattr(c, "handle_motion", -1))
                if getattr(c, "remeshing", None) is not None:
                    self.set_remeshing(cid, c.remeshing)
            for i, o in enumerate(spec.obstacles):
                oid = self.add_obstacle(o)
                m = spec.obstacle_motions[i] if getattr(spec, "obstacle_motions", None) else -1
                if m >= 0:
                    self.set_obstacle_motion(oid, m)
        if self.params.precond in (CS_PRECOND_LEARNED, CS_PRECOND_AMG_LEARNED) and os.path.exists(LPC_WEIGHTS):
            self.set_learned_weights(default_learned_weights())

    @classmethod
    def from_config(cls, config_path, root_dir=None, params=None):
        """A simulator for a reference configuration file (cs_load_scene); `params` supplies
        the PCG settings and flags the file does not hold."""
        self = cls.__new__(cls)
        self.L = load_library()
        h = C.c_void_p()
        info = SceneInfo()
        st = self.L.cs_load_scene(config_path.encode(), (root_dir or "").encode(),
                                  C.byref(params) if params is not None else None, C.byref(h), C.byref(info))
        if st != CS_OK:
            raise ClothError(st, "cs_load_scene", self.L.cs_scene_error().decode())
        self.h = h
        self.info = info
        self.params = params
        self._keep = []
        self.n_cloths = info.n_cloths
        return self

    def _check(self, st, what, ok=(CS_OK,)):
        if st not in ok:
            msg = self.L.cs_last_error(self.h)
            raise ClothError(st, what, msg.decode() if msg else "")
        return st

    def close(self):
        if getattr(self, "h", None):
            self.L.cs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        self.params = p
        self._check(self.L.cs_set_params(self.h, C.byref(p)), "cs_set_params")

    def add_cloth(self, mesh, mat):
        cid = C.c_int()
        m = mesh.c_struct()
        self._keep.append((mesh, mat))
        self._check(self.L.cs_add_cloth(self.h, C.byref(m), C.byref(mat), C.byref(cid)), "cs_add_cloth")
        self.n_cloths += 1
        return cid.value

    def add_obstacle(self, mesh):
        oid = C.c_int()
        m = mesh.c_struct()
        self._keep.append(mesh)
        self._check(self.L.cs_add_obstacle(self.h, C.byref(m), C.byref(oid)), "cs_add_obstacle")
        return oid.value

    def add_handles(self, cloth, nodes, motion=-1):
        a = np.ascontiguousarray(nodes, dtype=np.int32)
        if motion < 0:
            self._check(self.L.cs_add_handles(self.h, cloth, iptr(a), len(a)), "cs_add_handles")
        else:
            self._check(self.L.cs_add_moving_handles(self.h, cloth, iptr(a), len(a), motion), "cs_add_moving_handles")

    def add_motion(self, times, keys):
        """Keyframes (times, keys) as a motion of this context; returns its index."""
        t, k = _keys(times, keys)
        mid = C.c_int()
        self._check(self.L.cs_add_motion(self.h, fptr(t), k, len(t), C.byref(mid)), "cs_add_motion")
        return mid.value

    def set_obstacle_motion(self, obstacle, motion):
        self._check(self.L.cs_set_obstacle_motion(self.h, obstacle, motion), "cs_set_obstacle_motion")

    def topology(self, kind=CS_MESH_CLOTH, idx=0):
        """(face node ids (nf, 3), face uv-vertex ids (nf, 3), uv (nv, 2)), mesh-local (cs_get_topology)."""
        _, nf, _ = self.counts(kind, idx)
        nv = C.c_int()
        self._check(self.L.cs_get_topology(self.h, kind, idx, None, None, None, C.byref(nv)), "cs_get_topology")
        fn = np.zeros((nf, 3), dtype=np.int32)
        fv = np.zeros((nf, 3), dtype=np.int32)
        uv = np.zeros((nv.value, 2), dtype=np.float32)
        self._check(self.L.cs_get_topology(self.h, kind, idx, iptr(fn), iptr(fv), fptr(uv), C.byref(nv)),
                    "cs_get_topology")
        return fn, fv, uv

    def step_count(self):
        n = C.c_int()
        self._check(self.L.cs_get_step_count(self.h, C.byref(n)), "cs_get_step_count")
        return n.value

    def set_step_count(self, n):
        self._check(self.L.cs_set_step_count(self.h, n), "cs_set_step_count")

    def physics_step(self):
        st = Stats()
        s = self._check(self.L.cs_physics_step(self.h, C.byref(st)), "cs_physics_step", (CS_OK, CS_E_NOT_CONVERGED))
        return s, st

    def collision_step(self):
        st = Stats()
        s = self.L.cs_collision_step(self.h, C.byref(st))
        return s, st

    def step(self):
        st = Stats()
        s = self.L.cs_step(self.h, C.byref(st))
        return s, st

    def run(self, steps):
        st = Stats()
        s = self.L.cs_run(self.h, steps, C.byref(st))
        return s, st

    def counts(self, kind=CS_MESH_CLOTH, idx=0):
        nn, nf, ne = C.c_int(), C.c_int(), C.c_int()
        self._check(self.L.cs_get_counts(self.h, kind, idx, C.byref(nn), C.byref(nf), C.byref(ne)), "cs_get_counts")
        return nn.value, nf.value, ne.value

    def state(self, kind=CS_MESH_CLOTH, idx=0):
        nn = self.counts(kind, idx)[0]
        x, x0, v = (np.zeros((nn, 3), dtype=np.float32) for _ in range(3))
        self._check(self.L.cs_get_state(self.h, kind, idx, fptr(x), fptr(x0), fptr(v)), "cs_get_state")
        return x, x0, v

    def frame_begin(self, slot):
        """Snapshot every cloth's positions and velocities into download slot 0 or 1 and start
        their download; returns at once (cs_frame_begin)."""
        self._check(self.L.cs_frame_begin(self.h, slot), "cs_frame_begin")

    def frame_end(self, slot, cloth=0):
        """(x, v) of `cloth` in slot's snapshot, after waiting for its download (cs_frame_end)."""
        nul = C.POINTER(C.c_float)()
        n = C.c_int()
        self._check(self.L.cs_frame_end(self.h, slot, cloth, nul, nul, C.byref(n)), "cs_frame_end")
        x, v = np.zeros((n.value, 3), dtype=np.float32), np.zeros((n.value, 3), dtype=np.float32)
        self._check(self.L.cs_frame_end(self.h, slot, cloth, fptr(x), fptr(v), C.byref(n)), "cs_frame_end")
        return x, v

    def set_state(self, kind, idx, x=None, x0=None, v=None):
        conv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32)
        x, x0, v = conv(x), conv(x0), conv(v)
        nul = C.POINTER(C.c_float)()
        self._check(self.L.cs_set_state(self.h, kind, idx, fptr(x) if x is not None else nul,
                                        fptr(x0) if x0 is not None else nul, fptr(v) if v is not None else nul),
                    "cs_set_state")

    def geometry(self, kind=CS_MESH_CLOTH, idx=0):
        nn = self.counts(kind, idx)[0]
        n = np.zeros((nn, 3), dtype=np.float32)
        a = np.zeros(nn, dtype=np.float32)
        m = np.zeros(nn, dtype=np.float32)
        self._check(self.L.cs_get_geometry(self.h, kind, idx, fptr(n), fptr(a), fptr(m)), "cs_get_geometry")
        return n, a, m

    def lbvh_pairs(self, ccd, thickness, refit=False):
        """Candidate face pairs (cs_lbvh_pairs).  refit: refit the trees of the previous call (same
        ccd) to the current positions instead of building new ones (both calls below refit: a second
        refit on unchanged positions gives the same boxes)."""
        n = C.c_int()
        mode = int(bool(ccd)) | (2 if refit else 0)
        self._check(self.L.cs_lbvh_pairs(self.h, mode, thickness, None, 0, C.byref(n)), "cs_lbvh_pairs")
        pairs = np.zeros((max(n.value, 1), 2), dtype=np.int32)
        self._check(self.L.cs_lbvh_pairs(self.h, mode, thickness, iptr(pairs), n.value, C.byref(n)),
                    "cs_lbvh_pairs")
        return pairs[:n.value]

    def proximities(self, cap=1 << 20):
        out = np.zeros(cap, dtype=PROX_DTYPE)
        n = C.c_int()
        self._check(self.L.cs_find_proximities(self.h, out.ctypes.data_as(C.POINTER(ProximityRec)), cap, C.byref(n)),
                    "cs_find_proximities")
        return out[:n.value].copy()

    def assemble(self, cloth=0):
        nnzb = C.c_int()
        self._check(self.L.cs_assemble(self.h, cloth, None, None, None, 0, C.byref(nnzb), None), "cs_assemble")
        nn = self.counts(CS_MESH_CLOTH, cloth)[0]
        rp = np.zeros(nn + 1, dtype=np.int32)
        ci = np.zeros(nnzb.value, dtype=np.int32)
        vals = np.zeros((nnzb.value, 3, 3), dtype=np.float32)
        b = np.zeros(3 * nn, dtype=np.float32)
        self._check(self.L.cs_assemble(self.h, cloth, iptr(rp), iptr(ci), fptr(vals), nnzb.value, C.byref(nnzb),
                                       fptr(b)), "cs_assemble")
        return rp, ci, vals, b

    def solve(self, cloth=0):
        nn = self.counts(CS_MESH_CLOTH, cloth)[0]
        dv = np.zeros(3 * nn, dtype=np.float32)
        st = Stats()
        s = self._check(self.L.cs_solve(self.h, cloth, fptr(dv), C.byref(st)), "cs_solve", (CS_OK, CS_E_NOT_CONVERGED))
        return dv, st

    def impacts(self, thickness, cap=1 << 20):
        out = np.zeros(cap, dtype=IMPACT_DTYPE)
        n = C.c_int()
        self._check(self.L.cs_find_impacts(self.h, thickness, out.ctypes.data_as(C.POINTER(ImpactRec)), cap,
                                           C.byref(n)), "cs_find_impacts")
        return out[:n.value].copy()

    def precondition(self, cloth=0, r=None):
        """Level sizes of the multigrid hierarchy and, with r, z = one V-cycle applied to r."""
        sizes = (C.c_int * 64)()
        nl = C.c_int()
        if r is None:
            self._check(self.L.cs_precondition(self.h, cloth, None, None, sizes, 64, C.byref(nl)), "cs_precondition")
            return [sizes[i] for i in range(nl.value)], None
        rr = np.ascontiguousarray(r, dtype=np.float32).ravel()
        z = np.zeros_like(rr)
        self._check(self.L.cs_precondition(self.h, cloth, fptr(rr), fptr(z), sizes, 64, C.byref(nl)),
                    "cs_precondition")
        return [sizes[i] for i in range(nl.value)], z.reshape(-1, 3)

    def set_learned_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32).ravel()
        self._check(self.L.cs_set_learned_weights(self.h, fptr(w), len(w)), "cs_set_learned_weights")

    def learned_weights(self):
        w = np.zeros(self.L.cs_lpc_weight_count(), dtype=np.float32)
        self._check(self.L.cs_get_learned_weights(self.h, fptr(w), len(w)), "cs_get_learned_weights")
        return w

    def lpc_factor(self, cloth=0):
        """G of the last assembled system: diagonal blocks (n, 3, 3) and lower blocks (n_lower, 3, 3)
        in static-pattern order (rows ascending, columns ascending, j < i)."""
        nl = C.c_int()
        self._check(self.L.cs_lpc_factor(self.h, cloth, None, None, 0, C.byref(nl)), "cs_lpc_factor")
        n = self.counts(CS_MESH_CLOTH, cloth)[0]
        diag = np.zeros((n, 3, 3), dtype=np.float32)
        lower = np.zeros((max(nl.value, 1), 3, 3), dtype=np.float32)
        self._check(self.L.cs_lpc_factor(self.h, cloth, fptr(diag), fptr(lower), nl.value, C.byref(nl)),
                    "cs_lpc_factor")
        return diag, lower[:nl.value]

    def lpc_train_step(self, cloth=0, lr=0.0, grad=False, sol_weight=0.0, smooth_weight=0.0, seed=0):
        """Kaporin loss of the network's factor on the last assembled system (plus sol_weight times
        the solution term once the system is solved, plus smooth_weight times the smoothing term on
        the probe of `seed`), its gradient (when grad) and one Adam step when lr > 0."""
        loss = C.c_float()
        g = np.zeros(self.L.cs_lpc_weight_count(), dtype=np.float32) if grad else None
        gp = fptr(g) if grad else None
        if smooth_weight > 0.0:
            self._check(self.L.cs_lpc_train_step_smooth(self.h, cloth, lr, sol_weight, smooth_weight, seed,
                                                        C.byref(loss), gp), "cs_lpc_train_step_smooth")
        else:
            self._check(self.L.cs_lpc_train_step(self.h, cloth, lr, sol_weight, C.byref(loss), gp),
                        "cs_lpc_train_step")
        return (loss.value, g) if grad else loss.value

    def set_remeshing(self, cloth, r):
        """Remeshing parameters of one cloth (a Remeshing struct, or None to clear them)."""
        self._check(self.L.cs_set_remeshing(self.h, cloth, C.byref(r) if r is not None else None),
                    "cs_set_remeshing")

    def remesh(self):
        """The remeshing step of every cloth with remeshing parameters (cs_remesh)."""
        self._check(self.L.cs_remesh(self.h), "cs_remesh")

    def remesh_times(self):
        """(sizing, operators, rebuild, separation) milliseconds and the separation step's solve count
        of the last remesh (cs_remesh_times)."""
        t = np.zeros(4, np.float32)
        n = C.c_int()
        self._check(self.L.cs_remesh_times(self.h, fptr(t[0:]), fptr(t[1:]), fptr(t[2:]), fptr(t[3:]), C.byref(n)),
                    "cs_remesh_times")
        return tuple(float(v) for v in t) + (n.value,)

    def separation_step(self, backups):
        """The separation step on the current state; backups: per cloth (uv (F, 3, 2), x (F, 3, 3)),
        its faces before remeshing.  Returns the status (0, or CS_E_COLLISION_FAILED)."""
        if len(backups) != self.n_cloths:
            raise ValueError("separation_step needs one backup per cloth (%d, got %d)" % (self.n_cloths, len(backups)))
        counts = np.array([len(u) for u, _ in backups], np.int32)
        uv = np.ascontiguousarray(np.concatenate([np.asarray(u, np.float32).reshape(-1, 6) for u, _ in backups]))
        xx = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float32).reshape(-1, 9) for _, x in backups]))
        st = self.L.cs_separation_step(self.h, iptr(counts), fptr(uv), fptr(xx))
        self._check(st, "cs_separation_step", ok=(CS_OK, CS_E_COLLISION_FAILED))
        return st

    def remesh_sizing(self, cloth, r):
        """Face sizing (nf x 4) and node planes (nn x 8) the remeshing step would use now."""
        nn, nf, _ = self.counts(CS_MESH_CLOTH, cloth)
        fs = np.zeros((nf, 4), np.float32)
        pl = np.zeros((nn, 8), np.float32)
        self._check(self.L.cs_remesh_sizing(self.h, cloth, C.byref(r), fptr(fs), fptr(pl)), "cs_remesh_sizing")
        return fs, pl

    def smoother_omega(self, cloth=0):
        """Damping of the learned smoother of the last assembled system (CS_PRECOND_AMG_LEARNED)."""
        w = np.zeros(1, dtype=np.float32)
        self._check(self.L.cs_lpc_smoother_omega(self.h, cloth, fptr(w)), "cs_lpc_smoother_omega")
        return float(w[0])

    def batch_tuning(self):
        """(rows per workgroup, iterations per poll, solves timed) of the batched instance solve."""
        r, c, t = C.c_int(), C.c_int(), C.c_int()
        self._check(self.L.cs_batch_tuning(self.h, C.byref(r), C.byref(c), C.byref(t)), "cs_batch_tuning")
        return r.value, c.value, t.value

    def kernel_time(self, kernel_id, reset=False):
        """(summed ms, sampled launches) of a hot kernel; needs CS_FLAG_TIME_KERNELS in params.flags."""
        tot = C.c_double()
        cnt = C.c_int()
        self._check(self.L.cs_kernel_time(self.h, kernel_id, C.byref(tot), C.byref(cnt), int(reset)), "cs_kernel_time")
        return tot.value, cnt.value

    def impact_response(self, imps, deform, obstacle_mass=1e3):
        """The impact response (cs_impact_response) on the given impacts from the current state;
        returns its iteration count."""
        a = np.ascontiguousarray(imps, dtype=IMPACT_DTYPE)
        it = C.c_int()
        self._check(self.L.cs_impact_response(self.h, a.ctypes.data_as(C.POINTER(ImpactRec)), len(a), deform,
                                              obstacle_mass, C.byref(it)), "cs_impact_response")
        return it.value

    def independent_impacts(self, imps, deform):
        a = np.ascontiguousarray(imps, dtype=IMPACT_DTYPE)
        out = np.zeros(max(len(a), 1), dtype=IMPACT_DTYPE)
        n = C.c_int()
        self._check(self.L.cs_independent_impacts(self.h, a.ctypes.data_as(C.POINTER(ImpactRec)), len(a), deform,
                                                  out.ctypes.data_as(C.POINTER(ImpactRec)), C.byref(n)),
                    "cs_independent_impacts")
        return out[:n.value].copy()
