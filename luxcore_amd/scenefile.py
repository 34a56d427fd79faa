"""Scene files for the intersection device: PLY meshes and LuxCore .scn/.cfg
geometry, read by the native reader (libluxrays_ingest.so, include/lrh_ingest.h)
and handed to the device as it would receive a LuxRays DataSet.

    sf = load_scene("scenes/bigmonkey/bigmonkey-instances.scn")
    scene = scene_on_device(dev, sf)       # AUTO: MBVH if instances or motion, else BVH

Reference: Scene::ParseShapes / ParseObjects (src/slg/scene/parseshapes.cpp,
parseobjects.cpp), ExtTriangleMesh::LoadPly
(src/luxrays/core/exttrianglemeshfile.cpp:259-452), the DataSet AUTO policy
(src/luxrays/devices/oclintersectiondevice.cpp:42-56).
"""
import ctypes
import os

import numpy as np

from .device import LrhMesh, LrhObject
from .scenes import InstancedScene, Mesh

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "host", "libluxrays_ingest.so")
LRH_OK = 0
LRH_EIO = -7


class IngestError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class SceneFileInfo(ctypes.Structure):
    _fields_ = [("nMeshes", ctypes.c_uint), ("nObjects", ctypes.c_uint), ("nKeys", ctypes.c_uint),
                ("twoLevel", ctypes.c_uint), ("nVerts", ctypes.c_ulonglong), ("nTris", ctypes.c_ulonglong),
                ("nWorldTris", ctypes.c_ulonglong)]


class SceneFileCamera(ctypes.Structure):
    _fields_ = [("orig", ctypes.c_float * 3), ("target", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("fieldOfView", ctypes.c_float)]


_vp = ctypes.c_void_p
_SIGNATURES = {
    "lrh_ply_load": ([ctypes.c_char_p, ctypes.POINTER(_vp)], ctypes.c_int),
    "lrh_scenefile_load": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_uint,
                            ctypes.POINTER(_vp)], ctypes.c_int),
    "lrh_scenefile_parse": ([ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.POINTER(_vp)],
                            ctypes.c_int),
    "lrh_scenefile_get_info": ([_vp, ctypes.POINTER(SceneFileInfo)], ctypes.c_int),
    "lrh_scenefile_meshes": ([_vp, ctypes.POINTER(LrhMesh), ctypes.c_uint], ctypes.c_int),
    "lrh_scenefile_objects": ([_vp, ctypes.POINTER(LrhObject), ctypes.c_uint], ctypes.c_int),
    "lrh_scenefile_object_name": ([_vp, ctypes.c_uint, ctypes.POINTER(ctypes.c_char_p),
                                   ctypes.POINTER(ctypes.c_char_p)], ctypes.c_int),
    "lrh_scenefile_get_camera": ([_vp, ctypes.POINTER(SceneFileCamera)], ctypes.c_int),
    "lrh_scenefile_object_material": ([_vp, ctypes.c_uint, ctypes.POINTER(ctypes.c_float),
                                       ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "lrh_scenefile_destroy": ([_vp], ctypes.c_int),
    "lrh_ingest_last_error": ([], ctypes.c_char_p),
}
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(make -C luxcore_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != LRH_OK:
        raise IngestError(rc, lib().lrh_ingest_last_error().decode(errors="replace"))


class SceneFile:
    """A loaded scene: unique meshes, objects in DataSet order, camera.

    objects: dicts {name, mesh_name, mesh, kind (0 plain, 1 instance, 2
    motion), m (local->world, kind 1), times / keys (global->local, kind 2)}.
    """

    def __init__(self, handle):
        self._h = handle
        try:
            info = SceneFileInfo()
            _check(lib().lrh_scenefile_get_info(handle, ctypes.byref(info)))
            self.info = {f: getattr(info, f) for f, _ in SceneFileInfo._fields_}
            marr = (LrhMesh * max(1, info.nMeshes))()
            _check(lib().lrh_scenefile_meshes(handle, marr, info.nMeshes))
            self.meshes = []
            for i in range(info.nMeshes):
                m = marr[i]
                v = np.ctypeslib.as_array(m.verts, shape=(m.nVerts * 3,)).reshape(-1, 3).copy() if m.nVerts \
                    else np.zeros((0, 3), np.float32)
                t = np.ctypeslib.as_array(m.tris, shape=(m.nTris * 3,)).reshape(-1, 3).copy() if m.nTris \
                    else np.zeros((0, 3), np.uint32)
                self.meshes.append(Mesh(v.astype(np.float32), t.astype(np.uint32)))
            oarr = (LrhObject * max(1, info.nObjects))()
            _check(lib().lrh_scenefile_objects(handle, oarr, info.nObjects))
            self.objects = []
            for i in range(info.nObjects):
                o = oarr[i]
                name, mname = ctypes.c_char_p(), ctypes.c_char_p()
                _check(lib().lrh_scenefile_object_name(handle, i, ctypes.byref(name), ctypes.byref(mname)))
                d = {"name": name.value.decode(), "mesh_name": mname.value.decode(), "mesh": int(o.mesh),
                     "kind": int(o.kind), "m": np.eye(4, dtype=np.float32), "times": [], "keys": []}
                if o.kind == 1:
                    d["m"] = np.ctypeslib.as_array(o.m, shape=(16,)).reshape(4, 4).copy()
                elif o.kind == 2:
                    d["times"] = list(np.ctypeslib.as_array(o.keyTimes, shape=(o.nKeys,)).copy())
                    km = np.ctypeslib.as_array(o.keyMats, shape=(o.nKeys * 16,)).reshape(-1, 4, 4).copy()
                    d["keys"] = list(km)
                kd, le = (ctypes.c_float * 3)(), (ctypes.c_float * 3)()
                _check(lib().lrh_scenefile_object_material(handle, i, kd, le))
                d["kd"], d["le"] = np.array(kd, np.float32), np.array(le, np.float32)
                self.objects.append(d)
            cam = SceneFileCamera()
            _check(lib().lrh_scenefile_get_camera(handle, ctypes.byref(cam)))
            self.camera = (tuple(cam.orig), tuple(cam.target), tuple(cam.up), float(cam.fieldOfView))
        finally:
            lib().lrh_scenefile_destroy(handle)
            self._h = None

    @property
    def two_level(self):
        return bool(self.info["twoLevel"])

    def instanced(self):
        """InstancedScene over the unique meshes (for Device.scene_mbvh)."""
        n = len(self.objects)
        obj_mat = np.zeros((n, 4, 4), np.float32)
        key_off = np.zeros(n + 1, np.uint32)
        times, mats = [], []
        for i, o in enumerate(self.objects):
            if o["kind"] == 1:
                obj_mat[i] = o["m"]
            key_off[i + 1] = key_off[i] + len(o["times"])
            times += list(o["times"])
            mats += list(o["keys"])
        km = np.asarray(mats, np.float32).reshape(-1, 4, 4) if mats else np.zeros((0, 4, 4), np.float32)
        return InstancedScene(list(self.meshes), np.asarray([o["mesh"] for o in self.objects], np.uint32),
                              np.asarray([o["kind"] for o in self.objects], np.uint32), obj_mat, key_off,
                              np.asarray(times, np.float32), km)

    def world_bounds(self):
        """Axis-aligned world box of every object (instances through their
        matrix, motion objects at each key), in float64: for placing rays."""
        lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
        for o in self.objects:
            v = self.meshes[o["mesh"]].verts.astype(np.float64)
            if v.size == 0:
                continue
            if o["kind"] == 0:
                mats = [np.eye(4)]
            elif o["kind"] == 1:
                mats = [np.asarray(o["m"], np.float64)]
            else:
                mats = [np.linalg.inv(np.asarray(k, np.float64)) for k in o["keys"]]
            for M in mats:
                w = v @ M[:3, :3].T + M[:3, 3]
                lo, hi = np.minimum(lo, w.min(0)), np.maximum(hi, w.max(0))
        if not np.all(np.isfinite(lo)):
            return np.zeros(3), np.ones(3)
        return lo, hi

    def materials(self):
        """lrh_material records per DataSet mesh (= object), for lrh_shade."""
        from .device import MATERIAL_DTYPE
        out = np.zeros(len(self.objects), MATERIAL_DTYPE)
        for i, o in enumerate(self.objects):
            out[i]["kd"], out[i]["le"] = o["kd"], o["le"]
        return out

    def light_triangles(self):
        """lrh_light_tri records of every emitting plain object's triangles
        (world space = mesh space for plain objects)."""
        from .device import LIGHT_DTYPE
        parts = []
        for o in self.objects:
            if o["kind"] != 0 or not np.any(o["le"] > 0):
                continue
            m = self.meshes[o["mesh"]]
            v = m.verts[m.tris]                      # (T,3,3)
            rec = np.zeros(m.ntris, LIGHT_DTYPE)
            rec["p0"] = v[:, 0]
            rec["e1"] = v[:, 1] - v[:, 0]
            rec["e2"] = v[:, 2] - v[:, 0]
            rec["le"] = o["le"]
            parts.append(rec)
        return np.concatenate(parts) if parts else np.zeros(0, LIGHT_DTYPE)

    def dataset_meshes(self):
        """One mesh per object, in object order: the single-level DataSet."""
        return [self.meshes[o["mesh"]] for o in self.objects]


def load_ply(path):
    """One PLY file -> Mesh (ExtTriangleMesh::LoadPly)."""
    h = ctypes.c_void_p()
    _check(lib().lrh_ply_load(os.fsencode(path), ctypes.byref(h)))
    return SceneFile(h.value).meshes[0]


def load_scene(path, search_dirs=()):
    """A .scn file, or a .cfg naming one with scene.file."""
    h = ctypes.c_void_p()
    dirs = (ctypes.c_char_p * max(1, len(search_dirs)))(*[os.fsencode(d) for d in search_dirs])
    _check(lib().lrh_scenefile_load(os.fsencode(path), dirs, len(search_dirs), ctypes.byref(h)))
    return SceneFile(h.value)


def parse_scene(text, base_dir=""):
    """Scene properties given as text; relative file names resolve under base_dir."""
    h = ctypes.c_void_p()
    b = text.encode() if isinstance(text, str) else bytes(text)
    _check(lib().lrh_scenefile_parse(b, len(b), os.fsencode(base_dir), ctypes.byref(h)))
    return SceneFile(h.value)


def scene_on_device(dev, sf, **build_kw):
    """Build the scene on `dev` with the DataSet AUTO policy: two-level (MBVH)
    when any object is an instance or has motion, else single level (BVH)
    with mesh index = object index."""
    if sf.two_level:
        return dev.scene_mbvh(sf.instanced(), **build_kw)
    return dev.scene_bvh(sf.dataset_meshes(), **build_kw)
