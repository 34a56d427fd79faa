"""ctypes front end of liblrhip.so, the C-ABI of the HIP intersection device.

This is a thin host wrapper over include/lrhip.h for the tests, the bench and
the serving loop; it allocates device memory through the library itself
(lrh_alloc) and never falls back to anything else: if the shared library is
missing or no GPU is present, the calls raise.
"""
import ctypes
import os
import weakref

import numpy as np

from .raytypes import HIT_DTYPE, RAY_DTYPE
from .scenes import Mesh

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblrhip.so")

LRH_OK = 0
LRH_EINVAL, LRH_EHIP, LRH_ENOMEM, LRH_ERANGE, LRH_ENODEV, LRH_EDEVICE = -1, -2, -3, -4, -5, -6
LRH_H2D, LRH_D2H, LRH_D2D = 1, 2, 3
LRH_QUERY_CLOSEST, LRH_QUERY_ANY = 0, 1
TRACE_STATS = 11  # LRH_TRACE_STATS
ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "ERANGE", -5: "ENODEV", -6: "EDEVICE"}


class LrhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lrhip {ERRORS.get(code, code)}: {msg}")
        self.code = code


class LrhMesh(ctypes.Structure):
    _fields_ = [("verts", ctypes.POINTER(ctypes.c_float)), ("nVerts", ctypes.c_uint),
                ("tris", ctypes.POINTER(ctypes.c_uint)), ("nTris", ctypes.c_uint)]


class LrhObject(ctypes.Structure):
    _fields_ = [("mesh", ctypes.c_uint), ("kind", ctypes.c_uint), ("m", ctypes.POINTER(ctypes.c_float)),
                ("mInv", ctypes.POINTER(ctypes.c_float)), ("nKeys", ctypes.c_uint),
                ("keyTimes", ctypes.POINTER(ctypes.c_float)), ("keyMats", ctypes.POINTER(ctypes.c_float))]


class LrhBuildOpts(ctypes.Structure):
    _fields_ = [("layout", ctypes.c_uint), ("maxLeafSize", ctypes.c_uint), ("nThreads", ctypes.c_uint),
                ("travCost", ctypes.c_float), ("builder", ctypes.c_uint), ("splitBudget", ctypes.c_float)]


BUILD_HOST_SAH, BUILD_DEVICE_LBVH, BUILD_DEVICE_PLOC = 0, 1, 2


class LrhSceneInfo(ctypes.Structure):
    _fields_ = [("nodeCount", ctypes.c_ulonglong), ("nodeBytes", ctypes.c_ulonglong),
                ("primCount", ctypes.c_ulonglong), ("primBytes", ctypes.c_ulonglong),
                ("deviceBytes", ctypes.c_ulonglong), ("maxDepth", ctypes.c_uint), ("layout", ctypes.c_uint),
                ("buildSeconds", ctypes.c_double), ("uploadSeconds", ctypes.c_double)]


class LrhTraceOpts(ctypes.Structure):
    _fields_ = [("blockSize", ctypes.c_uint), ("ldsStack", ctypes.c_uint), ("blocksPerCU", ctypes.c_uint),
                ("refillThreshold", ctypes.c_uint),
                ("leafTris", ctypes.c_uint), ("gridPercent", ctypes.c_uint)]


class LrhCamera(ctypes.Structure):
    _fields_ = [("eye", ctypes.c_float * 3), ("look", ctypes.c_float * 3), ("up", ctypes.c_float * 3),
                ("vfovDeg", ctypes.c_float), ("mint", ctypes.c_float), ("width", ctypes.c_uint),
                ("height", ctypes.c_uint)]


class LrhShadeParams(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint), ("maxDepth", ctypes.c_uint), ("rrDepth", ctypes.c_uint),
                ("rrSurvive", ctypes.c_float), ("seed", ctypes.c_ulonglong), ("pixelBase", ctypes.c_uint),
                ("pad", ctypes.c_uint)]


class LrhShadeItem(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("rays", ctypes.c_void_p), ("hits", ctypes.c_void_p),
                ("shadowHits", ctypes.c_void_p), ("paths", ctypes.c_void_p), ("nextRays", ctypes.c_void_p),
                ("shadowRays", ctypes.c_void_p), ("params", LrhShadeParams)]


SHADE_BATCH_MAX = 16


# lrh_material, lrh_light_tri, lrh_path records
MATERIAL_DTYPE = np.dtype([("kd", "<f4", 3), ("le", "<f4", 3)])
LIGHT_DTYPE = np.dtype([("p0", "<f4", 3), ("e1", "<f4", 3), ("e2", "<f4", 3), ("le", "<f4", 3)])
PATH_DTYPE = np.dtype([("thr", "<f4", 3), ("rad", "<f4", 3), ("dl", "<f4", 3), ("depth", "<u4"),
                       ("alive", "<u4"), ("pad", "<u4")])
assert MATERIAL_DTYPE.itemsize == 24 and LIGHT_DTYPE.itemsize == 48 and PATH_DTYPE.itemsize == 48


class LrhServerConfig(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint), ("height", ctypes.c_uint), ("maxBounces", ctypes.c_uint),
                ("rrStart", ctypes.c_uint), ("rrSurvive", ctypes.c_float), ("hbmBudget", ctypes.c_ulonglong),
                ("maxRaysPerTick", ctypes.c_uint), ("cohorts", ctypes.c_uint), ("mode", ctypes.c_uint),
                ("extent", ctypes.c_float), ("light", ctypes.c_float * 3), ("opts", LrhTraceOpts)]


class LrhSessionSpec(ctypes.Structure):
    _fields_ = [("sid", ctypes.c_uint), ("scene", ctypes.c_uint), ("seed", ctypes.c_ulonglong),
                ("arrival", ctypes.c_double)]


class LrhServerStats(ctypes.Structure):
    _fields_ = [("sessionsDone", ctypes.c_ulonglong), ("sessionsQueued", ctypes.c_ulonglong),
                ("ticks", ctypes.c_ulonglong), ("wavefronts", ctypes.c_ulonglong), ("liveRays", ctypes.c_ulonglong),
                ("maxActive", ctypes.c_uint), ("sessionBuffersAllocated", ctypes.c_uint),
                ("busySeconds", ctypes.c_double), ("traceSeconds", ctypes.c_double),
                ("latencySamples", ctypes.c_ulonglong), ("p50Ms", ctypes.c_double), ("p99Ms", ctypes.c_double),
                ("maxMs", ctypes.c_double), ("queueP99Ms", ctypes.c_double)]


class LrhPlainBatchResult(ctypes.Structure):
    _fields_ = [("buffers", ctypes.c_uint), ("rays", ctypes.c_ulonglong), ("liveRays", ctypes.c_ulonglong),
                ("batchMs", ctypes.c_double), ("perBufferMs", ctypes.c_double)]


SERVE_SHADE, SERVE_BOUNCE = 0, 1


class LrhBatchItem(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("rays", ctypes.c_void_p), ("hits", ctypes.c_void_p),
                ("count", ctypes.c_uint), ("query", ctypes.c_uint)]


# every exported symbol of include/lrhip.h with its ctypes signature
_vp = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = {
    "lrh_last_error": ([], ctypes.c_char_p),
    "lrh_version": ([], ctypes.c_char_p),
    "lrh_device_count": ([ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "lrh_device_open": ([ctypes.c_int, _pp], ctypes.c_int),
    "lrh_device_close": ([_vp], ctypes.c_int),
    "lrh_device_info": ([_vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int),
                         ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "lrh_alloc": ([_vp, ctypes.c_size_t, _pp], ctypes.c_int),
    "lrh_free": ([_vp, _vp], ctypes.c_int),
    "lrh_memcpy": ([_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int, _vp], ctypes.c_int),
    "lrh_memcpy_async": ([_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int, _vp], ctypes.c_int),
    "lrh_host_alloc": ([_vp, ctypes.c_size_t, _pp], ctypes.c_int),
    "lrh_host_free": ([_vp, _vp], ctypes.c_int),
    "lrh_stream_create": ([_vp, _pp], ctypes.c_int),
    "lrh_stream_destroy": ([_vp, _vp], ctypes.c_int),
    "lrh_sync": ([_vp, _vp], ctypes.c_int),
    "lrh_used_memory": ([_vp], ctypes.c_size_t),
    "lrh_event_create": ([_vp, _pp], ctypes.c_int),
    "lrh_event_destroy": ([_vp], ctypes.c_int),
    "lrh_event_record": ([_vp, _vp], ctypes.c_int),
    "lrh_event_sync": ([_vp], ctypes.c_int),
    "lrh_event_elapsed": ([_vp, _vp, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "lrh_scene_create_bvh": ([_vp, ctypes.POINTER(LrhMesh), ctypes.c_uint, ctypes.POINTER(LrhBuildOpts), _pp],
                             ctypes.c_int),
    "lrh_scene_create_mbvh": ([_vp, ctypes.POINTER(LrhMesh), ctypes.c_uint, ctypes.POINTER(LrhObject),
                               ctypes.c_uint, ctypes.POINTER(LrhBuildOpts), _pp], ctypes.c_int),
    "lrh_scene_update": ([_vp, ctypes.POINTER(LrhObject), ctypes.c_uint], ctypes.c_int),
    "lrh_scene_refit": ([_vp, ctypes.POINTER(LrhObject), ctypes.c_uint], ctypes.c_int),
    "lrh_scene_object_boxes": ([_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_uint], ctypes.c_int),
    "lrh_scene_destroy": ([_vp], ctypes.c_int),
    "lrh_scene_get_info": ([_vp, ctypes.POINTER(LrhSceneInfo)], ctypes.c_int),
    "lrh_scene_download": ([_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t], ctypes.c_int),
    "lrh_trace": ([_vp, _vp, _vp, ctypes.c_uint, ctypes.c_uint, _vp], ctypes.c_int),
    "lrh_trace_ex": ([_vp, _vp, _vp, ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(LrhTraceOpts), _vp],
                     ctypes.c_int),
    "lrh_trace_stats": ([_vp, _vp, _vp, ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(LrhTraceOpts), _vp,
                         ctypes.POINTER(ctypes.c_ulonglong)], ctypes.c_int),
    "lrh_trace_noalloc": ([_vp, _vp, _vp, ctypes.c_uint, ctypes.c_uint, _vp], ctypes.c_int),
    "lrh_trace_batch": ([_vp, ctypes.POINTER(LrhBatchItem), ctypes.c_uint, ctypes.POINTER(LrhTraceOpts), _vp],
                        ctypes.c_int),
    "lrh_batch_create": ([_vp, ctypes.POINTER(LrhBatchItem), ctypes.c_uint, ctypes.POINTER(LrhTraceOpts), _pp],
                         ctypes.c_int),
    "lrh_batch_launch": ([_vp, _vp], ctypes.c_int),
    "lrh_batch_capture": ([_vp, _vp], ctypes.c_int),
    "lrh_batch_replay": ([_vp, _vp], ctypes.c_int),
    "lrh_batch_info": ([_vp, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_uint),
                        ctypes.POINTER(ctypes.c_uint)], ctypes.c_int),
    "lrh_batch_destroy": ([_vp], ctypes.c_int),
    "lrh_rays_camera": ([_vp, ctypes.POINTER(LrhCamera), _vp, _vp], ctypes.c_int),
    "lrh_rays_bounce": ([_vp, _vp, _vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_ulonglong, ctypes.c_uint,
                         ctypes.c_uint, ctypes.c_float, ctypes.POINTER(ctypes.c_float), _vp, _vp, _vp, _vp],
                        ctypes.c_int),
    "lrh_paths_init": ([_vp, _vp, ctypes.c_uint, _vp], ctypes.c_int),
    "lrh_shade": ([_vp, _vp, ctypes.c_uint, _vp, ctypes.c_uint, _vp, _vp, _vp, _vp, _vp, _vp,
                   ctypes.POINTER(LrhShadeParams), _vp, _vp], ctypes.c_int),
    "lrh_shade_batch": ([_vp, ctypes.c_uint, _vp, ctypes.c_uint, _vp, ctypes.c_uint, _vp, _vp], ctypes.c_int),
    "lrh_event_query": ([_vp], ctypes.c_int),
    "lrh_server_create": ([_vp, _pp, ctypes.c_uint, ctypes.POINTER(LrhServerConfig), _pp], ctypes.c_int),
    "lrh_server_submit": ([_vp, ctypes.POINTER(LrhSessionSpec), ctypes.c_uint], ctypes.c_int),
    "lrh_server_record": ([_vp, ctypes.c_uint], ctypes.c_int),
    "lrh_server_tick": ([_vp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "lrh_server_run": ([_vp, ctypes.c_double, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "lrh_server_get_stats": ([_vp, ctypes.POINTER(LrhServerStats)], ctypes.c_int),
    "lrh_server_latencies": ([_vp, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t,
                              ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "lrh_server_recorded": ([_vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, _vp, _vp,
                             ctypes.POINTER(ctypes.c_uint)], ctypes.c_int),
    "lrh_server_radiance": ([_vp, ctypes.c_uint, ctypes.POINTER(ctypes.c_float), ctypes.c_uint], ctypes.c_int),
    "lrh_server_capture": ([_vp, ctypes.c_ulonglong], ctypes.c_int),
    "lrh_server_plain_batch": ([_vp, ctypes.POINTER(LrhTraceOpts), ctypes.c_uint, ctypes.POINTER(LrhPlainBatchResult)],
                               ctypes.c_int),
    "lrh_server_destroy": ([_vp], ctypes.c_int),
    "lrh_autotune_build": ([_vp, ctypes.POINTER(ctypes.POINTER(LrhMesh)), ctypes.POINTER(ctypes.c_uint), ctypes.c_uint,
                            ctypes.POINTER(LrhBatchItem), ctypes.POINTER(ctypes.c_uint), ctypes.c_uint,
                            ctypes.POINTER(LrhBuildOpts), ctypes.c_uint, ctypes.POINTER(LrhTraceOpts), ctypes.c_uint,
                            ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(LrhTraceOpts),
                            ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "lrh_autotune_batch": ([_vp, ctypes.POINTER(LrhBatchItem), ctypes.c_uint, ctypes.POINTER(LrhTraceOpts),
                            ctypes.c_uint, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(LrhTraceOpts),
                            ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "lrh_device_stats": ([_vp, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_double)],
                         ctypes.c_int),
}

_lib = None


def lib():
    """Load liblrhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(make -C luxcore_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc):
    if rc != LRH_OK:
        raise LrhError(rc, lib().lrh_last_error().decode(errors="replace"))


def device_count():
    n = ctypes.c_int(0)
    rc = lib().lrh_device_count(ctypes.byref(n))
    return n.value if rc == LRH_OK else 0


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None


def _uptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint))


class DeviceBuffer:
    def __init__(self, dev, nbytes):
        self.dev = dev
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().lrh_alloc(dev.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def free(self):
        if self.ptr and self.dev.h:
            check(lib().lrh_free(self.dev.h, self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.dev.h:
                lib().lrh_free(self.dev.h, self.ptr)
        except Exception:
            pass
        self.ptr = None


class Event:
    def __init__(self, dev):
        p = ctypes.c_void_p()
        check(lib().lrh_event_create(dev.h, ctypes.byref(p)))
        self.h = p.value

    def record(self, stream=None):
        check(lib().lrh_event_record(self.h, stream))

    def sync(self):
        check(lib().lrh_event_sync(self.h))

    def query(self):
        rc = lib().lrh_event_query(self.h)
        if rc < 0:
            check(rc)
        return rc == 1

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        check(lib().lrh_event_elapsed(self.h, end.h, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if self.h:
                lib().lrh_event_destroy(self.h)
        except Exception:
            pass
        self.h = None


def _mesh_structs(meshes):
    keep = []
    arr = (LrhMesh * max(1, len(meshes)))()
    for i, m in enumerate(meshes):
        v = np.ascontiguousarray(m.verts, dtype=np.float32)
        t = np.ascontiguousarray(m.tris, dtype=np.uint32)
        keep += [v, t]
        arr[i].verts = _fptr(v)
        arr[i].nVerts = v.shape[0]
        arr[i].tris = _uptr(t)
        arr[i].nTris = t.shape[0]
    return arr, keep


def _object_structs(scene):
    """LrhObject array from an InstancedScene (kind 1: obj_mat local->world;
    kind 2: keys global->local)."""
    n = len(scene.obj_mesh)
    arr = (LrhObject * max(1, n))()
    keep = []
    mats = np.ascontiguousarray(scene.obj_mat, dtype=np.float32).reshape(-1, 16)
    kt = np.ascontiguousarray(scene.key_times, dtype=np.float32)
    km = np.ascontiguousarray(scene.key_mats, dtype=np.float32).reshape(-1, 16)
    keep += [mats, kt, km]
    for o in range(n):
        arr[o].mesh = int(scene.obj_mesh[o])
        arr[o].kind = int(scene.obj_kind[o])
        if arr[o].kind == 1:
            arr[o].m = ctypes.cast(mats.ctypes.data + 64 * o, ctypes.POINTER(ctypes.c_float))
        elif arr[o].kind == 2:
            k0, k1 = int(scene.key_off[o]), int(scene.key_off[o + 1])
            arr[o].nKeys = k1 - k0
            arr[o].keyTimes = ctypes.cast(kt.ctypes.data + 4 * k0, ctypes.POINTER(ctypes.c_float))
            arr[o].keyMats = ctypes.cast(km.ctypes.data + 64 * k0, ctypes.POINTER(ctypes.c_float))
    return arr, keep


def _build_opts(layout=0, max_leaf=4, threads=0, trav_cost=1.0, builder=BUILD_HOST_SAH, split_budget=0.0):
    o = LrhBuildOpts()
    o.builder = builder
    o.layout = layout
    o.maxLeafSize = max_leaf
    o.nThreads = threads
    o.travCost = trav_cost
    o.splitBudget = split_budget
    return o


def trace_opts(block=0, lds=0, blocks_per_cu=0, refill=0, leaf_tris=0, grid_percent=0):
    """lrh_trace_opts; zeros select the defaults."""
    o = LrhTraceOpts()
    o.blockSize, o.ldsStack, o.blocksPerCU, o.refillThreshold = block, lds, blocks_per_cu, refill
    o.leafTris, o.gridPercent = leaf_tris, grid_percent
    return o


class Scene:
    def __init__(self, dev, handle, two_level):
        self.dev = dev
        self.h = handle
        self.two_level = two_level
        if dev is not None:
            dev._scenes.add(self)

    def _alive(self):
        return self.h is not None and (self.dev is None or self.dev.h is not None)

    def info(self):
        i = LrhSceneInfo()
        check(lib().lrh_scene_get_info(self.h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in LrhSceneInfo._fields_}

    def update(self, inst_scene):
        arr, keep = _object_structs(inst_scene)
        check(lib().lrh_scene_update(self.h, arr, len(inst_scene.obj_mesh)))

    def refit(self, inst_scene):
        """lrh_scene_refit: new matrices / motion keys, root tree refitted on the GPU."""
        arr, keep = _object_structs(inst_scene)
        check(lib().lrh_scene_refit(self.h, arr, len(inst_scene.obj_mesh)))

    def object_boxes(self, n_objects):
        """(n, 2, 3) float32 culling boxes of the objects (lo, hi)."""
        out = np.zeros((n_objects, 6), np.float32)
        check(lib().lrh_scene_object_boxes(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n_objects))
        return out.reshape(n_objects, 2, 3)

    def download(self):
        inf = self.info()
        nodes = np.zeros(inf["nodeBytes"], np.uint8)
        prims = np.zeros(inf["primBytes"], np.uint8)
        check(lib().lrh_scene_download(self.h, nodes.ctypes.data, nodes.nbytes, prims.ctypes.data, prims.nbytes))
        return nodes, prims

    def close(self):
        if self._alive():
            check(lib().lrh_scene_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            if self._alive():
                lib().lrh_scene_destroy(self.h)
        except Exception:
            pass
        self.h = None


def host_scene_bvh(meshes, layout=0, max_leaf=4, threads=0, trav_cost=1.0, split_budget=0.0):
    """Build a scene on the host only (no GPU): for inspecting the builder."""
    arr, keep = _mesh_structs(meshes)
    opts = _build_opts(layout, max_leaf, threads, trav_cost, split_budget=split_budget)
    p = ctypes.c_void_p()
    check(lib().lrh_scene_create_bvh(None, arr, len(meshes), ctypes.byref(opts), ctypes.byref(p)))
    return Scene(None, p.value, False)


def host_scene_mbvh(inst_scene, layout=0, max_leaf=4, threads=0, trav_cost=1.0, split_budget=0.0):
    marr, keep1 = _mesh_structs(inst_scene.meshes)
    oarr, keep2 = _object_structs(inst_scene)
    opts = _build_opts(layout, max_leaf, threads, trav_cost, split_budget=split_budget)
    p = ctypes.c_void_p()
    check(lib().lrh_scene_create_mbvh(None, marr, len(inst_scene.meshes), oarr, len(inst_scene.obj_mesh),
                                      ctypes.byref(opts), ctypes.byref(p)))
    return Scene(None, p.value, True)


def _batch_items(items):
    arr = (LrhBatchItem * max(1, len(items)))()
    for i, (scene, rb, hb, count, query) in enumerate(items):
        arr[i].scene = scene.h if scene is not None else None
        arr[i].rays = rb.ptr if hasattr(rb, "ptr") else rb
        arr[i].hits = hb.ptr if hasattr(hb, "ptr") else hb
        arr[i].count = count
        arr[i].query = query
    return arr


class Batch:
    """A resident batched launch (lrh_batch_*): items are tuples
    (scene, ray buffer, hit buffer, count, query)."""

    def __init__(self, dev, items, opts=None):
        self.dev = dev
        self.items = list(items)  # keeps the buffers and scenes alive
        arr = _batch_items(self.items)
        p = ctypes.c_void_p()
        check(lib().lrh_batch_create(dev.h, arr, len(self.items), ctypes.byref(opts) if opts else None,
                                     ctypes.byref(p)))
        self.h = p.value
        r = ctypes.c_ulonglong()
        b = ctypes.c_uint()
        bs = ctypes.c_uint()
        check(lib().lrh_batch_info(self.h, ctypes.byref(r), ctypes.byref(b), ctypes.byref(bs)))
        self.rays, self.blocks, self.block_size = r.value, b.value, bs.value

    def launch(self, stream=None):
        check(lib().lrh_batch_launch(self.h, stream))

    def capture(self, stream=None):
        check(lib().lrh_batch_capture(self.h, stream))

    def replay(self, stream=None):
        check(lib().lrh_batch_replay(self.h, stream))

    def close(self):
        if self.h and self.dev.h:
            check(lib().lrh_batch_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            if self.h and self.dev.h:
                lib().lrh_batch_destroy(self.h)
        except Exception:
            pass
        self.h = None


class Device:
    """One GPU: buffers, streams, scenes and traces."""

    def __init__(self, ordinal=0):
        self._scenes = weakref.WeakSet()
        p = ctypes.c_void_p()
        check(lib().lrh_device_open(ordinal, ctypes.byref(p)))
        self.h = p.value
        name = ctypes.create_string_buffer(128)
        cu = ctypes.c_int()
        mem = ctypes.c_size_t()
        check(lib().lrh_device_info(self.h, name, 128, ctypes.byref(cu), ctypes.byref(mem)))
        self.name = name.value.decode()
        self.cu_count = cu.value
        self.total_mem = mem.value

    def close(self):
        """Closes the device; scenes and buffers created on it are released."""
        if self.h:
            for sc in list(self._scenes):
                sc.h = None  # released by lrh_device_close
            check(lib().lrh_device_close(self.h))
            self.h = None

    # -- memory --
    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def upload(self, arr, buf=None, stream=None):
        arr = np.ascontiguousarray(arr)
        if buf is None:
            buf = self.alloc(max(1, arr.nbytes))
        check(lib().lrh_memcpy(self.h, buf.ptr, arr.ctypes.data, arr.nbytes, LRH_H2D, stream))
        return buf

    def download(self, buf, dtype, count, stream=None):
        out = np.empty(count, dtype=dtype)
        check(lib().lrh_memcpy(self.h, out.ctypes.data, buf.ptr, out.nbytes, LRH_D2H, stream))
        return out

    def download_into(self, buf, out, stream=None):
        check(lib().lrh_memcpy(self.h, out.ctypes.data, buf.ptr, out.nbytes, LRH_D2H, stream))
        return out

    def stream(self):
        p = ctypes.c_void_p()
        check(lib().lrh_stream_create(self.h, ctypes.byref(p)))
        return p.value

    def destroy_stream(self, s):
        check(lib().lrh_stream_destroy(self.h, s))

    def sync(self, stream=None):
        check(lib().lrh_sync(self.h, stream))

    def event(self):
        return Event(self)

    def used_memory(self):
        return lib().lrh_used_memory(self.h)

    # -- scenes --
    def scene_bvh(self, meshes, layout=0, max_leaf=4, threads=0, trav_cost=1.0, builder=BUILD_HOST_SAH,
                  split_budget=0.0):
        """lrh_scene_create_bvh; builder=BUILD_DEVICE_LBVH or BUILD_DEVICE_PLOC builds the tree on the GPU; split_budget:
        extra references of early split clipping per triangle (0 = default, negative = off)."""
        arr, keep = _mesh_structs(meshes)
        opts = _build_opts(layout, max_leaf, threads, trav_cost, builder, split_budget)
        p = ctypes.c_void_p()
        check(lib().lrh_scene_create_bvh(self.h, arr, len(meshes), ctypes.byref(opts), ctypes.byref(p)))
        return Scene(self, p.value, False)

    def scene_mbvh(self, inst_scene, layout=0, max_leaf=4, threads=0, trav_cost=1.0, split_budget=0.0):
        marr, keep1 = _mesh_structs(inst_scene.meshes)
        oarr, keep2 = _object_structs(inst_scene)
        opts = _build_opts(layout, max_leaf, threads, trav_cost, split_budget=split_budget)
        p = ctypes.c_void_p()
        check(lib().lrh_scene_create_mbvh(self.h, marr, len(inst_scene.meshes), oarr, len(inst_scene.obj_mesh),
                                          ctypes.byref(opts), ctypes.byref(p)))
        return Scene(self, p.value, True)

    # -- tracing --
    def trace(self, scene, rays_buf, hits_buf, count, query=0, opts=None, stream=None):
        if opts is None:
            check(lib().lrh_trace(scene.h, rays_buf.ptr, hits_buf.ptr, count, query, stream))
        else:
            check(lib().lrh_trace_ex(scene.h, rays_buf.ptr, hits_buf.ptr, count, query, ctypes.byref(opts), stream))

    def trace_stats(self, scene, rays_buf, hits_buf, count, query=0, opts=None, stream=None):
        """(node visits, primitive tests, rays) of one trace (counting kernel variant)."""
        c = self.trace_counters(scene, rays_buf, hits_buf, count, query, opts, stream)
        return c["node_visits"], c["prim_tests"], c["rays"]

    def trace_counters(self, scene, rays_buf, hits_buf, count, query=0, opts=None, stream=None):
        """Every counter of lrh_trace_stats, by name."""
        out = (ctypes.c_ulonglong * TRACE_STATS)()
        check(lib().lrh_trace_stats(scene.h, rays_buf.ptr, hits_buf.ptr, count, query,
                                    ctypes.byref(opts) if opts else None, stream, out))
        return dict(zip(("node_visits", "prim_tests", "rays", "passes", "live_lane_passes", "node_passes",
                         "leaf_passes", "refills", "object_entries", "matrix_entries", "motion_entries"),
                        (int(x) for x in out)))

    def trace_batch(self, items, opts=None, stream=None):
        """One persistent launch over items (scene, rays, hits, count, query)."""
        arr = _batch_items(items)
        check(lib().lrh_trace_batch(self.h, arr, len(items), ctypes.byref(opts) if opts else None, stream))

    def batch(self, items, opts=None):
        return Batch(self, items, opts)

    def trace_numpy(self, scene, rays, query=0, opts=None, hits_init=None):
        """Upload rays, trace, download hits (synchronous convenience)."""
        rays = np.ascontiguousarray(rays)
        assert rays.dtype == RAY_DTYPE
        n = rays.shape[0]
        rb = self.upload(rays)
        if hits_init is not None:
            hb = self.upload(np.ascontiguousarray(hits_init, dtype=HIT_DTYPE))
        else:
            hb = self.alloc(max(1, n * HIT_DTYPE.itemsize))
        self.trace(scene, rb, hb, n, query, opts)
        self.sync()
        return self.download(hb, HIT_DTYPE, n)

    def camera_rays(self, rays_buf, eye, look, up=(0.0, 0.0, 1.0), vfov_deg=45.0, width=512, height=512,
                    mint=1e-4, stream=None):
        c = LrhCamera()
        c.eye[:] = list(eye)
        c.look[:] = list(look)
        c.up[:] = list(up)
        c.vfovDeg = vfov_deg
        c.mint = mint
        c.width, c.height = width, height
        check(lib().lrh_rays_camera(self.h, ctypes.byref(c), rays_buf.ptr, stream))

    def bounce_rays(self, scene, rays_buf, hits_buf, count, bounce, seed, next_buf, shadow_buf=None,
                    pixel_base=0, rr_start=2, rr_survive=0.7, light=(50.0, 50.0, 120.0), live_counter=None,
                    stream=None):
        lc = (ctypes.c_float * 3)(*light)
        check(lib().lrh_rays_bounce(scene.h, rays_buf.ptr, hits_buf.ptr, count, bounce, seed, pixel_base,
                                    rr_start, rr_survive, lc, next_buf.ptr, shadow_buf.ptr if shadow_buf else None,
                                    live_counter, stream))

    def paths_init(self, paths_buf, count, stream=None):
        check(lib().lrh_paths_init(self.h, paths_buf.ptr, count, stream))

    def shade(self, scene, materials_buf, n_materials, lights_buf, n_lights, rays_buf, hits_buf, shadow_hits_buf,
              paths_buf, next_buf, shadow_buf, count, max_depth=8, rr_depth=3, rr_survive=0.7, seed=1,
              pixel_base=0, live_counter=None, stream=None):
        """lrh_shade: one Lambertian + area-light shading step of a path wavefront."""
        p = LrhShadeParams(count, max_depth, rr_depth, rr_survive, seed, pixel_base, 0)
        check(lib().lrh_shade(scene.h, materials_buf.ptr if materials_buf else None, n_materials,
                              lights_buf.ptr if lights_buf else None, n_lights, rays_buf.ptr, hits_buf.ptr,
                              shadow_hits_buf.ptr if shadow_hits_buf else None, paths_buf.ptr, next_buf.ptr,
                              shadow_buf.ptr, ctypes.byref(p), live_counter, stream))

    def shade_batch(self, items, materials_buf, n_materials, lights_buf, n_lights, live_counter=None, stream=None):
        """lrh_shade_batch: several shading steps in one launch; items are
        (scene, rays, hits, shadow_hits or None, paths, next, shadow, count,
        max_depth, rr_depth, rr_survive, seed, pixel_base)."""
        arr = (LrhShadeItem * max(1, len(items)))()
        for k, (sc, rb, hb, shb, pb, nb, sb, cnt, md, rd, rs, seed, pbase) in enumerate(items):
            arr[k] = LrhShadeItem(sc.h, rb.ptr, hb.ptr, shb.ptr if shb else None, pb.ptr, nb.ptr, sb.ptr,
                                  LrhShadeParams(cnt, md, rd, rs, seed, pbase, 0))
        check(lib().lrh_shade_batch(arr if items else None, len(items), materials_buf.ptr if materials_buf else None,
                                    n_materials, lights_buf.ptr if lights_buf else None, n_lights, live_counter,
                                    stream))

    def stats(self):
        r = ctypes.c_ulonglong()
        s = ctypes.c_double()
        check(lib().lrh_device_stats(self.h, ctypes.byref(r), ctypes.byref(s)))
        return r.value, s.value
