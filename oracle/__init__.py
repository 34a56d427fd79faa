"""ctypes front end of the CPU oracle (liblroracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP intersection device.  Only
tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py
may import this module.  The restated reference functions and what pins each
one are listed in lroracle.h.
"""
import ctypes
import os
import subprocess

import numpy as np

from luxcore_amd.raytypes import HIT_DTYPE, RAY_DTYPE
from luxcore_amd.scenes import pack_meshes

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblroracle.so")
_lib = None

NODE_DTYPE = np.dtype([("a", "<u4", (6,)), ("nodeData", "<u4"), ("pad0", "<i4")])
BUILDERS = {"CLASSIC": 0, "SAH": 1}

_u32p = ctypes.POINTER(ctypes.c_uint32)
_f32p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_triangle_intersect.argtypes = [_vp, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
        L.orc_triangle_intersect.restype = ctypes.c_int
        L.orc_bbox_intersectp.argtypes = [_vp, _f32p, _f32p, _f32p, _f32p]
        L.orc_bbox_intersectp.restype = ctypes.c_int
        L.orc_epsilon.argtypes = [ctypes.c_float]
        L.orc_epsilon.restype = ctypes.c_float
        L.orc_epsilon_bbox.argtypes = [_f32p, _f32p]
        L.orc_epsilon_bbox.restype = ctypes.c_float
        L.orc_bvh_create.argtypes = [ctypes.c_uint, _u32p, _u32p, _f32p, _u32p, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int]
        L.orc_bvh_create.restype = _vp
        L.orc_bvh_destroy.argtypes = [_vp]
        L.orc_bvh_node_count.argtypes = [_vp]
        L.orc_bvh_node_count.restype = ctypes.c_uint
        L.orc_bvh_nodes.argtypes = [_vp]
        L.orc_bvh_nodes.restype = _vp
        L.orc_bvh_intersect.argtypes = [_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_ulonglong)]
        L.orc_bruteforce.argtypes = [ctypes.c_uint, _u32p, _u32p, _f32p, _u32p, _vp, _vp, ctypes.c_size_t,
                                     ctypes.c_int]
        L.orc_set_libm.argtypes = [ctypes.c_int]
        L.orc_set_libm.restype = None
        L.orc_matrix_inverse.argtypes = [_f32p, _f32p]
        L.orc_matrix_inverse.restype = ctypes.c_int
        L.orc_motion_sample.argtypes = [ctypes.c_uint, _f32p, _f32p, ctypes.c_float, _f32p]
        L.orc_motion_sample.restype = ctypes.c_int
        L.orc_motion_bound.argtypes = [ctypes.c_uint, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
        L.orc_motion_bound.restype = ctypes.c_int
        L.orc_mbvh_create.argtypes = [ctypes.c_uint, _u32p, _u32p, _f32p, _u32p, ctypes.c_uint, _u32p, _u32p,
                                      _f32p, _u32p, _f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_mbvh_create.restype = _vp
        L.orc_mbvh_destroy.argtypes = [_vp]
        L.orc_mbvh_leaf_bbox.argtypes = [_vp, ctypes.c_uint, _f32p, _f32p]
        L.orc_mbvh_intersect.argtypes = [_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_ulonglong)]
        L.orc_mbvh_bruteforce.argtypes = [_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _threads(n):
    return int(n) if n else (os.cpu_count() or 1)


def _f3(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(3))


def triangle_intersect(ray, p0, p1, p2):
    """Triangle::Intersect on one Ray record -> (hit, t, b1, b2)."""
    r = np.ascontiguousarray(np.asarray(ray, dtype=RAY_DTYPE).reshape(1))
    t, b1, b2 = ctypes.c_float(0), ctypes.c_float(0), ctypes.c_float(0)
    a, b, c = _f3(p0), _f3(p1), _f3(p2)
    hit = lib().orc_triangle_intersect(r.ctypes.data, _p(a, _f32p), _p(b, _f32p), _p(c, _f32p),
                                       ctypes.byref(t), ctypes.byref(b1), ctypes.byref(b2))
    return bool(hit), t.value, b1.value, b2.value


def bbox_intersectp(ray, pmin, pmax):
    r = np.ascontiguousarray(np.asarray(ray, dtype=RAY_DTYPE).reshape(1))
    t0, t1 = ctypes.c_float(0), ctypes.c_float(0)
    a, b = _f3(pmin), _f3(pmax)
    hit = lib().orc_bbox_intersectp(r.ctypes.data, _p(a, _f32p), _p(b, _f32p), ctypes.byref(t0), ctypes.byref(t1))
    return bool(hit), t0.value, t1.value


def epsilon(v):
    return lib().orc_epsilon(float(v))


def epsilon_bbox(pmin, pmax):
    a, b = _f3(pmin), _f3(pmax)
    return lib().orc_epsilon_bbox(_p(a, _f32p), _p(b, _f32p))


def set_libm(correctly_rounded=False):
    """Slerp's acosf/sinf: the host libm's (the default, as the reference
    calls them) or correctly rounded ones."""
    lib().orc_set_libm(1 if correctly_rounded else 0)


def matrix_inverse(m):
    a = np.ascontiguousarray(np.asarray(m, np.float32).reshape(16))
    out = np.zeros(16, np.float32)
    if not lib().orc_matrix_inverse(_p(a, _f32p), _p(out, _f32p)):
        raise ValueError("singular matrix")
    return out.reshape(4, 4)


def motion_sample(times, mats, time):
    t = np.ascontiguousarray(np.asarray(times, np.float32))
    m = np.ascontiguousarray(np.asarray(mats, np.float32).reshape(-1))
    out = np.zeros(16, np.float32)
    if not lib().orc_motion_sample(len(t), _p(t, _f32p), _p(m, _f32p), float(time), _p(out, _f32p)):
        raise ValueError("bad motion system")
    return out.reshape(4, 4)


def motion_bound(times, mats, pmin, pmax):
    t = np.ascontiguousarray(np.asarray(times, np.float32))
    m = np.ascontiguousarray(np.asarray(mats, np.float32).reshape(-1))
    a, b = _f3(pmin), _f3(pmax)
    omin, omax = np.zeros(3, np.float32), np.zeros(3, np.float32)
    if not lib().orc_motion_bound(len(t), _p(t, _f32p), _p(m, _f32p), _p(a, _f32p), _p(b, _f32p),
                                  _p(omin, _f32p), _p(omax, _f32p)):
        raise ValueError("bad motion system")
    return omin, omax


def _check_rays(rays):
    rays = np.ascontiguousarray(rays)
    if rays.dtype != RAY_DTYPE:
        raise TypeError("rays must use RAY_DTYPE")
    return rays


class OracleBVH:
    """BVHAccel restated on the CPU (CLASSIC builder or the binned-SAH
    stand-in for the default Embree builder)."""

    def __init__(self, meshes, builder="SAH", tree_type=4, cost_samples=0, isect_cost=80, trav_cost=10,
                 empty_bonus=0.5, threads=None):
        vo, to, v, t = pack_meshes(meshes)
        self._keep = (vo, to, v, t)
        self.h = lib().orc_bvh_create(len(meshes), _p(vo, _u32p), _p(to, _u32p), _p(v, _f32p), _p(t, _u32p),
                                      BUILDERS[builder], tree_type, cost_samples, isect_cost, trav_cost,
                                      empty_bonus, _threads(threads))
        if not self.h:
            raise RuntimeError("oracle BVH build failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bvh_destroy(self.h)
            self.h = None

    def nodes(self):
        n = lib().orc_bvh_node_count(self.h)
        if n == 0:
            return np.zeros(0, NODE_DTYPE)
        buf = (ctypes.c_char * (n * 32)).from_address(lib().orc_bvh_nodes(self.h))
        return np.frombuffer(bytes(buf), dtype=NODE_DTYPE).copy()

    def intersect(self, rays, threads=None, hits=None, counters=False):
        rays = _check_rays(rays)
        if hits is None:
            hits = np.zeros(rays.shape[0], HIT_DTYPE)
        c = (ctypes.c_ulonglong * 2)()
        lib().orc_bvh_intersect(self.h, rays.ctypes.data, hits.ctypes.data, rays.shape[0], _threads(threads),
                                c if counters else None)
        return (hits, (c[0], c[1])) if counters else hits


def bruteforce(meshes, rays, threads=None):
    rays = _check_rays(rays)
    vo, to, v, t = pack_meshes(meshes)
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    lib().orc_bruteforce(len(meshes), _p(vo, _u32p), _p(to, _u32p), _p(v, _f32p), _p(t, _u32p),
                         rays.ctypes.data, hits.ctypes.data, rays.shape[0], _threads(threads))
    return hits


class OracleMBVH:
    """MBVHAccel restated on the CPU over an ``InstancedScene``."""

    def __init__(self, scene, builder="SAH", tree_type=4, threads=None):
        vo, to, v, t = pack_meshes(scene.meshes)
        om = np.ascontiguousarray(scene.obj_mesh, np.uint32)
        ok = np.ascontiguousarray(scene.obj_kind, np.uint32)
        mat = np.ascontiguousarray(scene.obj_mat, np.float32).reshape(-1)
        koff = np.ascontiguousarray(scene.key_off, np.uint32)
        kt = np.ascontiguousarray(scene.key_times, np.float32)
        km = np.ascontiguousarray(scene.key_mats, np.float32).reshape(-1)
        if kt.size == 0:
            kt = np.zeros(1, np.float32)
            km = np.zeros(16, np.float32)
        self._keep = (vo, to, v, t, om, ok, mat, koff, kt, km)
        self.n_obj = len(om)
        self.h = lib().orc_mbvh_create(len(scene.meshes), _p(vo, _u32p), _p(to, _u32p), _p(v, _f32p),
                                       _p(t, _u32p), len(om), _p(om, _u32p), _p(ok, _u32p), _p(mat, _f32p),
                                       _p(koff, _u32p), _p(kt, _f32p), _p(km, _f32p), BUILDERS[builder],
                                       tree_type, _threads(threads))
        if not self.h:
            raise RuntimeError("oracle MBVH build failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mbvh_destroy(self.h)
            self.h = None

    def leaf_bbox(self, o):
        a, b = np.zeros(3, np.float32), np.zeros(3, np.float32)
        lib().orc_mbvh_leaf_bbox(self.h, o, _p(a, _f32p), _p(b, _f32p))
        return a, b

    def intersect(self, rays, threads=None, counters=False):
        rays = _check_rays(rays)
        hits = np.zeros(rays.shape[0], HIT_DTYPE)
        c = (ctypes.c_ulonglong * 2)()
        lib().orc_mbvh_intersect(self.h, rays.ctypes.data, hits.ctypes.data, rays.shape[0], _threads(threads),
                                 c if counters else None)
        return (hits, (c[0], c[1])) if counters else hits

    def bruteforce(self, rays, threads=None):
        rays = _check_rays(rays)
        hits = np.zeros(rays.shape[0], HIT_DTYPE)
        lib().orc_mbvh_bruteforce(self.h, rays.ctypes.data, hits.ctypes.data, rays.shape[0], _threads(threads))
        return hits
