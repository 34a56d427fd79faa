"""lrh_trace_noalloc, the trace entry point that allocates nothing (its ray
ticket counter and spill stack live in a static per-GPU scratch array), gives
exactly lrh_trace's hit records: called through the C-ABI twice in a row, the
second time with fewer rays, on a single-level and a two-level scene (with
motion), for closest and any hit, with masked rays left untouched.  The
single-level tree is deep enough that the no-alloc launch runs a smaller
persistent grid than lrh_trace (its spill stacks must fit the scratch)."""
import numpy as np
import pytest

from luxcore_amd import device as D
from luxcore_amd import scenes
from luxcore_amd.raytypes import HIT_DTYPE

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 128 * 1024 * 1024


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _poisoned(n):
    h = np.zeros(n, HIT_DTYPE)
    h["t"] = 7.5
    h["b1"] = -3.0
    h["meshIndex"] = 0x12345678
    h["triangleIndex"] = 0x9ABCDEF0
    return h


def _trace_both(dev, sc, rays, query):
    n = rays.shape[0]
    rb = dev.upload(rays)
    init = _poisoned(n)
    hb_ref = dev.upload(init)
    hb_new = dev.upload(init)
    D.check(D.lib().lrh_trace(sc.h, rb.ptr, hb_ref.ptr, n, query, None))
    dev.sync()
    D.check(D.lib().lrh_trace_noalloc(sc.h, rb.ptr, hb_new.ptr, n, query, None))
    dev.sync()
    ref = dev.download(hb_ref, HIT_DTYPE, n)
    got = dev.download(hb_new, HIT_DTYPE, n)
    return ref, got


def _check(rays, ref, got):
    assert ref.tobytes() == got.tobytes()
    masked = (rays["flags"] & 1) != 0
    assert masked.any()
    assert np.all(got["t"][masked] == 7.5) and np.all(got["meshIndex"][masked] == 0x12345678)
    assert (got["meshIndex"][~masked] != 0x12345678).all()


def test_noalloc_equals_trace_single_level(gpu_device):
    meshes = scenes.plain_scene(seed=1, n_random=20000)
    sc = gpu_device.scene_bvh(meshes)
    depth = sc.info()["maxDepth"]
    lane_bytes = 8 * max(1, 3 * (depth + 2) + 4 - 8)
    print("tree depth", depth, "lanes the scratch holds", (SCRATCH_BYTES - 256) // lane_bytes, flush=True)
    for query in (D.LRH_QUERY_CLOSEST, D.LRH_QUERY_ANY):
        for n in (1 << 18, (1 << 18) - 12345 * 7):   # second call: fewer rays
            _case("noalloc_bvh", query=query, rays=n)
            rays = scenes.plain_rays(n, seed=31 + n)
            rays["flags"][::13] = 1
            ref, got = _trace_both(gpu_device, sc, rays, query)
            _check(rays, ref, got)
    sc.close()


def test_noalloc_equals_trace_two_level(gpu_device):
    isc = scenes.instanced_scene(n_inst=60, motion_frac=0.3, freq=20, extent=40.0, seed=12)
    sc = gpu_device.scene_mbvh(isc)
    for query in (D.LRH_QUERY_CLOSEST, D.LRH_QUERY_ANY):
        for n in (1 << 17, 1000):   # second call: fewer rays
            _case("noalloc_mbvh", query=query, rays=n)
            rays = scenes.instanced_rays(n, extent=40.0, seed=41 + n)
            rays["flags"][::13] = 1
            ref, got = _trace_both(gpu_device, sc, rays, query)
            _check(rays, ref, got)
    sc.close()


def test_noalloc_rejects_bad_arguments(gpu_device):
    _case("noalloc_args")
    assert D.lib().lrh_trace_noalloc(None, None, None, 0, 0, None) == D.LRH_EINVAL
    sc = gpu_device.scene_bvh([scenes.box_mesh()])
    rb = gpu_device.alloc(48 * 4)
    hb = gpu_device.alloc(20 * 4)
    assert D.lib().lrh_trace_noalloc(sc.h, rb.ptr, hb.ptr, 4, 2, None) == D.LRH_EINVAL
    assert D.lib().lrh_trace_noalloc(sc.h, rb.ptr, hb.ptr, 0, 0, None) == 0
    sc.close()
