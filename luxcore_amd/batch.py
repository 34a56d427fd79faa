"""Batched tracing: many ray buffers over several resident scenes in one
persistent launch (optionally a captured hipGraph), and the launch-shape
autotuner.

The autotuner (lrh_autotune_batch, in the library) times candidate launch
shapes (block size x LDS stack depth x refill threshold) of the batched
kernel on the actual batch and keeps the fastest; results are cached on disk
per (device, workload signature) so a serving process tunes once.  Node layout and leaf size are build-time choices:
``tune_build`` rebuilds the scenes for each candidate and times the best
launch shape on each.
"""
import ctypes
import hashlib
import itertools
import json
import os
import time

import numpy as np

from . import device as D

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "lrhip", "autotune.txt")

# the library's built-in candidates (LRH_AUTOTUNE_SHAPES of them), in its order
SHAPES = [(b, l, r) for b, l, r in itertools.product((128, 256), (4, 8, 16), (4, 8, 12, 16, 24))]


def signature(dev, scenes_info, counts, extra=""):
    h = hashlib.sha1()
    h.update(dev.name.encode())
    h.update(json.dumps(scenes_info, sort_keys=True).encode())
    h.update(np.asarray(counts, np.int64).tobytes())
    h.update(extra.encode())
    return h.hexdigest()[:16]


def time_batch(dev, items, opts, reps=5):
    """Median ms of one batched launch (HIP events on the device stream)."""
    b = dev.batch(items, opts)
    for _ in range(2):
        b.launch()
    dev.sync()
    evs = [(dev.event(), dev.event()) for _ in range(reps)]
    for e0, e1 in evs:
        e0.record()
        b.launch()
        e1.record()
    dev.sync()
    ms = float(np.median([e0.elapsed_ms(e1) for e0, e1 in evs]))
    b.close()
    return ms


def autotune(dev, items, shapes=None, cache_path=DEFAULT_CACHE, key=None, log=None):
    """Best lrh_trace_opts for this batch from the library's autotuner
    (lrh_autotune_batch); returns (opts, table of per-shape ms; empty when the
    answer came from the cache)."""
    cand = shapes if shapes is not None else SHAPES
    arr = D._batch_items(items)
    carr = (D.LrhTraceOpts * len(cand))()
    for i, shape in enumerate(cand):
        b, l, r = shape[:3]
        carr[i] = D.trace_opts(b, l, 0, r, grid_percent=shape[3] if len(shape) > 3 else 0)
    best = D.LrhTraceOpts()
    best_ms = ctypes.c_float()
    ms = (ctypes.c_float * len(cand))()
    if cache_path:
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
    D.check(D.lib().lrh_autotune_batch(dev.h, arr, len(items), carr if shapes is not None else None,
                                       len(cand) if shapes is not None else 0,
                                       cache_path.encode() if cache_path else None, key.encode() if key else None,
                                       ctypes.byref(best), ctypes.byref(best_ms), ms))
    table = [{"block": sh[0], "lds": sh[1], "refill": sh[2], "grid_percent": sh[3] if len(sh) > 3 else 0,
              "ms": float(ms[i])} for i, sh in enumerate(cand) if not np.isnan(ms[i])]
    if log:
        for row in table:
            log(f"autotune: {row}")
    return D.trace_opts(best.blockSize, best.ldsStack, best.blocksPerCU, best.refillThreshold, best.leafTris,
                        best.gridPercent), table


def tune_build(dev, scene_meshes, items, builds=((0, 4), (1, 4), (2, 4), (2, 2)), shapes=None, log=None):
    """Build autotuner (lrh_autotune_build, in the library): each (layout,
    max_leaf) build -- node layout 0 (4-wide quantized), 1 (4-wide fp32) or
    2 (8-wide quantized), leaf size -- of the scenes is timed over the items
    with its best launch shape.  items: (scene index, ray buffer, hit buffer,
    count, query).  Returns (best build, results)."""
    shapes = shapes or [(256, 8, 8), (256, 8, 12), (256, 4, 12), (128, 8, 12)]
    keep = []
    mesh_arrs = (ctypes.POINTER(D.LrhMesh) * len(scene_meshes))()
    counts = (ctypes.c_uint * len(scene_meshes))()
    for j, meshes in enumerate(scene_meshes):
        arr, k = D._mesh_structs(meshes)
        keep += [arr, k]
        mesh_arrs[j] = ctypes.cast(arr, ctypes.POINTER(D.LrhMesh))
        counts[j] = len(meshes)
    arr = D._batch_items([(None, rb, hb, c, q) for (_, rb, hb, c, q) in items])
    idx = (ctypes.c_uint * len(items))(*[si for (si, _, _, _, _) in items])
    cand = (D.LrhBuildOpts * len(builds))()
    for i, (layout, leaf) in enumerate(builds):
        cand[i] = D._build_opts(layout=layout, max_leaf=leaf)
    sh = (D.LrhTraceOpts * len(shapes))()
    for i, (b, l, r) in enumerate(shapes):
        sh[i] = D.trace_opts(b, l, 0, r)
    best_c = ctypes.c_uint()
    best_o = D.LrhTraceOpts()
    ms = (ctypes.c_float * len(builds))()
    D.check(D.lib().lrh_autotune_build(dev.h, mesh_arrs, counts, len(scene_meshes), arr, idx, len(items), cand,
                                       len(builds), sh, len(shapes), ctypes.byref(best_c), ctypes.byref(best_o), ms))
    res = [{"layout": layout, "max_leaf": leaf, "ms": float(ms[i])} for i, (layout, leaf) in enumerate(builds)]
    if log:
        for r in res:
            log(f"build layout {r['layout']} leaf {r['max_leaf']}: {r['ms']:.3f} ms")
    best = dict(res[best_c.value], block=best_o.blockSize, lds=best_o.ldsStack, refill=best_o.refillThreshold)
    return best, res


class BatchedWorkload:
    """Resident scenes + ray buffers of a batched run (device memory)."""

    def __init__(self, dev, scene_meshes, buffers, extent=100.0, seed=300, layout=0, max_leaf=4):
        from . import scenes as S
        self.dev = dev
        self.scenes = [dev.scene_bvh(m, layout=layout, max_leaf=max_leaf) for m in scene_meshes]
        self.buffers = buffers
        self.rays = []
        self.ray_bufs = []
        self.hit_bufs = []
        for i, (si, count, query) in enumerate(buffers):
            r = S.buffer_rays(count, extent, seed + i)
            self.rays.append(r)
            self.ray_bufs.append(dev.upload(r))
            self.hit_bufs.append(dev.alloc(count * 20))

    def items(self):
        return [(self.scenes[si], self.ray_bufs[i], self.hit_bufs[i], count, query)
                for i, (si, count, query) in enumerate(self.buffers)]

    def total_rays(self):
        return sum(c for _, c, _ in self.buffers)
