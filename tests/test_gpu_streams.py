"""Traces enqueued on several streams at once (each stream has its own
ticket counter and spill stack) give the same hits as the same traces one
after another on the device's own stream."""
import numpy as np
import pytest

from luxcore_amd import device as D
from luxcore_amd import scenes
from luxcore_amd.raytypes import HIT_DTYPE

pytestmark = pytest.mark.gpu


def test_concurrent_streams_equal_sequential(gpu_device):
    print("[case] concurrent streams", flush=True)
    dev = gpu_device
    plain = dev.scene_bvh(scenes.plain_scene(seed=4, n_random=3000))
    inst = scenes.instanced_scene(n_inst=80, motion_frac=0.3, freq=12, extent=20.0, seed=5)
    two = dev.scene_mbvh(inst)
    work = [(plain, scenes.plain_rays(150_000, seed=11), D.LRH_QUERY_CLOSEST),
            (plain, scenes.plain_rays(90_000, seed=12), D.LRH_QUERY_ANY),
            (two, scenes.instanced_rays(120_000, extent=20.0, seed=13), D.LRH_QUERY_CLOSEST),
            (two, scenes.instanced_rays(60_000, extent=20.0, seed=14), D.LRH_QUERY_ANY)]
    ref = [dev.trace_numpy(sc, r, query=q) for sc, r, q in work]
    streams = [dev.stream() for _ in work]
    bufs = [(dev.upload(r), dev.alloc(r.shape[0] * HIT_DTYPE.itemsize)) for _, r, _ in work]
    for rep in range(4):
        for (sc, r, q), (rb, hb), st in zip(work, bufs, streams):
            dev.trace(sc, rb, hb, r.shape[0], q, stream=st)
    # a batched launch of all four on yet another stream, concurrently with more plain traces
    bst = dev.stream()
    outs = [dev.alloc(r.shape[0] * HIT_DTYPE.itemsize) for _, r, _ in work]
    dev.trace_batch([(sc, rb, ob, r.shape[0], q) for (sc, r, q), (rb, _), ob in zip(work, bufs, outs)],
                    stream=bst)
    for (sc, r, q), (rb, hb), st in zip(work, bufs, streams):
        dev.trace(sc, rb, hb, r.shape[0], q, stream=st)
    for st in streams + [bst]:
        dev.sync(st)
    for k, ((_, r, _), (rb, hb), ob) in enumerate(zip(work, bufs, outs)):
        got = dev.download(hb, HIT_DTYPE, r.shape[0])
        assert got.tobytes() == ref[k].tobytes(), k
        got_b = dev.download(ob, HIT_DTYPE, r.shape[0])
        assert got_b.tobytes() == ref[k].tobytes(), ("batch", k)
    for st in streams + [bst]:
        dev.destroy_stream(st)
    for rb, hb in bufs:
        rb.free()
        hb.free()
    for ob in outs:
        ob.free()
    plain.close()
    two.close()
