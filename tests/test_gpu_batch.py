"""Batched launches: one persistent kernel over many ray buffers of several
resident scenes (plain, as a resident batch, and replayed from a hipGraph)
must give exactly the per-buffer results; the autotuner picks and caches a
launch shape."""
import os

import numpy as np
import pytest

import oracle
from luxcore_amd import batch as B
from luxcore_amd import device as D
from luxcore_amd import scenes
from luxcore_amd.raytypes import HIT_DTYPE, NULL_INDEX
from tests.parity import assert_parity, compare

pytestmark = pytest.mark.gpu

SIZES = (2_000, 5_000, 20_000, 40_000, 60_000, 100_000, 150_000, 250_000)


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.fixture(scope="module")
def small_batch(gpu_device):
    meshes = scenes.batch_scenes(SIZES, seed=7)
    bufs = scenes.batch_buffers(n_buffers=24, lo=256, hi=65536, n_scenes=len(SIZES), seed=8)
    wl = B.BatchedWorkload(gpu_device, meshes, bufs, seed=9)
    return meshes, wl


def _eager(dev, wl):
    out = []
    for i, (si, count, query) in enumerate(wl.buffers):
        hb = dev.alloc(count * 20)
        dev.trace(wl.scenes[si], wl.ray_bufs[i], hb, count, query)
        out.append(hb)
    dev.sync()
    return [dev.download(hb, HIT_DTYPE, c) for hb, (_, c, _) in zip(out, wl.buffers)]


def _poison(dev, wl):
    for i, (_, count, _) in enumerate(wl.buffers):
        junk = np.zeros(count, HIT_DTYPE)
        junk["t"] = -1.0
        junk["meshIndex"] = 0xDEAD
        dev.upload(junk, wl.hit_bufs[i])


def _collect(dev, wl):
    return [dev.download(wl.hit_bufs[i], HIT_DTYPE, c) for i, (_, c, _) in enumerate(wl.buffers)]


def test_batch_equals_per_buffer_traces(gpu_device, small_batch):
    meshes, wl = small_batch
    _case("batch_vs_eager", buffers=len(wl.buffers), rays=wl.total_rays())
    ref = _eager(gpu_device, wl)
    _poison(gpu_device, wl)
    gpu_device.trace_batch(wl.items())
    gpu_device.sync()
    got = _collect(gpu_device, wl)
    for i in range(len(ref)):
        assert np.array_equal(ref[i], got[i]), i


def test_batch_against_oracle(gpu_device, small_batch):
    meshes, wl = small_batch
    _case("batch_vs_oracle")
    _poison(gpu_device, wl)
    gpu_device.trace_batch(wl.items())
    gpu_device.sync()
    got = _collect(gpu_device, wl)
    checked = 0
    for i, (si, count, query) in enumerate(wl.buffers):
        if query != 0 or checked >= 4:
            continue
        ref = oracle.OracleBVH(meshes[si]).intersect(wl.rays[i])
        st = compare(ref, got[i])
        print(i, st, flush=True)
        assert_parity(st)
        checked += 1
    assert checked >= 2


def test_resident_batch_and_graph_replays(gpu_device, small_batch):
    meshes, wl = small_batch
    _case("graph_replay")
    ref = _eager(gpu_device, wl)
    b = gpu_device.batch(wl.items(), D.trace_opts(256, 8, 0, 12))
    _poison(gpu_device, wl)
    b.launch()
    gpu_device.sync()
    got = _collect(gpu_device, wl)
    assert all(np.array_equal(a, c) for a, c in zip(ref, got))
    b.capture()
    for rep in range(4):
        _poison(gpu_device, wl)
        b.replay()
        gpu_device.sync()
        # eager work between replays
        _eager(gpu_device, wl) if rep == 1 else None
        got = _collect(gpu_device, wl)
        for i in range(len(ref)):
            assert np.array_equal(ref[i], got[i]), (rep, i)
    b.close()


def test_batch_masked_empty_and_two_level(gpu_device):
    _case("batch_mixed")
    dev = gpu_device
    inst = scenes.instanced_scene(n_inst=30, motion_frac=0.3, freq=10, extent=8.0, seed=51)
    two = dev.scene_mbvh(inst)
    one = dev.scene_bvh([scenes.displaced_sphere(freq=20, seed=2)])
    r1 = scenes.instanced_rays(30000, extent=8.0, seed=52)
    r1["flags"][::5] = 1
    r2 = scenes.buffer_rays(20000, extent=2.0, seed=53)
    r2["o"] -= 1.0
    rb1, rb2 = dev.upload(r1), dev.upload(r2)
    hb1, hb2, hb0 = dev.alloc(30000 * 20), dev.alloc(20000 * 20), dev.alloc(64)
    init = np.zeros(30000, HIT_DTYPE)
    init["t"] = 5.5
    dev.upload(init, hb1)
    dev.trace_batch([(two, rb1, hb1, 30000, 0), (one, rb2, hb2, 0, 0), (one, rb2, hb2, 20000, 1),
                     (two, rb1, hb0, 0, 1)])
    dev.sync()
    g1 = dev.download(hb1, HIT_DTYPE, 30000)
    g2 = dev.download(hb2, HIT_DTYPE, 20000)
    e1 = dev.trace_numpy(two, r1, hits_init=init)
    e2 = dev.trace_numpy(one, r2, query=1)
    assert np.array_equal(g1, e1) and np.array_equal(g2, e2)
    assert np.all(g1["t"][::5] == 5.5)
    assert (g2["meshIndex"] != NULL_INDEX).sum() > 1000


def test_autotuner_picks_and_caches(gpu_device, small_batch, tmp_path):
    _, wl = small_batch
    _case("autotune")
    path = str(tmp_path / "tune.json")
    shapes = [(256, 8, 12), (128, 8, 8), (256, 4, 16)]
    opts, table = B.autotune(gpu_device, wl.items(), shapes=shapes, reps=2, cache_path=path, key="k1")
    assert len(table) == 3 and os.path.exists(path)
    best = min(table, key=lambda r: r["ms"])
    assert (opts.blockSize, opts.ldsStack, opts.refillThreshold) == (best["block"], best["lds"], best["refill"])
    opts2, _ = B.autotune(gpu_device, wl.items(), shapes=shapes, reps=2, cache_path=path, key="k1")
    assert (opts2.blockSize, opts2.ldsStack, opts2.refillThreshold) == (best["block"], best["lds"], best["refill"])
