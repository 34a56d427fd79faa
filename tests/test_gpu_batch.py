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
    path = str(tmp_path / "tune.txt")
    shapes = [(256, 8, 12), (128, 8, 8), (256, 4, 16), (256, 8, 12, 50)]
    opts, table = B.autotune(gpu_device, wl.items(), shapes=shapes, cache_path=path, key="k1")
    assert len(table) == 4 and os.path.exists(path)
    best = min(table, key=lambda r: r["ms"])

    def fields(o):
        return o.blockSize, o.ldsStack, o.refillThreshold, o.gridPercent
    want = (best["block"], best["lds"], best["refill"], best["grid_percent"])
    assert fields(opts) == want
    # the second call answers from the cache (every field of the shape) and times nothing
    opts2, table2 = B.autotune(gpu_device, wl.items(), shapes=shapes, cache_path=path, key="k1")
    assert table2 == [] and fields(opts2) == want
    # another key is tuned and stored beside the first
    B.autotune(gpu_device, wl.items(), shapes=shapes[3:], cache_path=path, key="k2")
    with open(path) as f:
        assert sorted(ln.split()[0] for ln in f if ln.strip()) == ["k1", "k2"]
    opts3, _ = B.autotune(gpu_device, wl.items(), shapes=shapes, cache_path=path, key="k2")
    assert fields(opts3) == (256, 8, 12, 50)
    # a line of five fields (no grid share) reads as grid share 0
    with open(path, "a") as f:
        f.write("k3 128 4 16 1.5\n")
    opts4, table4 = B.autotune(gpu_device, wl.items(), shapes=shapes, cache_path=path, key="k3")
    assert table4 == [] and fields(opts4) == (128, 4, 16, 0)


def test_build_autotuner_times_every_layout(gpu_device, small_batch):
    """The node layout and leaf size are autotuned too: each candidate build
    of the scenes is timed over the same buffers with its best launch shape,
    and the fastest is returned."""
    meshes, wl = small_batch
    _case("tune_build")
    items = [(si, wl.ray_bufs[i], wl.hit_bufs[i], c, q) for i, (si, c, q) in enumerate(wl.buffers)]
    builds = ((0, 4), (1, 4), (2, 4), (2, 2))
    best, res = B.tune_build(gpu_device, meshes, items, builds=builds, shapes=[(256, 8, 8), (128, 16, 16)])
    print(best, res, flush=True)
    assert [(r["layout"], r["max_leaf"]) for r in res] == list(builds)
    assert all(r["ms"] > 0 for r in res)
    fastest = min(res, key=lambda r: r["ms"])
    assert (best["layout"], best["max_leaf"], best["ms"]) == (fastest["layout"], fastest["max_leaf"], fastest["ms"])
    assert (best["block"], best["lds"]) in ((256, 8), (128, 16))
    # a scene index outside the scenes is refused before anything is built
    bad = [(len(meshes), *items[0][1:])]
    with pytest.raises(D.LrhError, match="out of range"):
        B.tune_build(gpu_device, meshes, bad, builds=builds[:1])


def _more_keys(sc, n_keys=4, seed=71):
    """The same objects with n_keys motion keys each (the arrays of motion
    systems grow, so an edit with them moves the scene's device arrays)."""
    import copy
    rng = np.random.Generator(np.random.PCG64(seed))
    out = copy.deepcopy(sc)
    times, mats, off = [], [], [0]
    for o in range(len(sc.obj_kind)):
        if sc.obj_kind[o] == 2:
            k0 = int(sc.key_off[o])
            M0 = np.linalg.inv(sc.key_mats[k0].astype(np.float64))
            for k in range(n_keys):
                Mk = M0.copy()
                Mk[:3, 3] += rng.uniform(-2.0, 2.0, size=3)
                times.append(k / (n_keys - 1))
                mats.append(np.linalg.inv(Mk).astype(np.float32))
            off.append(off[-1] + n_keys)
        else:
            off.append(off[-1])
    out.key_off = np.asarray(off, np.uint32)
    out.key_times = np.asarray(times, np.float32)
    out.key_mats = np.asarray(mats, np.float32).reshape(-1, 4, 4)
    return out


def test_resident_batch_refuses_changed_or_destroyed_scene(gpu_device):
    """A resident batch keeps its scenes' device addresses: after an edit that
    moves them (more motion keys) it refuses to launch or replay, as it does
    after a scene is destroyed; a refit that moves nothing leaves it valid."""
    _case("batch_stale_scene")
    dev = gpu_device
    sc = scenes.instanced_scene(n_inst=40, motion_frac=0.5, freq=8, extent=20.0, seed=70)
    dsc = dev.scene_mbvh(sc)
    flat = dev.scene_bvh(scenes.plain_scene(seed=1, n_random=100))
    rays = scenes.instanced_rays(4096, extent=20.0, seed=72)
    rb = dev.upload(rays)
    hb = dev.alloc(rays.shape[0] * 20)
    hb2 = dev.alloc(rays.shape[0] * 20)
    b = dev.batch([(dsc, rb, hb, rays.shape[0], 0), (flat, rb, hb2, rays.shape[0], 0)])
    b.launch()
    b.capture()
    b.replay()
    dev.sync()
    # a refit with the same key counts: arrays rewritten in place, the tree
    # keeps its topology and depth, so the batch stays valid
    same = _more_keys(sc, n_keys=2, seed=73)
    dsc.refit(same)
    b.launch()
    dev.sync()
    got = dev.download(hb, HIT_DTYPE, rays.shape[0])
    ref = oracle.OracleMBVH(same).intersect(rays)
    assert_parity(compare(ref, got))
    # more keys: the motion arrays grow and move
    dsc.update(_more_keys(sc, n_keys=5, seed=74))
    with pytest.raises(D.LrhError, match="recreate the batch"):
        b.launch()
    with pytest.raises(D.LrhError, match="recreate the batch"):
        b.replay()
    b.close()
    # a fresh batch over the edited scene traces again
    b2 = dev.batch([(flat, rb, hb2, rays.shape[0], 0)])
    b2.launch()
    dev.sync()
    flat.close()
    with pytest.raises(D.LrhError, match="destroyed"):
        b2.launch()
    b2.close()
    dsc.close()
