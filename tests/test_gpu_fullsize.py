"""Parity at the benchmarked sizes, through properties that hold at any size.

The 26,238,080-triangle heightfield scene of the headline benchmark (1920x1080
primaries, one cosine bounce per hit, one shadow ray per hit) and the batched
workload (8 resident scenes, 11.76M triangles, 64 buffers of 4Ki..4Mi rays,
mixed closest/any hit) are traced whole; the checks are determinism,
independence of the node layout, any-hit/closest-hit agreement on misses,
masked rays leaving hit memory untouched, batched == per-buffer traces bit for
bit, and random subsets against the oracle built over the full scene.
"""
import os

import numpy as np
import pytest

import oracle
from luxcore_amd import batch as B
from luxcore_amd import scenes
from luxcore_amd.raytypes import HIT_DTYPE, NULL_INDEX, RAY_FLAGS_MASKED
from tests.parity import assert_parity, compare, tri_tie_checker

pytestmark = pytest.mark.gpu

THREADS = min(os.cpu_count() or 1, 16)


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _subset(n, k, seed):
    return np.sort(np.random.Generator(np.random.PCG64(seed)).choice(n, k, replace=False))


@pytest.fixture(scope="module")
def headline(gpu_device):
    meshes = scenes.heightfield_scene()
    assert sum(m.ntris for m in meshes) == 26_238_080
    sc = gpu_device.scene_bvh(meshes)
    prim = scenes.pinhole_rays(1920, 1080)
    hp = gpu_device.trace_numpy(sc, prim)
    bounce = scenes.bounce_rays(meshes, prim, hp, seed=5)
    shadow = scenes.shadow_rays(meshes, prim, hp, seed=6)
    yield meshes, sc, prim, hp, bounce, shadow
    sc.close()


def test_headline_scene_determinism_and_layouts(gpu_device, headline):
    meshes, sc, prim, hp, bounce, _ = headline
    _case("headline determinism + fp32 nodes", tris=26_238_080, rays=bounce.shape[0])
    assert bounce.shape[0] == 2_073_453
    a = gpu_device.trace_numpy(sc, bounce)
    assert np.array_equal(a, gpu_device.trace_numpy(sc, bounce))
    assert np.array_equal(hp, gpu_device.trace_numpy(sc, prim))
    # another node layout and leaf size: another traversal order, same closest hits
    s1 = gpu_device.scene_bvh(meshes, layout=1, max_leaf=2)
    try:
        for name, r, ref in (("bounce", bounce, a), ("primary", prim, hp)):
            got = gpu_device.trace_numpy(s1, r)
            st = compare(ref, got)
            print(name, st, flush=True)
            assert st["bad"] == 0 and st["mismatch"] - st["edge_band"] < 1e-5 * r.shape[0] + 5
    finally:
        s1.close()


def test_headline_scene_against_oracle_subset(gpu_device, headline):
    meshes, sc, prim, hp, bounce, _ = headline
    _case("headline oracle subsets", threads=THREADS)
    ob = oracle.OracleBVH(meshes, builder="SAH", threads=THREADS)
    a = gpu_device.trace_numpy(sc, bounce)
    for name, r, got, seed in (("bounce", bounce, a, 11), ("primary", prim, hp, 12)):
        sel = _subset(r.shape[0], 32_768, seed)
        ref = ob.intersect(r[sel], threads=THREADS)
        st = compare(ref, got[sel], tri_tie_checker(oracle, meshes, r[sel], ref, got[sel]))
        print(name, st, flush=True)
        assert st["hits"] > 1000
        assert_parity(st)


def test_headline_scene_shadow_and_masked(gpu_device, headline):
    _, sc, _, _, bounce, shadow = headline
    _case("headline shadow any/closest + masked", rays=shadow.shape[0])
    anyh = gpu_device.trace_numpy(sc, shadow, query=1)
    clo = gpu_device.trace_numpy(sc, shadow, query=0)
    assert np.array_equal(anyh["meshIndex"] == NULL_INDEX, clo["meshIndex"] == NULL_INDEX)
    hit = anyh["meshIndex"] != NULL_INDEX
    assert np.all(anyh["t"][hit] >= shadow["mint"][hit]) and np.all(anyh["t"][hit] <= shadow["maxt"][hit])
    assert np.all(clo["t"][hit] <= anyh["t"][hit])
    # every 97th bounce ray masked: those hit records keep their poison, the rest equal the unmasked trace
    r = bounce.copy()
    r["flags"][::97] |= RAY_FLAGS_MASKED
    init = np.zeros(r.shape[0], HIT_DTYPE)
    init["t"] = -7.0
    init["meshIndex"] = 0xBEEF
    got = gpu_device.trace_numpy(sc, r, hits_init=init)
    ref = gpu_device.trace_numpy(sc, bounce)
    m = np.zeros(r.shape[0], bool)
    m[::97] = True
    assert np.array_equal(got[m], init[m])
    assert np.array_equal(got[~m], ref[~m])


def test_batched_workload_full_size(gpu_device):
    meshes = scenes.batch_scenes()
    bufs = scenes.batch_buffers()
    total = sum(c for _, c, _ in bufs)
    _case("batched full size", tris=sum(m[0].ntris for m in meshes), buffers=len(bufs), rays=total)
    assert len(meshes) == 8 and len(bufs) == 64
    wl = B.BatchedWorkload(gpu_device, meshes, bufs)
    items = wl.items()
    eager = []
    for (sc, rb, _, c, q) in items:
        hb = gpu_device.alloc(c * HIT_DTYPE.itemsize)
        gpu_device.trace(sc, rb, hb, c, q)
        eager.append(hb)
    gpu_device.sync()
    b = gpu_device.batch(items)
    b.launch()
    gpu_device.sync()
    n_any = 0
    for i, (si, c, q) in enumerate(bufs):
        e = gpu_device.download(eager[i], HIT_DTYPE, c)
        g = gpu_device.download(wl.hit_bufs[i], HIT_DTYPE, c)
        assert np.array_equal(e, g), f"buffer {i} (scene {si}, {c} rays, query {q})"
        n_any += q
        eager[i].free()
    assert 0 < n_any < len(bufs)
    # one buffer of each scene against the oracle on a subset
    seen = set()
    for i, (si, c, q) in enumerate(bufs):
        if si in seen or q != 0:
            continue
        seen.add(si)
        sel = _subset(c, min(c, 4096), 100 + i)
        r = wl.rays[i][sel]
        got = gpu_device.download(wl.hit_bufs[i], HIT_DTYPE, c)[sel]
        ref = oracle.OracleBVH(meshes[si], threads=THREADS).intersect(r, threads=THREADS)
        st = compare(ref, got, tri_tie_checker(oracle, meshes[si], r, ref, got))
        print("scene", si, st, flush=True)
        assert_parity(st)
    b.close()
