"""GPU parity of the single-level device (BVHAccel path) against the oracle.

Every case prints its name/seed/shape first so a failing GPU run names it.
"""
import json
import os

import numpy as np
import pytest

import oracle
from luxcore_amd import device as D
from luxcore_amd import scenes
from luxcore_amd.raytypes import HIT_DTYPE, NULL_INDEX, RAY_DTYPE, RAY_FLAGS_MASKED, make_rays
from tests.parity import assert_parity, chunk_digests, compare, tri_tie_checker

pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.fixture(scope="module")
def plain():
    meshes = scenes.plain_scene(seed=1)
    rays = scenes.plain_rays(1 << 16, seed=2)
    ob = oracle.OracleBVH(meshes, builder="SAH")
    ref = ob.intersect(rays)
    return meshes, rays, ref


@pytest.mark.parametrize("layout,max_leaf", [(0, 4), (0, 1), (0, 8), (1, 4), (2, 4), (2, 1), (2, 8)])
def test_plain_closest_parity(gpu_device, plain, layout, max_leaf):
    meshes, rays, ref = plain
    _case("plain_closest", layout=layout, max_leaf=max_leaf, rays=rays.shape[0])
    sc = gpu_device.scene_bvh(meshes, layout=layout, max_leaf=max_leaf)
    got = gpu_device.trace_numpy(sc, rays)
    st = compare(ref, got, tri_tie_checker(oracle, meshes, rays, ref, got))
    print(st, flush=True)
    assert_parity(st)
    assert st["hits"] > 0.99 * rays.shape[0]


def test_plain_anyhit(gpu_device, plain):
    meshes, rays, ref = plain
    _case("plain_any", rays=rays.shape[0])
    sc = gpu_device.scene_bvh(meshes)
    got = gpu_device.trace_numpy(sc, rays, query=1)
    rhit = ref["meshIndex"] != NULL_INDEX
    ghit = got["meshIndex"] != NULL_INDEX
    assert np.array_equal(rhit, ghit)
    # every reported any-hit is a real hit of that triangle at that t
    idx = np.nonzero(ghit)[0][:2000]
    for i in idx:
        m, t = int(got["meshIndex"][i]), int(got["triangleIndex"][i])
        tv = meshes[m].tris[t]
        v = meshes[m].verts
        hit, tt, b1, b2 = oracle.triangle_intersect(rays[i], v[tv[0]], v[tv[1]], v[tv[2]])
        assert hit and tt == got["t"][i] and b1 == got["b1"][i] and b2 == got["b2"][i], i
        assert rays["mint"][i] <= tt < rays["maxt"][i]


def test_masked_rays_untouched(gpu_device, plain):
    meshes, rays, ref = plain
    _case("masked", rays=4096)
    r = rays[:4096].copy()
    r["flags"][::3] = RAY_FLAGS_MASKED
    init = np.zeros(r.shape[0], HIT_DTYPE)
    init["t"] = -7.0
    init["meshIndex"] = 0xABCDEF
    init["triangleIndex"] = 0x123456
    sc = gpu_device.scene_bvh(meshes)
    got = gpu_device.trace_numpy(sc, r, hits_init=init)
    masked = (r["flags"] & RAY_FLAGS_MASKED) != 0
    assert np.array_equal(got[masked], init[masked])
    st = compare(ref[:4096][~masked], got[~masked])
    assert_parity(st)


def test_finite_maxt_and_mint(gpu_device, plain):
    meshes, rays, _ = plain
    _case("maxt", rays=16384)
    r = rays[:16384].copy()
    rng = np.random.Generator(np.random.PCG64(11))
    r["maxt"] = rng.uniform(0.01, 3.0, r.shape[0]).astype(np.float32)
    r["mint"] = rng.uniform(0.0, 0.5, r.shape[0]).astype(np.float32)
    ob = oracle.OracleBVH(meshes, builder="CLASSIC")
    ref = ob.intersect(r)
    sc = gpu_device.scene_bvh(meshes)
    got = gpu_device.trace_numpy(sc, r)
    st = compare(ref, got, tri_tie_checker(oracle, meshes, r, ref, got))
    print(st, flush=True)
    assert_parity(st)
    miss = got["meshIndex"] == NULL_INDEX
    assert np.array_equal(got["t"][miss], r["maxt"][miss])


def test_empty_scene_all_miss(gpu_device):
    _case("empty")
    sc = gpu_device.scene_bvh([])
    r = scenes.plain_rays(1000, seed=3)
    r["maxt"][::2] = 5.0
    got = gpu_device.trace_numpy(sc, r)
    assert np.all(got["meshIndex"] == NULL_INDEX)
    assert np.all(got["triangleIndex"] == NULL_INDEX)
    assert np.array_equal(got["t"], r["maxt"])


def test_single_triangle_and_ragged_counts(gpu_device):
    _case("single_triangle")
    m = scenes.Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32))
    for n in (1, 63, 64, 65, 1000, 4097):
        rng = np.random.Generator(np.random.PCG64(n))
        o = np.stack([rng.uniform(-0.2, 1.2, n), rng.uniform(-0.2, 1.2, n), np.full(n, 1.0)], 1)
        r = make_rays(o, np.tile([0, 0, -1], (n, 1)))
        ref = oracle.bruteforce([m], r)
        got = gpu_device.trace_numpy(gpu_device.scene_bvh([m]), r)
        st = compare(ref, got)
        assert_parity(st)


def test_heightfield_bounce_parity(gpu_device):
    cells = 256
    _case("heightfield_small", cells=cells)
    meshes = scenes.heightfield_scene(cells=cells, n_spheres=8, seed=3)
    prim = scenes.pinhole_rays(320, 180)
    ob = oracle.OracleBVH(meshes, builder="SAH")
    sc = gpu_device.scene_bvh(meshes)
    for name, r in (("primary", prim),):
        ref = ob.intersect(r)
        got = gpu_device.trace_numpy(sc, r)
        st = compare(ref, got, tri_tie_checker(oracle, meshes, r, ref, got))
        print(name, st, flush=True)
        assert_parity(st)
    bounce = scenes.bounce_rays(meshes, prim, ref)
    ref_b = ob.intersect(bounce)
    got_b = gpu_device.trace_numpy(sc, bounce)
    st = compare(ref_b, got_b, tri_tie_checker(oracle, meshes, bounce, ref_b, got_b))
    print("bounce", st, flush=True)
    assert_parity(st)
    shadow = scenes.shadow_rays(meshes, prim, ref)
    ref_s = ob.intersect(shadow)
    got_s = gpu_device.trace_numpy(sc, shadow, query=1)
    assert np.array_equal(ref_s["meshIndex"] == NULL_INDEX, got_s["meshIndex"] == NULL_INDEX)


def test_stack_depths_give_identical_hits(gpu_device):
    """The per-lane stack keeps its top ldsStack entries in an LDS ring and
    spills the rest: every depth and launch shape must give the default's
    hits bit for bit (deep traversals spill)."""
    _case("stack_depths")
    meshes = scenes.heightfield_scene(cells=256, n_spheres=8, seed=3)
    prim = scenes.pinhole_rays(320, 180)
    sc = gpu_device.scene_bvh(meshes)
    h = gpu_device.trace_numpy(sc, prim)
    rays = np.concatenate([scenes.bounce_rays(meshes, prim, h), scenes.plain_rays(20000, seed=9)])
    base = gpu_device.trace_numpy(sc, rays)
    for block, lds, refill in ((256, 4, 8), (256, 16, 8), (128, 32, 16), (64, 64, 4), (256, 4, 1)):
        o = D.trace_opts(block, lds, 0, refill)
        for q in (0, 1):
            got = gpu_device.trace_numpy(sc, rays, query=q, opts=o)
            ref = base if q == 0 else gpu_device.trace_numpy(sc, rays, query=1)
            if q == 0:
                assert np.array_equal(got, ref), (block, lds, refill)
            else:
                assert np.array_equal(got["meshIndex"] == NULL_INDEX, ref["meshIndex"] == NULL_INDEX)
    sc.close()


def test_full_size_plain_properties(gpu_device):
    """1,048,576 rays on the 101,932-tri scene: properties that hold at any
    size (determinism, layout independence, any-hit/closest agreement)."""
    _case("plain_full", rays=1 << 20)
    meshes = scenes.plain_scene(seed=1)
    rays = scenes.plain_rays(1 << 20, seed=2)
    s0 = gpu_device.scene_bvh(meshes, layout=0)
    s1 = gpu_device.scene_bvh(meshes, layout=1, max_leaf=2)
    a = gpu_device.trace_numpy(s0, rays)
    b = gpu_device.trace_numpy(s0, rays)
    assert np.array_equal(a, b)
    c = gpu_device.trace_numpy(s1, rays)
    st = compare(a, c)
    print(st, flush=True)
    assert st["bad"] == 0 and st["mismatch"] - st["edge_band"] < 1e-5 * rays.shape[0] + 5
    anyh = gpu_device.trace_numpy(s0, rays, query=1)
    assert np.array_equal(a["meshIndex"] == NULL_INDEX, anyh["meshIndex"] == NULL_INDEX)
    # a random subset against the oracle
    sel = np.random.Generator(np.random.PCG64(5)).choice(rays.shape[0], 8192, replace=False)
    ref = oracle.OracleBVH(meshes).intersect(rays[sel])
    st = compare(ref, a[sel], tri_tie_checker(oracle, meshes, rays[sel], ref, a[sel]))
    assert_parity(st)


def test_full_size_plain_against_oracle_digests(gpu_device):
    """Every one of the 1,048,576 rays of the plain workload against the
    oracle: the device's hit records, chunk by chunk (4,096 rays), must equal
    the digests of the oracle's records (tests/golden/plain_full_digest.json,
    written by tests/golden/make_plain_full_digest.py); a chunk that differs
    is traced by the oracle here and held to the parity rule (edge band,
    exact ties)."""
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plain_full_digest.json")) as f:
        gold = json.load(f)
    _case("plain_full_digest", rays=gold["n_rays"])
    meshes = scenes.plain_scene(seed=1)
    rays = scenes.plain_rays(gold["n_rays"], seed=2)
    assert sum(m.ntris for m in meshes) == gold["triangles"]
    sc = gpu_device.scene_bvh(meshes)
    got = gpu_device.trace_numpy(sc, rays)
    sc.close()
    chunk = gold["chunk"]
    diff = [i for i, (a, b) in enumerate(zip(chunk_digests(got, chunk), gold["sha1"])) if a != b]
    print(f"{len(diff)} of {len(gold['sha1'])} chunks differ from the oracle's; hits {(got['meshIndex'] != NULL_INDEX).sum()}"
          f" (oracle {gold['hits']})", flush=True)
    ob = oracle.OracleBVH(meshes) if diff else None
    for i in diff:
        sl = slice(i * chunk, (i + 1) * chunk)
        ref = ob.intersect(rays[sl])
        st = compare(ref, got[sl], tri_tie_checker(oracle, meshes, rays[sl], ref, got[sl]))
        print(i, st, flush=True)
        assert_parity(st)
    # ties and edge-band rays are rare: a systematic difference is not
    assert len(diff) <= 8
