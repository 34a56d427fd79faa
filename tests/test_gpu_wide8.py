"""8-wide nodes (layout 2: one 128-byte L2 line per node, eight children
with 8-bit bounds, sorted by entry distance with a 19-comparator network):
the same hits as the oracle to the parity bar, the same hits for every
launch shape, fewer node visits than the 4-wide tree, the no-alloc entry
point equal to lrh_trace, and batches over 8-wide scenes equal to plain
traces (device builds of 8-wide trees are tested with the device builds)."""
import numpy as np
import pytest

import oracle
from luxcore_amd import device as D
from luxcore_amd import scenes
from luxcore_amd.raytypes import HIT_DTYPE, NULL_INDEX
from tests.parity import assert_parity, compare, tri_tie_checker

pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.fixture(scope="module")
def hf():
    meshes = scenes.heightfield_scene(cells=256, n_spheres=8, seed=3)
    prim = scenes.pinhole_rays(320, 180)
    ob = oracle.OracleBVH(meshes, builder="SAH")
    ref_p = ob.intersect(prim)
    bounce = scenes.bounce_rays(meshes, prim, ref_p)
    return meshes, prim, ob, ref_p, bounce


def test_wide8_heightfield_against_oracle(gpu_device, hf):
    meshes, prim, ob, ref_p, bounce = hf
    _case("wide8_heightfield")
    sc = gpu_device.scene_bvh(meshes, layout=2)
    assert sc.info()["layout"] == 2
    for name, r, ref in (("primary", prim, ref_p), ("bounce", bounce, ob.intersect(bounce))):
        got = gpu_device.trace_numpy(sc, r)
        st = compare(ref, got, tri_tie_checker(oracle, meshes, r, ref, got))
        print(name, st, flush=True)
        assert_parity(st)
    shadow = scenes.shadow_rays(meshes, prim, ref_p)
    ref_s = ob.intersect(shadow)
    got_s = gpu_device.trace_numpy(sc, shadow, query=1)
    assert np.array_equal(ref_s["meshIndex"] == NULL_INDEX, got_s["meshIndex"] == NULL_INDEX)
    sc.close()


def test_wide8_launch_shapes_and_masks(gpu_device, hf):
    """Every LDS ring depth and launch shape (a share of the persistent grid
    included) gives the default's hits bit for bit (8-wide visits push up to
    seven entries: deep traversals spill); masked rays keep their hit
    records; a grid share above 100% is refused."""
    meshes, prim, ob, ref_p, bounce = hf
    _case("wide8_shapes")
    sc = gpu_device.scene_bvh(meshes, layout=2)
    rays = np.concatenate([bounce, scenes.plain_rays(20000, seed=9)])
    rays["flags"][::17] = 1
    init = np.zeros(rays.shape[0], HIT_DTYPE)
    init["t"] = -3.0
    init["meshIndex"] = 0xABCDEF
    base = gpu_device.trace_numpy(sc, rays, hits_init=init)
    masked = (rays["flags"] & 1) != 0
    assert np.all(base["t"][masked] == -3.0) and np.all(base["meshIndex"][masked] == 0xABCDEF)
    for block, lds, refill, pct in ((256, 4, 8, 0), (256, 16, 8, 0), (128, 32, 16, 0), (64, 64, 4, 0),
                                    (256, 4, 1, 0), (256, 8, 12, 80), (64, 8, 8, 1)):
        o = D.trace_opts(block, lds, 0, refill, grid_percent=pct)
        got = gpu_device.trace_numpy(sc, rays, opts=o, hits_init=init)
        assert np.array_equal(got, base), (block, lds, refill, pct)
    with pytest.raises(D.LrhError, match="gridPercent"):
        gpu_device.trace_numpy(sc, rays[:64], opts=D.trace_opts(grid_percent=101))
    sc.close()


def test_wide8_fewer_visits_same_hits(gpu_device, hf):
    meshes, prim, ob, ref_p, bounce = hf
    _case("wide8_stats")
    s4 = gpu_device.scene_bvh(meshes, layout=0)
    s8 = gpu_device.scene_bvh(meshes, layout=2)
    n = bounce.shape[0]
    rb = gpu_device.upload(bounce)
    hb = gpu_device.alloc(n * HIT_DTYPE.itemsize)
    v4, p4, c4 = gpu_device.trace_stats(s4, rb, hb, n)
    h4 = gpu_device.download(hb, HIT_DTYPE, n)
    v8, p8, c8 = gpu_device.trace_stats(s8, rb, hb, n)
    h8 = gpu_device.download(hb, HIT_DTYPE, n)
    print("visits per ray 4-wide", v4 / c4, "8-wide", v8 / c8, "tests", p4 / c4, p8 / c8, flush=True)
    assert c4 == c8 == n
    assert v8 < 0.8 * v4
    st = compare(h4, h8)
    print(st, flush=True)
    assert_parity(st)
    assert s8.info()["nodeCount"] < s4.info()["nodeCount"]


def test_wide8_noalloc_equals_trace(gpu_device):
    _case("wide8_noalloc")
    meshes = scenes.plain_scene(seed=1, n_random=20000)
    sc = gpu_device.scene_bvh(meshes, layout=2)
    for n in (1 << 17, (1 << 17) - 999):  # second call: fewer rays
        rays = scenes.plain_rays(n, seed=40 + n)
        rb = gpu_device.upload(rays)
        ha = gpu_device.upload(np.zeros(n, HIT_DTYPE))
        hbuf = gpu_device.upload(np.zeros(n, HIT_DTYPE))
        D.check(D.lib().lrh_trace(sc.h, rb.ptr, ha.ptr, n, 0, None))
        D.check(D.lib().lrh_trace_noalloc(sc.h, rb.ptr, hbuf.ptr, n, 0, None))
        gpu_device.sync()
        assert gpu_device.download(ha, HIT_DTYPE, n).tobytes() == gpu_device.download(hbuf, HIT_DTYPE, n).tobytes()
    sc.close()


def test_wide8_batched_equals_plain_traces(gpu_device):
    """A batch over 8-wide scenes, single- and two-level, gives each buffer
    the hits of lrh_trace on it, closest and any hit; a batch mixing node
    layouts is refused."""
    _case("wide8_batched")
    meshes = [scenes.plain_scene(seed=1, n_random=3000), scenes.heightfield_scene(cells=128, n_spheres=2, seed=3)]
    scs = [gpu_device.scene_bvh(m, layout=2) for m in meshes]
    inst = scenes.instanced_scene(n_inst=40, motion_frac=0.25, freq=10, extent=8.0, seed=31)
    scs.append(gpu_device.scene_mbvh(inst, layout=2))  # two-level: the batch runs the two-level kernel
    items, refs = [], []
    for i, (si, n, q) in enumerate(((0, 30000, 0), (1, 20000, 1), (1, 45000, 0), (0, 7000, 1), (2, 25000, 0),
                                    (2, 9000, 1))):
        rays = scenes.plain_rays(n, seed=70 + i) if si < 2 else scenes.instanced_rays(n, extent=8.0, seed=70 + i)
        rb = gpu_device.upload(rays)
        hb = gpu_device.alloc(n * HIT_DTYPE.itemsize)
        items.append((scs[si], rb, hb, n, q))
        refs.append(gpu_device.trace_numpy(scs[si], rays, query=q))
    gpu_device.trace_batch(items)
    gpu_device.sync()
    for (sc, rb, hb, n, q), ref in zip(items, refs):
        got = gpu_device.download(hb, HIT_DTYPE, n)
        if q == 0:
            assert got.tobytes() == ref.tobytes()
        else:
            assert np.array_equal(got["meshIndex"] == NULL_INDEX, ref["meshIndex"] == NULL_INDEX)
    s4 = gpu_device.scene_bvh(meshes[0], layout=0)
    with pytest.raises(D.LrhError, match="layout"):
        gpu_device.trace_batch([items[0], (s4,) + items[2][1:]])
    for s in scs + [s4]:
        s.close()
