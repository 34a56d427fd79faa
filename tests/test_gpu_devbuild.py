"""Device-built trees (LRH_BUILD_DEVICE_LBVH, LRH_BUILD_DEVICE_PLOC): the
same 4-wide records as the host builder, a well-formed tree over every
triangle, deterministic, byte for byte the sequential CPU emulation of the
same kernels (tests/native/lbvh_emu.cpp), and traces equal to the oracle to
the parity bar (closest hit does not depend on the tree outside exact ties
and the edge band)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from luxcore_amd import device as D
from luxcore_amd import scenes
from luxcore_amd.raytypes import HIT_DTYPE, NULL_INDEX, machine_epsilon
from tests.parity import assert_parity, compare
from tests.test_abi_cpu import TRI, check_tree, node_dtype

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILDERS = {"lbvh": D.BUILD_DEVICE_LBVH, "ploc": D.BUILD_DEVICE_PLOC}


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _leaf_boxes(meshes, tris):
    lo = np.empty((tris.shape[0], 3), np.float32)
    hi = np.empty_like(lo)
    for mi, m in enumerate(meshes):
        sel = tris["mesh"] == mi
        if not sel.any():
            continue
        v = m.verts[m.tris[tris["tri"][sel]]]
        lo[sel] = v.min(axis=1)
        hi[sel] = v.max(axis=1)
    e = np.maximum(machine_epsilon(lo).max(1), machine_epsilon(hi).max(1))
    return (lo - e[:, None]).astype(np.float32), (hi + e[:, None]).astype(np.float32)


def _records_sorted(tris):
    order = np.lexsort((tris["tri"], tris["mesh"]))
    return tris[order]


@pytest.mark.parametrize("builder", list(BUILDERS))
@pytest.mark.parametrize("layout,max_leaf", [(0, 4), (0, 1), (0, 8), (1, 4), (2, 4)])
def test_device_build_tree_and_records(gpu_device, layout, max_leaf, builder):
    _case("devbuild_tree", layout=layout, max_leaf=max_leaf, builder=builder)
    meshes = scenes.plain_scene(seed=1, n_random=20000)
    dsc = gpu_device.scene_bvh(meshes, layout=layout, max_leaf=max_leaf, builder=BUILDERS[builder])
    info = dsc.info()
    nb, pb = dsc.download()
    nodes = nb.view(node_dtype(layout))
    tris = pb.view(TRI)
    ntri = sum(m.ntris for m in meshes)
    assert tris.shape[0] == ntri and nodes.shape[0] == info["nodeCount"]
    # records: the host builder's records, in another order
    hsc = D.host_scene_bvh(meshes, layout=layout, max_leaf=max_leaf, split_budget=-1.0)
    _, hpb = hsc.download()
    assert _records_sorted(tris).tobytes() == _records_sorted(hpb.view(TRI)).tobytes()
    # every triangle once; every child box contains its triangles' epsilon-expanded boxes
    lo, hi = _leaf_boxes(meshes, tris)

    def boxes_of_child(s, c):
        return lo[s:s + c].min(0), hi[s:s + c].max(0)
    check_tree(nodes, boxes_of_child, 0, nodes.shape[0], ntri, layout)
    # deterministic
    dsc2 = gpu_device.scene_bvh(meshes, layout=layout, max_leaf=max_leaf, builder=BUILDERS[builder])
    nb2, pb2 = dsc2.download()
    assert nb.tobytes() == nb2.tobytes() and pb.tobytes() == pb2.tobytes()
    dsc.close()
    dsc2.close()


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    exe = tmp_path_factory.mktemp("emu") / "lbvh_emu"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe),
                    os.path.join(ROOT, "tests", "native", "lbvh_emu.cpp")], check=True, timeout=180)
    return exe


@pytest.mark.parametrize("builder", list(BUILDERS))
@pytest.mark.parametrize("scene,layout,max_leaf", [("plain", 0, 4), ("plain", 1, 2), ("heightfield", 0, 4),
                                                   ("heightfield", 2, 4)])
def test_device_build_equals_cpu_emulation(gpu_device, emu, tmp_path, scene, layout, max_leaf, builder):
    """The kernels' tree is the sequential emulation's, node and record bytes
    alike: the per-element steps are shared and every step is deterministic
    (the heightfield's centre keys span the whole unsigned range, so a bound
    reduced at less than 32 bits would show)."""
    from tests.test_lbvh_cpu import run_emu
    _case("devbuild_emu", scene=scene, layout=layout, max_leaf=max_leaf, builder=builder)
    if scene == "plain":
        meshes = scenes.plain_scene(seed=1, n_random=20000)
    else:
        meshes = scenes.heightfield_scene(cells=256, n_spheres=4, seed=3)
    info, nodes, tris = run_emu(emu, tmp_path, meshes, max_leaf, layout, method=builder)
    assert info["err"] == 0
    dsc = gpu_device.scene_bvh(meshes, layout=layout, max_leaf=max_leaf, builder=BUILDERS[builder])
    nb, pb = dsc.download()
    assert dsc.info()["nodeCount"] == info["nodes"]
    assert nb.tobytes() == nodes.tobytes()
    assert pb.tobytes() == tris.tobytes()
    dsc.close()


@pytest.mark.parametrize("builder", list(BUILDERS))
@pytest.mark.parametrize("layout", [0, 1, 2])
def test_device_build_trace_parity(gpu_device, layout, builder):
    _case("devbuild_parity", layout=layout, builder=builder)
    meshes = scenes.plain_scene(seed=1, n_random=20000)
    rays = scenes.plain_rays(1 << 18, seed=21)
    rays["flags"][::11] = 1
    dsc = gpu_device.scene_bvh(meshes, layout=layout, builder=BUILDERS[builder])
    init = np.zeros(rays.shape[0], HIT_DTYPE)
    init["t"] = 3.25
    got = gpu_device.trace_numpy(dsc, rays, hits_init=init)
    live = (rays["flags"] & 1) == 0
    assert np.all(got["t"][~live] == 3.25)
    ref = oracle.OracleBVH(meshes).intersect(rays[live])
    st = compare(ref, got[live])
    print(st, flush=True)
    assert_parity(st)
    anyh = gpu_device.trace_numpy(dsc, rays[live], query=D.LRH_QUERY_ANY)
    assert np.array_equal(anyh["meshIndex"] != NULL_INDEX, ref["meshIndex"] != NULL_INDEX)
    dsc.close()


@pytest.mark.parametrize("builder", list(BUILDERS))
def test_device_build_small_and_ragged(gpu_device, builder):
    _case("devbuild_small", builder=builder)
    tri = scenes.Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32))
    empty = scenes.Mesh(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    box = scenes.box_mesh()
    rng = np.random.default_rng(5)
    # 300 copies of one triangle: every box equal (the clustering's stalled rounds)
    same = scenes.Mesh(np.tile(tri.verts, (300, 1)), np.arange(900, dtype=np.uint32).reshape(300, 3))
    for meshes in ([tri], [empty, tri], [box], [empty, box, empty, tri],
                   [scenes.random_triangles(9, (0, 0, 0), (10, 10, 10), rng)], [same, box]):
        for max_leaf in (1, 4, 8):
            dsc = gpu_device.scene_bvh(meshes, max_leaf=max_leaf, builder=BUILDERS[builder])
            rays = scenes.random_rays(4096, (-1, -1, -1), (11, 11, 11), rng)
            got = gpu_device.trace_numpy(dsc, rays)
            ref = oracle.bruteforce(meshes, rays)
            assert_parity(compare(ref, got))
            dsc.close()
    dsc = gpu_device.scene_bvh([empty], builder=BUILDERS[builder])
    got = gpu_device.trace_numpy(dsc, scenes.plain_rays(256, seed=3))
    assert np.all(got["meshIndex"] == NULL_INDEX)
    dsc.close()


@pytest.mark.parametrize("builder", list(BUILDERS))
def test_device_build_heightfield_equals_host_tree(gpu_device, builder):
    """Full 26M-triangle scene: the device-built tree's closest hits equal the
    host-built tree's (same rays) outside exact ties and the edge band."""
    _case("devbuild_heightfield", builder=builder)
    meshes = scenes.heightfield_scene()
    hsc = gpu_device.scene_bvh(meshes)
    dsc = gpu_device.scene_bvh(meshes, builder=BUILDERS[builder])
    print("host build", hsc.info()["buildSeconds"], "device build", dsc.info()["buildSeconds"],
          "upload", dsc.info()["uploadSeconds"], flush=True)
    prim = scenes.pinhole_rays(960, 540)
    h = gpu_device.trace_numpy(hsc, prim)
    bounce = scenes.bounce_rays(meshes, prim, h)
    a = gpu_device.trace_numpy(hsc, bounce)
    b = gpu_device.trace_numpy(dsc, bounce)
    st = compare(a, b)
    print(st, flush=True)
    assert_parity(st)
    hsc.close()
    dsc.close()
