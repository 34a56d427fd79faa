"""The device linear-BVH build's algorithm on the CPU: a sequential emulation
(tests/native/lbvh_emu.cpp) of the kernels of luxcore_amd/csrc/lrh_gpubuild.hip,
running the same per-element steps (luxcore_amd/csrc/lrh_lbvh.h) and the same
stable radix sort.  It must build a well-formed tree whose leaves hold every
triangle once in boxes that contain them, with records equal to the host
builder's; and codes out of order -- what a broken sort would leave -- must
be reported through the error word (sorted-order check, radix-tree range
search) and must not make any step loop forever.  The clustering method
(PLOC: rounds of mutual nearest neighbours among the Morton-ordered
clusters) is held to the same bar, must end on inputs where the neighbour
rounds stall (every box equal), and must give a tree of lower surface-area
cost than the radix tree's."""
import os
import subprocess

import numpy as np
import pytest

from luxcore_amd import device as D
from luxcore_amd import scenes
from luxcore_amd.raytypes import machine_epsilon
from tests.test_abi_cpu import NODE_F, TRI, check_tree, decode_q, node_dtype

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ERR_UNSORTED, ERR_RANGE = 1, 2


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    d = tmp_path_factory.mktemp("lbvh")
    exe = d / "lbvh_emu"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe),
                    os.path.join(ROOT, "tests", "native", "lbvh_emu.cpp")], check=True, timeout=180)
    return exe


def write_meshes(path, meshes):
    with open(path, "wb") as f:
        f.write(np.uint32(len(meshes)).tobytes())
        for m in meshes:
            v = np.ascontiguousarray(m.verts, np.float32)
            t = np.ascontiguousarray(m.tris, np.uint32)
            f.write(np.array([v.shape[0], t.shape[0]], np.uint32).tobytes())
            f.write(v.tobytes())
            f.write(t.tobytes())


def run_emu(emu, tmp_path, meshes, max_leaf, layout, mode="normal", method="lbvh"):
    inp = tmp_path / "in.bin"
    write_meshes(inp, meshes)
    out = tmp_path / f"out_{max_leaf}_{layout}_{mode}_{method}"
    out.mkdir()
    p = subprocess.run([str(emu), str(inp), str(out), str(max_leaf), str(layout), mode, method],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    f = (out / "info.txt").read_text().split()
    info = {f[i]: int(f[i + 1]) for i in range(0, len(f), 2)}
    nodes = np.fromfile(out / "nodes.bin", node_dtype(layout))
    tris = np.fromfile(out / "tris.bin", TRI)
    return info, nodes, tris


def leaf_boxes(meshes, tris):
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


def by_id(tris):
    return tris[np.lexsort((tris["tri"], tris["mesh"]))]


def sah_cost(nodes, layout, c_node=1.0, c_tri=1.0):
    """Surface-area cost of a 4-wide tree: each child box's area over the root
    box's, times c_node for an inner child or c_tri per triangle of a leaf."""
    if layout == 0:
        dec = decode_q(nodes)
    else:
        dec = {k: nodes[k] for k in ("lo_x", "hi_x", "lo_y", "hi_y", "lo_z", "hi_z")}
    ref = nodes["child"].astype(np.int64)
    used = ref != 0xFFFFFFFF
    d = [(dec["hi_" + a] - dec["lo_" + a]).astype(np.float64) for a in "xyz"]
    area = 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])
    leaf = (ref & 0x80000000) != 0
    weight = np.where(leaf, c_tri * (((ref >> 28) & 7) + 1), c_node)
    lo = [np.where(used[0], dec["lo_" + a][0], np.inf).min() for a in "xyz"]
    hi = [np.where(used[0], dec["hi_" + a][0], -np.inf).max() for a in "xyz"]
    e = [hi[k] - lo[k] for k in range(3)]
    root = 2.0 * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0])
    return float((area * weight)[used].sum() / root)


@pytest.mark.parametrize("method", ["lbvh", "ploc"])
@pytest.mark.parametrize("max_leaf,layout", [(4, 0), (1, 0), (8, 1), (4, 2), (1, 2)])
def test_emulated_device_build_is_a_valid_tree(emu, tmp_path, max_leaf, layout, method):
    meshes = scenes.plain_scene(seed=1, n_random=2000)
    info, nodes, tris = run_emu(emu, tmp_path, meshes, max_leaf, layout, method=method)
    assert info["err"] == 0 and info["nodes"] == nodes.shape[0] > 1
    n = sum(m.ntris for m in meshes)
    assert tris.shape[0] == n
    lo, hi = leaf_boxes(meshes, tris)

    def boxes_of_child(start, cnt):
        return lo[start:start + cnt].min(0), hi[start:start + cnt].max(0)
    check_tree(nodes, boxes_of_child, 0, nodes.shape[0], n, layout)
    # the records are the host builder's, one per triangle
    host = D.host_scene_bvh(meshes, layout=layout, max_leaf=max_leaf, threads=4, split_budget=-1.0)
    _, hp = host.download()
    assert np.array_equal(by_id(tris).view(np.uint8), by_id(hp.view(TRI)).view(np.uint8))
    # a binary radix tree over 63-bit codes collapsed 4-wide stays shallow,
    # and so does the clustering's tree of this scene
    assert info["depth"] < 40
    if method == "ploc":
        assert 2 <= info["rounds"] < 200


def test_clustering_tree_costs_less_than_radix_tree(emu, tmp_path):
    meshes = scenes.plain_scene(seed=1, n_random=2000)
    costs = {}
    for method in ("lbvh", "ploc"):
        info, nodes, _ = run_emu(emu, tmp_path, meshes, 4, 1, method=method)
        assert info["err"] == 0
        costs[method] = sah_cost(nodes, 1)
    host = D.host_scene_bvh(meshes, layout=1, max_leaf=4, threads=4, split_budget=-1.0)
    hn, _ = host.download()
    costs["sah"] = sah_cost(hn.view(NODE_F), 1)
    # (this scene: radix tree 1144, clustering 836, host SAH sweep 1028)
    assert costs["ploc"] < 0.8 * costs["lbvh"], costs
    assert costs["ploc"] < 1.1 * costs["sah"], costs


def test_clustering_of_equal_boxes_ends(emu, tmp_path):
    # every box equal: every merged area ties, the neighbour rounds merge one
    # pair each, and the forced rounds that follow them halve the clusters
    # (below 64 clusters one merge is progress enough: at most 63 more rounds)
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    n = 3000
    meshes = [scenes.Mesh(np.tile(v, (n, 1)), np.arange(3 * n, dtype=np.uint32).reshape(n, 3))]
    info, nodes, tris = run_emu(emu, tmp_path, meshes, 4, 0, method="ploc")
    assert info["err"] == 0 and tris.shape[0] == n
    assert info["rounds"] <= 2 * int(np.ceil(np.log2(n))) + 63, info
    lo, hi = leaf_boxes(meshes, tris)
    check_tree(nodes, lambda s, c: (lo[s:s + c].min(0), hi[s:s + c].max(0)), 0, nodes.shape[0], n, 0)


def test_small_scene_is_one_leaf(emu, tmp_path):
    meshes = [scenes.box_mesh()]
    info, nodes, tris = run_emu(emu, tmp_path, [scenes.Mesh(meshes[0].verts, meshes[0].tris[:3])], 4, 0)
    assert info["err"] == 0 and info["nodes"] == 1 and tris.shape[0] == 3
    assert nodes["child"][0][0] == 0x80000000 | (2 << 28)


@pytest.mark.parametrize("mode", ["unsorted", "reversed"])
def test_codes_out_of_order_are_reported_and_terminate(emu, tmp_path, mode):
    meshes = scenes.plain_scene(seed=1, n_random=2000)
    info, nodes, tris = run_emu(emu, tmp_path, meshes, 4, 0, mode)
    assert info["err"] & ERR_UNSORTED, info
    if mode == "unsorted":
        # the radix tree's range search fails instead of looping
        assert info["err"] & ERR_RANGE, info
    else:
        # descending codes are monotone: the radix tree still finds its ranges,
        # so only the explicit order check reports them
        assert not info["err"] & ERR_RANGE, info
    assert tris.shape[0] == 0               # nothing is written after an error
