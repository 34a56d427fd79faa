"""The C-ABI library loads and exports every function include/lrhip.h
declares; the host-side builder produces well-formed device records
(CPU only: no device call is made)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from luxcore_amd import device as D
from luxcore_amd import scenes
from luxcore_amd.raytypes import machine_epsilon

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE_Q = np.dtype([("org", "<f4", 3), ("e", "u1", 3), ("nchild", "u1"), ("qlo_x", "u1", 4), ("qhi_x", "u1", 4),
                   ("qlo_y", "u1", 4), ("qhi_y", "u1", 4), ("qlo_z", "u1", 4), ("qhi_z", "u1", 4),
                   ("child", "<u4", 4), ("pad", "<u4", 2)])
NODE_Q8 = np.dtype([("org", "<f4", 3), ("e", "u1", 3), ("nchild", "u1"), ("qlo_x", "u1", 8), ("qhi_x", "u1", 8),
                    ("qlo_y", "u1", 8), ("qhi_y", "u1", 8), ("qlo_z", "u1", 8), ("qhi_z", "u1", 8),
                    ("child", "<u4", 8), ("pad", "<u4", 8)])
NODE_F = np.dtype([("lo_x", "<f4", 4), ("hi_x", "<f4", 4), ("lo_y", "<f4", 4), ("hi_y", "<f4", 4),
                   ("lo_z", "<f4", 4), ("hi_z", "<f4", 4), ("child", "<u4", 4), ("pad", "<u4", 4)])
TRI = np.dtype([("p0", "<f4", 3), ("e1", "<f4", 3), ("e2", "<f4", 3), ("tri", "<u4"), ("mesh", "<u4"),
                ("pad", "<u4")])
INST = np.dtype([("root", "<u4"), ("kind", "<u4"), ("index", "<u4"), ("object", "<u4")])
assert NODE_Q.itemsize == 64 and NODE_F.itemsize == 128 and NODE_Q8.itemsize == 128 and TRI.itemsize == 48


def node_dtype(layout):
    """Node record of layout 0 (4-wide quantized), 1 (4-wide fp32), 2 (8-wide quantized)."""
    return {0: NODE_Q, 1: NODE_F, 2: NODE_Q8}[layout]


def header_functions():
    with open(os.path.join(ROOT, "include", "lrhip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(lrh_[a-z_0-9]+)\(", text, re.M)))


def test_library_exports_every_declared_function():
    names = header_functions()
    assert len(names) >= 30
    L = D.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(D.SIGNATURES), set(names) ^ set(D.SIGNATURES)
    assert D.lib().lrh_version().startswith(b"lrhip")


def test_bad_arguments_fail_cleanly():
    assert D.lib().lrh_autotune_build(None, None, None, 0, None, None, 0, None, 0, None, 0, None, None,
                                      None) == D.LRH_EINVAL
    assert D.lib().lrh_trace(None, None, None, 0, 0, None) == -1
    assert D.lib().lrh_scene_get_info(None, None) == -1
    assert b"null" in D.lib().lrh_last_error()
    with pytest.raises(D.LrhError):
        D.host_scene_bvh([scenes.Mesh(np.zeros((3, 3), np.float32), np.array([[0, 1, 5]], np.uint32))])
    with pytest.raises(D.LrhError):
        D.host_scene_bvh(scenes.plain_scene(n_random=10), max_leaf=9)


def decode_q(nodes):
    s = (nodes["e"].astype(np.uint32) << 23).view(np.float32)  # (N,3)
    out = {}
    for k, ax in enumerate("xyz"):
        o = nodes["org"][:, k:k + 1]
        sc = s[:, k:k + 1]
        out["lo_" + ax] = (o + nodes["qlo_" + ax].astype(np.float32) * sc).astype(np.float32)
        out["hi_" + ax] = (o + nodes["qhi_" + ax].astype(np.float32) * sc).astype(np.float32)
    return out


def check_tree(nodes, boxes_of_child, root, n_nodes, n_prims, layout):
    """Walk the tree: every primitive reached once, every child box contains
    the epsilon-expanded boxes of what it holds."""
    if layout != 1:
        dec = decode_q(nodes)
    else:
        dec = {k: nodes[k] for k in ("lo_x", "hi_x", "lo_y", "hi_y", "lo_z", "hi_z")}
    count = np.zeros(n_prims, np.int64)
    stack = [root]
    while stack:
        i = stack.pop()
        assert 0 <= i < n_nodes
        for c in range(nodes["child"].shape[1]):
            ref = int(nodes["child"][i][c])
            if ref == 0xFFFFFFFF:
                continue
            lo = np.array([dec["lo_x"][i][c], dec["lo_y"][i][c], dec["lo_z"][i][c]])
            hi = np.array([dec["hi_x"][i][c], dec["hi_y"][i][c], dec["hi_z"][i][c]])
            if ref & 0x80000000:
                start = ref & 0x0FFFFFFF
                cnt = ((ref >> 28) & 7) + 1
                count[start:start + cnt] += 1
                blo, bhi = boxes_of_child(start, cnt)
                assert np.all(lo <= blo) and np.all(hi >= bhi), (i, c)
            else:
                stack.append(ref)
    assert np.all(count == 1)


@pytest.mark.parametrize("layout,max_leaf", [(0, 4), (0, 1), (0, 8), (1, 4), (1, 2), (2, 4), (2, 1), (2, 8)])
def test_host_builder_records(layout, max_leaf):
    # one record per triangle (early split clipping off), boxes of whole triangles
    meshes = scenes.plain_scene(seed=1, n_random=2000)
    sc = D.host_scene_bvh(meshes, layout=layout, max_leaf=max_leaf, threads=4, split_budget=-1.0)
    info = sc.info()
    nb, pb = sc.download()
    nodes = nb.view(node_dtype(layout))
    tris = pb.view(TRI)
    assert nodes.shape[0] == info["nodeCount"] and tris.shape[0] == sum(m.ntris for m in meshes)
    # records carry p0 and the fp32 edges of the original triangle
    for i in range(0, tris.shape[0], 97):
        m = meshes[tris["mesh"][i]]
        v = m.verts[m.tris[tris["tri"][i]]]
        assert np.array_equal(tris["p0"][i], v[0])
        assert np.array_equal(tris["e1"][i], (v[1] - v[0]).astype(np.float32))
        assert np.array_equal(tris["e2"][i], (v[2] - v[0]).astype(np.float32))
    # epsilon-expanded boxes of the original triangles, in record order
    lo = np.empty((tris.shape[0], 3), np.float32)
    hi = np.empty_like(lo)
    for mi, m in enumerate(meshes):
        sel = tris["mesh"] == mi
        v = m.verts[m.tris[tris["tri"][sel]]]
        lo[sel] = v.min(axis=1)
        hi[sel] = v.max(axis=1)
    e = np.maximum(machine_epsilon(lo).max(1), machine_epsilon(hi).max(1))
    lo = (lo - e[:, None]).astype(np.float32)
    hi = (hi + e[:, None]).astype(np.float32)

    def boxes_of_child(s, c):
        return lo[s:s + c].min(0), hi[s:s + c].max(0)
    check_tree(nodes, boxes_of_child, 0, nodes.shape[0], tris.shape[0], layout)
    assert info["maxDepth"] < 40


def _grid(k):
    x, y = np.meshgrid(np.arange(k + 1, dtype=np.float32), np.arange(k + 1, dtype=np.float32))
    v = np.stack([x.ravel(), y.ravel(), np.zeros(x.size, np.float32)], 1)
    i = np.arange(k * (k + 1)).reshape(k, k + 1)[:, :k].ravel()
    t = np.concatenate([np.stack([i, i + 1, i + k + 2], 1), np.stack([i, i + k + 2, i + k + 1], 1)])
    return v, t.astype(np.uint32)


def leaf_boxes_of_records(nodes, root, n_prims, layout):
    """(lo, hi) of the leaf child slot that holds each record."""
    dec = decode_q(nodes) if layout != 1 else {k: nodes[k] for k in ("lo_x", "hi_x", "lo_y", "hi_y", "lo_z", "hi_z")}
    lo = np.full((n_prims, 3), np.nan, np.float32)
    hi = np.full((n_prims, 3), np.nan, np.float32)
    stack = [root]
    while stack:
        i = stack.pop()
        for c in range(nodes["child"].shape[1]):
            ref = int(nodes["child"][i][c])
            if ref == 0xFFFFFFFF:
                continue
            if ref & 0x80000000:
                start, cnt = ref & 0x0FFFFFFF, ((ref >> 28) & 7) + 1
                lo[start:start + cnt] = [dec["lo_x"][i][c], dec["lo_y"][i][c], dec["lo_z"][i][c]]
                hi[start:start + cnt] = [dec["hi_x"][i][c], dec["hi_y"][i][c], dec["hi_z"][i][c]]
            else:
                stack.append(ref)
    return lo, hi


@pytest.mark.parametrize("layout,budget", [(0, 0.0), (0, 4.0), (1, 1.0), (2, 4.0)])
def test_split_triangles_cover_their_triangles(layout, budget):
    """Early split clipping: triangles whose box is far above the median are
    cut into references with tighter boxes.  Every triangle keeps at least
    one record with its exact data, and every point of every triangle lies
    in the leaf box of one of its records (the culling stays conservative);
    triangles of typical size are not cut."""
    rng = np.random.default_rng(5)
    meshes = scenes.plain_scene(seed=2, n_random=3000)
    n_tris = sum(m.ntris for m in meshes)
    sc = D.host_scene_bvh(meshes, layout=layout, threads=4, split_budget=budget)
    nb, pb = sc.download()
    nodes = nb.view(node_dtype(layout))
    tris = pb.view(TRI)
    assert n_tris < tris.shape[0] <= n_tris * (1 + (budget or 4.0)) + 64
    lo, hi = leaf_boxes_of_records(nodes, 0, tris.shape[0], layout)
    assert not np.isnan(lo).any()
    gid = np.zeros(tris.shape[0], np.int64)
    base = np.cumsum([0] + [m.ntris for m in meshes])
    gid[:] = base[tris["mesh"]] + tris["tri"]
    counts = np.bincount(gid, minlength=n_tris)
    assert np.all(counts >= 1)
    # records carry the triangle's own data
    for i in rng.choice(tris.shape[0], 400, replace=False):
        v = meshes[tris["mesh"][i]].verts[meshes[tris["mesh"][i]].tris[tris["tri"][i]]]
        assert np.array_equal(tris["p0"][i], v[0]) and np.array_equal(tris["e1"][i], (v[1] - v[0]).astype(np.float32))
    # coverage: points of every cut triangle (vertices, edge midpoints, random interior points)
    cut = np.nonzero(counts > 1)[0]
    assert cut.size > 100
    order = np.argsort(gid, kind="stable")
    starts = np.searchsorted(gid[order], np.arange(n_tris))
    w = rng.dirichlet((1.0, 1.0, 1.0), size=60)
    w = np.vstack([np.eye(3), [[0.5, 0.5, 0], [0, 0.5, 0.5], [0.5, 0, 0.5]], w])
    for g in cut:
        mi = int(np.searchsorted(base, g, side="right") - 1)
        v = meshes[mi].verts[meshes[mi].tris[g - base[mi]]].astype(np.float64)
        pts = w @ v
        recs = order[starts[g]:starts[g] + counts[g]]
        inside = np.zeros(pts.shape[0], bool)
        for r in recs:
            inside |= np.all((pts >= lo[r]) & (pts <= hi[r]), axis=1)
        assert inside.all(), (g, pts[~inside][:3])
    # triangles of one size (a flat grid: every box area equal) are not cut
    g = scenes.Mesh(*_grid(32))
    h = D.host_scene_bvh([g], threads=4)
    assert h.info()["primCount"] == g.ntris


@pytest.mark.parametrize("layout", [0, 2])
def test_host_builder_two_level(layout):
    sc = scenes.instanced_scene(n_inst=30, motion_frac=0.3, freq=6, extent=10.0, seed=2)
    h = D.host_scene_mbvh(sc, layout=layout)
    nb, pb = h.download()
    nodes = nb.view(node_dtype(layout))
    ntri = sc.meshes[0].ntris
    tris = pb[: ntri * 48].view(TRI)
    insts = pb[ntri * 48:].view(INST)
    assert insts.shape[0] == 30
    assert sorted(insts["object"]) == list(range(30))
    assert np.array_equal(np.sort(insts["kind"]), np.sort(sc.obj_kind))
    # the root-leaf boxes equal the oracle's epsilon-expanded world boxes
    ob = oracle.OracleMBVH(sc)
    root_boxes = {}
    for o in range(30):
        root_boxes[o] = ob.leaf_bbox(o)

    def boxes_of_child(s, c):
        assert c == 1
        return root_boxes[int(insts["object"][s])]
    # root tree occupies the first slots; leaf trees follow
    dec_ok = True
    try:
        check_tree(nodes, boxes_of_child, 0, nodes.shape[0], 30, layout)
    except AssertionError:
        dec_ok = False
    assert dec_ok
    assert tris.shape[0] == ntri


def test_device_builder_needs_a_device():
    import ctypes
    arr, keep = D._mesh_structs([scenes.box_mesh()])
    opts = D._build_opts(builder=D.BUILD_DEVICE_LBVH)
    p = ctypes.c_void_p()
    assert D.lib().lrh_scene_create_bvh(None, arr, 1, ctypes.byref(opts), ctypes.byref(p)) == -1
    assert b"needs a device" in D.lib().lrh_last_error()
    opts.builder = 7
    assert D.lib().lrh_scene_create_bvh(None, arr, 1, ctypes.byref(opts), ctypes.byref(p)) == -1


def test_object_boxes_equal_oracle_leaf_boxes():
    """The culling box of every object (instance: transformed mesh box;
    motion: MotionSystem::Bound over 1,025 samples per segment; both
    MachineEpsilon-expanded) equals the oracle MBVHAccel's root-leaf box."""
    import oracle
    sc = scenes.instanced_scene(n_inst=60, motion_frac=0.5, freq=6, extent=20.0, seed=61)
    n = len(sc.obj_kind)
    got = D.host_scene_mbvh(sc).object_boxes(n)
    ob = oracle.OracleMBVH(sc)
    for o in range(n):
        lo, hi = ob.leaf_bbox(o)
        assert np.array_equal(got[o, 0], lo) and np.array_equal(got[o, 1], hi), o


def test_documented_limits_fail_cleanly():
    """Size limits are checked before anything touches a device: 2^31-1 rays
    per trace, query 0/1, maxLeaf 1..8 (0: default), layouts 0..2; updates
    and refits need a two-level device scene with the same objects."""
    L = D.lib()
    host = D.host_scene_bvh(scenes.plain_scene(n_random=10))
    fake = ctypes.c_void_p(1 << 20)
    assert L.lrh_trace(host.h, fake, fake, 0x80000000, 0, None) == D.LRH_ERANGE
    assert b"2^31-1" in L.lrh_last_error()
    assert L.lrh_trace(host.h, fake, fake, 16, 2, None) == -1
    assert L.lrh_trace(host.h, fake, fake, 16, 0, None) == -1  # host-only scene
    assert b"host-only" in L.lrh_last_error()
    assert L.lrh_scene_refit(host.h, None, 0) == -1
    assert L.lrh_scene_update(host.h, None, 0) == -1
    for bad in (dict(max_leaf=9), dict(layout=3)):  # 0 selects the default
        with pytest.raises(D.LrhError):
            D.host_scene_bvh(scenes.plain_scene(n_random=10), **bad)
    sc = scenes.instanced_scene(n_inst=6, motion_frac=0.5, freq=4, extent=5.0, seed=3)
    two = D.host_scene_mbvh(sc)
    with pytest.raises(D.LrhError):
        two.object_boxes(5)  # wrong object count
    arr, keep = D._object_structs(sc)
    assert L.lrh_scene_refit(two.h, arr, len(sc.obj_kind)) == -1  # host-only


def _bound_probe_scene(keys, p):
    """One motion object whose mesh is a single degenerate triangle at p (its
    local box is the point p)."""
    mesh = scenes.Mesh(np.repeat(p[None, :], 3, axis=0).astype(np.float32), np.array([[0, 1, 2]], np.uint32))
    return scenes.InstancedScene([mesh], np.zeros(1, np.uint32), np.array([2], np.uint32),
                                 np.zeros((1, 4, 4), np.float32), np.array([0, 2], np.uint32),
                                 np.array([0.0, 1.0], np.float32), keys)


@pytest.mark.parametrize("probe", ["A", "B"])
def test_host_builder_motion_bound_closed_form(probe):
    """The host builder's culling box of a moving object is MotionSystem::Bound
    (1,025 samples per segment, motionsystem.cpp:75-87) expanded by
    MachineEpsilon (mbvhaccel.cpp:192-193): checked against the closed form
    of a point rotated about z (tests/test_oracle_golden.py)."""
    from tests.test_oracle_golden import BOUND_PROBES, BOUND_TOL, _bound_err, bound_probe
    keys, p, expected = bound_probe(BOUND_PROBES[probe])
    h = D.host_scene_mbvh(_bound_probe_scene(keys, p))
    box = h.object_boxes(1)[0]
    lo, hi = oracle.motion_bound([0, 1], keys, p, p)
    e = oracle.epsilon_bbox(lo, hi)
    assert np.array_equal(box[0], (lo - e).astype(np.float32)) and np.array_equal(box[1], (hi + e).astype(np.float32))
    elo, ehi = expected(1024)
    assert _bound_err(box[0] + e, box[1] - e, (elo, ehi)) <= BOUND_TOL + 1e-6
    assert _bound_err(box[0] + e, box[1] - e, expected(1023)) > 2 * BOUND_TOL
