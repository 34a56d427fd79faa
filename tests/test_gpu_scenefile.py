"""Bundled LuxCore scenes through the device: scene files read by the product
reader (libluxrays_ingest.so), built with the DataSet AUTO policy and traced on
the GPU, against the oracle's own reading of the same files traced by the
oracle BVHAccel / MBVHAccel."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from luxcore_amd import scenefile as P
from luxcore_amd import scenes
from luxcore_amd.raytypes import HIT_DTYPE, NULL_INDEX
from oracle import scenefile as O
from tests.parity import assert_parity, compare
from tests.scene_fixtures import unpack

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "luxcore_amd", "host", "luxrays_hip_check")


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _world_box(sf):
    lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
    for o in sf.objects:
        v = sf.meshes[o["mesh"]].verts.astype(np.float64)
        if v.size == 0:
            continue
        mats = [np.eye(4)] if o["kind"] == 0 else [o["m"]] if o["kind"] == 1 else \
            [np.linalg.inv(np.asarray(k, np.float64)) for k in o["keys"]]
        for M in mats:
            w = v @ np.asarray(M, np.float64)[:3, :3].T + np.asarray(M, np.float64)[:3, 3]
            lo, hi = np.minimum(lo, w.min(0)), np.maximum(hi, w.max(0))
    return lo, hi


def _rays(sf, n_random, width, height, seed):
    rng = np.random.default_rng(seed)
    lo, hi = _world_box(sf)
    pad = 0.05 * (hi - lo)
    r = scenes.random_rays(n_random, lo - pad, hi + pad, rng)
    orig, target, up, fov = sf.camera
    if np.linalg.norm(np.subtract(target, orig)) > 0 and np.linalg.norm(up) > 0:
        cam = scenes.pinhole_rays(width, height, eye=orig, look=target, up=up, vfov_deg=fov)
        r = np.concatenate([cam, r])
    r["time"] = rng.random(r.shape[0], dtype=np.float32)
    return r


def _trace_both(dev, path, n_random=200000, width=480, height=270, seed=1):
    sf = P.load_scene(path)
    ref_sf = O.load_scene(path)
    rays = _rays(sf, n_random, width, height, seed)
    dsc = P.scene_on_device(dev, sf)
    got = dev.trace_numpy(dsc, rays)
    if ref_sf.two_level:
        ob = oracle.OracleMBVH(ref_sf.instanced())
    else:
        ob = oracle.OracleBVH(ref_sf.dataset_meshes())
    ref = ob.intersect(rays)
    motion = [i for i, o in enumerate(sf.objects) if o["kind"] == 2]
    st = compare(ref, got, motion_objects=motion)
    print(os.path.basename(path), sf.info, st, flush=True)
    assert_parity(st)
    anyh = dev.trace_numpy(dsc, rays, query=1)
    flips = np.nonzero((anyh["meshIndex"] != NULL_INDEX) != (ref["meshIndex"] != NULL_INDEX))[0]
    assert flips.size <= 1e-5 * rays.shape[0] + 5, flips[:10]
    dsc.close()
    return sf, rays, got, st


@pytest.mark.parametrize("name", ["bigmonkey.scn", "bigmonkey-instances.scn", "bigmonkey-motion.scn"])
def test_bigmonkey_scenes(gpu_device, tmp_path, name):
    _case("scenefile", scene=name)
    d = unpack("bigmonkey", tmp_path)
    sf, rays, got, st = _trace_both(gpu_device, os.path.join(d, name))
    assert sf.two_level == (name != "bigmonkey.scn")
    assert st["hits"] > 0.5 * rays.shape[0]


def test_simple_scene(gpu_device, tmp_path):
    _case("scenefile", scene="simple.scn")
    d = unpack("simple", tmp_path)
    sf, rays, got, st = _trace_both(gpu_device, os.path.join(d, "simple.scn"))
    assert st["hits"] > 0.3 * rays.shape[0]


def test_lightinstances_scene(gpu_device, tmp_path):
    """4,502 objects (instances of two 3,968-triangle meshes), 17.9M world
    triangles, one-level-per-instance MBVH."""
    _case("scenefile", scene="lightinstances/scene.scn")
    d = unpack("lightinstances", tmp_path)
    sf, rays, got, st = _trace_both(gpu_device, os.path.join(d, "scene.scn"), n_random=150000,
                                    width=320, height=180)
    assert sf.info["nObjects"] == 4502 and sf.two_level
    assert st["hits"] > 0.2 * rays.shape[0]


def test_cpp_classes_on_scene_file(gpu_device, tmp_path):
    """The C++ device classes fed by the scene reader give the same hits as
    the C ABI path."""
    _case("scenefile_cpp", scene="bigmonkey-motion.scn")
    d = unpack("bigmonkey", tmp_path)
    path = os.path.join(d, "bigmonkey-motion.scn")
    sf = P.load_scene(path)
    rays = _rays(sf, 50000, 160, 90, seed=4)
    work = tmp_path / "cpp"
    work.mkdir()
    rays.tofile(str(work / "rays.bin"))
    r = subprocess.run([CHECK, str(work), "--scene", path, "--serial", "64"], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "accel MBVH" in r.stdout
    cpp = np.fromfile(str(work / "hits.bin"), dtype=HIT_DTYPE)
    dsc = P.scene_on_device(gpu_device, sf)
    py = gpu_device.trace_numpy(dsc, rays)
    dsc.close()
    assert cpp.tobytes() == py.tobytes()
    serial = np.fromfile(str(work / "serial.bin"), dtype=HIT_DTYPE)
    assert serial.tobytes() == py[:64].tobytes()
