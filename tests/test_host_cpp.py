"""The C++ device classes (luxcore_amd/host: Context, DeviceDescription,
DataSet, HIPIntersectionDevice, Accelerator, HardwareIntersectionKernel,
HardwareDeviceBuffer) driven by
luxrays_hip_check the way a LuxRays application drives an intersection
device.  Hits must equal the ctypes path's bit for bit (same kernels) and
the oracle's to the parity bar; TraceRay must equal the buffer trace; Update
after a transform edit must equal a scene built from the edited objects."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle
from luxcore_amd import device as D
from luxcore_amd import scenes
from luxcore_amd.raytypes import HIT_DTYPE, RAY_DTYPE
from tests.parity import assert_parity, compare

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "luxcore_amd", "host", "luxrays_hip_check")


def write_meshes(d, meshes):
    with open(os.path.join(d, "meshes.bin"), "wb") as f:
        f.write(np.uint32(len(meshes)).tobytes())
        for m in meshes:
            v = np.ascontiguousarray(m.verts, np.float32)
            t = np.ascontiguousarray(m.tris, np.uint32)
            f.write(np.array([v.shape[0], t.shape[0]], np.uint32).tobytes())
            f.write(v.tobytes())
            f.write(t.tobytes())


def write_objects(path, sc):
    with open(path, "wb") as f:
        n = len(sc.obj_mesh)
        f.write(np.uint32(n).tobytes())
        for o in range(n):
            k0, k1 = int(sc.key_off[o]), int(sc.key_off[o + 1])
            f.write(np.array([sc.obj_mesh[o], sc.obj_kind[o]], np.uint32).tobytes())
            f.write(np.ascontiguousarray(sc.obj_mat[o], np.float32).tobytes())
            f.write(np.uint32(k1 - k0).tobytes())
            f.write(np.ascontiguousarray(sc.key_times[k0:k1], np.float32).tobytes())
            f.write(np.ascontiguousarray(sc.key_mats[k0:k1], np.float32).tobytes())


def run_check(d, *args, expect=0):
    p = subprocess.run([CHECK, str(d), *args], capture_output=True, text=True, timeout=300)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == expect, (p.returncode, p.stderr)
    return p


def read_hits(d, name="hits.bin"):
    return np.fromfile(os.path.join(d, name), dtype=HIT_DTYPE)


def _gpu_count():
    n = ctypes.c_int(0)
    if D.lib().lrh_device_count(ctypes.byref(n)) != 0:
        return 0
    return n.value


# ---------------------------------------------------------------- CPU checks

def test_check_program_built():
    assert os.access(CHECK, os.X_OK)
    assert os.path.exists(os.path.join(ROOT, "luxcore_amd", "host", "libluxrays_hip.so"))


def test_check_program_usage_and_missing_input(tmp_path):
    p = subprocess.run([CHECK], capture_output=True, text=True)
    assert p.returncode == 2 and "usage" in p.stderr
    p = subprocess.run([CHECK, str(tmp_path), "--bogus"], capture_output=True, text=True)
    assert p.returncode == 2
    p = subprocess.run([CHECK, str(tmp_path)], capture_output=True, text=True)
    assert p.returncode == 1 and "cannot open" in p.stderr


def test_truncated_mesh_file_is_an_error(tmp_path):
    with open(tmp_path / "meshes.bin", "wb") as f:
        f.write(np.array([1, 10, 4], np.uint32).tobytes())
    p = subprocess.run([CHECK, str(tmp_path)], capture_output=True, text=True)
    assert p.returncode == 1 and "truncated" in p.stderr


def test_no_gpu_fails_loudly(tmp_path):
    if _gpu_count() > 0:
        pytest.skip("a GPU is present")
    write_meshes(tmp_path, [scenes.box_mesh()])
    scenes.plain_rays(16, seed=1).tofile(tmp_path / "rays.bin")
    p = subprocess.run([CHECK, str(tmp_path)], capture_output=True, text=True)
    assert p.returncode == 1 and "lrhip:" in p.stderr


def test_device_registry():
    # DeviceType / DeviceDescription / Context type checks (include/luxrays/core/device.h:39-82,
    # src/luxrays/core/device.cpp:28-89, context.cpp:250-291) on synthetic descriptions:
    # dev0 NATIVE, dev1 OPENCL_CPU, dev2 OPENCL_GPU, dev3 CUDA_GPU, dev4 and dev5 HIP_GPU
    p = subprocess.run([CHECK, "--registry"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.splitlines()
    names = {int(l.split()[1]): l.split()[2] for l in lines if l.startswith("type ") and l.split()[1].isdigit()}
    assert names == {95: "ALL", 1: "NATIVE_THREAD", 30: "OPENCL_ALL", 2: "OPENCL_DEFAULT", 4: "OPENCL_CPU",
                     8: "OPENCL_GPU", 16: "OPENCL_UNKNOWN", 32: "CUDA_GPU", 64: "HIP_GPU"}
    assert "type error: Unknown device type in DeviceDescription::GetDeviceType(): 126" in lines
    # the bit values are the reference's, the HIP bit is the next free one
    assert "filter 95: dev0 dev1 dev2 dev3 dev4 dev5" in lines       # ALL keeps everything
    assert "filter 64: dev4 dev5" in lines                            # HIP_GPU
    assert "filter 126: dev1 dev2 dev3 dev4 dev5" in lines           # ALL_HARDWARE
    assert "filter 30: dev1 dev2" in lines                            # OPENCL_ALL
    assert "filterone all: dev2" in lines                             # first GPU description
    assert "filterone hip: dev4" in lines
    assert "filterone cpu-only:" in lines                             # no GPU: empty, as the reference ends
    assert "context error: Unknown device type in Context::CreateIntersectionDevices(): 1" in lines


def test_context_lists_devices():
    p = subprocess.run([CHECK, "--devices"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    n = _gpu_count()
    assert f"{n} device description(s)" in p.stdout
    assert f"[LuxRays] HIP device count: {n}" in p.stdout
    if n:
        assert "Device 0 type: HIP_GPU" in p.stdout and "selected HIP-" in p.stdout


# ---------------------------------------------------------------- GPU parity

@pytest.mark.gpu
def test_context_describes_the_gpu():
    print("[case] cpp context device descriptions", flush=True)
    p = subprocess.run([CHECK, "--devices"], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    n = _gpu_count()
    assert n >= 1 and f"{n} device description(s)" in p.stdout
    cus = [int(l.rsplit(" ", 1)[1]) for l in p.stdout.splitlines() if "compute units:" in l]
    assert len(cus) == n and all(c > 0 for c in cus)
    assert "selected HIP-" in p.stdout and "HIP_GPU" in p.stdout


@pytest.mark.gpu
def test_cpp_device_single_level(gpu_device, tmp_path):
    print("[case] cpp single-level", flush=True)
    meshes = scenes.plain_scene(seed=1, n_random=20000)
    rays = scenes.plain_rays(200_000, seed=11)
    rays["flags"][::7] = 1
    write_meshes(tmp_path, meshes)
    rays.tofile(tmp_path / "rays.bin")
    out = run_check(tmp_path, "--serial", "512")
    assert "accel BVH" in out.stdout
    got = read_hits(tmp_path)
    sc = gpu_device.scene_bvh(meshes)
    init = np.zeros(rays.shape[0], HIT_DTYPE)
    ref_dev = gpu_device.trace_numpy(sc, rays, hits_init=init)
    live = (rays["flags"] & 1) == 0
    assert np.array_equal(got[live], ref_dev[live])
    ref = oracle.OracleBVH(meshes).intersect(rays[live])
    st = compare(ref, got[live])
    print(st, flush=True)
    assert_parity(st)
    serial = read_hits(tmp_path, "serial.bin")
    s_live = live[:512]
    assert np.array_equal(serial[s_live], got[:512][s_live])
    sc.close()


@pytest.mark.gpu
def test_cpp_device_any_hit(gpu_device, tmp_path):
    print("[case] cpp any-hit", flush=True)
    meshes = scenes.plain_scene(seed=3, n_random=5000)
    rays = scenes.plain_rays(50_000, seed=12)
    write_meshes(tmp_path, meshes)
    rays.tofile(tmp_path / "rays.bin")
    run_check(tmp_path, "--anyhit")
    got = read_hits(tmp_path)
    sc = gpu_device.scene_bvh(meshes)
    ref = gpu_device.trace_numpy(sc, rays, query=D.LRH_QUERY_ANY)
    assert np.array_equal(got, ref)
    sc.close()


@pytest.mark.gpu
def test_cpp_device_two_level_and_update(gpu_device, tmp_path):
    print("[case] cpp two-level + update", flush=True)
    sc = scenes.instanced_scene(n_inst=60, motion_frac=0.3, freq=12, extent=10.0, seed=61)
    rays = scenes.instanced_rays(100_000, extent=10.0, seed=62)
    write_meshes(tmp_path, sc.meshes)
    write_objects(tmp_path / "objects.bin", sc)
    sc2 = scenes.instanced_scene(n_inst=60, motion_frac=0.3, freq=12, extent=10.0, seed=61)
    rng = np.random.Generator(np.random.PCG64(63))
    kinds = sc2.obj_kind
    sc2.obj_mat[kinds == 1, :3, 3] += rng.uniform(-1.0, 1.0, size=(int((kinds == 1).sum()), 3)).astype(np.float32)
    sc2.key_mats[:, :3, 3] += rng.uniform(-0.5, 0.5, size=(sc2.key_mats.shape[0], 3)).astype(np.float32)
    write_objects(tmp_path / "objects2.bin", sc2)
    rays.tofile(tmp_path / "rays.bin")
    out = run_check(tmp_path, "--serial", "256")
    assert "accel MBVH" in out.stdout
    got, got2 = read_hits(tmp_path), read_hits(tmp_path, "hits2.bin")
    dsc = gpu_device.scene_mbvh(sc)
    assert np.array_equal(got, gpu_device.trace_numpy(dsc, rays))
    dsc2 = gpu_device.scene_mbvh(sc2)
    assert np.array_equal(got2, gpu_device.trace_numpy(dsc2, rays))
    assert not np.array_equal(got, got2)
    assert np.array_equal(read_hits(tmp_path, "serial.bin"), got[:256])
    # the same edit through a GPU refit: same hits as the refitted C-ABI scene
    run_check(tmp_path, "--refit")
    got3 = read_hits(tmp_path, "hits2.bin")
    dsc.refit(sc2)
    assert np.array_equal(got3, gpu_device.trace_numpy(dsc, rays))
    dsc.close()
    dsc2.close()


@pytest.mark.gpu
def test_cpp_bvh_refuses_instances(gpu_device, tmp_path):
    print("[case] cpp bvh refuses instances", flush=True)
    sc = scenes.instanced_scene(n_inst=4, motion_frac=0.0, freq=4, extent=10.0, seed=64)
    write_meshes(tmp_path, sc.meshes)
    write_objects(tmp_path / "objects.bin", sc)
    scenes.instanced_rays(64, extent=10.0, seed=65).tofile(tmp_path / "rays.bin")
    p = run_check(tmp_path, "--accel", "bvh", expect=1)
    assert "BVHAccel supports only plain triangle meshes" in p.stderr


@pytest.mark.gpu
def test_cpp_wide8_two_level_and_grid_share(gpu_device, tmp_path):
    """hip.bvh.layout = WIDE8 on an MBVH scene and a launch on half the
    persistent grid (hip.kernel.gridpercent): the C-ABI 8-wide scene's hits."""
    print("[case] cpp wide8 two-level + grid share", flush=True)
    sc = scenes.instanced_scene(n_inst=30, motion_frac=0.3, freq=8, extent=10.0, seed=66)
    rays = scenes.instanced_rays(60_000, extent=10.0, seed=67)
    write_meshes(tmp_path, sc.meshes)
    write_objects(tmp_path / "objects.bin", sc)
    rays.tofile(tmp_path / "rays.bin")
    out = run_check(tmp_path, "--prop", "hip.bvh.layout = WIDE8", "--prop", "hip.kernel.gridpercent = 50")
    assert "accel MBVH" in out.stdout
    dsc = gpu_device.scene_mbvh(sc, layout=2)
    assert np.array_equal(read_hits(tmp_path), gpu_device.trace_numpy(dsc, rays))
    dsc.close()
