"""Two-level root refit on the GPU (lrh_scene_refit): after an edit of
instance matrices and motion keys, the object culling boxes equal the host's
(MotionSystem::Bound over 1,025 samples per segment, MachineEpsilon
expansion) and traces equal the oracle MBVHAccel built on the edited scene."""
import copy

import numpy as np
import pytest

import oracle
from luxcore_amd import device as D
from luxcore_amd import scenes
from tests.parity import assert_parity, compare

pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def moved(sc, seed, extent=100.0, step=5.0):
    """Same objects, kinds and key counts; new matrices and keys."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = copy.deepcopy(sc)
    for o in range(len(out.obj_kind)):
        if out.obj_kind[o] == 1:
            out.obj_mat[o] = scenes.random_trs(rng, extent)
        elif out.obj_kind[o] == 2:
            k0, k1 = int(out.key_off[o]), int(out.key_off[o + 1])
            M = scenes.random_trs(rng, extent).astype(np.float64)
            for k in range(k0, k1):
                Mk = M.copy()
                Mk[:3, 3] += rng.uniform(-step, step, size=3)
                if rng.random() < 0.5:
                    R = scenes._random_rotation(rng)
                    Mk[:3, :3] = Mk[:3, :3] @ R
                out.key_mats[k] = np.linalg.inv(Mk).astype(np.float32)
    return out


def _check_boxes(dev_scene, sc):
    n = len(sc.obj_kind)
    got = dev_scene.object_boxes(n)
    host = D.host_scene_mbvh(sc)
    want = host.object_boxes(n)
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]


def _check_trace(dev, dev_scene, sc, rays):
    ob = oracle.OracleMBVH(sc)
    ref = ob.intersect(rays)
    got = dev.trace_numpy(dev_scene, rays)
    st = compare(ref, got)
    print(st, flush=True)
    assert_parity(st)
    return st


@pytest.mark.parametrize("layout", [0, 1, 2])
def test_refit_boxes_and_trace(gpu_device, layout):
    _case("refit", layout=layout)
    sc = scenes.instanced_scene(n_inst=400, motion_frac=0.3, freq=20, extent=40.0, seed=31)
    rays = scenes.instanced_rays(1 << 18, extent=40.0, seed=32)
    dsc = gpu_device.scene_mbvh(sc, layout=layout)
    _check_boxes(dsc, sc)
    for step, seed in enumerate((33, 34, 35)):
        sc2 = moved(sc, seed, extent=40.0)
        dsc.refit(sc2)
        _check_boxes(dsc, sc2)
        st = _check_trace(gpu_device, dsc, sc2, rays)
        assert st["hits"] > 0.05 * rays.shape[0], step
    # an update (host rebuild) after refits, then a refit again
    sc3 = moved(sc, 36, extent=40.0)
    dsc.update(sc3)
    _check_boxes(dsc, sc3)
    _check_trace(gpu_device, dsc, sc3, rays)
    sc4 = moved(sc, 37, extent=40.0)
    dsc.refit(sc4)
    _check_boxes(dsc, sc4)
    _check_trace(gpu_device, dsc, sc4, rays)
    dsc.close()


def test_refit_full_size_motion(gpu_device):
    """1,000 instances of a 50,000-triangle sphere, 200 with motion, after a
    refit: boxes equal the host's, 1,048,576 rays equal the oracle."""
    _case("refit_full")
    sc = scenes.instanced_scene()
    rays = scenes.instanced_rays(1 << 20, seed=41)
    dsc = gpu_device.scene_mbvh(sc)
    sc2 = moved(sc, 42)
    dsc.refit(sc2)
    _check_boxes(dsc, sc2)
    _check_trace(gpu_device, dsc, sc2, rays)
    dsc.close()


def test_refit_rejects_bad_edits(gpu_device):
    _case("refit_errors")
    sc = scenes.instanced_scene(n_inst=20, motion_frac=0.5, freq=8, extent=10.0, seed=51)
    dsc = gpu_device.scene_mbvh(sc)
    bad = copy.deepcopy(sc)
    bad.obj_kind = bad.obj_kind.copy()
    i = int(np.nonzero(bad.obj_kind == 1)[0][0])
    bad.obj_kind[i] = 0
    with pytest.raises(D.LrhError):
        dsc.refit(bad)
    sing = copy.deepcopy(sc)
    j = int(np.nonzero(sing.obj_kind == 2)[0][0])
    sing.key_mats[int(sing.key_off[j])] = np.zeros((4, 4), np.float32)
    with pytest.raises(D.LrhError):
        dsc.refit(sing)
    flat = gpu_device.scene_bvh(scenes.plain_scene(seed=1, n_random=100))
    with pytest.raises(D.LrhError):
        D.check(D.lib().lrh_scene_refit(flat.h, None, 0))
    flat.close()
    # the scene still traces correctly after the rejected edits
    ok = moved(sc, 52, extent=10.0)
    dsc.refit(ok)
    _check_boxes(dsc, ok)
    _check_trace(gpu_device, dsc, ok, scenes.instanced_rays(50000, extent=10.0, seed=53))
    dsc.close()


def test_rejected_refit_keeps_last_accepted_boxes(gpu_device):
    """A good refit, then one with a singular motion sample: the rejected
    edit leaves the object boxes (and traces) of the good one."""
    _case("refit_rejected_keeps_boxes")
    sc = scenes.instanced_scene(n_inst=30, motion_frac=0.5, freq=8, extent=10.0, seed=61)
    dsc = gpu_device.scene_mbvh(sc)
    good = moved(sc, 62, extent=10.0)
    dsc.refit(good)
    before = dsc.object_boxes(len(sc.obj_kind)).copy()
    _check_boxes(dsc, good)
    sing = moved(sc, 63, extent=10.0)
    j = int(np.nonzero(sing.obj_kind == 2)[0][0])
    sing.key_mats[int(sing.key_off[j])] = np.zeros((4, 4), np.float32)
    with pytest.raises(D.LrhError):
        dsc.refit(sing)
    after = dsc.object_boxes(len(sc.obj_kind))
    assert np.array_equal(before, after)
    _check_boxes(dsc, good)
    _check_trace(gpu_device, dsc, good, scenes.instanced_rays(50000, extent=10.0, seed=64))
    dsc.close()


@pytest.mark.parametrize("probe", ["A", "B"])
def test_refit_motion_bound_closed_form(gpu_device, probe):
    """The GPU refit's box of a moving object against the closed form of a
    point rotated about z (MotionSystem::Bound, 1,025 samples, then
    MachineEpsilon), and equal to the host builder's bit for bit."""
    from tests.test_abi_cpu import _bound_probe_scene
    from tests.test_oracle_golden import BOUND_PROBES, BOUND_TOL, _bound_err, bound_probe
    _case("refit_bound_closed_form", probe=probe)
    keys, p, expected = bound_probe(BOUND_PROBES[probe])
    sc = _bound_probe_scene(keys, p)
    dsc = gpu_device.scene_mbvh(sc)
    dsc.refit(sc)
    box = dsc.object_boxes(1)[0]
    assert np.array_equal(box, D.host_scene_mbvh(sc).object_boxes(1)[0])
    lo, hi = oracle.motion_bound([0, 1], keys, p, p)
    e = oracle.epsilon_bbox(lo, hi)
    assert _bound_err(box[0] + e, box[1] - e, expected(1024)) <= BOUND_TOL + 1e-6
    assert _bound_err(box[0] + e, box[1] - e, expected(1023)) > 2 * BOUND_TOL
    dsc.close()
