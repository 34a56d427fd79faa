"""The device's shading step (lrh_shade) against the oracle's restatement,
step by step on a bundled scene: each step consumes the GPU's own rays and
hits, so both sides see the same inputs; path records and shadow rays must
match bit for bit, next-ray directions to fp32 rounding of cos/sin."""
import os

import numpy as np
import pytest

from luxcore_amd import device as D
from luxcore_amd import scenefile as P
from luxcore_amd.raytypes import HIT_DTYPE, RAY_DTYPE
from oracle import shade
from tests.scene_fixtures import unpack

pytestmark = pytest.mark.gpu

OVERRIDES = """
scene.materials.redmat.kd = 0.75 0.0 0.0
scene.materials.whitelight.emission = 8.0 8.0 8.0
scene.materials.whitelight.emission.gain = 1.0 1.0 1.0
"""


def _scene(tmp_path):
    d = unpack("bigmonkey", tmp_path)
    src = os.path.join(d, "bigmonkey.scn")
    path = os.path.join(d, "bigmonkey-lit.scn")
    with open(src) as f, open(path, "w") as g:
        g.write(f.read() + OVERRIDES)
    return P.load_scene(path)


def _same_rays(a, b, what):
    assert np.array_equal(a["flags"], b["flags"]), what
    live = a["flags"] == 0
    assert np.array_equal(a["o"][live], b["o"][live]), what
    assert np.array_equal(a["mint"][live], b["mint"][live]) and np.array_equal(a["maxt"][live], b["maxt"][live]), what
    return live


def test_shade_steps_equal_oracle(gpu_device, tmp_path):
    print("[case] shade bigmonkey", flush=True)
    sf = _scene(tmp_path)
    assert not sf.two_level
    mats, lights = sf.materials(), sf.light_triangles()
    assert len(lights) > 0 and mats[1]["kd"][1] == 0.0
    meshes = sf.dataset_meshes()
    dev = gpu_device
    scene = P.scene_on_device(dev, sf)
    W, H = 160, 120
    n = W * H
    orig, target, up, fov = sf.camera
    rb, nb, sb = dev.alloc(n * 48), dev.alloc(n * 48), dev.alloc(n * 48)
    hb, shb = dev.alloc(n * 20), dev.alloc(n * 20)
    pb = dev.alloc(n * D.PATH_DTYPE.itemsize)
    mb, lb = dev.upload(mats), dev.upload(lights)
    dev.camera_rays(rb, orig, target, up, fov, W, H)
    dev.paths_init(pb, n)
    dev.sync()
    max_depth, rr_depth, rr_s, seed = 5, 3, 0.7, 77
    have_shadow = False
    for depth in range(max_depth):
        dev.trace(scene, rb, hb, n, 0)
        if have_shadow:
            dev.trace(scene, sb, shb, n, 1)
        dev.sync()
        rays = dev.download(rb, RAY_DTYPE, n)
        hits = dev.download(hb, HIT_DTYPE, n)
        sh = dev.download(shb, HIT_DTYPE, n) if have_shadow else None
        before = dev.download(pb, D.PATH_DTYPE, n)
        dev.shade(scene, mb, len(mats), lb, len(lights), rb, hb, shb if have_shadow else None, pb, nb, sb, n,
                  max_depth=max_depth, rr_depth=rr_depth, rr_survive=rr_s, seed=seed)
        dev.sync()
        got_p = dev.download(pb, D.PATH_DTYPE, n)
        got_n = dev.download(nb, RAY_DTYPE, n)
        got_s = dev.download(sb, RAY_DTYPE, n)
        ref_p, ref_n, ref_s = shade.shade_step(meshes, mats, lights, rays, hits, sh, before, max_depth, rr_depth,
                                               rr_s, seed)
        print(f"depth {depth}: alive {int(got_p['alive'].sum())}, shadow {int((got_s['flags'] == 0).sum())}",
              flush=True)
        assert got_p.tobytes() == ref_p.tobytes(), depth
        live = _same_rays(got_n, ref_n, ("next", depth))
        assert np.max(np.abs(got_n["d"][live] - ref_n["d"][live]), initial=0.0) < 2e-6, depth
        assert got_s.tobytes() == ref_s.tobytes(), depth
        rb, nb = nb, rb
        have_shadow = True
    rad = dev.download(pb, D.PATH_DTYPE, n)["rad"]
    assert np.all(np.isfinite(rad)) and (rad.sum(axis=1) > 0).mean() > 0.5
    # the red monkey: more red than green where it is lit
    assert rad[:, 0].sum() > rad[:, 1].sum()
    for b in (rb, nb, sb, hb, shb, pb, mb, lb):
        b.free()
    scene.close()


def test_shade_rejects_bad_arguments(gpu_device, tmp_path):
    print("[case] shade errors", flush=True)
    sf = _scene(tmp_path)
    scene = P.scene_on_device(gpu_device, sf)
    b = gpu_device.alloc(48 * 16)
    with pytest.raises(D.LrhError):
        gpu_device.shade(scene, None, 0, None, 0, b, b, None, b, b, b, 16, max_depth=0)
    with pytest.raises(D.LrhError):
        gpu_device.shade(scene, None, 0, None, 0, b, b, None, b, b, b, 16, rr_survive=0.0)
    b.free()
    scene.close()


def test_shade_batch_equals_per_buffer_shading(gpu_device, tmp_path):
    """lrh_shade_batch: every item gets exactly lrh_shade's paths, next rays
    and shadow rays, and the live counter the same total; items of different
    sizes (one not a multiple of the block, one empty), with and without the
    previous step's shadow hits."""
    print("[case] shade batch", flush=True)
    sf = _scene(tmp_path)
    dev = gpu_device
    scene = P.scene_on_device(dev, sf)
    mats, lights = sf.materials(), sf.light_triangles()
    mb, lb = dev.upload(mats), dev.upload(lights)
    orig, target, up, fov = sf.camera
    sets = []
    for k, (w, h) in enumerate(((96, 64), (50, 37), (128, 96), (0, 0))):
        n = w * h
        cap = max(n, 1)
        s = {"n": n, "rb": dev.alloc(cap * 48), "hb": dev.alloc(cap * 20), "shb": dev.alloc(cap * 20),
             "pb": dev.alloc(cap * D.PATH_DTYPE.itemsize), "seed": 100 + k}
        if n:
            dev.camera_rays(s["rb"], orig, target, up, fov, w, h)
            dev.paths_init(s["pb"], n)
            dev.trace(scene, s["rb"], s["hb"], n, 0)
            # shadow hits of a previous step: the same rays traced as any hit
            dev.trace(scene, s["rb"], s["shb"], n, 1)
        s["shadow"] = k % 2 == 1
        sets.append(s)
    dev.sync()
    outs = {}
    for mode in ("one", "batch"):
        counter = dev.upload(np.zeros(2, np.uint32))
        items = []
        for k, s in enumerate(sets):
            cap = max(s["n"], 1)
            pb = dev.upload(dev.download(s["pb"], D.PATH_DTYPE, cap))
            nb, sb = dev.alloc(cap * 48), dev.alloc(cap * 48)
            items.append((scene, s["rb"], s["hb"], s["shb"] if s["shadow"] else None, pb, nb, sb, s["n"], 6, 3, 0.7,
                          s["seed"], 7 * k))
        if mode == "one":
            for it in items:
                sc, rb, hb, shb, pb, nb, sb, n, md, rd, rs, seed, pbase = it
                dev.shade(sc, mb, len(mats), lb, len(lights), rb, hb, shb, pb, nb, sb, n, max_depth=md,
                          rr_depth=rd, rr_survive=rs, seed=seed, pixel_base=pbase, live_counter=counter.ptr)
        else:
            dev.shade_batch(items, mb, len(mats), lb, len(lights), live_counter=counter.ptr)
        dev.sync()
        outs[mode] = ([(dev.download(it[4], D.PATH_DTYPE, it[7]).tobytes(), dev.download(it[5], RAY_DTYPE, it[7]).tobytes(),
                        dev.download(it[6], RAY_DTYPE, it[7]).tobytes()) if it[7] else () for it in items],
                      int(dev.download(counter, np.uint32, 2)[0]))
    assert outs["batch"][1] == outs["one"][1] > 0
    for k, (a, b) in enumerate(zip(outs["one"][0], outs["batch"][0])):
        assert a == b, k
    with pytest.raises(D.LrhError, match="LRH_SHADE_BATCH_MAX"):
        dev.shade_batch([items[0]] * (D.SHADE_BATCH_MAX + 1), mb, len(mats), lb, len(lights))
    scene.close()
