"""Served mode: ray generation kernels, admission by HBM budget, and a
session's results identical alone and inside a mix of concurrent sessions."""
import numpy as np
import pytest

from luxcore_amd import device as D
from luxcore_amd import scenes
from luxcore_amd import serve
from luxcore_amd.raytypes import HIT_DTYPE, NULL_INDEX, RAY_DTYPE

pytestmark = pytest.mark.gpu

SIZES = (5_000, 20_000, 60_000, 100_000)


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.fixture(scope="module")
def serve_scenes(gpu_device):
    meshes = scenes.batch_scenes(SIZES, seed=31)
    return meshes, [gpu_device.scene_bvh(m) for m in meshes]


def test_bounce_rays_invariants(gpu_device, serve_scenes):
    _case("bounce_invariants")
    dev = gpu_device
    meshes, scs = serve_scenes
    n = 128 * 128
    rb, nb, sb = dev.alloc(n * 48), dev.alloc(n * 48), dev.alloc(n * 48)
    hb = dev.alloc(n * 20)
    dev.camera_rays(rb, (20.0, 20.0, 90.0), (50.0, 50.0, 30.0), (0, 0, 1), 50.0, 128, 128)
    dev.trace(scs[3], rb, hb, n, 0)
    counter = dev.upload(np.zeros(2, np.uint32))
    light = (50.0, 50.0, 150.0)
    for bounce in (0, 5):
        dev.bounce_rays(scs[3], rb, hb, n, bounce, 1234, nb, sb, 0, 2, 0.7, light, counter.ptr)
        dev.sync()
        rays = dev.download(rb, RAY_DTYPE, n)
        hits = dev.download(hb, HIT_DTYPE, n)
        nxt = dev.download(nb, RAY_DTYPE, n)
        sh = dev.download(sb, RAY_DTYPE, n)
        hit = hits["meshIndex"] != NULL_INDEX
        assert hit.sum() > n // 4
        assert np.all(nxt["flags"][~hit] == 1) and np.all(sh["flags"][~hit] == 1)
        p = rays["o"][hit] + rays["d"][hit] * hits["t"][hit][:, None]
        alive = hit & (nxt["flags"] == 0)
        if bounce == 0:
            assert np.array_equal(alive, hit)
        else:
            frac = alive.sum() / hit.sum()
            assert 0.65 < frac < 0.75, frac
        assert np.allclose(nxt["o"][alive], (rays["o"] + rays["d"] * hits["t"][:, None])[alive], atol=1e-4)
        assert np.allclose(np.linalg.norm(nxt["d"][alive], axis=1), 1.0, atol=1e-5)
        # the new direction leaves on the side the ray came from
        m = meshes[3][0]
        idx = np.nonzero(alive)[0]
        tv = m.tris[hits["triangleIndex"][idx]]
        nrm = np.cross(m.verts[tv[:, 1]] - m.verts[tv[:, 0]], m.verts[tv[:, 2]] - m.verts[tv[:, 0]])
        side_in = np.sign(np.sum(nrm * rays["d"][idx], axis=1))
        side_out = np.sign(np.sum(nrm * nxt["d"][idx], axis=1))
        assert np.mean(side_in != side_out) > 0.999
        # shadow rays aim at the 1x1 light
        lit = hit & (sh["flags"] == 0)
        end = sh["o"][lit] + sh["d"][lit] * (sh["maxt"][lit] + 2 * sh["mint"][lit])[:, None]
        assert np.all(np.abs(end[:, 0] - light[0]) <= 0.5 + 1e-3) and np.all(np.abs(end[:, 2] - light[2]) < 1e-2)
        assert p.shape[0] == hit.sum()
    live = dev.download(counter, np.uint32, 2)[0]
    assert live > 0


def test_session_same_alone_and_in_mix(gpu_device, serve_scenes):
    _case("alone_vs_mix")
    dev = gpu_device
    _, scs = serve_scenes
    kw = dict(width=96, height=96, max_bounces=4, opts=D.trace_opts(256, 8, 0, 12))
    target = serve.SessionSpec(0, 2, 777, 0.0)
    alone = serve.Server(dev, scs, record={0}, **kw)
    alone.run([target])
    mix = serve.Server(dev, scs, record={0}, max_rays_per_tick=6 * 2 * 96 * 96, **kw)
    others = [serve.SessionSpec(i, i % 4, 1000 + i, 0.0005 * i) for i in range(1, 8)]
    mix.run([target] + others)
    a, b = alone.recorded_steps(0), mix.recorded_steps(0)
    # 5 closest-hit wavefronts (camera + 4 bounces), shadow rays from step 1 to 5
    assert [x is not None for x, _ in a] == [True] * 5 + [False]
    assert [y is not None for _, y in a] == [False] + [True] * 5
    # the shaded image is the same alone and in the mix, and something was lit
    ra, rb = alone.radiance(0), mix.radiance(0)
    assert np.array_equal(ra, rb)
    assert (ra.sum(axis=1) > 0).mean() > 0.05
    for k, (pa, pb) in enumerate(zip(a, b)):
        for x, y in ((pa[0], pb[0]), (pa[1], pb[1])):
            assert (x is None) == (y is None), k
            if x is not None:
                assert np.array_equal(x[0], y[0]), k   # rays
                assert np.array_equal(x[1], y[1]), k   # hits
    st = mix.stats()
    assert st["sessionsDone"] == 8 and st["maxActive"] >= 2
    alone.close()
    mix.close()


def test_admission_respects_hbm_budget(gpu_device, serve_scenes):
    _case("admission")
    dev = gpu_device
    _, scs = serve_scenes
    budget = 2 * serve.session_bytes(64, 64) + 100
    srv = serve.Server(dev, scs, width=64, height=64, max_bounces=3, hbm_budget=budget)
    specs = [serve.SessionSpec(i, i % 4, 50 + i, 0.0) for i in range(6)]
    srv.run(specs)
    st = srv.stats()
    assert st["sessionsDone"] == 6 and st["maxActive"] == 2 and st["sessionBuffersAllocated"] == 2
    # each session: camera wavefront + 3 bounce wavefronts + the last shadow wavefront
    assert st["latencySamples"] == 6 * 5 and len(srv.latencies_ms()) == 6 * 5
    srv.close()


def test_server_rejects_bad_arguments(gpu_device, serve_scenes):
    _case("server_errors")
    _, scs = serve_scenes
    srv = serve.Server(gpu_device, scs, width=32, height=32, max_bounces=1)
    with pytest.raises(D.LrhError):
        srv.submit([serve.SessionSpec(0, 99, 1, 0.0)])        # no such scene
    srv.submit([serve.SessionSpec(1, 0, 1, 0.0)])
    with pytest.raises(D.LrhError):
        srv.submit([serve.SessionSpec(1, 0, 2, 0.0)])         # id already used
    srv.run()
    assert srv.recorded(1, 0, 0) is None                       # not recorded
    with pytest.raises(D.LrhError):
        srv.plain_batch()                                      # nothing captured
    srv.close()
