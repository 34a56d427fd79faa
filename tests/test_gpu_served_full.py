"""The served workload at its configured size, checked against the oracle.

Eight resident scenes of 10k..5M triangles (11.76M in all); sessions of
512x512 camera rays, up to 8 bounce wavefronts (cosine-diffuse, Russian
roulette with survival 0.7 after bounce 2) and one shadow ray per bounce
(any-hit), Poisson arrivals, run by the library's server (lrh_server_*).
Three sessions (on the 5M-, 1M- and 50k-triangle scenes) are recorded: every
wavefront they traced in the mix -- rays generated and shaded on the GPU,
hits written by the batched launch -- is checked on a 4,096-ray sample
against the oracle BVHAccel of its scene, and each is run again alone: every
wavefront's rays and hits and the final image must be the same bit for bit.

Every case prints its name first so a failing GPU run names it.
"""
import numpy as np
import pytest

import oracle
from luxcore_amd import scenes
from luxcore_amd import serve
from luxcore_amd.raytypes import HIT_DTYPE, NULL_INDEX, RAY_FLAGS_MASKED
from tests.parity import EDGE_BAND, assert_parity, compare, tri_tie_checker

pytestmark = pytest.mark.gpu

N_SAMPLE = 4096
RECORDED = {0: 7, 1: 4, 2: 1}   # session id -> scene (5M, 1M, 50k triangles)


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.fixture(scope="module")
def c5(gpu_device):
    meshes = scenes.batch_scenes()
    assert sum(m[0].ntris for m in meshes) == 11_760_000
    return meshes, [gpu_device.scene_bvh(m) for m in meshes]


def _specs():
    rec = [serve.SessionSpec(sid, sc, 9000 + sid, 0.002 * i) for i, (sid, sc) in enumerate(RECORDED.items())]
    return rec + serve.poisson_specs(600.0, 0.05, 8, seed=505, first_sid=10)


def _min_bary(h):
    b1 = h["b1"].astype(np.float64)
    b2 = h["b2"].astype(np.float64)
    return np.minimum(np.minimum(b1, b2), 1.0 - b1 - b2)


def _check_closest(meshes, ob, rays, hits, rng):
    live = np.nonzero((rays["flags"] & RAY_FLAGS_MASKED) == 0)[0]
    idx = np.sort(rng.choice(live, min(N_SAMPLE, live.size), replace=False)) if live.size else live
    r, g = rays[idx], hits[idx]
    ref = ob.intersect(r)
    st = compare(ref, g, tri_tie_checker(oracle, meshes, r, ref, g))
    assert_parity(st)
    return st


def _check_shadow(meshes, ob, rays, hits, rng):
    """Any-hit: occluded exactly where the oracle finds a hit (edge band
    excepted); every reported hit is a real hit of that triangle at that t,
    b1, b2, within [mint, maxt), and not before the closest hit."""
    live = np.nonzero((rays["flags"] & RAY_FLAGS_MASKED) == 0)[0]
    idx = np.sort(rng.choice(live, min(N_SAMPLE, live.size), replace=False)) if live.size else live
    r, g = rays[idx], hits[idx]
    ref = ob.intersect(r)
    rhit, ghit = ref["meshIndex"] != NULL_INDEX, g["meshIndex"] != NULL_INDEX
    diff = np.nonzero(rhit != ghit)[0]
    for i in diff:
        h = ref if rhit[i] else g
        assert _min_bary(h[i:i + 1])[0] < EDGE_BAND, (int(idx[i]), ref[i], g[i])
    for i in np.nonzero(ghit)[0]:
        m, t = int(g["meshIndex"][i]), int(g["triangleIndex"][i])
        tv = meshes[m].tris[t]
        v = meshes[m].verts
        ok, tt, b1, b2 = oracle.triangle_intersect(r[i], v[tv[0]], v[tv[1]], v[tv[2]])
        assert ok and tt == g["t"][i] and b1 == g["b1"][i] and b2 == g["b2"][i], int(idx[i])
        assert r["mint"][i] <= tt < r["maxt"][i]
        if rhit[i]:
            assert ref["t"][i] <= tt
    return {"rays": int(idx.size), "occluded": int(ghit.sum()), "edge_band": int(diff.size)}


def test_served_sessions_against_oracle_and_alone(gpu_device, c5):
    meshes, scs = c5
    specs = _specs()
    _case("c5_mix", sessions=len(specs), recorded=sorted(RECORDED))
    mix = serve.Server(gpu_device, scs, record=set(RECORDED), max_rays_per_tick=3 << 20)
    mix.run(specs, time_limit=120.0)
    st = mix.stats()
    print(st, flush=True)
    assert st["sessionsDone"] == len(specs) and st["sessionsQueued"] == 0
    assert st["maxActive"] >= 4
    # every session traces 10 wavefronts: the camera, 8 bounces, the last shadow rays
    assert st["latencySamples"] == 10 * len(specs)
    rng = np.random.default_rng(77)
    totals = {"closest": 0, "shadow": 0, "hits": 0, "occluded": 0, "edge_band": 0, "ties": 0}
    mixed = {}
    for sid, si in RECORDED.items():
        ob = oracle.OracleBVH(meshes[si])
        steps = mix.recorded_steps(sid)
        mixed[sid] = (steps, mix.radiance(sid))
        assert [c is not None for c, _ in steps] == [True] * 9 + [False]
        assert [s is not None for _, s in steps] == [False] + [True] * 9
        for k, (clo, sha) in enumerate(steps):
            _case("c5_wavefront", sid=sid, scene=si, step=k)
            for pair, kind in ((clo, "closest"), (sha, "shadow")):
                if pair is None:
                    continue
                rays, hits = pair
                masked = (rays["flags"] & RAY_FLAGS_MASKED) != 0
                # masked rays leave their (cleared) hit records untouched
                assert not hits[masked].view(np.uint8).any()
                if kind == "closest":
                    s = _check_closest(meshes[si], ob, rays, hits, rng)
                    totals["hits"] += s["hits"]
                    totals["edge_band"] += s["edge_band"]
                    totals["ties"] += s["ties"]
                else:
                    s = _check_shadow(meshes[si], ob, rays, hits, rng)
                    totals["occluded"] += s["occluded"]
                    totals["edge_band"] += s["edge_band"]
                totals[kind] += s["rays"]
        # Russian roulette from bounce 3 on: survival 0.7 of the paths alive before
        live = [int(((c[0]["flags"] & RAY_FLAGS_MASKED) == 0).sum()) for c, _ in steps[:9]]
        assert live[0] == 512 * 512 and all(b <= a for a, b in zip(live, live[1:])), live
        del ob
    print("oracle-checked rays", totals, flush=True)
    assert totals["closest"] > 50_000 and totals["hits"] > 10_000 and totals["occluded"] > 0
    mix.close()
    lit = 0
    for sid, si in RECORDED.items():
        _case("c5_alone", sid=sid, scene=si)
        spec = [s for s in specs if s.sid == sid][0]
        alone = serve.Server(gpu_device, scs, record={sid})
        alone.run([serve.SessionSpec(sid, si, spec.seed, 0.0)], time_limit=60.0)
        steps_a = alone.recorded_steps(sid)
        steps_m, rad_m = mixed[sid]
        for k, (pa, pm) in enumerate(zip(steps_a, steps_m)):
            for x, y in ((pa[0], pm[0]), (pa[1], pm[1])):
                assert (x is None) == (y is None), (sid, k)
                if x is not None:
                    assert np.array_equal(x[0], y[0]), (sid, k, "rays")
                    assert np.array_equal(x[1].view(np.uint8), y[1].view(np.uint8)), (sid, k, "hits")
        assert np.array_equal(alone.radiance(sid), rad_m), sid
        lit += int((rad_m.sum(axis=1) > 0).sum())
        alone.close()
    # the soups occlude most shadow rays, but some light reaches the cameras
    assert lit > 0
