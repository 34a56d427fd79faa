"""Edge-case rays and scenes against the oracle: NaN components, zero and
axis-aligned directions, mint/maxt corner cases (negative mint, mint >= maxt,
maxt = 0), far-from-origin coordinates, degenerate and duplicated triangles,
masked rays mixed in.  The reference semantics decide every case: a hit is
kept iff mint <= t < best (best starts at maxt); a miss reports t = maxt."""
import numpy as np
import pytest

import oracle
from luxcore_amd import scenes
from luxcore_amd.raytypes import HIT_DTYPE, NULL_INDEX, RAY_DTYPE, make_rays
from tests.parity import assert_parity, compare, tri_tie_checker

pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _special_rays(rng, n=4096):
    r = scenes.random_rays(n, (0, 0, 0), (10, 10, 10), rng)
    k = n // 16
    # NaN in each origin / direction component
    for c in range(3):
        r["o"][c * k:(c + 1) * k // 2, c] = np.nan
        r["d"][(c + 3) * k:(c + 3) * k + k // 2, c] = np.nan
    # NaN mint (the CPU test !(t < mint) then accepts any t < maxt) and NaN maxt
    r["mint"][6 * k:7 * k] = np.nan
    r["maxt"][7 * k:8 * k] = np.nan
    # zero direction
    r["d"][8 * k:8 * k + 64] = 0.0
    # axis-aligned directions, signed zeros
    for j, d in enumerate(([1, 0, 0], [-1, 0, 0], [0, 1, -0.0], [0, 0, -1], [-0.0, -1, 0])):
        r["d"][9 * k + 32 * j:9 * k + 32 * (j + 1)] = np.array(d, np.float32)
    # origins exactly on the box planes with a zero direction component
    r["o"][10 * k:10 * k + 64, 0] = 0.0
    r["d"][10 * k:10 * k + 64, 0] = 0.0
    # negative mint: hits behind the origin count
    r["mint"][11 * k:12 * k] = -5.0
    # mint >= maxt, maxt == mint, maxt == 0
    r["mint"][12 * k:12 * k + 64] = 3.0
    r["maxt"][12 * k:12 * k + 64] = 1.0
    r["mint"][12 * k + 64:12 * k + 128] = 2.0
    r["maxt"][12 * k + 64:12 * k + 128] = 2.0
    r["maxt"][12 * k + 128:12 * k + 192] = 0.0
    # short finite segments
    r["maxt"][13 * k:14 * k] = rng.uniform(0.01, 3.0, k).astype(np.float32)
    # masked
    r["flags"][14 * k:15 * k:3] = 1
    return r


def _compare(dev, meshes, rays, ref=None):
    sc = dev.scene_bvh(meshes)
    init = np.zeros(rays.shape[0], HIT_DTYPE)
    init["t"] = -7.0
    got = dev.trace_numpy(sc, rays, hits_init=init)
    live = (rays["flags"] & 1) == 0
    assert np.all(got["t"][~live] == -7.0)
    if ref is None:
        ref = oracle.bruteforce(meshes, rays[live])
    st = compare(ref, got[live], tie_check=tri_tie_checker(oracle, meshes, rays[live], ref, got[live]))
    print(st, flush=True)
    assert_parity(st)
    sc.close()
    return got[live], ref


def test_special_rays_plain_scene(gpu_device):
    _case("special_rays")
    rng = np.random.default_rng(71)
    meshes = [scenes.box_mesh(), scenes.icosphere_l6(2.0, (5.0, 5.0, 5.0)),
              scenes.random_triangles(3000, (0, 0, 0), (10, 10, 10), rng)]
    rays = _special_rays(rng)
    got, ref = _compare(gpu_device, meshes, rays)
    live = rays[(rays["flags"] & 1) == 0]
    nan_od = np.isnan(live["o"]).any(1) | np.isnan(live["d"]).any(1)
    assert np.all(got["meshIndex"][nan_od] == NULL_INDEX)
    # negative mint produced hits behind some origins
    behind = (live["mint"] < 0) & (got["meshIndex"] != NULL_INDEX) & (got["t"] < 0)
    assert behind.sum() > 10
    # NaN mint still hits (reference: !(t < NaN) holds)
    nan_mint = np.isnan(live["mint"])
    assert (got["meshIndex"][nan_mint] != NULL_INDEX).sum() > 10


def test_special_rays_two_level(gpu_device):
    """The same special rays through MBVHAccel (instances and 2-key motion):
    NaN or out-of-range ray times as well (MotionSystem::Sample clamps to the
    end keys; a NaN time samples the last segment with NaN weights)."""
    _case("special_rays_two_level")
    rng = np.random.default_rng(73)
    sc = scenes.instanced_scene(n_inst=40, motion_frac=0.25, freq=10, extent=8.0, seed=74)
    rays = _special_rays(rng, n=8192)
    rays["o"] = (rays["o"] - np.float32(5.0)).astype(np.float32)
    rays["time"] = rng.random(rays.shape[0], dtype=np.float32)
    k = rays.shape[0] // 16
    rays["time"][15 * k:15 * k + k // 3] = np.nan
    rays["time"][15 * k + k // 3:15 * k + 2 * k // 3] = -1.0
    rays["time"][15 * k + 2 * k // 3:16 * k] = 2.0
    init = np.zeros(rays.shape[0], HIT_DTYPE)
    init["t"] = -7.0
    got = gpu_device.trace_numpy(gpu_device.scene_mbvh(sc), rays, hits_init=init)
    live = (rays["flags"] & 1) == 0
    assert np.all(got["t"][~live] == -7.0)
    ref = oracle.OracleMBVH(sc).intersect(rays[live])
    st = compare(ref, got[live])
    print(st, flush=True)
    assert_parity(st)
    assert st["hits"] > 500


def test_far_from_origin(gpu_device):
    _case("far_coordinates")
    rng = np.random.default_rng(72)
    off = np.array([1.0e5, -2.0e5, 3.0e5], np.float32)
    base = [scenes.icosphere_l6(2.0, (5.0, 5.0, 5.0)), scenes.random_triangles(2000, (0, 0, 0), (10, 10, 10), rng)]
    meshes = [scenes.Mesh((m.verts + off).astype(np.float32), m.tris) for m in base]
    rays = scenes.random_rays(8192, (0, 0, 0), (10, 10, 10), rng)
    rays["o"] = (rays["o"] + off).astype(np.float32)
    _compare(gpu_device, meshes, rays)


def test_degenerate_and_duplicate_triangles(gpu_device):
    _case("degenerate_duplicates")
    rng = np.random.default_rng(73)
    soup = scenes.random_triangles(4000, (0, 0, 0), (10, 10, 10), rng)
    v = soup.verts.copy()
    t = soup.tris.copy()
    # collapse every 10th triangle to a segment or a point
    t[::10, 2] = t[::10, 1]
    t[5::20, 1] = t[5::20, 0]
    t[5::20, 2] = t[5::20, 0]
    # an exact duplicate mesh (every hit is a tie between the two copies)
    meshes = [scenes.Mesh(v, t), scenes.Mesh(v.copy(), t.copy()), scenes.box_mesh()]
    rays = scenes.random_rays(8192, (0, 0, 0), (10, 10, 10), rng)
    got, ref = _compare(gpu_device, meshes, rays)
    hit = got["meshIndex"] != NULL_INDEX
    # the box encloses every origin: everything hits (ties between the copies allowed)
    assert hit.mean() > 0.99
