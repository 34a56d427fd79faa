"""The oracle against the closed-form fixtures of tests/golden (CPU only)."""
import json
import math
import os

import numpy as np
import pytest

import oracle
from luxcore_amd.raytypes import NULL_INDEX, make_rays
from luxcore_amd.scenes import Mesh

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _ray(c):
    maxt = np.inf if c["maxt"] == "inf" else c["maxt"]
    return make_rays([c["o"]], [c["d"]], mint=c["mint"], maxt=maxt)[0]


def _ulps(a, b):
    a = np.float32(a)
    b = np.float32(b)
    return abs(int(a.view(np.int32)) - int(b.view(np.int32)))


@pytest.mark.parametrize("case", load("triangle_kat.json"), ids=lambda c: c["name"])
def test_triangle_known_answers(case):
    hit, t, b1, b2 = oracle.triangle_intersect(_ray(case), case["p0"], case["p1"], case["p2"])
    assert hit == case["hit"]
    if hit:
        tol = 0 if case["exact"] else 2
        assert _ulps(t, case["t"]) <= tol
        assert _ulps(b1, case["b1"]) <= tol + (0 if case["b1"] != 0 else 0)
        assert _ulps(b2, case["b2"]) <= tol


def test_accelerator_rejects_hit_at_maxt():
    """Triangle::Intersect is inclusive at maxt but BVHAccel keeps only
    t < best with best = maxt (bvhaccel.cpp:173, 233): a hit at exactly maxt
    is a miss at the accelerator level, with t = maxt."""
    m = Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32))
    r = make_rays([[0.25, 0.25, 1.0], [0.25, 0.25, 1.0]], [[0, 0, -1], [0, 0, -1]], mint=0.0,
                  maxt=np.array([1.0, 1.5], np.float32))
    for builder in ("CLASSIC", "SAH"):
        h = oracle.OracleBVH([m], builder=builder).intersect(r)
        assert h["meshIndex"][0] == NULL_INDEX and h["t"][0] == np.float32(1.0)
        assert h["meshIndex"][1] == 0 and h["t"][1] == np.float32(1.0)
    h = oracle.bruteforce([m], r)
    assert h["meshIndex"][0] == NULL_INDEX and h["meshIndex"][1] == 0


@pytest.mark.parametrize("case", load("bbox_kat.json"), ids=lambda c: c["name"])
def test_bbox_known_answers(case):
    hit, t0, t1 = oracle.bbox_intersectp(_ray(case), case["min"], case["max"])
    assert hit == case["hit"]
    if hit:
        assert t0 == case["t0"] and t1 == case["t1"]


def test_epsilon_known_answers():
    g = load("epsilon_kat.json")
    for c in g["values"]:
        assert oracle.epsilon(c["v"]) == np.float32(c["E"]), c
    for c in g["boxes"]:
        assert oracle.epsilon_bbox(c["min"], c["max"]) == np.float32(c["E"]), c
    # E(1.0) is 128 ulps of 1.0 = 2^-16
    assert oracle.epsilon(1.0) == 2.0 ** -16


def test_matrix_inverse_known_answers():
    g = load("transform_kat.json")
    for c in g["inverse"]:
        inv = oracle.matrix_inverse(np.array(c["m"], np.float32).reshape(4, 4))
        assert np.array_equal(inv, np.array(c["inv"], np.float32).reshape(4, 4))
    rng = np.random.Generator(np.random.PCG64(0))
    for _ in range(20):
        m = (rng.normal(size=(4, 4)) + 4 * np.eye(4)).astype(np.float32)
        inv = oracle.matrix_inverse(m)
        assert np.allclose(m.astype(np.float64) @ inv.astype(np.float64), np.eye(4), atol=1e-5)
    with pytest.raises(ValueError):
        oracle.matrix_inverse(np.zeros((4, 4), np.float32))


@pytest.mark.parametrize("case", load("transform_kat.json")["motion"], ids=lambda c: c["name"])
def test_motion_known_answers(case):
    keys = np.array(case["keys"], np.float32).reshape(-1, 4, 4)
    m = oracle.motion_sample(case["times"], keys, case["t"])
    exp = np.array(case["m"], np.float32).reshape(4, 4)
    if case["exact"]:
        assert np.array_equal(m, exp), (m, exp)
    else:
        assert np.allclose(m, exp, atol=2e-6), (m, exp)


def test_motion_rotation_angle_is_linear():
    """Slerp between identity and a rotation about z by theta gives a
    rotation by t*theta (closed form), for several t."""
    th = 1.2
    c, s = math.cos(th), math.sin(th)
    R = np.array([[c, s, 0, 0], [-s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32)
    keys = np.stack([np.eye(4, dtype=np.float32), R])
    for t in (0.1, 0.25, 0.5, 0.9):
        m = oracle.motion_sample([0.0, 1.0], keys, t)
        a = t * th
        exp = np.array([[math.cos(a), math.sin(a), 0, 0], [-math.sin(a), math.cos(a), 0, 0], [0, 0, 1, 0],
                        [0, 0, 0, 1]])
        assert np.allclose(m, exp, atol=3e-6)


def test_motion_trs_endpoints_roundtrip():
    """Keys T*R*S (shear-free): sampling at the key times returns the keys
    exactly; midway the rotation part stays orthonormal and the scale lerps."""
    rng = np.random.Generator(np.random.PCG64(4))
    from luxcore_amd.scenes import random_trs
    for _ in range(10):
        a, b = random_trs(rng), random_trs(rng)
        keys = np.stack([a, b])
        assert np.array_equal(oracle.motion_sample([0, 1], keys, 0.0), a)
        assert np.array_equal(oracle.motion_sample([0, 1], keys, 1.0), b)
        mid = oracle.motion_sample([0, 1], keys, 0.5).astype(np.float64)
        sa = np.linalg.norm(a[0, :3].astype(np.float64))
        sb = np.linalg.norm(b[0, :3].astype(np.float64))
        rows = np.linalg.norm(mid[:3, :3], axis=1)
        assert np.allclose(rows, 0.5 * (sa + sb), rtol=1e-5)
        assert np.allclose(mid[:3, 3], 0.5 * (a[:3, 3].astype(np.float64) + b[:3, 3]), rtol=1e-5, atol=1e-4)


def test_motion_bound_contains_samples():
    rng = np.random.Generator(np.random.PCG64(6))
    from luxcore_amd.scenes import random_trs
    a, b = random_trs(rng, 10.0), random_trs(rng, 10.0)
    keys = np.stack([np.linalg.inv(a.astype(np.float64)), np.linalg.inv(b.astype(np.float64))]).astype(np.float32)
    lo, hi = oracle.motion_bound([0, 1], keys, [-1, -1, -1], [1, 1, 1])
    corners = np.array([[x, y, z, 1] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float64)
    for t in np.linspace(0, 1, 17):
        g2l = oracle.motion_sample([0, 1], keys, t).astype(np.float64)
        pts = (np.linalg.inv(g2l) @ corners.T).T[:, :3]
        assert np.all(pts >= lo - 1e-3) and np.all(pts <= hi + 1e-3)


# MotionSystem::Bound (src/luxrays/core/geometry/motionsystem.cpp:75-87,
# 356-365) is the union over i = 0..1024 of Sample(Lerp(i/1024, start, end))
# inverted and applied to the box.  Closed form: a point at radius 1000 from
# the z axis, rotated about z through 170 degrees over [0, 1] (keys are
# global->local, so they are the inverse rotations).  Its bound is the box of
# the 1,025 points at angles -85 + 170 i/1024 degrees (plus the point's own
# angle), computed here in float64.  Probe A puts the largest x exactly on
# sample i = 513, which neither 1,023 nor 16 samples contain; probe B puts it
# midway between samples 512 and 513, where a continuous (swept) bound would
# reach the full radius.
BOUND_R = 1000.0
BOUND_TH = np.deg2rad(170.0)
BOUND_TOL = 2.5e-4   # 4 ulps at 1000; the float evaluation stays within 1.2 ulps


def _rotz(a):
    c, s = np.cos(a), np.sin(a)
    m = np.eye(4)
    m[0, 0], m[0, 1], m[1, 0], m[1, 1] = c, -s, s, c
    return m


def bound_probe(tstar):
    """(keys, point, expected box for N samples) of a probe whose largest x
    is reached at time tstar."""
    keys = np.stack([np.linalg.inv(_rotz(-BOUND_TH / 2)), np.linalg.inv(_rotz(BOUND_TH / 2))]).astype(np.float32)
    alpha = BOUND_TH / 2 - BOUND_TH * tstar
    p = np.array([BOUND_R * np.cos(alpha), BOUND_R * np.sin(alpha), 0.0], np.float32)

    def expected(n=1024, swept=False):
        ang = -BOUND_TH / 2 + BOUND_TH * np.arange(n + 1) / n + np.arctan2(float(p[1]), float(p[0]))
        r = float(np.hypot(float(p[0]), float(p[1])))
        x, y = r * np.cos(ang), r * np.sin(ang)
        lo, hi = np.array([x.min(), y.min(), 0.0]), np.array([x.max(), y.max(), 0.0])
        if swept:
            hi[0] = r   # the arc passes angle 0
        return lo, hi
    return keys, p, expected


BOUND_PROBES = {"A": 513 / 1024, "B": 512.5 / 1024}


def _bound_err(lo, hi, e):
    return max(np.abs(np.asarray(lo, np.float64) - e[0]).max(), np.abs(np.asarray(hi, np.float64) - e[1]).max())


@pytest.mark.parametrize("probe", sorted(BOUND_PROBES))
def test_motion_bound_closed_form(probe):
    keys, p, expected = bound_probe(BOUND_PROBES[probe])
    lo, hi = oracle.motion_bound([0, 1], keys, p, p)
    assert _bound_err(lo, hi, expected(1024)) <= BOUND_TOL, (lo, hi, expected(1024))
    # what a wrong sample count or a swept bound would give is outside the tolerance
    wrong = {"A": [expected(1023), expected(16)], "B": [expected(1023), expected(swept=True)]}[probe]
    for e in wrong:
        assert _bound_err(lo, hi, e) > 2 * BOUND_TOL
