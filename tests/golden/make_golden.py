"""Writes the known-answer fixtures of tests/golden/*.json.

The reference ships no golden vectors for its intersection path (its tests are
image diffs with self-generated references), so every expected value here is
fixed by mathematics, independently of the restated code:

* triangle: dyadic rays and triangles; the expected (t, b1, b2) solve
  o + t*d = p0 + b1*(p1-p0) + b2*(p2-p0) exactly (Cramer's rule over
  rationals), so they are exact in fp32; hit/miss follows from the signs of
  b0, b1, b2, t and the [mint, maxt] window;
* bbox: slab entry/exit distances of axis-aligned boxes, computed per axis by
  hand, including zero direction components;
* epsilon: E(v) = clamp(128 * ulp(v), 1e-5, 0.1) for binade-interior v;
* matrix/motion: inverses of diagonal/translation matrices and the
  endpoints / translation / scale / rotation-about-z midpoints of motion
  systems.

Run: python tests/golden/make_golden.py   (rewrites the JSON files)
"""
import json
import math
import os
from fractions import Fraction as Fr

HERE = os.path.dirname(os.path.abspath(__file__))


def det3(a, b, c):
    return (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0])
            + a[2] * (b[0] * c[1] - b[1] * c[0]))


def solve(o, d, p0, p1, p2):
    """t, b1, b2 with o + t d = p0 + b1 e1 + b2 e2 (exact rationals)."""
    e1 = [Fr(p1[i]) - Fr(p0[i]) for i in range(3)]
    e2 = [Fr(p2[i]) - Fr(p0[i]) for i in range(3)]
    rhs = [Fr(o[i]) - Fr(p0[i]) for i in range(3)]
    nd = [-Fr(x) for x in d]
    # columns: e1, e2, -d ; unknowns b1, b2, t
    cols = [e1, e2, nd]
    M = [[cols[c][r] for c in range(3)] for r in range(3)]
    D = det3(*M)
    if D == 0:
        return None
    res = []
    for k in range(3):
        Mk = [[(rhs[r] if c == k else cols[c][r]) for c in range(3)] for r in range(3)]
        res.append(det3(*Mk) / D)
    b1, b2, t = res
    return t, b1, b2, D


def tri_case(name, o, d, p0, p1, p2, mint=0.0, maxt=math.inf):
    s = solve(o, d, p0, p1, p2)
    exact = True
    if s is None:
        hit, t, b1, b2 = False, None, None, None
    else:
        t, b1, b2, D = s
        # a power-of-two determinant makes the single division exact in fp32
        a = abs(D)
        exact = a.numerator & (a.numerator - 1) == 0 and a.denominator & (a.denominator - 1) == 0
        b0 = 1 - b1 - b2
        hit = b1 >= 0 and b2 >= 0 and b0 >= 0 and t >= Fr(mint) and (maxt == math.inf or t <= Fr(maxt))
    return {"name": name, "o": o, "d": d, "p0": p0, "p1": p1, "p2": p2, "mint": mint,
            "maxt": "inf" if maxt == math.inf else maxt, "hit": bool(hit),
            "t": float(t) if hit else None, "b1": float(b1) if hit else None, "b2": float(b2) if hit else None,
            "exact": bool(exact)}


def triangles():
    U = ([0, 0, 0], [1, 0, 0], [0, 1, 0])
    return [
        tri_case("unit_center", [0.25, 0.25, 1], [0, 0, -1], *U),
        tri_case("unit_outside", [0.75, 0.75, 1], [0, 0, -1], *U),
        tri_case("unit_parallel", [0.25, 0.25, 0], [1, 0, 0], *U),
        tri_case("unit_behind", [0.25, 0.25, -1], [0, 0, -1], *U),
        tri_case("unit_t_eq_mint", [0.25, 0.25, 1], [0, 0, -1], *U, mint=1.0),
        tri_case("unit_t_eq_maxt", [0.25, 0.25, 1], [0, 0, -1], *U, maxt=1.0),
        tri_case("unit_t_beyond_maxt", [0.25, 0.25, 1], [0, 0, -1], *U, maxt=0.5),
        tri_case("unit_hypotenuse", [0.5, 0.5, 1], [0, 0, -1], *U),
        tri_case("unit_vertex", [0, 0, 1], [0, 0, -1], *U),
        tri_case("unit_from_below", [0.125, 0.5, -2], [0, 0, 0.5], *U),
        tri_case("scaled", [2, 3, 7], [0, 0, -2], [1, 2, 3], [3, 2, 3], [1, 6, 3]),
        tri_case("slanted", [0.5, 0.25, 4], [0.25, 0.5, -2], [0, 0, 0], [4, 0, 0], [0, 4, 1]),
        tri_case("reversed_winding", [0.25, 0.25, 1], [0, 0, -1], [0, 0, 0], [0, 1, 0], [1, 0, 0]),
    ]


def bboxes():
    inf = "inf"
    return [
        {"name": "straight", "o": [0.5, 0.5, -1], "d": [0, 0, 1], "min": [0, 0, 0], "max": [1, 1, 1],
         "mint": 0.0, "maxt": inf, "hit": True, "t0": 1.0, "t1": 2.0},
        {"name": "diagonal", "o": [-1, -1, -1], "d": [1, 1, 1], "min": [0, 0, 0], "max": [1, 1, 1],
         "mint": 0.0, "maxt": inf, "hit": True, "t0": 1.0, "t1": 2.0},
        {"name": "miss_side", "o": [2, 0.5, -1], "d": [0, 0, 1], "min": [0, 0, 0], "max": [1, 1, 1],
         "mint": 0.0, "maxt": inf, "hit": False},
        {"name": "touch_t0_eq_t1", "o": [-1, 0, 0.5], "d": [1, 1, 0], "min": [0, 0, 0], "max": [1, 1, 1],
         "mint": 0.0, "maxt": inf, "hit": True, "t0": 1.0, "t1": 1.0},
        {"name": "origin_on_slab_zero_dir", "o": [0, 0.5, -1], "d": [0, 0, 1], "min": [0, 0, 0],
         "max": [1, 1, 1], "mint": 0.0, "maxt": inf, "hit": True, "t0": 1.0, "t1": 2.0},
        {"name": "inside", "o": [0.5, 0.5, 0.5], "d": [0, 0, 1], "min": [0, 0, 0], "max": [1, 1, 1],
         "mint": 0.0, "maxt": inf, "hit": True, "t0": 0.0, "t1": 0.5},
        {"name": "behind", "o": [0.5, 0.5, 2], "d": [0, 0, 1], "min": [0, 0, 0], "max": [1, 1, 1],
         "mint": 0.0, "maxt": inf, "hit": False},
        {"name": "maxt_short", "o": [0.5, 0.5, -1], "d": [0, 0, 1], "min": [0, 0, 0], "max": [1, 1, 1],
         "mint": 0.0, "maxt": 0.5, "hit": False},
        {"name": "negative_dir", "o": [0.5, 3, 0.5], "d": [0, -2, 0], "min": [0, 0, 0], "max": [1, 1, 1],
         "mint": 0.0, "maxt": inf, "hit": True, "t0": 1.0, "t1": 1.5},
    ]


def ulp(v):
    return 2.0 ** (math.floor(math.log2(abs(v))) - 23)


def epsilons():
    out = []
    for v in [1.0, -1.0, 2.0, 100.0, 0.5, 3.0, 1024.0, 1e6, 0.0, 1e-30, 65536.0]:
        if v == 0.0 or abs(v) < 1e-30 * 2:
            e = 1e-5
        else:
            e = min(max(128 * ulp(v), 1e-5), 0.1)
        out.append({"v": v, "E": e})
    boxes = [{"min": [-1, -2, 0], "max": [100, 0.5, 3], "E": 2.0 ** -10},
             {"min": [0, 0, 0], "max": [1, 1, 1], "E": 2.0 ** -16},
             {"min": [0.001, 0.001, 0.001], "max": [0.002, 0.002, 0.002], "E": 1e-5}]
    return {"values": out, "boxes": boxes}


def mat(rows):
    return [float(x) for r in rows for x in r]


def transforms():
    diag = [[2, 0, 0, 0], [0, 4, 0, 0], [0, 0, 8, 0], [0, 0, 0, 1]]
    diag_inv = [[0.5, 0, 0, 0], [0, 0.25, 0, 0], [0, 0, 0.125, 0], [0, 0, 0, 1]]
    tr = [[1, 0, 0, 3], [0, 1, 0, -5], [0, 0, 1, 0.5], [0, 0, 0, 1]]
    tr_inv = [[1, 0, 0, -3], [0, 1, 0, 5], [0, 0, 1, -0.5], [0, 0, 0, 1]]
    c, s = math.cos(math.pi / 4), math.sin(math.pi / 4)
    I = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]]
    rz90 = [[0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]]
    rz45 = [[c, s, 0, 0], [-s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]]
    t0 = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]]
    t1 = [[1, 0, 0, 2], [0, 1, 0, 4], [0, 0, 1, 8], [0, 0, 0, 1]]
    tmid = [[1, 0, 0, 1], [0, 1, 0, 2], [0, 0, 1, 4], [0, 0, 0, 1]]
    s1 = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]]
    s3 = [[3, 0, 0, 0], [0, 3, 0, 0], [0, 0, 3, 0], [0, 0, 0, 1]]
    s2 = [[2, 0, 0, 0], [0, 2, 0, 0], [0, 0, 2, 0], [0, 0, 0, 1]]
    return {
        "inverse": [{"m": mat(diag), "inv": mat(diag_inv)}, {"m": mat(tr), "inv": mat(tr_inv)}],
        "motion": [
            {"name": "translation_mid", "times": [0, 1], "keys": [mat(t0), mat(t1)], "t": 0.5, "m": mat(tmid),
             "exact": True},
            {"name": "translation_before", "times": [0, 1], "keys": [mat(t0), mat(t1)], "t": -1.0, "m": mat(t0),
             "exact": True},
            {"name": "translation_after", "times": [0, 1], "keys": [mat(t0), mat(t1)], "t": 3.0, "m": mat(t1),
             "exact": True},
            {"name": "scale_mid", "times": [0, 1], "keys": [mat(s1), mat(s3)], "t": 0.5, "m": mat(s2), "exact": True},
            {"name": "rotz_mid", "times": [0, 1], "keys": [mat(I), mat(rz90)], "t": 0.5, "m": mat(rz45),
             "exact": False},
            {"name": "static_keys", "times": [0, 1], "keys": [mat(tr), mat(tr)], "t": 0.3, "m": mat(tr),
             "exact": True},
            {"name": "three_keys_second_segment", "times": [0, 1, 2], "keys": [mat(t0), mat(t0), mat(t1)],
             "t": 1.5, "m": mat(tmid), "exact": True},
        ],
    }


def main():
    for name, data in (("triangle_kat.json", triangles()), ("bbox_kat.json", bboxes()),
                       ("epsilon_kat.json", epsilons()), ("transform_kat.json", transforms())):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, indent=1)
        print("wrote", name)


if __name__ == "__main__":
    main()
