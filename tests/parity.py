"""Hit-record comparison used by the parity tests.

Rule (from the acceptance bar of the device): hit/miss and (meshIndex,
triangleIndex) must be identical except for rays whose hit barycentrics lie
within 1e-6 of a triangle edge (the edge band, counted) or whose two hits are
an exact tie in t (both triangles accept the ray at the same t; which one is
kept depends on visit order).  Where the indices agree, t, b1 and b2 must
agree within 1e-5 relative.
"""
import numpy as np

from luxcore_amd.raytypes import NULL_INDEX

EDGE_BAND = 1e-6
REL_TOL = 1e-5

# Motion band.  A motion object's ray is moved into object space by a matrix
# sampled with acosf/sinf (quaternion slerp, quaternion.cpp:143-164); the
# device's last bits differ from the host libm's (glibc itself is not
# correctly rounded: ~6% of acosf and ~1% of sinf results), and with the ray
# origin ~100 object radii away a 1-2 ulp matrix difference moves the object-
# space ray by ~1e-4 of a triangle's size.  For hits on motion objects the
# comparison therefore uses an edge band of MOTION_EDGE in barycentric space
# and |dt| <= MOTION_T_ABS + REL_TOL*|t|, |db| <= MOTION_B_ABS; those rays are
# counted separately ("motion_band").
MOTION_EDGE = 1e-2
MOTION_T_ABS = 1e-3
MOTION_B_ABS = 1e-2


def _min_bary(h):
    b1 = h["b1"].astype(np.float64)
    b2 = h["b2"].astype(np.float64)
    return np.minimum(np.minimum(b1, b2), 1.0 - b1 - b2)


def compare(ref, got, tie_check=None, rel_tol=REL_TOL, motion_objects=None, mint=None):
    """Returns a dict of counts; tie_check(i) -> True when ray i is an exact
    tie (the got triangle is hit at the reference t).  motion_objects: mesh
    (object) indices whose hits are compared within the motion band; mint:
    per-ray mint (a motion ray whose only hit lies within MOTION_T_ABS of mint
    may flip to a miss)."""
    assert ref.shape == got.shape
    rhit = ref["meshIndex"] != NULL_INDEX
    ghit = got["meshIndex"] != NULL_INDEX
    same_idx = (ref["meshIndex"] == got["meshIndex"]) & (ref["triangleIndex"] == got["triangleIndex"])
    same_idx &= rhit == ghit
    mism = np.nonzero(~same_idx)[0]
    edge = np.zeros(ref.shape[0], bool)
    edge[rhit] |= _min_bary(ref[rhit]) < EDGE_BAND
    edge[ghit] |= _min_bary(got[ghit]) < EDGE_BAND
    motion = np.zeros(ref.shape[0], bool)
    if motion_objects is not None and len(motion_objects):
        mo = np.asarray(sorted(motion_objects), np.uint32)
        motion = np.isin(ref["meshIndex"], mo) | np.isin(got["meshIndex"], mo)
    medge = np.zeros(ref.shape[0], bool)
    medge[rhit] |= _min_bary(ref[rhit]) < MOTION_EDGE
    medge[ghit] |= _min_bary(got[ghit]) < MOTION_EDGE
    n_edge = n_tie = n_bad = n_motion = 0
    bad = []
    for i in mism:
        if edge[i]:
            n_edge += 1
        elif motion[i] and (medge[i] or (rhit[i] and ghit[i] and abs(float(ref["t"][i]) - float(got["t"][i]))
                                         <= MOTION_T_ABS + rel_tol * abs(float(ref["t"][i])))
                            or (mint is not None and (rhit[i] != ghit[i]) and
                                min(float(ref["t"][i]), float(got["t"][i])) <= float(mint[i]) + MOTION_T_ABS)):
            n_motion += 1
        elif rhit[i] and ghit[i] and ref["t"][i] == got["t"][i] and (tie_check is None or tie_check(i)):
            n_tie += 1
        else:
            n_bad += 1
            if len(bad) < 10:
                bad.append(int(i))
    both = np.nonzero(same_idx & rhit & ~motion)[0]
    mboth = np.nonzero(same_idx & rhit & motion)[0]

    def rel(a, b):
        a = a.astype(np.float64)
        b = b.astype(np.float64)
        return np.abs(a - b) / np.maximum(np.abs(a), 1e-30)
    worst = {}
    for f in ("t", "b1", "b2"):
        r = rel(ref[f][both], got[f][both]) if both.size else np.zeros(1)
        # b1/b2 near 0 are compared absolutely at the tolerance scale
        if f != "t" and both.size:
            r = np.minimum(r, np.abs(ref[f][both].astype(np.float64) - got[f][both].astype(np.float64)))
        worst[f] = float(r.max()) if r.size else 0.0
    # motion hits with matching indices: excess over the motion tolerances
    motion_excess = 0.0
    if mboth.size:
        dt = np.abs(ref["t"][mboth].astype(np.float64) - got["t"][mboth])
        lim = MOTION_T_ABS + rel_tol * np.abs(ref["t"][mboth].astype(np.float64))
        db = np.maximum(np.abs(ref["b1"][mboth].astype(np.float64) - got["b1"][mboth]),
                        np.abs(ref["b2"][mboth].astype(np.float64) - got["b2"][mboth]))
        motion_excess = float(max((dt - lim).max(), (db - MOTION_B_ABS).max(), 0.0))
    mt_r, mt_g = ref["t"][~rhit & ~ghit], got["t"][~rhit & ~ghit]
    miss_t_ok = bool(np.all((mt_g == mt_r) | (np.isnan(mt_g) & np.isnan(mt_r))))  # t = maxt, NaN included
    return {"rays": int(ref.shape[0]), "hits": int(rhit.sum()), "mismatch": int(mism.size),
            "edge_band": n_edge, "ties": n_tie, "motion_band": n_motion, "bad": n_bad, "bad_examples": bad,
            "worst_rel": worst, "motion_hits": int(mboth.size), "motion_excess": motion_excess,
            "miss_t_ok": miss_t_ok, "edge_band_rays": int(edge.sum())}


def assert_parity(stats, rel_tol=REL_TOL):
    assert stats["bad"] == 0, stats
    assert stats["miss_t_ok"], stats
    assert stats.get("motion_excess", 0.0) == 0.0, stats
    for f, v in stats["worst_rel"].items():
        assert v <= rel_tol, (f, stats)


def tri_tie_checker(oracle_mod, meshes, rays, ref, got):
    """tie check: the got triangle accepts the ray at exactly the ref t."""
    def check(i):
        m = int(got["meshIndex"][i])
        t = int(got["triangleIndex"][i])
        tv = meshes[m].tris[t]
        v = meshes[m].verts
        hit, tt, _, _ = oracle_mod.triangle_intersect(rays[i], v[tv[0]], v[tv[1]], v[tv[2]])
        return hit and tt == ref["t"][i]
    return check
