"""Hit-record comparison used by the parity tests.

Rule (from the acceptance bar of the device): hit/miss and (meshIndex,
triangleIndex) must be identical except for rays whose hit barycentrics lie
within 1e-6 of a triangle edge (the edge band, counted) or whose two hits are
an exact tie in t (both triangles accept the ray at the same t; which one is
kept depends on visit order).  Where the indices agree, t, b1 and b2 must
agree within 1e-5 relative.
"""
import hashlib

import numpy as np

from luxcore_amd.raytypes import NULL_INDEX

EDGE_BAND = 1e-6
REL_TOL = 1e-5


def _min_bary(h):
    b1 = h["b1"].astype(np.float64)
    b2 = h["b2"].astype(np.float64)
    return np.minimum(np.minimum(b1, b2), 1.0 - b1 - b2)


def compare(ref, got, tie_check=None, rel_tol=REL_TOL, motion_objects=None):
    """Returns a dict of counts; tie_check(i) -> True when ray i is an exact
    tie (the got triangle is hit at the reference t).  motion_objects: object
    indices of moving objects; their hits are held to the same bar and also
    counted and measured on their own (motion_hits, motion_worst_dt,
    motion_worst_db: largest |dt| and |db| where the indices agree)."""
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
    n_edge = n_tie = n_bad = 0
    bad = []
    for i in mism:
        if edge[i]:
            n_edge += 1
        elif rhit[i] and ghit[i] and ref["t"][i] == got["t"][i] and (tie_check is None or tie_check(i)):
            n_tie += 1
        else:
            n_bad += 1
            if len(bad) < 10:
                bad.append(int(i))
    both = np.nonzero(same_idx & rhit)[0]

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
    mboth = np.nonzero(same_idx & rhit & motion)[0]
    m_dt = m_db = 0.0
    if mboth.size:
        m_dt = float(np.abs(ref["t"][mboth].astype(np.float64) - got["t"][mboth]).max())
        m_db = float(max(np.abs(ref["b1"][mboth].astype(np.float64) - got["b1"][mboth]).max(),
                         np.abs(ref["b2"][mboth].astype(np.float64) - got["b2"][mboth]).max()))
    mt_r, mt_g = ref["t"][~rhit & ~ghit], got["t"][~rhit & ~ghit]
    miss_t_ok = bool(np.all((mt_g == mt_r) | (np.isnan(mt_g) & np.isnan(mt_r))))  # t = maxt, NaN included
    return {"rays": int(ref.shape[0]), "hits": int(rhit.sum()), "mismatch": int(mism.size),
            "edge_band": n_edge, "ties": n_tie, "bad": n_bad, "bad_examples": bad,
            "worst_rel": worst, "motion_hits": int(mboth.size), "motion_worst_dt": m_dt, "motion_worst_db": m_db,
            "miss_t_ok": miss_t_ok, "edge_band_rays": int(edge.sum())}


def assert_parity(stats, rel_tol=REL_TOL):
    assert stats["bad"] == 0, stats
    assert stats["miss_t_ok"], stats
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


def chunk_digests(hits, chunk):
    """SHA-1 of each run of `chunk` hit records, normalized: a miss keeps its
    t and null indices, its barycentrics count as zero (the oracle and the
    device may leave different values there)."""
    h = np.array(hits, copy=True)
    miss = h["meshIndex"] == NULL_INDEX
    h["b1"][miss] = 0.0
    h["b2"][miss] = 0.0
    return [hashlib.sha1(h[i:i + chunk].tobytes()).hexdigest() for i in range(0, h.shape[0], chunk)]
