"""One shading step of a path wavefront, restated on the CPU with numpy fp32
(operation for operation, no fused multiply-add): the checker for the
device's lrh_shade.

TEST INFRASTRUCTURE ONLY: only tests/ may import this module.

The step is the Lambertian + triangle-area-light reduction of SLG's
AdvancePaths micro-kernels (include/slg/engines/pathoclbase/kernels/
pathoclbase_kernels_micro.cl: MK_HIT_OBJECT :204-330, MK_DL_ILLUMINATE /
MK_DL_SAMPLE_BSDF :450-560, MK_GENERATE_NEXT_VERTEX_RAY :1045-1075) that
include/lrhip.h documents.  What pins it (tests/test_shade_cpu.py): the
next-event estimate of a Lambertian point below a rectangular light
averages to kd * Le * F, F the closed-form differential form factor of a
parallel rectangle, within its standard error; shadow-ray bookkeeping,
depth-0 emission and the cosine-weighted directions (mean cosine 2/3) are
checked on the same setup.
"""
import numpy as np

from luxcore_amd.raytypes import RAY_DTYPE

F = np.float32
NULL = np.uint32(0xFFFFFFFF)
INV_PI = F(0.31830988)
TWO_PI = F(6.2831853)


def _pcg(v):
    """The PCG-based integer hash of Jarzynski & Olano, "Hash Functions for
    GPU Rendering" (JCGT 2020), modulo 2^32."""
    with np.errstate(over="ignore"):
        v = np.asarray(v, np.uint32)
        s = (v * np.uint32(747796405) + np.uint32(2891336453)).astype(np.uint32)
        w = (((s >> ((s >> np.uint32(28)) + np.uint32(4))) ^ s) * np.uint32(277803737)).astype(np.uint32)
        return ((w >> np.uint32(22)) ^ w).astype(np.uint32)


def rnd(seed, pixel, bounce, dim):
    """(pcg(seed_lo ^ pcg(seed_hi ^ pcg(pixel ^ pcg(bounce*16 + dim)))) >> 8) / 2^24"""
    lo = np.uint32(seed & 0xFFFFFFFF)
    hi = np.uint32((seed >> 32) & 0xFFFFFFFF)
    b = (np.asarray(bounce, np.uint32) * np.uint32(16) + np.uint32(dim)).astype(np.uint32)
    h = _pcg(lo ^ _pcg(hi ^ _pcg(np.asarray(pixel, np.uint32) ^ _pcg(b))))
    return (h >> np.uint32(8)).astype(F) * F(1.0 / 16777216.0)


def machine_eps(v):
    v = np.asarray(v, F)
    n = (v.view(np.uint32) + np.uint32(0x80)).view(F)
    e = np.abs(n - v)
    return np.where(e > F(1e-5), np.where(e < F(1e-1), e, F(1e-1)), F(1e-5)).astype(F)


def _cross(a, b):
    return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], 1).astype(F)


def _dot(a, b):
    return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1] + a[:, 2] * b[:, 2]).astype(F)


def _masked(n):
    r = np.zeros(n, RAY_DTYPE)
    r["d"][:, 2] = 1.0
    r["flags"] = 1
    return r


def _cosine_dir(n, u1, u2):
    r = np.sqrt(u1).astype(F)
    phi = (TWO_PI * u2).astype(F)
    lx = (r * np.cos(phi).astype(F)).astype(F)
    ly = (r * np.sin(phi).astype(F)).astype(F)
    lz = np.sqrt(np.maximum(F(0), F(1) - u1)).astype(F)
    sg = np.where(n[:, 2] >= 0, F(1), F(-1)).astype(F)
    a = (F(-1) / (sg + n[:, 2])).astype(F)
    b = (n[:, 0] * n[:, 1] * a).astype(F)
    t0 = np.stack([F(1) + sg * n[:, 0] * n[:, 0] * a, sg * b, -sg * n[:, 0]], 1).astype(F)
    t1 = np.stack([b, sg + n[:, 1] * n[:, 1] * a, -n[:, 1]], 1).astype(F)
    v = (lx[:, None] * t0 + ly[:, None] * t1 + lz[:, None] * n).astype(F)
    inv = (F(1) / np.sqrt(_dot(v, v))).astype(F)
    return (v * inv[:, None]).astype(F)


def shade_step(meshes, materials, lights, rays, hits, shadow_hits, paths, max_depth, rr_depth, rr_survive, seed,
               pixel_base=0):
    """Returns (paths, next_rays, shadow_rays) of one step; meshes are the
    scene's meshes in DataSet order (records p0, e1 = p1-p0, e2 = p2-p0)."""
    n = rays.shape[0]
    paths = paths.copy()
    pend = np.any(paths["dl"] != 0, axis=1)
    # pad: 1 once a step has seen the path ended (the device then skips it
    # in later steps, its masked ray slots being written already)
    paths["pad"][paths["alive"] == 0] = 1
    if shadow_hits is not None:
        add = pend & (shadow_hits["meshIndex"] == NULL)
        paths["rad"][add] = (paths["rad"][add] + paths["dl"][add]).astype(F)
    paths["dl"] = 0
    nxt, shd = _masked(n), _masked(n)
    mi = hits["meshIndex"]
    ti = hits["triangleIndex"]
    valid = (paths["alive"] != 0) & ((rays["flags"] & 1) == 0) & (mi < len(meshes))
    ntri = np.array([m.ntris for m in meshes] + [0], np.int64)
    valid &= ti < ntri[np.minimum(mi, len(meshes))]
    paths["alive"][~valid] = 0
    idx = np.nonzero(valid)[0]
    if idx.size == 0:
        return paths, nxt, shd
    r, h, pth = rays[idx], hits[idx], paths[idx]
    kd = np.zeros((idx.size, 3), F)
    le = np.zeros((idx.size, 3), F)
    ok = h["meshIndex"] < len(materials)
    kd[ok] = materials["kd"][h["meshIndex"][ok]]
    le[ok] = materials["le"][h["meshIndex"][ok]]
    p0 = np.zeros((idx.size, 3), F)
    e1 = np.zeros((idx.size, 3), F)
    e2 = np.zeros((idx.size, 3), F)
    for m_i, m in enumerate(meshes):
        sel = h["meshIndex"] == m_i
        if not sel.any():
            continue
        v = m.verts[m.tris[h["triangleIndex"][sel]]]
        p0[sel] = v[:, 0]
        e1[sel] = (v[:, 1] - v[:, 0]).astype(F)
        e2[sel] = (v[:, 2] - v[:, 0]).astype(F)
    nrm = _cross(e1, e2)
    nl = np.sqrt(_dot(nrm, nrm)).astype(F)
    s = (np.where(_dot(nrm, r["d"]) > 0, F(-1), F(1)) / np.maximum(nl, F(1e-30))).astype(F)
    nrm = (nrm * s[:, None]).astype(F)
    p = (r["o"] + r["d"] * h["t"][:, None]).astype(F)
    eps = np.maximum(machine_eps(p[:, 0]), np.maximum(machine_eps(p[:, 1]), machine_eps(p[:, 2]))).astype(F)
    pix = (np.uint32(pixel_base) + idx.astype(np.uint32)).astype(np.uint32)
    depth = pth["depth"]
    d0 = depth == 0
    pth["rad"][d0] = (pth["rad"][d0] + pth["thr"][d0] * le[d0]).astype(F)
    diffuse = np.any(kd > 0, axis=1)
    shd_sub = _masked(idx.size)
    if len(lights):
        u0, u1, u2 = (rnd(seed, pix, depth, k) for k in (3, 4, 5))
        j = np.minimum((u0 * F(len(lights))).astype(np.uint32), np.uint32(len(lights) - 1))
        L = lights[j]
        su = np.sqrt(u1).astype(F)
        a = (su * (F(1) - u2)).astype(F)
        b = (su * u2).astype(F)
        q = ((L["p0"] + a[:, None] * L["e1"]) + b[:, None] * L["e2"]).astype(F)
        w = (q - p).astype(F)
        d2 = _dot(w, w)
        dist = np.sqrt(d2).astype(F)
        wn = (w / dist[:, None]).astype(F)
        ln = _cross(L["e1"], L["e2"])
        lnl = np.sqrt(_dot(ln, ln)).astype(F)
        cos_s = _dot(nrm, wn)
        cos_l = (np.abs(_dot(ln, wn)) / lnl).astype(F)
        use = diffuse & (cos_s > 0) & (cos_l > 0) & (dist > F(4) * eps)
        g = ((((cos_s * cos_l).astype(F) * ((F(0.5) * lnl).astype(F) * F(len(lights))).astype(F)) / d2).astype(F)
             * INV_PI).astype(F)
        dl = (((pth["thr"] * kd).astype(F) * L["le"]).astype(F) * g[:, None]).astype(F)
        pth["dl"][use] = dl[use]
        shd_sub["o"][use] = p[use]
        shd_sub["d"][use] = wn[use]
        shd_sub["mint"][use] = eps[use]
        shd_sub["maxt"][use] = (dist - F(2) * eps)[use]
        shd_sub["time"][use] = r["time"][use]
        shd_sub["flags"][use] = 0
    depth1 = (depth + 1).astype(np.uint32)
    alive = (depth1 < max_depth) & diffuse
    rr = alive & (depth1 >= rr_depth)
    surv = rnd(seed, pix, depth, 2) < F(rr_survive)
    keep = rr & surv
    pth["thr"][keep] = (pth["thr"][keep] / F(rr_survive)).astype(F)
    alive &= ~(rr & ~surv)
    nxt_sub = _masked(idx.size)
    if alive.any():
        dirs = _cosine_dir(nrm, rnd(seed, pix, depth, 0), rnd(seed, pix, depth, 1))
        pth["thr"][alive] = (pth["thr"][alive] * kd[alive]).astype(F)
        nxt_sub["o"][alive] = p[alive]
        nxt_sub["d"][alive] = dirs[alive]
        nxt_sub["mint"][alive] = eps[alive]
        nxt_sub["maxt"][alive] = np.inf
        nxt_sub["time"][alive] = r["time"][alive]
        nxt_sub["flags"][alive] = 0
    pth["depth"] = depth1
    pth["alive"] = alive.astype(np.uint32)
    paths[idx] = pth
    nxt[idx] = nxt_sub
    shd[idx] = shd_sub
    return paths, nxt, shd
