"""Seeded synthetic scenes and ray sets for the intersection workloads.

Every generator takes a ``seed`` and draws from numpy's PCG64, so a scene or
ray set is reproducible bit for bit.  Positions are fp32.  Directions that are
"uniform on the sphere" use z = 1 - 2u, phi = 2 pi v.

Workloads (triangle counts are exact):

* ``plain_scene``: a [0,10]^3 box (12 tris) + a level-6 icosphere (81,920
  tris, r=2, centred) + 20,000 random triangles in the box = 101,932 tris,
  three meshes; rays: origins uniform in the box.
* ``heightfield_scene``: a 3,240 x 3,240-cell fBm heightfield over
  [-50,50]^2 (2 tris/cell, 20,995,200 tris) + 64 level-6 icospheres =
  26,238,080 tris; rays: 1920x1080 pinhole primaries, one cosine bounce per
  hit (the incoherent set), shadow rays to a 1x1 area light.
* ``instanced_scene``: a 50,000-tri displaced geodesic sphere instanced
  1,000 times (T*R*S, scale 0.5-2, translation in a 200^3 box), 20% of the
  instances with 2-key motion.
"""
from dataclasses import dataclass, field

import numpy as np

from .raytypes import RAY_DTYPE, make_rays, machine_epsilon_point


@dataclass
class Mesh:
    verts: np.ndarray   # (V,3) float32
    tris: np.ndarray    # (T,3) uint32, indices local to this mesh

    @property
    def ntris(self):
        return int(self.tris.shape[0])


def pack_meshes(meshes):
    """Concatenate meshes into (vertOff, triOff, verts, tris) arrays; the
    triangle indices stay local to their mesh (mesh index = list order)."""
    vo = np.zeros(len(meshes) + 1, dtype=np.uint32)
    to = np.zeros(len(meshes) + 1, dtype=np.uint32)
    for i, m in enumerate(meshes):
        vo[i + 1] = vo[i] + m.verts.shape[0]
        to[i + 1] = to[i] + m.tris.shape[0]
    if meshes:
        v = np.ascontiguousarray(np.concatenate([m.verts for m in meshes]).astype(np.float32))
        t = np.ascontiguousarray(np.concatenate([m.tris for m in meshes]).astype(np.uint32))
    else:
        v = np.zeros((0, 3), np.float32)
        t = np.zeros((0, 3), np.uint32)
    return vo, to, v, t


# ---------------------------------------------------------------------------
# primitives
# ---------------------------------------------------------------------------
def box_mesh(lo=(0.0, 0.0, 0.0), hi=(10.0, 10.0, 10.0)):
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    v = np.array([[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]], [lo[0], hi[1], lo[2]],
                  [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]], [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]],
                 dtype=np.float32)
    t = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [3, 7, 6], [3, 6, 2], [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], dtype=np.uint32)
    return Mesh(v, t)


_ICO_V = None


def _icosahedron():
    global _ICO_V
    p = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0], [0, -1, p], [0, 1, p],
                  [0, -1, -p], [0, 1, -p], [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


def geodesic_sphere(freq, radius=1.0, center=(0.0, 0.0, 0.0), displace=None):
    """Icosahedron with every face split freq x freq, projected onto the
    sphere: 20*freq^2 triangles, 10*freq^2+2 shared vertices.  freq=64 has
    the triangle and vertex counts of a level-6 icosphere (81,920 / 40,962);
    freq=50 gives exactly 50,000 triangles.  ``displace(dirs)`` optionally
    returns a radial scale per unit direction."""
    V, F = _icosahedron()
    n = int(freq)
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    keep = ii + jj <= n
    ii = ii[keep].astype(np.int64)
    jj = jj[keep].astype(np.int64)
    # local grid index -> flat index within a face
    lut = -np.ones((n + 1, n + 1), dtype=np.int64)
    lut[ii, jj] = np.arange(ii.size)
    # small triangles of one face in (i,j) grid coordinates
    a_i, a_j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    up = (a_i + a_j) <= n - 1
    ui, uj = a_i[up], a_j[up]
    tri_up = np.stack([lut[ui, uj], lut[ui + 1, uj], lut[ui, uj + 1]], axis=1)
    dn = (a_i + a_j) <= n - 2
    di, dj = a_i[dn], a_j[dn]
    tri_dn = np.stack([lut[di + 1, dj], lut[di + 1, dj + 1], lut[di, dj + 1]], axis=1)
    local_tris = np.concatenate([tri_up, tri_dn])
    npf = ii.size
    A = V[F[:, 0]][:, None, :]
    B = V[F[:, 1]][:, None, :]
    C = V[F[:, 2]][:, None, :]
    P = A + (B - A) * (ii[None, :, None] / n) + (C - A) * (jj[None, :, None] / n)
    P = P.reshape(-1, 3)
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    tris = (local_tris[None, :, :] + (np.arange(20) * npf)[:, None, None]).reshape(-1, 3)
    # merge the vertices shared along face edges
    key = np.round(P * (1 << 20)).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    order = np.argsort(first)
    remap = np.empty_like(order)
    remap[order] = np.arange(order.size)
    verts = P[first[order]]
    tris = remap[inv[tris]]
    r = radius * (displace(verts) if displace is not None else 1.0)
    if np.ndim(r) == 1:
        r = r[:, None]
    verts = verts * r + np.asarray(center, dtype=np.float64)
    return Mesh(verts.astype(np.float32), tris.astype(np.uint32))


def icosphere_l6(radius=2.0, center=(0.0, 0.0, 0.0)):
    return geodesic_sphere(64, radius, center)


def random_triangles(n, lo, hi, rng):
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    v = (lo + (hi - lo) * rng.random((3 * n, 3), dtype=np.float32)).astype(np.float32)
    t = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return Mesh(v, t)


def uniform_sphere_dirs(n, rng):
    u = rng.random(n, dtype=np.float32)
    v = rng.random(n, dtype=np.float32)
    z = (1.0 - 2.0 * u).astype(np.float32)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z)).astype(np.float32)
    phi = (2.0 * np.pi * v).astype(np.float32)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1).astype(np.float32)


def random_rays(n, lo, hi, rng, mint=1e-4, maxt=np.inf, time=None):
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    o = (lo + (hi - lo) * rng.random((n, 3), dtype=np.float32)).astype(np.float32)
    d = uniform_sphere_dirs(n, rng)
    r = make_rays(o, d, mint=mint, maxt=maxt)
    if time is not None:
        r["time"] = time
    return r


# ---------------------------------------------------------------------------
# plain scene: box + icosphere + random soup (101,932 tris), 1 Mi rays
# ---------------------------------------------------------------------------
def plain_scene(seed=1, n_random=20000):
    rng = np.random.Generator(np.random.PCG64(seed))
    return [box_mesh(), icosphere_l6(2.0, (5.0, 5.0, 5.0)), random_triangles(n_random, (0, 0, 0), (10, 10, 10), rng)]


def plain_rays(n=1 << 20, seed=2):
    rng = np.random.Generator(np.random.PCG64(seed))
    return random_rays(n, (0, 0, 0), (10, 10, 10), rng)


# ---------------------------------------------------------------------------
# heightfield scene (26,238,080 tris)
# ---------------------------------------------------------------------------
def _value_noise(x, y, freq, rng):
    """Bilinear value noise with smoothstep weights on a (freq+2)^2 lattice."""
    lat = rng.random((freq + 2, freq + 2), dtype=np.float32) * 2.0 - 1.0
    fx = x * freq
    fy = y * freq
    ix = np.floor(fx).astype(np.int64)
    iy = np.floor(fy).astype(np.int64)
    tx = (fx - ix).astype(np.float32)
    ty = (fy - iy).astype(np.float32)
    tx = tx * tx * (3 - 2 * tx)
    ty = ty * ty * (3 - 2 * ty)
    ix = np.clip(ix, 0, freq)
    iy = np.clip(iy, 0, freq)
    a = lat[ix, iy]
    b = lat[ix + 1, iy]
    c = lat[ix, iy + 1]
    d = lat[ix + 1, iy + 1]
    return (a + (b - a) * tx) + ((c + (d - c) * tx) - (a + (b - a) * tx)) * ty


def heightfield_mesh(cells=3240, extent=50.0, amplitude=5.0, seed=3, octaves=6):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = cells + 1
    g = np.linspace(0.0, 1.0, n, dtype=np.float32)
    X, Y = np.meshgrid(g, g, indexing="ij")
    z = np.zeros_like(X)
    amp, freq, tot = 1.0, 4, 0.0
    for _ in range(octaves):
        z += np.float32(amp) * _value_noise(X, Y, freq, rng)
        tot += amp
        amp *= 0.5
        freq *= 2
    z *= np.float32(amplitude / tot)
    verts = np.empty((n * n, 3), dtype=np.float32)
    verts[:, 0] = (X * (2 * extent) - extent).reshape(-1)
    verts[:, 1] = (Y * (2 * extent) - extent).reshape(-1)
    verts[:, 2] = z.reshape(-1)
    del X, Y, z
    i, j = np.meshgrid(np.arange(cells, dtype=np.uint32), np.arange(cells, dtype=np.uint32), indexing="ij")
    v00 = (i * n + j).reshape(-1)
    v10 = v00 + np.uint32(n)
    v01 = v00 + np.uint32(1)
    v11 = v10 + np.uint32(1)
    tris = np.empty((2 * cells * cells, 3), dtype=np.uint32)
    tris[0::2, 0] = v00
    tris[0::2, 1] = v10
    tris[0::2, 2] = v11
    tris[1::2, 0] = v00
    tris[1::2, 1] = v11
    tris[1::2, 2] = v01
    return Mesh(verts, tris)


def heightfield_scene(cells=3240, n_spheres=64, seed=3):
    """One heightfield mesh + n_spheres icosphere meshes (mesh index order)."""
    meshes = [heightfield_mesh(cells=cells, seed=seed)]
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    base = icosphere_l6(1.0)
    extent = 50.0 * cells / 3240.0 if cells < 3240 else 50.0
    for _ in range(n_spheres):
        r = np.float32(rng.uniform(1.0, 3.0))
        c = np.array([rng.uniform(-0.9, 0.9) * extent, rng.uniform(-0.9, 0.9) * extent,
                      rng.uniform(0.0, 10.0)], dtype=np.float32)
        meshes.append(Mesh((base.verts * r + c).astype(np.float32), base.tris.copy()))
    return meshes


def pinhole_rays(width=1920, height=1080, eye=(0.0, -30.0, 45.0), look=(0.0, 8.0, 0.0),
                 up=(0.0, 0.0, 1.0), vfov_deg=36.0):
    """Pixel-centre primaries of a pinhole camera (coherent set)."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(look, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, np.float64))
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd)
    th = np.tan(np.radians(vfov_deg) / 2)
    tw = th * width / height
    px = (np.arange(width) + 0.5) / width * 2 - 1
    py = 1 - (np.arange(height) + 0.5) / height * 2
    PX, PY = np.meshgrid(px, py)
    d = fwd[None, :] + (PX.reshape(-1, 1) * tw) * right[None, :] + (PY.reshape(-1, 1) * th) * upv[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.broadcast_to(eye, d.shape)
    return make_rays(o, d, mint=1e-4)


def _onb(n):
    """Orthonormal basis around unit normals n (n,3) (branchless, Duff et al. 2017)."""
    sign = np.where(n[:, 2] >= 0, 1.0, -1.0).astype(np.float32)
    a = -1.0 / (sign + n[:, 2])
    b = n[:, 0] * n[:, 1] * a
    t = np.stack([1.0 + sign * n[:, 0] * n[:, 0] * a, sign * b, -sign * n[:, 0]], axis=1)
    s = np.stack([b, sign + n[:, 1] * n[:, 1] * a, -n[:, 1]], axis=1)
    return t.astype(np.float32), s.astype(np.float32)


def hit_points_normals(meshes, rays, hits):
    """World hit points and geometric normals (facing the incoming ray) of
    the hit records with meshIndex != NULL."""
    hit = hits["meshIndex"] != 0xFFFFFFFF
    idx = np.nonzero(hit)[0]
    mi = hits["meshIndex"][idx].astype(np.int64)
    ti = hits["triangleIndex"][idx].astype(np.int64)
    p0 = np.empty((idx.size, 3), np.float32)
    p1 = np.empty_like(p0)
    p2 = np.empty_like(p0)
    for m in np.unique(mi):
        sel = mi == m
        tv = meshes[m].tris[ti[sel]]
        p0[sel] = meshes[m].verts[tv[:, 0]]
        p1[sel] = meshes[m].verts[tv[:, 1]]
        p2[sel] = meshes[m].verts[tv[:, 2]]
    n = np.cross(p1 - p0, p2 - p0).astype(np.float32)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    d = rays["d"][idx]
    flip = np.sum(n * d, axis=1) > 0
    n[flip] = -n[flip]
    p = rays["o"][idx] + rays["d"][idx] * hits["t"][idx][:, None]
    return idx, p.astype(np.float32), n.astype(np.float32)


def bounce_rays(meshes, rays, hits, seed=5):
    """One cosine-weighted diffuse bounce from every hit (incoherent set).
    mint = MachineEpsilon::E(origin), as Ray(origin, dir) sets it
    (include/luxrays/core/geometry/ray.h:41-44)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    idx, p, n = hit_points_normals(meshes, rays, hits)
    u1 = rng.random(idx.size, dtype=np.float32)
    u2 = rng.random(idx.size, dtype=np.float32)
    r = np.sqrt(u1)
    phi = 2.0 * np.pi * u2
    x = r * np.cos(phi)
    y = r * np.sin(phi)
    z = np.sqrt(np.maximum(0.0, 1.0 - u1))
    t, s = _onb(n)
    d = (x[:, None] * t + y[:, None] * s + z[:, None] * n).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    out = make_rays(p, d, mint=machine_epsilon_point(p), maxt=np.inf)
    return out


def shadow_rays(meshes, rays, hits, light_center=(0.0, 0.0, 30.0), seed=6):
    """Any-hit rays from each hit toward a uniform point on a 1x1 light
    parallel to the xy plane; maxt stops just short of the light."""
    rng = np.random.Generator(np.random.PCG64(seed))
    idx, p, n = hit_points_normals(meshes, rays, hits)
    lc = np.asarray(light_center, np.float32)
    lp = lc + np.stack([rng.random(idx.size, dtype=np.float32) - 0.5,
                        rng.random(idx.size, dtype=np.float32) - 0.5,
                        np.zeros(idx.size, np.float32)], axis=1)
    v = (lp - p).astype(np.float32)
    dist = np.linalg.norm(v, axis=1).astype(np.float32)
    d = (v / dist[:, None]).astype(np.float32)
    eps = machine_epsilon_point(p)
    return make_rays(p, d, mint=eps, maxt=(dist - eps * 2).astype(np.float32))


# ---------------------------------------------------------------------------
# instanced / motion scene
# ---------------------------------------------------------------------------
def displaced_sphere(freq=50, seed=7):
    rng = np.random.Generator(np.random.PCG64(seed))
    k = rng.uniform(2.0, 7.0, size=(4, 3))
    ph = rng.uniform(0, 2 * np.pi, size=(4, 3))

    def disp(dirs):
        s = np.zeros(dirs.shape[0])
        for a in range(4):
            s += 0.05 * np.sin(k[a, 0] * dirs[:, 0] + ph[a, 0]) * np.sin(k[a, 1] * dirs[:, 1] + ph[a, 1]) \
                * np.sin(k[a, 2] * dirs[:, 2] + ph[a, 2])
        return 1.0 + s
    return geodesic_sphere(freq, 1.0, (0.0, 0.0, 0.0), displace=disp)


def _random_rotation(rng):
    """Uniform random rotation (Shoemake's quaternion method)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def random_trs(rng, extent=100.0, smin=0.5, smax=2.0):
    """Local->world T*R*S with a uniform rotation, uniform scale and a
    translation uniform in [-extent, extent]^3 (shear-free)."""
    R = _random_rotation(rng)
    s = rng.uniform(smin, smax)
    T = rng.uniform(-extent, extent, size=3)
    M = np.eye(4)
    M[:3, :3] = R * s
    M[:3, 3] = T
    return M.astype(np.float32)


@dataclass
class InstancedScene:
    """Base meshes plus objects: kind 0 plain, 1 instance (local->world
    matrix), 2 motion (keys: times + global->local matrices)."""
    meshes: list
    obj_mesh: np.ndarray
    obj_kind: np.ndarray
    obj_mat: np.ndarray           # (nObj,4,4) local->world (kind 1)
    key_off: np.ndarray           # (nObj+1,)
    key_times: np.ndarray         # (nKeys,)
    key_mats: np.ndarray          # (nKeys,4,4) global->local
    extra: dict = field(default_factory=dict)


def instanced_scene(n_inst=1000, motion_frac=0.2, freq=50, extent=100.0, seed=8):
    rng = np.random.Generator(np.random.PCG64(seed))
    base = displaced_sphere(freq=freq, seed=seed + 1)
    n_mot = int(round(n_inst * motion_frac))
    is_mot = np.zeros(n_inst, bool)
    is_mot[rng.choice(n_inst, n_mot, replace=False)] = True
    obj_mesh = np.zeros(n_inst, np.uint32)
    obj_kind = np.where(is_mot, 2, 1).astype(np.uint32)
    obj_mat = np.zeros((n_inst, 4, 4), np.float32)
    key_off = np.zeros(n_inst + 1, np.uint32)
    times, mats = [], []
    for o in range(n_inst):
        M0 = random_trs(rng, extent)
        if is_mot[o]:
            M1 = random_trs(rng, extent)
            # small motion: the end key is the start key moved a little
            M1[:3, 3] = M0[:3, 3] + rng.uniform(-5.0, 5.0, size=3).astype(np.float32)
            R1 = _random_rotation(rng) if rng.random() < 0.5 else np.eye(3)
            blend = M0[:3, :3] @ R1.astype(np.float32) if rng.random() < 0.5 else M0[:3, :3]
            M1[:3, :3] = blend
            for t, M in ((0.0, M0), (1.0, M1)):
                times.append(t)
                mats.append(np.linalg.inv(M.astype(np.float64)).astype(np.float32))
            key_off[o + 1] = key_off[o] + 2
        else:
            obj_mat[o] = M0
            key_off[o + 1] = key_off[o]
    key_times = np.asarray(times, np.float32)
    key_mats = np.asarray(mats, np.float32).reshape(-1, 4, 4) if mats else np.zeros((0, 4, 4), np.float32)
    return InstancedScene([base], obj_mesh, obj_kind, obj_mat, key_off, key_times, key_mats)


def instanced_rays(n=1 << 22, extent=100.0, seed=9):
    rng = np.random.Generator(np.random.PCG64(seed))
    r = random_rays(n, (-extent,) * 3, (extent,) * 3, rng)
    r["time"] = rng.random(n, dtype=np.float32)
    return r


# ---------------------------------------------------------------------------
# batched workload: 8 resident scenes, 64 ray buffers
# ---------------------------------------------------------------------------
BATCH_SCENE_TRIS = (10_000, 50_000, 200_000, 500_000, 1_000_000, 2_000_000, 3_000_000, 5_000_000)


def soup_scene(ntris, extent=100.0, seed=0, mean_edge=0.05):
    """Random soup: centres uniform in [0, extent]^3, triangle size lognormal
    (sigma 0.5) with mean edge length mean_edge * extent."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sigma = 0.5
    mu = np.log(mean_edge * extent) - sigma * sigma / 2
    L = rng.lognormal(mu, sigma, ntris).astype(np.float32)
    c = (rng.random((ntris, 3), dtype=np.float32) * extent).astype(np.float32)
    u = rng.normal(size=(ntris, 3, 3)).astype(np.float32)
    u /= np.linalg.norm(u, axis=2, keepdims=True)
    # vertices on a sphere of radius L/sqrt(3) around the centre: edges ~ L
    v = c[:, None, :] + u * (L / np.sqrt(3.0))[:, None, None]
    return Mesh(v.reshape(-1, 3).astype(np.float32), np.arange(3 * ntris, dtype=np.uint32).reshape(-1, 3))


def sphere_grid_scene(ntris, extent=100.0, seed=0, freq=10):
    """Displaced geodesic spheres (20*freq^2 tris each) on a jittered grid."""
    rng = np.random.Generator(np.random.PCG64(seed))
    per = 20 * freq * freq
    n = max(1, ntris // per)
    base = displaced_sphere(freq=freq, seed=seed + 1)
    k = int(np.ceil(n ** (1.0 / 3.0)))
    cell = extent / k
    idx = np.arange(n)
    grid = np.stack([idx % k, (idx // k) % k, idx // (k * k)], axis=1).astype(np.float32)
    centres = (grid + 0.5 + rng.uniform(-0.2, 0.2, (n, 3))) * cell
    radii = rng.uniform(0.2, 0.45, n) * cell
    v = base.verts[None, :, :] * radii[:, None, None] + centres[:, None, :]
    t = base.tris[None, :, :] + (np.arange(n, dtype=np.uint32) * base.verts.shape[0])[:, None, None]
    return Mesh(v.reshape(-1, 3).astype(np.float32), t.reshape(-1, 3).astype(np.uint32))


def batch_scenes(sizes=BATCH_SCENE_TRIS, extent=100.0, seed=100):
    """Scene i has sizes[i] tris: even i a random soup, odd i a sphere grid."""
    out = []
    for i, n in enumerate(sizes):
        m = soup_scene(n, extent, seed + i) if i % 2 == 0 else sphere_grid_scene(n, extent, seed + i)
        out.append([m])
    return out


def batch_buffers(n_buffers=64, lo=4096, hi=4 << 20, n_scenes=8, extent=100.0, seed=200):
    """Buffer specs (scene index, ray count, query): sizes log-uniform in
    [lo, hi], scenes and closest/any queries uniform."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_buffers)).astype(np.int64)
    scene = rng.integers(0, n_scenes, n_buffers)
    query = rng.integers(0, 2, n_buffers)
    return [(int(scene[i]), int(sizes[i]), int(query[i])) for i in range(n_buffers)]


def buffer_rays(count, extent=100.0, seed=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    return random_rays(count, (0, 0, 0), (extent, extent, extent), rng)
