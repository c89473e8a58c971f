"""Seeded synthetic workloads (the reference ships no scenes: its scene-file
front end is a stub, src/main.rs:800-849).  Every generator is deterministic
in its seed; sizes are those of the ray-query and render configurations.

Scene dict keys (consumed by the renderer and by the CPU oracle):
  verts (n,3) f32 world space, idx (f,3) u32, face_mat (f,) u32,
  mats (m,4) f32 [kind, r, g, b] (kind 0 = matte/Lambertian, 1 = mirror),
  face_light (f,) i32 area-light id or -1, area_le (a,3) f32,
  point_lights (p,6) f32 [pos, intensity]
"""
import math

import numpy as np

F32 = np.float32


# ---------------------------------------------------------------------------
# random triangles: centroids U[-1,1]^3, vertices c + l_k u_k with
# l_k ~ U(lmin, lmax) and u_k uniform on the sphere
# ---------------------------------------------------------------------------
def random_triangles(n, seed=0xC0, lmin=0.01, lmax=0.05, lo=-1.0, hi=1.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(lo, hi, (n, 1, 3))
    u = rng.normal(size=(n, 3, 3))
    u /= np.linalg.norm(u, axis=2, keepdims=True)
    ln = rng.uniform(lmin, lmax, (n, 3, 1))
    v = (c + ln * u).astype(F32).reshape(-1, 3)
    return v, np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)


def pinhole(xres, yres, fov=60.0, cam_pos=(0.0, 0.0, -3.0)):
    """Camera looking along +z from cam_pos (cam_to_world = Translate); the
    screen window spans the longer image axis as pbrt's default frame does."""
    frame = xres / yres
    if frame > 1:
        screen = [-frame, frame, -1.0, 1.0]
    else:
        screen = [-1.0, 1.0, -1.0 / frame, 1.0 / frame]
    c2w = np.eye(4, dtype=F32)
    c2w[:3, 3] = np.asarray(cam_pos, F32)
    return dict(cam_to_world=c2w, fov=float(fov), screen=[float(s) for s in screen],
                xres=int(xres), yres=int(yres))


def ray_query_scene(ntris=16384, seed=0xC0):
    """Plain ray queries: random triangles + point light at (2,2,-2)."""
    v, idx = random_triangles(ntris, seed)
    f = idx.shape[0]
    return dict(verts=v, idx=idx, face_mat=np.zeros(f, np.uint32),
                mats=np.array([[0, 0.5, 0.5, 0.5]], F32), face_light=np.full(f, -1, np.int32),
                area_le=np.zeros((0, 3), F32),
                point_lights=np.array([[2, 2, -2, 10, 10, 10]], F32))


# ---------------------------------------------------------------------------
# icosphere + fBm displacement
# ---------------------------------------------------------------------------
def icosphere(level):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                  [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(level):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.minimum(e[:, 0], e[:, 1]) * nv + np.maximum(e[:, 0], e[:, 1])
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // nv, uk % nv
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv.reshape(3, -1) + nv  # midpoints of edges 01, 12, 20
        v = np.concatenate([v, mid])
        f0, f1, f2 = f[:, 0], f[:, 1], f[:, 2]
        m01, m12, m20 = m[0], m[1], m[2]
        f = np.concatenate([np.stack([f0, m01, m20], 1), np.stack([f1, m12, m01], 1),
                            np.stack([f2, m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def _grad_noise(p, seed):
    """Gradient noise on the integer lattice (hash-chosen unit gradients,
    quintic fade), vectorised over (n,3) points."""
    pi = np.floor(p).astype(np.int64)
    pf = p - pi
    w = pf * pf * pf * (pf * (pf * 6 - 15) + 10)

    def hash3(ix, iy, iz):
        h = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
        h = (h ^ (h >> 13)) * 1274126177
        return (h ^ (h >> 16)) & 0xFFFFFFFF

    out = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                h = hash3(pi[:, 0] + dx, pi[:, 1] + dy, pi[:, 2] + dz)
                th = (h & 0xFFFF) / 65536.0 * 2 * math.pi
                z = ((h >> 16) & 0xFFFF) / 65536.0 * 2 - 1
                r = np.sqrt(np.maximum(0.0, 1 - z * z))
                g = np.stack([r * np.cos(th), r * np.sin(th), z], 1)
                d = pf - np.array([dx, dy, dz], np.float64)
                val = np.sum(g * d, axis=1)
                wx = w[:, 0] if dx else 1 - w[:, 0]
                wy = w[:, 1] if dy else 1 - w[:, 1]
                wz = w[:, 2] if dz else 1 - w[:, 2]
                out = out + wx * wy * wz * val
    return out


def fbm(p, octaves=4, seed=1, freq=2.0, lacunarity=2.0, gain=0.5):
    s = np.zeros(p.shape[0])
    amp, f, norm = 1.0, freq, 0.0
    for o in range(octaves):
        s += amp * _grad_noise(p * f, seed + o)
        norm += amp
        amp *= gain
        f *= lacunarity
    return s / norm


def displaced_icosphere(level, amplitude=0.05, seed=1, center=(0, 0, 0), scale=1.0):
    v, f = icosphere(level)
    n = fbm(v, 4, seed)
    n = n / max(1e-9, np.abs(n).max())
    v = v * (1.0 + amplitude * n)[:, None]
    v = v * scale + np.asarray(center, np.float64)
    return v.astype(F32), f.astype(np.uint32)


def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3)."""
    return np.array([p0, p1, p2, p3], F32), np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def _merge(parts):
    verts, idx, base = [], [], 0
    for v, f in parts:
        verts.append(v)
        idx.append(f + base)
        base += v.shape[0]
    return np.concatenate(verts).astype(F32), np.concatenate(idx).astype(np.uint32)


def direct_lighting_scene(level=8, seed=1, light_radiance=10.0):
    """Displaced icosphere (albedo 0.5) over a ground quad (albedo 0.8) lit by
    a downward-facing 2-triangle quad area light of radiance light_radiance."""
    sph = displaced_icosphere(level, 0.05, seed)
    g = 1.2
    ground = quad([-20, -g, -20], [-20, -g, 20], [20, -g, 20], [20, -g, -20])
    # light at y = 2.5: winding chosen so the triangle normal points down (-y)
    light = quad([-0.75, 2.5, -0.75], [-0.75, 2.5, 0.75], [0.75, 2.5, 0.75], [0.75, 2.5, -0.75])
    v, idx = _merge([sph, ground, light])
    nf = idx.shape[0]
    ns, ng = sph[1].shape[0], 2
    face_mat = np.zeros(nf, np.uint32)
    face_mat[ns:ns + ng] = 1
    face_mat[ns + ng:] = 2
    face_light = np.full(nf, -1, np.int32)
    face_light[ns + ng:] = 0
    mats = np.array([[0, 0.5, 0.5, 0.5], [0, 0.8, 0.8, 0.8], [0, 0.0, 0.0, 0.0]], F32)
    return dict(verts=v, idx=idx, face_mat=face_mat, mats=mats, face_light=face_light,
                area_le=np.array([[light_radiance] * 3], F32), point_lights=np.zeros((0, 6), F32))


def spheres_scene(n_spheres=64, level=6, seed=2, mirror_frac=0.3, light_radiance=10.0):
    """n_spheres icospheres at U[-10,10]^3 with scales U(0.3,1.5); a share
    mirror_frac of them mirrors, the rest Lambertian; one quad area light."""
    rng = np.random.default_rng(seed)
    base_v, base_f = icosphere(level)
    base_v = base_v.astype(F32)
    parts, mats_of = [], []
    for k in range(n_spheres):
        c = rng.uniform(-10, 10, 3)
        s = rng.uniform(0.3, 1.5)
        parts.append(((base_v * s + c).astype(F32), base_f.astype(np.uint32)))
        mats_of.append(1 if rng.uniform() < mirror_frac else 0)
    light = quad([-3, 14, -3], [-3, 14, 3], [3, 14, 3], [3, 14, -3])  # normal points down
    parts.append(light)
    v, idx = _merge(parts)
    nf = idx.shape[0]
    per = base_f.shape[0]
    mats = np.array([[0, 0.7, 0.7, 0.7], [1, 0.9, 0.9, 0.9], [0, 0.0, 0.0, 0.0]], F32)
    face_mat = np.zeros(nf, np.uint32)
    for k, m in enumerate(mats_of):
        face_mat[k * per:(k + 1) * per] = m
    face_mat[n_spheres * per:] = 2
    face_light = np.full(nf, -1, np.int32)
    face_light[n_spheres * per:] = 0
    return dict(verts=v, idx=idx, face_mat=face_mat, mats=mats, face_light=face_light,
                area_le=np.array([[light_radiance] * 3], F32), point_lights=np.zeros((0, 6), F32))


def analytic_spheres_scene(n_spheres=12, seed=4, mirror_frac=0.3, light_radiance=10.0, partial=True):
    """Sphere shapes (Shape::sphere) of radius U(0.3, 1) at U[-4,4] x
    U[-1,3] x U[-4,4] above a ground quad (albedo 0.8), Lambertian (albedo
    0.7) or mirror (0.9), under a downward-facing quad area light; with
    partial, some spheres are clipped in z and phi, some scaled non-uniformly
    or mirrored (reverse_orientation / a handedness-swapping transform)."""
    from .api import Sphere, Transform
    rng = np.random.default_rng(seed)
    g = 1.2
    ground = quad([-20, -g, -20], [-20, -g, 20], [20, -g, 20], [20, -g, -20])
    light = quad([-1.5, 6.0, -1.5], [-1.5, 6.0, 1.5], [1.5, 6.0, 1.5], [1.5, 6.0, -1.5])
    v, idx = _merge([ground, light])
    mats = np.array([[0, 0.8, 0.8, 0.8], [0, 0.0, 0.0, 0.0], [0, 0.7, 0.7, 0.7], [1, 0.9, 0.9, 0.9]], F32)
    spheres, smat = [], []
    for k in range(n_spheres):
        c = [rng.uniform(-4, 4), rng.uniform(-1, 3), rng.uniform(-4, 4)]
        r = float(rng.uniform(0.3, 1.0))
        xf = Transform.translate(c)
        z0, z1, phimax, ro = -r, r, 360.0, False
        if partial and k % 4 == 1:
            z0, z1, phimax = -0.6 * r, 0.8 * r, 270.0
        if partial and k % 4 == 2:
            xf = xf * Transform.scale(1.0, 1.4, 0.8)
        if partial and k % 4 == 3:
            xf = xf * Transform.scale(-1.0, 1.0, 1.0)  # swaps handedness
            ro = True
        spheres.append(Sphere(xf, xf.inverse(), ro, r, z0, z1, phimax))
        smat.append(3 if rng.uniform() < mirror_frac else 2)
    return dict(verts=v, idx=idx, face_mat=np.array([0, 0, 1, 1], np.uint32), mats=mats,
                face_light=np.array([-1, -1, 0, 0], np.int32), area_le=np.array([[light_radiance] * 3], F32),
                point_lights=np.zeros((0, 6), F32), spheres=spheres, sphere_mat=np.array(smat, np.uint32))


_DISPLACED_POOL = {}


def displaced_pool(count=16, level=6, seed=100):
    """`count` level-`level` icospheres displaced by fBm with seeds seed..,
    made once per process (the instances of the sphere fields below)."""
    key = (count, level, seed)
    if key not in _DISPLACED_POOL:
        _DISPLACED_POOL[key] = [displaced_icosphere(level, 0.05, seed + k) for k in range(count)]
    return _DISPLACED_POOL[key]


def sphere_field(n_spheres, seed=3, extent=10.0, level=6, light_radiance=10.0):
    """n_spheres displaced level-`level` icospheres (instances of a pool of 16
    fBm-displaced ones) at U[-extent, extent]^3 with scales U(0.3, 1.5) *
    extent / 10, Lambertian albedo 0.5 or 0.8, under one quad area light
    above the field (normal pointing down)."""
    rng = np.random.default_rng(seed)
    pool = displaced_pool(level=level)
    nv, nf = pool[0][0].shape[0], pool[0][1].shape[0]
    k = rng.integers(0, len(pool), n_spheres)
    c = rng.uniform(-extent, extent, (n_spheres, 3))
    sc = rng.uniform(0.3, 1.5, n_spheres) * (extent / 10.0)
    verts = np.empty((n_spheres * nv + 4, 3), F32)
    idx = np.empty((n_spheres * nf + 2, 3), np.uint32)
    for j in range(n_spheres):
        v, f = pool[k[j]]
        verts[j * nv:(j + 1) * nv] = (v * F32(sc[j]) + c[j].astype(F32)).astype(F32)
        idx[j * nf:(j + 1) * nf] = f + np.uint32(j * nv)
    h, w = 1.4 * extent, 0.3 * extent
    lv, lf = quad([-w, h, -w], [-w, h, w], [w, h, w], [w, h, -w])
    verts[n_spheres * nv:] = lv
    idx[n_spheres * nf:] = lf + np.uint32(n_spheres * nv)
    nfaces = idx.shape[0]
    face_mat = np.repeat(rng.integers(0, 2, n_spheres).astype(np.uint32), nf)
    face_mat = np.concatenate([face_mat, np.full(2, 2, np.uint32)])
    face_light = np.full(nfaces, -1, np.int32)
    face_light[n_spheres * nf:] = 0
    mats = np.array([[0, 0.5, 0.5, 0.5], [0, 0.8, 0.8, 0.8], [0, 0.0, 0.0, 0.0]], F32)
    return dict(verts=verts, idx=idx, face_mat=face_mat, mats=mats, face_light=face_light,
                area_le=np.array([[light_radiance] * 3], F32), point_lights=np.zeros((0, 6), F32))


def look_at(xres, yres, eye, target, fov=60.0, up=(0.0, 1.0, 0.0)):
    """Perspective camera at eye looking at target (pbrt's LookAt: camera +z
    toward the target, +y up), as a camera dict like pinhole's."""
    eye, target, up = (np.asarray(a, np.float64) for a in (eye, target, up))
    d = target - eye
    d /= np.linalg.norm(d)
    r = np.cross(up / np.linalg.norm(up), d)
    if np.linalg.norm(r) < 1e-9:
        r = np.cross(np.array([1.0, 0.0, 0.0]), d)
    r /= np.linalg.norm(r)
    u = np.cross(d, r)
    c2w = np.eye(4, dtype=F32)
    c2w[:3, 0], c2w[:3, 1], c2w[:3, 2], c2w[:3, 3] = r, u, d, eye
    cam = pinhole(xres, yres, fov)
    cam["cam_to_world"] = c2w
    return cam


def sub_image_camera(cam, xres, yres):
    """The same view at another resolution."""
    c = dict(cam)
    return pinhole(xres, yres, cam["fov"], tuple(cam["cam_to_world"][:3, 3]))
