"""Pins for the parts of the oracle the reference leaves undefined -- the
direct-lighting and path Li, the area light, the point light with a
normalised wi (deviations D1, D2, D5 of the survey's register) -- against
what the mathematics fixes, independently of the oracle's own formulas:

* a Lambertian wall lit by a point light: every sample's radiance is
  rho/pi * I * |cos| / d^2 at the hit point, computed here in float64 from
  the camera ray;
* a closed, uniformly emitting Lambertian sphere seen from inside (a furnace):
  the irradiance at any wall point is pi * Le, so direct lighting has mean
  Le * (1 + rho), and a path tracer with next-event estimation at vertices
  0..D has mean Le * sum_{k=0}^{D+1} rho^k, with or without Russian roulette.
A dropped cosine, a wrong pdf (area vs solid angle), an unnormalised wi, a
missing 1/pi or a biased roulette weight fails one of these.
"""
import numpy as np

import oracle_lib as ol
from pbrt_rust_amd import scenes

F32 = np.float32


def camera_rays(cam, nx, ny, seed):
    """the oracle's own camera rays for every (pixel, sample)"""
    L = ol.lib()
    xres, yres = cam["xres"], cam["yres"]
    spp = nx * ny
    r2c = np.zeros(16, F32)
    L.orc_camera_raster_to_camera(xres, yres, ol.f32(cam["screen"]), cam["fov"], r2c)
    xy = np.zeros((yres * xres * spp, 2), F32)
    buf = np.zeros(2 * spp, F32)
    for y in range(yres):
        for x in range(xres):
            L.orc_pixel_samples(seed, xres, x, y, nx, ny, 1, buf)
            xy[(y * xres + x) * spp:(y * xres + x + 1) * spp] = buf.reshape(spp, 2)
    o = np.zeros(3 * xy.shape[0], F32)
    d = np.zeros_like(o)
    L.orc_camera_rays(r2c, ol.f32(cam["cam_to_world"]).reshape(16), xy.shape[0], xy.reshape(-1), o, d)
    return o.reshape(-1, 3).astype(np.float64), d.reshape(-1, 3).astype(np.float64)


def test_point_light_on_lambertian_wall_closed_form():
    wall_z, rho = 2.0, 0.6
    I = np.array([10.0, 5.0, 2.0])
    lp = np.array([0.3, 0.2, 0.5])
    v, idx = scenes.quad([-5, -5, wall_z], [5, -5, wall_z], [5, 5, wall_z], [-5, 5, wall_z])
    scene = dict(verts=v, idx=idx, face_mat=np.zeros(2, np.uint32), mats=np.array([[0, rho, rho, rho]], F32),
                 face_light=np.full(2, -1, np.int32), area_le=np.zeros((0, 3), F32),
                 point_lights=np.array([np.concatenate([lp, I])], F32))
    cam = scenes.pinhole(24, 20)
    integ = dict(kind="direct", nx=2, ny=2, seed=11, light_samples=1)
    res = ol.render(scene, cam, integ, nthreads=4, want_samples=True)
    got = res["samples"].reshape(-1, 3).astype(np.float64)
    o, d = camera_rays(cam, 2, 2, 11)
    t = (wall_z - o[:, 2]) / d[:, 2]
    p = o + d * t[:, None]
    w = lp - p
    d2 = np.sum(w * w, axis=1)
    cos = np.abs(w[:, 2]) / np.sqrt(d2)
    want = rho / np.pi * I[None, :] * (cos / d2)[:, None]
    assert np.all(got > 0)
    np.testing.assert_allclose(got, want, rtol=2e-5)


def inward_sphere(level):
    v, f = scenes.icosphere(level)
    v = v.astype(np.float64)
    f = f.astype(np.int64).copy()
    # geometric normal of the refined triple (p1,p2,p3) = (P[c],P[b],P[a]):
    # normalize((p2-p1) x (p3-p2)); make it point to the centre
    n = np.cross(v[f[:, 1]] - v[f[:, 2]], v[f[:, 0]] - v[f[:, 1]])
    c = v[f].mean(1)
    out = np.sum(n * c, axis=1) > 0
    f[out] = f[out][:, [1, 0, 2]]
    n = np.cross(v[f[:, 1]] - v[f[:, 2]], v[f[:, 0]] - v[f[:, 1]])
    assert np.all(np.sum(n * c, axis=1) < 0)
    return v.astype(F32), f.astype(np.uint32)


def furnace_scene(rho, le, level=3):
    v, f = inward_sphere(level)
    nf = f.shape[0]
    return dict(verts=v, idx=f, face_mat=np.zeros(nf, np.uint32), mats=np.array([[0, rho, rho, rho]], F32),
                face_light=np.zeros(nf, np.int32), area_le=np.array([[le, le, le]], F32),
                point_lights=np.zeros((0, 6), F32))


def furnace_mean(integ, rho=0.5, le=1.0):
    scene = furnace_scene(rho, le)
    cam = scenes.pinhole(32, 32, fov=60.0, cam_pos=(0.0, 0.0, 0.0))
    res = ol.render(scene, cam, integ, nthreads=8, want_samples=True)
    s = res["samples"].reshape(-1, 3).astype(np.float64)
    assert np.all(np.isfinite(s))
    return s.mean(), s.std() / np.sqrt(s.shape[0])


def test_furnace_direct_lighting():
    rho, le = 0.5, 1.0
    m, se = furnace_mean(dict(kind="direct", nx=4, ny=4, seed=3, light_samples=4), rho, le)
    want = le * (1 + rho)
    assert abs(m - want) < max(4 * se, 0.01 * want), (m, want, se)


def test_furnace_path_tracing_without_and_with_roulette():
    rho, le, D = 0.5, 1.0, 3
    want = le * sum(rho ** k for k in range(D + 2))
    m, se = furnace_mean(dict(kind="path", nx=4, ny=4, seed=5, max_depth=D, rr_start=10), rho, le)
    assert abs(m - want) < max(4 * se, 0.01 * want), (m, want, se)
    # roulette from bounce 2 on: unbiased, noisier
    m, se = furnace_mean(dict(kind="path", nx=4, ny=4, seed=6, max_depth=D, rr_start=1), rho, le)
    assert abs(m - want) < max(4 * se, 0.015 * want), (m, want, se)
    # depth 0: emission plus one light sample, the direct-lighting mean
    m, se = furnace_mean(dict(kind="path", nx=4, ny=4, seed=7, max_depth=0, rr_start=10), rho, le)
    assert abs(m - le * (1 + rho)) < max(4 * se, 0.01 * le * (1 + rho)), m
