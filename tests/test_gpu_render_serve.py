"""The render-request server (pbrtx_render_server_*): every image it returns
is, bit for bit, the image the same request gives alone (SamplerRenderer on
the batched path, itself pinned to the oracle) -- in both modes, with
requests of several scenes, sizes and sample counts in flight at once, under
a tight HBM admission budget and with concurrent submitters -- and a small
request is checked against the oracle directly."""
import threading

import numpy as np
import pytest

import oracle_lib as ol
import pbrt_rust_amd as px
from pbrt_rust_amd import scenes

pytestmark = pytest.mark.gpu
PATH = px.PathIntegrator(5, 3)


@pytest.fixture(scope="module")
def world():
    return [
        (scenes.spheres_scene(n_spheres=8, level=3, seed=5), (0.0, 0.0, -25.0)),
        (scenes.spheres_scene(n_spheres=40, level=3, seed=8), (0.0, 0.0, 0.0)),
        (scenes.direct_lighting_scene(level=3), (0.0, 0.0, -3.0)),
    ]


# (scene, xres, yres, nx, ny, seed)
REQUESTS = [(0, 32, 24, 1, 1, 3), (1, 48, 32, 2, 2, 4), (2, 64, 48, 2, 2, 5), (1, 20, 10, 3, 1, 6),
            (0, 96, 64, 2, 2, 7), (2, 1, 1, 1, 1, 8), (1, 128, 96, 1, 1, 9), (2, 40, 30, 4, 4, 10),
            (0, 17, 13, 1, 3, 11), (1, 64, 64, 2, 2, 12)]


def camera(world, req):
    s, xres, yres = req[:3]
    return scenes.pinhole(xres, yres, fov=60.0, cam_pos=world[s][1])


def alone(world, req):
    s, xres, yres, nx, ny, seed = req
    r = px.SamplerRenderer(world[s][0], camera(world, req), px.StratifiedSampler(nx, ny, seed), PATH)
    img, st = r.render("batched")
    r.close()
    return img, st


def serve(world, reqs, mode, arena_bytes=256 << 20, batch=16384, threads=1):
    srv = px.RenderServer(arena_bytes=arena_bytes, batch_samples=batch, mode=mode)
    assert [srv.add_scene(w[0], PATH) for w in world] == [0, 1, 2]
    out = [None] * len(reqs)
    tickets = [None] * len(reqs)

    def submit(idx):
        for i in idx:
            s, xres, yres, nx, ny, seed = reqs[i]
            tickets[i] = srv.submit(s, camera(world, reqs[i]), px.StratifiedSampler(nx, ny, seed))

    parts = [list(range(k, len(reqs), threads)) for k in range(threads)]
    th = [threading.Thread(target=submit, args=(p,)) for p in parts]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i, t in enumerate(tickets):
        out[i] = srv.wait(t)
    c = srv.counters()
    srv.close()
    return out, c


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_images_same_alone_and_in_mix(world, mode):
    out, c = serve(world, REQUESTS, mode)
    cam_rays = 0
    shadow = ext = 0
    for req, (img, st) in zip(REQUESTS, out):
        ref, rst = alone(world, req)
        assert img.shape == ref.shape
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), req
        assert st["n_samples"] == req[1] * req[2] * req[3] * req[4]
        assert st["latency_ms"] >= 0 and st["batches"] >= 1
        cam_rays += rst["camera_rays"]
        shadow += rst["shadow_rays"]
        ext += rst["extension_rays"]
    assert c["requests_done"] == len(REQUESTS)
    assert c["samples_done"] == cam_rays == c["camera_rays"]
    assert c["shadow_rays"] == shadow and c["extension_rays"] == ext


def test_tight_arena_and_concurrent_submitters(world):
    # an arena that holds about two of the larger requests: admission waits
    # for retirements; four threads submit at once
    reqs = REQUESTS * 2
    out, c = serve(world, reqs, "batched", arena_bytes=2 << 20, batch=8192, threads=4)
    for req, (img, st) in zip(reqs, out):
        ref, _ = alone(world, req)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), req
    assert c["requests_done"] == len(reqs)
    assert c["steps"] > 1 and c["batches"] >= c["steps"]


def test_served_image_matches_oracle(world):
    req = (1, 48, 32, 2, 2, 4)
    out, _ = serve(world, [req, REQUESTS[2], REQUESTS[6]], "batched", batch=4096)
    img = out[0][0]
    integ = dict(kind="path", nx=2, ny=2, seed=4, max_depth=5, rr_start=3)
    ref = ol.render(world[1][0], camera(world, req), integ, nthreads=16)
    assert np.array_equal(img.view(np.uint32), ref["rgb"].view(np.uint32))


def test_request_errors(world):
    srv = px.RenderServer(arena_bytes=1 << 20, batch_samples=4096)
    sid = srv.add_scene(world[0][0], PATH)
    cam = scenes.pinhole(8, 8)
    with pytest.raises(px.PbrtxError):
        srv.submit(sid + 5, cam, px.StratifiedSampler(1, 1, 0))  # unknown scene
    with pytest.raises(px.PbrtxError):
        srv.submit(sid, scenes.pinhole(512, 512), px.StratifiedSampler(4, 4, 0))  # larger than the arena
    with pytest.raises(px.PbrtxError):
        srv.submit(sid, dict(cam, xres=0), px.StratifiedSampler(1, 1, 0))  # empty image
    with pytest.raises(px.PbrtxError):
        srv.add_scene(world[2][0], px.DirectLightingIntegrator(1))  # the server runs the path integrator
    img, st = srv.wait(srv.submit(sid, cam, px.StratifiedSampler(1, 1, 0)))
    assert img.shape == (8, 8, 3)
    srv.close()
