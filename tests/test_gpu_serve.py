"""The request server (pbrtx_server_*): continuous batching over several
device-resident scenes returns, for every request, exactly what the same
request gives alone (and what the oracle gives), in both modes, under a tight
HBM admission budget and with concurrent submitters."""
import threading

import numpy as np
import pytest

import pbrt_rust_amd as px
from pbrt_rust_amd import scenes
from gpu_util import F32, mesh_pair

pytestmark = pytest.mark.gpu


def scene_rays(bvh, n, seed):
    """origins uniform in the scene's box, directions uniform on the sphere"""
    rng = np.random.default_rng(seed)
    lo, hi = bvh.world_bound()
    o = rng.uniform(lo, hi, (n, 3)).astype(F32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(F32)
    return o, d


@pytest.fixture(scope="module")
def world():
    out = []
    for k, ntri in enumerate((2000, 30000, 120000)):
        v, idx = scenes.random_triangles(ntri, 40 + k)
        out.append(mesh_pair(v, idx))
    return out


def make_requests(world, sizes, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for j, n in enumerate(sizes):
        s = int(rng.integers(0, len(world)))
        kind = "closest" if rng.random() < 0.7 else "any"
        o, d = scene_rays(world[s][0], n, seed * 1000 + j)
        maxt = F32(np.finfo(F32).max) if kind == "closest" else F32(0.3)
        reqs.append((s, kind, px.RayBatch(o, d, 0.0, maxt)))
    return reqs


def alone(world, s, kind, rays):
    bvh = world[s][0]
    r = px.RayBatch.from_records(rays.a.copy())
    return bvh.intersect(r, "batched") if kind == "closest" else bvh.intersect_p(r, "batched")


def same(kind, a, b):
    if kind == "closest":
        return np.array_equal(a.view(np.uint32), b.view(np.uint32))
    return np.array_equal(a, b)


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_results_same_alone_and_in_mix(world, mode):
    sizes = [1, 63, 64, 65, 1000, 70001, 250000, 3, 131072, 9999]
    reqs = make_requests(world, sizes, 5)
    srv = px.RayServer(arena_bytes=64 << 20, batch_rays=1 << 18, mode=mode)
    ids = [srv.add_scene(w[0]) for w in world]
    assert ids == [0, 1, 2]
    tickets = [srv.submit(s, kind, rays) for s, kind, rays in reqs]
    for (s, kind, rays), t in zip(reqs, tickets):
        out, st = srv.wait(t)
        ref = alone(world, s, kind, rays)
        assert same(kind, out, ref), (s, kind, len(rays))
        assert st["latency_ms"] >= 0 and st["launches"] >= 1
    c = srv.counters()
    assert c["requests_done"] == len(reqs) and c["rays_done"] == sum(sizes)
    srv.close()


def test_batched_results_match_oracle(world):
    reqs = make_requests(world, [4096, 20000, 777], 9)
    srv = px.RayServer(arena_bytes=32 << 20, batch_rays=8192)
    for w in world:
        srv.add_scene(w[0])
    tickets = [srv.submit(s, kind, rays) for s, kind, rays in reqs]
    for (s, kind, rays), t in zip(reqs, tickets):
        out, st = srv.wait(t)
        orc = world[s][1]
        o = np.stack([rays.a["ox"], rays.a["oy"], rays.a["oz"]], 1)
        d = np.stack([rays.a["dx"], rays.a["dy"], rays.a["dz"]], 1)
        if kind == "closest":
            ref = orc.intersect(o, d, 0.0, rays.a["tmax"][0], nthreads=8)
            assert np.array_equal(out["prim"], ref["prim"])
            h = ref["prim"] != px.NO_HIT
            assert np.array_equal(out["t"][h].view(np.uint32), ref["t"][h].view(np.uint32))
        else:
            ref, _ = orc.intersect_p(o, d, 0.0, rays.a["tmax"][0], nthreads=8)
            assert np.array_equal(out, ref.astype(bool))
        # big requests are split over several steps of 8192 rays
        if len(rays) > 8192:
            assert st["launches"] > 1
    srv.close()


def test_admission_budget_and_concurrent_submitters(world):
    # an arena that holds about two requests at a time: requests wait for
    # space, then complete; a request larger than the arena is refused
    n = 100000
    srv = px.RayServer(arena_bytes=8 << 20, batch_rays=1 << 16)
    for w in world:
        srv.add_scene(w[0])
    o, d = scene_rays(world[1][0], n, 3)
    big = px.RayBatch(np.zeros((400000, 3), F32), np.ones((400000, 3), F32))
    with pytest.raises(px.PbrtxError):
        srv.submit(1, "closest", big)
    rays = px.RayBatch(o, d)
    ref = alone(world, 1, "closest", rays)
    errors = []

    def client(k):
        try:
            for _ in range(3):
                out, st = srv.wait(srv.submit(1, "closest", rays))
                if not same("closest", out, ref):
                    errors.append(k)
        except Exception as e:  # surfaced below
            errors.append(repr(e))

    th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors
    assert srv.counters()["requests_done"] == 12
    srv.close()


def test_pinned_buffers(world):
    n = 50000
    o, d = scene_rays(world[2][0], n, 8)
    rays = px.RayBatch(o, d)
    pin = px.PinnedArray((n,), px.RAY_DTYPE)
    pin.array[:] = rays.a
    out = px.PinnedArray(((n + 31) // 32,), np.uint32)
    srv = px.RayServer(arena_bytes=16 << 20)
    for w in world:
        srv.add_scene(w[0])
    res, _ = srv.wait(srv.submit(2, "any", pin.array, out.array))
    assert np.array_equal(res, alone(world, 2, "any", rays))
    srv.close()


def test_scene_added_while_serving(world):
    # pbrtx_server_add_scene while the worker is tracing: the worker reads the
    # scene table under the server's lock, so requests in flight and requests
    # on the new scene all give their results alone
    srv = px.RayServer(arena_bytes=64 << 20, batch_rays=1 << 16)
    srv.add_scene(world[0][0])
    reqs = make_requests(world[:1], [200000, 150000, 90000], 21)
    tickets = [srv.submit(s, kind, rays) for s, kind, rays in reqs]
    late = []
    for k in (1, 2):
        assert srv.add_scene(world[k][0]) == k
        o, d = scene_rays(world[k][0], 50000, 70 + k)
        rays = px.RayBatch(o, d)
        late.append((k, "closest", rays, srv.submit(k, "closest", rays)))
    for (s, kind, rays), t in zip(reqs, tickets):
        out, _ = srv.wait(t)
        assert same(kind, out, alone(world, s, kind, rays))
    for s, kind, rays, t in late:
        out, _ = srv.wait(t)
        assert same(kind, out, alone(world, s, kind, rays))
    srv.close()
