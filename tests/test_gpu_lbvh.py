"""The tree built on the GPU (pbrtx_scene_build_gpu, a linear BVH) against
the oracle (the reference's SAH tree on the CPU): closest hits give the same
primitive, t, barycentrics and maxt write-back bit for bit except on rays
where the reference's own answer depends on the order in which it meets the
tree -- a closest t shared by several primitives (a tie), or a hit whose own
box the slab test rejects at maxt = its t (a hit on its box's boundary, culled
or not by the pop-time test depending on what was hit before) -- both counted
by the oracle's brute force; occlusion bits are the same for every ray.
Meshes with shared edges and rays aimed at vertices make such rays happen."""
import numpy as np
import pytest

import oracle_lib as ol
import pbrt_rust_amd as px
from pbrt_rust_amd import scenes
from gpu_util import F32, FMAX, random_rays

pytestmark = pytest.mark.gpu


def pair(v, idx, spheres=()):
    prims = [px.Primitive.simple(px.TriangleMesh(None, None, False, idx, v))] if len(idx) else []
    prims += [px.Primitive.simple(s) for s in spheres]
    gpu = px.BVHAccelerator(prims, 1, "lbvh")
    orc = ol.OracleScene()
    if len(idx):
        orc.add_mesh(v, idx)
    for s in spheres:
        orc.add_sphere(s.o2w.m, s.o2w.m_inv, s.radius, s.z0, s.z1, s.phimax, int(s.reverse_orientation))
    orc.build("sah", 1)
    return gpu, orc


def check(gpu, orc, o, d, mint=0.0, maxt=FMAX, occ_maxt=None):
    n = o.shape[0]
    for path in ("plain", "batched"):
        rays = px.RayBatch(o, d, mint, maxt)
        hits = gpu.intersect(rays, path)
        ref = orc.intersect(o, d, mint, maxt, nthreads=8)
        mi, ma = np.broadcast_to(F32(mint), (n,)), np.broadcast_to(F32(maxt), (n,))
        bp, bt, ties = orc.brute(o, d, mi, ma)
        sensitive = (ties > 0) | (orc.box_misses(o, d, mi, ma) > 0)
        same = hits["prim"] == ref["prim"]
        # a different primitive only where the reference's answer is order-sensitive
        assert np.all(sensitive[~same]), np.nonzero(~same & ~sensitive)[0][:5]
        h = same & (ref["prim"] != px.NO_HIT)
        for k in ("t", "b1", "b2"):
            assert np.array_equal(hits[k][h].view(np.uint32), ref[k][h].view(np.uint32)), k
        fixed = ~sensitive
        assert np.array_equal(rays.maxt[fixed].view(np.uint32), ref["maxt"][fixed].view(np.uint32))
        om = maxt if occ_maxt is None else occ_maxt
        occ = gpu.intersect_p(px.RayBatch(o, d, mint, om), path)
        ref_occ, _ = orc.intersect_p(o, d, mint, om, nthreads=8)
        assert np.array_equal(occ, ref_occ.astype(bool))
    return int((~same).sum()), int(sensitive.sum())


def test_random_triangles():
    v, idx = scenes.random_triangles(5000, 21)
    gpu, orc = pair(v, idx)
    o, d = random_rays(20000, 22)
    diff, tied = check(gpu, orc, o, d, occ_maxt=F32(0.6))
    assert gpu.build_ms is not None and gpu.build_ms > 0


def test_shared_edges_and_vertex_rays():
    # an icosphere: neighbouring faces share edges and vertices; rays aimed
    # at vertices and edge midpoints hit several faces at one t
    v, idx = scenes.icosphere(3)
    gpu, orc = pair(v, idx)
    rng = np.random.default_rng(3)
    targets = np.concatenate([v, 0.5 * (v[idx[:, 0]] + v[idx[:, 1]])]).astype(F32)
    o = (rng.normal(size=(targets.shape[0], 3)) * 3).astype(F32)
    d = (targets - o).astype(F32)
    o2, d2 = random_rays(5000, 23)
    o = np.concatenate([o, o2]).astype(F32)
    d = np.concatenate([d, d2]).astype(F32)
    diff, sens = check(gpu, orc, o, d, occ_maxt=F32(0.9))
    assert sens > 0  # the case is exercised
    print("icosphere: %d rays, %d order-sensitive (tie or boundary hit), %d of them answered differently"
          % (o.shape[0], sens, diff))


def test_spheres_and_small_scenes():
    xf = px.Transform.translate([0.2, 0.1, 0.3])
    sph = [px.Sphere(xf, xf.inverse(), False, 0.4, -0.4, 0.4, 360.0),
           px.Sphere(px.Transform(), px.Transform(), False, 0.25, -0.1, 0.2, 300.0)]
    v, idx = scenes.random_triangles(300, 24)
    gpu, orc = pair(v, idx, sph)
    o, d = random_rays(8000, 25)
    check(gpu, orc, o, d, occ_maxt=F32(0.7))
    for m in (1, 2, 3):  # one leaf, one interior node, uneven
        v, idx = scenes.random_triangles(m, 30 + m, lmin=0.3, lmax=0.6)
        gpu, orc = pair(v, idx)
        o, d = random_rays(3000, 40 + m)
        check(gpu, orc, o, d)


def test_render_and_serve_with_the_gpu_tree():
    # a direct-lighting frame over the GPU-built tree: every sample equals the
    # oracle's unless its camera or shadow rays met a tie (none expected on
    # this fBm-displaced sphere at these sample positions); and a ray server
    # over the GPU-built scene answers as the scene alone
    scene = scenes.direct_lighting_scene(level=4)
    cam = scenes.pinhole(64, 48)
    integ = dict(kind="direct", nx=2, ny=2, seed=7, light_samples=4)
    ref = ol.render(scene, cam, integ, nthreads=16, want_samples=True)
    for path in ("plain", "batched"):
        r = px.SamplerRenderer(scene, cam, px.StratifiedSampler(2, 2, 7), px.DirectLightingIntegrator(4),
                               split="lbvh")
        img, st = r.render(path)
        samples = r.sample_radiance().reshape(ref["samples"].shape)
        r.close()
        bad = np.any(samples.view(np.uint32) != ref["samples"].view(np.uint32), axis=-1)
        assert bad.mean() <= 1e-3, bad.sum()
        assert float(np.sqrt(np.mean((img.astype(np.float64) - ref["rgb"]) ** 2))) <= 1e-4
    v, idx = scenes.random_triangles(20000, 26)
    gpu, orc = pair(v, idx)
    srv = px.RayServer(arena_bytes=64 << 20, batch_rays=1 << 16)
    sid = srv.add_scene(gpu)
    o, d = random_rays(70000, 27)
    rays = px.RayBatch(o, d)
    out, _ = srv.wait(srv.submit(sid, "closest", rays))
    alone = gpu.intersect(px.RayBatch(o, d), "batched")
    assert np.array_equal(out.view(np.uint32), alone.view(np.uint32))
    srv.close()
