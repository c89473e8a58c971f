"""GPU parity of Intersectable::intersect / intersect_p through the C ABI:
both traversal paths against the CPU oracle, bit for bit (prim id, t, b1, b2,
the maxt write-back, occlusion bits), on the reference's own fixtures,
random scenes, edge cases and a deep tree that uses the spill stack."""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol
import pbrt_rust_amd as px
from pbrt_rust_amd import scenes
import xf32
from gpu_util import FMAX, F32, check_any, check_closest, mesh_pair, random_rays

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))
PATHS = ["plain", "batched"]


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("method", ["sah", "middle", "equal"])
def test_eight_spheres_prim_ids(path, method):
    g = GOLD["spheres8"]
    prims = []
    for c in g["centers"]:
        t = px.Transform.translate(c)
        prims.append(px.Primitive.simple(px.Sphere(t, t.inverse(), False, 1.0, -1.0, 1.0, 360.0)))
    bvh = px.Aggregate.bvh(prims, 1, method)
    o = [r["o"] for r in g["rays"]]
    d = [r["d"] for r in g["rays"]]
    hits = bvh.intersect(px.RayBatch(o, d), path)
    assert hits["prim"].tolist() == [r["prim"] for r in g["rays"]]
    assert bvh.intersect_p(px.RayBatch(o, d), path).all()


@pytest.mark.parametrize("path", PATHS)
def test_tetrahedron_maxt_side_effect(path):
    g = GOLD["tetrahedron"]
    for method, mp in g["bvh_ray"]["builds"]:
        bvh = px.BVHAccelerator([px.Primitive.simple(px.TriangleMesh(None, None, False, g["indices"],
                                                                     g["points"]))], mp, method)
        rays = px.RayBatch([g["bvh_ray"]["o"]], [g["bvh_ray"]["d"]], 0.0)
        assert bvh.intersect_p(rays, path)[0]
        bvh.intersect(rays, path)
        assert rays.maxt[0] == g["bvh_ray"]["maxt_after"]


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("split", ["sah", "middle", "equal"])
def test_random_scene_parity(path, split):
    v, idx = scenes.random_triangles(3000, 11)
    bvh, orc = mesh_pair(v, idx, split)
    o, d = random_rays(20000, 12)
    hits, ref = check_closest(bvh, orc, o, d, path=path)
    assert (hits["prim"] != px.NO_HIT).mean() > 0.1
    check_any(bvh, orc, o, d, path=path)
    # segments: bounded maxt as VisibilityTester::segment produces
    check_any(bvh, orc, o, d, mint=1e-3, maxt=0.5, path=path)
    check_closest(bvh, orc, o, d, mint=0.1, maxt=0.7, path=path)


@pytest.mark.parametrize("path", PATHS)
def test_ray_query_camera_parity(path):
    # the plain ray-query workload at 128x128: camera rays, closest hit, then
    # one shadow segment per hit to the point light (2,2,-2), eps = t*5e-4
    L = ol.lib()
    s = scenes.ray_query_scene(16384)
    bvh, orc = mesh_pair(s["verts"], s["idx"], "sah")
    cam = scenes.pinhole(128, 128)
    r2c = np.zeros(16, F32)
    L.orc_camera_raster_to_camera(128, 128, ol.f32(cam["screen"]), cam["fov"], r2c)
    xy = np.zeros((128 * 128, 2), F32)
    for y in range(128):
        for x in range(128):
            p = np.zeros(2, F32)
            L.orc_pixel_samples(0xC0, 128, x, y, 1, 1, 1, p)
            xy[y * 128 + x] = p
    o = np.zeros(3 * xy.shape[0], F32); d = np.zeros_like(o)
    L.orc_camera_rays(r2c, ol.f32(cam["cam_to_world"]).reshape(16), xy.shape[0], xy.reshape(-1), o, d)
    o = o.reshape(-1, 3); d = d.reshape(-1, 3)
    hits, ref = check_closest(bvh, orc, o, d, path=path)
    h = hits["prim"] != px.NO_HIT
    assert h.mean() > 0.1
    t = hits["t"][h]
    p = (o[h] + d[h] * t[:, None]).astype(F32)
    eps = (t * F32(5e-4)).astype(F32)
    lp = np.array([2, 2, -2], F32)
    dist = np.sqrt(np.sum((p - lp) ** 2, axis=1)).astype(F32)
    dirn = ((lp - p) * (F32(1) / dist)[:, None]).astype(F32)
    check_any(bvh, orc, p, dirn, eps, dist, path=path)


@pytest.mark.parametrize("path", PATHS)
def test_edge_cases(path):
    v, idx = scenes.random_triangles(500, 21)
    bvh, orc = mesh_pair(v, idx)
    for n in (1, 2, 31, 33, 63, 64, 65, 257):
        o, d = random_rays(n, 100 + n)
        check_closest(bvh, orc, o, d, path=path)
        check_any(bvh, orc, o, d, path=path)
    # axis-aligned and zero-component directions (1/d = inf, NaN slabs)
    rng = np.random.default_rng(5)
    o = rng.uniform(-1.2, 1.2, (3000, 3)).astype(F32)
    d = np.zeros((3000, 3), F32)
    ax = rng.integers(0, 3, 3000)
    d[np.arange(3000), ax] = rng.choice([-1.0, 1.0, -0.0], 3000)
    d[:1000, (ax[:1000] + 1) % 3] = -0.0
    o[::7] = np.round(o[::7], 1)  # origins on slab planes
    check_closest(bvh, orc, o, d, path=path)
    check_any(bvh, orc, o, d, path=path)
    # empty intervals and starts beyond every hit
    o, d = random_rays(1000, 7)
    check_closest(bvh, orc, o, d, mint=5.0, maxt=1.0, path=path)
    check_any(bvh, orc, o, d, mint=1e6, maxt=FMAX, path=path)
    # zero rays is a no-op
    assert bvh.intersect(px.RayBatch(np.zeros((0, 3)), np.zeros((0, 3))), path).shape == (0,)


def cube_grid(n=8, step=0.5, size=0.25):
    """Axis-aligned cubes on a dyadic grid: node boxes whose faces lie on
    the planes that the rays below start on and graze."""
    cv = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], F32) * F32(size)
    cf = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1],
                   [2, 3, 7], [2, 7, 6], [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]], np.int32)
    vs, fs = [], []
    rng = np.random.default_rng(3)
    for i in range(n):
        for j in range(n):
            for k in range(n):
                if rng.random() < 0.4:
                    fs.append(cf + 8 * len(vs))
                    vs.append(cv + np.array([i, j, k], F32) * F32(step) - F32(n * step / 2))
    return np.concatenate(vs), np.concatenate(fs)


@pytest.mark.parametrize("path", PATHS)
def test_grazing_rays_on_box_planes(path):
    # Origins on the cubes' face planes, directions with components from
    # 2^-46 to 2^-30 (1/d on both sides of the 2^40 limit of the ray-space
    # record test) and origins beyond 2^40: near-parallel rays that skim
    # node faces, mixed in one wave with ordinary rays.
    v, idx = cube_grid()
    bvh, orc = mesh_pair(v, idx)
    rng = np.random.default_rng(41)
    n = 6000
    o = (np.round(rng.uniform(-2.2, 2.2, (n, 3)) * 4) / 4).astype(F32)
    d = rng.normal(size=(n, 3)).astype(F32)
    k = rng.integers(0, 3, n)
    sel = rng.random(n) < 0.6
    tiny = (rng.choice([-1.0, 1.0], n) * 2.0 ** rng.uniform(-46, -30, n)).astype(F32)
    d[sel, k[sel]] = tiny[sel]
    sel2 = rng.random(n) < 0.2
    d[sel2, (k[sel2] + 1) % 3] = tiny[::-1][sel2]
    far = rng.random(n) < 0.05
    o[far] = (o[far] * F32(2.0 ** 42)).astype(F32)
    check_closest(bvh, orc, o, d, path=path)
    check_any(bvh, orc, o, d, path=path)
    check_any(bvh, orc, o, d, mint=1e-3, maxt=0.75, path=path)


@pytest.mark.parametrize("path", PATHS)
def test_mixed_finite_and_infinite_inverse_directions(path):
    # Rays with finite and with infinite 1/d components (NaN slab bounds)
    # interleaved in one wave must match the oracle.
    v, idx = scenes.random_triangles(2000, 8)
    bvh, orc = mesh_pair(v, idx)
    rng = np.random.default_rng(77)
    n = 4096
    o, d = random_rays(n, 78)
    k = rng.integers(0, 3, n)
    sel = rng.random(n) < 0.3
    # subnormal components: 1/d overflows to +-inf
    d[sel, k[sel]] = rng.choice([1e-40, -1e-40, 0.0, -0.0, 1e-45], int(sel.sum())).astype(F32)
    o[::5] = np.round(o[::5], 1)
    check_closest(bvh, orc, o, d, path=path)
    check_any(bvh, orc, o, d, path=path)
    # NaN tmin is ignored by the slab max (Rust f32::max) on both sides
    check_any(bvh, orc, o, d, mint=float("nan"), maxt=2.0, path=path)


@pytest.mark.parametrize("path", PATHS)
def test_deep_tree_uses_spill_stack(path):
    # triangles in the planes x = 16^k: every SAH split peels off the farthest
    # one (all other centroids share bucket 0), giving a 29-level chain that is
    # deeper than both LDS stacks, so the global spill stack is exercised
    n = 31
    xs = 16.0 ** np.arange(n)
    v = np.concatenate([np.array([[x, 0, 0], [x, 1, 0], [x, 0, 1]]) for x in xs]).astype(F32)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    bvh, orc = mesh_pair(v, idx)
    assert bvh.info()["depth"] == 29
    rng = np.random.default_rng(1)
    m = 4000
    k = rng.integers(0, n, m)
    tmin = np.where(k == 0, 0.0, xs[np.maximum(k - 1, 0)] * 1.5).astype(F32)
    o = np.stack([np.full(m, -1.0), rng.uniform(0.05, 0.45, m), rng.uniform(0.05, 0.45, m)], 1).astype(F32)
    d = np.tile(np.array([[1, 0, 0]], F32), (m, 1))
    d[::3, 1] = F32(1e-9)
    hits, ref = check_closest(bvh, orc, o, d, mint=tmin, path=path)
    assert len(np.unique(hits["prim"])) > 25
    check_any(bvh, orc, o, d, mint=tmin, path=path)
    px.synchronize()  # raises if a traversal stack overflowed


@pytest.mark.parametrize("path", PATHS)
def test_mixed_triangles_and_spheres(path):
    v, idx = scenes.random_triangles(300, 31)
    prims = [px.Primitive.simple(px.TriangleMesh(None, None, False, idx, v))]
    orc = ol.OracleScene()
    orc.add_mesh(v, idx)
    rng = np.random.default_rng(4)
    for k in range(20):
        c = rng.uniform(-1, 1, 3)
        t = px.Transform.translate(c) * px.Transform.scale(0.1, 0.1, 0.1)
        prims.append(px.Primitive.simple(px.Sphere(t, t.inverse(), False, 1.0, -1.0, 1.0, 360.0)))
        orc.add_sphere(t.m, t.m_inv, 1.0, -1.0, 1.0, 360.0)
    bvh = px.BVHAccelerator(prims, 1, "sah")
    orc.build("sah", 1)
    o, d = random_rays(5000, 8)
    hits, ref = check_closest(bvh, orc, o, d, path=path)
    assert (hits["prim"] >= 300).sum() > 10   # some sphere hits
    check_any(bvh, orc, o, d, path=path)


def test_paths_agree_at_full_size():
    # size-independent property at the ray-query configuration's full size:
    # 16,384 triangles, 1M rays -- plain and batched give identical records,
    # and the occlusion bit equals "closest hit exists"
    s = scenes.ray_query_scene(16384)
    bvh = px.BVHAccelerator([px.Primitive.simple(px.TriangleMesh(None, None, False, s["idx"], s["verts"]))], 1, "sah")
    o, d = random_rays(1 << 20, 77)
    ra, rb = px.RayBatch(o, d), px.RayBatch(o, d)
    ha = bvh.intersect(ra, "plain")
    hb = bvh.intersect(rb, "batched")
    assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32))
    assert np.array_equal(ra.maxt.view(np.uint32), rb.maxt.view(np.uint32))
    occ = bvh.intersect_p(px.RayBatch(o, d), "batched")
    assert np.array_equal(occ, ha["prim"] != px.NO_HIT)
    assert np.array_equal(bvh.intersect_p(px.RayBatch(o, d), "plain"), occ)


@pytest.mark.parametrize("stack", [8, 24])
def test_stack_depths_match_oracle(stack):
    # the per-ray kernel at another LDS stack depth: the 29-level chain (deep
    # spills at 8 entries, none at 24) and incoherent rays on a random scene
    try:
        for kind in ("closest", "any"):
            px.set_tuning("batched", kind, 256, 0, 10, "ray", stack=stack)
        n = 31
        xs = 16.0 ** np.arange(n)
        v = np.concatenate([np.array([[x, 0, 0], [x, 1, 0], [x, 0, 1]]) for x in xs]).astype(F32)
        idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
        bvh, orc = mesh_pair(v, idx)
        rng = np.random.default_rng(2)
        m = 3000
        k = rng.integers(0, n, m)
        tmin = np.where(k == 0, 0.0, xs[np.maximum(k - 1, 0)] * 1.5).astype(F32)
        o = np.stack([np.full(m, -1.0), rng.uniform(0.05, 0.45, m), rng.uniform(0.05, 0.45, m)], 1).astype(F32)
        d = np.tile(np.array([[1, 0, 0]], F32), (m, 1))
        check_closest(bvh, orc, o, d, mint=tmin)
        check_any(bvh, orc, o, d, mint=tmin)
        v, idx = scenes.random_triangles(20000, 12)
        bvh, orc = mesh_pair(v, idx)
        o, d = random_rays(30000, 13)
        check_closest(bvh, orc, o, d)
        check_any(bvh, orc, o, d, maxt=0.7)
        px.synchronize()  # raises if a traversal stack overflowed
    finally:
        px.reset_tuning()


@pytest.fixture
def packet_occlusion():
    px.set_tuning("batched", "any", 256, 0, 0, "packet")
    yield
    px.reset_tuning()


@pytest.mark.skip(reason="hung once on the GPU (no output for 180 s) while a build with early block exits and "
                         "multi-packet work claims in the traversal kernels ran it; the cause was not shown, so the "
                         "packet occlusion kernel was removed and pbrtx_set_tuning refuses the scheme for "
                         "occlusion (tests/test_capi_symbols.py checks the refusal)")
def test_packet_occlusion_scheme_matches_oracle(packet_occlusion):
    # one stack per wave of 64 rays: incoherent rays, the point-light shadow
    # segments of the ray-query workload, the 8-sphere fixture, ragged counts
    v, idx = scenes.random_triangles(3000, 11)
    bvh, orc = mesh_pair(v, idx)
    for n in (1, 63, 64, 65, 20000):
        o, d = random_rays(n, 300 + n)
        check_any(bvh, orc, o, d, path="batched")
        check_any(bvh, orc, o, d, mint=1e-3, maxt=0.5, path="batched")
    s = scenes.ray_query_scene(16384)
    bvh, orc = mesh_pair(s["verts"], s["idx"], "sah")
    rng = np.random.default_rng(3)
    p = rng.uniform(-1, 1, (50000, 3)).astype(F32)
    lp = np.array([2, 2, -2], F32)
    dist = np.sqrt(np.sum((p - lp) ** 2, axis=1)).astype(F32)
    dirn = ((lp - p) * (F32(1) / dist)[:, None]).astype(F32)
    occ = check_any(bvh, orc, p, dirn, F32(1e-4), dist, path="batched")
    assert 0.05 < occ.mean() < 0.95
    g = GOLD["spheres8"]
    prims = []
    for c in g["centers"]:
        t = px.Transform.translate(c)
        prims.append(px.Primitive.simple(px.Sphere(t, t.inverse(), False, 1.0, -1.0, 1.0, 360.0)))
    assert px.Aggregate.bvh(prims, 1, "sah").intersect_p(
        px.RayBatch([r["o"] for r in g["rays"]], [r["d"] for r in g["rays"]]), "batched").all()


@pytest.fixture
def exact_records():
    px.set_tuning("batched", "closest", 256, 0, 0, "ray_exact")
    px.set_tuning("batched", "any", 256, 0, 10, "ray_exact")
    yield
    px.reset_tuning()


def test_exact_record_scheme_matches_oracle(exact_records):
    # the batched kernels on the exact 128-byte records (the default reads the
    # compressed ones): the same fixtures and edge cases
    test_random_scene_parity("batched", "sah")
    test_random_scene_parity("batched", "equal")
    test_edge_cases("batched")
    test_mixed_finite_and_infinite_inverse_directions("batched")
    test_deep_tree_uses_spill_stack("batched")
    test_mixed_triangles_and_spheres("batched")
    test_eight_spheres_prim_ids("batched", "middle")


@pytest.fixture
def packet_closest():
    px.set_tuning("batched", "closest", 256, 0, 0, "packet")
    yield
    px.reset_tuning()


def test_packet_closest_scheme_matches_oracle(packet_closest):
    # one stack per group of equal direction signs in a wave of 64 rays:
    # coherent camera rays, incoherent rays (several sign groups per wave),
    # zero and subnormal components, spheres, the maxt write-back, and a tree
    # too deep for the per-wave stack (falls back to the per-ray kernel)
    test_ray_query_camera_parity("batched")
    test_random_scene_parity("batched", "sah")
    test_random_scene_parity("batched", "middle")
    test_edge_cases("batched")
    test_mixed_finite_and_infinite_inverse_directions("batched")
    # groups with every ray inside the 2^40 limits take the sign-ordered box
    # test, groups with one ray outside them the reference's swap
    test_grazing_rays_on_box_planes("batched")
    test_mixed_triangles_and_spheres("batched")
    test_eight_spheres_prim_ids("batched", "equal")
    test_tetrahedron_maxt_side_effect("batched")
    test_deep_tree_uses_spill_stack("batched")


@pytest.fixture(params=[1, 4, 16])
def occlusion_lanes(request):
    px.set_tuning("batched", "any", 256, 0, 60, "ray", lanes=request.param)
    yield request.param
    px.reset_tuning()


def test_occlusion_lanes_match_oracle(occlusion_lanes):
    # Batched occlusion launches of up to 2^20 rays take 8 lanes per ray by
    # default; here 1 (the per-ray kernel), 4 and 16: the same fixtures and
    # edge cases, the 29-level chain, spheres, and the full-size random soup
    # whose rays pass many boxes (the group stacks run up to their bound);
    # the overflow flag stays clear
    test_random_scene_parity("batched", "sah")
    test_random_scene_parity("batched", "middle")
    test_edge_cases("batched")
    test_grazing_rays_on_box_planes("batched")
    test_mixed_finite_and_infinite_inverse_directions("batched")
    test_deep_tree_uses_spill_stack("batched")
    test_mixed_triangles_and_spheres("batched")
    test_eight_spheres_prim_ids("batched", "sah")
    test_paths_agree_at_full_size()
    px.synchronize()


def test_occlusion_group_stack_fills(occlusion_lanes):
    # 2000 thin triangles along the diagonal of the square [-1,1]^2, stacked
    # in z: every box spans the square, so a ray along z away from the
    # diagonal passes every box and misses every triangle.  Its frontier is
    # the whole tree: the group stacks fill to their bound and continue a
    # depth-first walk from there.  Rays near the diagonal's wide end hit.
    n = 2000
    rng = np.random.default_rng(9)
    z = np.sort(rng.uniform(-1, 1, n))
    w = 0.05
    v = np.concatenate([np.array([[-1, -1, zz], [1, 1, zz], [-1 + w, -1, zz]]) for zz in z]).astype(F32)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    bvh, orc = mesh_pair(v, idx)
    m = 4000
    xy = rng.uniform(-1, 1, (m, 2))
    near = rng.random(m) < 0.3
    xy[near, 0] = rng.uniform(-0.99, -0.97, int(near.sum()))
    xy[near, 1] = xy[near, 0] - rng.uniform(0.0, 0.01, int(near.sum()))
    o = np.concatenate([xy, np.full((m, 1), -2.0)], 1).astype(F32)
    d = np.concatenate([rng.normal(0, 1e-3, (m, 2)), np.ones((m, 1))], 1).astype(F32)
    occ = check_any(bvh, orc, o, d, path="batched")
    assert 0.1 < occ.mean() < 0.5
    px.synchronize()  # raises if a traversal stack overflowed


def test_deep_chain_occlusion(occlusion_lanes):
    # triangles in the planes x = 2^k, split at the middle: a reference tree
    # 61 levels deep (its own records for closest hit; the occlusion records
    # come from a tree built over the same leaves).  Rays along +x inside
    # every box: ones that hit the k-th triangle past tmin, and ones outside
    # the triangles (y + z > 1) that visit the whole chain
    n = 100
    xs = 2.0 ** np.arange(n)
    v = np.concatenate([np.array([[x, 0, 0], [x, 1, 0], [x, 0, 1]]) for x in xs]).astype(F32)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    bvh, orc = mesh_pair(v, idx, "middle")
    assert bvh.info()["depth"] == 61
    rng = np.random.default_rng(n)
    m = 3000
    k = rng.integers(0, n, m)
    tmin = np.where(k == 0, 0.0, xs[np.maximum(k - 1, 0)] * 1.5).astype(F32)
    yz = rng.uniform(0.05, 0.45, (m, 2))
    out = rng.random(m) < 0.5
    yz[out] = rng.uniform(0.55, 0.95, (int(out.sum()), 2))
    o = np.concatenate([np.full((m, 1), -1.0), yz], 1).astype(F32)
    d = np.tile(np.array([[1, 0, 0]], F32), (m, 1))
    occ = check_any(bvh, orc, o, d, mint=tmin, path="batched")
    assert occ[~out].mean() > 0.9 and not occ[out].any()
    px.synchronize()  # raises if a traversal stack overflowed


@pytest.mark.parametrize("max_prims,split", [(4, "sah"), (8, "equal"), (3, "middle")])
def test_occlusion_leaves_of_several_prims(max_prims, split):
    # leaves of several prims: the occlusion records' own tree is built over
    # the reference's leaves, each leaf kept whole; occlusion at one and at
    # eight lanes per ray, and closest hit, against the oracle
    v, idx = scenes.random_triangles(20000, 40 + max_prims)
    bvh, orc = mesh_pair(v, idx, split, max_prims)
    o, d = random_rays(30000, 41)
    try:
        for lanes in (1, 8):
            px.set_tuning("batched", "any", 256, 0, 60, "ray", lanes=lanes)
            check_any(bvh, orc, o, d, path="batched")
            check_any(bvh, orc, o, d, mint=1e-3, maxt=0.4, path="batched")
    finally:
        px.reset_tuning()
    check_closest(bvh, orc, o, d, path="batched")
    px.synchronize()
