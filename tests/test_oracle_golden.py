"""The CPU oracle against the known answers pbrt_rust's own tests hold
(tests/golden/reference_vectors.json) plus closed forms and brute force.
These pin the oracle before it is trusted as the GPU path's checker."""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as ol
import xf32

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))
L = ol.lib()
F = np.float32


def bbox_int(bmin, bmax, o, d, mint, maxt=np.finfo(np.float32).max):
    t0 = ctypes.c_float(); t1 = ctypes.c_float()
    hit = L.orc_bbox_intersect(ol.f32(bmin), ol.f32(bmax), ol.f32(o), ol.f32(d), mint, maxt,
                               ctypes.byref(t0), ctypes.byref(t1))
    return (t0.value, t1.value) if hit else None


def test_bbox_slab_vectors():
    g = GOLD["bbox_intersect"]
    bmin, bmax = g["box"]
    for c in g["cases"]:
        got = bbox_int(bmin, bmax, c["o"], c["d"], c["mint"])
        if c["hit"] is None:
            assert got is None, c
        else:
            assert got == tuple(float(x) for x in c["hit"]), (c, got)
    dg = g["diagonal"]
    d = np.asarray(dg["d_unnormalized"], F)
    d = d * (F(1) / F(math.sqrt(float(np.sum(d * d)))))
    t0, t1 = bbox_int(bmin, bmax, dg["o"], d, 0.0)
    assert t0 == 0.0
    assert abs(t1 - math.sqrt(12.0)) < dg["tol"]


def test_bbox_slab_matches_closed_form_random():
    # for axis-aligned rays the slab interval is exact (pmin - o)/d, (pmax - o)/d
    rng = np.random.default_rng(3)
    for _ in range(200):
        lo = rng.uniform(-2, 0, 3).astype(F); hi = lo + rng.uniform(0.1, 2, 3).astype(F)
        axis = rng.integers(0, 3)
        o = rng.uniform(-0.5, 0.5, 3).astype(F) * (hi - lo) + (lo + hi) / F(2)
        o[axis] = lo[axis] - F(3)
        d = np.zeros(3, F); d[axis] = 1
        got = bbox_int(lo, hi, o, d, 0.0)
        assert got is not None
        assert got[0] == F(lo[axis] - o[axis]) and got[1] == F(hi[axis] - o[axis])


def quad(a, b, c):
    t0 = ctypes.c_float(); t1 = ctypes.c_float()
    ok = L.orc_quadratic(a, b, c, ctypes.byref(t0), ctypes.byref(t1))
    return (t0.value, t1.value) if ok else None


def test_quadratic_vectors():
    for c in GOLD["quadratic"]["cases"]:
        got = quad(*c["abc"])
        assert got == (None if c["roots"] is None else tuple(float(x) for x in c["roots"])), c
    for i in range(2, 200):
        assert quad(1.0, -float(i), 0.0) == (0.0, float(i))
    s3 = F(math.sqrt(3.0))
    assert quad(-1.0, 8.0, -13.0) == (float(F(-s3) + F(4)), float(s3 + F(4)))


def test_partition_by_vectors():
    g = GOLD["partition_by"]
    for xs in g["median_split_cases"]:
        v = np.array(xs, np.int32)
        L.orc_partition_by_i32(v, v.size)
        m = v.size // 2
        assert sorted(v.tolist()) == sorted(xs)
        assert v[:m].max() <= v[m:].min()
    for c in g["exact_cases"]:
        v = np.array(c["in"], np.int32)
        L.orc_partition_by_i32(v, v.size)
        assert v.tolist() == c["out"], c
    v = np.array(g["points_first_x"]["xs"], F)
    L.orc_partition_by_f32(v, v.size)
    assert v[0] == g["points_first_x"]["first"]


def test_partition_by_random_is_permutation_and_three_way_first_pass():
    # The reference recurses on a sub-slice with that sub-slice's own midpoint
    # (utils/mod.rs:164-168), so the result is not always a median split; the
    # oracle keeps that behaviour.  What holds: a permutation, and with
    # distinct keys of size <= 2 the result is sorted.
    rng = np.random.default_rng(5)
    not_median = 0
    for n in range(1, 80):
        xs = rng.integers(-5, 5, n).astype(np.int32)
        v = xs.copy()
        L.orc_partition_by_i32(v, n)
        m = n // 2
        assert sorted(v.tolist()) == sorted(xs.tolist())
        if 0 < m < n and v[:m].max() > v[m:].min():
            not_median += 1
    assert not_median > 0   # the quirk is present in the restatement


def test_crop_and_sub_windows():
    for c in GOLD["crop_window"]["cases"]:
        out = np.zeros(4, F)
        L.orc_get_crop_window(c["num"], c["count"], c["aspect"], out)
        assert out.tolist() == c["out"], c
    for c in GOLD["sub_window"]["cases"]:
        out = np.zeros(4, np.int32)
        L.orc_compute_sub_window(*c["window"], c["num"], c["count"], out)
        assert out.tolist() == c["out"], c


def test_stratified_nojitter():
    out = np.zeros(4, F)
    L.orc_stratified_1d_nojitter(out, 4)
    assert out.tolist() == GOLD["stratified_nojitter"]["1d_n4"]
    out = np.zeros(8, F)
    L.orc_stratified_2d_nojitter(out, 2, 2)
    assert out.tolist() == GOLD["stratified_nojitter"]["2d_2x2"]


def test_film_extents():
    for c in GOLD["film_extents"]["cases"]:
        out = np.zeros(8, np.int32)
        crop = np.array(c["crop"], F)
        L.orc_film_extents(c["res"][0], c["res"][1], c["width"][0], c["width"][1], crop, out)
        assert out[:4].tolist() == c["sample_extent"]
        assert out[4:].tolist() == c["pixel_extent"]


def test_matrix_inverse_lu():
    rng = np.random.default_rng(11)
    for _ in range(50):
        m = rng.uniform(-2, 2, (4, 4)).astype(F) + np.eye(4, dtype=F) * F(3)
        out = np.zeros(16, F)
        assert L.orc_matrix_inverse(ol.f32(m).reshape(16), out) == 0
        inv = out.reshape(4, 4).astype(np.float64)
        assert np.allclose(inv @ m.astype(np.float64), np.eye(4), atol=2e-5)
    # singular
    out = np.zeros(16, F)
    assert L.orc_matrix_inverse(np.zeros(16, F), out) == -1
    # a pure translation inverts exactly
    m, mi = xf32.translate([1.5, -2.25, 3.0])
    assert L.orc_matrix_inverse(ol.f32(m).reshape(16), out) == 0
    assert np.array_equal(out.reshape(4, 4), mi)


def tet():
    g = GOLD["tetrahedron"]
    return np.array(g["points"], F), np.array(g["indices"], np.uint32)


def test_triangle_refine_order_area_and_hits():
    g = GOLD["tetrahedron"]
    pts, idx = tet()
    area = ctypes.c_float()
    for k, v in enumerate(g["refined_v"]):
        vo = np.zeros(3, np.uint32)
        L.orc_mesh_triangle_query(None, ol.f32(pts).reshape(-1), 4, idx, 12, k, np.zeros(3, F),
                                  np.ones(3, F), 0.0, 1e30, ctypes.byref(area), vo)
        assert vo.tolist() == v
        if k > 0:
            assert area.value == g["area_tris_1_2_3"]
    sm, _ = xf32.scale(2, 2, 2)
    for k in (1, 2, 3):
        vo = np.zeros(3, np.uint32)
        L.orc_mesh_triangle_query(sm.reshape(16).ctypes.data, ol.f32(pts).reshape(-1), 4, idx, 12, k,
                                  np.zeros(3, F), np.ones(3, F), 0.0, 1e30, ctypes.byref(area), vo)
        assert area.value == g["area_tris_1_2_3_scaled2"]
    for c in g["tri0_rays"]:
        vo = np.zeros(3, np.uint32)
        hit = L.orc_mesh_triangle_query(None, ol.f32(pts).reshape(-1), 4, idx, 12, 0, ol.f32(c["o"]),
                                        ol.f32(c["d"]), 0.0, float(np.finfo(F).max),
                                        ctypes.byref(area), vo)
        assert bool(hit) == c["hit"], c


def test_bvh_tetrahedron_maxt():
    g = GOLD["tetrahedron"]["bvh_ray"]
    pts, idx = tet()
    for method, mp in g["builds"]:
        s = ol.OracleScene()
        s.add_mesh(pts, idx)
        s.build(method, mp)
        occ, _ = s.intersect_p([g["o"]], [g["d"]], g["mint"], np.finfo(F).max)
        assert occ[0] == 1
        r = s.intersect([g["o"]], [g["d"]], g["mint"], np.finfo(F).max)
        assert r["maxt"][0] == g["maxt_after"]
        assert r["t"][0] == g["maxt_after"]


def spheres_scene(centers, method):
    g = GOLD["spheres8"]["sphere"]
    s = ol.OracleScene()
    for c in centers:
        m, mi = xf32.translate(c)
        s.add_sphere(m, mi, g["radius"], g["z0"], g["z1"], g["phimax"])
    s.build(method, 1)
    return s


def test_bvh_eight_spheres_prim_ids():
    g = GOLD["spheres8"]
    for method in g["methods"]:
        s = spheres_scene(g["centers"], method)
        b, off, meta, leaf, order = s.nodes()
        leaves = sorted(off[leaf == 1].tolist())
        assert leaves == g["leaf_offsets_sorted"]
        assert np.all(meta[leaf == 1] == 1)
        for r in g["rays"]:
            res = s.intersect([r["o"]], [r["d"]], 0.0, np.finfo(F).max)
            assert res["prim"][0] == r["prim"], (method, r)


def test_bvh_node_bounds_vectors():
    for method in ("middle", "equal", "sah"):
        g = GOLD["bvh_bounds"][method]
        s = spheres_scene(g["centers"], method)
        b, off, meta, leaf, order = s.nodes()
        for k, (lo, hi) in g["nodes"].items():
            assert b[int(k)].tolist() == lo + hi, (method, k, b[int(k)])


def test_sphere_shape_vectors():
    g = GOLD["sphere_shape"]
    for c in g["hit_cases"]:
        t = xf32.translate(c["translate"])
        xf = xf32.mul(t, xf32.scale(c["scale"], c["scale"], c["scale"]))
        out = np.zeros(16, F)
        hit = L.orc_sphere_intersect(ol.f32(xf[0]).reshape(16), ol.f32(xf[1]).reshape(16), 0,
                                     c["radius"], c["z"][0], c["z"][1], c["phimax"], ol.f32(c["o"]),
                                     ol.f32(c["d"]), 0.0, float(np.finfo(F).max), out)
        assert bool(hit) == c["hit"], c
    c = g["info_case"]
    m, mi = xf32.translate(c["translate"])
    out = np.zeros(16, F)
    assert L.orc_sphere_intersect(m.reshape(16), mi.reshape(16), 0, c["radius"], c["z"][0], c["z"][1],
                                  c["phimax"], ol.f32(c["o"]), ol.f32(c["d"]), 0.0,
                                  float(np.finfo(F).max), out) == 1
    assert out[0] == c["t"]
    assert out[1] == F(F(0.5) * F(5e-4))
    assert out[2:5].tolist() == c["p"]
    assert out[5] == 0.0 and out[7] == 0.0 and abs(out[6] - 1.0) < c["tol"]
    assert out[8] == c["u"]
    assert abs(out[9] - c["v"]) < c["tol"]
    assert abs(out[10] - c["dpdu_x"]) < c["tol"] and out[11] == 0.0 and out[12] == 0.0
    assert out[13] == 0.0 and out[14] == 0.0 and abs(out[15] - c["dpdv_z"]) < c["tol"]


def test_projection_vectors():
    g = GOLD["projection"]
    xr, yr = g["film"]
    m, mi = xf32.translate([0, 0, -1])
    out = np.zeros(48, F)
    assert L.orc_projection(xr, yr, m.reshape(16), mi.reshape(16).ctypes.data,
                            ol.f32([0, 640, 0, 480]), out) == 0
    assert np.array_equal(out[:16].reshape(4, 4), np.array(g["raster_to_screen_640"], F))
    assert L.orc_projection(xr, yr, m.reshape(16), mi.reshape(16).ctypes.data, ol.f32([0, 1, 0, 1]), out) == 0
    assert np.abs(out[:16].reshape(4, 4) - np.array(g["raster_to_screen_unit"], F)).max() < g["mat_tol"]
    for c in g["raster_to_camera"]:
        assert L.orc_projection(xr, yr, m.reshape(16), mi.reshape(16).ctypes.data, ol.f32(c["screen"]), out) == 0
        r2c = out[32:].reshape(4, 4)
        v = xf32.xf_vector(r2c, c["v"])
        if "exact" in c:
            assert v.tolist() == c["exact"], (c, v)
        else:
            assert float(np.sum((v - np.array(c["approx"], F)) ** 2)) < c["tol_sq"]


def test_pinhole_camera_closed_form():
    xres, yres = 64, 48
    frame = xres / yres
    screen = ol.f32([-frame, frame, -1, 1])
    r2c = np.zeros(16, F)
    assert L.orc_camera_raster_to_camera(xres, yres, screen, 60.0, r2c) == 0
    c2w, _ = xf32.translate([0, 0, -3])
    xy = ol.f32([[xres / 2, yres / 2], [0, 0], [xres, yres], [17.25, 30.5]]).reshape(-1)
    o = np.zeros(12, F); d = np.zeros(12, F)
    L.orc_camera_rays(r2c, ol.f32(c2w).reshape(16), 4, xy, o, d)
    o = o.reshape(4, 3); d = d.reshape(4, 3)
    assert np.all(o == np.array([0, 0, -3], F))
    assert np.allclose(d[0], [0, 0, 1], atol=1e-6)
    tanh = math.tan(math.radians(30))
    for (x, y), dd in zip(xy.reshape(4, 2), d):
        px = (x / xres * 2 - 1) * frame * tanh
        py = -(y / yres * 2 - 1) * tanh
        ref = np.array([px, py, 1.0]); ref /= np.linalg.norm(ref)
        assert np.allclose(dd, ref, atol=2e-6), (x, y, dd, ref)


def random_tris(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (n, 1, 3))
    u = rng.normal(size=(n, 3, 3)); u /= np.linalg.norm(u, axis=2, keepdims=True)
    ln = rng.uniform(0.01, 0.05, (n, 3, 1)) * 4
    v = (c + ln * u).astype(F).reshape(-1, 3)
    return v, np.arange(3 * n, dtype=np.uint32)


@pytest.mark.parametrize("method", ["sah", "middle", "equal"])
def test_bvh_closest_hit_equals_brute_force(method):
    pts, idx = random_tris(400, 7)
    s = ol.OracleScene(); s.add_mesh(pts, idx); s.build(method, 1)
    rng = np.random.default_rng(8)
    n = 1500
    o = rng.uniform(-1.5, 1.5, (n, 3)).astype(F)
    tgt = rng.uniform(-1, 1, (n, 3)).astype(F)
    d = (tgt - o).astype(F)
    res = s.intersect(o, d, 0.0, np.finfo(F).max)
    bp, bt, ties = s.brute(o, d, 0.0, np.finfo(F).max)
    hit = bp != 0xFFFFFFFF
    assert np.array_equal(res["prim"] != 0xFFFFFFFF, hit)
    assert np.array_equal(res["t"][hit], bt[hit])
    uniq = hit & (ties == 0)
    assert np.array_equal(res["prim"][uniq], bp[uniq])
    occ, _ = s.intersect_p(o, d, 0.0, np.finfo(F).max)
    assert np.array_equal(occ.astype(bool), hit)


def test_bvh_structure_invariants():
    pts, idx = random_tris(1000, 9)
    s = ol.OracleScene(); s.add_mesh(pts, idx); s.build("sah", 4)
    b, off, meta, leaf, order = s.nodes()
    assert b.shape[0] == 2 * 1000 - 1          # SAH makes 1-prim leaves
    assert np.all(meta[leaf == 1] == 1)
    assert sorted(order.tolist()) == list(range(1000))
    tris = pts.reshape(-1, 3, 3)
    for i in np.nonzero(leaf)[0]:
        p = tris[order[off[i]]]
        assert np.all(p.min(0) >= b[i, :3]) and np.all(p.max(0) <= b[i, 3:])
    for i in np.nonzero(leaf == 0)[0]:
        c1, c2 = i + 1, off[i]
        assert np.array_equal(np.minimum(b[c1, :3], b[c2, :3]), b[i, :3])
        assert np.array_equal(np.maximum(b[c1, 3:], b[c2, 3:]), b[i, 3:])


def test_rng_and_pixel_samples_in_strata():
    xy = np.zeros(2 * 16, F)
    L.orc_pixel_samples(7, 100, 3, 5, 4, 4, 1, xy)
    xy = xy.reshape(16, 2)
    for s in range(16):
        sx, sy = s % 4, s // 4
        assert 3 + sx / 4 <= xy[s, 0] <= 3 + (sx + 1) / 4
        assert 5 + sy / 4 <= xy[s, 1] <= 5 + (sy + 1) / 4
    L.orc_pixel_samples(7, 100, 3, 5, 4, 4, 0, xy.reshape(-1))
    assert xy.reshape(16, 2)[0].tolist() == [3.125, 5.125]
    u = np.zeros(4, F)
    vals = []
    for k in range(2000):
        L.orc_rng4(1, k, 0, 0, u)
        vals.append(u.copy())
    vals = np.array(vals)
    assert vals.min() >= 0 and vals.max() < 1
    assert abs(vals.mean() - 0.5) < 0.02
