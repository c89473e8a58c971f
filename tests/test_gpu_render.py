"""GPU renders against the oracle's SamplerRenderer restatement at equal spp
and identical per-pixel sample seeds.  The bar: per-sample radiance and the
film identical bit for bit (the oracle and the kernels run the same IEEE
operations in the same order); the stated image tolerance RMSE <= 1e-4 is
asserted as well, and the GPU image is checked to be the same for both
traversal paths, for every chunking, and across repeated frames."""
import numpy as np
import pytest

import oracle_lib as ol
import pbrt_rust_amd as px
from pbrt_rust_amd import scenes

pytestmark = pytest.mark.gpu
RMSE_TOL = 1e-4


def rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))


def oracle_render(scene, cam, integ, nthreads=16):
    return ol.render(scene, cam, integ, nthreads=nthreads, want_samples=True)


def gpu_render(scene, cam, integ, path, chunk=None):
    sampler = px.StratifiedSampler(integ["nx"], integ["ny"], integ["seed"])
    if integ["kind"] == "path":
        it = px.PathIntegrator(integ["max_depth"], integ["rr_start"])
    else:
        it = px.DirectLightingIntegrator(integ["light_samples"])
    r = px.SamplerRenderer(scene, cam, sampler, it, chunk=chunk)
    img, st = r.render(path)
    spp = integ["nx"] * integ["ny"]
    samples = r.sample_radiance().reshape(cam["yres"], cam["xres"], spp, 3)
    r.close()
    return img, samples, st


def assert_same(ref, img, samples):
    assert np.array_equal(samples.view(np.uint32), ref["samples"].view(np.uint32)), \
        np.argwhere(samples != ref["samples"])[:5]
    assert np.array_equal(img.view(np.uint32), ref["rgb"].view(np.uint32))
    assert rmse(img, ref["rgb"]) <= RMSE_TOL


DIRECT = dict(kind="direct", nx=2, ny=2, seed=7, light_samples=4)
PATH = dict(kind="path", nx=2, ny=2, seed=9, max_depth=5, rr_start=3)


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_direct_lighting_area_light_matches_oracle(path):
    scene = scenes.direct_lighting_scene(level=4)
    cam = scenes.pinhole(64, 48)
    ref = oracle_render(scene, cam, DIRECT)
    img, samples, st = gpu_render(scene, cam, DIRECT, path)
    assert_same(ref, img, samples)
    assert st["camera_rays"] == 64 * 48 * 4
    assert st["shadow_rays"] == ref["stats"][1]
    assert img.mean() > 0.01


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_direct_lighting_point_light_matches_oracle(path):
    scene = scenes.ray_query_scene(4096)
    cam = scenes.pinhole(48, 48)
    integ = dict(kind="direct", nx=1, ny=1, seed=0xC0, light_samples=1)
    ref = oracle_render(scene, cam, integ)
    img, samples, st = gpu_render(scene, cam, integ, path)
    assert_same(ref, img, samples)
    assert st["shadow_rays"] == ref["stats"][1]


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_path_tracing_matches_oracle(path):
    scene = scenes.spheres_scene(n_spheres=8, level=3, seed=5)
    cam = scenes.pinhole(48, 32, fov=60.0, cam_pos=(0.0, 0.0, -25.0))
    ref = oracle_render(scene, cam, PATH)
    img, samples, st = gpu_render(scene, cam, PATH, path)
    assert_same(ref, img, samples)
    assert st["shadow_rays"] == ref["stats"][1]
    assert st["extension_rays"] == ref["stats"][2]


@pytest.mark.parametrize("chunk", [None, 5000])
def test_octant_sorted_queues_match_oracle(chunk):
    # the extension queues sorted by direction octant (tiles of 4096 entries:
    # several tiles, a ragged last one) change no sample, also in graph mode
    scene = scenes.spheres_scene(n_spheres=40, level=3, seed=8)
    cam = scenes.pinhole(96, 64, fov=60.0, cam_pos=(0.0, 0.0, 0.0))
    ref = oracle_render(scene, cam, PATH)
    r = px.SamplerRenderer(scene, cam, px.StratifiedSampler(2, 2, PATH["seed"]), px.PathIntegrator(5, 3), chunk=chunk)
    try:
        r.set_octant_sort(True)
        for graph in (False, True, True):
            r.set_graph(graph)
            img, st = r.render("batched")
            samples = r.sample_radiance().reshape(cam["yres"], cam["xres"], 4, 3)
            assert_same(ref, img, samples)
            assert st["extension_rays"] == ref["stats"][2]
        assert ref["stats"][2] > 3 * 4096  # the queues span several tiles
    finally:
        r.close()


@pytest.mark.parametrize("kind", ["direct", "path"])
def test_large_frames_batched_equals_plain(kind):
    # thousands of look-back tiles in flight (queue offsets found many tiles
    # back): the batched frame must still equal the plain one bit for bit
    if kind == "path":
        scene = scenes.spheres_scene(n_spheres=40, level=3, seed=8)
        cam = scenes.pinhole(512, 512, fov=60.0, cam_pos=(0.0, 0.0, 0.0))
        it = px.PathIntegrator(5, 3)
    else:
        scene = scenes.direct_lighting_scene(level=4)
        cam = scenes.pinhole(512, 512)
        it = px.DirectLightingIntegrator(4)
    r = px.SamplerRenderer(scene, cam, px.StratifiedSampler(2, 2, 11), it)
    try:
        a, sa = r.render("batched")
        b, sb = r.render("plain")
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert sa["shadow_rays"] == sb["shadow_rays"] and sa["extension_rays"] == sb["extension_rays"]
        if kind == "path":
            r.set_octant_sort(True)
            c, _ = r.render("batched")
            assert np.array_equal(c.view(np.uint32), b.view(np.uint32))
    finally:
        r.close()


def test_chunking_and_repeat_frames_are_identical():
    scene = scenes.spheres_scene(n_spheres=6, level=3, seed=6)
    cam = scenes.pinhole(40, 30, cam_pos=(0.0, 0.0, -25.0))
    base, _, _ = gpu_render(scene, cam, PATH, "batched")
    for chunk in (4, 100, 333, 4800):
        img, _, _ = gpu_render(scene, cam, PATH, "batched", chunk=chunk)
        assert np.array_equal(img.view(np.uint32), base.view(np.uint32)), chunk
    plain, _, _ = gpu_render(scene, cam, PATH, "plain", chunk=1000)
    assert np.array_equal(plain.view(np.uint32), base.view(np.uint32))
    sampler = px.StratifiedSampler(2, 2, PATH["seed"])
    r = px.SamplerRenderer(scene, cam, sampler, px.PathIntegrator(5, 3))
    a, _ = r.render("batched")
    b, _ = r.render("batched")
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("nx,ny", [(1, 1), (2, 1), (2, 2), (3, 3), (4, 4), (8, 8)])
def test_camera_tile_packets(nx, ny):
    # camera-ray packets take pixel tiles when a launch covers whole rows
    # that the tiles divide (spp a power of two), else runs of 64 samples:
    # the whole frame, chunks of whole tile rows (launches starting at row
    # 8, 16, ...) and ragged chunks all give the oracle's samples
    scene = scenes.direct_lighting_scene(level=3)
    cam = scenes.pinhole(64, 48)
    integ = dict(kind="direct", nx=nx, ny=ny, seed=21, light_samples=1)
    ref = oracle_render(scene, cam, integ)
    spp = nx * ny
    for chunk in (None, 64 * spp * 8, 64 * spp * 8 + spp, 777):
        img, samples, _ = gpu_render(scene, cam, integ, "batched", chunk=chunk)
        assert_same(ref, img, samples)


def test_film_boundary_samples():
    # 1x1 spp with strata of width 1: with jitter the sample can sit exactly
    # on a pixel edge only when u = 0; a large image makes the gather path
    # across neighbours run and must still match the oracle
    scene = scenes.direct_lighting_scene(level=2)
    cam = scenes.pinhole(96, 64)
    integ = dict(kind="direct", nx=1, ny=1, seed=1234, light_samples=2)
    ref = oracle_render(scene, cam, integ)
    img, samples, _ = gpu_render(scene, cam, integ, "batched")
    assert_same(ref, img, samples)


def test_camera_ray_schemes_give_the_same_frame():
    # every camera-ray scheme (chosen by samples per pixel unless set) must
    # give the same samples, at few and at many samples per pixel
    scene = scenes.direct_lighting_scene(level=4)
    cam = scenes.pinhole(64, 48)
    ref = oracle_render(scene, cam, DIRECT)
    many = dict(DIRECT, nx=6, ny=6, light_samples=1)
    cam2 = scenes.pinhole(24, 16)
    ref2 = oracle_render(scene, cam2, many)
    try:
        for scheme in (None, "ray", "ray_exact", "packet"):
            if scheme:
                px.set_tuning("batched", "camera", 256, 0, 0, scheme)
            img, samples, st = gpu_render(scene, cam, DIRECT, "batched")
            assert_same(ref, img, samples)
            img, samples, st = gpu_render(scene, cam2, many, "batched")
            assert_same(ref2, img, samples)
    finally:
        px.reset_tuning()


@pytest.mark.parametrize("integ", [DIRECT, PATH], ids=["direct", "path"])
def test_graph_replays_match_eager_frames(integ):
    # graph mode: every replay, with other eager work and synchronizations in
    # between and a tuning change (recapture) halfway, gives the eager frame
    if integ["kind"] == "path":
        scene = scenes.spheres_scene(n_spheres=8, level=3, seed=5)
        cam = scenes.pinhole(48, 32, fov=60.0, cam_pos=(0.0, 0.0, -25.0))
    else:
        scene = scenes.direct_lighting_scene(level=4)
        cam = scenes.pinhole(64, 48)
    ref = oracle_render(scene, cam, integ)
    other = px.SamplerRenderer(scenes.direct_lighting_scene(level=3), scenes.pinhole(32, 24),
                               px.StratifiedSampler(2, 2, 3), px.DirectLightingIntegrator(2))
    sampler = px.StratifiedSampler(integ["nx"], integ["ny"], integ["seed"])
    it = px.PathIntegrator(integ["max_depth"], integ["rr_start"]) if integ["kind"] == "path" \
        else px.DirectLightingIntegrator(integ["light_samples"])
    r = px.SamplerRenderer(scene, cam, sampler, it)
    spp = integ["nx"] * integ["ny"]
    try:
        for path in ("batched", "plain"):
            eager, _ = r.render(path)
            r.set_graph(True)
            for k in range(4):
                if k == 2:
                    px.set_tuning("batched", "any", 128, 0, 10, "ray")
                img, st = r.render(path)
                samples = r.sample_radiance().reshape(cam["yres"], cam["xres"], spp, 3)
                assert np.array_equal(img.view(np.uint32), eager.view(np.uint32)), (path, k)
                assert_same(ref, img, samples)
                assert st["ms_total"] > 0
                other.render("batched")
                px.synchronize()
            r.set_graph(False)
            px.reset_tuning()
    finally:
        px.reset_tuning()
        r.close()
        other.close()


FILM_FILTERS = [px.Filter.mean(1.5, 1.5), px.Filter.triangle(2.0, 1.25), px.Filter.gaussian(2.0, 2.0, 2.0),
                px.Filter.mitchell(2.0, 2.0, 1 / 3, 1 / 3), px.Filter.lanczos(3.0, 3.0, 3.0),
                px.Filter.triangle(1.4999999, 0.9999)]  # half-widths just under the reach steps


@pytest.mark.parametrize("f", FILM_FILTERS, ids=["mean1.5", "triangle", "gaussian", "mitchell", "lanczos", "steps"])
@pytest.mark.parametrize("integ", [DIRECT, PATH], ids=["direct", "path"])
def test_film_filters_match_oracle(f, integ):
    # Film::add_sample with each filter's weight table: the image equals the
    # oracle's bit for bit, on both paths, in graph replays, and back to the
    # box filter after a change
    if integ["kind"] == "path":
        scene = scenes.spheres_scene(n_spheres=8, level=3, seed=5)
        cam = scenes.pinhole(40, 28, fov=60.0, cam_pos=(0.0, 0.0, -25.0))
    else:
        scene = scenes.direct_lighting_scene(level=4)
        cam = scenes.pinhole(37, 29)
    ref = ol.render(scene, cam, integ, nthreads=16, filt=f.args())
    box = ol.render(scene, cam, integ, nthreads=16)
    sampler = px.StratifiedSampler(integ["nx"], integ["ny"], integ["seed"])
    it = px.PathIntegrator(integ["max_depth"], integ["rr_start"]) if integ["kind"] == "path" \
        else px.DirectLightingIntegrator(integ["light_samples"])
    r = px.SamplerRenderer(scene, cam, sampler, it, filter=f)
    try:
        for path in ("plain", "batched"):
            img, _ = r.render(path)
            assert np.array_equal(img.view(np.uint32), ref["rgb"].view(np.uint32)), path
        r.set_graph(True)
        for _ in range(3):
            img, _ = r.render("batched")
            assert np.array_equal(img.view(np.uint32), ref["rgb"].view(np.uint32))
        r.set_filter(px.Filter.mean(0.5, 0.5))
        img, _ = r.render("batched")
        assert np.array_equal(img.view(np.uint32), box["rgb"].view(np.uint32))
    finally:
        r.close()


@pytest.mark.parametrize("integ", [dict(DIRECT, sampler=1, nx=3, ny=2), dict(PATH, sampler=1, nx=4, ny=4)],
                         ids=["direct-6to8spp", "path-16spp"])
def test_low_discrepancy_sampler_matches_oracle(integ):
    # LDSampler: spp rounded up to a power of two, per-pixel scrambled (0,2)
    # points shuffled; every sample's radiance and the film equal the
    # oracle's on both paths, and with a filter
    if integ["kind"] == "path":
        scene = scenes.spheres_scene(n_spheres=8, level=3, seed=5)
        cam = scenes.pinhole(40, 28, fov=60.0, cam_pos=(0.0, 0.0, -25.0))
        it = px.PathIntegrator(integ["max_depth"], integ["rr_start"])
    else:
        scene = scenes.direct_lighting_scene(level=4)
        cam = scenes.pinhole(37, 29)
        it = px.DirectLightingIntegrator(integ["light_samples"])
    ref = oracle_render(scene, cam, integ)
    spp = ref["samples"].shape[2]
    sampler = px.LDSampler(integ["nx"] * integ["ny"], integ["seed"])
    assert sampler.maximum_sample_count() == spp
    r = px.SamplerRenderer(scene, cam, sampler, it)
    try:
        for path in ("plain", "batched"):
            img, st = r.render(path)
            samples = r.sample_radiance().reshape(cam["yres"], cam["xres"], spp, 3)
            assert_same(ref, img, samples)
            assert st["camera_rays"] == cam["xres"] * cam["yres"] * spp
        f = px.Filter.gaussian(2.0, 2.0, 2.0)
        reff = ol.render(scene, cam, integ, nthreads=16, filt=f.args())
        r.set_filter(f)
        img, _ = r.render("batched")
        assert np.array_equal(img.view(np.uint32), reff["rgb"].view(np.uint32))
    finally:
        r.close()


def test_scene_file_renders_as_the_oracle():
    # a scene file through the scene-description API: the renderer it makes
    # gives the oracle's image of the same scene, camera, sampler and filter
    import test_pbrt_api as tp
    from pbrt_rust_amd import pbrt_api as pa

    class Keep(pa.Pbrt):
        pass

    p = Keep()
    r = pa.parse(tp.SCENE, p)
    try:
        img, st = r.render("batched")
        integ = dict(kind="direct", nx=2, ny=2, seed=0, light_samples=2)
        ref = ol.render(p.scene(), p.camera_dict(), integ, nthreads=16, filt=p.filter_obj().args())
        assert np.array_equal(img.view(np.uint32), ref["rgb"].view(np.uint32))
        assert st["shadow_rays"] == ref["stats"][1] and img.mean() > 0.0
    finally:
        r.close()


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_mirror_closed_form_scenes_match_oracle(path):
    # the scenes whose oracle radiance tests/test_oracle_mirror.py pins by
    # closed forms: mirror -> Lambertian under a point light, and mirror ->
    # area light (emission after a specular bounce); the device gives the
    # oracle's samples bit for bit
    import test_oracle_mirror as tm
    v, idx = tm._two_quads(False)
    point = dict(verts=v, idx=idx, face_mat=np.array([1, 1, 0, 0], np.uint32),
                 mats=np.array([[0, 0.6, 0.6, 0.6], [1, 0.9, 0.9, 0.9]], np.float32),
                 face_light=np.full(4, -1, np.int32), area_le=np.zeros((0, 3), np.float32),
                 point_lights=np.array([[0, 4, 5, 20, 10, 5]], np.float32))
    emit = dict(point, face_light=np.array([-1, -1, 0, 0], np.int32),
                area_le=np.array([[4, 2, 1]], np.float32), point_lights=np.zeros((0, 6), np.float32))
    for scene in (point, emit):
        for integ in (tm.PATH1, dict(tm.PATH1, max_depth=3, rr_start=1)):
            ref = oracle_render(scene, tm.CAM, integ)
            img, samples, _ = gpu_render(scene, tm.CAM, integ, path)
            assert_same(ref, img, samples)
            assert img.max() > 0


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_headline_settings_on_a_band(path):
    # the direct-lighting headline's settings -- stratified 4x4 spp, 4 light
    # samples per hit, the quad area light, fBm-displaced icosphere and ground
    # -- at level 6 (81,920 sphere triangles) and 256x256: a band of rows of
    # the GPU frame equals the oracle's render of those rows, sample for
    # sample and pixel for pixel
    scene = scenes.direct_lighting_scene(level=6)
    cam = scenes.pinhole(256, 256)
    integ = dict(kind="direct", nx=4, ny=4, seed=1, light_samples=4)
    y0, y1 = 112, 144
    ref = ol.render(scene, cam, integ, y0=y0, y1=y1, nthreads=16, want_samples=True)
    sampler = px.StratifiedSampler(4, 4, 1)
    r = px.SamplerRenderer(scene, cam, sampler, px.DirectLightingIntegrator(4))
    img, st = r.render(path)
    spp = 16
    samples = r.sample_radiance(y1 * 256 * spp)[y0 * 256 * spp:].reshape(y1 - y0, 256, spp, 3)
    r.close()
    assert np.array_equal(samples.view(np.uint32), ref["samples"].view(np.uint32))
    assert np.array_equal(img[y0:y1].view(np.uint32), ref["rgb"].view(np.uint32))
    assert st["camera_rays"] == 256 * 256 * spp and st["shadow_rays"] > 0


def test_write_image_and_scene_file_run(tmp_path):
    # Film::write_image of a rendered frame: the .pfm holds the frame's rgb
    # bit for bit, the .png write_img's encoding of it; pbrt_api.run renders
    # a scene file and writes the Film's "filename"
    import test_image_out as ti
    import test_pbrt_api as tp
    from pbrt_rust_amd import pbrt_api as pa
    scene = scenes.direct_lighting_scene(level=3)
    cam = scenes.pinhole(40, 24)
    r = px.SamplerRenderer(scene, cam, px.StratifiedSampler(2, 2, 3), px.DirectLightingIntegrator(2))
    with pytest.raises(px.PbrtxError):
        r.write_image(str(tmp_path / "early.pfm"))  # no frame yet
    img, _ = r.render("batched")
    r.write_image(str(tmp_path / "f.pfm"))
    r.write_image(str(tmp_path / "f.png"))
    r.close()
    assert np.array_equal(ti.read_pfm(str(tmp_path / "f.pfm")).view(np.uint32), img.view(np.uint32))
    assert np.array_equal(ti.read_png(str(tmp_path / "f.png")), ti.want8(img))
    out = str(tmp_path / "scene.png")
    text = tp.SCENE.replace('Film "image"', 'Film "image" "string filename" "%s"' % out)
    fn = tmp_path / "scene.pbrt"
    fn.write_text(text)
    assert pa.run([str(fn)]) == [out]
    p = pa.Pbrt()
    pa.parse(text, p).close()
    want = ol.render(p.scene(), p.camera_dict(), dict(kind="direct", nx=2, ny=2, seed=0, light_samples=2),
                     nthreads=16, filt=p.filter_obj().args())
    assert np.array_equal(ti.read_png(out), ti.want8(want["rgb"]))


@pytest.mark.parametrize("integ", [dict(DIRECT, sampler=2, nx=3, ny=1), dict(PATH, sampler=2, nx=2, ny=2)],
                         ids=["direct", "path"])
def test_halton_sampler_matches_oracle(integ):
    # HaltonSampler over the image as one window: pixels receive varying
    # numbers of samples (K slots, the rest empty); every slot's radiance and
    # the film equal the oracle's on both paths, with the box film and with a
    # filter, on a non-square image
    if integ["kind"] == "path":
        scene = scenes.spheres_scene(n_spheres=8, level=3, seed=5)
        cam = scenes.pinhole(40, 28, fov=60.0, cam_pos=(0.0, 0.0, -25.0))
        it = px.PathIntegrator(integ["max_depth"], integ["rr_start"])
    else:
        scene = scenes.direct_lighting_scene(level=4)
        cam = scenes.pinhole(37, 29)
        it = px.DirectLightingIntegrator(integ["light_samples"])
    ref = oracle_render(scene, cam, integ)
    K, cnt, _ = ol.halton_table(cam["xres"], cam["yres"], integ["nx"] * integ["ny"])
    assert ref["samples"].shape[2] == K and cnt.min() < K  # some slots empty
    r = px.SamplerRenderer(scene, cam, px.HaltonSampler(integ["nx"] * integ["ny"], integ["seed"]), it)
    try:
        assert r.spp == K
        for path in ("plain", "batched"):
            img, st = r.render(path)
            samples = r.sample_radiance().reshape(cam["yres"], cam["xres"], K, 3)
            assert_same(ref, img, samples)
            empty = np.arange(K)[None, None, :] >= cnt[:, :, None]
            assert np.all(samples[empty] == 0)
        f = px.Filter.mitchell(2.0, 2.0, 1 / 3, 1 / 3)
        reff = ol.render(scene, cam, integ, nthreads=16, filt=f.args())
        r.set_filter(f)
        for path in ("plain", "batched"):
            img, _ = r.render(path)
            assert np.array_equal(img.view(np.uint32), reff["rgb"].view(np.uint32))
    finally:
        r.close()


@pytest.mark.parametrize("integ", [DIRECT, PATH], ids=["direct", "path"])
def test_sphere_shapes_match_oracle(integ):
    # Sphere shapes in the renderer (full, clipped, non-uniformly scaled and
    # handedness-swapped, Lambertian and mirror) on both paths.  Hit ids and
    # hit/miss are exact; the sphere's shading frame takes acos and sin,
    # which the device evaluates with its own math library and the oracle
    # with the host C library, so radiance is compared to a tolerance: the
    # per-sample relative difference of 1e-4 for all but 0.2% of samples (a
    # path that leaves a sphere a few ulps apart can meet a different
    # surface), and the image within RMSE 1e-5 of the oracle's (measured on
    # MI355X: no sample off, RMSE 4e-9 direct and 2e-8 path)
    scene = scenes.analytic_spheres_scene()
    cam = scenes.pinhole(64, 48, fov=60.0, cam_pos=(0.0, 1.0, -10.0))
    ref = oracle_render(scene, cam, integ)
    for path in ("plain", "batched"):
        img, samples, st = gpu_render(scene, cam, integ, path)
        want = ref["samples"].astype(np.float64)
        got = samples.astype(np.float64)
        close = np.all(np.abs(got - want) <= 1e-4 * np.abs(want) + 1e-6, axis=-1)
        frac_off = 1.0 - close.mean()
        e = rmse(img, ref["rgb"])
        print("spheres %s %s: %.5f of samples off, image rmse %.3g, mean %.4f"
              % (integ["kind"], path, frac_off, e, float(ref["rgb"].mean())))
        assert frac_off <= 0.002
        assert e <= 1e-5
        assert img.mean() > 0.05


def test_scene_file_with_spheres_renders_near_the_oracle():
    # a scene file with sphere shapes (ReverseOrientation, a clipped sphere,
    # a mirror sphere) through the scene-description API renders within the
    # sphere tolerance (image RMSE 1e-5) of the oracle's image of the same scene
    from pbrt_rust_amd import pbrt_api as pa
    text = """LookAt 0 1 -6  0 0 0  0 1 0
    Camera "perspective" "float fov" [ 45 ]
    Film "image" "integer xresolution" [ 48 ] "integer yresolution" 40
    Sampler "stratified" "integer xsamples" 2 "integer ysamples" 2 "bool jitter" "true"
    SurfaceIntegrator "path" "integer maxdepth" [ 3 ]
    WorldBegin
      AttributeBegin
        AreaLightSource "diffuse" "rgb L" [ 8 8 8 ]
        Translate 0 4 0
        Shape "trianglemesh" "point P" [ -1 0 -1  -1 0 1  1 0 1  1 0 -1 ] "integer indices" [ 0 1 2  0 2 3 ]
      AttributeEnd
      Material "matte" "rgb Kd" [ 0.6 0.6 0.6 ]
      Shape "trianglemesh" "point P" [ -9 -1 -9  -9 -1 9  9 -1 9  9 -1 -9 ] "integer indices" [ 0 1 2 0 2 3 ]
      AttributeBegin
        Translate -1.2 0 0
        Shape "sphere" "float radius" 0.9
      AttributeEnd
      AttributeBegin
        Translate 1.2 0 0.5
        ReverseOrientation
        Material "mirror" "rgb Kr" [ 0.9 0.9 0.9 ]
        Shape "sphere" "float radius" 0.8 "float zmin" -0.5 "float phimax" 300
      AttributeEnd
    WorldEnd"""
    p = pa.Pbrt()
    r = pa.parse(text, p)
    try:
        img, st = r.render("batched")
        integ = dict(kind="path", nx=2, ny=2, seed=0, max_depth=3, rr_start=3)
        ref = ol.render(p.scene(), p.camera_dict(), integ, nthreads=16)
        e = rmse(img, ref["rgb"])
        print("scene file with spheres: image rmse %.3g, mean %.4f" % (e, float(ref["rgb"].mean())))
        assert e <= 1e-5 and img.mean() > 0.02
    finally:
        r.close()
