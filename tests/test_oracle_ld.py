"""The oracle's LDSampler image samples (sampler/lds.rs:17-70,
sampler/utils.rs:5-69): van_der_corput and sobol2 as the reference writes
them -- its van_der_corput's last swap shifts by 2, so bits move or drop, and
its sobol2 flips on a zero bit -- pinned by values worked by hand from those
lines; a pixel's samples are the scrambled (0,2) points, shuffled; renders
with the sampler round spp up to a power of two."""
import numpy as np
import pytest

import oracle_lib as ol
from pbrt_rust_amd import scenes

# (n, scramble) -> value, worked through the reference's shifts by hand
VDC = [((0, 0), 0.0), ((1, 0), 0.0), ((2, 0), 0.125), ((3, 0), 0.125), ((4, 0), 0.25),
       ((0, 0xFFFFFFFF), 1.0 - 2.0 ** -24), ((0, 0x12345678), 0x123456 / 2.0 ** 24)]
SOBOL = [((0, 0), 0.0), ((1, 0), 0.0), ((2, 0), 0.5), ((3, 0), 0.0), ((0, 0xFFFFFF00), 1.0 - 2.0 ** -24),
         ((2, 0x80000000), 0.0)]


@pytest.mark.parametrize("args,want", VDC)
def test_van_der_corput_as_written(args, want):
    assert ol.lib().orc_ld_vdc(*args) == np.float32(want)


@pytest.mark.parametrize("args,want", SOBOL)
def test_sobol2_as_written(args, want):
    assert ol.lib().orc_ld_sobol2(*args) == np.float32(want)


@pytest.mark.parametrize("spp", [1, 4, 16, 64])
def test_pixel_samples_are_the_scrambled_points_shuffled(spp):
    L = ol.lib()
    seed = 7
    orders = set()
    for pixel in (0, 1, 5, 1234):
        pts = np.zeros(2 * spp, np.float32)
        L.orc_ld_pixel_points(seed, pixel, spp, pts)
        pts = pts.reshape(-1, 2)
        s0 = L.orc_ld_draw(seed, pixel, 0) & 0xFFFFFFFF
        s1 = L.orc_ld_draw(seed, pixel, 1) & 0xFFFFFFFF
        base = np.array([(L.orc_ld_vdc(i, s0), L.orc_ld_sobol2(i, s1)) for i in range(spp)], np.float32)
        assert sorted(map(tuple, pts)) == sorted(map(tuple, base))
        assert ((pts >= 0) & (pts < 1)).all()
        orders.add(tuple(np.lexsort(pts.T)))
    if spp >= 4:
        assert len(orders) > 1  # shuffled per pixel


def test_render_rounds_spp_to_a_power_of_two_and_samples_inside_their_pixel():
    sc = scenes.direct_lighting_scene(level=2)
    cam = scenes.pinhole(6, 5)
    integ = dict(kind="direct", nx=3, ny=2, seed=3, light_samples=1, sampler=1)
    res = ol.render(sc, cam, integ, nthreads=4, want_samples=True)
    assert res["samples"].shape == (5, 6, 8, 3)
    pts = np.zeros(2 * 8, np.float32)
    ol.lib().orc_ld_pixel_points(3, 2 * 6 + 4, 8, pts)
    assert res["rgb"].shape == (5, 6, 3) and np.isfinite(res["rgb"]).all()
    assert (pts >= 0).all() and (pts < 1).all()
