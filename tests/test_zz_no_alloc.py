"""pbrtx_intersect_p_noalloc: the occlusion query with its scratch in a fixed
device array instead of allocated memory gives exactly the bits of
pbrtx_intersect_p (batched path), on two calls in a row -- the second with a
smaller input, so stale counters or flags from the first would show -- and on
a deep tree whose spill stacks cut the persistent grid."""
import ctypes

import numpy as np
import pytest

import pbrt_rust_amd as px
from pbrt_rust_amd import scenes
from pbrt_rust_amd._lib import DeviceBuffer, check
from gpu_util import F32, mesh_pair, random_rays

pytestmark = pytest.mark.gpu


def bits(bvh, o, d, mint, maxt, noalloc):
    n = o.shape[0]
    rays = px.RayBatch(o, d, mint, maxt)
    dr = DeviceBuffer.from_array(rays.a)
    nw = (n + 31) // 32
    db = DeviceBuffer(nw * 4)
    # a stale pattern in the output: both entry points must write every word
    check(px.lib().pbrtx_memcpy_htod(db.ptr, np.full(nw, 0xA5A5A5A5, np.uint32).ctypes.data, nw * 4))
    if noalloc:
        check(px.lib().pbrtx_intersect_p_noalloc(bvh.scene, dr.ptr, n, db.ptr, None))
    else:
        bvh.intersect_p_device(dr.ptr, n, db.ptr, "batched")
    check(px.lib().pbrtx_synchronize())
    return db.download(np.uint32, nw)


def test_noalloc_occlusion_matches_intersect_p():
    v, idx = scenes.random_triangles(40000, 31)
    bvh, _ = mesh_pair(v, idx)
    for n, seed, maxt in ((200000, 1, 0.6), (7001, 2, 0.9)):
        o, d = random_rays(n, seed)
        want = bits(bvh, o, d, 1e-4, maxt, False)
        got = bits(bvh, o, d, 1e-4, maxt, True)
        assert np.array_equal(got, want), (n, np.nonzero(got != want)[0][:5])
        assert 0 < np.unpackbits(want.view(np.uint8)).sum() < n


def test_noalloc_on_a_deep_tree_and_empty_input():
    # triangles at x = 2^(k/4) (k < 160): each middle split of the centroid
    # range peels off the last few, so the tree is deep and its spill stacks
    # cut the persistent grid to what the fixed scratch holds
    k = np.arange(160)
    x = (2.0 ** (k / 4.0)).astype(F32)
    w = (0.05 * x).astype(F32)
    v = np.zeros((3 * k.size, 3), F32)
    v[0::3, 0] = x
    v[1::3, 0] = x + w
    v[1::3, 1] = 1.0
    v[2::3, 0] = x + 2 * w
    idx = np.arange(3 * k.size, dtype=np.uint32).reshape(-1, 3)
    bvh, _ = mesh_pair(v, idx, "middle")
    rng = np.random.default_rng(5)
    m = 30000
    xs = 2.0 ** rng.uniform(0, 39.5, m)
    o = np.stack([xs, rng.uniform(0.05, 0.6, m), np.full(m, -1.0)], 1).astype(F32)
    d = np.tile(np.array([[0, 0, 1]], F32), (m, 1))
    want = bits(bvh, o, d, 0.0, 2.0, False)
    got = bits(bvh, o, d, 0.0, 2.0, True)
    assert np.array_equal(got, want)
    assert np.unpackbits(want.view(np.uint8)).sum() > 0
    assert px.lib().pbrtx_intersect_p_noalloc(bvh.scene, None, 0, None, None) == 0
    assert px.lib().pbrtx_intersect_p_noalloc(None, None, 5, None, None) < 0
