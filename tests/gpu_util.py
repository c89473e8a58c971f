"""Shared helpers of the GPU parity tests."""
import numpy as np

import oracle_lib as ol
import pbrt_rust_amd as px
from pbrt_rust_amd import scenes

F32 = np.float32
FMAX = float(np.finfo(np.float32).max)


def mesh_pair(v, idx, split="sah", max_prims=1):
    bvh = px.BVHAccelerator([px.Primitive.simple(px.TriangleMesh(None, None, False, idx, v))], max_prims, split)
    orc = ol.OracleScene()
    orc.add_mesh(v, idx)
    orc.build(split, max_prims)
    return bvh, orc


def random_rays(n, seed, lo=-1.5, hi=1.5, target=1.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(F32)
    t = rng.uniform(-target, target, (n, 3)).astype(F32)
    d = (t - o).astype(F32)
    return o, d


def check_closest(bvh, orc, o, d, mint=0.0, maxt=FMAX, path="batched", nthreads=8):
    rays = px.RayBatch(o, d, mint, maxt)
    hits = bvh.intersect(rays, path)
    ref = orc.intersect(o, d, mint, maxt, nthreads=nthreads)
    assert np.array_equal(hits["prim"], ref["prim"]), np.nonzero(hits["prim"] != ref["prim"])[0][:10]
    h = ref["prim"] != px.NO_HIT
    for k in ("t", "b1", "b2"):
        assert np.array_equal(hits[k][h].view(np.uint32), ref[k][h].view(np.uint32)), k
    assert np.array_equal(rays.maxt.view(np.uint32), ref["maxt"].view(np.uint32))
    return hits, ref


def check_any(bvh, orc, o, d, mint=0.0, maxt=FMAX, path="batched", nthreads=8):
    occ = bvh.intersect_p(px.RayBatch(o, d, mint, maxt), path)
    ref, _ = orc.intersect_p(o, d, mint, maxt, nthreads=nthreads)
    assert np.array_equal(occ, ref.astype(bool)), np.nonzero(occ != ref.astype(bool))[0][:10]
    return occ
