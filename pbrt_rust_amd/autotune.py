"""Autotuner for the batched traversal kernels' launch shape.

For the renderer's camera-ray (closest-hit) and occlusion kernels separately
it sweeps threads per block, persistent blocks per CU, the idle-lane refill
threshold and the traversal scheme -- a stack per ray or one per wave of 64
consecutive rays -- on a real frame of a renderer, keeps the fastest
configuration of each (timed with HIP events on the render stream), and
stores both as JSON so later runs (bench, serving) start from them:

    python -m pbrt_rust_amd.autotune --out tuning.json
"""
import argparse
import itertools
import json
import os
import sys
import time

from . import api

DEFAULT_SPACE = {
    "block": [64, 128, 256],
    "blocks_per_cu": [0, 2, 4, 8],
    "refill_pct": [0, 10, 25, 40, 60],
}
# the packet scheme has no refill threshold
PACKET_SPACE = {"block": [64, 128, 256], "blocks_per_cu": [0, 2, 4, 8, 16]}
# on a direct-lighting frame the closest-hit stage traces the camera rays
STAGE = {"camera": "ms_trace_closest", "any": "ms_trace_any"}
KINDS = ("camera", "any")


def sweep(renderer, space=None, repeats=2, log=None):
    """Tune each kernel kind on `renderer` (batched path).  Returns
    {kind: (results, best)} and leaves the best shapes applied."""
    space = space or DEFAULT_SPACE
    out = {}
    for kind in KINDS:
        cands = [(b, c, f, "ray") for b, c, f in
                 itertools.product(space["block"], space["blocks_per_cu"], space["refill_pct"])]
        cands += [(b, c, 0, "packet") for b, c in
                  itertools.product(PACKET_SPACE["block"], PACKET_SPACE["blocks_per_cu"])]
        results = []
        for block, bpc, refill, scheme in cands:
            try:
                api.set_tuning("batched", kind, block, bpc, refill, scheme)
            except Exception as e:
                results.append(dict(block=block, blocks_per_cu=bpc, refill_pct=refill, scheme=scheme, error=str(e)))
                continue
            renderer.render("batched", want_image=False)
            ms = min(renderer.render("batched", want_image=False)[1][STAGE[kind]] for _ in range(repeats))
            r = dict(kind=kind, block=block, blocks_per_cu=bpc, refill_pct=refill, scheme=scheme, ms=ms)
            results.append(r)
            if log:
                log(r)
        best = min((r for r in results if "error" not in r), key=lambda r: r["ms"])
        apply({kind: best})
        out[kind] = (results, best)
    return out


def apply(cfg):
    for kind in ("closest", "any", "camera"):
        b = cfg.get(kind)
        if b:
            api.set_tuning("batched", kind, b["block"], b["blocks_per_cu"], b["refill_pct"], b.get("scheme", "ray"))


def load(path):
    """Apply a stored tuning file; returns its best shapes or None."""
    if not path or not os.path.exists(path):
        return None
    cfg = json.load(open(path)).get("best", {})
    apply(cfg)
    return cfg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="tuning.json")
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--spp-side", type=int, default=4)
    args = ap.parse_args()
    from . import scenes
    scene = scenes.direct_lighting_scene(level=args.level)
    cam = scenes.pinhole(args.res, args.res)
    n = args.res * args.res * args.spp_side ** 2
    r = api.SamplerRenderer(scene, cam, api.StratifiedSampler(args.spp_side, args.spp_side, 1),
                            api.DirectLightingIntegrator(4), chunk=n)
    t = time.time()
    res = sweep(r, log=lambda x: print(json.dumps(x), flush=True))
    best = {k: v[1] for k, v in res.items()}
    json.dump(dict(workload="direct lighting %dx%d %d spp level %d" % (args.res, args.res, args.spp_side ** 2,
                                                                        args.level),
                   results={k: v[0] for k, v in res.items()}, best=best, seconds=time.time() - t),
              open(args.out, "w"), indent=1)
    print("best", json.dumps(best))
    return 0


if __name__ == "__main__":
    sys.exit(main())
