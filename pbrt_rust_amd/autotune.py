"""Autotuner for the batched traversal kernels and the path tracer's queues.

For the renderer's camera-ray (closest-hit) and occlusion kernels separately
it sweeps threads per block, persistent blocks per CU, the idle-lane refill
threshold, the LDS stack depth and the traversal scheme -- a stack per ray
or one per wave of 64 consecutive rays -- on a real frame of a renderer,
keeps the fastest configuration of each (timed with HIP events on the render
stream), and stores them as JSON so later runs (bench, serving) start from
them.  With --path it also tunes the path tracer: the extension-ray kernel's
shape and stack depth and the queue capacity (samples per pass) on the
path-tracing workload.

    python -m pbrt_rust_amd.autotune --out tuning.json [--path]
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
    "stack": [8, 16, 24],
}
# queue capacities (samples per pass) tried for the path tracer; 0 = the
# whole frame in one pass
CHUNKS = [1 << 22, 1 << 24, 1 << 26, 0]
# the packet scheme has no refill threshold
# camera packets run a fixed schedule over any grid: more blocks than fit at
# once are started as others finish
PACKET_SPACE = {"block": [64, 128, 256], "blocks_per_cu": [0, 8, 16, 32, 64, 128]}
# on a direct-lighting frame the closest-hit stage traces the camera rays
STAGE = {"camera": "ms_trace_closest", "any": "ms_trace_any"}
KINDS = ("camera", "any")


def sweep(renderer, space=None, repeats=2, log=None):
    """Tune each kernel kind on `renderer` (batched path).  Returns
    {kind: (results, best)} and leaves the best shapes applied."""
    space = space or DEFAULT_SPACE
    out = {}
    for kind in KINDS:
        out[kind] = sweep_kind(renderer, kind, STAGE[kind], space, repeats, log)
    return out


def sweep_kind(renderer, kind, stage, space=None, repeats=2, log=None, packets=None):
    """Tune one kernel kind on a renderer's frames, timing `stage`.  The
    packet scheme is swept for camera rays only (packets=None); occlusion has
    no packet kernel and the library refuses it."""
    space = space or DEFAULT_SPACE
    if packets is None:
        packets = kind == "camera"
    if packets and kind == "any":
        raise ValueError("occlusion has no packet scheme")
    cands = [(b, c, f, "ray", k) for b, c, f, k in
             itertools.product(space["block"], space["blocks_per_cu"], space["refill_pct"], space["stack"])]
    if packets:
        cands += [(b, c, 0, "packet", 16) for b, c in
                  itertools.product(PACKET_SPACE["block"], PACKET_SPACE["blocks_per_cu"])]
    results = []
    for block, bpc, refill, scheme, stack in cands:
        try:
            api.set_tuning("batched", kind, block, bpc, refill, scheme, stack=stack)
        except Exception as e:
            results.append(dict(block=block, blocks_per_cu=bpc, refill_pct=refill, scheme=scheme, stack=stack,
                                error=str(e)))
            continue
        renderer.render("batched", want_image=False)
        ms = min(renderer.render("batched", want_image=False)[1][stage] for _ in range(repeats))
        r = dict(kind=kind, block=block, blocks_per_cu=bpc, refill_pct=refill, scheme=scheme, stack=stack, ms=ms)
        results.append(r)
        if log:
            log(r)
    best = min((r for r in results if "error" not in r), key=lambda r: r["ms"])
    apply({kind: best})
    return results, best


def sweep_chunk(renderer, n_samples, chunks=None, repeats=2, log=None):
    """Queue capacity (samples per pass) of a renderer: the fastest frame."""
    results = []
    for c in chunks or CHUNKS:
        c = n_samples if c == 0 or c > n_samples else c
        renderer.set_chunk(c)
        renderer.render("batched", want_image=False)
        ms = min(renderer.render("batched", want_image=False)[1]["ms_total"] for _ in range(repeats))
        r = dict(chunk=c, ms=ms)
        results.append(r)
        if log:
            log(r)
    best = min(results, key=lambda r: r["ms"])
    renderer.set_chunk(best["chunk"])
    return results, best


def apply(cfg):
    for kind in ("closest", "any", "camera"):
        b = cfg.get(kind)
        if b:
            api.set_tuning("batched", kind, b["block"], b["blocks_per_cu"], b["refill_pct"], b.get("scheme", "ray"),
                           stack=b.get("stack", 16))


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
    ap.add_argument("--path", action="store_true", help="also tune the path tracer (extension rays, queue capacity)")
    args = ap.parse_args()
    from . import scenes
    scene = scenes.direct_lighting_scene(level=args.level)
    cam = scenes.pinhole(args.res, args.res)
    n = args.res * args.res * args.spp_side ** 2
    r = api.SamplerRenderer(scene, cam, api.StratifiedSampler(args.spp_side, args.spp_side, 1),
                            api.DirectLightingIntegrator(4), chunk=n)
    t = time.time()
    say = lambda x: print(json.dumps(x), flush=True)  # noqa: E731
    res = sweep(r, log=say)
    r.close()
    best = {k: v[1] for k, v in res.items()}
    results = {k: v[0] for k, v in res.items()}
    if args.path:
        from . import workloads
        pr, _ = workloads.renderer("path_tracing")
        n = workloads.make("path_tracing")[4]
        results["closest"], best["closest"] = sweep_kind(pr, "closest", "ms_trace_closest", log=say, packets=False)
        results["chunk"], best["chunk"] = sweep_chunk(pr, n, log=say)
        pr.close()
    json.dump(dict(workload="direct lighting %dx%d %d spp level %d%s" % (
        args.res, args.res, args.spp_side ** 2, args.level, "; path tracing 1920x1080 64 spp" if args.path else ""),
                   results=results, best=best, seconds=time.time() - t),
              open(args.out, "w"), indent=1)
    print("best", json.dumps(best))
    return 0


if __name__ == "__main__":
    sys.exit(main())
