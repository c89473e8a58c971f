"""Benchmark: Mrays/s of the batched (wavefront) HIP path vs the plain
one-kernel-per-query path on the direct-lighting render workload.

Workload (one step = one frame, all work on the GPU, inputs resident in HBM):
  1024x1024 pixels, stratified 4x4 = 16 spp, SamplerRenderer + direct-lighting
  integrator with 4 light samples per hit; level-8 icosphere (1,310,720
  triangles) displaced by 4-octave fBm (amplitude 0.05), 2-triangle ground
  plane, 2-triangle quad area light of radiance 10; Lambertian albedo 0.5/0.8.
  Rays per frame = camera rays + shadow rays actually traced.

Prints one JSON line (see the keys below).  Only N=1 exists: the path is
single-GPU by design (rays are independent; multi-GPU sharding is described in
the project notes, not run).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "trace_any_direct1024_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--spp-side", type=int, default=4)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bounded CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tuning", default=None, help="autotuner JSON to apply (pbrt_rust_amd.autotune)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run one device copy of this many bytes (counter calibration) and exit")
    ap.add_argument("--verbose", action="store_true")
    return ap.parse_args()


def copy_calibration(nbytes):
    import pbrt_rust_amd as px
    from pbrt_rust_amd._lib import check
    a = px.DeviceBuffer(nbytes)
    b = px.DeviceBuffer(nbytes)
    check(px.lib().pbrtx_device_copy(b.ptr, a.ptr, nbytes, None))
    px.synchronize()
    print(json.dumps({"copy_bytes": nbytes}))


def log(args, *a):
    if args.verbose:
        print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (N=1); multi-GPU runs are not part of it.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_bytes:
        copy_calibration(args.copy_bytes)
        return
    import pbrt_rust_amd as px
    from pbrt_rust_amd import autotune, scenes
    if args.tuning:
        autotune.load(args.tuning)

    t0 = time.time()
    scene = scenes.direct_lighting_scene(level=args.level)
    cam = scenes.pinhole(args.res, args.res)
    nx = ny = args.spp_side
    seed = 1
    integ = px.DirectLightingIntegrator(4)
    n_samples = args.res * args.res * nx * ny
    r = px.SamplerRenderer(scene, cam, px.StratifiedSampler(nx, ny, seed), integ, chunk=n_samples)
    log(args, "scene + BVH ready in %.1fs" % (time.time() - t0))

    def run(path, steps, warmup):
        for _ in range(warmup):
            r.render(path, want_image=False)
        px.synchronize()
        sts = []
        t = time.perf_counter()
        for _ in range(steps):
            _, st = r.render(path, want_image=False)
            sts.append(st)
        px.synchronize()
        wall = (time.perf_counter() - t) / steps
        return wall, sts

    wall_b, sts_b = run("batched", args.steps, args.warmup)
    wall_p, sts_p = run("plain", args.plain_steps, 1)
    st = sts_b[-1]
    rays = st["camera_rays"] + st["shadow_rays"]
    ms_b = float(np.mean([s["ms_total"] for s in sts_b]))
    ms_p = float(np.mean([s["ms_total"] for s in sts_p]))
    mrays_b = rays / (ms_b * 1e-3) / 1e6
    mrays_p = rays / (ms_p * 1e-3) / 1e6
    log(args, "batched %.3f ms/frame (wall %.3f)  plain %.3f ms/frame" % (ms_b, wall_b * 1e3, ms_p))

    # ---- roofline of the dominant kernel (per-launch averages, HIP events) ----
    any_ms = float(np.mean([s["ms_trace_any"] / max(1, s["launches_any"]) for s in sts_b]))
    cl_ms = float(np.mean([s["ms_trace_closest"] / max(1, s["launches_closest"]) for s in sts_b]))
    dominant = "trace_any" if any_ms >= cl_ms else "trace_closest"
    # algorithmic bytes per ray from the instrumented oracle traversal on a
    # sample of the same rays: 32 B per node visited + 36 B per triangle
    # tested + 32 B ray in + 16 B hit out (4 B bit for occlusion)
    alg = algorithmic_bytes(scene, cam, nx, ny, seed, r, px)
    if dominant == "trace_any":
        per_launch_bytes = alg["any_bytes_per_ray"] * st["shadow_rays"]
        kernel_ms = any_ms
    else:
        per_launch_bytes = alg["closest_bytes_per_ray"] * st["camera_rays"]
        kernel_ms = cl_ms
    achieved = per_launch_bytes / (kernel_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        try:
            traffic = json.load(open(PROFILE_TRAFFIC)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene, cam, nx, ny, seed, args.cpu_seconds)

    out = {
        "metric": "Mrays/s batched vs plain GPU kernel; p50/p99 request latency at fixed rate",
        "value": round(mrays_b, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_b, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": {
            "workload": "direct-lighting render %dx%d, %d spp stratified, 4 light samples, level-%d "
                        "displaced icosphere + ground + quad area light" % (args.res, args.res, nx * ny, args.level),
            "triangles": int(scene["idx"].shape[0]),
            "camera_rays": int(st["camera_rays"]),
            "shadow_rays": int(st["shadow_rays"]),
            "path": "batched",
        },
        "plain_mrays_per_s": round(mrays_p, 2),
        "batched_over_plain": round(mrays_b / mrays_p, 3),
        "wall_ms_per_step": round(wall_b * 1e3, 4),
        "stage_ms": {"trace_closest": round(float(np.mean([s["ms_trace_closest"] for s in sts_b])), 4),
                     "trace_any": round(float(np.mean([s["ms_trace_any"] for s in sts_b])), 4),
                     "other": round(float(np.mean([s["ms_other"] for s in sts_b])), 4)},
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "alg_bytes_per_launch": int(per_launch_bytes),
            "kernel_ms_per_launch": round(kernel_ms, 4),
            "nodes_per_ray": alg["nodes_per_ray"],
        },
        "cpu_baseline": cpu,
        "device": px.device_name(),
    }
    print(json.dumps(out), flush=True)
    r.close()


def algorithmic_bytes(scene, cam, nx, ny, seed, renderer, px, n_pixels=4096):
    """Mean reference-traversal work per ray on a sample of the frame's rays,
    counted by the oracle (tests/ helpers are test infrastructure; the oracle
    here is only the counter, never what is measured)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    L = ol.lib()
    o = ol.OracleScene()
    o.add_mesh(scene["verts"], scene["idx"])
    o.build("sah", 1)
    rng = np.random.default_rng(0)
    xres, yres = cam["xres"], cam["yres"]
    pix = rng.choice(xres * yres, n_pixels, replace=False)
    r2c = np.zeros(16, np.float32)
    L.orc_camera_raster_to_camera(xres, yres, ol.f32(cam["screen"]), cam["fov"], r2c)
    xy = np.zeros((n_pixels, 2), np.float32)
    buf = np.zeros(2 * nx * ny, np.float32)
    for k, p in enumerate(pix):
        L.orc_pixel_samples(seed, xres, int(p % xres), int(p // xres), nx, ny, 1, buf)
        xy[k] = buf[:2]
    oo = np.zeros(3 * n_pixels, np.float32)
    dd = np.zeros(3 * n_pixels, np.float32)
    L.orc_camera_rays(r2c, ol.f32(cam["cam_to_world"]).reshape(16), n_pixels, xy.reshape(-1), oo, dd)
    res = o.intersect(oo, dd, 0.0, np.finfo(np.float32).max, nthreads=os.cpu_count() or 1)
    n_cl, t_cl = res["stats"][0] / n_pixels, res["stats"][1] / n_pixels
    # shadow segments toward the light's centre from every hit
    hit = res["prim"] != 0xFFFFFFFF
    t = res["t"][hit]
    p = (oo.reshape(-1, 3)[hit] + dd.reshape(-1, 3)[hit] * t[:, None]).astype(np.float32)
    light_c = scene["verts"][scene["idx"][scene["face_light"] >= 0].reshape(-1)].mean(0).astype(np.float32)
    w = light_c - p
    dist = np.sqrt((w * w).sum(1)).astype(np.float32)
    dirn = (w / dist[:, None]).astype(np.float32)
    occ, st = o.intersect_p(p, dirn, (t * np.float32(5e-4)).astype(np.float32), (dist * np.float32(0.999)),
                            nthreads=os.cpu_count() or 1)
    n_sh = max(1, p.shape[0])
    n_an, t_an = st[0] / n_sh, st[1] / n_sh
    return {"closest_bytes_per_ray": 32 * n_cl + 36 * t_cl + 32 + 16,
            "any_bytes_per_ray": 32 * n_an + 36 * t_an + 32 + 4 / 32.0,
            "nodes_per_ray": {"closest": round(float(n_cl), 2), "any": round(float(n_an), 2),
                              "tris_closest": round(float(t_cl), 2), "tris_any": round(float(t_an), 2)}}


def cpu_baseline(scene, cam, nx, ny, seed, seconds):
    """The oracle's render (CPU restatement of the same path) timed on the
    host cores over a bounded band of image rows."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    cores = os.cpu_count() or 1
    integ = dict(kind="direct", nx=nx, ny=ny, seed=seed, light_samples=4)
    # calibrate: 2 rows first, then as many rows as fit the budget
    y0 = cam["yres"] // 2
    t = time.perf_counter()
    res = ol.render(scene, cam, integ, y0=y0, y1=y0 + 2, nthreads=cores)
    first = time.perf_counter() - t
    build_s = res["stats"][3] * 1e-6
    per_row = max(1e-4, (first - build_s) / 2)
    rows = int(max(2, min(cam["yres"] - y0, (seconds - build_s) / per_row)))
    res = ol.render(scene, cam, integ, y0=y0, y1=y0 + rows, nthreads=cores)
    st = res["stats"]
    trace_s = st[4] * 1e-9
    rays = int(st[0] + st[1])
    return {"value": round(rays / trace_s / 1e6, 4), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": "oracle render of rows %d..%d of the same frame (%d camera + %d shadow rays), "
                      "BVH build excluded" % (y0, y0 + rows, int(st[0]), int(st[1]))}


if __name__ == "__main__":
    main()
