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
# HIP maps every stream of a process onto GPU_MAX_HW_QUEUES hardware queues
# (4 by default).  The request servers overlap uploads, traversal and
# downloads on several streams (the ray server six, the render server
# sixteen batch streams), and the renderer holds streams of its own: with 4 queues
# the ray server's compute stream can share a queue with its copy streams,
# whose waits then hold its kernels back (at 70% of the saturated rate it
# fell behind: 0.91 of the offered rate, p50 161 ms; with 8 queues 1.00, p50
# 2.9 ms), and the render server gained a few per cent more from 16 queues
# than from 8 (32 halved every server's throughput).  Read by the HIP
# runtime when it starts, so raised to 16 before the library loads (a larger
# setting is kept).
try:
    _hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
except ValueError:
    _hwq = 4
if _hwq < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3  # MI355X vector / matrix FP32
PEAK_CLOCK_HZ = 2.4e9  # MI355X peak engine clock
SIMDS = 1024  # 256 CUs x 4 SIMDs
# HBM bytes per dispatch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this
# workload, scaled by a device-copy calibration (written next to its summary)
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic_direct1024.json")
PROFILE_VALU = os.path.join(ROOT, "profiles", "valu_direct1024.json")
KERNEL_ANY = "k_trace_wide<true, false, true, 16>"
KERNEL_CLOSEST = "k_closest_packet<true, false>"  # c1: camera rays made in the packet kernel


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--spp-side", type=int, default=4)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="bounded CPU baseline sample (also the source of the per-ray work figures)")
    ap.add_argument("--tuning", default=None, help="autotuner JSON to apply (pbrt_rust_amd.autotune)")
    ap.add_argument("--no-serve", action="store_true", help="skip the request-serving measurement")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the ray-query and path-tracing measurements")
    ap.add_argument("--pt-spp-side", type=int, default=8, help="path-tracing workload: sqrt(spp)")
    ap.add_argument("--serve-pool", type=int, default=48, help="distinct requests in the serving pool")
    ap.add_argument("--serve-seconds", type=float, default=3.0,
                    help="arrival span of the fixed-rate ray-serving run (requests = rate x span)")
    ap.add_argument("--serve-load", type=float, default=0.7,
                    help="fixed-rate ray-serving load as a fraction of the batched server's saturated rate")
    ap.add_argument("--serve-passes", type=int, default=5,
                    help="timed saturated passes over the ray-serving pool per server (the median is reported)")
    ap.add_argument("--serve-batch", type=int, default=1 << 22, help="rays per serving step")
    ap.add_argument("--serve-arena-gb", type=int, default=8)
    ap.add_argument("--no-render-serve", action="store_true", help="skip the render-request serving measurement")
    ap.add_argument("--rs-small", action="store_true",
                    help="render serving without the 64M-triangle scene (16 scenes only)")
    ap.add_argument("--rs-pool", type=int, default=32, help="distinct render requests in the pool")
    ap.add_argument("--rs-requests", type=int, default=96, help="render requests in the fixed-rate run")
    ap.add_argument("--rs-batch", type=int, default=1 << 23, help="paths per render-serving batch")
    ap.add_argument("--rs-passes", type=int, default=3,
                    help="saturated passes of the batched render server (the median is reported)")
    ap.add_argument("--rs-arena-gb", type=int, default=40)
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run one device copy and one 64-byte gather over this many bytes (a power of two, "
                         "counter calibration) and exit")
    ap.add_argument("--verbose", action="store_true")
    return ap.parse_args()


def copy_calibration(nbytes):
    """one device copy of nbytes and one 64-byte gather over nbytes (2^k x
    64), for the counter calibration passes"""
    import pbrt_rust_amd as px
    from pbrt_rust_amd._lib import check
    L = px.lib()
    a = px.DeviceBuffer(nbytes)
    b = px.DeviceBuffer(nbytes)
    check(L.pbrtx_device_copy(b.ptr, a.ptr, nbytes, None))
    px.synchronize()
    nt = int(L.pbrtx_gather_probe_threads())
    out = px.DeviceBuffer(nt * 4)
    check(L.pbrtx_gather_probe(a.ptr, nbytes, out.ptr, nt, None))
    px.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "gather_bytes": nbytes}))


def measured_peaks():
    """Achievable ceilings on this GPU: a streaming device copy of 2 GiB
    (bytes read + written per second) and the fma-chain probe (non-packed
    FP32 fmas), best of a few timed runs after a warm-up."""
    import ctypes
    import pbrt_rust_amd as px
    from pbrt_rust_amd._lib import check
    L = px.lib()
    ms = ctypes.c_float()

    def timed(fn, reps):
        fn()
        px.synchronize()
        best = float("inf")
        for _ in range(reps):
            check(L.pbrtx_timer_start(None))
            fn()
            check(L.pbrtx_timer_stop(None, ctypes.byref(ms)))
            best = min(best, ms.value)
        return best

    nbytes = 1 << 31
    a, b = px.DeviceBuffer(nbytes), px.DeviceBuffer(nbytes)
    copy_ms = timed(lambda: check(L.pbrtx_device_copy(b.ptr, a.ptr, nbytes, None)), 5)
    a.free()
    b.free()
    # scattered 64-byte reads of a 2 GiB buffer, each record once (the
    # traversal's request size): the achievable rate of such reads
    g = px.DeviceBuffer(nbytes)
    gout = px.DeviceBuffer(int(L.pbrtx_gather_probe_threads()) * 4)
    gather_ms = timed(lambda: check(L.pbrtx_gather_probe(g.ptr, nbytes, gout.ptr, int(L.pbrtx_gather_probe_threads()),
                                                         None)), 5)
    g.free()
    gout.free()
    threads = int(L.pbrtx_fma_probe_threads())
    out = px.DeviceBuffer(threads * 4)
    flops = ctypes.c_uint64()
    iters = 1 << 15
    fma_ms = timed(lambda: check(L.pbrtx_fma_probe(out.ptr, threads, iters, ctypes.byref(flops), None)), 3)
    out.free()
    return {"hbm_copy_gbs": round(2 * nbytes / (copy_ms * 1e-3) / 1e9, 1),
            "gather64_gbs": round(nbytes / (gather_ms * 1e-3) / 1e9, 1),
            "fp32_fma_tflops": round(flops.value / (fma_ms * 1e-3) / 1e12, 2),
            "how": "2 GiB device copy (read + write); 2 GiB of 64-byte records read once each in a scattered "
                   "order; %d threads x 8 chains x %d fmas" % (threads, iters)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def log(args, *a):
    """progress lines on stderr (the result is the one JSON line on stdout)"""
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

    from pbrt_rust_amd import workloads
    t0 = time.time()
    scene, cam, sampler, integ, n_samples = workloads.make("direct", spp_side=args.spp_side, res=args.res,
                                                           level=args.level)
    nx, ny, seed = sampler.xs, sampler.ys, sampler.seed
    r = px.SamplerRenderer(scene, cam, sampler, integ, chunk=n_samples)
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
    # the headline frame once more with its image, outside the clock: the
    # CPU-baseline leg checks a band of its rows against the oracle
    gpu_img, _ = r.render("batched")
    band = cpu_band(scene, cam, dict(kind="direct", nx=nx, ny=ny, seed=seed, light_samples=4), args.cpu_seconds)
    gpu_samples = r.sample_radiance(band[1] * cam["xres"] * nx * ny)[band[0] * cam["xres"] * nx * ny:]
    # the headline renderer's streams and buffers are not needed past here
    r.close()
    st = sts_b[-1]
    rays = st["camera_rays"] + st["shadow_rays"]
    ms_b = float(np.mean([s["ms_total"] for s in sts_b]))
    ms_p = float(np.mean([s["ms_total"] for s in sts_p]))
    mrays_b = rays / (ms_b * 1e-3) / 1e6
    mrays_p = rays / (ms_p * 1e-3) / 1e6
    log(args, "batched %.3f ms/frame (wall %.3f)  plain %.3f ms/frame" % (ms_b, wall_b * 1e3, ms_p))

    serve = None
    if not args.no_serve:
        serve = serving(args, px, scenes, lambda m: log(args, m))
    workloads = None
    if not args.no_workloads:
        workloads = other_workloads(args, px, scenes, lambda m: log(args, m))
    render_serve = None
    if not args.no_render_serve:
        render_serve = render_serving(args, px, scenes, lambda m: log(args, m))

    # ---- CPU baseline (bounded sample of the same frame, after the GPU
    # sections so that its all-core run does not slow their host side); it
    # also yields the reference traversal's per-ray work used for the
    # algorithmic figures ----
    cpu, work, parity = cpu_baseline(scene, cam, nx, ny, seed, band, gpu_img, gpu_samples)
    del gpu_samples
    log(args, "cpu baseline %s" % json.dumps(cpu))
    log(args, "parity %s" % json.dumps(parity))

    peaks = measured_peaks()
    log(args, "measured ceilings: %s" % peaks)
    # ---- roofline of the dominant kernel (per-launch averages, HIP events on
    # the render stream) ----
    any_ms = float(np.mean([s["ms_trace_any"] / max(1, s["launches_any"]) for s in sts_b]))
    cl_ms = float(np.mean([s["ms_trace_closest"] / max(1, s["launches_closest"]) for s in sts_b]))
    dominant = "trace_any" if any_ms >= cl_ms else "trace_closest"
    # bytes touched per ray (the reference traversal's reads, counted on the
    # oracle's traversal of a band of the same frame): 32 B per node popped +
    # 36 B per triangle tested + 32 B ray in + 16 B hit out (1 bit for
    # occlusion); the launch's sum counts every ray's reads, most of them
    # served by L1, L2 and MALL, so it is NOT an HBM figure.  flops per ray
    # ~ 15 per node + 50 per triangle.
    if dominant == "trace_any":
        nodes, tris, out_b, n_rays = work["any_nodes"], work["any_tris"], 1 / 8.0, st["shadow_rays"]
        kernel_ms = any_ms
        kname = KERNEL_ANY
    else:
        nodes, tris, out_b, n_rays = work["closest_nodes"], work["closest_tris"], 16.0, st["camera_rays"]
        kernel_ms = cl_ms
        kname = KERNEL_CLOSEST
    touched_per_ray = 32 * nodes + 36 * tris + 32 + out_b
    touched = touched_per_ray * n_rays
    flops_launch = (15 * nodes + 50 * tris) * n_rays
    # HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE passes,
    # reads scaled by the 64-byte gather calibration (the traversal's request
    # size), writes by the copy calibration
    traffic = prof_t = calib = None
    if os.path.exists(PROFILE_TRAFFIC):
        try:
            tj = json.load(open(PROFILE_TRAFFIC))
            prof_t = tj["kernels"][kname]
            traffic = prof_t["hbm_bytes"]
            calib = {k: tj[k] for k in ("fetch_factor", "write_factor", "copy_fetch_factor") if k in tj}
        except (KeyError, ValueError):
            traffic = prof_t = None
    # VALU issue: SQ_INSTS_VALU of the kernel's dispatch (committed counter
    # pass) against the issue slots of the live launch time at the peak
    # engine clock (each of the 1024 SIMDs issues at most one wave64 VALU
    # instruction every 4 cycles), so the fraction cannot pass 1; lanes
    # active per VALU instruction from SQ_THREAD_CYCLES_VALU over
    # SQ_ACTIVE_INST_VALU x 64
    issue = None
    if os.path.exists(PROFILE_VALU):
        try:
            kv = json.load(open(PROFILE_VALU))["kernels"][kname]
            slots = kernel_ms * 1e-3 * PEAK_CLOCK_HZ * SIMDS / 4.0
            issue = {"valu_insts_per_launch": int(kv["valu"]), "salu_insts_per_launch": int(kv["salu"]),
                     "slots_per_launch": int(slots), "frac": round(kv["valu"] / slots, 4),
                     "lane_efficiency": None if kv.get("lane_efficiency") is None else round(kv["lane_efficiency"], 3),
                     "clock_ghz_in_profile": None if kv.get("clock_ghz") is None else round(kv["clock_ghz"], 3),
                     "l2_hit_rate": None if kv.get("l2_hit_rate") is None else round(kv["l2_hit_rate"], 3),
                     "peak_clock_ghz": PEAK_CLOCK_HZ / 1e9}
        except (KeyError, ValueError):
            issue = None
    compulsory = n_rays * (32 + out_b)
    achieved = None if traffic is None else traffic / (kernel_ms * 1e-3) / 1e9
    hbm_frac = None if achieved is None else achieved / HBM_PEAK_GBS
    binding = None
    if hbm_frac is not None and issue is not None:
        binding = "valu_issue" if issue["frac"] >= hbm_frac else "hbm"

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
            "kernel_symbol": kname,
            "achieved": None if achieved is None else round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": None if hbm_frac is None else round(hbm_frac, 4),
            "traffic": None if traffic is None else int(traffic),
            "traffic_source": "rocprofv3 FETCH_SIZE x 64-byte-gather calibration + WRITE_SIZE x copy calibration, "
                              "per dispatch (profiles/traffic_direct1024.md)",
            "calibration": calib if traffic is not None else None,
            "kernel_ms_per_launch": round(kernel_ms, 4),
            "rays_per_launch": int(n_rays),
            "binding": binding,
            "issue": issue,
            "compulsory_bytes_per_launch": int(compulsory),
            "bytes_touched": {"per_ray": round(touched_per_ray, 2), "per_launch": int(touched),
                              "gb_per_s": round(touched / (kernel_ms * 1e-3) / 1e9, 1),
                              "note": "the reference traversal's reads summed over the launch's rays, served by "
                                      "L1/L2/MALL as well as HBM: not an HBM rate"},
            "valu_flops": {"per_launch": int(flops_launch),
                           "tflops": round(flops_launch / (kernel_ms * 1e-3) / 1e12, 3),
                           "frac_of_fp32_peak": round(flops_launch / (kernel_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4)},
            "measured_peaks": peaks,
            "frac_of_measured_gather": None if achieved is None or not peaks.get("gather64_gbs") else
            round(achieved / peaks["gather64_gbs"], 4),
            "work_per_ray": {k: round(v, 2) for k, v in work.items()},
        },
        "serving": serve,
        "render_serving": render_serve,
        "workloads": workloads,
        "cpu_baseline": cpu,
        "parity": parity,
        "device": px.device_name(),
    }
    print(json.dumps(out), flush=True)


def frame_rate(px, r, path, steps, warmup):
    """ms per frame (HIP events on the render stream) and rays per frame"""
    for _ in range(warmup):
        r.render(path, want_image=False)
    sts = [r.render(path, want_image=False)[1] for _ in range(steps)]
    st = sts[-1]
    rays = st["camera_rays"] + st["shadow_rays"] + st["extension_rays"]
    ms = float(np.mean([s["ms_total"] for s in sts]))
    return ms, rays, st


def other_workloads(args, px, scenes, log_fn):
    """The ray-query workload (16,384 random triangles, 512x512 pinhole, one
    jittered sample, closest hit then one shadow segment to the point light
    at (2,2,-2)) and the path-tracing workload (64 level-6 icospheres, 5.2M
    triangles, 70% Lambertian / 30% mirror, quad light, 1920x1080, 64 spp,
    depth 5, roulette after bounce 3), batched and plain."""
    from pbrt_rust_amd import workloads
    out = {}
    r, sc = workloads.renderer("ray_queries")
    mse, rays, st = frame_rate(px, r, "batched", 20, 3)
    msp, _, _ = frame_rate(px, r, "plain", 20, 3)
    # the small frame is launch-bound: the batched frames replayed from a
    # captured graph
    r.set_graph(True)
    msb, _, _ = frame_rate(px, r, "batched", 20, 3)
    r.set_graph(False)
    r.close()
    out["ray_queries"] = {"workload": "16,384 random triangles, 512x512, 1 spp, closest hit + point-light shadow rays",
                          "rays_per_frame": int(rays), "camera_rays": int(st["camera_rays"]),
                          "shadow_rays": int(st["shadow_rays"]),
                          "batched_ms": round(msb, 4), "batched_eager_ms": round(mse, 4), "plain_ms": round(msp, 4),
                          "batched_mrays_per_s": round(rays / msb / 1e3, 1),
                          "plain_mrays_per_s": round(rays / msp / 1e3, 1), "batched_over_plain": round(msp / msb, 3)}
    log_fn("ray queries %s" % json.dumps(out["ray_queries"]))
    side = args.pt_spp_side
    r, sc = workloads.renderer("path_tracing", spp_side=side)
    msb, rays, st = frame_rate(px, r, "batched", 2, 1)
    msp, _, _ = frame_rate(px, r, "plain", 1, 1)
    r.close()
    out["path_tracing"] = {"workload": "64 level-6 icospheres (%d tris), 1920x1080, %d spp, depth 5, roulette after "
                                       "bounce 3" % (int(sc["idx"].shape[0]), side * side),
                           "rays_per_frame": int(rays), "camera_rays": int(st["camera_rays"]),
                           "shadow_rays": int(st["shadow_rays"]), "extension_rays": int(st["extension_rays"]),
                           "batched_ms": round(msb, 3), "plain_ms": round(msp, 3),
                           "batched_mrays_per_s": round(rays / msb / 1e3, 1),
                           "plain_mrays_per_s": round(rays / msp / 1e3, 1), "batched_over_plain": round(msp / msb, 3)}
    log_fn("path tracing %s" % json.dumps(out["path_tracing"]))
    return out


def serving(args, px, scenes, log_fn):
    """Ray-query serving: 8 resident scenes of random triangles (10K-2M,
    log-uniform), requests of 1K-4M rays (log-uniform), 70% closest hit / 30%
    shadow segments, origins uniform in the scene box and directions uniform
    on the sphere.  Saturated throughput with every request queued at once
    (the median of several passes; the first, cold pass runs about a fifth
    slower), then p50/p99 latency under Poisson arrivals at 70% of the
    batched server's saturated rate (--serve-load) over --serve-seconds of
    arrivals, for the continuous-batching server and for the
    one-plain-kernel-per-request server at the same rate; achieved = rays
    served / time until the last request is done, beside the offered rate."""
    import threading
    rng = np.random.default_rng(0x5E4E)
    counts = np.exp(rng.uniform(np.log(1e4), np.log(2e6), 8)).astype(np.int64)
    bvhs = []
    for k, c in enumerate(counts):
        v, idx = scenes.random_triangles(int(c), 0xC3 + k)
        bvhs.append(px.BVHAccelerator([px.Primitive.simple(px.TriangleMesh(None, None, False, idx, v))], 4, "sah"))
    # request pool in page-locked memory
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(4e6), args.serve_pool)).astype(np.int64)
    pool = []
    for j, n in enumerate(sizes):
        s = int(rng.integers(0, len(bvhs)))
        kind = "closest" if rng.random() < 0.7 else "any"
        lo, hi = bvhs[s].world_bound()
        r = px.PinnedArray((int(n),), px.RAY_DTYPE)
        a = r.array
        o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
        d = rng.normal(size=(n, 3))
        d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        a["ox"], a["oy"], a["oz"] = o[:, 0], o[:, 1], o[:, 2]
        a["dx"], a["dy"], a["dz"] = d[:, 0], d[:, 1], d[:, 2]
        a["tmin"] = 0.0
        diag = float(np.linalg.norm(hi - lo))
        a["tmax"] = np.finfo(np.float32).max if kind == "closest" else \
            rng.uniform(0.0, 0.5 * diag, n).astype(np.float32)
        out = px.PinnedArray((int(n),), px.HIT_DTYPE) if kind == "closest" else \
            px.PinnedArray((int((n + 31) // 32),), np.uint32)
        pool.append((s, kind, r, out))
    mean_rays = float(np.mean(sizes))

    def make(mode):
        srv = px.RayServer(arena_bytes=args.serve_arena_gb << 30, batch_rays=args.serve_batch, mode=mode)
        for b in bvhs:
            srv.add_scene(b)
        for s in range(len(bvhs)):  # first launch per scene (scratch sizing) outside the clock
            w = pool[0][2].array[:64].copy()
            srv.wait(srv.submit(s, "closest", w))
        return srv

    def saturated(mode):
        # every request of the pool queued at once, waited for, timed; the
        # median over passes (one pass is some tens of milliseconds)
        srv = make(mode)
        rates = []
        for _ in range(max(1, args.serve_passes)):
            t = time.perf_counter()
            tick = [srv.submit(s, kind, r.array, out.array) for s, kind, r, out in pool]
            for tk in tick:
                srv.wait(tk)
            dt = time.perf_counter() - t
            rates.append(float(sizes.sum()) / dt / 1e6)
        srv.close()
        return float(np.median(rates)), [round(x, 1) for x in rates]

    def at_rate(mode, req_per_s, nreq):
        srv = make(mode)
        arr = np.cumsum(np.random.default_rng(7).exponential(1.0 / req_per_s, nreq))
        tickets = [None] * nreq
        ready = threading.Semaphore(0)

        def producer():
            t0 = time.perf_counter()
            for i in range(nreq):
                while True:
                    lag = arr[i] - (time.perf_counter() - t0)
                    if lag <= 0:
                        break
                    if lag > 2e-3:
                        time.sleep(lag - 1e-3)
                s, kind, r, out = pool[i % len(pool)]
                tickets[i] = srv.submit(s, kind, r.array, out.array)
                ready.release()

        th = threading.Thread(target=producer)
        c0 = srv.counters()
        t = time.perf_counter()
        th.start()
        lat = []
        for i in range(nreq):
            ready.acquire()
            _, st = srv.wait(tickets[i])
            lat.append(st["latency_ms"])
        th.join()
        dt = time.perf_counter() - t
        c1 = srv.counters()
        rays = sum(int(sizes[i % len(pool)]) for i in range(nreq))
        srv.close()
        lat = np.array(lat)
        k = max(1, nreq // 10)
        return {"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_ms": round(float(lat.mean()), 3), "achieved_mrays_per_s": round(rays / dt / 1e6, 2),
                "gpu_busy_share": round((c1["busy_ms"] - c0["busy_ms"]) / (dt * 1e3), 3),
                "steps": int(c1["steps"] - c0["steps"]), "launches": int(c1["launches"] - c0["launches"]),
                "p50_ms_by_arrival_decile": [round(float(np.median(lat[i:i + k])), 2) for i in range(0, nreq, k)][:10]}

    sat_b, passes_b = saturated("batched")
    sat_p, passes_p = saturated("plain")
    log_fn("serving saturated: batched %.1f Mrays/s, plain %.1f Mrays/s" % (sat_b, sat_p))
    rate = args.serve_load * sat_b
    req_per_s = rate * 1e6 / mean_rays
    nreq = max(len(pool), int(round(args.serve_seconds * req_per_s)))
    lat_b = at_rate("batched", req_per_s, nreq)
    lat_p = at_rate("plain", req_per_s, nreq)
    for lat in (lat_b, lat_p):
        lat["achieved_over_offered"] = round(lat["achieved_mrays_per_s"] / rate, 3)
    log_fn("serving at %.1f Mrays/s: batched %s plain %s" % (rate, lat_b, lat_p))
    return {
        "workload": "8 resident scenes of random triangles (%d tris total), %d-request pool of 1K-4M rays "
                    "(log-uniform), 70%% closest / 30%% shadow, random origins and directions, host page-locked "
                    "buffers (latency includes both copies)" % (int(counts.sum()), len(pool)),
        "saturated_mrays_per_s": {"batched": round(sat_b, 2), "plain": round(sat_p, 2)},
        "saturated_passes_mrays_per_s": {"batched": passes_b, "plain": passes_p},
        "served_over_plain": round(sat_b / sat_p, 3),
        "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
        "rate_mrays_per_s": round(rate, 2),
        "rate_requests_per_s": round(req_per_s, 2),
        "requests": nreq,
        "latency": {"batched": lat_b, "plain": lat_p},
    }


def render_serving(args, px, scenes, log_fn):
    """Render-request serving: 16 resident scenes of 15 level-6 displaced
    icospheres each (1.23M triangles, 19.7M in all; host SAH trees) plus one
    64M-triangle scene of 782 level-6 displaced icospheres whose tree is built
    on the GPU at admission (left out with --rs-small);
    requests of square images with side log-uniform in 128..2048, spp from
    {1, 4, 16, 64}, path depth 5 with roulette after bounce 3, a look-at
    camera on a sphere around the scene.  Saturated throughput with every
    request queued at once, then p50/p99 latency under Poisson arrivals at 70%
    of the batched server's saturated request rate, for the batched server
    (progressive tiles of the requests of a scene in shared wavefront
    batches) and for the plain one (one request at a time on the plain path).
    Latency includes the film and its copy to page-locked host memory."""
    import concurrent.futures as cf
    import threading
    t0 = time.time()
    integ = px.PathIntegrator(5, 3)
    srv = px.RenderServer(arena_bytes=args.rs_arena_gb << 30, batch_samples=args.rs_batch, mode="batched")
    defs = [(15, 10 + k, 10.0) for k in range(16)]
    if not args.rs_small:
        defs.append((782, 99, 40.0))
    tris = [0] * len(defs)

    admit_s = [0.0] * len(defs)

    def add(k):
        # the 64M-triangle scene's tree is built on the GPU (linear BVH), the
        # others' on the host (the reference's SAH build)
        n, seed, ext = defs[k]
        sc = scenes.sphere_field(n, seed=seed, extent=ext)
        tris[k] = int(sc["idx"].shape[0])
        ta = time.perf_counter()
        sid = srv.add_scene(sc, integ, split="lbvh" if k == 16 else "sah")
        admit_s[k] = time.perf_counter() - ta
        log_fn("render serving: scene %d (%d triangles) resident after %.1fs (add_scene %.2fs)"
               % (sid, tris[k], time.time() - t0, admit_s[k]))
        return sid

    scenes.displaced_pool()
    with cf.ThreadPoolExecutor(4) as ex:  # the host builds release the interpreter lock
        ids = list(ex.map(add, range(16)))
    if not args.rs_small:
        ids.append(add(16))
    sid = dict(zip(range(len(defs)), ids))  # server ids come in completion order
    rng = np.random.default_rng(0xC4)
    pool = []
    for j in range(args.rs_pool):
        s = int(rng.integers(0, len(defs)))
        side = int(round(np.exp(rng.uniform(np.log(128), np.log(2048)))))
        k = int(rng.choice([1, 2, 4, 8]))
        u = rng.normal(size=3)
        eye = u / np.linalg.norm(u) * defs[s][2] * 3.2
        cam = scenes.look_at(side, side, eye, (0.0, 0.0, 0.0), fov=60.0)
        out = px.PinnedArray((side, side, 3), np.float32)
        pool.append((sid[s], cam, px.StratifiedSampler(k, k, 1000 + j), out))
    samples = [p[1]["xres"] * p[1]["yres"] * p[2].xs * p[2].ys for p in pool]
    log_fn("render serving: %d scenes, %d-request pool (%.1fM samples) ready in %.1fs"
           % (len(defs), len(pool), sum(samples) / 1e6, time.time() - t0))

    def rays_of(c):
        return c["camera_rays"] + c["shadow_rays"] + c["extension_rays"]

    # first batch per scene (trace scratch sizing) outside the clock
    warm = scenes.look_at(64, 64, (0.0, 0.0, -40.0), (0.0, 0.0, 0.0))
    for s in ids:
        srv.wait(srv.submit(s, warm, px.StratifiedSampler(1, 1, 0)))

    def saturated(mode, passes):
        # every request of the pool queued at once; the median of `passes`
        # passes (one pass of the batched server is about half a second, and
        # its end waits on the last request's tail)
        srv.set_mode(mode)
        rates, rps = [], []
        for _ in range(passes):
            c0 = srv.counters()
            t = time.perf_counter()
            tick = [srv.submit(s, cam, smp, out.array) for s, cam, smp, out in pool]
            for tk in tick:
                srv.wait(tk)
            dt = time.perf_counter() - t
            c1 = srv.counters()
            rates.append((rays_of(c1) - rays_of(c0)) / dt / 1e6)
            rps.append(len(pool) / dt)
        k = int(np.argsort(rates)[len(rates) // 2])
        return rates[k], rps[k], [round(x, 1) for x in rates]

    def at_rate(mode, req_per_s, nreq):
        srv.set_mode(mode)
        arr = np.cumsum(np.random.default_rng(11).exponential(1.0 / req_per_s, nreq))
        tickets = [None] * nreq
        ready = threading.Semaphore(0)

        def producer():
            t0_ = time.perf_counter()
            for i in range(nreq):
                while True:
                    lag = arr[i] - (time.perf_counter() - t0_)
                    if lag <= 0:
                        break
                    if lag > 2e-3:
                        time.sleep(lag - 1e-3)
                s, cam, smp, out = pool[i % len(pool)]
                tickets[i] = srv.submit(s, cam, smp, out.array)
                ready.release()

        c0 = srv.counters()
        th = threading.Thread(target=producer)
        t = time.perf_counter()
        th.start()
        lat = []
        for i in range(nreq):
            ready.acquire()
            _, st = srv.wait(tickets[i])
            lat.append(st["latency_ms"])
        th.join()
        dt = time.perf_counter() - t
        c1 = srv.counters()
        lat = np.array(lat)
        return {"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_ms": round(float(lat.mean()), 3),
                "achieved_mrays_per_s": round((rays_of(c1) - rays_of(c0)) / dt / 1e6, 2)}

    sat_b, rps_b, passes_b = saturated("batched", args.rs_passes)
    sat_p, rps_p, passes_p = saturated("plain", 1)
    log_fn("render serving saturated: batched %.1f Mrays/s (%.2f req/s), plain %.1f Mrays/s (%.2f req/s)"
           % (sat_b, rps_b, sat_p, rps_p))
    req_per_s = 0.7 * rps_b
    lat_b = at_rate("batched", req_per_s, args.rs_requests)
    lat_p = at_rate("plain", req_per_s, args.rs_requests)
    log_fn("render serving at %.2f req/s: batched %s plain %s" % (req_per_s, lat_b, lat_p))
    c = srv.counters()
    srv.close()
    for p in pool:
        p[3].free()
    return {
        "workload": "%d resident scenes of level-6 displaced icospheres (%d triangles in all, largest %d), "
                    "%d-request pool: square images of side 128-2048 (log-uniform), spp from {1,4,16,64}, "
                    "path depth 5, roulette after bounce 3, look-at cameras; films copied to page-locked host "
                    "memory (latency includes the copy)" % (len(defs), sum(tris), max(tris), len(pool)),
        "saturated_mrays_per_s": {"batched": round(sat_b, 2), "plain": round(sat_p, 2)},
        "saturated_requests_per_s": {"batched": round(rps_b, 3), "plain": round(rps_p, 3)},
        "saturated_passes_mrays_per_s": {"batched": passes_b, "plain": passes_p},
        "served_over_plain": round(sat_b / sat_p, 3),
        "rate_requests_per_s": round(req_per_s, 3),
        "requests": args.rs_requests,
        "latency": {"batched": lat_b, "plain": lat_p},
        "batch_paths": args.rs_batch,
        "scene_bytes": int(c["scene_bytes"]),
        "add_scene_s": {"small_scenes_max": round(max(admit_s[:16]), 3),
                        "large_scene_gpu_tree": None if args.rs_small else round(admit_s[16], 3),
                        "note": "wall time of add_scene: tree build, uploads and the shading tables; the large "
                                "scene's tree is built on the GPU (linear BVH), the small ones' on the host (SAH, "
                                "four host builds at a time)"},
    }


def cpu_cores():
    """the host share given to this process (the pool sets
    OMP_NUM_THREADS), else every core"""
    return int(os.environ.get("OMP_NUM_THREADS", 0)) or os.cpu_count() or 1


def cpu_band(scene, cam, integ, seconds):
    """The rows [y0, y1) of the frame the CPU baseline renders: from the
    middle row down, as many as fit the time budget (timed on two rows, the
    scene build excluded)."""
    ol = _oracle()
    y0 = cam["yres"] // 2
    t = time.perf_counter()
    res = ol.render(scene, cam, integ, y0=y0, y1=y0 + 2, nthreads=cpu_cores())
    first = time.perf_counter() - t
    build_s = res["stats"][3] * 1e-6
    per_row = max(1e-4, (first - build_s) / 2)
    rows = int(max(2, min(cam["yres"] - y0, (seconds - build_s) / per_row)))
    return y0, y0 + rows


def cpu_baseline(scene, cam, nx, ny, seed, band, gpu_img, gpu_samples):
    """The oracle's render (CPU restatement of the same path) timed on the
    host cores over a bounded band of image rows of the same frame, and the
    GPU frame's same rows checked against it bit for bit: every sample's
    radiance and every film pixel.  Also returns the reference traversal's
    work per ray on that band (nodes popped and prims tested; occlusion
    queries counted up to their first hit), the per-unit figures the roofline
    uses."""
    ol = _oracle()
    cores = cpu_cores()
    integ = dict(kind="direct", nx=nx, ny=ny, seed=seed, light_samples=4)
    y0, y1 = band
    res = ol.render(scene, cam, integ, y0=y0, y1=y1, nthreads=cores, want_samples=True)
    st = [int(v) for v in res["stats"]]
    rows = y1 - y0
    spp = nx * ny
    want_s = res["samples"].reshape(-1, 3)
    got_s = gpu_samples.reshape(-1, 3)
    bad_s = int(np.count_nonzero(np.any(want_s.view(np.uint32) != got_s.view(np.uint32), axis=1)))
    got_f = gpu_img[y0:y1]
    bad_f = int(np.count_nonzero(np.any(res["rgb"].view(np.uint32) != got_f.view(np.uint32), axis=2)))
    parity = {"rows": [y0, y1], "samples": int(want_s.shape[0]), "pixels": int(rows * cam["xres"]),
              "mismatches": bad_s + bad_f, "sample_mismatches": bad_s, "pixel_mismatches": bad_f,
              "rmse": float(np.sqrt(np.mean((res["rgb"].astype(np.float64) - got_f.astype(np.float64)) ** 2))),
              "against": "oracle render of the same rows (per-sample radiance and film, bit for bit)"}
    assert want_s.shape[0] == rows * cam["xres"] * spp
    trace_s = st[4] * 1e-9
    rays = st[0] + st[1]
    # one thread on two rows, for the per-core rate
    one = ol.render(scene, cam, integ, y0=y0, y1=y0 + 2, nthreads=1)
    s1 = [int(v) for v in one["stats"]]
    cpu = {"value": round(rays / trace_s / 1e6, 4), "unit": "Mrays/s", "cores": cores, "kind": "port",
           "sample": "oracle render of rows %d..%d of the same frame (%d camera + %d shadow rays, "
                     "shading included), BVH build excluded" % (y0, y1, st[0], st[1]),
           "one_thread_mrays_per_s": round((s1[0] + s1[1]) / (s1[4] * 1e-9) / 1e6, 4),
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}
    work = {"closest_nodes": st[5] / max(1, st[0]), "closest_tris": st[6] / max(1, st[0]),
            "any_nodes": st[7] / max(1, st[1]), "any_tris": st[8] / max(1, st[1])}
    return cpu, work, parity


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    return oracle_lib


if __name__ == "__main__":
    main()
