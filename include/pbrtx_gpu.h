/*
 * pbrtx_gpu.h -- C ABI of the MI355X (gfx950) ray-query and sampling path.
 *
 * Drop-in boundary for pbrt_rust's aggregate / renderer interfaces:
 *   Intersectable::{intersect, intersect_p}   src/intersection.rs:61-66
 *   BVHAccelerator::{new, intersect}          src/primitive/aggregates/bvh.rs:335-422
 *   Aggregate dispatch                        src/primitive/aggregates/mod.rs:47-63
 *   Renderer::render / SamplerRenderer        src/renderer.rs:8-26, src/sampler_renderer.rs:146-206
 * Every function returns 0 (PBRTX_OK) or a negative PBRTX_ERR_* code and never
 * aborts; pbrtx_last_error() gives the message of the calling thread's last
 * failure.  Pointers named d_* are device (HBM) pointers; `stream` is a
 * hipStream_t (NULL = the default stream); queries are asynchronous on it.
 */
#ifndef PBRTX_GPU_H
#define PBRTX_GPU_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define PBRTX_OK 0
#define PBRTX_ERR_ARG (-1)
#define PBRTX_ERR_HIP (-2)
#define PBRTX_ERR_OOM (-3)
#define PBRTX_ERR_BUILD (-4)
#define PBRTX_ERR_STATE (-5)

/* traversal paths */
#define PBRTX_PATH_PLAIN 0   /* one lane per ray, reference node layout */
#define PBRTX_PATH_BATCHED 1 /* persistent waves, four-wide records */

/* Ray record (32 B): Ray{o, d, mint, maxt} of src/ray.rs:7-61 without the
   time / depth fields the queries do not read. */
typedef struct pbrtx_ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} pbrtx_ray;

/* Hit record (16 B): the closest hit's primitive, t and barycentrics
   (Intersection of src/intersection.rs:14-23 reduced to what shading needs;
   dg is rebuilt from these).  prim = input prim id (mesh faces in face order,
   then spheres) or 0xFFFFFFFF on a miss. */
typedef struct pbrtx_hit {
    uint32_t prim;
    float t, b1, b2;
} pbrtx_hit;

/* Sphere primitive (Shape::sphere, src/shape/sphere.rs:28-44). */
typedef struct pbrtx_sphere {
    float o2w[16];     /* object-to-world Transform m, row-major */
    float o2w_inv[16]; /* its m_inv */
    float radius, z0, z1, phimax_deg;
    int32_t reverse_orientation;
} pbrtx_sphere;

typedef struct pbrtx_bvh pbrtx_bvh;     /* host-side build result */
typedef struct pbrtx_scene pbrtx_scene; /* device-resident BVH + primitives */

const char* pbrtx_last_error(void);

/* ---- BVH build: BVHAccelerator::new (bvh.rs:335-362) --------------------- */
/* Builds the reference's tree on the host over mesh faces (world-space
   vertices, index triples) followed by spheres.  split: 0 "sah", 1 "middle",
   2 "equal" (bvh.rs:341-349).  Faces become primitives in face order
   (Primitive::fully_refine + Mesh::refine, primitive/mod.rs:47-62,
   mesh.rs:324-338). */
int pbrtx_bvh_build(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t nfaces,
                    const pbrtx_sphere* spheres, uint32_t nspheres, int32_t split,
                    uint32_t max_prims, pbrtx_bvh** out);
/* node count, deepest level, prim count, host build time */
int pbrtx_bvh_info(const pbrtx_bvh* bvh, uint32_t* n_nodes, uint32_t* depth, uint32_t* n_prims,
                   double* build_ms);
/* the flattened PackedBVHNode array (bvh.rs:260-326): bounds[6n], off[n]
   (second child or prim offset), meta[n] (axis or prim count), leaf[n], and
   the BVH-ordered prim ids order[n_prims] */
int pbrtx_bvh_export(const pbrtx_bvh* bvh, float* bounds, uint32_t* off, uint32_t* meta,
                     uint8_t* leaf, uint32_t* order);
/* The device layouts the build produced, for inspection: counts of the
   four-wide records (128 B each) and of the compressed records (64 B each,
   0 when the scene has none), then copies of them and of the compressed
   layout's 64-byte prim records (any pointer may be NULL). */
int pbrtx_bvh_layout_info(const pbrtx_bvh* bvh, uint32_t* n_wide, uint32_t* n_compressed, uint32_t* n_prims);
int pbrtx_bvh_export_layouts(const pbrtx_bvh* bvh, void* wide, void* compressed, void* prims64);
/* The compressed occlusion records (64 B each, root first; leaves referenced
   as in the compressed layout) that any-hit queries traverse: *n receives
   their count (0 when the scene has none and any-hit queries use the
   compressed layout) and *depth their record levels; records may be NULL.
   No counterpart in the reference, whose any-hit query is the closest-hit
   walk of the same binary tree (intersection.rs:62-64, bvh.rs:377-421).
   Their tree is the builder's own over that tree's leaves, which reaches
   the same leaves.  For inspection and tests. */
int pbrtx_bvh_occlusion_records(const pbrtx_bvh* bvh, uint32_t* n, uint32_t* depth, void* records);
int pbrtx_bvh_free(pbrtx_bvh* bvh);

/* ---- device scene -------------------------------------------------------- */
/* A scene whose tree is built on the GPU: a linear BVH (Morton-ordered
   primitives, Karras radix tree) over mesh faces then spheres, laid out for
   both traversal paths.  Replaces BVHAccelerator::new's host build
   (bvh.rs:189-258) where admission time matters (serving): its tree is not
   the reference's, so closest-hit queries return the reference's hit except
   on rays where the reference's own answer depends on its visiting order:
   several primitives sharing the closest t, or a hit on its box's boundary
   that the pop-time box test (bvh.rs:390-395) culls at a smaller maxt.
   Occlusion bits are the reference's exactly.  No
   host tree is kept (pbrtx_bvh_* inspection does not apply).  *build_ms (may
   be NULL) receives the build's wall time, uploads included.  A render
   description selects it with split = 3. */
int pbrtx_scene_build_gpu(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t nfaces,
                          const pbrtx_sphere* spheres, uint32_t nspheres, pbrtx_scene** out, double* build_ms);
/* Copies both node layouts and the primitive records to HBM.  The handle is
   immutable afterwards and may be shared by host threads. */
int pbrtx_scene_create(const pbrtx_bvh* bvh, pbrtx_scene** out);
int pbrtx_scene_destroy(pbrtx_scene* scene);
int pbrtx_scene_device_bytes(const pbrtx_scene* scene, uint64_t* bytes);

/* ---- queries ------------------------------------------------------------- */
/* Intersectable::intersect over n rays (BVHAccelerator::intersect,
   bvh.rs:377-421): d_hits[i] = closest hit; on a hit d_rays[i].tmax = t, the
   reference's maxt side effect (geometric.rs:64). */
int pbrtx_intersect(pbrtx_scene* scene, pbrtx_ray* d_rays, uint32_t n, pbrtx_hit* d_hits,
                    int32_t path, void* stream);
/* Intersectable::intersect_p (intersection.rs:63-65): bit i of
   d_occluded[ceil(n/32)] = 1 iff ray i hits anything in [tmin, tmax]. */
int pbrtx_intersect_p(pbrtx_scene* scene, const pbrtx_ray* d_rays, uint32_t n,
                      uint32_t* d_occluded, int32_t path, void* stream);

/* pbrtx_intersect_p on the batched path without allocating: no counterpart
   in the reference.  For callers that cannot call hipMalloc (code under HIP
   graph capture, tight serving loops, harnesses that forbid allocation).  Its
   scratch (work counters, spilled traversal stacks) is a fixed 64 MiB device
   array per GPU, held until the process ends; the persistent grid is cut to
   the blocks whose spill stacks fit it.  Fails with PBRTX_ERR_ARG when one
   block's spill does not fit (with 256-thread blocks, a tree of more than
   about 10,900 record levels).
   Every call on a GPU shares that GPU's array: calls that overlap on one GPU
   (different streams or threads) race. */
int pbrtx_intersect_p_noalloc(pbrtx_scene* scene, const pbrtx_ray* d_rays, uint32_t n,
                              uint32_t* d_occluded, void* stream);

/* Launch shape of a path's closest-hit (kind 0), occlusion (kind 1) or
   renderer camera-ray (kind 2: closest hit of the coherent camera rays)
   kernel: block in {64,128,256}; blocks_per_cu 0 = occupancy maximum
   (batched path), else up to 64 (up to 1024 for kind 2); refill_pct = idle-lane refill threshold (batched path);
   scheme 0 = a traversal stack per ray over the compressed records (the
   exact ones when a scene has none), 1 = one stack per wave of 64
   consecutive rays in equal-sign groups (batched path, closest hit and
   camera rays only: occlusion (kind 1) refuses it with PBRTX_ERR_ARG), 2 = a
   stack per ray over the exact records.  Process-wide setting. */
int pbrtx_set_tuning(int32_t path, int32_t kind, int32_t block, int32_t blocks_per_cu, int32_t refill_pct,
                     int32_t scheme);
/* LDS traversal-stack entries per lane (8, 16 or 24; default 16; deeper
   entries spill to global memory) of the batched path's per-ray kernel for
   that kind.  Process-wide setting, kept by pbrtx_set_tuning. */
int pbrtx_set_stack_depth(int32_t path, int32_t kind, int32_t entries);
/* Lanes per ray of the batched path's occlusion queries: 0 (the default)
   chooses by launch size, several lanes per ray for launches too small to
   fill the GPU; 1 always one; 4, 8 or 16 always that many (when the
   scene has the compressed records and its record depth fits the group
   stack, else one).  No reference counterpart: a launch setting of this
   implementation.  Process-wide setting, kept by pbrtx_set_tuning. */
int pbrtx_set_ray_lanes(int32_t lanes);
/* Back to the built-in launch settings (the autotuner's choices on the
   direct-lighting and path-tracing frames).  Unless set explicitly, the
   batched path's camera rays use one stack per wave of 64 rays. */
int pbrtx_reset_tuning(void);

/* ---- renderer: SamplerRenderer::render (sampler_renderer.rs:146-206) ------ */
/* Scene + camera + sampler + integrator description.  Triangles (world-space
   vertices, index triples) and spheres, each with a material: mats holds 4 floats per
   material (kind 0 = matte / Lambertian, 1 = mirror; rgb reflectance).
   face_light: area-light id of an emitting face or -1; area_le: radiance of
   each area light (3 floats); point_lights: 6 floats each (position,
   intensity).  Camera: Camera::perspective (camera/mod.rs:105-135) with a
   static cam_to_world; StratifiedSampler nx x ny samples per pixel with
   jitter (stratified.rs:60-127) driven by a counter-based generator keyed by
   (seed, pixel, sample, dimension).  integrator 0 = direct lighting
   (light_samples per area light), 1 = path (max_depth, roulette once the
   bounce index exceeds rr_start).  split / max_prims as pbrtx_bvh_build. */
typedef struct pbrtx_render_desc {
    const float* verts; uint32_t nverts;
    const uint32_t* idx; uint32_t nfaces;
    const uint32_t* face_mat; const float* mats; uint32_t nmats;
    const int32_t* face_light; const float* area_le; uint32_t n_area;
    const float* point_lights; uint32_t n_point;
    float cam_to_world[16]; float fov_deg; float screen[4];
    uint32_t xres, yres, nx, ny, seed;
    int32_t integrator;
    uint32_t light_samples;
    uint32_t max_depth;
    uint32_t rr_start;
    int32_t split; uint32_t max_prims;  /* split 3: the tree built on the GPU (pbrtx_scene_build_gpu) */
    /* 0 StratifiedSampler (nx x ny strata, jittered; stratified.rs:60-127),
       1 LDSampler (nx*ny samples per pixel rounded up to a power of two;
       lds.rs:17-70), its scrambles and shuffles drawn from the counter-based
       generator, 2 HaltonSampler (nx*ny samples per pixel on average;
       halton.rs:18-107) over the image as one window.  Requests of the
       render server are stratified. */
    int32_t sampler;
    /* spheres (Shape::sphere, main.rs:314-323, sphere.rs:28-44) after the
       faces: prim id nfaces + k, material sphere_mat[k]; shaded with the
       sphere's differential geometry (sphere.rs:135-175, helpers.rs:17-43);
       not emitters.  NULL / 0 for none. */
    const pbrtx_sphere* spheres; uint32_t nspheres;
    const uint32_t* sphere_mat;
} pbrtx_render_desc;

typedef struct pbrtx_render_stats {
    uint64_t camera_rays, shadow_rays, extension_rays; /* rays traced, by kind */
    double ms_total;                 /* whole frame, HIP events on the render stream */
    double ms_trace_closest, ms_trace_any, ms_other; /* per stage, summed over launches */
    double build_ms;                 /* host BVH build at create time */
    uint32_t launches_closest, launches_any;
} pbrtx_render_stats;

typedef struct pbrtx_renderer pbrtx_renderer;
/* builds the BVH, uploads scene, lights and camera */
int pbrtx_renderer_create(const pbrtx_render_desc* desc, pbrtx_renderer** out);
/* samples per work chunk (whole pixels; bounds the per-chunk buffers) */
int pbrtx_renderer_set_chunk(pbrtx_renderer* r, uint32_t samples_per_chunk);
/* graph mode (enable != 0): after one ordinary frame of a path, a frame's
   device work is captured once as a hipGraph and replayed (captured again
   after a tuning change); per-stage times are then not measured (ms_total
   is).  Cuts the host launch cost of small frames. */
int pbrtx_renderer_set_graph(pbrtx_renderer* r, int32_t enable);
/* path tracer, batched path (enable != 0): the extension-ray queue of every
   bounce is octant-sorted (stable, within tiles of 4096 queue entries) before
   it is traced.  Images are unchanged: a ray's result does not depend on
   its queue position.  No counterpart in the reference (its rays are traced
   one at a time); off by default. */
int pbrtx_renderer_set_octant_sort(pbrtx_renderer* r, int32_t enable);
/* The film's filter (Film::image with a Filter, film.rs:77-122; filter.rs:44-86):
   kind 0 mean (box), 1 triangle, 2 gaussian (p0 = alpha), 3 mitchell
   (p0 = B, p1 = C), 4 lanczos (p0 = tau); xw, yw the half-widths.  Samples
   are added as Film::add_sample (film.rs:192-249) does, with the 16 x 16
   weight table, in pixel and sample order.  The default is Filter::mean(0.5,
   0.5). */
int pbrtx_renderer_set_filter(pbrtx_renderer* r, int32_t kind, float xw, float yw, float p0, float p1);
/* Filter::evaluate (filter.rs:89-124) and the film's weight table
   (film.rs:89-97, 256 floats, row-major in y) of a filter given as above;
   host only, no device needed. */
int pbrtx_filter_eval(int32_t kind, float xw, float yw, float p0, float p1, float x, float y, float* out);
int pbrtx_filter_table(int32_t kind, float xw, float yw, float p0, float p1, float* out256);
/* renders one frame with the given traversal path; rgb_out (host, may be
   NULL) receives xres*yres*3 floats: the film's final rgb */
int pbrtx_renderer_render(pbrtx_renderer* r, int32_t path, float* rgb_out, pbrtx_render_stats* stats);
/* Sample slots per pixel of the renderer's frames: nx*ny for the stratified
   sampler, that rounded up to a power of two for LDSampler, and for
   HaltonSampler the most samples any pixel receives (a pixel's samples take
   its first slots in sequence order; its other slots are empty: zero
   radiance, no film weight). */
int pbrtx_renderer_sample_slots(pbrtx_renderer* r, uint32_t* slots_per_pixel);
/* Film::write_image (film.rs:316-354, write_img :15-33) of the last frame:
   the film's final rgb to `filename` by its extension -- .pfm (the floats),
   .ppm (8-bit binary) or .png (8-bit RGB) -- the 8-bit forms with write_img's
   encoding (255 p^(1/2.2) + 0.5, clamped to [0, 255], truncated).
   PBRTX_ERR_STATE before the first frame, PBRTX_ERR_ARG for another
   extension. */
int pbrtx_renderer_write_image(pbrtx_renderer* r, const char* filename);
/* the same for a host rgb image (xres*yres*3 floats, rows top to bottom);
   host only, no device needed */
int pbrtx_write_image(const char* filename, const float* rgb, uint32_t xres, uint32_t yres);
/* write_img's per-channel encoding (film.rs:22-24) of n floats; host only */
int pbrtx_image_encode8(const float* rgb, uint64_t n, uint8_t* out);
/* radiance of the first `count` samples of the last frame (3 floats each;
   sample index = pixel * spp + sample, pixel = y * xres + x, spp as
   pbrtx_renderer_sample_slots gives it) */
int pbrtx_renderer_sample_radiance(pbrtx_renderer* r, float* out, uint64_t count);
int pbrtx_renderer_destroy(pbrtx_renderer* r);

/* ---- request server ------------------------------------------------------ */
/* A server of ray-query requests (Intersectable::intersect / intersect_p,
   intersection.rs:61-66, over BVHAccelerator scenes) against many
   device-resident scenes.  The reference has no server: this is the serving
   loop around the same queries.  One worker thread owns the GPU work.
   PBRTX_SERVE_BATCHED: requests are admitted in arrival order while their
   device buffers (32 B per ray + results) fit arena_bytes; each step traces
   up to batch_rays rays taken from every admitted request, one launch per
   (scene, kind), so requests that arrive while others run join the next step
   and large requests progress in chunks.  PBRTX_SERVE_PLAIN: one request at a
   time, one plain-path kernel per request.  A ray's result does not depend on
   the other requests in flight.
   The worker overlaps uploads, traversal and downloads on six streams
   (compute, three side streams for a step's launches, upload, download),
   which overlap only on different hardware queues: HIP maps all of a
   process's streams onto GPU_MAX_HW_QUEUES queues (4 by default), so a
   process that also holds other streams (a renderer, a second server) should
   set it to 8 or more before HIP starts (measured: at 70% of the saturated
   rate, 0.91 of the offered rate served with 4 queues beside a renderer, 1.00
   with 8). */
#define PBRTX_SERVE_BATCHED 0
#define PBRTX_SERVE_PLAIN 1
typedef struct pbrtx_server pbrtx_server;
typedef struct pbrtx_server_config {
    uint64_t arena_bytes;        /* HBM for in-flight requests (admission budget) */
    uint64_t scene_budget_bytes; /* cap on resident scene bytes, 0 = none */
    uint32_t batch_rays;         /* rays per step (batched mode) */
    int32_t mode;                /* PBRTX_SERVE_* */
} pbrtx_server_config;
/* per-request record; times in ms since the server was created */
typedef struct pbrtx_request_stats {
    uint64_t n_rays;
    int32_t status;              /* 0 or the error code of a failed request */
    uint32_t launches;           /* launches that carried part of it */
    double submit_ms, admit_ms, first_launch_ms, done_ms;
} pbrtx_request_stats;
typedef struct pbrtx_server_stats {
    uint64_t rays_done, requests_done, steps, launches;
    double busy_ms;              /* time spent in GPU steps */
    double elapsed_ms;           /* since creation */
    uint64_t arena_bytes, scene_bytes;
} pbrtx_server_counters;
int pbrtx_server_create(const pbrtx_server_config* cfg, pbrtx_server** out);
/* registers a scene (it must outlive the server); fails with
   PBRTX_ERR_OOM past scene_budget_bytes */
int pbrtx_server_add_scene(pbrtx_server* srv, const pbrtx_scene* scene, uint32_t* scene_id);
/* kind 0: closest hit, out = n pbrtx_hit; kind 1: occlusion, out =
   ceil(n/32) bit words.  rays and out are host memory that must stay valid
   until pbrtx_server_wait returns for the ticket. */
int pbrtx_server_submit(pbrtx_server* srv, uint32_t scene_id, int32_t kind, const pbrtx_ray* rays, uint32_t n,
                        void* out, uint64_t* ticket);
/* blocks until the request is done; fills *stats (may be NULL) and retires
   the ticket */
int pbrtx_server_wait(pbrtx_server* srv, uint64_t ticket, pbrtx_request_stats* stats);
int pbrtx_server_stats(pbrtx_server* srv, pbrtx_server_counters* stats);
/* finishes the requests already submitted, then stops */
int pbrtx_server_destroy(pbrtx_server* srv);

/* ---- render-request server ----------------------------------------------- */
/* A server of render requests (Renderer::render, renderer.rs:8-26, with the
   path integrator) against many device-resident scenes; the reference has no
   server.  One worker thread owns the GPU work.  PBRTX_SERVE_BATCHED:
   requests are admitted in arrival order while their device state (16 B per
   sample + 12 B per pixel) fits arena_bytes; each step gives every admitted
   request of a scene a share of up to batch_samples paths (progressive tiles:
   its next samples in pixel order) and runs them as one wavefront batch whose traversal
   launches all requests share; up to sixteen batches run side by side on
   their own streams, each with work buffers sized at creation for
   batch_samples paths (about 200 B a path).  PBRTX_SERVE_PLAIN: one request
   at a time on the plain path.  An image does not depend on the other
   requests in flight. */
typedef struct pbrtx_render_server pbrtx_render_server;
typedef struct pbrtx_render_server_config {
    uint64_t arena_bytes;    /* HBM for admitted requests (admission budget) */
    uint32_t batch_samples;  /* paths per batch (>= 4096) */
    int32_t mode;            /* PBRTX_SERVE_* */
} pbrtx_render_server_config;
/* One render: camera (as pbrtx_render_desc), image size and stratified
   sampler (nx x ny samples per pixel, seed). */
typedef struct pbrtx_render_request {
    uint32_t scene;
    float cam_to_world[16]; float fov_deg; float screen[4];
    uint32_t xres, yres, nx, ny, seed;
} pbrtx_render_request;
/* per-request record; times in ms since the server was created */
typedef struct pbrtx_render_request_stats {
    uint64_t n_samples;
    int32_t status;          /* 0 or the error code of a failed request */
    uint32_t batches;        /* batches that carried part of it */
    double submit_ms, admit_ms, first_launch_ms, done_ms;
} pbrtx_render_request_stats;
typedef struct pbrtx_render_server_counters {
    uint64_t samples_done, requests_done, steps, batches;
    uint64_t camera_rays, shadow_rays, extension_rays; /* rays traced */
    double busy_ms;          /* summed batch times (HIP events) */
    double elapsed_ms;       /* since creation */
    uint64_t arena_bytes, scene_bytes;
} pbrtx_render_server_counters;
int pbrtx_render_server_create(const pbrtx_render_server_config* cfg, pbrtx_render_server** out);
/* registers a scene: the geometry, material and light fields of desc (the
   camera, image and sampler fields are ignored); desc->integrator must be 1
   (path), and its max_depth / rr_start apply to every request of the scene.
   Builds the BVH on the host and uploads it. */
int pbrtx_render_server_add_scene(pbrtx_render_server* srv, const pbrtx_render_desc* desc, uint32_t* scene_id);
/* switches between PBRTX_SERVE_BATCHED and PBRTX_SERVE_PLAIN; only while no
   request is pending (every submitted ticket waited for) */
int pbrtx_render_server_set_mode(pbrtx_render_server* srv, int32_t mode);
/* rgb_out: host memory for xres*yres*3 floats (the film's final rgb, as
   pbrtx_renderer_render); it must stay valid until the wait returns */
int pbrtx_render_server_submit(pbrtx_render_server* srv, const pbrtx_render_request* req, float* rgb_out,
                               uint64_t* ticket);
/* blocks until the request is done; fills *stats (may be NULL) and retires
   the ticket */
int pbrtx_render_server_wait(pbrtx_render_server* srv, uint64_t ticket, pbrtx_render_request_stats* stats);
int pbrtx_render_server_stats(pbrtx_render_server* srv, pbrtx_render_server_counters* stats);
/* finishes the requests already submitted, then stops */
int pbrtx_render_server_destroy(pbrtx_render_server* srv);

/* ---- device memory and timing plumbing ----------------------------------- */
int pbrtx_device_malloc(void** d_ptr, uint64_t bytes);
int pbrtx_device_free(void* d_ptr);
int pbrtx_memcpy_htod(void* d_dst, const void* src, uint64_t bytes);
int pbrtx_memcpy_dtoh(void* dst, const void* d_src, uint64_t bytes);
int pbrtx_synchronize(void);
/* device-to-device copy by a plain streaming kernel (also the reference
   point for calibrating memory-traffic counters: it moves exactly `bytes`
   in and `bytes` out) */
int pbrtx_device_copy(void* d_dst, const void* d_src, uint64_t bytes, void* stream);
/* FP32 vector probe: pbrtx_fma_probe_threads() threads each run eight
   independent chains of `iters` fused multiply-adds and write one float to
   d_out (out_floats >= that count); *flops receives the flops launched.
   Time it with the timer below for the achievable non-packed FP32 rate. */
uint64_t pbrtx_fma_probe_threads(void);
int pbrtx_fma_probe(float* d_out, uint64_t out_floats, uint32_t iters, uint64_t* flops, void* stream);
/* 64-byte gather probe: reads each 64-byte record of d_buf (bytes = 64 x 2^k,
   k <= 32) exactly once, in a scattered order (neighbouring threads read
   far-apart records, as traversal lanes do), four 16-byte loads a record.
   The reference point for calibrating memory-traffic counters on the
   traversal's request size: it reads exactly `bytes` and writes (almost)
   nothing.  d_out: pbrtx_gather_probe_threads() words.  No counterpart in
   the reference (measurement plumbing). */
uint64_t pbrtx_gather_probe_threads(void);
int pbrtx_gather_probe(const void* d_buf, uint64_t bytes, uint32_t* d_out, uint64_t out_words, void* stream);
/* HIP-event timer on a stream: start, then stop returns elapsed ms */
int pbrtx_timer_start(void* stream);
int pbrtx_timer_stop(void* stream, float* ms);
int pbrtx_device_name(char* buf, uint32_t len);
/* page-locked host memory (request buffers whose copies should overlap
   device work) */
int pbrtx_host_malloc(void** ptr, uint64_t bytes);
int pbrtx_host_free(void* ptr);

#ifdef __cplusplus
}
#endif
#endif
