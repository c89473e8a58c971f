// render.hpp -- the renderer's device parameter blocks and the host-side
// pieces shared by the frame renderer (render.hip) and the render-request
// server (render_serve.hip): scene set-up, camera set-up, work buffers and
// the path tracer's per-batch launch sequence.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "../../include/pbrtx_gpu.h"
#include "camera.hpp"
#include "capi_internal.hpp"

namespace px {

struct FaceShade {  // per face, precomputed with the hit-time formulas
    float nn[3];    // dg.nn = normalize(dpdu x dpdv)  (mesh.rs:224-254, diff_geom.rs:52-60)
    float tn[3];    // normalize(dpdu)                  (bsdf/mod.rs:70-73)
    uint32_t mat;
    int32_t light;
};
// per sphere: what shading a hit needs (Sphere::intersect, sphere.rs:135-175)
struct SphereShade {
    float w2o[16], o2w[16];  // world->object and object->world m (row-major)
    float radius, thetamin, thetamax, phimax;
    uint32_t flip;           // reverse_orientation ^ transform_swaps_handedness
    uint32_t mat;
    uint32_t pad0, pad1;
};
struct AreaTriRec {
    float p1[3], p2[3], p3[3], nn[3];
    float area, pad0, pad1, pad2;
};
struct MatRec {
    uint32_t kind;
    float r, g, b;
};

// A served batch maps its paths to the samples of several requests: batch
// indices [start, next segment's start) are samples g0, g0+1, ... of request
// req.  Starts ascend from 0.
struct RenderSeg {
    uint32_t start, req;
    uint64_t g0;
};
constexpr uint32_t MAX_RENDER_SEGS = 256;  // a path keeps its segment in 8 bits
// A served request as the kernels see it: its camera and sampler, and where
// its sample radiance goes.
struct ReqView {
    CamParams cam;
    float4* sample_l;
};

struct RenderParams {
    float r2c[16], c2w[16];
    uint32_t xres, yres, nx, ny, spp, seed;
    uint32_t n_point, n_area, nlights, light_samples, max_depth, rr_start, S;
    float dx, dy;
    const FaceShade* faces;
    const MatRec* mats;
    const AreaTriRec* atris;
    const uint32_t* a_first;
    const uint32_t* a_count;
    const float* a_total;
    const float* a_le;     // 3 per area light
    const float* points;   // 6 per point light
    const SphereShade* sph;  // prims nfaces + k are spheres k
    uint32_t nfaces;
    // served batches only (null for a frame): the batch's segments and the
    // requests they name; the camera fields above are then unused
    const RenderSeg* segs;
    const ReqView* reqs;
    uint32_t nsegs;
    // film filter (frames): its 16 x 16 weight table, widths and their
    // inverses, and how many pixels a sample reaches either way; a null table
    // is the box filter of half-width 0.5
    const float* ftab;
    float fxw, fyw, fixw, fiyw;
    int32_t frx, fry;
    // its staging, sample-major (index s * pixels + pixel): each sample's
    // position within its pixel and its xyz
    float2* fpos;
    float4* fxyz;
    // LDSampler and HaltonSampler frames: each sample's position within its
    // pixel (pixel-major); null for the stratified sampler
    const float2* spos;
    // HaltonSampler frames: how many of each pixel's spp slots hold a sample
    // (the rest are empty); null otherwise
    const uint32_t* scount;
};

// Filter (filter.rs:4-124) of a film: kind 0 mean, 1 triangle, 2 gaussian
// (p0 = alpha), 3 mitchell (p0 = b, p1 = c), 4 lanczos (p0 = tau).
struct FilmFilter {
    int32_t kind = 0;
    float xw = 0.5f, yw = 0.5f, p0 = 0.0f, p1 = 0.0f;
};
constexpr int FILTER_TABLE_DIM = 16;  // film.rs:12
// Filter::evaluate and Film::image's weight table (film.rs:89-97); false for
// an unknown kind or a width that is not positive and finite.
bool filter_valid(const FilmFilter& f);
float filter_evaluate(const FilmFilter& f, float x, float y);
void filter_weight_table(const FilmFilter& f, float* out256);

struct PathState {
    float beta[3];
    float L[3];
    uint32_t gid;    // batch index of the path
    uint32_t flags;  // 1 specular, 2 done after this bounce's shadow ray, 4 finished;
                     // bits 8-15: the path's segment in a served batch
};

template <typename T>
struct DevArr {
    T* p = nullptr;
    size_t n = 0;
    int ensure(size_t count) {
        if (count <= n) return PBRTX_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T) + 16);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(render buffer)");
        n = count;
        return PBRTX_OK;
    }
    int upload(const std::vector<T>& v) {
        int rc = ensure(v.size() ? v.size() : 1);
        if (rc) return rc;
        if (v.empty()) return PBRTX_OK;
        hipError_t e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpy(render upload)");
        return PBRTX_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// A scene ready to render: the BVH on the device plus the shading tables;
// base holds the scene's part of RenderParams (no camera).
struct RenderScene {
    HostBvh bvh;
    DeviceScene scene{};
    RenderParams base{};
    uint32_t nfaces = 0;
    int integrator = 0;
    double build_ms = 0.0;
    DevArr<FaceShade> faces;
    DevArr<SphereShade> sph;
    DevArr<MatRec> mats;
    DevArr<AreaTriRec> atris;
    DevArr<uint32_t> a_first, a_count;
    DevArr<float> a_total, a_le, points;
    void release();
};
// Builds the BVH and uploads the scene part of d (camera fields unused).
int render_scene_setup(RenderScene* S, const pbrtx_render_desc* d);
// The camera, film and sampler part of RenderParams (Camera::perspective,
// StratifiedSampler); checks the image is not empty or too large.
int camera_setup(RenderParams* P, const float* cam_to_world, float fov_deg, const float* screen, uint32_t xres,
                 uint32_t yres, uint32_t nx, uint32_t ny, uint32_t seed);
CamParams cam_params(const RenderParams& P, uint64_t g0);

// Work buffers of one pass over a batch of samples.
struct WorkBufs {
    DevArr<Ray> rays, rays2, rays3, sh;
    DevArr<Hit> hits;
    DevArr<Hit> hits2;  // served batches: bounce 0's hits at their queue positions
    DevArr<float4> contrib, lbase;
    DevArr<uint32_t> qidx, pid, pid2, pid3, shpid, bcounts;
    DevArr<uint8_t> occ;  // occlusion of each queued shadow ray: a byte per ray
    DevArr<PathState> st;
    void release();
};

struct EventPool {
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    hipEvent_t next() {
        if (used == ev.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            ev.push_back(e);
        }
        return ev[used++];
    }
    ~EventPool() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

using EventPairs = std::vector<std::pair<hipEvent_t, hipEvent_t>>;
struct Timing {
    EventPairs closest, any, other;
};
// Per-stage event pairs on a stream; a null pool records nothing.
struct Stager {
    EventPool* pool = nullptr;
    Timing* T = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t begin() {
        if (!pool) return nullptr;
        hipEvent_t e = pool->next();
        return (e && hipEventRecord(e, st) == hipSuccess) ? e : nullptr;
    }
    hipError_t end(hipEvent_t a, EventPairs Timing::*which) {
        if (!pool || !a) return hipSuccess;
        hipEvent_t b = pool->next();
        if (!b) return hipErrorOutOfMemory;
        hipError_t e = hipEventRecord(b, st);
        if (e == hipSuccess) (T->*which).push_back({a, b});
        return e;
    }
};

// A second stream for the path tracer's shadow rays: bounce b's occlusion
// launch and shadow accumulation run on it beside bounce b + 1's extension
// traversal (a path's next vertex waits for its shadow result, its next ray
// does not).  ev: fork / join events; sg: per-stage timing on st.
struct ShadowStream {
    hipStream_t st = nullptr;
    TraceScratch* scr = nullptr;
    EventPool* ev = nullptr;
    Stager* sg = nullptr;
};

// Sizes W for batches of up to cap paths (no allocation afterwards).
int reserve_path_work(WorkBufs& W, uint32_t cap, bool osort);

// Renderer counters (u32 words): [1], [2] the two extension queues' counts,
// [5] this bounce's shadow rays, [6,7] running total of shadow rays (lo, hi),
// [8,9] running total of extension rays.
constexpr uint32_t N_COUNTERS = 16;

// The path tracer over one batch of n paths on stream st: samples g0 + i of
// the frame (P.segs null), or the samples P's segment table maps the batch
// to.  path: PATH_PLAIN (every slot traced each bounce) or PATH_BATCHED
// (compacted queues, persistent traversal); osort: octant-sort the extension
// queues (batched); sample_l: the frame's sample radiance (frames only).
// Adds the batch's rays to cnt[6..9]; enqueues device work only.
int enqueue_path_batch(const RenderParams& P, const SceneView& sv, WorkBufs& W, uint32_t* cnt, uint64_t g0,
                       uint32_t n, int path, bool osort, float4* sample_l, TraceScratch* scr, hipStream_t st,
                       Stager& sg, const ShadowStream* ss = nullptr);
// Film::add_sample + write_image over a frame's sample radiance.
hipError_t enqueue_film(const RenderParams& P, const float4* sample_l, float* rgb, hipStream_t st);
// Spill bytes of the traversal launches of a path (largest over the kinds).
size_t render_spill_bytes(const SceneView& sv, int path);

}  // namespace px
