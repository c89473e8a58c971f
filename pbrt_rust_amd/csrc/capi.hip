// capi.hip -- extern "C" entry points of libpbrtx_gpu (include/pbrtx_gpu.h).
#include <hip/hip_runtime.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>

#include "../../include/pbrtx_gpu.h"
#include "bvh_build.hpp"
#include "capi_internal.hpp"
#include "trace.hpp"

using namespace px;

struct pbrtx_bvh {
    HostBvh h;
};

namespace px {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
int hip_fail(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? PBRTX_ERR_OOM : PBRTX_ERR_HIP;
}

uint32_t g_tuning_gen = 1;
// [path][kind]: closest hit, occlusion, camera rays of the renderer (coherent:
// the batched path walks them as packets)
// Defaults from the autotuner (python -m pbrt_rust_amd.autotune --path) on
// the 1024x1024 direct-lighting and 1920x1080 path-tracing frames: batched
// closest hit 256 threads, occupancy grid, refill below 60% busy lanes, 8
// LDS stack entries; occlusion the same with 16 entries.
static const TraceTuning k_default_tuning[2][3] = {
    {TraceTuning{256, 0, 0, SCHEME_RAY}, TraceTuning{256, 0, 0, SCHEME_RAY}, TraceTuning{256, 0, 0, SCHEME_RAY}},
    {TraceTuning{256, 0, 60, SCHEME_RAY, 8}, TraceTuning{256, 0, 60, SCHEME_RAY, 16},
     TraceTuning{128, 8, 0, SCHEME_PACKET}}};
TraceTuning g_tuning[2][3] = {
    {k_default_tuning[0][0], k_default_tuning[0][1], k_default_tuning[0][2]},
    {k_default_tuning[1][0], k_default_tuning[1][1], k_default_tuning[1][2]}};
bool g_camera_auto = true;

// Camera rays of a frame with spp samples per pixel.  Unless set explicitly,
// the batched path traces them with one stack per wave made in the
// traversal (a wave's 64 camera rays are neighbouring samples), 256 threads
// and 64 blocks per CU: eight times the waves that fit at once, so the
// fixed schedule hands each wave few packets and the hardware starts the
// waves that wait as others finish.  Camera stage of the 1024x1024 16-spp
// frame at 16 / 32 / 64 / 128 / 256 blocks per CU: 1.55 / 1.38 / 1.36 /
// 1.40 / 1.40 ms; the 1920x1080 64-spp path-tracing frame 14.8 / 14.4 /
// 14.1 / 14.1 / 14.1 ms; the 512x512 1-spp frame 0.195 / 0.196 / 0.196 /
// 0.199 / 0.244 ms.  (Earlier: a stack per ray against one per wave, 7.36
// against 7.09 ms on the 16-spp frame.)
TraceTuning camera_tuning(int path, uint32_t spp) {
    (void)spp;
    if (path == PATH_BATCHED && g_camera_auto) return TraceTuning{256, 64, 0, SCHEME_PACKET};
    return g_tuning[path][2];
}

namespace {
std::mutex g_scr_mu;
std::map<hipStream_t, TraceScratch> g_scratch;
}  // namespace

// Scratch of a stream, grown to at least `spill` bytes.  Calls on one stream
// are ordered, so one scratch per stream is race free.
int stream_scratch(hipStream_t st, size_t spill, TraceScratch** out) {
    std::lock_guard<std::mutex> lk(g_scr_mu);
    TraceScratch& s = g_scratch[st];
    if (!s.counter) {
        hipError_t e = hipMalloc((void**)&s.counter, TRACE_COUNTER_WORDS * sizeof(uint32_t));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(counter)");
        e = fill_u32(s.counter, 0u, TRACE_COUNTER_WORDS, st);
        if (e != hipSuccess) return hip_fail(e, "fill");
    }
    if (spill > s.spill_bytes) {
        if (s.spill) {
            hipError_t e = hipStreamSynchronize(st);
            if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
            (void)hipFree(s.spill);
            s.spill = nullptr;
            s.spill_bytes = 0;
        }
        hipError_t e = hipMalloc(&s.spill, spill);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(spill)");
        s.spill_bytes = spill;
    }
    *out = &s;
    return PBRTX_OK;
}

void release_stream_scratch(hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_scr_mu);
    auto it = g_scratch.find(st);
    if (it == g_scratch.end()) return;
    if (it->second.counter) (void)hipFree(it->second.counter);
    if (it->second.spill) (void)hipFree(it->second.spill);
    g_scratch.erase(it);
}

int upload(void** d, const void* h, size_t bytes) {
    *d = nullptr;
    if (bytes == 0) return PBRTX_OK;
    hipError_t e = hipMalloc(d, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(scene)");
    e = hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(scene)");
    return PBRTX_OK;
}

int scene_from_host(const HostBvh& h, DeviceScene* s) {
    memset(s, 0, sizeof(*s));
    int rc;
    if ((rc = upload(&s->d_rnodes, h.rnodes.data(), h.rnodes.size() * sizeof(RNode)))) return rc;
    if ((rc = upload(&s->d_qnodes, h.qnodes.data(), h.qnodes.size() * sizeof(QNode)))) return rc;
    if ((rc = upload(&s->d_prims, h.prims.data(), h.prims.size() * sizeof(PrimRec)))) return rc;
    if ((rc = upload(&s->d_spheres, h.spheres.data(), h.spheres.size() * sizeof(SphereRec)))) return rc;
    if ((rc = upload(&s->d_cnodes, h.cnodes.data(), h.cnodes.size() * sizeof(CNode)))) return rc;
    if ((rc = upload(&s->d_prims64, h.prims64.data(), h.prims64.size() * sizeof(PrimRec64)))) return rc;
    if ((rc = upload(&s->d_order, h.order.data(), h.order.size() * sizeof(uint32_t)))) return rc;
    if ((rc = upload(&s->d_ocnodes, h.ocnodes.data(), h.ocnodes.size() * sizeof(CNode)))) return rc;
    s->bytes = h.rnodes.size() * sizeof(RNode) + h.qnodes.size() * sizeof(QNode) +
               h.prims.size() * sizeof(PrimRec) + h.spheres.size() * sizeof(SphereRec) +
               h.cnodes.size() * sizeof(CNode) + h.prims64.size() * sizeof(PrimRec64) +
               h.order.size() * sizeof(uint32_t) + h.ocnodes.size() * sizeof(CNode);
    SceneView& v = s->view;
    v.rnodes = (const RNode*)s->d_rnodes;
    v.qnodes = (const QNode*)s->d_qnodes;
    v.prims = (const PrimRec*)s->d_prims;
    v.spheres = (const SphereRec*)s->d_spheres;
    // the compressed layout exists for a tree of records (a root that is a
    // leaf has none) whose bounds are all finite
    const bool cq = !h.cnodes.empty() && !h.prims64.empty();
    v.cnodes = cq ? (const CNode*)s->d_cnodes : nullptr;
    v.prims64 = cq ? (const PrimRec64*)s->d_prims64 : nullptr;
    const bool oc = cq && !h.ocnodes.empty();
    v.ocnodes = oc ? (const CNode*)s->d_ocnodes : v.cnodes;
    v.order = (const uint32_t*)s->d_order;
    v.n_rnodes = (uint32_t)h.rnodes.size();
    v.n_qnodes = (uint32_t)h.qnodes.size();
    v.n_prims = (uint32_t)h.prims.size();
    v.n_spheres = (uint32_t)h.spheres.size();
    v.qroot = h.qroot;
    v.ocroot = oc ? 0 : h.qroot;
    v.depth = h.depth;
    v.qdepth = h.qdepth;
    v.odepth = oc ? h.odepth : h.qdepth;
    for (int k = 0; k < 3; ++k) { v.root_lo[k] = h.root_lo[k]; v.root_hi[k] = h.root_hi[k]; }
    return PBRTX_OK;
}

void scene_free(DeviceScene* s) {
    if (s->d_rnodes) (void)hipFree(s->d_rnodes);
    if (s->d_qnodes) (void)hipFree(s->d_qnodes);
    if (s->d_prims) (void)hipFree(s->d_prims);
    if (s->d_spheres) (void)hipFree(s->d_spheres);
    if (s->d_cnodes) (void)hipFree(s->d_cnodes);
    if (s->d_prims64) (void)hipFree(s->d_prims64);
    if (s->d_order) (void)hipFree(s->d_order);
    if (s->d_ocnodes) (void)hipFree(s->d_ocnodes);
    memset(s, 0, sizeof(*s));
}

int run_closest(const SceneView& v, Ray* rays, uint32_t n, Hit* hits, int path, hipStream_t st) {
    if (path != PATH_PLAIN && path != PATH_BATCHED) return fail(PBRTX_ERR_ARG, "unknown path");
    const TraceTuning& tun = g_tuning[path][0];
    TraceScratch* scr;
    int rc = stream_scratch(st, trace_spill_bytes(v, path, tun), &scr);
    if (rc) return rc;
    hipError_t e = trace_closest(v, rays, n, hits, path, tun, scr, st);
    if (e != hipSuccess) return hip_fail(e, "trace_closest");
    return PBRTX_OK;
}

int run_any(const SceneView& v, const Ray* rays, uint32_t n, uint32_t* occ, int path, hipStream_t st) {
    if (path != PATH_PLAIN && path != PATH_BATCHED) return fail(PBRTX_ERR_ARG, "unknown path");
    const TraceTuning& tun = g_tuning[path][1];
    TraceScratch* scr;
    int rc = stream_scratch(st, trace_spill_bytes(v, path, tun), &scr);
    if (rc) return rc;
    hipError_t e = trace_any(v, rays, n, occ, path, tun, scr, st);
    if (e != hipSuccess) return hip_fail(e, "trace_any");
    return PBRTX_OK;
}

}  // namespace px

namespace {
hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
std::mutex g_ev_mu;
}  // namespace

extern "C" {

const char* pbrtx_last_error(void) { return g_err.c_str(); }

int pbrtx_bvh_build(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t nfaces,
                    const pbrtx_sphere* spheres, uint32_t nspheres, int32_t split, uint32_t max_prims,
                    pbrtx_bvh** out) {
    if (!out) return fail(PBRTX_ERR_ARG, "out is null");
    *out = nullptr;
    if (nspheres && !spheres) return fail(PBRTX_ERR_ARG, "spheres is null");
    static_assert(sizeof(pbrtx_sphere) == sizeof(SphereIn), "sphere layout");
    pbrtx_bvh* b = new (std::nothrow) pbrtx_bvh();
    if (!b) return fail(PBRTX_ERR_OOM, "host allocation");
    std::string err;
    if (!build_bvh(verts, nverts, idx, nfaces, reinterpret_cast<const SphereIn*>(spheres), nspheres,
                   split, max_prims, &b->h, &err)) {
        delete b;
        return fail(PBRTX_ERR_BUILD, err);
    }
    *out = b;
    return PBRTX_OK;
}

int pbrtx_bvh_info(const pbrtx_bvh* b, uint32_t* n_nodes, uint32_t* depth, uint32_t* n_prims,
                   double* build_ms) {
    if (!b) return fail(PBRTX_ERR_ARG, "bvh is null");
    if (n_nodes) *n_nodes = (uint32_t)b->h.rnodes.size();
    if (depth) *depth = b->h.depth;
    if (n_prims) *n_prims = (uint32_t)b->h.prims.size();
    if (build_ms) *build_ms = b->h.build_ms;
    return PBRTX_OK;
}

int pbrtx_bvh_export(const pbrtx_bvh* b, float* bounds, uint32_t* off, uint32_t* meta, uint8_t* leaf,
                     uint32_t* order) {
    if (!b) return fail(PBRTX_ERR_ARG, "bvh is null");
    const auto& n = b->h.rnodes;
    for (size_t i = 0; i < n.size(); ++i) {
        bool lf = (n[i].meta & 3u) == 3u;
        if (bounds)
            for (int k = 0; k < 3; ++k) { bounds[6 * i + k] = n[i].lo[k]; bounds[6 * i + 3 + k] = n[i].hi[k]; }
        if (off) off[i] = n[i].off;
        if (meta) meta[i] = lf ? (n[i].meta >> 2) : (n[i].meta & 3u);
        if (leaf) leaf[i] = lf ? 1 : 0;
    }
    if (order) memcpy(order, b->h.order.data(), b->h.order.size() * sizeof(uint32_t));
    return PBRTX_OK;
}

int pbrtx_bvh_layout_info(const pbrtx_bvh* b, uint32_t* n_wide, uint32_t* n_compressed, uint32_t* n_prims) {
    if (!b) return fail(PBRTX_ERR_ARG, "bvh is null");
    if (n_wide) *n_wide = (uint32_t)b->h.qnodes.size();
    if (n_compressed) *n_compressed = (uint32_t)b->h.cnodes.size();
    if (n_prims) *n_prims = (uint32_t)b->h.prims.size();
    return PBRTX_OK;
}

int pbrtx_bvh_occlusion_records(const pbrtx_bvh* b, uint32_t* n, uint32_t* depth, void* records) {
    if (!b) return fail(PBRTX_ERR_ARG, "bvh is null");
    const HostBvh& h = b->h;
    if (n) *n = (uint32_t)h.ocnodes.size();
    if (depth) *depth = h.odepth;
    if (records) memcpy(records, h.ocnodes.data(), h.ocnodes.size() * sizeof(CNode));
    return PBRTX_OK;
}

int pbrtx_bvh_export_layouts(const pbrtx_bvh* b, void* wide, void* compressed, void* prims64) {
    if (!b) return fail(PBRTX_ERR_ARG, "bvh is null");
    const HostBvh& h = b->h;
    if (wide) memcpy(wide, h.qnodes.data(), h.qnodes.size() * sizeof(QNode));
    if (compressed) memcpy(compressed, h.cnodes.data(), h.cnodes.size() * sizeof(CNode));
    if (prims64) memcpy(prims64, h.prims64.data(), h.prims64.size() * sizeof(PrimRec64));
    return PBRTX_OK;
}

int pbrtx_bvh_free(pbrtx_bvh* b) {
    delete b;
    return PBRTX_OK;
}

int pbrtx_scene_create(const pbrtx_bvh* b, pbrtx_scene** out) {
    if (!b || !out) return fail(PBRTX_ERR_ARG, "null argument");
    *out = nullptr;
    pbrtx_scene* s = new (std::nothrow) pbrtx_scene();
    if (!s) return fail(PBRTX_ERR_OOM, "host allocation");
    int rc = scene_from_host(b->h, &s->s);
    if (rc) {
        scene_free(&s->s);
        delete s;
        return rc;
    }
    *out = s;
    return PBRTX_OK;
}

int pbrtx_scene_destroy(pbrtx_scene* s) {
    if (!s) return PBRTX_OK;
    scene_free(&s->s);
    delete s;
    return PBRTX_OK;
}

int pbrtx_scene_device_bytes(const pbrtx_scene* s, uint64_t* bytes) {
    if (!s || !bytes) return fail(PBRTX_ERR_ARG, "null argument");
    *bytes = s->s.bytes;
    return PBRTX_OK;
}

int pbrtx_intersect(pbrtx_scene* s, pbrtx_ray* d_rays, uint32_t n, pbrtx_hit* d_hits, int32_t path,
                    void* stream) {
    if (!s) return fail(PBRTX_ERR_ARG, "scene is null");
    if (n == 0) return PBRTX_OK;
    if (!d_rays || !d_hits) return fail(PBRTX_ERR_ARG, "null ray or hit buffer");
    return run_closest(s->s.view, reinterpret_cast<Ray*>(d_rays), n, reinterpret_cast<Hit*>(d_hits), path,
                       (hipStream_t)stream);
}

int pbrtx_intersect_p(pbrtx_scene* s, const pbrtx_ray* d_rays, uint32_t n, uint32_t* d_occ, int32_t path,
                      void* stream) {
    if (!s) return fail(PBRTX_ERR_ARG, "scene is null");
    if (n == 0) return PBRTX_OK;
    if (!d_rays || !d_occ) return fail(PBRTX_ERR_ARG, "null ray or bit buffer");
    return run_any(s->s.view, reinterpret_cast<const Ray*>(d_rays), n, d_occ, path, (hipStream_t)stream);
}

int pbrtx_intersect_p_noalloc(pbrtx_scene* s, const pbrtx_ray* d_rays, uint32_t n, uint32_t* d_occ, void* stream) {
    if (!s) return fail(PBRTX_ERR_ARG, "scene is null");
    if (n == 0) return PBRTX_OK;
    if (!d_rays || !d_occ) return fail(PBRTX_ERR_ARG, "null ray or bit buffer");
    const TraceTuning& tun = g_tuning[PATH_BATCHED][1];
    const size_t per_block = trace_any_block_spill(s->s.view, tun);
    if (per_block > NOALLOC_SCRATCH_BYTES - NOALLOC_COUNTER_BYTES)
        return fail(PBRTX_ERR_ARG, "scene too deep for the fixed scratch: one block's spill stacks need " +
                                       std::to_string(per_block) + " bytes, the scratch holds " +
                                       std::to_string(NOALLOC_SCRATCH_BYTES - NOALLOC_COUNTER_BYTES));
    TraceScratch scr;
    hipError_t e = hipSuccess;
    if (noalloc_scratch(&scr, &e)) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    hipStream_t st = (hipStream_t)stream;
    // the array keeps the last call's contents: the counter block (work
    // counters, overflow flag) starts at zero
    e = fill_u32(scr.counter, 0u, TRACE_COUNTER_WORDS, st);
    if (e == hipSuccess) e = trace_any(s->s.view, reinterpret_cast<const Ray*>(d_rays), n, d_occ, PATH_BATCHED, tun, &scr, st);
    if (e != hipSuccess) return hip_fail(e, "trace_any (fixed scratch)");
    return PBRTX_OK;
}

int pbrtx_set_tuning(int32_t path, int32_t kind, int32_t block, int32_t blocks_per_cu, int32_t refill_pct,
                     int32_t scheme) {
    if (path != 0 && path != 1) return fail(PBRTX_ERR_ARG, "unknown path");
    if (kind < 0 || kind > 2) return fail(PBRTX_ERR_ARG, "kind must be 0 (closest hit), 1 (occlusion) or 2 (camera)");
    if (block != 64 && block != 128 && block != 256) return fail(PBRTX_ERR_ARG, "block must be 64, 128 or 256");
    // the camera kernel's fixed schedule takes any grid (more blocks than fit
    // at once are run as others finish); the persistent kernels' spill
    // stacks are sized by their grid
    if (blocks_per_cu < 0 || blocks_per_cu > (kind == 2 ? 1024 : 64))
        return fail(PBRTX_ERR_ARG, "blocks_per_cu out of range");
    if (refill_pct < 0 || refill_pct > 100) return fail(PBRTX_ERR_ARG, "refill_pct out of range");
    if (scheme != SCHEME_RAY && scheme != SCHEME_PACKET && scheme != SCHEME_RAY_EXACT)
        return fail(PBRTX_ERR_ARG, "unknown traversal scheme");
    if (scheme == SCHEME_PACKET && path != 1) return fail(PBRTX_ERR_ARG, "the packet scheme exists for the batched path only");
    if (scheme == SCHEME_PACKET && kind == 1)
        return fail(PBRTX_ERR_ARG, "occlusion has no packet scheme: its kernel was removed after it hung once with the "
                                   "cause not found");
    TraceTuning t{block, blocks_per_cu, refill_pct, scheme};
    t.stack = g_tuning[path][kind].stack;
    t.lanes = g_tuning[path][kind].lanes;
    g_tuning[path][kind] = t;
    if (kind == 2 && path == PATH_BATCHED) g_camera_auto = false;  // an explicit choice
    ++g_tuning_gen;  // captured frames are captured again
    return PBRTX_OK;
}

int pbrtx_reset_tuning(void) {
    for (int p = 0; p < 2; ++p)
        for (int k = 0; k < 3; ++k) g_tuning[p][k] = k_default_tuning[p][k];
    g_camera_auto = true;
    ++g_tuning_gen;
    return PBRTX_OK;
}

int pbrtx_set_stack_depth(int32_t path, int32_t kind, int32_t entries) {
    if (path != 0 && path != 1) return fail(PBRTX_ERR_ARG, "unknown path");
    if (kind < 0 || kind > 2) return fail(PBRTX_ERR_ARG, "kind must be 0 (closest hit), 1 (occlusion) or 2 (camera)");
    if (entries != 8 && entries != 16 && entries != 24) return fail(PBRTX_ERR_ARG, "stack depth must be 8, 16 or 24");
    g_tuning[path][kind].stack = entries;
    if (kind == 2 && path == PATH_BATCHED) g_camera_auto = false;
    ++g_tuning_gen;
    return PBRTX_OK;
}

int pbrtx_set_ray_lanes(int32_t lanes) {
    if (lanes != 0 && lanes != 1 && lanes != 4 && lanes != 8 && lanes != 16)
        return fail(PBRTX_ERR_ARG, "lanes per ray must be 0 (by launch size), 1, 4, 8 or 16");
    g_tuning[PATH_BATCHED][1].lanes = lanes;
    ++g_tuning_gen;
    return PBRTX_OK;
}

int pbrtx_device_malloc(void** p, uint64_t bytes) {
    if (!p) return fail(PBRTX_ERR_ARG, "null pointer");
    hipError_t e = hipMalloc(p, bytes ? bytes : 4);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    return PBRTX_OK;
}
int pbrtx_host_malloc(void** p, uint64_t bytes) {
    if (!p) return fail(PBRTX_ERR_ARG, "null pointer");
    hipError_t e = hipHostMalloc(p, bytes ? bytes : 4, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc");
    return PBRTX_OK;
}
int pbrtx_host_free(void* p) {
    if (!p) return PBRTX_OK;
    hipError_t e = hipHostFree(p);
    if (e != hipSuccess) return hip_fail(e, "hipHostFree");
    return PBRTX_OK;
}
int pbrtx_device_free(void* p) {
    if (!p) return PBRTX_OK;
    hipError_t e = hipFree(p);
    if (e != hipSuccess) return hip_fail(e, "hipFree");
    return PBRTX_OK;
}
int pbrtx_memcpy_htod(void* d, const void* h, uint64_t bytes) {
    hipError_t e = hipMemcpy(d, h, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy H2D");
    return PBRTX_OK;
}
int pbrtx_memcpy_dtoh(void* h, const void* d, uint64_t bytes) {
    hipError_t e = hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy D2H");
    return PBRTX_OK;
}
int pbrtx_synchronize(void) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
    // report a kernel-side stack overflow flag, if any
    std::lock_guard<std::mutex> lk(g_scr_mu);
    for (auto& kv : g_scratch) {
        if (!kv.second.counter) continue;
        uint32_t flag = 0;
        e = hipMemcpy(&flag, kv.second.counter + 1, 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpy flag");
        if (flag) return fail(PBRTX_ERR_STATE, "traversal stack deeper than the scene's recorded depth");
    }
    return PBRTX_OK;
}

int pbrtx_device_copy(void* d_dst, const void* d_src, uint64_t bytes, void* stream) {
    if (!d_dst || !d_src || (bytes % 16) != 0) return fail(PBRTX_ERR_ARG, "copy needs device pointers and a multiple of 16 bytes");
    hipError_t e = copy_bytes(d_dst, d_src, bytes, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "copy");
    return PBRTX_OK;
}

int pbrtx_fma_probe(float* d_out, uint64_t out_floats, uint32_t iters, uint64_t* flops, void* stream) {
    const uint64_t n = fma_chain_threads();
    if (!d_out || out_floats < n) return fail(PBRTX_ERR_ARG, "fma probe needs a device buffer of pbrtx_fma_probe_threads() floats");
    hipError_t e = fma_chain(d_out, iters, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "fma probe");
    if (flops) *flops = n * 8ull * 2ull * iters;
    return PBRTX_OK;
}
uint64_t pbrtx_fma_probe_threads(void) { return fma_chain_threads(); }

int pbrtx_gather_probe(const void* d_buf, uint64_t bytes, uint32_t* d_out, uint64_t out_words, void* stream) {
    const uint64_t recs = bytes / 64;
    if (!d_buf || !d_out || bytes % 64 != 0 || recs == 0 || (recs & (recs - 1)) != 0 || recs > (1ull << 32))
        return fail(PBRTX_ERR_ARG, "gather probe needs a device buffer of 64 x 2^k bytes (k <= 32)");
    if (out_words < gather_probe_threads())
        return fail(PBRTX_ERR_ARG, "gather probe needs pbrtx_gather_probe_threads() output words");
    hipError_t e = gather_probe(d_buf, recs, d_out, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "gather probe");
    return PBRTX_OK;
}
uint64_t pbrtx_gather_probe_threads(void) { return gather_probe_threads(); }

int pbrtx_timer_start(void* stream) {
    std::lock_guard<std::mutex> lk(g_ev_mu);
    hipError_t e;
    if (!g_ev0 && (e = hipEventCreate(&g_ev0)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    if (!g_ev1 && (e = hipEventCreate(&g_ev1)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    if ((e = hipEventRecord(g_ev0, (hipStream_t)stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    return PBRTX_OK;
}
int pbrtx_timer_stop(void* stream, float* ms) {
    std::lock_guard<std::mutex> lk(g_ev_mu);
    if (!g_ev0 || !g_ev1 || !ms) return fail(PBRTX_ERR_STATE, "timer not started");
    hipError_t e;
    if ((e = hipEventRecord(g_ev1, (hipStream_t)stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    if ((e = hipEventSynchronize(g_ev1)) != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    if ((e = hipEventElapsedTime(ms, g_ev0, g_ev1)) != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
    return PBRTX_OK;
}
int pbrtx_device_name(char* buf, uint32_t len) {
    if (!buf || len == 0) return fail(PBRTX_ERR_ARG, "null buffer");
    int dev = 0;
    hipDeviceProp_t p;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&p, dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return PBRTX_OK;
}

}  // extern "C"
