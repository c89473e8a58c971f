// render_serve.hip -- a server of render requests over device-resident scenes.
//
// Clients submit render requests -- a scene, a camera, an image size and a
// stratified sampler: Renderer::render (renderer.rs:8-26) with the path
// integrator -- of any size against any registered scene.  One worker thread
// owns the GPU work:
//   admission  FIFO while the request's device state -- its per-sample
//              radiance (16 B per sample) and its film (12 B per pixel) --
//              fits a fixed HBM arena;
//   batching   each step gives every admitted request of a scene a share of
//              up to batch_samples paths (progressive tiles: the request's
//              next samples in pixel order) and runs the scene's shares as ONE batch of the wavefront path
//              tracer: every traversal launch of the batch is shared by the
//              requests in it, and a segment table maps the batch's paths to
//              (request, sample).  Batches of different scenes run side by
//              side on the server's slot streams, so a request that arrives
//              while others render joins the next batch of its scene;
//   retirement a request whose samples are all done gets its film computed
//              and copied out, and its arena space freed.
// PBRTX_SERVE_PLAIN renders one request at a time, whole, on the plain path
// (one traversal kernel per query, every path slot traced each bounce): the
// baseline the batched mode is measured against.  A sample's radiance is a
// pure function of (scene, request, sample): every image is the same alone
// and in any mix.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "arena.hpp"
#include "render.hpp"

namespace px {
namespace {

using Clock = std::chrono::steady_clock;
// Batches in flight, one stream each.  Saturated render serving (17 scenes,
// 8M-sample batches) at 2 / 3 / 4 / 5 / 6 / 8 / 12 slots with 8 hardware
// queues: 5.36 / 6.66-6.71 / 7.52-7.86 / 7.48 / 7.93 / 8.07-8.43 /
// 8.08-8.18 Grays/s; at 8 / 12 / 16 / 24 / 32 slots with 16 queues:
// 8.36-8.44 / 8.89-8.96 / 9.26-9.79 / 9.35 / 9.83 Grays/s (with 32 queues
// everything ran at half speed); paired in the full benchmark, six passes
// each: 16 slots 8.79-9.96 (mean 9.19), 8 slots 8.61-9.46 (mean 9.04).
// A batch's later bounces are small launches
// that leave most of the GPU to the others.  Each slot's work buffers are
// sized once for batch_samples.
constexpr int NSLOT = 16;
constexpr uint64_t SHARE_MIN = 65536;  // smallest per-request share of a batch (samples)

struct RRequest {
    uint64_t ticket = 0;
    uint32_t scene = 0;
    RenderParams P{};       // the scene's parameters with this request's camera
    uint64_t n = 0;         // samples
    uint32_t npix = 0;
    float* h_rgb = nullptr;
    size_t off = 0, bytes = 0;
    float4* d_sl = nullptr;  // sample radiance
    float* d_rgb = nullptr;  // film rgb
    uint64_t next = 0;       // samples handed to batches
    uint64_t done = 0;       // samples of finished batches
    uint32_t batches = 0;
    hipEvent_t out = nullptr;  // film copied out
    bool copying = false;
    int status = 0;          // published (under the lock): 0 pending, 2 done, < 0 error code
    int state = 0;           // the worker's own: 0 queued, 1 admitted, 2 done, < 0 error code
    double t_submit = 0, t_admit = 0, t_first = -1, t_done = 0;
};

size_t rrequest_bytes(uint64_t n, uint32_t npix) {
    return round_up(n * sizeof(float4), GRAIN) + round_up((size_t)npix * 3 * sizeof(float), GRAIN);
}

// One stream's batch state.
struct Slot {
    hipStream_t st = nullptr;
    WorkBufs W;
    DevArr<uint32_t> cnt;
    RenderSeg* h_segs = nullptr;  // pinned staging
    ReqView* h_reqs = nullptr;
    uint32_t* h_cnt = nullptr;
    RenderSeg* d_segs = nullptr;
    ReqView* d_reqs = nullptr;
    hipEvent_t start = nullptr, end = nullptr;
    std::vector<std::pair<RRequest*, uint32_t>> parts;  // the batch's shares
    uint32_t total = 0;
    bool busy = false;
};

}  // namespace

struct RenderServer {
    pbrtx_render_server_config cfg{};
    int device = 0;
    Clock::time_point t0;
    Arena arena;
    std::vector<std::unique_ptr<RenderScene>> scenes;
    uint64_t scene_bytes = 0;
    size_t spill_need = 0;
    Slot slots[NSLOT];
    std::deque<int> inflight;  // busy slots in launch order

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<RRequest*> queued;   // submitted, not admitted
    std::vector<RRequest*> active;  // admitted, not finished (worker-owned)
    std::unordered_map<uint64_t, std::unique_ptr<RRequest>> all;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    // statistics
    uint64_t samples_done = 0, requests_done = 0, steps = 0, batches = 0;
    uint64_t camera_rays = 0, shadow_rays = 0, extension_rays = 0;
    double busy_ms = 0;
    std::string worker_error;

    double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

    void run();
    bool admit_one(RRequest* r);
    int launch_batches();
    int retire(bool wait_oldest);
    int serve_plain(RRequest* r);
    int start_film(RRequest* r, hipStream_t st);
    void finish(RRequest* r, int status);
    void publish();
    std::vector<RRequest*> to_publish;  // finished, not yet published
    void poll_copies();
};

// Ends r on the worker: frees its device resources and queues its result for
// publication.  The worker touches r no more once publish() has run, since
// a waiter may then retire (and free) it.
void RenderServer::finish(RRequest* r, int status) {
    if (r->bytes) arena.release(r->off, r->bytes);
    r->bytes = 0;
    if (r->out) (void)hipEventDestroy(r->out);
    r->out = nullptr;
    r->state = status;
    r->t_done = now_ms();
    to_publish.push_back(r);
}

// Drops finished requests from the active list, then makes their results
// visible to waiters.
void RenderServer::publish() {
    active.erase(std::remove_if(active.begin(), active.end(), [](RRequest* r) { return r->state != 1; }), active.end());
    if (to_publish.empty()) return;
    std::lock_guard<std::mutex> lk(mu);
    for (RRequest* r : to_publish) {
        if (r->state == 2) {
            samples_done += r->n;
            ++requests_done;
        }
        r->status = r->state;
    }
    to_publish.clear();
    cv_done.notify_all();
}

bool RenderServer::admit_one(RRequest* r) {
    size_t off;
    if (!arena.alloc(r->bytes, &off)) return false;
    r->off = off;
    char* p = arena.base() + off;
    r->d_sl = (float4*)p;
    r->d_rgb = (float*)(p + round_up(r->n * sizeof(float4), GRAIN));
    hipError_t e = hipEventCreateWithFlags(&r->out, hipEventDisableTiming);
    {
        std::lock_guard<std::mutex> lk(mu);
        r->t_admit = now_ms();
        r->state = 1;
    }
    if (e != hipSuccess) finish(r, PBRTX_ERR_HIP);
    return true;
}

// The film of r and its copy to the host, on st after everything before it.
int RenderServer::start_film(RRequest* r, hipStream_t st) {
    hipError_t e = enqueue_film(r->P, r->d_sl, r->d_rgb, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r->h_rgb, r->d_rgb, (size_t)r->npix * 3 * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(r->out, st);
    if (e != hipSuccess) return hip_fail(e, "film");
    r->copying = true;
    return PBRTX_OK;
}

void RenderServer::poll_copies() {
    for (RRequest* r : active) {
        if (r->state == 1 && r->copying) {
            hipError_t q = hipEventQuery(r->out);
            if (q == hipSuccess) finish(r, 2);
            else if (q != hipErrorNotReady) finish(r, PBRTX_ERR_HIP);
        }
    }
}

// Launches one batch per scene with unscheduled samples, oldest request's
// scene first, while slots are free.  Returns the number of batches launched.
int RenderServer::launch_batches() {
    // scenes in the order of their oldest admitted request with work left
    std::vector<uint32_t> order;
    {
        std::vector<RRequest*> cand;
        for (RRequest* r : active)
            if (r->state == 1 && r->next < r->n) cand.push_back(r);
        std::sort(cand.begin(), cand.end(), [](RRequest* a, RRequest* b) { return a->ticket < b->ticket; });
        for (RRequest* r : cand)
            if (std::find(order.begin(), order.end(), r->scene) == order.end()) order.push_back(r->scene);
    }
    int launched = 0;
    for (uint32_t sc : order) {
        int si = -1;
        for (int k = 0; k < NSLOT; ++k)
            if (!slots[k].busy) { si = k; break; }
        if (si < 0) break;
        Slot& S = slots[si];
        std::vector<RRequest*> reqs;
        for (RRequest* r : active)
            if (r->state == 1 && r->scene == sc && r->next < r->n) reqs.push_back(r);
        std::sort(reqs.begin(), reqs.end(), [](RRequest* a, RRequest* b) { return a->ticket < b->ticket; });
        if (reqs.size() > MAX_RENDER_SEGS) reqs.resize(MAX_RENDER_SEGS);
        const uint64_t cap = cfg.batch_samples;
        const uint64_t share = std::max<uint64_t>(cap / reqs.size(), std::min<uint64_t>(SHARE_MIN, cap / 8));
        // progressive tiles: the next samples of each request, an equal share
        // each (at least SHARE_MIN or an eighth of the batch), then what is
        // left of the batch to the oldest requests
        S.parts.clear();
        uint64_t total = 0;
        for (int pass = 0; pass < 2 && total < cap; ++pass) {
            for (RRequest* r : reqs) {
                if (total >= cap) break;
                uint64_t have = 0;
                for (auto& pt : S.parts)
                    if (pt.first == r) have += pt.second;
                const uint64_t left = r->n - r->next - have;
                if (left == 0) continue;
                const uint64_t take = std::min<uint64_t>(left, pass == 0 ? std::min(share, cap - total) : cap - total);
                if (take == 0) continue;
                bool merged = false;
                for (auto& pt : S.parts)
                    if (pt.first == r) { pt.second += (uint32_t)take; merged = true; }
                if (!merged) S.parts.push_back({r, (uint32_t)take});
                total += take;
            }
        }
        if (total == 0) continue;
        // segment table and request views
        uint32_t start = 0;
        for (size_t k = 0; k < S.parts.size(); ++k) {
            RRequest* r = S.parts[k].first;
            S.h_segs[k] = RenderSeg{start, (uint32_t)k, r->next};
            S.h_reqs[k] = ReqView{cam_params(r->P, 0), r->d_sl};
            start += S.parts[k].second;
            r->next += S.parts[k].second;
            r->batches++;
        }
        const RenderScene* scn;
        size_t spill;
        {
            std::lock_guard<std::mutex> lk(mu);  // add_scene may grow the table
            scn = scenes[sc].get();
            spill = spill_need;
        }
        RenderParams P = scn->base;
        P.segs = S.d_segs;
        P.reqs = S.d_reqs;
        P.nsegs = (uint32_t)S.parts.size();
        const double t_launch = now_ms();
        for (auto& pt : S.parts)
            if (pt.first->t_first < 0) pt.first->t_first = t_launch;
        hipError_t e = hipEventRecord(S.start, S.st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(S.d_segs, S.h_segs, S.parts.size() * sizeof(RenderSeg), hipMemcpyHostToDevice, S.st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(S.d_reqs, S.h_reqs, S.parts.size() * sizeof(ReqView), hipMemcpyHostToDevice, S.st);
        if (e == hipSuccess) e = fill_u32(S.cnt.p, 0u, N_COUNTERS, S.st);
        if (e != hipSuccess) return hip_fail(e, "batch set-up");
        TraceScratch* scr;
        int rc = stream_scratch(S.st, spill, &scr);
        if (rc) return rc;
        // (the shadow rays stay on the batch's stream: with three batches
        // in flight, a second stream per batch as in the frame renderer
        // served fewer rays, 6.24-6.29 against 6.51-6.61 Grays/s)
        Stager none;
        if (S.parts.size() == 1) {
            // one request: the frame path over its samples [g0, g0 + total)
            // with the request's own parameters, where the traversal makes
            // the camera rays itself (no segment table, no init pass); a
            // sample's radiance is the same either way
            RRequest* r = S.parts[0].first;
            rc = enqueue_path_batch(r->P, scn->scene.view, S.W, S.cnt.p, S.h_segs[0].g0, (uint32_t)total,
                                    PATH_BATCHED, false, r->d_sl, scr, S.st, none);
        } else {
            rc = enqueue_path_batch(P, scn->scene.view, S.W, S.cnt.p, 0, (uint32_t)total, PATH_BATCHED, false,
                                    nullptr, scr, S.st, none);
        }
        if (rc) return rc;
        e = hipMemcpyAsync(S.h_cnt, S.cnt.p, N_COUNTERS * sizeof(uint32_t), hipMemcpyDeviceToHost, S.st);
        if (e == hipSuccess) e = hipEventRecord(S.end, S.st);
        if (e != hipSuccess) return hip_fail(e, "batch end");
        S.total = (uint32_t)total;
        S.busy = true;
        inflight.push_back(si);
        ++batches;
        ++launched;
    }
    if (launched) ++steps;
    return launched;
}

// Retires finished batches in launch order (waiting for the oldest first if
// asked): their samples count as done, and a request with all samples done
// gets its film and copy-out on that batch's stream.
int RenderServer::retire(bool wait_oldest) {
    if (wait_oldest && !inflight.empty()) {
        hipError_t e = hipEventSynchronize(slots[inflight.front()].end);
        if (e != hipSuccess) return hip_fail(e, "batch");
    }
    while (!inflight.empty()) {
        Slot& S = slots[inflight.front()];
        hipError_t q = hipEventQuery(S.end);
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) return hip_fail(q, "batch");
        float ms = 0;
        if (hipEventElapsedTime(&ms, S.start, S.end) == hipSuccess) busy_ms += ms;
        camera_rays += S.total;
        shadow_rays += (uint64_t)S.h_cnt[6] | ((uint64_t)S.h_cnt[7] << 32);
        extension_rays += (uint64_t)S.h_cnt[8] | ((uint64_t)S.h_cnt[9] << 32);
        for (auto& pt : S.parts) {
            RRequest* r = pt.first;
            r->done += pt.second;
            if (r->done == r->n && r->state == 1) {
                int rc = start_film(r, S.st);
                if (rc) return rc;
            }
        }
        S.parts.clear();
        S.busy = false;
        inflight.pop_front();
    }
    return PBRTX_OK;
}

int RenderServer::serve_plain(RRequest* r) {
    Slot& S = slots[0];
    const RenderScene* scn;
    size_t spill;
    {
        std::lock_guard<std::mutex> lk(mu);  // add_scene may grow the table
        scn = scenes[r->scene].get();
        spill = spill_need;
    }
    TraceScratch* scr;
    int rc = stream_scratch(S.st, spill, &scr);
    if (rc) return rc;
    r->t_first = now_ms();
    hipError_t e = hipEventRecord(S.start, S.st);
    if (e == hipSuccess) e = fill_u32(S.cnt.p, 0u, N_COUNTERS, S.st);
    if (e != hipSuccess) return hip_fail(e, "plain set-up");
    Stager none;
    for (uint64_t g0 = 0; g0 < r->n; g0 += cfg.batch_samples) {
        const uint32_t n = (uint32_t)std::min<uint64_t>(cfg.batch_samples, r->n - g0);
        rc = enqueue_path_batch(r->P, scn->scene.view, S.W, S.cnt.p, g0, n, PATH_PLAIN, false, r->d_sl, scr, S.st,
                                none);
        if (rc) return rc;
        r->batches++;
        ++batches;
    }
    e = hipMemcpyAsync(S.h_cnt, S.cnt.p, N_COUNTERS * sizeof(uint32_t), hipMemcpyDeviceToHost, S.st);
    if (e == hipSuccess) e = hipEventRecord(S.end, S.st);
    if (e != hipSuccess) return hip_fail(e, "plain render");
    rc = start_film(r, S.st);
    if (rc) return rc;
    e = hipEventSynchronize(r->out);
    if (e != hipSuccess) return hip_fail(e, "plain render");
    float ms = 0;
    if (hipEventElapsedTime(&ms, S.start, S.end) == hipSuccess) busy_ms += ms;
    camera_rays += r->n;
    shadow_rays += (uint64_t)S.h_cnt[6] | ((uint64_t)S.h_cnt[7] << 32);
    extension_rays += (uint64_t)S.h_cnt[8] | ((uint64_t)S.h_cnt[9] << 32);
    r->next = r->done = r->n;
    ++steps;
    return PBRTX_OK;
}

void RenderServer::run() {
    (void)hipSetDevice(device);
    while (true) {
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queued.empty() || !active.empty(); });
            if (stop && queued.empty() && active.empty()) break;
        }
        // admission, in arrival order, as the arena frees up
        while (true) {
            RRequest* r = nullptr;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (queued.empty()) break;
                r = queued.front();
            }
            if (cfg.mode == PBRTX_SERVE_PLAIN && !active.empty()) break;  // one request at a time
            if (!admit_one(r)) break;
            {
                std::lock_guard<std::mutex> lk(mu);
                queued.pop_front();
            }
            if (r->state == 1) active.push_back(r);
        }
        int rc = PBRTX_OK;
        if (cfg.mode == PBRTX_SERVE_PLAIN) {
            if (!active.empty()) {
                RRequest* r = active.front();
                rc = serve_plain(r);
                if (rc == PBRTX_OK && r->state == 1) finish(r, 2);
            }
        } else {
            int launched = launch_batches();
            if (launched < 0) rc = launched;
            if (rc == PBRTX_OK) rc = retire(false);
            if (rc == PBRTX_OK) poll_copies();
            if (rc == PBRTX_OK && launched == 0) {
                // nothing new to launch: wait for the oldest batch, else give
                // the film copies a moment
                if (!inflight.empty()) {
                    rc = retire(true);
                    if (rc == PBRTX_OK) poll_copies();
                } else {
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
                }
            }
        }
        if (rc < 0) {
            // a device error ends every admitted request with that code
            std::string msg = pbrtx_last_error();
            for (Slot& S : slots) (void)hipStreamSynchronize(S.st);
            inflight.clear();
            for (Slot& S : slots) {
                S.busy = false;
                S.parts.clear();
            }
            for (RRequest* r : active)
                if (r->state == 1) finish(r, rc);
            active.clear();
            std::lock_guard<std::mutex> lk(mu);
            worker_error = msg;
        }
        publish();
    }
    publish();
    for (Slot& S : slots) (void)hipStreamSynchronize(S.st);
}

}  // namespace px

struct pbrtx_render_server {
    px::RenderServer s;
};

extern "C" {

int pbrtx_render_server_create(const pbrtx_render_server_config* cfg, pbrtx_render_server** out) {
    using namespace px;
    if (!cfg || !out) return fail(PBRTX_ERR_ARG, "null argument");
    *out = nullptr;
    if (cfg->mode != PBRTX_SERVE_BATCHED && cfg->mode != PBRTX_SERVE_PLAIN) return fail(PBRTX_ERR_ARG, "unknown mode");
    if (cfg->arena_bytes < (1u << 20)) return fail(PBRTX_ERR_ARG, "arena_bytes below 1 MiB");
    if (cfg->batch_samples < 4096) return fail(PBRTX_ERR_ARG, "batch_samples below 4096");
    pbrtx_render_server* p = new (std::nothrow) pbrtx_render_server();
    if (!p) return fail(PBRTX_ERR_OOM, "host allocation");
    RenderServer& s = p->s;
    s.cfg = *cfg;
    s.t0 = Clock::now();
    hipError_t e = hipGetDevice(&s.device);
    if (e == hipSuccess) e = s.arena.init(cfg->arena_bytes);
    int rc = PBRTX_OK;
    for (int k = 0; k < NSLOT && e == hipSuccess && rc == PBRTX_OK; ++k) {
        Slot& S = s.slots[k];
        e = hipStreamCreateWithFlags(&S.st, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&S.start);
        if (e == hipSuccess) e = hipEventCreate(&S.end);
        if (e == hipSuccess) e = hipHostMalloc((void**)&S.h_segs, MAX_RENDER_SEGS * sizeof(RenderSeg), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&S.h_reqs, MAX_RENDER_SEGS * sizeof(ReqView), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&S.h_cnt, N_COUNTERS * sizeof(uint32_t), hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc((void**)&S.d_segs, MAX_RENDER_SEGS * sizeof(RenderSeg));
        if (e == hipSuccess) e = hipMalloc((void**)&S.d_reqs, MAX_RENDER_SEGS * sizeof(ReqView));
        if (e == hipSuccess) rc = S.cnt.ensure(N_COUNTERS);
        // sized once: no allocation while serving
        if (e == hipSuccess && rc == PBRTX_OK) rc = reserve_path_work(S.W, cfg->batch_samples, false);
    }
    if (e != hipSuccess || rc != PBRTX_OK) {
        if (e != hipSuccess) rc = hip_fail(e, "render server setup");
        pbrtx_render_server_destroy(p);
        return rc;
    }
    s.worker = std::thread([&s] { s.run(); });
    *out = p;
    return PBRTX_OK;
}

int pbrtx_render_server_set_mode(pbrtx_render_server* p, int32_t mode) {
    using namespace px;
    if (!p) return fail(PBRTX_ERR_ARG, "null server");
    if (mode != PBRTX_SERVE_BATCHED && mode != PBRTX_SERVE_PLAIN) return fail(PBRTX_ERR_ARG, "unknown mode");
    RenderServer& s = p->s;
    std::lock_guard<std::mutex> lk(s.mu);
    if (!s.queued.empty() || !s.all.empty()) return fail(PBRTX_ERR_STATE, "requests in flight");
    s.cfg.mode = mode;
    return PBRTX_OK;
}

int pbrtx_render_server_add_scene(pbrtx_render_server* p, const pbrtx_render_desc* scene, uint32_t* scene_id) {
    using namespace px;
    if (!p || !scene || !scene_id) return fail(PBRTX_ERR_ARG, "null argument");
    if (scene->integrator != 1) return fail(PBRTX_ERR_ARG, "the render server runs the path integrator (integrator 1)");
    auto sc = std::make_unique<RenderScene>();
    int rc = render_scene_setup(sc.get(), scene);
    if (rc) {
        sc->release();
        return rc;
    }
    RenderServer& s = p->s;
    std::lock_guard<std::mutex> lk(s.mu);
    const SceneView& v = sc->scene.view;
    s.spill_need = std::max(s.spill_need, std::max(render_spill_bytes(v, PATH_BATCHED), render_spill_bytes(v, PATH_PLAIN)));
    s.scene_bytes += sc->scene.bytes;
    s.scenes.push_back(std::move(sc));
    *scene_id = (uint32_t)(s.scenes.size() - 1);
    return PBRTX_OK;
}

int pbrtx_render_server_submit(pbrtx_render_server* p, const pbrtx_render_request* req, float* rgb_out,
                               uint64_t* ticket) {
    using namespace px;
    if (!p || !req || !rgb_out || !ticket) return fail(PBRTX_ERR_ARG, "null argument");
    RenderServer& s = p->s;
    auto r = std::make_unique<RRequest>();
    r->scene = req->scene;
    {
        std::lock_guard<std::mutex> lk(s.mu);
        if (req->scene >= s.scenes.size()) return fail(PBRTX_ERR_ARG, "unknown scene id");
        r->P = s.scenes[req->scene]->base;
    }
    int rc = camera_setup(&r->P, req->cam_to_world, req->fov_deg, req->screen, req->xres, req->yres, req->nx,
                          req->ny, req->seed);
    if (rc) return rc;
    r->npix = req->xres * req->yres;
    r->n = (uint64_t)r->npix * r->P.spp;
    if (r->n > 0xFFFFFFFFull) return fail(PBRTX_ERR_ARG, "more than 2^32 - 1 samples in one request");
    r->h_rgb = rgb_out;
    r->bytes = rrequest_bytes(r->n, r->npix);
    if (r->bytes > s.arena.capacity()) return fail(PBRTX_ERR_OOM, "request larger than the server's arena");
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.stop) return fail(PBRTX_ERR_STATE, "server is shutting down");
    r->ticket = s.next_ticket++;
    r->t_submit = s.now_ms();
    *ticket = r->ticket;
    s.queued.push_back(r.get());
    s.all.emplace(r->ticket, std::move(r));
    s.cv_work.notify_one();
    return PBRTX_OK;
}

int pbrtx_render_server_wait(pbrtx_render_server* p, uint64_t ticket, pbrtx_render_request_stats* st) {
    using namespace px;
    if (!p) return fail(PBRTX_ERR_ARG, "null server");
    RenderServer& s = p->s;
    std::unique_lock<std::mutex> lk(s.mu);
    auto it = s.all.find(ticket);
    if (it == s.all.end()) return fail(PBRTX_ERR_ARG, "unknown ticket");
    RRequest* r = it->second.get();
    s.cv_done.wait(lk, [&] { return r->status == 2 || r->status < 0; });
    if (st) {
        st->n_samples = r->n;
        st->status = r->status == 2 ? PBRTX_OK : r->status;
        st->batches = r->batches;
        st->submit_ms = r->t_submit;
        st->admit_ms = r->t_admit;
        st->first_launch_ms = r->t_first;
        st->done_ms = r->t_done;
    }
    const int status = r->status;
    s.all.erase(it);
    if (status < 0) return fail(status, "request failed: " + s.worker_error);
    return PBRTX_OK;
}

int pbrtx_render_server_stats(pbrtx_render_server* p, pbrtx_render_server_counters* st) {
    using namespace px;
    if (!p || !st) return fail(PBRTX_ERR_ARG, "null argument");
    RenderServer& s = p->s;
    std::lock_guard<std::mutex> lk(s.mu);
    st->samples_done = s.samples_done;
    st->requests_done = s.requests_done;
    st->steps = s.steps;
    st->batches = s.batches;
    st->camera_rays = s.camera_rays;
    st->shadow_rays = s.shadow_rays;
    st->extension_rays = s.extension_rays;
    st->busy_ms = s.busy_ms;
    st->elapsed_ms = s.now_ms();
    st->arena_bytes = s.arena.capacity();
    st->scene_bytes = s.scene_bytes;
    return PBRTX_OK;
}

int pbrtx_render_server_destroy(pbrtx_render_server* p) {
    using namespace px;
    if (!p) return PBRTX_OK;
    RenderServer& s = p->s;
    {
        std::lock_guard<std::mutex> lk(s.mu);
        s.stop = true;
        s.cv_work.notify_all();
    }
    if (s.worker.joinable()) s.worker.join();
    for (Slot& S : s.slots) {
        if (S.st) (void)hipStreamSynchronize(S.st);
        S.W.release();
        S.cnt.release();
        if (S.h_segs) (void)hipHostFree(S.h_segs);
        if (S.h_reqs) (void)hipHostFree(S.h_reqs);
        if (S.h_cnt) (void)hipHostFree(S.h_cnt);
        if (S.d_segs) (void)hipFree(S.d_segs);
        if (S.d_reqs) (void)hipFree(S.d_reqs);
        if (S.start) (void)hipEventDestroy(S.start);
        if (S.end) (void)hipEventDestroy(S.end);
        if (S.st) {
            release_stream_scratch(S.st);
            (void)hipStreamDestroy(S.st);
        }
    }
    for (auto& sc : s.scenes) sc->release();
    delete p;
    return PBRTX_OK;
}

}  // extern "C"
