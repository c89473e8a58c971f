// serve.hip -- a request server over device-resident BVHs.
//
// Clients submit ray-query requests of any size (closest hit or occlusion,
// Intersectable::{intersect, intersect_p}, intersection.rs:61-66) against any
// of the registered scenes; one worker thread owns all GPU work:
//   admission  FIFO while the request's device buffers (rays + results) fit a
//              fixed HBM arena; the rays are copied in on a copy stream, so
//              new requests land while earlier ones are traced;
//   batching   every step takes up to a quota of rays from each admitted
//              request and traces all chunks of one (scene, kind) in a single
//              persistent-kernel launch over a segment table, so requests that
//              arrive while others are on the GPU join the next step
//              (continuous batching) and big requests progress in chunks;
//   retirement a request whose rays have all been traced has its results
//              copied out and its arena space freed at once.
// PBRTX_SERVE_PLAIN instead runs each request alone, in arrival order, with
// one plain-path kernel per query (the baseline the batched mode is measured
// against).  Every ray's result is a pure function of (scene, ray): the same
// alone and in any mix.
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

#include "../../include/pbrtx_gpu.h"
#include "arena.hpp"
#include "capi_internal.hpp"

namespace px {
namespace {

// Streams beside the compute stream for the launches of one step.
constexpr int SIDE_STREAMS = 3;

using Clock = std::chrono::steady_clock;
constexpr uint32_t CHUNK_MIN = 65536;  // smallest per-request share of a step (rays)

struct Request {
    uint64_t ticket = 0;
    uint32_t scene = 0;
    int kind = 0;  // 0 closest hit, 1 occlusion
    uint32_t n = 0;
    const pbrtx_ray* h_rays = nullptr;
    void* h_out = nullptr;
    size_t off = 0, bytes = 0;
    Ray* d_rays = nullptr;
    Hit* d_hits = nullptr;
    uint32_t* d_occ = nullptr;
    hipEvent_t ready = nullptr;  // rays resident (upload stream)
    hipEvent_t out = nullptr;    // results copied out (download stream)
    bool is_ready = false;
    bool copying = false;
    uint32_t launched = 0;       // rays handed to launches so far
    uint32_t launches = 0;
    int status = 0;              // published (under the lock): 0 pending, 2 done, < 0 error code
    int state = 0;               // the worker's own: 0 queued, 1 admitted, 2 done, < 0 error code
    double t_submit = 0, t_admit = 0, t_first = -1, t_done = 0;
};

size_t out_bytes(int kind, uint32_t n) {
    return kind == 0 ? (size_t)n * sizeof(Hit) : (((size_t)n + 31) / 32) * 4;
}
size_t request_bytes(int kind, uint32_t n) {
    return round_up((size_t)n * sizeof(Ray), GRAIN) + round_up(out_bytes(kind, n), GRAIN);
}

// A launched step: its events and the half of the segment staging it uses.
struct Step {
    hipEvent_t start = nullptr, end = nullptr;
    int half = 0;
};

}  // namespace

struct Server {
    pbrtx_server_config cfg{};
    int device = 0;
    Clock::time_point t0;
    Arena arena;
    std::vector<const DeviceScene*> scenes;  // guarded by mu (add_scene runs on client threads)
    uint64_t scene_bytes = 0;
    size_t spill_need = 0;                   // guarded by mu
    // compute, upload and download streams (all default priority)
    hipStream_t s_comp = nullptr, s_up = nullptr, s_down = nullptr;
    // a step's launches (one per scene and kind) go to s_comp and these in
    // turn, so small launches of different scenes run side by side; each
    // waits for the step's segment upload (ev_up), and s_comp for each of
    // them (ev_side) before the step's end
    hipStream_t s_side[SIDE_STREAMS] = {};
    hipEvent_t ev_up = nullptr, ev_side[SIDE_STREAMS] = {};
    TraceSeg* h_segs = nullptr;  // pinned staging of the segment tables, two halves
    TraceSeg* d_segs = nullptr;
    uint32_t seg_cap = 0;        // per half
    std::deque<Step> inflight;   // at most two steps
    std::vector<Step> spare;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queued;  // submitted, not admitted
    std::vector<Request*> active; // admitted, not finished (worker-owned)
    std::unordered_map<uint64_t, std::unique_ptr<Request>> all;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    // statistics
    uint64_t rays_done = 0, steps = 0, launches = 0, requests_done = 0;
    double busy_ms = 0;
    std::string worker_error;

    double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

    void run();
    bool admit_one(Request* r);
    int launch_step();
    int retire_steps(bool wait_oldest);
    int serve_plain(Request* r);
    void start_copy_out(Request* r, hipEvent_t after);
    void finish(Request* r, int status);
    void publish();
    std::vector<Request*> to_publish;  // finished, not yet published
};

// Ends r on the worker: frees its device resources and queues its result for
// publication.  The worker touches r no more once publish() has run, since
// a waiter may then retire (and free) it.
void Server::finish(Request* r, int status) {
    if (r->bytes) arena.release(r->off, r->bytes);
    r->bytes = 0;
    if (r->ready) (void)hipEventDestroy(r->ready);
    if (r->out) (void)hipEventDestroy(r->out);
    r->ready = r->out = nullptr;
    r->state = status;
    r->t_done = now_ms();
    to_publish.push_back(r);
}

// Drops finished requests from the active list, then makes their results
// visible to waiters.
void Server::publish() {
    active.erase(std::remove_if(active.begin(), active.end(), [](Request* r) { return r->state != 1; }), active.end());
    if (to_publish.empty()) return;
    std::lock_guard<std::mutex> lk(mu);
    for (Request* r : to_publish) {
        if (r->state == 2) {
            rays_done += r->n;
            ++requests_done;
        }
        r->status = r->state;
    }
    to_publish.clear();
    cv_done.notify_all();
}

bool Server::admit_one(Request* r) {
    size_t off;
    if (!arena.alloc(r->bytes, &off)) return false;
    r->off = off;
    char* p = arena.base() + off;
    r->d_rays = (Ray*)p;
    r->d_hits = (Hit*)(p + round_up((size_t)r->n * sizeof(Ray), GRAIN));
    r->d_occ = (uint32_t*)r->d_hits;
    hipError_t e = hipMemcpyAsync(r->d_rays, r->h_rays, (size_t)r->n * sizeof(Ray), hipMemcpyHostToDevice, s_up);
    if (e == hipSuccess && r->kind == 1) e = fill_u32(r->d_occ, 0u, ((size_t)r->n + 31) / 32, s_up);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ready, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->out, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(r->ready, s_up);
    {
        std::lock_guard<std::mutex> lk(mu);
        r->t_admit = now_ms();
        r->state = 1;
    }
    if (e != hipSuccess) finish(r, PBRTX_ERR_HIP);
    return true;
}

// The results of r leave after event `after` (its last step) on the download
// stream, without the worker waiting for the step.
void Server::start_copy_out(Request* r, hipEvent_t after) {
    hipError_t e = hipStreamWaitEvent(s_down, after, 0);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r->h_out, r->kind == 0 ? (void*)r->d_hits : (void*)r->d_occ, out_bytes(r->kind, r->n),
                           hipMemcpyDeviceToHost, s_down);
    if (e == hipSuccess) e = hipEventRecord(r->out, s_down);
    if (e != hipSuccess) {
        finish(r, PBRTX_ERR_HIP);
        return;
    }
    r->copying = true;
}

// Launches one continuous-batching step over the admitted requests whose rays
// are resident.  Returns the number of rays launched (0: nothing to launch).
int Server::launch_step() {
    std::vector<Request*> cand;
    for (Request* r : active) {
        if (r->state != 1) continue;
        if (!r->is_ready) {
            hipError_t q = hipEventQuery(r->ready);
            if (q == hipSuccess) r->is_ready = true;
            else if (q != hipErrorNotReady) { finish(r, PBRTX_ERR_HIP); continue; }
        }
        if (r->is_ready && r->launched < r->n) cand.push_back(r);
    }
    if (cand.empty()) return 0;
    const uint32_t cap = cfg.batch_rays;
    uint32_t quota = (uint32_t)std::max<uint64_t>(CHUNK_MIN, cap / cand.size());
    quota = (quota + 63) / 64 * 64;  // chunks start at multiples of 32 in a request
    // one launch per (scene, kind) over the chunks of its requests
    std::map<std::pair<uint32_t, int>, std::vector<std::pair<Request*, uint32_t>>> groups;
    uint32_t total = 0, nseg = 0;
    for (Request* r : cand) {
        if (total >= cap || nseg >= seg_cap) break;
        uint32_t take = std::min(r->n - r->launched, std::min(quota, cap - total));
        if (take < r->n - r->launched) take = take / 64 * 64;
        if (take == 0) continue;
        groups[{r->scene, r->kind}].push_back({r, take});
        total += take;
        ++nseg;
    }
    if (groups.empty()) return 0;
    // pbrtx_server_add_scene may grow the scene table meanwhile (under mu)
    std::vector<const DeviceScene*> scn;
    size_t spill;
    {
        std::lock_guard<std::mutex> lk(mu);
        scn = scenes;
        spill = spill_need;
    }
    Step st;
    if (!spare.empty()) {
        st = spare.back();
        spare.pop_back();
    } else {
        hipError_t e = hipEventCreate(&st.start);
        if (e == hipSuccess) e = hipEventCreate(&st.end);
        if (e != hipSuccess) return hip_fail(e, "step events");
    }
    // the half of the staging buffers not used by the step still in flight
    st.half = inflight.empty() ? 0 : 1 - inflight.back().half;
    TraceSeg* hs = h_segs + (size_t)st.half * seg_cap;
    TraceSeg* ds = d_segs + (size_t)st.half * seg_cap;
    const double t_launch = now_ms();
    hipError_t e = hipEventRecord(st.start, s_comp);
    if (e != hipSuccess) return hip_fail(e, "step start");
    uint32_t si = 0;
    std::vector<Request*> last;
    struct Launch { const DeviceScene* sc; bool any; uint32_t s0, nseg, rays; };
    std::vector<Launch> launches_v;
    for (auto& g : groups) {
        const DeviceScene* sc = scn[g.first.first];
        const bool any = g.first.second == 1;
        const uint32_t s0 = si;
        uint32_t start = 0;
        for (auto& rt : g.second) {
            Request* r = rt.first;
            TraceSeg& sg = hs[si++];
            sg.rays = r->d_rays + r->launched;
            sg.hits = any ? nullptr : r->d_hits + r->launched;
            sg.occ = any ? r->d_occ + r->launched / 32 : nullptr;
            sg.start = start;
            sg.count = rt.second;
            start += rt.second;
            r->launched += rt.second;
            r->launches++;
            if (r->t_first < 0) r->t_first = t_launch;
            if (r->launched == r->n) last.push_back(r);
        }
        launches_v.push_back({sc, any, s0, si - s0, start});
    }
    // one upload of the step's segment table
    e = hipMemcpyAsync(ds, hs, si * sizeof(TraceSeg), hipMemcpyHostToDevice, s_comp);
    if (e != hipSuccess) return hip_fail(e, "segment upload");
    const int nside = (int)std::min<size_t>(SIDE_STREAMS, launches_v.size() - 1);
    if (nside > 0) {
        e = hipEventRecord(ev_up, s_comp);
        for (int k = 0; k < nside && e == hipSuccess; ++k) e = hipStreamWaitEvent(s_side[k], ev_up, 0);
        if (e != hipSuccess) return hip_fail(e, "step fork");
    }
    for (size_t g = 0; g < launches_v.size(); ++g) {
        const Launch& L = launches_v[g];
        const hipStream_t stg = g % (size_t)(nside + 1) == 0 ? s_comp : s_side[g % (size_t)(nside + 1) - 1];
        const TraceTuning& tun = g_tuning[PATH_BATCHED][L.any ? 1 : 0];
        TraceScratch* scr;
        int rc = stream_scratch(stg, spill, &scr);
        if (rc) return rc;
        e = trace_segments(L.sc->view, L.any, ds + L.s0, L.nseg, L.rays, tun, scr, stg);
        if (e != hipSuccess) return hip_fail(e, "trace_segments");
        ++launches;
    }
    for (int k = 0; k < nside && e == hipSuccess; ++k) {
        e = hipEventRecord(ev_side[k], s_side[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(s_comp, ev_side[k], 0);
    }
    if (e != hipSuccess) return hip_fail(e, "step join");
    e = hipEventRecord(st.end, s_comp);
    if (e != hipSuccess) return hip_fail(e, "step end");
    for (Request* r : last) start_copy_out(r, st.end);
    inflight.push_back(st);
    ++steps;
    return (int)total;
}

// Accounts finished steps (all of them, or waits for the oldest first) and
// finishes requests whose results have arrived.
int Server::retire_steps(bool wait_oldest) {
    if (wait_oldest && !inflight.empty()) {
        hipError_t e = hipEventSynchronize(inflight.front().end);
        if (e != hipSuccess) return hip_fail(e, "step");
    }
    while (!inflight.empty()) {
        Step& st = inflight.front();
        hipError_t q = hipEventQuery(st.end);
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) return hip_fail(q, "step");
        float ms = 0;
        if (hipEventElapsedTime(&ms, st.start, st.end) == hipSuccess) busy_ms += ms;
        spare.push_back(st);
        inflight.pop_front();
    }
    for (Request* r : active) {
        if (r->state == 1 && r->copying) {
            hipError_t q = hipEventQuery(r->out);
            if (q == hipSuccess) finish(r, 2);
            else if (q != hipErrorNotReady) finish(r, PBRTX_ERR_HIP);
        }
    }
    return PBRTX_OK;
}

int Server::serve_plain(Request* r) {
    hipError_t e = hipEventSynchronize(r->ready);
    if (e != hipSuccess) return hip_fail(e, "ray upload");
    const DeviceScene* sc;
    {
        std::lock_guard<std::mutex> lk(mu);  // add_scene may grow the table
        sc = scenes[r->scene];
    }
    Step st;
    if (!spare.empty()) {
        st = spare.back();
        spare.pop_back();
    } else {
        e = hipEventCreate(&st.start);
        if (e == hipSuccess) e = hipEventCreate(&st.end);
        if (e != hipSuccess) return hip_fail(e, "step events");
    }
    r->t_first = now_ms();
    r->launches = 1;
    e = hipEventRecord(st.start, s_comp);
    if (e != hipSuccess) return hip_fail(e, "step start");
    int rc = r->kind == 0 ? run_closest(sc->view, r->d_rays, r->n, r->d_hits, PATH_PLAIN, s_comp)
                          : run_any(sc->view, r->d_rays, r->n, r->d_occ, PATH_PLAIN, s_comp);
    if (rc) return rc;
    e = hipEventRecord(st.end, s_comp);
    if (e == hipSuccess) e = hipEventSynchronize(st.end);
    if (e != hipSuccess) return hip_fail(e, "plain query");
    float ms = 0;
    if (hipEventElapsedTime(&ms, st.start, st.end) == hipSuccess) busy_ms += ms;
    spare.push_back(st);
    ++steps;
    ++launches;
    r->launched = r->n;
    start_copy_out(r, st.end);
    if (r->state != 1) return PBRTX_OK;
    e = hipEventSynchronize(r->out);
    if (e != hipSuccess) return hip_fail(e, "result copy");
    return PBRTX_OK;
}

void Server::run() {
    (void)hipSetDevice(device);
    while (true) {
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queued.empty() || !active.empty(); });
            if (stop && queued.empty() && active.empty()) break;
        }
        // admission, in arrival order, as the arena frees up
        while (true) {
            Request* r = nullptr;
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
                Request* r = active.front();
                rc = serve_plain(r);
                if (rc == PBRTX_OK && r->state == 1) finish(r, 2);
            }
        } else {
            // keep two steps in flight: launch while the previous one runs
            int launched = 0;
            if (inflight.size() < 2) {
                launched = launch_step();
                if (launched < 0) rc = launched;
            }
            if (rc == PBRTX_OK) rc = retire_steps(false);
            if (rc == PBRTX_OK && launched == 0) {
                // nothing new to launch: wait for the oldest step, else give
                // the copies a moment
                if (!inflight.empty()) rc = retire_steps(true);
                else std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
        }
        if (rc < 0) {
            // a device error ends every admitted request with that code
            std::string msg = pbrtx_last_error();
            (void)hipStreamSynchronize(s_comp);
            for (hipStream_t st : s_side) (void)hipStreamSynchronize(st);
            (void)hipStreamSynchronize(s_down);
            inflight.clear();
            for (Request* r : active)
                if (r->state == 1) finish(r, rc);
            active.clear();
            std::lock_guard<std::mutex> lk(mu);
            worker_error = msg;
        }
        publish();
    }
    publish();
    (void)hipStreamSynchronize(s_comp);
    for (hipStream_t st : s_side) (void)hipStreamSynchronize(st);
    (void)hipStreamSynchronize(s_down);
}

}  // namespace px

struct pbrtx_server {
    px::Server s;
};

extern "C" {

int pbrtx_server_create(const pbrtx_server_config* cfg, pbrtx_server** out) {
    using namespace px;
    if (!cfg || !out) return fail(PBRTX_ERR_ARG, "null argument");
    *out = nullptr;
    if (cfg->mode != PBRTX_SERVE_BATCHED && cfg->mode != PBRTX_SERVE_PLAIN) return fail(PBRTX_ERR_ARG, "unknown mode");
    if (cfg->arena_bytes < (1u << 20)) return fail(PBRTX_ERR_ARG, "arena_bytes below 1 MiB");
    if (cfg->batch_rays < 64) return fail(PBRTX_ERR_ARG, "batch_rays below 64");
    pbrtx_server* p = new (std::nothrow) pbrtx_server();
    if (!p) return fail(PBRTX_ERR_OOM, "host allocation");
    Server& s = p->s;
    s.cfg = *cfg;
    s.t0 = Clock::now();
    hipError_t e = hipGetDevice(&s.device);
    if (e == hipSuccess) e = s.arena.init(cfg->arena_bytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s.s_comp, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s.s_up, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s.s_down, hipStreamNonBlocking);
    for (int k = 0; k < SIDE_STREAMS && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&s.s_side[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.ev_side[k], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s.ev_up, hipEventDisableTiming);
    // at most one chunk per admitted request per step (a step stops taking
    // chunks when its half of the staging is full)
    s.seg_cap = 4096;
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&s.h_segs, 2 * s.seg_cap * sizeof(TraceSeg), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void**)&s.d_segs, 2 * s.seg_cap * sizeof(TraceSeg));
    if (e != hipSuccess) {
        int rc = hip_fail(e, "server setup");
        pbrtx_server_destroy(p);
        return rc;
    }
    s.worker = std::thread([&s] { s.run(); });
    *out = p;
    return PBRTX_OK;
}

int pbrtx_server_add_scene(pbrtx_server* p, const pbrtx_scene* scene, uint32_t* scene_id) {
    using namespace px;
    if (!p || !scene || !scene_id) return fail(PBRTX_ERR_ARG, "null argument");
    Server& s = p->s;
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.cfg.scene_budget_bytes && s.scene_bytes + scene->s.bytes > s.cfg.scene_budget_bytes)
        return fail(PBRTX_ERR_OOM, "scene exceeds the server's resident-scene budget");
    const TraceTuning& a = g_tuning[PATH_BATCHED][0];
    const TraceTuning& b = g_tuning[PATH_BATCHED][1];
    const size_t need = std::max(std::max(trace_spill_bytes(scene->s.view, PATH_BATCHED, a),
                                          trace_spill_bytes(scene->s.view, PATH_BATCHED, b)),
                                 std::max(trace_spill_bytes(scene->s.view, PATH_PLAIN, g_tuning[PATH_PLAIN][0]),
                                          trace_spill_bytes(scene->s.view, PATH_PLAIN, g_tuning[PATH_PLAIN][1])));
    s.spill_need = std::max(s.spill_need, need);
    s.scenes.push_back(&scene->s);
    s.scene_bytes += scene->s.bytes;
    *scene_id = (uint32_t)(s.scenes.size() - 1);
    return PBRTX_OK;
}

int pbrtx_server_submit(pbrtx_server* p, uint32_t scene_id, int32_t kind, const pbrtx_ray* rays, uint32_t n,
                        void* out, uint64_t* ticket) {
    using namespace px;
    if (!p || !ticket) return fail(PBRTX_ERR_ARG, "null argument");
    if (kind != 0 && kind != 1) return fail(PBRTX_ERR_ARG, "kind must be 0 (closest hit) or 1 (occlusion)");
    if (n == 0 || !rays || !out) return fail(PBRTX_ERR_ARG, "empty request or null buffer");
    Server& s = p->s;
    auto r = std::make_unique<Request>();
    r->scene = scene_id;
    r->kind = kind;
    r->n = n;
    r->h_rays = rays;
    r->h_out = out;
    r->bytes = request_bytes(kind, n);
    if (r->bytes > s.arena.capacity()) return fail(PBRTX_ERR_OOM, "request larger than the server's arena");
    std::lock_guard<std::mutex> lk(s.mu);
    if (scene_id >= s.scenes.size()) return fail(PBRTX_ERR_ARG, "unknown scene id");
    if (s.stop) return fail(PBRTX_ERR_STATE, "server is shutting down");
    r->ticket = s.next_ticket++;
    r->t_submit = s.now_ms();
    *ticket = r->ticket;
    s.queued.push_back(r.get());
    s.all.emplace(r->ticket, std::move(r));
    s.cv_work.notify_one();
    return PBRTX_OK;
}

int pbrtx_server_wait(pbrtx_server* p, uint64_t ticket, pbrtx_request_stats* st) {
    using namespace px;
    if (!p) return fail(PBRTX_ERR_ARG, "null server");
    Server& s = p->s;
    std::unique_lock<std::mutex> lk(s.mu);
    auto it = s.all.find(ticket);
    if (it == s.all.end()) return fail(PBRTX_ERR_ARG, "unknown ticket");
    Request* r = it->second.get();
    s.cv_done.wait(lk, [&] { return r->status == 2 || r->status < 0; });
    if (st) {
        st->n_rays = r->n;
        st->status = r->status == 2 ? PBRTX_OK : r->status;
        st->launches = r->launches;
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

int pbrtx_server_stats(pbrtx_server* p, pbrtx_server_counters* st) {
    using namespace px;
    if (!p || !st) return fail(PBRTX_ERR_ARG, "null argument");
    Server& s = p->s;
    std::lock_guard<std::mutex> lk(s.mu);
    st->rays_done = s.rays_done;
    st->requests_done = s.requests_done;
    st->steps = s.steps;
    st->launches = s.launches;
    st->busy_ms = s.busy_ms;
    st->elapsed_ms = s.now_ms();
    st->arena_bytes = s.arena.capacity();
    st->scene_bytes = s.scene_bytes;
    return PBRTX_OK;
}

int pbrtx_server_destroy(pbrtx_server* p) {
    using namespace px;
    if (!p) return PBRTX_OK;
    Server& s = p->s;
    {
        std::lock_guard<std::mutex> lk(s.mu);
        s.stop = true;
        s.cv_work.notify_all();
    }
    if (s.worker.joinable()) s.worker.join();
    std::vector<hipStream_t> streams = {s.s_comp, s.s_up, s.s_down};
    for (hipStream_t st : s.s_side) streams.push_back(st);
    for (hipStream_t st : streams)
        if (st) (void)hipStreamSynchronize(st);
    for (Step& st : s.spare) {
        (void)hipEventDestroy(st.start);
        (void)hipEventDestroy(st.end);
    }
    if (s.ev_up) (void)hipEventDestroy(s.ev_up);
    for (hipEvent_t ev : s.ev_side)
        if (ev) (void)hipEventDestroy(ev);
    if (s.h_segs) (void)hipHostFree(s.h_segs);
    if (s.d_segs) (void)hipFree(s.d_segs);
    for (hipStream_t st : streams)
        if (st) {
            release_stream_scratch(st);
            (void)hipStreamDestroy(st);
        }
    delete p;
    return PBRTX_OK;
}

}  // extern "C"
