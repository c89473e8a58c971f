// render.hip -- SamplerRenderer::render on gfx950 (sampler_renderer.rs:61-206).
//
// Wavefront organisation: every stage is a kernel over a batch of samples or
// paths, ray queries go through the traversal kernels of trace.hip.
//   direct lighting: camera rays -> closest hit -> shade (light samples,
//     shadow segments, contributions) -> occlusion -> accumulate -> film
//   path tracing: camera rays, then per bounce closest hit -> shade (one
//     light sample + BSDF sample + roulette) -> occlusion -> accumulate.
// PATH_PLAIN keeps one slot per sample / path and traces every slot each
// stage (dead slots carry an empty interval); PATH_BATCHED compacts live
// shadow and extension rays into dense queues with wave-aggregated atomics
// (ballot + prefix rank) and traces them with the persistent kernel, whose
// ray count is read from device memory.
//
// The arithmetic of every stage is that of the CPU restatement: the same
// IEEE operations in the same order (no contraction), the same
// counter-based sample values, so images match sample for sample.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "render.hpp"

namespace px {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr int RB = 256;  // render kernel block

__device__ __forceinline__ f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }

__device__ __forceinline__ bool black3(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ __forceinline__ f3 mulc(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 divs3(f3 a, float f) { return mk3(a.x / f, a.y / f, a.z / f); }

// One light sample at p (normal n, outgoing wo, material mat): the shadow
// segment and its contribution f*Li*|wi.n|/pdf; false when nothing would be
// traced.  PointLight::sample_l (point.rs:28-35; wi normalised), the quad area
// light (uniform by area, pdf d^2/(cos*A)), VisibilityTester::segment
// (visibility_tester.rs:16-24), BSDF::f with pbrt's flag test (bsdf/mod.rs:132-149).
__device__ bool sample_light(const RenderParams& P, uint32_t li, f3 p, float eps, f3 n, f3 wo,
                             const MatRec& mat, float u0, float u1, float u2, Ray* out, f3* c) {
    f3 Li, wi, target;
    float pdf, eps2;
    if (li < P.n_point) {
        const float* q = P.points + 6 * li;
        f3 pos = ld3(q);
        f3 w = sub3(pos, p);
        float d2 = lensq3(w);
        Li = divs3(ld3(q + 3), d2);
        wi = normalize3(w);
        pdf = 1.0f;
        target = pos;
        eps2 = 0.0f;
    } else {
        const uint32_t la = li - P.n_point;
        const uint32_t first = P.a_first[la], cnt = P.a_count[la];
        const float total = P.a_total[la];
        float tgt = u0 * total;
        uint32_t k = 0;
        float acc = P.atris[first].area;
        while (k + 1 < cnt && tgt >= acc) { ++k; acc = acc + P.atris[first + k].area; }
        const AreaTriRec& t = P.atris[first + k];
        float su1 = sqrtf(u1);
        float b1 = 1.0f - su1;
        float b2 = u2 * su1;
        float b0 = (1.0f - b1) - b2;
        f3 ps = add3(add3(mul3(ld3(t.p1), b1), mul3(ld3(t.p2), b2)), mul3(ld3(t.p3), b0));
        f3 w = sub3(ps, p);
        float d2 = lensq3(w);
        float dist = sqrtf(d2);
        wi = mul3(w, 1.0f / dist);
        float cosl = dot3(ld3(t.nn), neg3(wi));
        if (!(cosl > 0.0f)) return false;
        Li = ld3(P.a_le + 3 * la);
        pdf = d2 / (cosl * total);
        target = ps;
        eps2 = 1e-3f;
    }
    if (black3(Li) || pdf == 0.0f) return false;
    f3 f = mk3(0.0f, 0.0f, 0.0f);
    if (mat.kind == 0u && dot3(wi, n) * dot3(wo, n) > 0.0f) f = mul3(mk3(mat.r, mat.g, mat.b), 1.0f / PI_F);
    if (black3(f)) return false;
    float dist = sqrtf(lensq3(sub3(p, target)));
    f3 dir = mul3(sub3(target, p), 1.0f / dist);
    *out = Ray{p.x, p.y, p.z, eps, dir.x, dir.y, dir.z, (1.0f - eps2) * dist};
    *c = divs3(mul3(mulc(f, Li), fabsf(dot3(wi, n))), pdf);
    return true;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ unsigned long long lt_mask() {
    uint32_t l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}
// wave-aggregated slot allocation: every lane of the wave calls it
__device__ __forceinline__ uint32_t wave_alloc(bool want, uint32_t* counter) {
    const unsigned long long m = __ballot(want);
    if (!m) return NONE;
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, 0);
    return want ? base + (uint32_t)__popcll(m & lt_mask()) : NONE;
}

__device__ __forceinline__ void set_dead(Ray* r) { *r = Ray{0, 0, 0, 1.0f, 0, 0, 1.0f, 0.0f}; }

// The shading geometry of a hit: the point, dg.nn, the BSDF frame's tn =
// normalize(dpdu) (bsdf/mod.rs:70-73), the material and the emitting area
// light (-1 for none).  A face's come precomputed (FaceShade; the point is
// ray.point_at(t), ray.rs:60).  A sphere's (Sphere::intersect,
// sphere.rs:135-175, and compute_dg, helpers.rs:17-43): the object-space hit
// of the ray at t, theta = acos(clamp(z / r)) and sin(theta) from the
// device's acosf and sinf (the reference's come from the host C library, so
// these may differ in the last place: sphere shading matches the reference
// to a tolerance, not bit for bit), dp/du and dp/dv, the world point
// o2w(p_hit) and dg.nn = normalize(o2w(dpdu) x o2w(dpdv)), flipped when
// reverse_orientation ^ transform_swaps_handedness (diff_geom.rs:52-60).
struct HitShade {
    f3 p, nn, tn;
    uint32_t mat;
    int32_t light;
};
__device__ __forceinline__ void sphere_shade(const SphereShade& s, const Ray& r, float t, HitShade* out) {
    const f3 o = xf_point(s.w2o, mk3(r.ox, r.oy, r.oz));
    const f3 d = xf_vector(s.w2o, mk3(r.dx, r.dy, r.dz));
    const f3 ph = add3(o, mul3(d, t));
    float c = ph.z / s.radius;  // f32::clamp(-1, 1), a NaN kept
    c = c < -1.0f ? -1.0f : (c > 1.0f ? 1.0f : c);
    const float theta = acosf(c);
    const float zradius = sqrtf(ph.x * ph.x + ph.y * ph.y);
    const float inv_z = 1.0f / zradius;
    const float cos_phi = ph.x * inv_z, sin_phi = ph.y * inv_z;
    const f3 dpdu = mk3(-s.phimax * ph.y, s.phimax * ph.x, 0.0f);
    const f3 dpdv = mul3(mk3(ph.z * cos_phi, ph.z * sin_phi, -s.radius * sinf(theta)), s.thetamax - s.thetamin);
    out->p = xf_point(s.o2w, ph);
    const f3 du = xf_vector(s.o2w, dpdu), dv = xf_vector(s.o2w, dpdv);
    f3 n = normalize3(cross3(du, dv));
    if (s.flip) n = mul3(n, -1.0f);
    out->nn = n;
    out->tn = normalize3(du);
    out->mat = s.mat;
    out->light = -1;
}
__device__ __forceinline__ HitShade hit_shade(const RenderParams& P, const Ray& r, const Hit& h) {
    HitShade out;
    if (h.prim >= P.nfaces) {
        sphere_shade(P.sph[h.prim - P.nfaces], r, h.t, &out);
        return out;
    }
    const FaceShade fs = P.faces[h.prim];
    out.p = add3(mk3(r.ox, r.oy, r.oz), mul3(mk3(r.dx, r.dy, r.dz), h.t));  // ray.rs:60
    out.nn = ld3(fs.nn);
    out.tn = ld3(fs.tn);
    out.mat = fs.mat;
    out.light = fs.light;
    return out;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(RB) void k_camera(RenderParams P, uint64_t g0, uint32_t n, Ray* rays) {
    for (uint32_t i = blockIdx.x * RB + threadIdx.x; i < n; i += gridDim.x * RB) {
        Ray r;
        camera_ray_of(P, g0 + i, &r);
        rays[i] = r;
    }
}

// Direct lighting shade (deviation D1 of the build: emitted light plus, per
// light, the mean over its samples of the unoccluded contributions).
// Slots per sample: point lights first, then light_samples per area light.
// Exclusive prefix sum of v over the RB threads of a block (wave shuffles,
// then one LDS round for the four wave totals).
template <int NB = RB>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[NB / 64];
    const uint32_t lane = __lane_id(), wid = threadIdx.x / 64;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (uint32_t w = 0; w < NB / 64; ++w) {
        const uint32_t s = wsum[w];
        if (w < wid) before += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return before + x - v;
}


// Shade modes: SHADE_PLAIN writes every slot in place (dead slots carry an
// empty interval) and counts live rays with one atomic per block;
// SHADE_QUEUE writes the live rays densely: each block takes a run of the
// queue for its live rays with one atomic on the count, in whatever order
// the blocks get there, and orders its rays within the run.  A ray's result
// does not depend on where it sits in the queue (the accumulation finds it
// through qidx), so no block waits on another.  (A single-pass scan over the
// blocks in block order, with decoupled look-back, had taken the
// 1024x1024 direct-lighting frame's shading stage to 0.85 ms: the block,
// alone on its CU, waited for its predecessors' prefixes.)  With at most
// SH_CACHE slots per sample the light samples are kept in registers between
// the count and the write.
enum ShadeMode : int { SHADE_PLAIN = 0, SHADE_QUEUE = 1 };
// The path tracer's vertex kernel: PSHADE_PLAIN writes every path's slot,
// PSHADE_QUEUE fills two dense queues (shadow and next rays), a run of each
// per tile.
enum PathShadeMode : int { PSHADE_PLAIN = 0, PSHADE_QUEUE = 1 };
constexpr uint32_t SH_CACHE = 4;

// Threads per block of the shade kernels: a block is one tile of queue
// output, and every tile costs atomics on shared counters (with block-order
// scans, 256 / 512 / 1024 threads: direct-lighting frame 8.63 / 8.41 / 8.32
// ms, path-tracing frame 24.3 / 23.0 / 22.6 ms).
constexpr int SHADE_B = 512;
constexpr int VERTEX_B = 1024;

// list (queue form, with its count at n_dev): the samples to shade, those
// whose camera rays hit (the others ended in the miss pass); entry q is
// sample list[q], and lbase, qidx and the accumulation are by entry.
__global__ __launch_bounds__(SHADE_B) void k_shade_direct(RenderParams P, uint64_t g0, uint32_t n_host, const Ray* rays,
                                                     const Hit* hits, Ray* sh, float4* contrib,
                                                     uint32_t* qidx, float4* lbase, uint32_t* sh_count, int mode,
                                                     const uint32_t* list, const uint32_t* n_dev) {
    const uint32_t S = P.S;
    const uint32_t n = n_dev ? *n_dev : n_host;
    __shared__ uint32_t s_excl;
    const uint32_t bid = blockIdx.x;
    // a launch sized for every sample: the blocks past the list's count have
    // nothing to do
    if ((uint64_t)bid * SHADE_B >= n) return;
    const uint32_t ent = bid * SHADE_B + threadIdx.x;  // entry
    const bool live = ent < n;
    const uint32_t i = live ? (list ? list[ent] : ent) : 0u;  // its sample
    Hit h = live ? hits[i] : Hit{NO_HIT, 0, 0, 0};
    const bool hit = live && h.prim != NO_HIT;
    f3 p = mk3(0, 0, 0), nn = mk3(0, 0, 0), wo = mk3(0, 0, 0);
    float eps = 0.0f;
    MatRec mat{1u, 0, 0, 0};
    uint32_t pixel = 0, s = 0;
    if (hit) {
        // the camera ray: read, or made again as the traversal made it
        Ray r;
        if (rays) r = rays[i];
        else camera_ray_of(P, g0 + i, &r);
        const HitShade hs = hit_shade(P, r, h);
        p = hs.p;
        nn = hs.nn;
        wo = neg3(mk3(r.dx, r.dy, r.dz));
        eps = h.t * 5e-4f;  // mesh.rs:262, sphere.rs:174
        mat = P.mats[hs.mat];
        f3 le = mk3(0, 0, 0);
        if (hs.light >= 0 && dot3(nn, wo) > 0.0f) le = ld3(P.a_le + 3 * hs.light);
        lbase[ent] = make_float4(le.x, le.y, le.z, 0.0f);
        uint64_t g = g0 + i;
        pixel = (uint32_t)(g / P.spp); s = (uint32_t)(g % P.spp);
    } else if (live) {
        lbase[ent] = make_float4(0, 0, 0, 0);
    }
    auto slot_sample = [&](uint32_t k, Ray* sr, f3* c) -> bool {
        if (!hit) return false;
        uint32_t li, j;
        if (k < P.n_point) { li = k; j = 0; }
        else { li = P.n_point + (k - P.n_point) / P.light_samples; j = (k - P.n_point) % P.light_samples; }
        float u[4] = {0, 0, 0, 0};
        if (li >= P.n_point) rng4(P.seed, pixel, s, 1 + (li - P.n_point) * P.light_samples + j, u);
        return sample_light(P, li, p, eps, nn, wo, mat, u[0], u[1], u[2], sr, c);
    };
    if (mode == SHADE_PLAIN) {
        uint32_t live_rays = 0;
        for (uint32_t k = 0; k < S; ++k) {
            Ray sr;
            f3 c = mk3(0, 0, 0);
            const bool valid = slot_sample(k, &sr, &c);
            const size_t slot = (size_t)ent * S + k;
            if (!valid) set_dead(&sr);
            if (live) {
                sh[slot] = sr;
                qidx[(size_t)k * n + ent] = valid ? (uint32_t)slot : NONE;
            }
            if (valid) contrib[slot] = make_float4(c.x, c.y, c.z, 0.0f);
            live_rays += valid ? 1u : 0u;
        }
        uint32_t tot;
        block_exclusive_scan<SHADE_B>(live_rays, &tot);
        if (threadIdx.x == 0 && tot) atomicAdd(sh_count, tot);
        return;
    }
    // SHADE_QUEUE: count (keeping the samples when they fit), scan, write.
    // Within a block's part of the queue the rays are ordered by (light
    // sample k, sample), so the stores of one k are one contiguous run across
    // the lanes; consecutive rays still start in neighbouring pixels.  The
    // queue index of slot k of sample i is stored at k*n + i (contiguous per
    // wave as well); the contribution of a traced ray sits at its queue
    // position.
    // A shadow ray starts at the shading point with tmin = eps for every
    // slot: only its direction and tmax are kept, with the contribution.
    const bool cached = S <= SH_CACHE;
    f3 cd[SH_CACHE], cc[SH_CACHE];
    float ct[SH_CACHE];
    uint32_t vm = 0, cnt = 0;
    for (uint32_t k = 0; k < S; ++k) {
        Ray sr{};
        f3 c = mk3(0, 0, 0);
        const bool valid = slot_sample(k, &sr, &c);
        if (valid) {
            ++cnt;
            vm |= 1u << (k & 31u);
        }
        if (cached) {
#pragma unroll
            for (uint32_t kk = 0; kk < SH_CACHE; ++kk)
                if (kk == k) { cd[kk] = mk3(sr.dx, sr.dy, sr.dz); ct[kk] = sr.tmax; cc[kk] = c; }
        }
    }
    // rank of each lane's slot k among the block's valid slots k, and the
    // block's count of valid slots k (wave ballots, then the four waves)
    __shared__ uint32_t s_cnt[SH_CACHE][SHADE_B / 64];
    const uint32_t wid = threadIdx.x / 64;
    const unsigned long long ltm = lt_mask();
    uint32_t before[SH_CACHE];
    if (cached) {
#pragma unroll
        for (uint32_t k = 0; k < SH_CACHE; ++k) {
            const unsigned long long m = __ballot(k < S && ((vm >> k) & 1u));
            before[k] = (uint32_t)__popcll(m & ltm);
            if (__lane_id() == 0) s_cnt[k][wid] = (uint32_t)__popcll(m);
        }
    }
    uint32_t tot;
    const uint32_t local = block_exclusive_scan<SHADE_B>(cnt, &tot);  // also orders the s_cnt writes
    uint32_t run = 0;  // valid slots of earlier k in this block
    if (cached) {
#pragma unroll
        for (uint32_t k = 0; k < SH_CACHE; ++k) {
            uint32_t wb = 0, bt = 0;
            for (uint32_t w2 = 0; w2 < SHADE_B / 64; ++w2) {
                const uint32_t c = s_cnt[k][w2];
                if (w2 < wid) wb += c;
                bt += c;
            }
            before[k] += wb + run;
            run += bt;
        }
    }
    if (threadIdx.x == 0) s_excl = tot ? atomicAdd(sh_count, tot) : 0u;  // the block's run of the queue
    __syncthreads();
    uint32_t pos = s_excl + local;  // sample-major order when the slots are not kept
    for (uint32_t k = 0; k < S; ++k) {
        Ray sr;
        f3 c = mk3(0, 0, 0);
        bool valid;
        if (cached) {
            valid = (vm >> k) & 1u;
#pragma unroll
            for (uint32_t kk = 0; kk < SH_CACHE; ++kk)
                if (kk == k) { sr = Ray{p.x, p.y, p.z, eps, cd[kk].x, cd[kk].y, cd[kk].z, ct[kk]}; c = cc[kk]; }
        } else {
            valid = slot_sample(k, &sr, &c);
        }
        uint32_t q = NONE;
        if (valid) {
            if (cached) {
                uint32_t b = 0;
#pragma unroll
                for (uint32_t kk = 0; kk < SH_CACHE; ++kk)
                    if (kk == k) b = before[kk];
                q = s_excl + b;
            } else {
                q = pos++;
            }
            sh[q] = sr;
            contrib[q] = make_float4(c.x, c.y, c.z, 0.0f);
        }
        if (live) qidx[(size_t)k * n + ent] = q;
    }
}

// list (with its count at n_dev): entry i is sample list[i] (the shade
// kernel's entries); otherwise entry i is sample i.
__global__ __launch_bounds__(RB) void k_accum_direct(RenderParams P, uint64_t g0, uint32_t n_host, const float4* lbase,
                                                     const float4* contrib, const uint32_t* qidx,
                                                     const uint8_t* occ, float4* sample_l, const uint32_t* list,
                                                     const uint32_t* n_dev) {
    const uint32_t n = n_dev ? *n_dev : n_host;
    for (uint32_t i = blockIdx.x * RB + threadIdx.x; i < n; i += gridDim.x * RB) {
        const float4 lb = lbase[i];
        f3 L = mk3(lb.x, lb.y, lb.z);
        // with at most SH_CACHE slots, every slot's queue index, then its
        // occlusion byte, then its contribution are loaded before any is
        // used (three rounds of independent loads instead of a chain per
        // slot); the sums run in the same order either way
        float4 cv[SH_CACHE];
        if (P.S <= SH_CACHE) {
            uint32_t q[SH_CACHE];
#pragma unroll
            for (uint32_t k = 0; k < SH_CACHE; ++k) q[k] = k < P.S ? qidx[(size_t)k * n + i] : NONE;
            bool lit[SH_CACHE];
#pragma unroll
            for (uint32_t k = 0; k < SH_CACHE; ++k) lit[k] = q[k] != NONE && !occ[q[k]];
#pragma unroll
            for (uint32_t k = 0; k < SH_CACHE; ++k) cv[k] = lit[k] ? contrib[q[k]] : make_float4(0, 0, 0, -1.0f);
        }
        uint32_t k = 0;
        for (uint32_t li = 0; li < P.nlights; ++li) {
            const uint32_t ns = li < P.n_point ? 1u : P.light_samples;
            f3 Ld = mk3(0, 0, 0);
            for (uint32_t j = 0; j < ns; ++j, ++k) {
                float4 c;
                if (P.S <= SH_CACHE) {
                    c = cv[0];
#pragma unroll
                    for (uint32_t kk = 1; kk < SH_CACHE; ++kk)
                        if (kk == k) c = cv[kk];
                    if (c.w < 0.0f) continue;  // no ray, or occluded
                } else {
                    const uint32_t q = qidx[(size_t)k * n + i];
                    if (q == NONE) continue;
                    if (occ[q]) continue;  // occluded
                    c = contrib[q];
                }
                Ld = add3(Ld, mk3(c.x, c.y, c.z));
            }
            L = add3(L, divs3(Ld, (float)ns));
        }
        sample_l[g0 + (list ? list[i] : i)] = make_float4(L.x, L.y, L.z, 0.0f);
    }
}

// hi:lo += *add (a 64-bit running total in two words); then (when given) *add and *clr are cleared for the next
// chunk, which saves that chunk a clearing launch.
__global__ void k_add_counter(uint32_t* lo, uint32_t* add, uint32_t* hi, bool clear_add, uint32_t* clr) {
    uint64_t v = ((uint64_t)*hi << 32) | (uint64_t)*lo;
    v += *add;
    *lo = (uint32_t)v;
    *hi = (uint32_t)(v >> 32);
    if (clear_add) *add = 0u;
    if (clr) *clr = 0u;
}

// ---- path tracing ------------------------------------------------------------
// Where a path's sample radiance goes and what keys its random numbers:
// sample g0 + gid of the frame, or -- in a served batch (P.segs) -- the
// sample of the request that the path's segment (flags bits 8-15) maps
// batch index gid to.
struct SampleRef {
    uint64_t g;
    uint32_t spp, seed;
    float4* out;
};
__device__ __forceinline__ SampleRef sample_ref(const RenderParams& P, uint64_t g0, const PathState& ps,
                                                float4* sample_l) {
    if (P.segs) {
        const RenderSeg sg = P.segs[ps.flags >> 8];
        const ReqView* rq = P.reqs + sg.req;
        const uint64_t g = sg.g0 + (ps.gid - sg.start);
        return SampleRef{g, rq->cam.spp, rq->cam.seed, rq->sample_l + g};
    }
    const uint64_t g = g0 + ps.gid;
    return SampleRef{g, P.spp, P.seed, sample_l + g};
}

// Camera rays and initial path state of the batch's paths.
__global__ __launch_bounds__(RB) void k_path_init(RenderParams P, uint64_t g0, uint32_t n, Ray* rays,
                                                  PathState* st, uint32_t* pid_q, uint32_t* q_count) {
    for (uint32_t i = blockIdx.x * RB + threadIdx.x; i < n; i += gridDim.x * RB) {
        Ray r;
        uint32_t flags = 0;
        if (P.segs) {
            // the segment holding batch index i: the last one starting at or before it
            uint32_t lo = 0, hi = P.nsegs - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (P.segs[mid].start <= i) lo = mid;
                else hi = mid - 1;
            }
            const RenderSeg sg = P.segs[lo];
            camera_ray_of(P.reqs[sg.req].cam, sg.g0 + (i - sg.start), &r);
            flags = lo << 8;
        } else {
            camera_ray_of(P, g0 + i, &r);
        }
        rays[i] = r;
        st[i] = PathState{{1.0f, 1.0f, 1.0f}, {0.0f, 0.0f, 0.0f}, i, flags};
        if (pid_q) pid_q[i] = i;
    }
    if (q_count && blockIdx.x == 0 && threadIdx.x == 0) *q_count = n;
}

// One path vertex (PathIntegrator, deviation D1): emission on the first or a
// specular bounce, one light sample (no MIS), BSDF sample, roulette.  Path
// i of the bounce: queue form, the path pid_q[i] with ray rays_in[i] and hit
// hits[i]; plain form (pid_q unused), path i.  fresh (rays_in null): the
// first vertex of new paths whose camera rays the traversal made itself --
// path i, initial state, the camera ray made again here.  Writes the path's
// state (and its radiance when it ends with no shadow ray pending) and
// returns the rays it wants traced.
struct VertexOut {
    Ray sray, nray;
    f3 contrib;
    uint32_t pid;
    bool shadow, next;
};
__device__ __forceinline__ void path_vertex(const RenderParams& P, uint64_t g0, uint32_t bounce, uint32_t i,
                                            bool queued, bool fresh, const uint32_t* pid_q, const Ray* rays_in,
                                            const Hit* hits, PathState* st, float4* sample_l, VertexOut& o) {
    // queue form: the path of queue entry i; fresh paths come from the list
    // of camera rays that hit (pid_q), their hit records at the path index
    const uint32_t pid = queued ? pid_q[i] : i;
    PathState ps;
    if (fresh) ps = PathState{{1.0f, 1.0f, 1.0f}, {0.0f, 0.0f, 0.0f}, pid, 0u};
    else ps = st[pid];
    const bool active = queued || !(ps.flags & 4u);  // plain form: bit 4 marks a finished path
    Hit h = active ? hits[fresh ? pid : i] : Hit{NO_HIT, 0, 0, 0};
    bool want_shadow = false, want_next = false, finish = false;
    Ray& sray = o.sray;
    Ray& nray = o.nray;
    f3 contrib = mk3(0, 0, 0);
    if (active && h.prim == NO_HIT) {
        finish = true;  // a miss ends the path (no infinite lights)
    } else if (active) {
        Ray r;
        if (fresh) camera_ray_of(P, g0 + pid, &r);
        else r = rays_in[i];
        const f3 d = mk3(r.dx, r.dy, r.dz);
        const HitShade hs = hit_shade(P, r, h);
        const f3 p = hs.p;
        const f3 nn = hs.nn;
        const f3 wo = neg3(d);
        const float eps = h.t * 5e-4f;
        f3 beta = ld3(ps.beta), L = ld3(ps.L);
        const bool specular = (ps.flags & 1u) != 0u;
        if (bounce == 0 || specular) {
            f3 le = mk3(0, 0, 0);
            if (hs.light >= 0 && dot3(nn, wo) > 0.0f) le = ld3(P.a_le + 3 * hs.light);
            L = add3(L, mulc(beta, le));
        }
        const MatRec mat = P.mats[hs.mat];
        const SampleRef sr = sample_ref(P, g0, ps, sample_l);
        const uint32_t pixel = (uint32_t)(sr.g / sr.spp), s = (uint32_t)(sr.g % sr.spp);
        float ua[4], ub[4];
        rng4(sr.seed, pixel, s, 1 + 2 * bounce, ua);
        rng4(sr.seed, pixel, s, 2 + 2 * bounce, ub);
        if (P.nlights > 0 && mat.kind == 0u) {
            uint32_t ln = (uint32_t)(ua[0] * (float)P.nlights);
            if (ln > P.nlights - 1) ln = P.nlights - 1;
            f3 c;
            if (sample_light(P, ln, p, eps, nn, wo, mat, ua[1], ua[2], ua[3], &sray, &c)) {
                want_shadow = true;
                contrib = mulc(beta, mul3(c, (float)P.nlights));
            }
        }
        bool cont = bounce != P.max_depth;
        if (cont) {
            // BSDF frame (bsdf/mod.rs:69-84) and world_to_local (:103-105)
            const f3 tn = hs.tn;
            const f3 sn = cross3(nn, tn);
            const f3 wol = mk3(dot3(wo, sn), dot3(wo, tn), dot3(wo, nn));
            f3 wil, f;
            float pdf;
            bool spec;
            if (mat.kind == 0u) {
                float dx, dy;
                concentric_disk(ub[0], ub[1], &dx, &dy);
                float z = sqrtf(fmaxf(0.0f, (1.0f - dx * dx) - dy * dy));
                wil = mk3(dx, dy, z);
                if (wol.z < 0.0f) wil.z = -wil.z;
                pdf = (wol.z * wil.z > 0.0f) ? fabsf(wil.z) * (1.0f / PI_F) : 0.0f;
                f = mul3(mk3(mat.r, mat.g, mat.b), 1.0f / PI_F);
                spec = false;
            } else {
                // SpecularReflection::sample_f (specular.rs:34-40), Fresnel::noop
                wil = mk3(-wol.x, -wol.y, wol.z);
                pdf = 1.0f;
                f = divs3(mk3(mat.r, mat.g, mat.b), fabsf(wil.z));
                spec = true;
            }
            if (black3(f) || pdf == 0.0f) {
                cont = false;
            } else {
                // local_to_world: bsdf/mod.rs:107-111
                const f3 wi = mk3(sn.x * wil.x + tn.x * wil.y + nn.x * wil.z,
                                  sn.y * wil.x + tn.y * wil.y + nn.y * wil.z,
                                  sn.z * wil.x + tn.z * wil.y + nn.z * wil.z);
                ps.flags = spec ? (ps.flags | 1u) : (ps.flags & ~1u);
                beta = mulc(beta, divs3(mul3(f, fabsf(dot3(wi, nn))), pdf));
                if (bounce > P.rr_start) {
                    const float y = 0.212671f * beta.x + 0.715160f * beta.y + 0.072169f * beta.z;
                    const float q = (y < 0.5f) ? y : 0.5f;
                    if (ub[2] > q) cont = false;
                    else beta = divs3(beta, q);
                }
                if (cont) nray = Ray{p.x, p.y, p.z, eps, wi.x, wi.y, wi.z, FLT_MAX};
            }
        }
        ps.beta[0] = beta.x; ps.beta[1] = beta.y; ps.beta[2] = beta.z;
        ps.L[0] = L.x; ps.L[1] = L.y; ps.L[2] = L.z;
        want_next = cont;
        finish = !cont;
    }
    // paths that end with no shadow ray pending are final now
    if (active) {
        bool keep = true;
        if (finish && !want_shadow) {
            *sample_ref(P, g0, ps, sample_l).out = make_float4(ps.L[0], ps.L[1], ps.L[2], 0.0f);
            ps.flags |= 4u;
            keep = !queued;  // queue form: a finished path is in no queue, its state is not read again
        } else if (finish) {
            ps.flags |= 2u;
        }
        if (keep) st[pid] = ps;
    }
    o.contrib = contrib;
    o.pid = pid;
    o.shadow = want_shadow;
    o.next = want_next;
}

// Bounce 0 of fused batches: a camera ray that missed ends its path with
// zero radiance (no infinite lights), written here; the paths whose rays hit
// are listed densely, and only they go through the vertex kernel.  A block
// takes a tile of MISS_B * MISS_PER paths; round k of it reads paths tile +
// k*MISS_B + thread (coalesced).  The tile's hits take a run of the list
// with one atomic on the count (tiles in whatever order they get there), in
// path order within the run: round-major, then wave, then lane.  SEG (a served batch of several
// requests): a path's radiance goes to its segment's request, and a hit's
// path starts here -- its state (segment in the flags), its camera ray and
// its hit record stored at its queue position, as the vertex kernel reads a
// queue.
constexpr uint32_t MISS_B = 1024, MISS_PER = 8;
template <bool SEG>
__global__ __launch_bounds__(MISS_B) void k_path_miss0(RenderParams P, uint64_t g0, uint32_t n, const Hit* hits,
                                                       float4* sample_l, uint32_t* list, uint32_t* count,
                                                       Ray* rays_q, Hit* hits_q, PathState* st) {
    constexpr uint32_t W = MISS_B / 64;
    __shared__ uint32_t s_excl;
    __shared__ uint32_t s_pre[MISS_PER * W];  // hits before (round, wave) in path order
    const uint32_t base = blockIdx.x * (MISS_B * MISS_PER);
    const uint32_t wid = threadIdx.x / 64, lane = __lane_id();
    const unsigned long long ltm = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const CamSeg* segs = reinterpret_cast<const CamSeg*>(P.segs);
    unsigned long long m[MISS_PER];
#pragma unroll
    for (uint32_t k = 0; k < MISS_PER; ++k) {
        const uint32_t i = base + k * MISS_B + threadIdx.x;
        bool hit = false;
        if (i < n) {
            hit = hits[i].prim != NO_HIT;
            if (!hit) {
                float4* out;
                if (SEG) {
                    const CamSeg sg = segs[cam_seg_of(segs, P.nsegs, i)];
                    out = P.reqs[sg.cam].sample_l + sg.g0 + (i - sg.start);
                } else {
                    out = sample_l + g0 + i;
                }
                *out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        }
        m[k] = __ballot(hit);
        if (lane == 0) s_pre[k * W + wid] = (uint32_t)__popcll(m[k]);
    }
    __syncthreads();
    // exclusive prefix over the (round, wave) counts, by the first wave: two
    // entries a lane
    if (threadIdx.x < 64) {
        static_assert(MISS_PER * W == 128, "two entries a lane");
        const uint32_t a = s_pre[2 * lane], b = s_pre[2 * lane + 1];
        uint32_t x = a + b;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= (uint32_t)off) x += y;
        }
        const uint32_t tot = __shfl(x, 63);
        const uint32_t ex = x - (a + b);
        s_pre[2 * lane] = ex;
        s_pre[2 * lane + 1] = ex + a;
        if (lane == 0) s_excl = tot ? atomicAdd(count, tot) : 0u;  // the tile's run of the list
    }
    __syncthreads();
    const uint32_t e = s_excl;
#pragma unroll
    for (uint32_t k = 0; k < MISS_PER; ++k)
        if ((m[k] >> lane) & 1ull) {
            const uint32_t i = base + k * MISS_B + threadIdx.x;
            const uint32_t q = e + s_pre[k * W + wid] + (uint32_t)__popcll(m[k] & ltm);
            list[q] = i;
            if (SEG) {
                const uint32_t s = cam_seg_of(segs, P.nsegs, i);
                const CamSeg sg = segs[s];
                Ray r;
                camera_ray_of(P.reqs[sg.cam].cam, sg.g0 + (i - sg.start), &r);
                rays_q[q] = r;
                hits_q[q] = hits[i];
                st[i] = PathState{{1.0f, 1.0f, 1.0f}, {0.0f, 0.0f, 0.0f}, i, s << 8};
            }
        }
}

// The path vertex kernel.  PSHADE_QUEUE: persistent blocks take tiles of
// VERTEX_B * PATH_PER paths (PATH_PER consecutive paths a thread) by ticket,
// and write the tile's shadow and next rays densely, each queue's rays in
// path order within a run that the tile takes with one atomic on that
// queue's count (tiles in whatever order they get there: a ray's result does
// not depend on its queue position, and the path id travels with it).  A
// ticket is a device-scope atomic on one address, so fewer, larger tiles.
// PSHADE_PLAIN: one path per thread, every path's slot written.
constexpr uint32_t PATH_PER = 2;
__global__ __launch_bounds__(VERTEX_B) void k_path_shade(RenderParams P, uint64_t g0, uint32_t bounce, uint32_t cap,
                                                   const uint32_t* q_count, const uint32_t* pid_q,
                                                   const Ray* rays_in, const Hit* hits, PathState* st,
                                                   Ray* rays_out, uint32_t* pid_out, uint32_t* out_count,
                                                   Ray* sh, uint32_t* sh_pid, float4* shc,
                                                   uint32_t* sh_count, float4* sample_l, uint32_t* ticket,
                                                   int mode) {
    const uint32_t n = q_count ? *q_count : cap;
    const bool fresh = rays_in == nullptr;
    if (mode == PSHADE_PLAIN) {
        const uint32_t i = blockIdx.x * VERTEX_B + threadIdx.x;
        VertexOut o;
        o.shadow = o.next = false;
        o.contrib = mk3(0, 0, 0);
        if (i < n) path_vertex(P, g0, bounce, i, false, fresh, pid_q, rays_in, hits, st, sample_l, o);
        if (i < n) {
            if (!o.shadow) set_dead(&o.sray);
            if (!o.next) set_dead(&o.nray);
            sh[i] = o.sray;
            shc[i] = make_float4(o.contrib.x, o.contrib.y, o.contrib.z, o.shadow ? 1.0f : 0.0f);
            rays_out[i] = o.nray;
        }
        uint32_t ts, tn;
        block_exclusive_scan<VERTEX_B>(o.shadow ? 1u : 0u, &ts);
        block_exclusive_scan<VERTEX_B>(o.next ? 1u : 0u, &tn);
        if (threadIdx.x == 0) {
            if (ts) atomicAdd(sh_count, ts);
            if (tn) atomicAdd(out_count, tn);
        }
        return;
    }
    __shared__ uint32_t s_bid, s_es, s_en;
    constexpr uint32_t TILE = VERTEX_B * PATH_PER;
    const uint32_t ntiles = (n + TILE - 1) / TILE;
    while (true) {
        if (threadIdx.x == 0) s_bid = atomicAdd(ticket, 1u);
        __syncthreads();
        const uint32_t bid = s_bid;
        __syncthreads();  // every thread has read s_bid before the next tile
        if (bid >= ntiles) break;
        const uint32_t first = bid * TILE + threadIdx.x * PATH_PER;
        VertexOut o[PATH_PER];
        uint32_t cs = 0, cn = 0;
#pragma unroll
        for (uint32_t k = 0; k < PATH_PER; ++k) {
            o[k].shadow = o[k].next = false;
            if (first + k < n) path_vertex(P, g0, bounce, first + k, true, fresh, pid_q, rays_in, hits, st, sample_l, o[k]);
            cs += o[k].shadow ? 1u : 0u;
            cn += o[k].next ? 1u : 0u;
        }
        // one scan for both queues: shadow count in the low half, next in
        // the high half (a tile holds at most 1024 of each)
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<VERTEX_B>(cs | (cn << 16), &tot);
        const uint32_t ts = tot & 0xFFFFu, tn = tot >> 16;
        // the tile's runs of the two queues
        if (threadIdx.x == 0) s_es = ts ? atomicAdd(sh_count, ts) : 0u;
        if (threadIdx.x == 64) s_en = tn ? atomicAdd(out_count, tn) : 0u;
        __syncthreads();
        uint32_t qs = s_es + (ex & 0xFFFFu), qn = s_en + (ex >> 16);
#pragma unroll
        for (uint32_t k = 0; k < PATH_PER; ++k) {
            if (o[k].shadow) {
                sh[qs] = o[k].sray;
                sh_pid[qs] = o[k].pid;
                shc[qs] = make_float4(o[k].contrib.x, o[k].contrib.y, o[k].contrib.z, 0.0f);
                ++qs;
            }
            if (o[k].next) {
                rays_out[qn] = o[k].nray;
                pid_out[qn] = o[k].pid;
                ++qn;
            }
        }
    }
}

// Adds the unoccluded light samples of this bounce; finishes paths that ended.
__global__ __launch_bounds__(RB) void k_path_shadow(RenderParams P, uint64_t g0, uint32_t cap,
                                                    const uint32_t* sh_count, const uint32_t* sh_pid,
                                                    const float4* shc, const uint8_t* occ, PathState* st,
                                                    float4* sample_l) {
    const uint32_t n = sh_count ? *sh_count : cap;
    for (uint32_t i = blockIdx.x * RB + threadIdx.x; i < n; i += gridDim.x * RB) {
        const uint32_t pid = sh_pid ? sh_pid[i] : i;
        const float4 c = shc[i];
        PathState ps = st[pid];
        if (!sh_pid) {
            if (ps.flags & 4u) continue;   // plain form: finished earlier
            if (c.w == 0.0f) continue;     // no light sample this bounce
        }
        if (!occ[i]) {  // unoccluded
            ps.L[0] = ps.L[0] + c.x; ps.L[1] = ps.L[1] + c.y; ps.L[2] = ps.L[2] + c.z;
        }
        if (ps.flags & 2u) {
            *sample_ref(P, g0, ps, sample_l).out = make_float4(ps.L[0], ps.L[1], ps.L[2], 0.0f);
            ps.flags = (ps.flags & ~2u) | 4u;
        }
        st[pid] = ps;
    }
}

// Paths still alive when the last bounce has been processed (plain form keeps
// them in place) or finished in the plain form: nothing to do -- every path
// writes its radiance when it ends.

// Octant sort of the extension-ray queue: a stable partition by direction
// octant within tiles of OS_TILE consecutive queue entries.  The rays and
// their path ids are copied to the output queue in (tile, octant, position)
// order, so a wave of the next traversal takes rays of one octant -- one
// near-first visiting order, the same direction signs -- wherever a tile
// holds 64 or more of them, while rays stay near their queue neighbours
// (nearby pixels).  Octant bit a is the sign bit of d[a] (the sign of 1/d
// that picks the traversal's near child, -0 included).  Positions depend
// only on the input order (which itself depends on the order in which the
// vertex kernel's tiles took their runs; no result depends on either).
constexpr uint32_t OS_BLOCK = 1024, OS_PER = 4, OS_TILE = OS_BLOCK * OS_PER;
__device__ __forceinline__ uint32_t octant_of(const Ray& r) {
    return (__float_as_uint(r.dx) >> 31) | ((__float_as_uint(r.dy) >> 31) << 1) | ((__float_as_uint(r.dz) >> 31) << 2);
}
__global__ __launch_bounds__(OS_BLOCK) void k_octant_sort(const uint32_t* count, const Ray* rin, const uint32_t* pin,
                                                          Ray* rout, uint32_t* pout) {
    const uint32_t n = *count;
    constexpr uint32_t NW = OS_BLOCK / 64;
    __shared__ uint32_t s_cnt[OS_PER][NW][8];  // per (round, wave, octant): count, then exclusive offset
    __shared__ uint32_t s_base[8];
    const uint32_t lane = __lane_id(), wid = threadIdx.x >> 6;
    const unsigned long long lt = lt_mask();
    for (uint32_t tile = blockIdx.x; (uint64_t)tile * OS_TILE < n; tile += gridDim.x) {
        const uint32_t t0 = tile * OS_TILE;
        Ray r[OS_PER];
        uint32_t p[OS_PER], oc[OS_PER], rank[OS_PER];
#pragma unroll
        for (uint32_t k = 0; k < OS_PER; ++k) {
            const uint32_t i = t0 + k * OS_BLOCK + threadIdx.x;
            oc[k] = 8u;  // past the end: no octant
            rank[k] = 0u;
            if (i < n) {
                r[k] = rin[i];
                p[k] = pin[i];
                oc[k] = octant_of(r[k]);
            }
#pragma unroll
            for (uint32_t o = 0; o < 8; ++o) {
                const unsigned long long m = __ballot(oc[k] == o);
                if (oc[k] == o) rank[k] = (uint32_t)__popcll(m & lt);
                if (lane == 0) s_cnt[k][wid][o] = (uint32_t)__popcll(m);
            }
        }
        __syncthreads();
        // per octant: exclusive offsets over (round, wave), which is input order
        if (threadIdx.x < 8) {
            const uint32_t o = threadIdx.x;
            uint32_t acc = 0;
            for (uint32_t k = 0; k < OS_PER; ++k)
                for (uint32_t w = 0; w < NW; ++w) {
                    const uint32_t c = s_cnt[k][w][o];
                    s_cnt[k][w][o] = acc;
                    acc += c;
                }
            s_base[o] = acc;  // the octant's total, turned into its base below
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t acc = 0;
            for (uint32_t o = 0; o < 8; ++o) {
                const uint32_t c = s_base[o];
                s_base[o] = acc;
                acc += c;
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < OS_PER; ++k) {
            if (oc[k] < 8u) {
                const uint32_t dst = t0 + s_base[oc[k]] + s_cnt[k][wid][oc[k]] + rank[k];
                rout[dst] = r[k];
                pout[dst] = p[k];
            }
        }
        __syncthreads();  // the tables are reused by the next tile
    }
}

// Film: box filter of half-width 0.5 (Filter::mean, filter.rs:89;
// Film::add_sample film.rs:192-249), gathered per pixel from the samples of
// the 3x3 neighbourhood in row-major order; write_image's conversion with the
// buffer and splat fixed (film.rs:316-354).
__device__ __forceinline__ bool covers(float ixf, int32_t x) {
    float dimg = ixf - 0.5f;
    int32_t a = (int32_t)ceilf(dimg - 0.5f), b = (int32_t)floorf(dimg + 0.5f);
    return a <= x && x <= b;
}
// A sample at image position ix of pixel column qx (qx <= ix <= qx + 1)
// covers qx, and also qx - 1 when ix == qx and qx + 1 when ix == qx + 1;
// likewise in y.  Each pixel's edge bits say which neighbours its samples
// reach: 1 left (ix == qx), 2 right (ix == qx + 1), 4 up (iy == qy), 8 down
// (iy == qy + 1).  A 16 x 16 tile of pixels finds the bits of itself and its
// one-pixel ring first, then each pixel gathers its own samples and those
// of the neighbours whose bits reach it -- the same additions, in the same
// order, as a gather over all nine neighbourhoods, without evaluating the
// sample positions of neighbours that cannot contribute.
constexpr int FT = 16;  // film tile side (FT * FT = RB threads)
// the samples of a pixel the film adds: all spp, or a Halton pixel's count
__device__ __forceinline__ uint32_t film_samples(const RenderParams& P, uint32_t pixel) {
    return P.scount ? P.scount[pixel] : P.spp;
}
static_assert(FT * FT == RB, "one film tile per render block");
__global__ __launch_bounds__(RB) void k_film(RenderParams P, const float4* sample_l, float* rgb) {
    __shared__ uint8_t s_edge[FT + 2][FT + 2];
    const int32_t tx0 = (int32_t)blockIdx.x * FT, ty0 = (int32_t)blockIdx.y * FT;
    for (uint32_t h = threadIdx.x; h < (FT + 2) * (FT + 2); h += RB) {
        const int32_t qx = tx0 - 1 + (int32_t)(h % (FT + 2)), qy = ty0 - 1 + (int32_t)(h / (FT + 2));
        uint8_t e = 0;
        if (qx >= 0 && qx < (int32_t)P.xres && qy >= 0 && qy < (int32_t)P.yres) {
            const uint32_t ns = film_samples(P, (uint32_t)qy * P.xres + (uint32_t)qx);
            for (uint32_t s = 0; s < ns; ++s) {
                float ix, iy;
                sample_pos(P, (uint32_t)qx, (uint32_t)qy, s, &ix, &iy);
                if (ix == (float)qx) e |= 1u;
                if (ix == (float)(qx + 1)) e |= 2u;
                if (iy == (float)qy) e |= 4u;
                if (iy == (float)(qy + 1)) e |= 8u;
            }
        }
        s_edge[h / (FT + 2)][h % (FT + 2)] = e;
    }
    __syncthreads();
    const int32_t x = tx0 + (int32_t)(threadIdx.x % FT), y = ty0 + (int32_t)(threadIdx.x / FT);
    if (x >= (int32_t)P.xres || y >= (int32_t)P.yres) return;
    float xyz[3] = {0.0f, 0.0f, 0.0f};
    float wsum = 0.0f;
    for (int32_t dy = -1; dy <= 1; ++dy) {
        const int32_t qy = y + dy;
        if (qy < 0 || qy >= (int32_t)P.yres) continue;
        for (int32_t dx = -1; dx <= 1; ++dx) {
            const int32_t qx = x + dx;
            if (qx < 0 || qx >= (int32_t)P.xres) continue;
            const uint32_t e = s_edge[qy - ty0 + 1][qx - tx0 + 1];
            // can a sample of (qx, qy) reach (x, y)?
            const bool reach_x = dx == 0 || (dx < 0 ? (e & 2u) : (e & 1u));
            const bool reach_y = dy == 0 || (dy < 0 ? (e & 8u) : (e & 4u));
            if (!reach_x || !reach_y) continue;
            const uint32_t qpix = (uint32_t)qy * P.xres + (uint32_t)qx;
            const float4* sl = sample_l + (size_t)qpix * P.spp;
            const uint32_t ns = film_samples(P, qpix);
            // rgb_to_xyz (spectrum.rs:37-41), weight 1
            auto add = [&](const float4 v) {
                const float c0 = 0.412453f * v.x + 0.357580f * v.y + 0.180423f * v.z;
                const float c1 = 0.212671f * v.x + 0.715160f * v.y + 0.072169f * v.z;
                const float c2 = 0.019334f * v.x + 0.119193f * v.y + 0.950227f * v.z;
                const float wt = 1.0f;
                xyz[0] = xyz[0] + wt * c0;
                xyz[1] = xyz[1] + wt * c1;
                xyz[2] = xyz[2] + wt * c2;
                wsum = wsum + wt;
            };
            if (dx == 0 && dy == 0) {
                // every own sample covers the pixel: read them eight at a time
                // (one 128-byte line a thread), add them in order
                uint32_t s = 0;
                for (; s + 8 <= ns; s += 8) {
                    float4 v[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) v[k] = sl[s + k];
#pragma unroll
                    for (int k = 0; k < 8; ++k) add(v[k]);
                }
                for (; s < ns; ++s) add(sl[s]);
                continue;
            }
            for (uint32_t s = 0; s < ns; ++s) {
                float ix, iy;
                sample_pos(P, (uint32_t)qx, (uint32_t)qy, s, &ix, &iy);
                if (covers(ix, x) && covers(iy, y)) add(sl[s]);
            }
        }
    }
    // xyz_to_rgb: spectrum.rs:31-35
    float o0 = 3.240479f * xyz[0] - 1.37150f * xyz[1] - 0.498535f * xyz[2];
    float o1 = -0.969256f * xyz[0] + 1.875991f * xyz[1] + 0.041556f * xyz[2];
    float o2 = 0.055648f * xyz[0] - 0.204043f * xyz[1] + 1.057311f * xyz[2];
    if (wsum != 0.0f) {
        const float inv = 1.0f / wsum;
        o0 = fmaxf(o0 * inv, 0.0f); o1 = fmaxf(o1 * inv, 0.0f); o2 = fmaxf(o2 * inv, 0.0f);
    }
    const size_t pix = (size_t)y * P.xres + (size_t)x;
    rgb[3 * pix] = o0; rgb[3 * pix + 1] = o1; rgb[3 * pix + 2] = o2;
}

// Film with a tabulated filter (Film::add_sample film.rs:192-249 with the
// weight table of film.rs:89-97): each pixel gathers, in row-major order of
// the pixels within the filter's reach and sample order within each, the
// samples whose raster extent (film.rs:197-208) covers it, weighted by the
// table entry of its offsets (film.rs:214-229) -- the additions the
// reference makes for that pixel when the frame's samples are added in
// pixel and sample order.  Every sample is read by up to (2R+1)^2 pixels, so
// its position within its pixel and its xyz are staged once, sample-major:
// the 64 pixels of a wave (a 16 x 4 block of the tile) then read sample s of
// 64 neighbouring pixels as four contiguous runs.
__device__ __forceinline__ void film_extent(float v, float w, int32_t n, int32_t* lo, int32_t* hi) {
    const float d = v - 0.5f;
    *lo = max(0, (int32_t)ceilf(d - w));
    *hi = min(n - 1, (int32_t)floorf(d + w));
}
__device__ __forceinline__ int32_t film_offset(int32_t p, float v, float inv_w) {
    const float d = v - 0.5f;
    const float a = floorf(fabsf(((float)p - d) * inv_w * (float)FILTER_TABLE_DIM));
    if (!(a < (float)(FILTER_TABLE_DIM - 1))) return a == a ? FILTER_TABLE_DIM - 1 : 0;  // `as usize` saturates
    return (int32_t)a;
}
// LDSampler image samples of every pixel (ld_shuffle_scrambled_2d,
// sampler/utils.rs:52-69), one pixel a thread, in place in out (pixel-major):
// the scrambled (0,2) points, then the pixel's shuffle (RNG::shuffle,
// rng.rs:23-32).  Draws 0 and 1 give the scrambles and draws 2 .. spp+1 the
// one-point shuffles that change nothing; the shuffle takes draws spp+2 ..
__global__ __launch_bounds__(RB) void k_ld_points(uint32_t seed, uint32_t npix, uint32_t spp, float2* out) {
    for (uint32_t pixel = blockIdx.x * RB + threadIdx.x; pixel < npix; pixel += gridDim.x * RB) {
        float2* pts = out + (size_t)pixel * spp;
        const uint32_t sc0 = (uint32_t)ld_draw(seed, pixel, 0), sc1 = (uint32_t)ld_draw(seed, pixel, 1);
        for (uint32_t i = 0; i < spp; ++i) pts[i] = make_float2(ld_vdc(i, sc0), ld_sobol2(i, sc1));  // sample02
        for (uint32_t i = 0; i < spp; ++i) {
            const uint32_t other = i + (uint32_t)(ld_draw(seed, pixel, spp + 2 + i) % (uint64_t)(spp - i));
            const float2 t = pts[i];
            pts[i] = pts[other];
            pts[other] = t;
        }
    }
}

// HaltonSampler (halton.rs:48-107) over the whole image as one window: sample
// k of the sequence lies at (lerp(0, delta, rinv(k, 3)), lerp(0, delta,
// rinv(k, 2))) and is kept when inside the image; a pixel's samples take its
// slots in sequence order.  radical_inverse (montecarlo.rs:8-21) in f64,
// Lerp for f32 (utils/mod.rs:15-18), as the reference computes them.
__device__ double radical_inverse(uint64_t n, uint64_t b) {
    double v = 0.0, aib = 1.0;
    const double inv_base = 1.0 / (double)b;
    while (n > 0) {
        const double d = (double)(n % b);
        n /= b;
        aib *= inv_base;
        v += d * aib;
    }
    return v;
}
__device__ __forceinline__ bool halton_point(uint64_t k, uint32_t xres, uint32_t yres, float delta, float* ix,
                                             float* iy) {
    const float u = (float)radical_inverse(k, 3), v = (float)radical_inverse(k, 2);
    const float x0 = 0.0f, y0 = 0.0f;
    *ix = x0 * (1.0f - u) + (x0 + delta) * u;
    *iy = y0 * (1.0f - v) + (y0 + delta) * v;
    return !(*ix >= (float)xres || *iy >= (float)yres);
}
// pass 0: count each pixel's samples; pass 1: give each its slot (any order)
__global__ __launch_bounds__(RB) void k_halton_bin(uint64_t wanted, uint32_t xres, uint32_t yres, float delta,
                                                   uint32_t* cnt, uint64_t* slots, uint32_t K) {
    for (uint64_t k = (uint64_t)blockIdx.x * RB + threadIdx.x; k < wanted; k += (uint64_t)gridDim.x * RB) {
        float ix, iy;
        if (!halton_point(k, xres, yres, delta, &ix, &iy)) continue;
        const uint32_t pixel = (uint32_t)iy * xres + (uint32_t)ix;
        const uint32_t at = atomicAdd(cnt + pixel, 1u);
        if (slots && at < K) slots[(size_t)pixel * K + at] = k;
    }
}
// per pixel: its slots in sequence order, then each sample's position within
// the pixel (exact: ix - px is exact for ix in [px, px + 1), and px + that
// gives ix back); empty slots at 0
__global__ __launch_bounds__(RB) void k_halton_pos(uint32_t npix, uint32_t xres, uint32_t yres, float delta,
                                                   const uint32_t* cnt, uint64_t* slots, uint32_t K, float2* spos) {
    for (uint32_t p = blockIdx.x * RB + threadIdx.x; p < npix; p += gridDim.x * RB) {
        uint64_t* sl = slots + (size_t)p * K;
        const uint32_t n = min(cnt[p], K);
        for (uint32_t i = 1; i < n; ++i) {
            const uint64_t v = sl[i];
            uint32_t j = i;
            for (; j > 0 && sl[j - 1] > v; --j) sl[j] = sl[j - 1];
            sl[j] = v;
        }
        const float px = (float)(p % xres), py = (float)(p / xres);
        for (uint32_t s = 0; s < K; ++s) {
            float ix = px, iy = py;
            if (s < n) halton_point(sl[s], xres, yres, delta, &ix, &iy);
            spos[(size_t)p * K + s] = make_float2(ix - px, iy - py);
        }
    }
}

__global__ __launch_bounds__(RB) void k_film_stage(RenderParams P, const float4* sample_l, uint64_t n) {
    const uint64_t npix = (uint64_t)P.xres * P.yres;
    for (uint64_t g = (uint64_t)blockIdx.x * RB + threadIdx.x; g < n; g += (uint64_t)gridDim.x * RB) {
        const uint32_t pixel = (uint32_t)(g / P.spp), s = (uint32_t)(g % P.spp);
        float2 f;
        if (P.spos) {
            f = P.spos[g];
        } else {
            float u[4];
            rng4(P.seed, pixel, s, 0, u);  // sample_pos's jitter
            const uint32_t sx = s % P.nx, sy = s / P.nx;
            f = make_float2(((float)sx + u[0]) * P.dx, ((float)sy + u[1]) * P.dy);
        }
        const float4 v = sample_l[g];
        // rgb_to_xyz (spectrum.rs:37-41)
        const float c0 = 0.412453f * v.x + 0.357580f * v.y + 0.180423f * v.z;
        const float c1 = 0.212671f * v.x + 0.715160f * v.y + 0.072169f * v.z;
        const float c2 = 0.019334f * v.x + 0.119193f * v.y + 0.950227f * v.z;
        const uint64_t t = (uint64_t)s * npix + pixel;
        P.fpos[t] = f;
        P.fxyz[t] = make_float4(c0, c1, c2, 0.0f);
    }
}
__global__ __launch_bounds__(RB) void k_film_filter(RenderParams P, float* rgb) {
    __shared__ float s_tab[FILTER_TABLE_DIM * FILTER_TABLE_DIM];
    for (uint32_t k = threadIdx.x; k < FILTER_TABLE_DIM * FILTER_TABLE_DIM; k += RB) s_tab[k] = P.ftab[k];
    __syncthreads();
    const int32_t x = (int32_t)(blockIdx.x * FT + threadIdx.x % FT), y = (int32_t)(blockIdx.y * FT + threadIdx.x / FT);
    if (x >= (int32_t)P.xres || y >= (int32_t)P.yres) return;
    const int32_t xres = (int32_t)P.xres, yres = (int32_t)P.yres;
    float xyz[3] = {0.0f, 0.0f, 0.0f};
    float wsum = 0.0f;
    for (int32_t qy = max(0, y - P.fry); qy <= min(yres - 1, y + P.fry); ++qy) {
        for (int32_t qx = max(0, x - P.frx); qx <= min(xres - 1, x + P.frx); ++qx) {
            const size_t npix = (size_t)P.xres * P.yres, q = (size_t)qy * P.xres + (size_t)qx;
            const uint32_t ns = film_samples(P, (uint32_t)q);
            for (uint32_t s = 0; s < ns; ++s) {
                // sample_pos: the position within the pixel plus the pixel's coordinates
                const float2 f = P.fpos[(size_t)s * npix + q];
                const float ix = f.x + (float)qx, iy = f.y + (float)qy;
                int32_t x0, x1, y0, y1;
                film_extent(ix, P.fxw, xres, &x0, &x1);
                film_extent(iy, P.fyw, yres, &y0, &y1);
                if (x < x0 || x > x1 || y < y0 || y > y1) continue;
                const float wt = s_tab[film_offset(y, iy, P.fiyw) * FILTER_TABLE_DIM + film_offset(x, ix, P.fixw)];
                const float4 c = P.fxyz[(size_t)s * npix + q];
                xyz[0] = xyz[0] + wt * c.x;
                xyz[1] = xyz[1] + wt * c.y;
                xyz[2] = xyz[2] + wt * c.z;
                wsum = wsum + wt;
            }
        }
    }
    // xyz_to_rgb: spectrum.rs:31-35
    float o0 = 3.240479f * xyz[0] - 1.37150f * xyz[1] - 0.498535f * xyz[2];
    float o1 = -0.969256f * xyz[0] + 1.875991f * xyz[1] + 0.041556f * xyz[2];
    float o2 = 0.055648f * xyz[0] - 0.204043f * xyz[1] + 1.057311f * xyz[2];
    if (wsum != 0.0f) {
        const float inv = 1.0f / wsum;
        o0 = fmaxf(o0 * inv, 0.0f); o1 = fmaxf(o1 * inv, 0.0f); o2 = fmaxf(o2 * inv, 0.0f);
    }
    const size_t pix = (size_t)y * P.xres + (size_t)x;
    rgb[3 * pix] = o0; rgb[3 * pix + 1] = o1; rgb[3 * pix + 2] = o2;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// ---- Filter::evaluate (filter.rs:89-124), sinc_1d (utils/mod.rs:207-217) and
// the film's weight table (film.rs:89-97)
bool filter_valid(const FilmFilter& f) {
    return f.kind >= 0 && f.kind <= 4 && f.xw > 0.0f && f.yw > 0.0f && std::isfinite(f.xw) && std::isfinite(f.yw);
}
namespace {
float sinc_1d(float x, float tau) {
    float v = std::fabs(x);
    if (v < 1e-5f) return 1.0f;
    if (v >= 1.0f) return 0.0f;
    v = v * 3.14159265358979323846f;
    const float vtau = v * tau;
    const float sv = std::sin(vtau) / vtau;
    return sv * std::sin(v) / v;
}
float mitchell_1d(float v, float b, float c) {
    const float t = std::fabs(v * 2.0f);
    float p;
    if (t >= 2.0f) p = 0.0f;
    else if (t > 1.0f)
        p = (-b - 6.0f * c) * t * t * t + (6.0f * b + 30.0f * c) * t * t + (-12.0f * b - 48.0f * c) * t +
            (8.0f * b + 24.0f * c);
    else
        p = (12.0f - 9.0f * b - 6.0f * c) * t * t * t + (-18.0f + 12.0f * b + 6.0f * c) * t * t + (6.0f - 2.0f * b);
    return (1.0f / 6.0f) * p;
}
}  // namespace
float filter_evaluate(const FilmFilter& f, float x, float y) {
    const float ixw = 1.0f / f.xw, iyw = 1.0f / f.yw;  // FilterBase::new
    switch (f.kind) {
        case 0: return 1.0f;
        case 1: {
            const float dx = (f.xw - std::fabs(x)) * ixw, dy = (f.yw - std::fabs(y)) * iyw;
            return std::fmax(dx, 0.0f) * std::fmax(dy, 0.0f);
        }
        case 2: {
            const float a = f.p0;
            const float ex = std::exp(-a * f.xw * f.xw), ey = std::exp(-a * f.yw * f.yw);  // Filter::gaussian
            return std::fmax(std::exp(-a * x * x) - ex, 0.0f) * std::fmax(std::exp(-a * y * y) - ey, 0.0f);
        }
        case 3: return mitchell_1d(x * ixw, f.p0, f.p1) * mitchell_1d(y * iyw, f.p0, f.p1);
        default: return sinc_1d(x * ixw, f.p0) * sinc_1d(y * iyw, f.p0);
    }
}
void filter_weight_table(const FilmFilter& f, float* out) {
    for (int y = 0; y < FILTER_TABLE_DIM; ++y) {
        const float fy = ((float)y + 0.5f) * f.yw / (float)FILTER_TABLE_DIM;
        for (int x = 0; x < FILTER_TABLE_DIM; ++x) {
            const float fx = ((float)x + 0.5f) * f.xw / (float)FILTER_TABLE_DIM;
            out[y * FILTER_TABLE_DIM + x] = filter_evaluate(f, fx, fy);
        }
    }
}

namespace {

// ---- Camera::perspective + Projection::new (camera/mod.rs:105-135,
// projective.rs:48-72) with the LU inverse of matrix4x4.rs:66-160.
struct HM4 {
    float m[4][4];
};
HM4 hid() {
    HM4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = i == j ? 1.0f : 0.0f;
    return r;
}
HM4 hmul(const HM4& a, const HM4& b) {
    HM4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] +
                        a.m[i][3] * b.m[3][j];
    return r;
}
bool hinv(const HM4& a, HM4& out) {
    float s[4];
    for (int i = 0; i < 4; ++i) {
        float acc = 0.0f;
        for (int j = 0; j < 4; ++j) acc = fmaxf(fabsf(a.m[i][j]), acc);
        s[i] = acc;
    }
    HM4 lu = a;
    int piv[4] = {0, 1, 2, 3};
    for (int k = 0; k < 3; ++k) {
        float c = 0.0f;
        int p = k;
        for (int i = k; i < 4; ++i) {
            float cc = fabsf(lu.m[i][k] / s[i]);
            if (cc > c) { c = cc; p = i; }
        }
        piv[k] = p;
        if (c == 0.0f) return false;
        if (p != k)
            for (int j = k; j < 4; ++j) { float t = lu.m[k][j]; lu.m[k][j] = lu.m[p][j]; lu.m[p][j] = t; }
        for (int i = k + 1; i < 4; ++i) {
            float mi = lu.m[i][k] / lu.m[k][k];
            lu.m[i][k] = mi;
            for (int j = k + 1; j < 4; ++j) lu.m[i][j] = lu.m[i][j] - mi * lu.m[k][j];
        }
    }
    if (fabsf(lu.m[3][3]) < 1.0e-6f) return false;
    HM4 rows;
    for (int c = 0; c < 4; ++c) {
        float b[4] = {0, 0, 0, 0};
        b[c] = 1.0f;
        for (int k = 0; k < 3; ++k) {
            if (piv[k] != k) { float t = b[piv[k]]; b[piv[k]] = b[k]; b[k] = t; }
            for (int i = k + 1; i < 4; ++i) b[i] = b[i] - lu.m[i][k] * b[k];
        }
        b[3] = b[3] / lu.m[3][3];
        for (int i = 2; i >= 0; --i) {
            float sum = 0.0f;
            for (int j = i + 1; j < 4; ++j) sum = sum + lu.m[i][j] * b[j];
            b[i] = (b[i] - sum) / lu.m[i][i];
        }
        for (int j = 0; j < 4; ++j) rows.m[c][j] = b[j];
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out.m[i][j] = rows.m[j][i];
    return true;
}
struct HXf {
    HM4 m, mi;
};
HXf hscale(float x, float y, float z) {
    HXf t{hid(), hid()};
    t.m.m[0][0] = x; t.m.m[1][1] = y; t.m.m[2][2] = z;
    t.mi.m[0][0] = 1.0f / x; t.mi.m[1][1] = 1.0f / y; t.mi.m[2][2] = 1.0f / z;
    return t;
}
HXf htranslate(float x, float y, float z) {
    HXf t{hid(), hid()};
    t.m.m[0][3] = x; t.m.m[1][3] = y; t.m.m[2][3] = z;
    t.mi.m[0][3] = -x; t.mi.m[1][3] = -y; t.mi.m[2][3] = -z;
    return t;
}
HXf hxmul(const HXf& a, const HXf& b) { return HXf{hmul(a.m, b.m), hmul(b.mi, a.mi)}; }
HXf hxinv(const HXf& a) { return HXf{a.mi, a.m}; }

bool perspective_r2c(uint32_t xres, uint32_t yres, const float screen[4], float fov, float out[16]) {
    const float znear = 1e-2f, zfar = 1000.0f;
    HM4 p = hid();
    p.m[2][2] = zfar / (zfar - znear);
    p.m[2][3] = -(zfar * znear) / (zfar - znear);
    p.m[3][2] = 1.0f;
    p.m[3][3] = 0.0f;
    HM4 pinv;
    if (!hinv(p, pinv)) return false;
    const float inv_tan = 1.0f / tanf((fov * PI_F / 180.0f) / 2.0f);
    HXf persp = hxmul(hscale(inv_tan, inv_tan, 1.0f), HXf{p, pinv});
    HXf s2r = hxmul(hxmul(hscale((float)xres, (float)yres, 1.0f),
                          hscale(1.0f / (screen[1] - screen[0]), 1.0f / (screen[2] - screen[3]), 1.0f)),
                    htranslate(-screen[0], -screen[3], 0.0f));
    HXf r2c = hxmul(hxinv(persp), hxinv(s2r));
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[4 * i + j] = r2c.m.m[i][j];
    return true;
}

int grid_for(size_t n) {
    size_t g = (n + RB - 1) / RB;
    size_t cap = (size_t)device_cu_count() * 16;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

}  // namespace

#define RCHK(x)                                  \
    do {                                         \
        hipError_t e_ = (x);                     \
        if (e_ != hipSuccess) return hip_fail(e_, #x); \
    } while (0)
#define ICHK(x)              \
    do {                     \
        int r_ = (x);        \
        if (r_) return r_;   \
    } while (0)

// ---- scene and camera set-up -------------------------------------------------
void RenderScene::release() {
    scene_free(&scene);
    faces.release(); sph.release(); mats.release(); atris.release(); a_first.release(); a_count.release();
    a_total.release(); a_le.release(); points.release();
}

void WorkBufs::release() {
    rays.release(); rays2.release(); rays3.release(); sh.release(); hits.release(); hits2.release(); contrib.release();
    lbase.release(); qidx.release(); occ.release(); pid.release(); pid2.release(); pid3.release();
    shpid.release(); bcounts.release(); st.release();
}

// Triangle::intersect's dg with the default uvs (mesh.rs:212-254) for the
// face (p1, p2, p3) = the vertices of idx[2], idx[1], idx[0]: dg.nn =
// normalize(dpdu x dpdv) and normalize(dpdu) (bsdf/mod.rs:70-73).  The
// host's area-light records and the device's face records use it alike.
PX_HD void face_frame(f3 p1, f3 p2, f3 p3, f3* nn, f3* tn) {
    const float du1 = 0.0f - 1.0f, du2 = 1.0f - 1.0f, dv1 = 0.0f - 1.0f, dv2 = 0.0f - 1.0f;
    const f3 dp1 = sub3(p1, p3), dp2 = sub3(p2, p3);
    const float det = du1 * dv2 - dv1 * du2;
    const float inv_det = 1.0f / det;
    const f3 dpdu = mul3(sub3(mul3(dp1, dv2), mul3(dp2, dv1)), inv_det);
    const f3 dpdv = mul3(add3(mul3(dp1, -du2), mul3(dp2, du1)), inv_det);
    *nn = normalize3(cross3(dpdu, dpdv));
    *tn = normalize3(dpdu);
}

__global__ __launch_bounds__(RB) void k_face_shade(const float* verts, const uint32_t* idx, uint32_t nfaces,
                                                   const uint32_t* face_mat, const int32_t* face_light,
                                                   FaceShade* out) {
    for (uint32_t f = blockIdx.x * RB + threadIdx.x; f < nfaces; f += gridDim.x * RB) {
        auto V = [&](uint32_t i) { return mk3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]); };
        const f3 p1 = V(idx[3 * f + 2]), p2 = V(idx[3 * f + 1]), p3 = V(idx[3 * f]);
        f3 nn, tn;
        face_frame(p1, p2, p3, &nn, &tn);
        FaceShade sh;
        sh.nn[0] = nn.x; sh.nn[1] = nn.y; sh.nn[2] = nn.z;
        sh.tn[0] = tn.x; sh.tn[1] = tn.y; sh.tn[2] = tn.z;
        sh.mat = face_mat[f];
        sh.light = face_light ? face_light[f] : -1;
        out[f] = sh;
    }
}

// The per-face shading records, computed on the device from uploaded copies
// of the mesh (the host loop over every face took seconds on 64M faces).
static int face_shade_device(const pbrtx_render_desc* d, DevArr<FaceShade>* faces) {
    int rc = faces->ensure(d->nfaces ? d->nfaces : 1);
    if (rc || d->nfaces == 0) return rc;
    DevArr<float> v;
    DevArr<uint32_t> ix, fm;
    DevArr<int32_t> fl;
    auto put = [](auto* arr, const auto* src, size_t count) -> int {
        int r = arr->ensure(count);
        if (r) return r;
        hipError_t e = hipMemcpy(arr->p, src, count * sizeof(*src), hipMemcpyHostToDevice);
        return e == hipSuccess ? PBRTX_OK : hip_fail(e, "hipMemcpy(mesh for face records)");
    };
    if (!(rc = put(&v, d->verts, 3 * (size_t)d->nverts)) && !(rc = put(&ix, d->idx, 3 * (size_t)d->nfaces)) &&
        !(rc = put(&fm, d->face_mat, d->nfaces)) && !(d->face_light && (rc = put(&fl, d->face_light, d->nfaces)))) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>(((uint64_t)d->nfaces + RB - 1) / RB, 65536);
        hipLaunchKernelGGL(k_face_shade, dim3(grid), dim3(RB), 0, nullptr, v.p, ix.p, d->nfaces, fm.p,
                           d->face_light ? fl.p : nullptr, faces->p);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);  // before the temporaries are freed
        if (e != hipSuccess) rc = hip_fail(e, "k_face_shade");
    }
    v.release(); ix.release(); fm.release(); fl.release();
    return rc;
}

int render_scene_setup(RenderScene* S, const pbrtx_render_desc* d) {
    if (!d->verts || !d->idx || !d->face_mat || !d->mats) return fail(PBRTX_ERR_ARG, "null scene arrays");
    if (d->integrator != 0 && d->integrator != 1) return fail(PBRTX_ERR_ARG, "unknown integrator");
    if (d->integrator == 0 && d->n_area > 0 && d->light_samples == 0)
        return fail(PBRTX_ERR_ARG, "light_samples must be >= 1");
    for (uint32_t f = 0; f < d->nfaces; ++f)
        if (d->face_mat[f] >= d->nmats) return fail(PBRTX_ERR_ARG, "face material out of range");
    if (d->nspheres && (!d->spheres || !d->sphere_mat)) return fail(PBRTX_ERR_ARG, "null sphere arrays");
    std::vector<SphereShade> ssh(d->nspheres);
    for (uint32_t k = 0; k < d->nspheres; ++k) {
        const pbrtx_sphere& q = d->spheres[k];
        if (d->sphere_mat[k] >= d->nmats) return fail(PBRTX_ERR_ARG, "sphere material out of range");
        if (!(q.radius > 0.0f)) return fail(PBRTX_ERR_ARG, "sphere radius must be positive");
        // Sphere::new (sphere.rs:28-44) with Transform::inverse swapping m and
        // m_inv (world2object), and swaps_handedness of object2world
        // (transform.rs): the theta range from the host's acos, as the
        // reference computes it at construction
        SphereShade& h = ssh[k];
        memcpy(h.w2o, q.o2w_inv, sizeof(h.w2o));
        memcpy(h.o2w, q.o2w, sizeof(h.o2w));
        const float r = q.radius;
        const float zmin = std::min(std::max(fminf(q.z0, q.z1), -r), r);
        const float zmax = std::min(std::max(fmaxf(q.z0, q.z1), -r), r);
        h.radius = r;
        h.thetamin = std::acos(zmin / r);
        h.thetamax = std::acos(zmax / r);
        h.phimax = std::min(std::max(q.phimax_deg, 0.0f), 360.0f) * PI_F / 180.0f;
        const float* a = q.o2w;
        const float det = a[0] * (a[5] * a[10] - a[6] * a[9]) - a[1] * (a[4] * a[10] - a[6] * a[8]) +
                          a[2] * (a[4] * a[9] - a[5] * a[8]);
        h.flip = ((q.reverse_orientation != 0) != (0.0f > det)) ? 1u : 0u;
        h.mat = d->sphere_mat[k];
        h.pad0 = h.pad1 = 0;
    }
    static_assert(sizeof(pbrtx_sphere) == sizeof(SphereIn), "sphere layout");
    int rc;
    if (d->split == SPLIT_GPU_LBVH) {
        // the tree built on the device (a linear BVH, bvh_gpu.hip)
        if ((rc = gpu_build_scene(d->verts, d->nverts, d->idx, d->nfaces, reinterpret_cast<const SphereIn*>(d->spheres),
                                  d->nspheres, &S->scene, &S->build_ms)))
            return rc;
    } else {
        std::string err;
        if (!build_bvh(d->verts, d->nverts, d->idx, d->nfaces, reinterpret_cast<const SphereIn*>(d->spheres),
                       d->nspheres, d->split, d->max_prims, &S->bvh, &err))
            return fail(PBRTX_ERR_BUILD, err);
        S->build_ms = S->bvh.build_ms;
        if ((rc = scene_from_host(S->bvh, &S->scene))) return rc;
        S->bvh = HostBvh{};  // the device copy is all that renders need
    }
    S->nfaces = d->nfaces;
    if ((rc = face_shade_device(d, &S->faces))) return rc;
    // the emitting faces' records on the host (face_frame is the device's
    // formula: the same bits)
    auto V = [&](uint32_t i) { return mk3(d->verts[3 * i], d->verts[3 * i + 1], d->verts[3 * i + 2]); };
    std::vector<std::vector<AreaTriRec>> per_light(d->n_area);
    for (uint32_t f = 0; d->face_light && f < d->nfaces; ++f) {
        const int32_t light = d->face_light[f];
        if (light < 0) continue;
        if ((uint32_t)light >= d->n_area) return fail(PBRTX_ERR_ARG, "face light out of range");
        const f3 p1 = V(d->idx[3 * f + 2]), p2 = V(d->idx[3 * f + 1]), p3 = V(d->idx[3 * f]);
        f3 nn, tn;
        face_frame(p1, p2, p3, &nn, &tn);
        AreaTriRec t{};
        t.p1[0] = p1.x; t.p1[1] = p1.y; t.p1[2] = p1.z;
        t.p2[0] = p2.x; t.p2[1] = p2.y; t.p2[2] = p2.z;
        t.p3[0] = p3.x; t.p3[1] = p3.y; t.p3[2] = p3.z;
        t.nn[0] = nn.x; t.nn[1] = nn.y; t.nn[2] = nn.z;
        t.area = 0.5f * sqrtf(lensq3(cross3(sub3(p2, p1), sub3(p3, p1))));  // mesh.rs:100-103
        per_light[(size_t)light].push_back(t);
    }
    std::vector<AreaTriRec> at;
    std::vector<uint32_t> afirst, acount;
    std::vector<float> atotal, ale;
    for (uint32_t l = 0; l < d->n_area; ++l) {
        if (per_light[l].empty()) return fail(PBRTX_ERR_ARG, "area light without triangles");
        afirst.push_back((uint32_t)at.size());
        acount.push_back((uint32_t)per_light[l].size());
        float total = 0.0f;
        for (auto& t : per_light[l]) { total = total + t.area; at.push_back(t); }
        atotal.push_back(total);
        for (int c = 0; c < 3; ++c) ale.push_back(d->area_le[3 * l + c]);
    }
    std::vector<MatRec> mats(d->nmats);
    for (uint32_t m = 0; m < d->nmats; ++m)
        mats[m] = MatRec{(uint32_t)d->mats[4 * m], d->mats[4 * m + 1], d->mats[4 * m + 2], d->mats[4 * m + 3]};
    std::vector<float> pts(d->point_lights ? d->point_lights : (const float*)nullptr,
                           d->point_lights ? d->point_lights + 6 * d->n_point : (const float*)nullptr);
    if ((rc = S->sph.upload(ssh)) || (rc = S->mats.upload(mats)) ||
        (rc = S->atris.upload(at)) ||
        (rc = S->a_first.upload(afirst)) || (rc = S->a_count.upload(acount)) || (rc = S->a_total.upload(atotal)) ||
        (rc = S->a_le.upload(ale)) || (rc = S->points.upload(pts)))
        return rc;
    RenderParams& P = S->base;
    P.n_point = d->n_point; P.n_area = d->n_area; P.nlights = d->n_point + d->n_area;
    P.light_samples = d->light_samples ? d->light_samples : 1;
    P.max_depth = d->max_depth; P.rr_start = d->rr_start;
    P.S = P.n_point + P.n_area * P.light_samples;
    P.faces = S->faces.p; P.mats = S->mats.p; P.atris = S->atris.p;
    P.sph = d->nspheres ? S->sph.p : nullptr;
    P.nfaces = d->nfaces;
    P.a_first = S->a_first.p; P.a_count = S->a_count.p; P.a_total = S->a_total.p; P.a_le = S->a_le.p;
    P.points = S->points.p;
    P.segs = nullptr; P.reqs = nullptr; P.nsegs = 0;
    S->integrator = d->integrator;
    return PBRTX_OK;
}

int camera_setup(RenderParams* P, const float* cam_to_world, float fov_deg, const float* screen, uint32_t xres,
                 uint32_t yres, uint32_t nx, uint32_t ny, uint32_t seed) {
    if (nx == 0 || ny == 0 || xres == 0 || yres == 0) return fail(PBRTX_ERR_ARG, "empty image");
    if ((uint64_t)xres * yres * nx * ny >= (1ull << 40)) return fail(PBRTX_ERR_ARG, "image too large");
    if (!perspective_r2c(xres, yres, screen, fov_deg, P->r2c)) return fail(PBRTX_ERR_ARG, "singular projection");
    memcpy(P->c2w, cam_to_world, sizeof(P->c2w));
    P->xres = xres; P->yres = yres; P->nx = nx; P->ny = ny; P->spp = nx * ny; P->seed = seed;
    P->dx = 1.0f / (float)nx; P->dy = 1.0f / (float)ny;
    return PBRTX_OK;
}

CamParams cam_params(const RenderParams& P, uint64_t g0) {
    CamParams cp;
    memcpy(cp.r2c, P.r2c, sizeof(cp.r2c));
    memcpy(cp.c2w, P.c2w, sizeof(cp.c2w));
    cp.xres = P.xres; cp.nx = P.nx; cp.spp = P.spp; cp.seed = P.seed;
    cp.dx = P.dx; cp.dy = P.dy; cp.g0 = g0;
    cp.spos = P.spos;
    cp.scount = P.scount;
    return cp;
}

size_t render_spill_bytes(const SceneView& sv, int path) {
    const size_t a = trace_spill_bytes(sv, path, g_tuning[path][0]);
    const size_t b = trace_spill_bytes(sv, path, g_tuning[path][1]);
    // camera rays: either scheme the frame's samples per pixel may pick
    const size_t c = trace_spill_bytes(sv, path, camera_tuning(path, 1));
    const size_t d = trace_spill_bytes(sv, path, camera_tuning(path, 1u << 20));
    return std::max(std::max(a, b), std::max(c, d));
}

hipError_t enqueue_film(const RenderParams& P, const float4* sample_l, float* rgb, hipStream_t st) {
    const dim3 grid((P.xres + FT - 1) / FT, (P.yres + FT - 1) / FT);
    if (P.ftab) {
        const uint64_t n = (uint64_t)P.xres * P.yres * P.spp;
        const uint64_t blocks = std::min<uint64_t>((n + RB - 1) / RB, 65536);
        hipLaunchKernelGGL(k_film_stage, dim3((unsigned)blocks), dim3(RB), 0, st, P, sample_l, n);
        hipLaunchKernelGGL(k_film_filter, grid, dim3(RB), 0, st, P, rgb);
    } else {
        hipLaunchKernelGGL(k_film, grid, dim3(RB), 0, st, P, sample_l, rgb);
    }
    return hipGetLastError();
}

// ---- the path tracer's launch sequence over one batch ---------------------------
int reserve_path_work(WorkBufs& W, uint32_t cap, bool osort) {
    ICHK(W.rays.ensure(cap));
    ICHK(W.hits.ensure(cap));
    ICHK(W.st.ensure(cap));
    ICHK(W.rays2.ensure(cap));
    ICHK(W.sh.ensure(cap));
    ICHK(W.contrib.ensure(cap));
    ICHK(W.occ.ensure((size_t)cap + 64));
    ICHK(W.pid.ensure(cap));
    ICHK(W.pid2.ensure(cap));
    ICHK(W.shpid.ensure(cap));
    ICHK(W.bcounts.ensure(2));
    if (osort) {
        ICHK(W.rays3.ensure(cap));
        ICHK(W.pid3.ensure(cap));
    }
    return PBRTX_OK;
}

int enqueue_path_batch(const RenderParams& P, const SceneView& sv, WorkBufs& W, uint32_t* cnt, uint64_t g0,
                       uint32_t n, int path, bool osort_req, float4* sample_l, TraceScratch* scr, hipStream_t st,
                       Stager& sg, const ShadowStream* ss) {
    if (n == 0) return PBRTX_OK;
    const bool batched = path == PATH_BATCHED;
    // shadow rays on their own stream (batched queues only)
    const bool split = batched && ss && ss->st && ss->scr && ss->ev;
    hipEvent_t joined = nullptr;  // the last shadow work enqueued on ss->st
    const TraceTuning& tunC = g_tuning[path][0];
    const TraceTuning& tunA = g_tuning[path][1];
    const TraceTuning tunCam = camera_tuning(path, P.spp);  // camera rays (coherent)
    const bool osort = batched && osort_req;
    ICHK(reserve_path_work(W, n, osort));
    // batched frames: the packet kernel makes the camera rays and the first
    // vertex makes them again and starts the paths (no init pass); served
    // batches (several cameras) start from an init pass
    const bool fused = batched && !P.segs && camera_fusable(sv, tunCam);
    // served batches of several requests: camera rays made in the packet
    // kernel from the segment table too, and the misses ended the same way
    const bool fused_seg = batched && P.segs && camera_fusable(sv, tunCam);
    CamTable ct;
    if (fused_seg) {
        static_assert(sizeof(RenderSeg) == sizeof(CamSeg) && offsetof(RenderSeg, req) == offsetof(CamSeg, cam) &&
                          offsetof(RenderSeg, g0) == offsetof(CamSeg, g0),
                      "segment layouts");
        static_assert(offsetof(ReqView, cam) == 0, "camera first in a request view");
        ct.segs = reinterpret_cast<const CamSeg*>(P.segs);
        ct.cams = reinterpret_cast<const uint8_t*>(P.reqs);
        ct.nsegs = P.nsegs;
        ct.stride = (uint32_t)sizeof(ReqView);
        ICHK(W.hits2.ensure(n));
    }
    hipEvent_t a = sg.begin();
    if (fused || fused_seg) {
        // fused: the camera kernel lists the hits and counts them here
        RCHK(fill_u32(cnt + 1, fused ? 0u : n, 1, st));
    } else {
        hipLaunchKernelGGL(k_path_init, dim3(grid_for(n)), dim3(RB), 0, st, P, g0, n, W.rays.p, W.st.p,
                           batched ? W.pid.p : nullptr, batched ? cnt + 1 : nullptr);
        RCHK(hipGetLastError());
    }
    RCHK(sg.end(a, &Timing::other));
    Ray* rin = W.rays.p;
    Ray* rout = W.rays2.p;
    Ray* rspare = W.rays3.p;
    uint32_t* pin = W.pid.p;
    uint32_t* pout = W.pid2.p;
    uint32_t* pspare = W.pid3.p;
    uint32_t* cin = cnt + 1;
    uint32_t* cout = cnt + 2;
    for (uint32_t bounce = 0; bounce <= P.max_depth; ++bounce) {
        a = sg.begin();
        const TraceTuning& tb = (bounce == 0 && !P.segs) ? tunCam : tunC;  // frames' bounce 0: camera rays
        if (fused && bounce == 0) {
            // a missed camera ray's path ends in the traversal (zero
            // radiance); the hits' path ids become the queue
            const CamOut co{sample_l + g0, pin, cin};
            RCHK(trace_camera(sv, cam_params(P, g0), n, W.hits.p, tb, scr, st, nullptr, &co));
        } else if (fused_seg && bounce == 0) {
            RCHK(trace_camera(sv, CamParams{}, n, W.hits.p, tunCam, scr, st, &ct));
        } else if (batched) {
            RCHK(trace_closest(sv, rin, n, W.hits.p, path, tb, scr, st, cin));
        } else {
            RCHK(trace_closest(sv, rin, n, W.hits.p, path, tb, scr, st));
        }
        RCHK(sg.end(a, &Timing::closest));
        if (fused_seg && bounce == 0) {
            // the misses end here; the hits' path ids become the queue
            a = sg.begin();
            const uint32_t mb = (uint32_t)(((uint64_t)n + MISS_B * MISS_PER - 1) / (MISS_B * MISS_PER));
            RCHK(fill_u32(cin, 0u, 1, st));  // the list's count
            hipLaunchKernelGGL(k_path_miss0<true>, dim3(mb), dim3(MISS_B), 0, st, P, g0, n, W.hits.p, sample_l,
                               pin, cin, rin, W.hits2.p, W.st.p);
            RCHK(hipGetLastError());
            RCHK(sg.end(a, &Timing::other));
        }
        // the previous bounce's shadow work has read the shadow queue and
        // added its results to the path states the vertex kernel reads
        if (joined) RCHK(hipStreamWaitEvent(st, joined, 0));
        a = sg.begin();
        {
            const uint32_t nb = (n + VERTEX_B - 1) / VERTEX_B;
            // the two queues' counts and (queue form) the tile ticket
            ICHK(W.bcounts.ensure(2));
            RCHK(zero3_u32(cout, 1, cnt + 5, 1, W.bcounts.p, batched ? 1 : 0, st));
            // queue form: persistent blocks over the live paths' tiles
            const uint32_t grid = batched ? std::min<uint32_t>(nb, (uint32_t)device_cu_count() * 8u) : nb;
            hipLaunchKernelGGL(k_path_shade, dim3(grid), dim3(VERTEX_B), 0, st, P, g0, bounce, n,
                               batched ? (const uint32_t*)cin : (const uint32_t*)nullptr,
                               batched ? (const uint32_t*)pin : (const uint32_t*)nullptr,
                               (fused && bounce == 0) ? (const Ray*)nullptr : rin,
                               (fused_seg && bounce == 0) ? W.hits2.p : W.hits.p, W.st.p, rout,
                               batched ? pout : nullptr, cout, W.sh.p, batched ? W.shpid.p : nullptr, W.contrib.p,
                               cnt + 5, sample_l, W.bcounts.p, (int)(batched ? PSHADE_QUEUE : PSHADE_PLAIN));
            RCHK(hipGetLastError());
        }
        RCHK(sg.end(a, &Timing::other));
        // shadow rays: occlusion, then the unoccluded samples added
        hipStream_t sst = split ? ss->st : st;
        TraceScratch* sscr = split ? ss->scr : scr;
        Stager& ssg = split ? *ss->sg : sg;
        if (split) {
            hipEvent_t fork = ss->ev->next();
            if (!fork) return fail(PBRTX_ERR_OOM, "events");
            RCHK(hipEventRecord(fork, st));
            RCHK(hipStreamWaitEvent(sst, fork, 0));
        }
        a = ssg.begin();
        if (batched)
            RCHK(trace_any_bytes(sv, W.sh.p, n, W.occ.p, path, tunA, sscr, sst, cnt + 5));
        else
            RCHK(trace_any_bytes(sv, W.sh.p, n, W.occ.p, path, tunA, sscr, sst));
        RCHK(ssg.end(a, &Timing::any));
        a = ssg.begin();
        hipLaunchKernelGGL(k_path_shadow, dim3(grid_for(n)), dim3(RB), 0, sst, P, g0, n,
                           batched ? cnt + 5 : (const uint32_t*)nullptr,
                           batched ? W.shpid.p : (const uint32_t*)nullptr, W.contrib.p, W.occ.p, W.st.p, sample_l);
        RCHK(hipGetLastError());
        hipLaunchKernelGGL(k_add_counter, dim3(1), dim3(1), 0, sst, cnt + 6, cnt + 5, cnt + 7, false, (uint32_t*)nullptr);
        RCHK(hipGetLastError());
        if (split) {
            RCHK(ssg.end(a, &Timing::other));
            joined = ss->ev->next();
            if (!joined) return fail(PBRTX_ERR_OOM, "events");
            RCHK(hipEventRecord(joined, sst));
            a = sg.begin();
        }
        if (bounce < P.max_depth)
            hipLaunchKernelGGL(k_add_counter, dim3(1), dim3(1), 0, st, cnt + 8, cout, cnt + 9, false, (uint32_t*)nullptr);
        RCHK(hipGetLastError());
        if (osort && bounce < P.max_depth) {
            const uint32_t tiles = (n + OS_TILE - 1) / OS_TILE;
            const uint32_t grid = std::min<uint32_t>(tiles, (uint32_t)device_cu_count() * 2u);
            hipLaunchKernelGGL(k_octant_sort, dim3(grid), dim3(OS_BLOCK), 0, st, (const uint32_t*)cout,
                               (const Ray*)rout, (const uint32_t*)pout, rspare, pspare);
            RCHK(hipGetLastError());
            std::swap(rout, rspare);
            std::swap(pout, pspare);
        }
        RCHK(sg.end(a, &Timing::other));
        std::swap(rin, rout);
        std::swap(pin, pout);
        std::swap(cin, cout);
    }
    // the batch ends when its last shadow work has
    if (joined) RCHK(hipStreamWaitEvent(st, joined, 0));
    return PBRTX_OK;
}

// ---- the frame renderer ----------------------------------------------------------
struct Renderer {
    RenderScene sc;
    RenderParams P{};
    uint64_t n_samples = 0;
    uint32_t chunk = 1u << 22;
    WorkBufs W;
    DevArr<float4> sample_l;
    DevArr<uint32_t> counters;
    DevArr<float> rgb;
    DevArr<float> ftab;  // the film filter's weight table (none: box 0.5)
    DevArr<float2> spos;  // LDSampler, HaltonSampler: the samples' positions within their pixels
    DevArr<uint32_t> scount;  // HaltonSampler: samples of each pixel
    DevArr<float2> fpos;  // and its staging of the frame's samples
    DevArr<float4> fxyz;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // the path tracer's shadow rays
    EventPool events;
    bool octant_sort = false;  // path tracer, batched: octant-sort the extension queue
    bool have_frame = false;   // rgb holds a finished frame
    // graph mode, per path: warm after one eager frame; the captured frame
    // and the tuning generation it was captured under
    bool use_graph = false;
    bool warm[2] = {false, false};
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    uint32_t gen[2] = {0, 0};
    uint64_t gcamera[2] = {0, 0};
    void drop_graphs() {
        for (int k = 0; k < 2; ++k) {
            if (gexec[k]) (void)hipGraphExecDestroy(gexec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            gexec[k] = nullptr;
            graph[k] = nullptr;
            warm[k] = false;
        }
    }
};

// HaltonSampler::new over the whole image (halton.rs:18-29): max(dx, dy)^2 *
// spp sequence samples, those inside the image kept; K slots a pixel, K the
// most any pixel received.
static int halton_setup(Renderer* R, const pbrtx_render_desc* d) {
    const uint32_t spp = d->nx * d->ny, xres = d->xres, yres = d->yres;
    const uint64_t side = std::max(xres, yres), wanted = side * side * spp;
    if (wanted >= (1ull << 36)) return fail(PBRTX_ERR_ARG, "Halton sampler: image too large");
    const float dx = (float)xres, dy = (float)yres;
    const float delta = dy > dx ? dy : dx;  // dy.max(dx)
    const uint32_t npix = xres * yres;
    ICHK(R->scount.ensure(npix));
    hipStream_t st = R->stream;
    RCHK(fill_u32(R->scount.p, 0u, npix, st));
    const uint32_t gk = (uint32_t)std::min<uint64_t>((wanted + RB - 1) / RB, 65536u);
    hipLaunchKernelGGL(k_halton_bin, dim3(gk), dim3(RB), 0, st, wanted, xres, yres, delta, R->scount.p,
                       (uint64_t*)nullptr, 0u);
    RCHK(hipGetLastError());
    std::vector<uint32_t> cnt(npix);
    RCHK(hipMemcpyAsync(cnt.data(), R->scount.p, npix * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    RCHK(hipStreamSynchronize(st));
    uint32_t K = 0;
    for (uint32_t c : cnt) K = std::max(K, c);
    if (K == 0) return fail(PBRTX_ERR_ARG, "Halton sampler: no samples");
    if ((uint64_t)npix * K >= (1ull << 40)) return fail(PBRTX_ERR_ARG, "image too large");
    DevArr<uint64_t> slots;
    ICHK(slots.ensure((size_t)npix * K));
    ICHK(R->spos.ensure((size_t)npix * K));
    RCHK(fill_u32(R->scount.p, 0u, npix, st));
    hipLaunchKernelGGL(k_halton_bin, dim3(gk), dim3(RB), 0, st, wanted, xres, yres, delta, R->scount.p, slots.p, K);
    RCHK(hipGetLastError());
    const uint32_t gp = std::min<uint32_t>((npix + RB - 1) / RB, 65536u);
    hipLaunchKernelGGL(k_halton_pos, dim3(gp), dim3(RB), 0, st, npix, xres, yres, delta, R->scount.p, slots.p, K,
                       R->spos.p);
    RCHK(hipGetLastError());
    RCHK(hipStreamSynchronize(st));
    slots.release();
    R->P.nx = K; R->P.ny = 1; R->P.spp = K;
    R->P.dx = 1.0f / (float)K; R->P.dy = 1.0f;
    R->P.spos = R->spos.p;
    R->P.scount = R->scount.p;
    return PBRTX_OK;
}

static int renderer_setup(Renderer* R, const pbrtx_render_desc* d) {
    if (!d->verts || !d->idx || !d->face_mat || !d->mats) return fail(PBRTX_ERR_ARG, "null scene arrays");
    if (d->nx == 0 || d->ny == 0 || d->xres == 0 || d->yres == 0) return fail(PBRTX_ERR_ARG, "empty image");
    ICHK(render_scene_setup(&R->sc, d));
    // the renderer's own stream (default priority; a frame ends by waiting
    // for its last event, so host reads after a frame see its results)
    {
        hipError_t e = hipStreamCreateWithFlags(&R->stream, hipStreamNonBlocking);
        if (e == hipSuccess && d->integrator == 1) e = hipStreamCreateWithFlags(&R->stream2, hipStreamNonBlocking);
        if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    }
    R->P = R->sc.base;
    ICHK(camera_setup(&R->P, d->cam_to_world, d->fov_deg, d->screen, d->xres, d->yres, d->nx, d->ny, d->seed));
    if (d->sampler == 1) {
        // LDSampler::new (lds.rs:17-24): samples per pixel rounded up to a
        // power of two; positions made once, as they depend on (seed, pixel)
        const uint32_t spp = next_pow2(d->nx * d->ny);
        if ((uint64_t)d->xres * d->yres * spp >= (1ull << 40)) return fail(PBRTX_ERR_ARG, "image too large");
        R->P.nx = spp; R->P.ny = 1; R->P.spp = spp;
        R->P.dx = 1.0f / (float)spp; R->P.dy = 1.0f;
        const uint32_t npix = d->xres * d->yres;
        ICHK(R->spos.ensure((size_t)npix * spp));
        const uint32_t blocks = std::min<uint32_t>((npix + RB - 1) / RB, 65536u);
        hipLaunchKernelGGL(k_ld_points, dim3(blocks), dim3(RB), 0, R->stream, d->seed, npix, spp, R->spos.p);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(R->stream);
        if (e != hipSuccess) return hip_fail(e, "LDSampler positions");
        R->P.spos = R->spos.p;
    } else if (d->sampler == 2) {
        ICHK(halton_setup(R, d));
    } else if (d->sampler != 0) {
        return fail(PBRTX_ERR_ARG, "unknown sampler (0 stratified, 1 low-discrepancy, 2 Halton)");
    }
    R->n_samples = (uint64_t)d->xres * d->yres * R->P.spp;
    return PBRTX_OK;
}

// Enqueues one frame's device work on the renderer's stream: every kernel and
// fill from the counter reset to the film, nothing that waits on the host, so
// it can be captured as a graph.  sg: per-stage events (or none).
static int enqueue_frame(Renderer* R, int path, TraceScratch* scr, Stager& sg, uint64_t* camera_rays,
                         const ShadowStream* ss) {
    RenderParams& P = R->P;
    hipStream_t st = R->stream;
    WorkBufs& W = R->W;
    const uint32_t spp = P.spp;
    // chunk: whole pixels, at most R->chunk samples
    uint64_t per = (uint64_t)R->chunk / spp * spp;
    if (per < spp) per = spp;
    if (per > R->n_samples) per = R->n_samples;
    const uint32_t cap = (uint32_t)per;
    ICHK(R->sample_l.ensure(R->n_samples));
    ICHK(W.rays.ensure(cap));
    ICHK(W.hits.ensure(cap));
    ICHK(R->counters.ensure(N_COUNTERS));
    const bool batched = path == PATH_BATCHED;
    const TraceTuning& tunA = g_tuning[path][1];
    const TraceTuning tunCam = camera_tuning(path, P.spp);  // camera rays (coherent)
    uint32_t* cnt = R->counters.p;
    RCHK(fill_u32(cnt, 0u, N_COUNTERS, st));
    const SceneView& sv = R->sc.scene.view;
    uint64_t total_camera = 0;
    for (uint64_t g0 = 0; g0 < R->n_samples; g0 += per) {
        const uint32_t n = (uint32_t)((R->n_samples - g0) < per ? (R->n_samples - g0) : per);
        total_camera += n;
        if (R->sc.integrator != 0) {
            ICHK(enqueue_path_batch(P, sv, W, cnt, g0, n, path, R->octant_sort, R->sample_l.p, scr, st, sg, ss));
            continue;
        }
        const size_t slots = (size_t)n * P.S;
        ICHK(W.sh.ensure(slots ? slots : 1));
        ICHK(W.contrib.ensure(slots ? slots : 1));
        ICHK(W.qidx.ensure(slots ? slots : 1));
        ICHK(W.lbase.ensure(n));
        ICHK(W.occ.ensure((size_t)slots + 64));
        hipEvent_t a;
        // batched: the packet kernel makes the camera rays itself and the
        // shade kernel makes them again (rays = null): no ray array
        const bool fused = batched && camera_fusable(sv, tunCam);
        Ray* crays = fused ? nullptr : W.rays.p;
        if (!fused) {
            a = sg.begin();
            hipLaunchKernelGGL(k_camera, dim3(grid_for(n)), dim3(RB), 0, st, P, g0, n, W.rays.p);
            RCHK(hipGetLastError());
            RCHK(sg.end(a, &Timing::other));
        }
        a = sg.begin();
        if (fused)
            RCHK(trace_camera(sv, cam_params(P, g0), n, W.hits.p, tunCam, scr, st));
        else
            RCHK(trace_closest(sv, W.rays.p, n, W.hits.p, path, tunCam, scr, st));
        RCHK(sg.end(a, &Timing::closest));
        // batched: a sample whose camera ray missed ends here (zero radiance);
        // the others are listed, and shading and accumulation run over the
        // list.  Not for small chunks, where the extra launch costs more than
        // it saves (512x512 at 1 spp: 0.348 -> 0.354 ms a frame).
        const uint32_t* list = nullptr;
        const uint32_t* nlist = nullptr;
        // (the count of the list of hits and that of the shadow rays start
        // at zero: cleared with the frame's counters, and again at the end
        // of each chunk)
        const bool listed = batched && n >= (1u << 20);
        if (listed) {
            a = sg.begin();
            ICHK(W.pid.ensure(n));
            const uint32_t mb = (uint32_t)(((uint64_t)n + MISS_B * MISS_PER - 1) / (MISS_B * MISS_PER));
            hipLaunchKernelGGL(k_path_miss0<false>, dim3(mb), dim3(MISS_B), 0, st, P, g0, n, W.hits.p, R->sample_l.p,
                               W.pid.p, cnt + 1, (Ray*)nullptr, (Hit*)nullptr, (PathState*)nullptr);
            RCHK(hipGetLastError());
            list = W.pid.p;
            nlist = cnt + 1;
            RCHK(sg.end(a, &Timing::other));
        }
        a = sg.begin();
        {
            const uint32_t nb = (n + SHADE_B - 1) / SHADE_B;
            hipLaunchKernelGGL(k_shade_direct, dim3(nb), dim3(SHADE_B), 0, st, P, g0, n, crays, W.hits.p, W.sh.p,
                               W.contrib.p, W.qidx.p, W.lbase.p, cnt + 5, (int)(batched ? SHADE_QUEUE : SHADE_PLAIN),
                               list, nlist);
            RCHK(hipGetLastError());
        }
        RCHK(sg.end(a, &Timing::other));
        if (slots) {
            a = sg.begin();
            if (batched)
                RCHK(trace_any_bytes(sv, W.sh.p, (uint32_t)slots, W.occ.p, path, tunA, scr, st, cnt + 5));
            else
                RCHK(trace_any_bytes(sv, W.sh.p, (uint32_t)slots, W.occ.p, path, tunA, scr, st));
            RCHK(sg.end(a, &Timing::any));
        }
        a = sg.begin();
        hipLaunchKernelGGL(k_accum_direct, dim3(grid_for(n)), dim3(RB), 0, st, P, g0, n, W.lbase.p, W.contrib.p,
                           W.qidx.p, W.occ.p, R->sample_l.p, list, nlist);
        RCHK(hipGetLastError());
        // running total of shadow rays; the shadow-ray and list counts
        // cleared for the next chunk
        hipLaunchKernelGGL(k_add_counter, dim3(1), dim3(1), 0, st, cnt + 6, cnt + 5, cnt + 7, true, cnt + 1);
        RCHK(hipGetLastError());
        RCHK(sg.end(a, &Timing::other));
    }
    ICHK(R->rgb.ensure((size_t)P.xres * P.yres * 3));
    {
        hipEvent_t a = sg.begin();
        RCHK(enqueue_film(P, R->sample_l.p, R->rgb.p, st));
        RCHK(sg.end(a, &Timing::other));
    }
    *camera_rays = total_camera;
    return PBRTX_OK;
}

// One frame: eagerly with per-stage timing, or -- graph mode, after one eager
// frame of that path has sized every buffer -- as a captured graph replayed
// (captured again when the launch tuning changed since).
static int render_frame(Renderer* R, int path, pbrtx_render_stats* stats) {
    hipStream_t st = R->stream;
    TraceScratch* scr;
    ICHK(stream_scratch(st, render_spill_bytes(R->sc.scene.view, path), &scr));
    TraceScratch* scr2 = nullptr;
    if (R->stream2) ICHK(stream_scratch(R->stream2, render_spill_bytes(R->sc.scene.view, path), &scr2));
    R->events.used = 0;
    Timing T;
    uint64_t total_camera = 0;
    hipEvent_t e_start = R->events.next(), e_end = R->events.next();
    if (!e_start || !e_end) return fail(PBRTX_ERR_OOM, "events");
    const bool replay = R->use_graph && R->warm[path];
    if (replay && (!R->gexec[path] || R->gen[path] != g_tuning_gen)) {
        if (R->gexec[path]) { (void)hipGraphExecDestroy(R->gexec[path]); R->gexec[path] = nullptr; }
        if (R->graph[path]) { (void)hipGraphDestroy(R->graph[path]); R->graph[path] = nullptr; }
        RCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        Stager none, none2;
        const ShadowStream ss{R->stream2, scr2, &R->events, &none2};
        const int rc = enqueue_frame(R, path, scr, none, &R->gcamera[path], &ss);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        RCHK(ec);
        R->graph[path] = g;
        RCHK(hipGraphInstantiate(&R->gexec[path], g, nullptr, nullptr, 0));
        R->gen[path] = g_tuning_gen;
    }
    RCHK(hipEventRecord(e_start, st));
    if (replay) {
        RCHK(hipGraphLaunch(R->gexec[path], st));
        total_camera = R->gcamera[path];
    } else {
        Stager sg{&R->events, &T, st}, sg2{&R->events, &T, R->stream2};
        const ShadowStream ss{R->stream2, scr2, &R->events, &sg2};
        ICHK(enqueue_frame(R, path, scr, sg, &total_camera, &ss));
        R->warm[path] = true;
    }
    RCHK(hipEventRecord(e_end, st));
    RCHK(hipEventSynchronize(e_end));
    uint32_t* cnt = R->counters.p;
    if (stats) {
        float ms = 0.0f;
        RCHK(hipEventElapsedTime(&ms, e_start, e_end));
        stats->ms_total = ms;
        auto sum = [&](const EventPairs& v, double* out) -> int {
            double t = 0.0;
            for (auto& pr : v) {
                float x = 0.0f;
                RCHK(hipEventElapsedTime(&x, pr.first, pr.second));
                t += x;
            }
            *out = t;
            return PBRTX_OK;
        };
        ICHK(sum(T.closest, &stats->ms_trace_closest));
        ICHK(sum(T.any, &stats->ms_trace_any));
        ICHK(sum(T.other, &stats->ms_other));
        stats->launches_closest = (uint32_t)T.closest.size();
        stats->launches_any = (uint32_t)T.any.size();
        uint32_t c[10];
        RCHK(hipMemcpy(c, cnt, sizeof(c), hipMemcpyDeviceToHost));
        stats->camera_rays = total_camera;
        stats->shadow_rays = (uint64_t)c[6] | ((uint64_t)c[7] << 32);
        stats->extension_rays = (uint64_t)c[8] | ((uint64_t)c[9] << 32);
        stats->build_ms = R->sc.build_ms;
    }
    R->have_frame = true;
    return PBRTX_OK;
}

}  // namespace px

using namespace px;

struct pbrtx_renderer {
    Renderer r;
};

extern "C" {

int pbrtx_renderer_create(const pbrtx_render_desc* desc, pbrtx_renderer** out) {
    if (!desc || !out) return fail(PBRTX_ERR_ARG, "null argument");
    *out = nullptr;
    pbrtx_renderer* h = new (std::nothrow) pbrtx_renderer();
    if (!h) return fail(PBRTX_ERR_OOM, "host allocation");
    int rc = renderer_setup(&h->r, desc);
    if (rc) {
        pbrtx_renderer_destroy(h);
        return rc;
    }
    *out = h;
    return PBRTX_OK;
}

int pbrtx_renderer_set_chunk(pbrtx_renderer* h, uint32_t samples_per_chunk) {
    if (!h || samples_per_chunk == 0) return fail(PBRTX_ERR_ARG, "bad chunk");
    h->r.chunk = samples_per_chunk;
    h->r.drop_graphs();  // the frame's shape changed
    return PBRTX_OK;
}

int pbrtx_renderer_set_octant_sort(pbrtx_renderer* h, int32_t enable) {
    if (!h) return fail(PBRTX_ERR_ARG, "renderer is null");
    h->r.octant_sort = enable != 0;
    h->r.drop_graphs();  // the frame's kernels changed
    return PBRTX_OK;
}

// How many pixels either way a sample reaches with half-width w: a sample
// of pixel q lies in [q, q+1], so its extent [ceil(d - w), floor(d + w)]
// with d = ix - 1/2 stays within floor(w + 1/2) of q -- unless rounding d +- w
// crosses an integer, which needs w + 1/2 within an ulp of the coordinates
// (2^-7 below 2^16) under the next integer: one more then.
static int32_t filter_reach(float w) {
    const float r = std::floor(w + 0.5f);
    return (int32_t)r + ((w + 0.5f) - r > 1.0f - 0x1p-6f ? 1 : 0);
}

int pbrtx_renderer_set_filter(pbrtx_renderer* h, int32_t kind, float xw, float yw, float p0, float p1) {
    if (!h) return fail(PBRTX_ERR_ARG, "renderer is null");
    FilmFilter f;
    f.kind = kind; f.xw = xw; f.yw = yw; f.p0 = p0; f.p1 = p1;
    if (!filter_valid(f)) return fail(PBRTX_ERR_ARG, "unknown filter kind or width not positive");
    Renderer& R = h->r;
    if (R.P.xres > 65536u || R.P.yres > 65536u || xw > 4096.0f || yw > 4096.0f)
        return fail(PBRTX_ERR_ARG, "filtered films: at most 65536 pixels a side and half-widths up to 4096");
    R.drop_graphs();  // the frame's film launch changes
    if (kind == 0 && xw == 0.5f && yw == 0.5f) {
        R.P.ftab = nullptr;  // the box filter's own film kernel
        return PBRTX_OK;
    }
    std::vector<float> t(FILTER_TABLE_DIM * FILTER_TABLE_DIM);
    filter_weight_table(f, t.data());
    ICHK(R.ftab.upload(t));
    ICHK(R.fpos.ensure(R.n_samples));
    ICHK(R.fxyz.ensure(R.n_samples));
    R.P.ftab = R.ftab.p;
    R.P.fpos = R.fpos.p;
    R.P.fxyz = R.fxyz.p;
    R.P.fxw = xw; R.P.fyw = yw;
    R.P.fixw = 1.0f / xw; R.P.fiyw = 1.0f / yw;
    R.P.frx = filter_reach(xw);
    R.P.fry = filter_reach(yw);
    return PBRTX_OK;
}

int pbrtx_filter_eval(int32_t kind, float xw, float yw, float p0, float p1, float x, float y, float* out) {
    FilmFilter f;
    f.kind = kind; f.xw = xw; f.yw = yw; f.p0 = p0; f.p1 = p1;
    if (!filter_valid(f) || !out) return fail(PBRTX_ERR_ARG, "unknown filter kind, width not positive or null output");
    *out = filter_evaluate(f, x, y);
    return PBRTX_OK;
}

int pbrtx_filter_table(int32_t kind, float xw, float yw, float p0, float p1, float* out256) {
    FilmFilter f;
    f.kind = kind; f.xw = xw; f.yw = yw; f.p0 = p0; f.p1 = p1;
    if (!filter_valid(f) || !out256) return fail(PBRTX_ERR_ARG, "unknown filter kind, width not positive or null output");
    filter_weight_table(f, out256);
    return PBRTX_OK;
}

int pbrtx_renderer_set_graph(pbrtx_renderer* h, int32_t enable) {
    if (!h) return fail(PBRTX_ERR_ARG, "renderer is null");
    h->r.use_graph = enable != 0;
    if (!enable) h->r.drop_graphs();
    return PBRTX_OK;
}

int pbrtx_renderer_render(pbrtx_renderer* h, int32_t path, float* rgb_out, pbrtx_render_stats* stats) {
    if (!h) return fail(PBRTX_ERR_ARG, "renderer is null");
    if (path != PATH_PLAIN && path != PATH_BATCHED) return fail(PBRTX_ERR_ARG, "unknown path");
    if (stats) memset(stats, 0, sizeof(*stats));
    int rc = render_frame(&h->r, path, stats);
    if (rc) return rc;
    if (rgb_out) {
        const size_t bytes = (size_t)h->r.P.xres * h->r.P.yres * 3 * sizeof(float);
        hipError_t e = hipMemcpy(rgb_out, h->r.rgb.p, bytes, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpy(rgb)");
    }
    return PBRTX_OK;
}

int pbrtx_renderer_sample_slots(pbrtx_renderer* h, uint32_t* slots_per_pixel) {
    if (!h || !slots_per_pixel) return fail(PBRTX_ERR_ARG, "null argument");
    *slots_per_pixel = h->r.P.spp;
    return PBRTX_OK;
}

int pbrtx_renderer_write_image(pbrtx_renderer* h, const char* filename) {
    if (!h || !filename) return fail(PBRTX_ERR_ARG, "null argument");
    Renderer& R = h->r;
    if (!R.have_frame) return fail(PBRTX_ERR_STATE, "no frame rendered yet");
    std::vector<float> rgb((size_t)R.P.xres * R.P.yres * 3);
    hipError_t e = hipMemcpy(rgb.data(), R.rgb.p, rgb.size() * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(rgb)");
    return pbrtx_write_image(filename, rgb.data(), R.P.xres, R.P.yres);
}

int pbrtx_renderer_sample_radiance(pbrtx_renderer* h, float* out, uint64_t count) {
    if (!h || !out) return fail(PBRTX_ERR_ARG, "null argument");
    if (count > h->r.n_samples || !h->r.sample_l.p) return fail(PBRTX_ERR_ARG, "count exceeds the frame");
    std::vector<float4> tmp(count);
    hipError_t e = hipMemcpy(tmp.data(), h->r.sample_l.p, count * sizeof(float4), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(sample radiance)");
    for (uint64_t i = 0; i < count; ++i) {
        out[3 * i] = tmp[i].x; out[3 * i + 1] = tmp[i].y; out[3 * i + 2] = tmp[i].z;
    }
    return PBRTX_OK;
}

int pbrtx_renderer_destroy(pbrtx_renderer* h) {
    if (!h) return PBRTX_OK;
    Renderer& R = h->r;
    (void)hipDeviceSynchronize();
    R.drop_graphs();
    R.sc.release();
    R.W.release();
    R.sample_l.release(); R.counters.release(); R.rgb.release(); R.ftab.release(); R.fpos.release(); R.fxyz.release(); R.spos.release(); R.scount.release();
    for (hipStream_t st : {R.stream, R.stream2})
        if (st) {
            (void)hipStreamSynchronize(st);
            release_stream_scratch(st);
            (void)hipStreamDestroy(st);
        }
    delete h;
    return PBRTX_OK;
}

}  // extern "C"
