// trace.hip -- BVH traversal kernels for gfx950 (closest hit and occlusion).
//
// Both paths reproduce BVHAccelerator::intersect (bvh.rs:375-422) exactly:
// near-first order chosen by the sign of 1/d on the node's split axis, boxes
// culled against the *current* maxt, leaf prims tested in order, a later
// equal-t hit replacing an earlier one (mesh.rs:71).  Occlusion is the
// reference's intersect(r).is_some() (intersection.rs:63-65) with an early
// exit, which yields the same bit.
//
// PATH_PLAIN   one lane per ray, static schedule, the 32-byte reference node
//              record, box test at pop time (exactly the reference loop), a
//              per-lane LDS stack of node indices with a global spill.
// PATH_BATCHED persistent waves pulling rays from a global counter and
//              refilling idle lanes (wave ballot + prefix rank), 128-byte
//              four-wide records (two reference levels per fetch).  The four
//              grandchild boxes are tested when their record is fetched; a
//              passing slot is pushed with its entry distance in the
//              reference's visiting order and re-checked against the current
//              tmax when popped -- the reference's pop-time test.  Skipping
//              the middle level's box test changes nothing: a box contains
//              its children's boxes, and the rounded slab distances are
//              monotone in the box bounds, so whenever a grandchild passes,
//              its parent passes too (at the same or a larger tmax).
//              Occlusion walks its own records (the same leaves under the
//              builder's own tree): by the same monotonicity a leaf whose
//              exact box passes has every box above it pass, in any tree,
//              so the walk reaches exactly the leaves the reference's does
//              and its answer, which no visiting order changes, is the same.
#include <mutex>

#include "camera.hpp"
#include "trace.hpp"

namespace px {
namespace {

constexpr int STK_PLAIN = 24;   // LDS entries per lane (u32 node index)
// LDS traversal stack entries per lane of the persistent per-ray kernel
// (ref + entry distance; deeper entries spill to global memory): a launch
// parameter, one kernel instance per supported depth
constexpr int STK_DEFAULT = 16;
inline bool stack_supported(int e) { return e == 8 || e == 16 || e == 24; }
constexpr int MAX_BLOCK = 256;

__device__ __forceinline__ float as_f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t as_u(float f) { return __float_as_uint(f); }

// Slab interval of one box (bbox.rs:185-209): returns entry tn and exit tf
// over the three axes and [tmin, tmax].  The reference rejects as soon as
// t0 > t1 after any axis; t0 only grows and t1 only shrinks, so testing once
// after the three axes gives the same answer.  The swap keeps the
// reference's NaN behaviour (a NaN slab bound is ignored by max / min).
__device__ __forceinline__ bool slab(float lx, float hx, float ly, float hy, float lz, float hz,
                                     float ox, float oy, float oz, float ix, float iy, float iz,
                                     float tmin, float tmax, float* tn_out) {
    float ax = (lx - ox) * ix, bx = (hx - ox) * ix;
    float ay = (ly - oy) * iy, by = (hy - oy) * iy;
    float az = (lz - oz) * iz, bz = (hz - oz) * iz;
    float nx = ax > bx ? bx : ax, fx = ax > bx ? ax : bx;
    float ny = ay > by ? by : ay, fy = ay > by ? ay : by;
    float nz = az > bz ? bz : az, fz = az > bz ? az : bz;
    float t0 = fmaxf(fmaxf(fmaxf(tmin, nx), ny), nz);
    float t1 = fminf(fminf(fminf(tmax, fx), fy), fz);
    *tn_out = t0;
    return !(t0 > t1);
}

// The same interval with each axis's planes given near-first by the sign of
// 1/d.  For 1/d >= 0 (or NaN) the reference's compare-and-swap keeps lo, hi
// in that order for any ray.  For 1/d < 0 it swaps them exactly when no slab
// distance is NaN: with a finite 1/d, a finite origin and finite planes, (b -
// o) * id is monotone in b and the swap result equals hi, lo up to the sign
// of a zero distance, which no comparison here sees.  Callers take this form
// only for such rays (cq_fast_ray in a scene passing cq_fast_scene).
__device__ __forceinline__ bool slab_ord(float nx, float fx, float ny, float fy, float nz, float fz,
                                         float ox, float oy, float oz, float ix, float iy, float iz,
                                         float tmin, float tmax) {
    const float ax = (nx - ox) * ix, bx = (fx - ox) * ix;
    const float ay = (ny - oy) * iy, by = (fy - oy) * iy;
    const float az = (nz - oz) * iz, bz = (fz - oz) * iz;
    const float t0 = fmaxf(fmaxf(fmaxf(tmin, ax), ay), az);
    const float t1 = fminf(fminf(fminf(tmax, bx), by), bz);
    return !(t0 > t1);
}

struct RayState {
    float ox, oy, oz, dx, dy, dz, ix, iy, iz, tmin, tmax;
};

// Triangle::get_intersection_point (mesh.rs:41-72) with e1, e2 precomputed.
__device__ __forceinline__ bool tri_test(const float4 a, const float4 b, const float4 c,
                                         const RayState& r, float* t, float* b1o, float* b2o) {
    const f3 p1 = mk3(a.x, a.y, a.z);
    const f3 e1 = mk3(a.w, b.x, b.y);
    const f3 e2 = mk3(b.z, b.w, c.x);
    const f3 d = mk3(r.dx, r.dy, r.dz);
    const f3 s1 = cross3(d, e2);
    const float div = dot3(s1, e1);
    if (div == 0.0f) return false;
    const float inv = 1.0f / div;
    const f3 s = sub3(mk3(r.ox, r.oy, r.oz), p1);
    const float b1 = dot3(s1, s) * inv;
    if (b1 < 0.0f || b1 > 1.0f) return false;
    const f3 s2 = cross3(s, e1);
    const float b2 = dot3(d, s2) * inv;
    if (b2 < 0.0f || (b1 + b2) > 1.0f) return false;
    const float tt = dot3(e2, s2) * inv;
    if (tt < r.tmin || tt > r.tmax) return false;
    *t = tt; *b1o = b1; *b2o = b2;
    return true;
}

// Sphere::get_intersection_point (sphere.rs:46-107) on the object-space ray
// (world2object.t(r), sphere.rs:131,137); utils::quadratic (utils/mod.rs:56-92).
__device__ bool sphere_test(const SphereRec& s, const RayState& r, float* t_out) {
    const float* m = s.w2o;
    float x = r.ox, y = r.oy, z = r.oz;
    float xt = m[0] * x + m[1] * y + m[2] * z + m[3];
    float yt = m[4] * x + m[5] * y + m[6] * z + m[7];
    float zt = m[8] * x + m[9] * y + m[10] * z + m[11];
    float w = m[12] * x + m[13] * y + m[14] * z + m[15];
    f3 o = (w != 1.0f) ? mk3(xt / w, yt / w, zt / w) : mk3(xt, yt, zt);
    f3 d = mk3(m[0] * r.dx + m[1] * r.dy + m[2] * r.dz, m[4] * r.dx + m[5] * r.dy + m[6] * r.dz,
               m[8] * r.dx + m[9] * r.dy + m[10] * r.dz);
    float a = lensq3(d);
    float b = 2.0f * dot3(d, o);
    float c = lensq3(o) - s.radius * s.radius;
    float disc = b * b - 4.0f * a * c;
    float t0, t1;
    if (disc < 0.0f) return false;
    if (fabsf(disc) < 1e-6f) {
        if (a == 0.0f) return false;
        t0 = t1 = -b / (2.0f * a);
    } else {
        float root = sqrtf(disc);
        float q = (b < 0.0f) ? -0.5f * (b - root) : -0.5f * (b + root);
        if (a == 0.0f || q == 0.0f) return false;
        float r0 = q / a, r1 = c / q;
        t0 = r0 < r1 ? r0 : r1;
        t1 = r0 < r1 ? r1 : r0;
    }
    if (t0 > r.tmax || t1 < r.tmin) return false;
    float th = t0;
    if (t0 < r.tmin) {
        th = t1;
        if (th > r.tmax) return false;
    }
    for (int pass = 0; pass < 2; ++pass) {
        f3 h = add3(o, mul3(d, th));
        if (h.x == 0.0f && h.y == 0.0f) h.x = 1e-5f * s.radius;
        float ang = atan2f(h.y, h.x);
        if (ang < 0.0f) ang += 2.0f * PI_F;
        bool bad = (h.z > -s.radius && h.z < s.zmin) || (h.z < s.radius && h.z > s.zmax) ||
                   (ang > s.phimax);
        if (!bad) { *t_out = th; return true; }
        if (pass == 1 || th == t1 || t1 > r.tmax) return false;
        th = t1;
    }
    return false;
}

// Test one primitive record; on a hit tmax becomes t (geometric.rs:64).
__device__ __forceinline__ bool prim_test(const SceneView& sv, uint32_t pos, RayState& r,
                                          uint32_t* prim, float* b1, float* b2, bool* last) {
    const float4* p = reinterpret_cast<const float4*>(sv.prims + pos);
    const float4 a = p[0], b = p[1], c = p[2];
    const uint32_t kind = as_u(c.w);
    *last = (kind & 0x100u) != 0u;
    float t, u, v;
    bool hit;
    if ((kind & 0xFFu) == 0u) {
        hit = tri_test(a, b, c, r, &t, &u, &v);
    } else {
        hit = sphere_test(sv.spheres[as_u(c.y)], r, &t);
        u = 0.0f; v = 0.0f;
    }
    if (hit) {
        r.tmax = t;
        *prim = as_u(c.z);
        *b1 = u; *b2 = v;
    }
    return hit;
}

__device__ __forceinline__ void load_ray(const Ray* rays, uint32_t i, RayState& r) {
    const float4* p = reinterpret_cast<const float4*>(rays + i);
    const float4 a = p[0], b = p[1];
    r.ox = a.x; r.oy = a.y; r.oz = a.z; r.tmin = a.w;
    r.dx = b.x; r.dy = b.y; r.dz = b.z; r.tmax = b.w;
    // bvh.rs:382-383 (and the per-axis 1/d of bbox.rs:191: the same value)
    r.ix = 1.0f / r.dx; r.iy = 1.0f / r.dy; r.iz = 1.0f / r.dz;
}
__device__ __forceinline__ void set_ray(const Ray& a, RayState& r) {
    r.ox = a.ox; r.oy = a.oy; r.oz = a.oz; r.tmin = a.tmin;
    r.dx = a.dx; r.dy = a.dy; r.dz = a.dz; r.tmax = a.tmax;
    r.ix = 1.0f / r.dx; r.iy = 1.0f / r.dy; r.iz = 1.0f / r.dz;
}

// ---------------------------------------------------------------------------
// PATH_PLAIN
// ---------------------------------------------------------------------------
template <bool ANY>
__global__ __launch_bounds__(MAX_BLOCK) void k_trace_plain(SceneView sv, Ray* rays, uint32_t n,
                                                           Hit* hits, uint32_t* occ, uint8_t* occ8,
                                                           uint32_t* spill, uint32_t spill_levels,
                                                           uint32_t* err) {
    __shared__ uint32_t stk[STK_PLAIN][MAX_BLOCK];
    const uint32_t tid = threadIdx.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + tid;
    const uint32_t total = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += total) {
        const uint32_t i = base + tid;
        bool hit = false;
        if (i < n) {
            RayState r;
            load_ray(rays, i, r);
            const bool neg0 = r.ix < 0.0f, neg1 = r.iy < 0.0f, neg2 = r.iz < 0.0f;
            uint32_t hprim = NO_HIT;
            float hb1 = 0.0f, hb2 = 0.0f;
            int sp = 0;
            auto push = [&](uint32_t v) {
                if (sp < STK_PLAIN) stk[sp][tid] = v;
                else if ((uint32_t)(sp - STK_PLAIN) < spill_levels)
                    spill[(size_t)(sp - STK_PLAIN) * total + gtid] = v;
                else { atomicOr(err, 1u); return; }
                ++sp;
            };
            auto pop = [&]() -> uint32_t {
                --sp;
                if (sp < STK_PLAIN) return stk[sp][tid];
                return spill[(size_t)(sp - STK_PLAIN) * total + gtid];
            };
            push(0u);
            while (sp > 0) {
                const uint32_t ni = pop();
                const float4* q = reinterpret_cast<const float4*>(sv.rnodes + ni);
                const float4 qa = q[0], qb = q[1];
                float tn;
                if (!slab(qa.x, qa.w, qa.y, qb.x, qa.z, qb.y, r.ox, r.oy, r.oz, r.ix, r.iy, r.iz,
                          r.tmin, r.tmax, &tn))
                    continue;
                const uint32_t off = as_u(qb.z), meta = as_u(qb.w);
                if ((meta & 3u) == 3u) {
                    const uint32_t cnt = meta >> 2;
                    for (uint32_t k = 0; k < cnt; ++k) {
                        bool last;
                        if (prim_test(sv, off + k, r, &hprim, &hb1, &hb2, &last)) {
                            hit = true;
                            if (ANY) break;
                        }
                    }
                    if (ANY && hit) break;
                } else if (meta == 0 ? neg0 : (meta == 1 ? neg1 : neg2)) {
                    push(ni + 1);
                    push(off);
                } else {
                    push(off);
                    push(ni + 1);
                }
            }
            if (!ANY) {
                hits[i] = Hit{hit ? hprim : NO_HIT, hit ? r.tmax : 0.0f, hb1, hb2};
                if (hit) rays[i].tmax = r.tmax;
            }
        }
        if (ANY && occ8) {
            if (i < n) occ8[i] = hit ? 1u : 0u;  // a byte per ray
        } else if (ANY) {
            // 64 consecutive rays per wave: one ballot, two words
            const unsigned long long m = __ballot(hit);
            const uint32_t lane = __lane_id();
            const uint32_t w0 = (base + (tid & ~63u)) >> 5;
            if (lane == 0 && (size_t)w0 * 32 < n) occ[w0] = (uint32_t)(m & 0xFFFFFFFFull);
            if (lane == 32 && (size_t)(w0 + 1) * 32 < n) occ[w0 + 1] = (uint32_t)(m >> 32);
        }
    }
}

// ---------------------------------------------------------------------------
// PATH_BATCHED
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t sel4i(int i, int32_t a, int32_t b, int32_t c, int32_t d) {
    return i == 0 ? a : (i == 1 ? b : (i == 2 ? c : d));
}
__device__ __forceinline__ float sel4f(int i, float a, float b, float c, float d) {
    return i == 0 ? a : (i == 1 ? b : (i == 2 ? c : d));
}
__device__ __forceinline__ bool sel4b(int i, bool a, bool b, bool c, bool d) {
    return i == 0 ? a : (i == 1 ? b : (i == 2 ? c : d));
}

// Segment of a served launch: launch rays [start, start+count) are the
// chunk's rays at `rays`; results go to hits / to the bits of occ (whose bit 0
// is the chunk's first ray: chunks start at multiples of 32 of a request).
// Segment starts are ascending.
__device__ __forceinline__ uint32_t find_seg(const TraceSeg* segs, uint32_t nseg, uint32_t id) {
    uint32_t lo = 0, hi = nseg;  // last segment with start <= id
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (segs[mid].start <= id) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Decoded bound of slot k (byte k of word w) of a compressed record:
// o + q * s rounded once (q * s is exact: an 8-bit integer times a power of
// two no larger than 2^119), the value the host checked against the exact box.
__device__ __forceinline__ float cq_bound(uint32_t w, int k, float o, float s) {
    const float q = (float)((w >> (8 * k)) & 0xFFu);
    return __builtin_fmaf(q, s, o);
}

// Slab test of an interior record of the compressed layout, with the near
// plane chosen by the sign of 1/d (nb_* near bounds, fb_* far bounds).  These
// boxes only need to be conservative -- the exact leaf box is tested when the
// leaf is popped -- and per axis this interval contains the reference's
// (bbox.rs:185-209): equal to it whenever no product is NaN, and wider in the
// one case where they differ (d = -0 with the ray origin on the box's low
// face, where the reference's swap test sees a NaN).  Its entry distance is
// therefore never above the exact one, so the pop-time re-check stays
// conservative as well.
__device__ __forceinline__ bool slab_signed(float nbx, float fbx, float nby, float fby, float nbz, float fbz,
                                            float ox, float oy, float oz, float ix, float iy, float iz,
                                            float tmin, float tmax, float* tn_out) {
    const float nx = (nbx - ox) * ix, fx = (fbx - ox) * ix;
    const float ny = (nby - oy) * iy, fy = (fby - oy) * iy;
    const float nz = (nbz - oz) * iz, fz = (fbz - oz) * iz;
    const float t0 = fmaxf(fmaxf(fmaxf(tmin, nx), ny), nz);
    const float t1 = fminf(fminf(fminf(tmax, fx), fy), fz);
    *tn_out = t0;
    return !(t0 > t1);
}
__device__ __forceinline__ float cq_scale(uint32_t ebyte) { return as_f(ebyte << 23); }

// The same record test in ray space: the distance to the plane of byte q on
// one axis is (o + q*s - ro) * id, evaluated as fma(q, A, B +- P) with
// A = s*id and B = (o - ro)*id computed once per record and axis.  Rounding
// makes this differ from both the decoded-box distance and the reference's
// distance to the exact plane, so the near distances are lowered and the far
// ones raised by a margin P that bounds every difference: with u = 2^-24,
//   P = 2^-20 * (|B| + |o|*|id| + 255*|A|) + 2^-100
// is at least 6u|B| (rounding of B, of B-P and of the fma, plus the
// reference's two roundings), u*|o|*|id| (the decoded plane o + q*s is
// rounded once: the exact sum may lie past the exact plane by half an ulp),
// 4u*255*|A| and the underflow of every step.  The interval is then never
// narrower than the reference's for the exact box, as the filter requires.
// Valid when nothing overflows and no NaN appears: the ray's 1/d and origin
// within 2^40 in magnitude, tmin and tmax not NaN, and the scene within 2^40
// (so every record origin and scale is); other rays use slab_signed.
constexpr float CQ_FAST_LIM = 0x1p40f;
__device__ __forceinline__ bool cq_fast_ray(const RayState& r) {
    const float m = fmaxf(fmaxf(fabsf(r.ix), fabsf(r.iy)), fmaxf(fabsf(r.iz), fabsf(r.ox)));
    return m <= CQ_FAST_LIM && fmaxf(fabsf(r.oy), fabsf(r.oz)) <= CQ_FAST_LIM && r.tmin == r.tmin &&
           r.tmax == r.tmax;
}
__device__ __forceinline__ bool cq_fast_scene(const SceneView& sv) {
    float m = 0.0f;
    for (int a = 0; a < 3; ++a) m = fmaxf(m, fmaxf(fabsf(sv.root_lo[a]), fabsf(sv.root_hi[a])));
    return m <= CQ_FAST_LIM;  // false for a NaN bound as well
}
struct CqAxis {
    float a, bn, bf;  // A, B - P, B + P
};
__device__ __forceinline__ CqAxis cq_axis_ray(float o, uint32_t ebyte, float ro, float id) {
    const float A = cq_scale(ebyte) * id;  // exact unless subnormal
    const float B = (o - ro) * id;
    const float m = __builtin_fmaf(255.0f, fabsf(A), __builtin_fmaf(fabsf(o), fabsf(id), fabsf(B)));
    const float P = __builtin_fmaf(m, 0x1p-20f, 0x1p-100f);
    return CqAxis{A, B - P, B + P};
}
__device__ __forceinline__ bool slab_ray(uint32_t nqx, uint32_t fqx, uint32_t nqy, uint32_t fqy, uint32_t nqz,
                                         uint32_t fqz, int k, const CqAxis& X, const CqAxis& Y, const CqAxis& Z,
                                         float tmin, float tmax, float* tn_out) {
    const float nx = __builtin_fmaf((float)((nqx >> (8 * k)) & 0xFFu), X.a, X.bn);
    const float fx = __builtin_fmaf((float)((fqx >> (8 * k)) & 0xFFu), X.a, X.bf);
    const float ny = __builtin_fmaf((float)((nqy >> (8 * k)) & 0xFFu), Y.a, Y.bn);
    const float fy = __builtin_fmaf((float)((fqy >> (8 * k)) & 0xFFu), Y.a, Y.bf);
    const float nz = __builtin_fmaf((float)((nqz >> (8 * k)) & 0xFFu), Z.a, Z.bn);
    const float fz = __builtin_fmaf((float)((fqz >> (8 * k)) & 0xFFu), Z.a, Z.bf);
    const float t0 = fmaxf(fmaxf(tmin, nx), fmaxf(ny, nz));
    const float t1 = fminf(fminf(tmax, fx), fminf(fy, fz));
    *tn_out = t0;
    return !(t0 > t1);
}

// Test one record of the compressed layout whose first three words are in
// a, b, c (the fourth is loaded here); on a hit tmax becomes t and *pos_hit
// the record's position.
__device__ __forceinline__ bool prim_test64(const SceneView& sv, uint32_t pos, const float4 a, const float4 b,
                                            const float4 c, uint32_t kind, RayState& r, uint32_t* pos_hit,
                                            float* b1, float* b2) {
    float t, u = 0.0f, v = 0.0f;
    bool hit;
    if ((kind & 0xFFu) == 0u) hit = tri_test(a, b, c, r, &t, &u, &v);
    else hit = sphere_test(sv.spheres[as_u(a.x)], r, &t);
    if (hit) {
        r.tmax = t;
        *pos_hit = pos;
        *b1 = u; *b2 = v;
    }
    return hit;
}

constexpr uint32_t LEAF_WAIT_PCT = 25;

// SEG: the rays come from a segment table (served launch) instead of one array.
// CQ: the compressed records (CNode, PrimRec64) instead of QNode / PrimRec.
// Camera-ray packets dealt to one XCD at a time (see k_closest_packet).
constexpr uint32_t CAM_RUN = 16;

// Work claims of the persistent per-ray kernel (see k_trace_wide): rays per
// claim and counters (parts of the ray range), WIDE_PART_STRIDE words apart.
// On the 1024x1024 direct-lighting frame's 26.5M shadow rays the occlusion
// kernel took 6.7 ms with one counter and claims of exactly what idle lanes
// needed, 5.2 ms with claims of 256, and 5.5 ms with claims of 256 from
// eight per-XCD parts (faster on the path tracer's shadow rays, 3.4 -> 3.2
// ms, slower on its extension rays): one part.
constexpr uint32_t WIDE_CHUNK = 256;
constexpr uint32_t WIDE_PARTS = 1;
constexpr uint32_t WIDE_PART_STRIDE = 32;
// where the part counters sit in a stream's scratch counter block (words;
// word 0 is the packet kernels' counter, word 1 the stack-overflow flag)
constexpr uint32_t WIDE_PARTS_AT = 32;
static_assert(WIDE_PARTS_AT + WIDE_PARTS * WIDE_PART_STRIDE <= TRACE_COUNTER_WORDS, "scratch counter block");

template <bool ANY, bool SEG, bool CQ, int STK_WIDE>
__global__ __launch_bounds__(MAX_BLOCK) __attribute__((amdgpu_waves_per_eu(8))) void k_trace_wide(SceneView sv, Ray* rays, uint32_t n_host,
                                                          const uint32_t* n_dev, Hit* hits,
                                                          uint32_t* occ, uint8_t* occ8, uint32_t* counter,
                                                          uint2* spill, uint32_t spill_levels,
                                                          uint32_t refill, uint32_t* err,
                                                          const TraceSeg* segs, uint32_t nseg) {
    const uint32_t n = n_dev ? *n_dev : n_host;
    // A launch sized for a queue's capacity: blocks beyond what the queue's
    // count can occupy (64 rays a wave) leave before touching the work
    // counter, one address that every wave of the chip would otherwise hit.
    if ((uint64_t)blockIdx.x * blockDim.x >= (uint64_t)n) return;
    __shared__ int32_t s_ref[STK_WIDE][MAX_BLOCK];
    __shared__ float s_tn[ANY ? 1 : STK_WIDE][MAX_BLOCK];
    const uint32_t tid = threadIdx.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + tid;
    const uint32_t total = gridDim.x * blockDim.x;
    // Work is claimed in chunks (WIDE_CHUNK rays with plenty of rays per
    // lane, else 64) from WIDE_PARTS counters, one per contiguous part of
    // the rays: a wave starts on its XCD's part (blocks are dealt to the 8
    // XCDs in turn), so neighbouring rays share an L2 and the claims spread
    // over eight addresses, and moves on to the next part when its part is
    // used up.  A part's counter only grows: a part found used up stays so.
    const uint32_t nparts = (gridDim.x % WIDE_PARTS == 0u) ? WIDE_PARTS : 1u;
    const uint32_t chunk = (uint64_t)n >= 8ull * total ? WIDE_CHUNK : 64u;
    uint32_t cur_part = blockIdx.x % nparts, tries = 0;
    uint32_t res_b = 0, res_e = 0;  // the wave's claimed ids not handed out yet
    auto part_start = [&](uint32_t q) { return (uint32_t)((uint64_t)n * q / nparts); };
    const uint32_t lane = __lane_id();
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    RayState r{};
    bool neg0 = false, neg1 = false, neg2 = false;
    uint32_t negm = 0;  // bit a: 1/d < 0 on axis a
    uint32_t rid = 0, hprim = NO_HIT;
    uint32_t rseg = 0;  // SEG: segment of the ray; rid is then its index in the segment
    float hb1 = 0.0f, hb2 = 0.0f;
    int32_t cur = 0;
    bool have_cur = false;  // false: take the next stack entry
    int sp = 0;
    bool has = false, drained = false;
    // CQ: this ray, and every busy ray of the wave, may take the ray-space
    // record test (cq_fast_ray); decided again after each refill
    const bool fast_scene = CQ && cq_fast_scene(sv);
    bool fastr = false, wfast = false;
    // An entry pushed at record time is re-checked against tmax when popped
    // (for occlusion tmax never shrinks before the exit, so that check and the
    // stored distance are not needed there).
    auto push = [&](int32_t ref, float tn) {
        if (sp < STK_WIDE) {
            s_ref[sp][tid] = ref;
            if (!ANY) s_tn[sp][tid] = tn;
            ++sp;
        } else if ((uint32_t)(sp - STK_WIDE) < spill_levels) {
            spill[(size_t)(sp - STK_WIDE) * total + gtid] = make_uint2((uint32_t)ref, as_u(tn));
            ++sp;
        } else {
            atomicOr(err, 1u);  // deeper than the scene's recorded depth: never expected
        }
    };

    while (true) {
        // ---- refill idle lanes with the next rays (one atomic per wave)
        const unsigned long long need = __ballot(!has && !drained);
        if (need) {
            // ids to the lanes in need until each has one or every part is
            // used up: the wave's chunk first, then fresh chunks
            unsigned long long pend = need;
            uint32_t id = 0;
            bool got = false;
            while (pend) {
                const uint32_t avail = res_e - res_b;
                if (avail == 0) {
                    if (tries >= nparts) break;
                    uint32_t b = 0;
                    if (lane == 0) b = atomicAdd(counter + WIDE_PART_STRIDE * cur_part, chunk);
                    b = __builtin_amdgcn_readfirstlane(__shfl(b, 0));
                    const uint32_t p0 = part_start(cur_part), plen = part_start(cur_part + 1) - p0;
                    if (b >= plen) {
                        cur_part = cur_part + 1 == nparts ? 0u : cur_part + 1;
                        ++tries;
                        continue;
                    }
                    res_b = p0 + b;
                    res_e = p0 + (plen - b < chunk ? plen : b + chunk);
                    continue;
                }
                const uint32_t pre = (uint32_t)__popcll(pend & lt_mask);
                const bool mine = ((pend >> lane) & 1ull) && pre < avail;
                if (mine) {
                    id = res_b + pre;
                    got = true;
                }
                const unsigned long long taken = __ballot(mine);
                res_b += (uint32_t)__popcll(taken);
                pend &= ~taken;
            }
            if (!has && !drained) {
                if (got) {
                    if (SEG) {
                        rseg = find_seg(segs, nseg, id);
                        rid = id - segs[rseg].start;
                        load_ray(segs[rseg].rays, rid, r);
                    } else {
                        rid = id;
                        load_ray(rays, id, r);
                    }
                    neg0 = r.ix < 0.0f; neg1 = r.iy < 0.0f; neg2 = r.iz < 0.0f;
                    if (CQ) fastr = cq_fast_ray(r);
                    negm = (neg0 ? 1u : 0u) | (neg1 ? 2u : 0u) | (neg2 ? 4u : 0u);
                    hprim = NO_HIT; hb1 = 0.0f; hb2 = 0.0f;
                    sp = 0;
                    float tn;
                    // the root's own box (popped first by the reference)
                    have_cur = slab(sv.root_lo[0], sv.root_hi[0], sv.root_lo[1], sv.root_hi[1],
                                    sv.root_lo[2], sv.root_hi[2], r.ox, r.oy, r.oz, r.ix, r.iy, r.iz,
                                    r.tmin, r.tmax, &tn);
                    cur = (ANY && CQ) ? sv.ocroot : sv.qroot;
                    has = true;
                } else {
                    drained = true;
                }
            }
        }
        if (!__any(has)) break;
        if (CQ) wfast = fast_scene && __ballot(has && !fastr) == 0ull;
        // ---- traverse until too few lanes are busy
        while (has) {
            if (!have_cur) {
                while (sp > 0) {
                    --sp;
                    int32_t ref;
                    float tn = 0.0f;
                    if (sp < STK_WIDE) {
                        ref = s_ref[sp][tid];
                        if (!ANY) tn = s_tn[sp][tid];
                    } else {
                        const uint2 e = spill[(size_t)(sp - STK_WIDE) * total + gtid];
                        ref = (int32_t)e.x; tn = as_f(e.y);
                    }
                    if (ANY || tn <= r.tmax) { cur = ref; have_cur = true; break; }
                }
                if (!have_cur) {
                    // ray finished
                    if (ANY && occ8) occ8[rid] = 0u;  // no hit
                    if (!ANY) {
                        Hit* hp = SEG ? segs[rseg].hits : hits;
                        Ray* rp = SEG ? segs[rseg].rays : rays;
                        if (CQ && hprim != NO_HIT) hprim = sv.order[hprim];  // position -> prim id
                        hp[rid] = Hit{hprim, hprim != NO_HIT ? r.tmax : 0.0f, hb1, hb2};
                        if (hprim != NO_HIT) rp[rid].tmax = r.tmax;
                    }
                    has = false;
                }
            }
            // A lane at a leaf waits while fewer than LEAF_WAIT_PCT % of the
            // wave's tracing lanes are at one (and some are at records): every
            // step that some lane spends on a leaf costs the whole wave the
            // leaf code, so leaves are taken in bunches.  On the 1024x1024
            // direct-lighting frame, occlusion 3.87 -> 3.45 ms at 25 % (15 %:
            // 3.52, 35 %: 3.52, 50 %: 3.62); the path-tracing frame gains too.
            // Tried and slower there: no waiting, with each leaf reached put
            // in a per-wave queue in LDS and the whole wave taking 64 queued
            // leaves at a time, one per lane, on rays read across lanes
            // (exact, but 3.19 -> 3.58-3.78 ms: a round lasts as long as its
            // largest leaf, and lanes whose rays wait on queued leaves idle).
            const unsigned long long at_rec = __ballot(has && cur >= 0);
            const unsigned long long at_leaf = __ballot(has && cur < 0);
            const bool leaf_now = at_rec == 0ull || (uint32_t)__popcll(at_leaf) * 100u >=
                                                        LEAF_WAIT_PCT * (uint32_t)__popcll(at_leaf | at_rec);
            if (has && cur >= 0) {
                int4 RF;
                uint32_t meta;
                float t0, t1, t2, t3;
                bool p0, p1, p2, p3;
                if constexpr (CQ) {
                    // occlusion reads the order-free records
                    const float4* q = reinterpret_cast<const float4*>((ANY ? sv.ocnodes : sv.cnodes) + cur);
                    const float4 W0 = q[0], W1 = q[1], W2 = q[2], W3 = q[3];
                    const uint32_t em = as_u(W0.w);
                    // near / far bytes of all four slots by the ray's signs
                    const uint32_t nqx = neg0 ? as_u(W1.y) : as_u(W1.x), fqx = neg0 ? as_u(W1.x) : as_u(W1.y);
                    const uint32_t nqy = neg1 ? as_u(W1.w) : as_u(W1.z), fqy = neg1 ? as_u(W1.z) : as_u(W1.w);
                    const uint32_t nqz = neg2 ? as_u(W2.y) : as_u(W2.x), fqz = neg2 ? as_u(W2.x) : as_u(W2.y);
                    RF = make_int4((int)as_u(W2.z), (int)as_u(W2.w), (int)as_u(W3.x), (int)as_u(W3.y));
                    meta = em >> 24;
                    if (wfast) {  // every busy ray of the wave: cq_fast_ray
                        const CqAxis X = cq_axis_ray(W0.x, em & 0xFFu, r.ox, r.ix);
                        const CqAxis Y = cq_axis_ray(W0.y, (em >> 8) & 0xFFu, r.oy, r.iy);
                        const CqAxis Z = cq_axis_ray(W0.z, (em >> 16) & 0xFFu, r.oz, r.iz);
// every slot tested and the empty ones masked after, with no branch per
// slot (occlusion on the 1024x1024 direct-lighting frame: 3.27-3.30 -> 3.19-3.21
// ms).  Tried and slower there: the margin P bounded once per ray (from the
// scene box: 3.53 ms) or from a per-record size term kept in the record
// (3.67 ms); their extra live values at 64 VGPRs cost more than the five
// instructions per axis they save.  The same form in the camera packet
// kernel was slower (1.355 -> 1.39 ms on that frame's camera rays).
#define PX_CQ_SLOT(K, I, P, T) \
    P = (RF.K != QEMPTY) & slab_ray(nqx, fqx, nqy, fqy, nqz, fqz, I, X, Y, Z, r.tmin, r.tmax, &T);
                        PX_CQ_SLOT(x, 0, p0, t0)
                        PX_CQ_SLOT(y, 1, p1, t1)
                        PX_CQ_SLOT(z, 2, p2, t2)
                        PX_CQ_SLOT(w, 3, p3, t3)
#undef PX_CQ_SLOT
                    } else {
                        const float sx = cq_scale(em & 0xFFu), sy = cq_scale((em >> 8) & 0xFFu),
                                    sz = cq_scale((em >> 16) & 0xFFu);
#define PX_CQ_SLOT(K, I, P, T)                                                                                \
    P = RF.K != QEMPTY &&                                                                                     \
        slab_signed(cq_bound(nqx, I, W0.x, sx), cq_bound(fqx, I, W0.x, sx), cq_bound(nqy, I, W0.y, sy),       \
                    cq_bound(fqy, I, W0.y, sy), cq_bound(nqz, I, W0.z, sz), cq_bound(fqz, I, W0.z, sz), r.ox,  \
                    r.oy, r.oz, r.ix, r.iy, r.iz, r.tmin, r.tmax, &T);
                        PX_CQ_SLOT(x, 0, p0, t0)
                        PX_CQ_SLOT(y, 1, p1, t1)
                        PX_CQ_SLOT(z, 2, p2, t2)
                        PX_CQ_SLOT(w, 3, p3, t3)
#undef PX_CQ_SLOT
                    }
                } else {
                    const float4* q = reinterpret_cast<const float4*>(sv.qnodes + cur);
                    const float4 LX = q[0], HX = q[1], LY = q[2], HY = q[3], LZ = q[4], HZ = q[5];
                    RF = reinterpret_cast<const int4*>(q)[6];
                    meta = reinterpret_cast<const uint32_t*>(q)[28];
                    p0 = RF.x != QEMPTY && slab(LX.x, HX.x, LY.x, HY.x, LZ.x, HZ.x, r.ox, r.oy, r.oz,
                                                r.ix, r.iy, r.iz, r.tmin, r.tmax, &t0);
                    p1 = RF.y != QEMPTY && slab(LX.y, HX.y, LY.y, HY.y, LZ.y, HZ.y, r.ox, r.oy, r.oz,
                                                r.ix, r.iy, r.iz, r.tmin, r.tmax, &t1);
                    p2 = RF.z != QEMPTY && slab(LX.z, HX.z, LY.z, HY.z, LZ.z, HZ.z, r.ox, r.oy, r.oz,
                                                r.ix, r.iy, r.iz, r.tmin, r.tmax, &t2);
                    p3 = RF.w != QEMPTY && slab(LX.w, HX.w, LY.w, HY.w, LZ.w, HZ.w, r.ox, r.oy, r.oz,
                                                r.ix, r.iy, r.iz, r.tmin, r.tmax, &t3);
                }
                if (ANY && sp + 4 <= STK_WIDE) {
                    // Occlusion: the answer does not depend on the order in
                    // which a ray meets its leaves, so the passing slots go
                    // on the stack in slot order, each write unconditional
                    // and the top advanced past the passing ones; the lowest
                    // passing slot, written last, is taken at once.  (On the
                    // 1024x1024 direct-lighting frame's shadow rays: 4.16 ms
                    // in the reference's order with branches, 3.81 ms so.)
                    int k = sp;
                    s_ref[k][tid] = RF.w; k += p3 ? 1 : 0;
                    s_ref[k][tid] = RF.z; k += p2 ? 1 : 0;
                    s_ref[k][tid] = RF.y; k += p1 ? 1 : 0;
                    s_ref[k][tid] = RF.x; k += p0 ? 1 : 0;
                    have_cur = k > sp;
                    cur = p0 ? RF.x : (p1 ? RF.y : (p2 ? RF.z : RF.w));
                    sp = have_cur ? k - 1 : sp;
                } else {
                    // visiting order of the reference (bvh.rs:409-415 applied at A,
                    // then at B and C): near half first, near child first in each.
                    // The slots [B0 B1 C0 C1] are permuted in registers into that
                    // order by three conditional swaps (selects, no branches);
                    // occlusion near a full stack keeps slot order.
                    int32_t r0 = RF.x, r1 = RF.y, r2 = RF.z, r3 = RF.w;
                    if (!ANY) {
                        const bool nA = (negm >> (meta & 3u)) & 1u;
                        const bool nB = (negm >> ((meta >> 2) & 3u)) & 1u;
                        const bool nC = (negm >> ((meta >> 4) & 3u)) & 1u;
#define PX_SWAP(C, I, J)                                                                 \
    {                                                                                     \
        const bool pa = C ? p##J : p##I, pb = C ? p##I : p##J;                            \
        const int32_t ra = C ? r##J : r##I, rb = C ? r##I : r##J;                         \
        const float ta = C ? t##J : t##I, tb = C ? t##I : t##J;                           \
        p##I = pa; p##J = pb; r##I = ra; r##J = rb; t##I = ta; t##J = tb;                 \
    }
                        PX_SWAP(nB, 0, 1)
                        PX_SWAP(nC, 2, 3)
                        PX_SWAP(nA, 0, 2)
                        PX_SWAP(nA, 1, 3)
#undef PX_SWAP
                    }
                    // the first passing slot is taken now, the others pushed so
                    // that they pop in order
                    bool pend = false;
                    int32_t pref = 0;
                    float ptn = 0.0f;
#define PX_VISIT(K)                                                                      \
    if (p##K) {                                                                           \
        if (pend) push(pref, ptn);                                                        \
        pend = true;                                                                      \
        pref = r##K;                                                                      \
        ptn = t##K;                                                                       \
    }
                    PX_VISIT(3)
                    PX_VISIT(2)
                    PX_VISIT(1)
                    PX_VISIT(0)
#undef PX_VISIT
                    have_cur = pend;
                    cur = pref;
                }
            } else if (has && leaf_now) {
                // leaf: prims in order until the record flagged last
                uint32_t pos = (uint32_t)(~cur);
                bool hitnow = false;
                if constexpr (CQ) {
                    // the leaf's exact box at the current tmax: the reference's
                    // pop-time test (the record boxes above it are wider)
                    const float4* p = reinterpret_cast<const float4*>(sv.prims64 + pos);
                    float4 P2 = p[2], P3 = p[3];
                    float tn;
                    if (slab(P2.y, P3.x, P2.z, P3.y, P2.w, P3.z, r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, r.tmin, r.tmax,
                             &tn)) {
                        while (true) {
                            const float4 P0 = p[0], P1 = p[1];
                            const uint32_t kind = as_u(P3.w);
                            if (prim_test64(sv, pos, P0, P1, P2, kind, r, &hprim, &hb1, &hb2)) {
                                hitnow = true;
                                if (ANY) break;
                            }
                            if (kind & 0x100u) break;
                            ++pos;
                            p += 4;
                            P2 = p[2];
                            P3 = p[3];
                        }
                    }
                } else {
                    while (true) {
                        bool last;
                        if (prim_test(sv, pos, r, &hprim, &hb1, &hb2, &last)) {
                            hitnow = true;
                            if (ANY) break;
                        }
                        if (last) break;
                        ++pos;
                    }
                }
                have_cur = false;
                if (ANY && hitnow) {
                    if (occ8) {
                        occ8[rid] = 1u;  // a byte per ray: one plain store, no atomic
                    } else {
                        uint32_t* ow = SEG ? segs[rseg].occ : occ;
                        atomicOr(ow + (rid >> 5), 1u << (rid & 31u));
                    }
                    has = false;
                }
            }
            // leave to refill when the wave has gone sparse
            const unsigned long long busy = __ballot(has);
            if ((uint32_t)__popcll(busy) * 100u < refill * 64u) break;
        }
    }
}

// Occlusion with G lanes per ray, for launches with too few rays to fill the
// chip: there the per-ray kernel lasts as long as its slowest ray's chain of
// dependent record loads, at about one wave per SIMD.  A group of G lanes of
// one wave takes one ray and one stack in LDS.  Each step up to G entries
// come off the top of the stack, one per lane; each lane tests its record's
// slots or its leaf against the group's ray, and the passing slots go back
// on the stack.  The answer does not depend on the order in which a ray
// meets its leaves (its tmax does not change before the exit), so this
// tests the same boxes and prims as the per-ray walk, several at a time, and
// the group stops at the first hit any lane finds.
// Stack bound: a step takes k >= 2 entries only while top + 3k <= lim, with
// lim = GSTK - 3 * odepth, and one entry otherwise.  A run of one-entry steps
// is a depth-first walk; from a top of at most lim it stays within lim + 3 *
// odepth = GSTK (three pending slots per record level below the entry it
// started from, the fourth child taken), and every run starts from a top of
// at most lim (the first step, or one after a step of k >= 2).
// SEG: the rays come from a segment table (served launch) instead of one array.
constexpr int GSTK = 128;
template <int G, bool SEG>
__global__ __launch_bounds__(MAX_BLOCK) __attribute__((amdgpu_waves_per_eu(8))) void k_occlude_group(
    SceneView sv, const Ray* rays, uint32_t n_host, const uint32_t* n_dev, uint32_t* occ, uint8_t* occ8,
    uint32_t lim, uint32_t* err, const TraceSeg* segs, uint32_t nseg) {
    static_assert(G >= 2 && G <= 64 && (64 % G) == 0, "a group lies within one wave");
    // (the grid covers n_host rays: a count on the device never exceeds it)
    const uint32_t n = n_dev ? min(*n_dev, n_host) : n_host;
    const uint32_t gpb = blockDim.x / G;  // groups per block
    if ((uint64_t)blockIdx.x * gpb >= (uint64_t)n) return;
    __shared__ int32_t s_stk[MAX_BLOCK / G][GSTK];
    const uint32_t tid = threadIdx.x, lane = __lane_id();
    const uint32_t g = tid / G, gl = tid % G;
    const uint32_t id = blockIdx.x * gpb + g;
    uint32_t rid = id, rseg = 0;  // SEG: rid is the ray's index in segment rseg
    const uint32_t gbase = lane & ~(uint32_t)(G - 1);
    const unsigned long long gm = (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << gbase;  // the group's lanes
    const unsigned long long lowg = gm & (lane == 0 ? 0ull : (~0ull >> (64 - lane)));  // ... below this one
    int32_t* stk = s_stk[g];
    RayState r{};
    bool done = id >= n;
    uint32_t top = 0;
    if (!done) {
        if (SEG) {
            rseg = find_seg(segs, nseg, id);
            rid = id - segs[rseg].start;
            load_ray(segs[rseg].rays, rid, r);
        } else {
            load_ray(rays, rid, r);
        }
        float tn;
        // the root's own box (popped first by the reference)
        if (slab(sv.root_lo[0], sv.root_hi[0], sv.root_lo[1], sv.root_hi[1], sv.root_lo[2], sv.root_hi[2], r.ox,
                 r.oy, r.oz, r.ix, r.iy, r.iz, r.tmin, r.tmax, &tn)) {
            top = 1;
            if (gl == 0) stk[0] = sv.ocroot;
        }
        done = top == 0;
    }
    const bool neg0 = r.ix < 0.0f, neg1 = r.iy < 0.0f, neg2 = r.iz < 0.0f;
    // every ray of the wave may take the ray-space record test (cq_fast_ray)
    const bool wfast = cq_fast_scene(sv) && __ballot(id < n && !cq_fast_ray(r)) == 0ull;
    bool hit = false;
    while (__any(!done)) {
        // entries for this step: up to G, fewer near the stack's bound
        uint32_t k = 0;
        if (!done) {
            k = top < (uint32_t)G ? top : (uint32_t)G;
            const uint32_t room = top + 6u <= lim ? (lim - top) / 3u : 1u;
            k = k < room ? k : room;
        }
        const bool mine = gl < k;
        const int32_t ref = mine ? stk[top - 1u - gl] : 0;
        top -= k;
        bool p0 = false, p1 = false, p2 = false, p3 = false, hitnow = false;
        int4 RF = make_int4(0, 0, 0, 0);
        if (mine && ref >= 0) {
            const float4* q = reinterpret_cast<const float4*>(sv.ocnodes + ref);
            const float4 W0 = q[0], W1 = q[1], W2 = q[2], W3 = q[3];
            const uint32_t em = as_u(W0.w);
            const uint32_t nqx = neg0 ? as_u(W1.y) : as_u(W1.x), fqx = neg0 ? as_u(W1.x) : as_u(W1.y);
            const uint32_t nqy = neg1 ? as_u(W1.w) : as_u(W1.z), fqy = neg1 ? as_u(W1.z) : as_u(W1.w);
            const uint32_t nqz = neg2 ? as_u(W2.y) : as_u(W2.x), fqz = neg2 ? as_u(W2.x) : as_u(W2.y);
            RF = make_int4((int)as_u(W2.z), (int)as_u(W2.w), (int)as_u(W3.x), (int)as_u(W3.y));
            float t0, t1, t2, t3;
            if (wfast) {
                const CqAxis X = cq_axis_ray(W0.x, em & 0xFFu, r.ox, r.ix);
                const CqAxis Y = cq_axis_ray(W0.y, (em >> 8) & 0xFFu, r.oy, r.iy);
                const CqAxis Z = cq_axis_ray(W0.z, (em >> 16) & 0xFFu, r.oz, r.iz);
#define PX_CQ_SLOT(K, I, P, T) \
    P = RF.K != QEMPTY && slab_ray(nqx, fqx, nqy, fqy, nqz, fqz, I, X, Y, Z, r.tmin, r.tmax, &T);
                PX_CQ_SLOT(x, 0, p0, t0)
                PX_CQ_SLOT(y, 1, p1, t1)
                PX_CQ_SLOT(z, 2, p2, t2)
                PX_CQ_SLOT(w, 3, p3, t3)
#undef PX_CQ_SLOT
            } else {
                const float sx = cq_scale(em & 0xFFu), sy = cq_scale((em >> 8) & 0xFFu),
                            sz = cq_scale((em >> 16) & 0xFFu);
#define PX_CQ_SLOT(K, I, P, T)                                                                                \
    P = RF.K != QEMPTY &&                                                                                     \
        slab_signed(cq_bound(nqx, I, W0.x, sx), cq_bound(fqx, I, W0.x, sx), cq_bound(nqy, I, W0.y, sy),       \
                    cq_bound(fqy, I, W0.y, sy), cq_bound(nqz, I, W0.z, sz), cq_bound(fqz, I, W0.z, sz), r.ox,  \
                    r.oy, r.oz, r.ix, r.iy, r.iz, r.tmin, r.tmax, &T);
                PX_CQ_SLOT(x, 0, p0, t0)
                PX_CQ_SLOT(y, 1, p1, t1)
                PX_CQ_SLOT(z, 2, p2, t2)
                PX_CQ_SLOT(w, 3, p3, t3)
#undef PX_CQ_SLOT
            }
        } else if (mine) {
            // leaf: its exact box, then prims in order until the record flagged last
            uint32_t pos = (uint32_t)(~ref);
            const float4* p = reinterpret_cast<const float4*>(sv.prims64 + pos);
            float4 P2 = p[2], P3 = p[3];
            float tn;
            if (slab(P2.y, P3.x, P2.z, P3.y, P2.w, P3.z, r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, r.tmin, r.tmax, &tn)) {
                while (true) {
                    const float4 P0 = p[0], P1 = p[1];
                    const uint32_t kind = as_u(P3.w);
                    uint32_t hp;
                    float hb1, hb2;
                    if (prim_test64(sv, pos, P0, P1, P2, kind, r, &hp, &hb1, &hb2)) {
                        hitnow = true;
                        break;
                    }
                    if (kind & 0x100u) break;
                    ++pos;
                    p += 4;
                    P2 = p[2];
                    P3 = p[3];
                }
            }
        }
        // passing slots onto the group's stack: each lane's offset is the
        // count of its group's lower lanes' passing slots (three ballots of
        // the bits of each lane's count)
        const uint32_t c = (p0 ? 1u : 0u) + (p1 ? 1u : 0u) + (p2 ? 1u : 0u) + (p3 ? 1u : 0u);
        const unsigned long long c0 = __ballot(c & 1u), c1 = __ballot(c & 2u), c2 = __ballot(c & 4u);
        const uint32_t pre = (uint32_t)__popcll(c0 & lowg) + 2u * (uint32_t)__popcll(c1 & lowg) +
                             4u * (uint32_t)__popcll(c2 & lowg);
        const uint32_t tot = (uint32_t)__popcll(c0 & gm) + 2u * (uint32_t)__popcll(c1 & gm) +
                             4u * (uint32_t)__popcll(c2 & gm);
        const bool ghit = (__ballot(hitnow) & gm) != 0ull;
        if (!done && !ghit) {
            if (top + tot > (uint32_t)GSTK) {
                if (gl == 0) atomicOr(err, 1u);  // outside the bound above: never expected
                done = true;
            } else {
                uint32_t w = top + pre;
                if (p0) stk[w++] = RF.x;
                if (p1) stk[w++] = RF.y;
                if (p2) stk[w++] = RF.z;
                if (p3) stk[w] = RF.w;
                top += tot;
            }
        }
        hit = hit || ghit;
        done = done || ghit || top == 0u;
    }
    if (id < n && gl == 0) {
        if (occ8) {
            occ8[rid] = hit ? 1u : 0u;
        } else if (hit) {
            uint32_t* ow = SEG ? segs[rseg].occ : occ;
            atomicOr(ow + (rid >> 5), 1u << (rid & 31u));
        }
    }
}

constexpr int PSTK = 96;  // stack entries per wave of the packet closest-hit kernel below
constexpr int PWAVES = MAX_BLOCK / 64;

// Read-only scene data seen through the constant address space: a load from
// a wave-uniform address then goes through the scalar cache (s_load) rather
// than 64 vector lanes.
typedef __attribute__((address_space(4))) const float cfloat;
__device__ __forceinline__ float4 cload4(const void* p, uint32_t k) {
    cfloat* q = (cfloat*)p + 4 * k;
    return make_float4(q[0], q[1], q[2], q[3]);
}

// Batched closest hit, wave-coherent ("packet") form, for coherent rays such
// as a camera's.  A wave takes 64 consecutive rays and splits them into
// groups of equal direction signs; the reference's visiting order
// (bvh.rs:409-415) depends on the ray only through those signs, so one group
// walks the records in exactly each member's order with ONE stack of (record,
// lane mask) entries, every record and prim record read once per wave from a
// wave-uniform address.  Each lane tests boxes and prims against its own
// tmax, so a lane reaches and tests exactly what the reference's traversal of
// its ray would, in the same order (a later equal-t hit replaces an earlier
// one, mesh.rs:71).  The pop-time test of an interior record is left out: a
// record popped after the lane's tmax has dropped below its entry distance
// only tests its children again, and they fail (a child box lies inside its
// parent's, and the rounded slab distances are monotone in the bounds), so
// that visit changes nothing.  A leaf's pop-time test is kept: each lane
// tests the leaf's exact box (carried by the 64-byte prim records) against
// its current tmax before the leaf's prims.  Interior records are the exact
// 128-byte ones.  GEN: ray i is the camera ray of sample cam.g0 + i, made
// here instead of read (no ray array; only the hits are written).
// co.list (GEN, a frame's bounce 0 of path tracing): a camera ray that
// misses ends its path here -- zero radiance written to co.miss[i], no hit
// record -- and each packet's hits take a run of co.list (their sample
// indices) with one atomic on co.count.
template <bool GEN, bool SEGS = false>
__global__ __launch_bounds__(MAX_BLOCK) __attribute__((amdgpu_waves_per_eu(8))) void k_closest_packet(SceneView sv, Ray* rays, uint32_t n_host,
                                                              const uint32_t* n_dev, Hit* hits, uint32_t* counter,
                                                              uint32_t* err, CamParams cam, CamTable ct, CamOut co) {
    const uint32_t n = n_dev ? *n_dev : n_host;
    __shared__ int32_t s_ref[PWAVES][PSTK];
    __shared__ unsigned long long s_msk[PWAVES][PSTK];
    const bool fast_scene = cq_fast_scene(sv);  // planes within 2^40: slab_ord may apply
    const uint32_t lane = __lane_id();
    const uint32_t w = threadIdx.x >> 6;
    // GEN (camera rays, their count known at launch): a fixed schedule with
    // no shared counter.  Packets go in runs of CAM_RUN to the 8 XCDs in turn
    // (blocks are dealt to the XCDs in turn), and an XCD's waves take its
    // packets in turn: neighbouring pixels share an L2, and every XCD gets
    // a share of every part of the image.  Otherwise waves claim packets
    // from the launch's counter.
    const uint32_t wpb = blockDim.x >> 6;
    const uint32_t nxcd = (gridDim.x % 8u == 0u) ? 8u : 1u;
    const uint64_t npk = ((uint64_t)n + 63u) / 64u;
    const uint32_t xcd = blockIdx.x % nxcd;
    const uint32_t wstride = (gridDim.x / nxcd) * wpb;
    uint64_t k = (uint64_t)(blockIdx.x / nxcd) * wpb + w;  // this wave's next packet among its XCD's
    while (true) {
        uint32_t base = 0;
        if constexpr (GEN) {
            const uint64_t pk = (k / CAM_RUN * nxcd + xcd) * CAM_RUN + k % CAM_RUN;
            if (pk >= npk) break;
            base = (uint32_t)(pk * 64u);
            k += wstride;
        } else {
            if (lane == 0) base = atomicAdd(counter, 64u);
            base = __builtin_amdgcn_readfirstlane(__shfl(base, 0));
            if (base >= n) break;
        }
        uint32_t i = base + lane;
        if (GEN && cam.tile_lw != 0xFFu) {
            // the packet is a tile of 64 / spp pixels, tile_w wide, all their
            // samples (whole rows of tiles: the launch covers whole rows)
            const uint32_t tw = 1u << cam.tile_lw;
            const uint32_t th_l = 6u - cam.spp_l - cam.tile_lw;
            const uint32_t ntx = cam.xres >> cam.tile_lw;
            const uint32_t tile = base >> 6;
            const uint32_t tx = tile % ntx, ty = tile / ntx;
            const uint32_t q = lane >> cam.spp_l, sidx = lane & ((1u << cam.spp_l) - 1u);
            const uint32_t x = (tx << cam.tile_lw) + (q & (tw - 1u)), y = (ty << th_l) + (q >> cam.tile_lw);
            i = ((y * cam.xres + x) << cam.spp_l) + sidx;
        }
        const bool valid = i < n;
        RayState r{};
        bool alive = false;
        uint32_t negm = 0;
        if (GEN && SEGS) {
            // several cameras: each distinct segment of the packet in turn,
            // its camera read with scalar loads
            const uint32_t sidx = valid ? cam_seg_of(ct.segs, ct.nsegs, i) : 0u;
            unsigned long long pend = __ballot(valid);
            while (pend) {
                const uint32_t s0 = __builtin_amdgcn_readfirstlane(__shfl(sidx, (int)__builtin_ctzll(pend)));
                const unsigned long long mine = __ballot(valid && sidx == s0);
                pend &= ~mine;
                const CamSeg sg = ct.segs[s0];
                const CamParams* cp = reinterpret_cast<const CamParams*>(ct.cams + (size_t)sg.cam * ct.stride);
                if ((mine >> lane) & 1ull) {
                    Ray g;
                    camera_ray_of(*cp, sg.g0 + (i - sg.start), &g);
                    set_ray(g, r);
                }
            }
        }
        if (valid) {
            if constexpr (GEN) {
                if (!SEGS) {
                    Ray g;
                    camera_ray_of(cam, cam.g0 + i, &g);
                    set_ray(g, r);
                }
            } else {
                load_ray(rays, i, r);
            }
            negm = (r.ix < 0.0f ? 1u : 0u) | (r.iy < 0.0f ? 2u : 0u) | (r.iz < 0.0f ? 4u : 0u);
            float tn;
            alive = slab(sv.root_lo[0], sv.root_hi[0], sv.root_lo[1], sv.root_hi[1], sv.root_lo[2],
                         sv.root_hi[2], r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, r.tmin, r.tmax, &tn);
        }
        uint32_t hprim = NO_HIT;
        float hb1 = 0.0f, hb2 = 0.0f;
        const unsigned long long alive_m = __ballot(alive);
        // lanes that may not take the sign-ordered box test (slab_ord)
        const unsigned long long slow_m = __ballot(alive && !cq_fast_ray(r));
        unsigned long long todo = __ballot(valid);
        while (todo) {
            // the next group: lanes with the signs of the first remaining lane
            const uint32_t lead = (uint32_t)__builtin_ctzll(todo);
            const uint32_t g = __builtin_amdgcn_readfirstlane(__shfl(negm, (int)lead));
            const unsigned long long grp = __ballot(valid && negm == g) & todo;
            todo &= ~grp;
            const unsigned long long live = grp & alive_m;
            if (!live) continue;
            const bool gfast = fast_scene && (live & slow_m) == 0ull;
            int sp = 0;
            int32_t cur = sv.qroot;
            unsigned long long cm = live;  // lanes that reached the current entry
            while (true) {
                const bool act = (cm >> lane) & 1ull;
                if (cur >= 0) {
                    const QNode* q = sv.qnodes + cur;
                    const float4 LX = cload4(q, 0), HX = cload4(q, 1), LY = cload4(q, 2), HY = cload4(q, 3),
                                 LZ = cload4(q, 4), HZ = cload4(q, 5), R6 = cload4(q, 6);
                    const uint32_t meta = as_u(cload4(q, 7).x);
                    int32_t r0 = (int32_t)as_u(R6.x), r1 = (int32_t)as_u(R6.y), r2 = (int32_t)as_u(R6.z),
                            r3 = (int32_t)as_u(R6.w);
                    bool p0, p1, p2, p3;
                    if (gfast) {
                        // planes ordered by the group's signs (wave-uniform selects)
                        const float4 NX = (g & 1u) ? HX : LX, FX = (g & 1u) ? LX : HX;
                        const float4 NY = (g & 2u) ? HY : LY, FY = (g & 2u) ? LY : HY;
                        const float4 NZ = (g & 4u) ? HZ : LZ, FZ = (g & 4u) ? LZ : HZ;
                        p0 = act && r0 != QEMPTY && slab_ord(NX.x, FX.x, NY.x, FY.x, NZ.x, FZ.x, r.ox, r.oy, r.oz,
                                                             r.ix, r.iy, r.iz, r.tmin, r.tmax);
                        p1 = act && r1 != QEMPTY && slab_ord(NX.y, FX.y, NY.y, FY.y, NZ.y, FZ.y, r.ox, r.oy, r.oz,
                                                             r.ix, r.iy, r.iz, r.tmin, r.tmax);
                        p2 = act && r2 != QEMPTY && slab_ord(NX.z, FX.z, NY.z, FY.z, NZ.z, FZ.z, r.ox, r.oy, r.oz,
                                                             r.ix, r.iy, r.iz, r.tmin, r.tmax);
                        p3 = act && r3 != QEMPTY && slab_ord(NX.w, FX.w, NY.w, FY.w, NZ.w, FZ.w, r.ox, r.oy, r.oz,
                                                             r.ix, r.iy, r.iz, r.tmin, r.tmax);
                    } else {
                        float t;
                        p0 = act && r0 != QEMPTY && slab(LX.x, HX.x, LY.x, HY.x, LZ.x, HZ.x, r.ox, r.oy, r.oz,
                                                         r.ix, r.iy, r.iz, r.tmin, r.tmax, &t);
                        p1 = act && r1 != QEMPTY && slab(LX.y, HX.y, LY.y, HY.y, LZ.y, HZ.y, r.ox, r.oy, r.oz,
                                                         r.ix, r.iy, r.iz, r.tmin, r.tmax, &t);
                        p2 = act && r2 != QEMPTY && slab(LX.z, HX.z, LY.z, HY.z, LZ.z, HZ.z, r.ox, r.oy, r.oz,
                                                         r.ix, r.iy, r.iz, r.tmin, r.tmax, &t);
                        p3 = act && r3 != QEMPTY && slab(LX.w, HX.w, LY.w, HY.w, LZ.w, HZ.w, r.ox, r.oy, r.oz,
                                                         r.ix, r.iy, r.iz, r.tmin, r.tmax, &t);
                    }
                    unsigned long long m0 = __ballot(p0), m1 = __ballot(p1), m2 = __ballot(p2), m3 = __ballot(p3);
                    // the group's visiting order: near half first, near child first
                    const bool nA = (g >> (meta & 3u)) & 1u;
                    const bool nB = (g >> ((meta >> 2) & 3u)) & 1u;
                    const bool nC = (g >> ((meta >> 4) & 3u)) & 1u;
#define PX_SWAP2(C, I, J)                                                                \
    if (C) {                                                                              \
        const unsigned long long mb = m##I; m##I = m##J; m##J = mb;                       \
        const int32_t rb = r##I; r##I = r##J; r##J = rb;                                  \
    }
                    PX_SWAP2(nB, 0, 1)
                    PX_SWAP2(nC, 2, 3)
                    PX_SWAP2(nA, 0, 2)
                    PX_SWAP2(nA, 1, 3)
#undef PX_SWAP2
                    // nearest passing slot next, the others pushed farthest first
                    bool pend = false;
                    int32_t pref = 0;
                    unsigned long long pm = 0;
#define PX_VISIT2(K)                                                                     \
    if (m##K) {                                                                           \
        if (pend) {                                                                       \
            if (sp < PSTK) { s_ref[w][sp] = pref; s_msk[w][sp] = pm; ++sp; }              \
            else atomicOr(err, 1u);                                                       \
        }                                                                                 \
        pend = true; pref = r##K; pm = m##K;                                              \
    }
                    PX_VISIT2(3)
                    PX_VISIT2(2)
                    PX_VISIT2(1)
                    PX_VISIT2(0)
#undef PX_VISIT2
                    if (pend) {
                        cur = __builtin_amdgcn_readfirstlane(pref);
                        cm = pm;
                        continue;
                    }
                } else {
                    // leaf: the reference's pop-time box test at each lane's
                    // tmax, then its prims in order until the record flagged last
                    uint32_t pos = (uint32_t)(~cur);
                    const PrimRec64* pr = sv.prims64 + pos;
                    float4 c = cload4(pr, 2), d = cload4(pr, 3);
                    bool go;
                    if (gfast) {
                        go = act && slab_ord((g & 1u) ? d.x : c.y, (g & 1u) ? c.y : d.x, (g & 2u) ? d.y : c.z,
                                             (g & 2u) ? c.z : d.y, (g & 4u) ? d.z : c.w, (g & 4u) ? c.w : d.z,
                                             r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, r.tmin, r.tmax);
                    } else {
                        float tn;
                        go = act && slab(c.y, d.x, c.z, d.y, c.w, d.z, r.ox, r.oy, r.oz, r.ix, r.iy, r.iz, r.tmin,
                                         r.tmax, &tn);
                    }
                    if (__ballot(go)) {
                        while (true) {
                            const float4 a = cload4(pr, 0), b = cload4(pr, 1);
                            const uint32_t kind = as_u(d.w);
                            if (go) {
                                float tt, u = 0.0f, v = 0.0f;
                                bool h;
                                if ((kind & 0xFFu) == 0u) h = tri_test(a, b, c, r, &tt, &u, &v);
                                else h = sphere_test(sv.spheres[as_u(a.x)], r, &tt);
                                if (h) { r.tmax = tt; hprim = pos; hb1 = u; hb2 = v; }
                            }
                            if (kind & 0x100u) break;
                            ++pos;
                            ++pr;
                            c = cload4(pr, 2);
                            d = cload4(pr, 3);
                        }
                    }
                }
                if (sp == 0) break;
                --sp;
                cur = __builtin_amdgcn_readfirstlane(s_ref[w][sp]);
                cm = s_msk[w][sp];
            }
        }
        if (valid) {
            if (hprim != NO_HIT) hprim = sv.order[hprim];  // position -> prim id
            if (GEN && co.list && hprim == NO_HIT) co.miss[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            else hits[i] = Hit{hprim, hprim != NO_HIT ? r.tmax : 0.0f, hb1, hb2};
            if (!GEN && hprim != NO_HIT) rays[i].tmax = r.tmax;
        }
        if (GEN && co.list) {
            const bool h = valid && hprim != NO_HIT;
            const unsigned long long hm = __ballot(h);
            if (hm) {
                uint32_t b = 0;
                if (lane == 0) b = atomicAdd(co.count, (uint32_t)__popcll(hm));
                b = __builtin_amdgcn_readfirstlane(__shfl(b, 0));
                const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
                if (h) co.list[b + (uint32_t)__popcll(hm & lt)] = i;
            }
        }
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
// Three word ranges zeroed by one launch (the counters a stage needs
// cleared, one launch instead of one each).
__global__ void k_zero3_u32(uint32_t* a, size_t na, uint32_t* b, size_t nb, uint32_t* c, size_t nc) {
    const size_t n = na + nb + nc;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        if (i < na) a[i] = 0u;
        else if (i < na + nb) b[i - na] = 0u;
        else c[i - na - nb] = 0u;
    }
}

// Streaming copy, four 16-byte loads in flight a thread before their
// stores, marked non-temporal (read once, written once).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__global__ void k_copy16(uint4* dst_, const uint4* src_, size_t n) {
    u32x4* dst = reinterpret_cast<u32x4*>(dst_);
    const u32x4* src = reinterpret_cast<const u32x4*>(src_);
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const u32x4 a = __builtin_nontemporal_load(src + i), b = __builtin_nontemporal_load(src + i + stride);
        const u32x4 c = __builtin_nontemporal_load(src + i + 2 * stride);
        const u32x4 d = __builtin_nontemporal_load(src + i + 3 * stride);
        __builtin_nontemporal_store(a, dst + i);
        __builtin_nontemporal_store(b, dst + i + stride);
        __builtin_nontemporal_store(c, dst + i + 2 * stride);
        __builtin_nontemporal_store(d, dst + i + 3 * stride);
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

// Random 64-byte gathers: thread i reads record (i * GATHER_MUL) mod n of a
// buffer of n = 2^k records (an odd multiplier permutes [0, n), so every
// record is read exactly once and neighbouring lanes read far-apart records,
// as traversal lanes read records), four 16-byte loads each.  The loaded
// words are folded into one value that is stored only if it equals a
// constant no folded data here can give by chance often enough to matter
// (the store keeps the loads alive; its slot is the thread's own).
constexpr uint32_t GATHER_MUL = 0x9E3779B1u;
__global__ __launch_bounds__(256) void k_gather64(const uint4* buf, uint32_t n_mask, uint32_t n, uint32_t* out) {
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t acc = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint4* p = buf + 4ull * ((i * GATHER_MUL) & n_mask);
        const uint4 a = p[0], b = p[1], c = p[2], d = p[3];
        acc += (a.x ^ a.y ^ a.z ^ a.w) + (b.x ^ b.y ^ b.z ^ b.w) + (c.x ^ c.y ^ c.z ^ c.w) + (d.x ^ d.y ^ d.z ^ d.w);
    }
    if (acc == 0x5EEDF00Du) out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// Eight independent fma chains a thread: a measure of the achievable FP32
// vector rate (one fma per lane per instruction; no packed math).
__global__ __launch_bounds__(256) void k_fma_chain(float* out, uint32_t iters, float a, float b) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = (float)(gid & 1023u) * 1e-3f + (float)k;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = __builtin_fmaf(x[k], a, b);
    }
    float t = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) t += x[k];
    out[gid] = t;
}

int g_cu_count = 0;

constexpr int PLAIN_BLOCKS_PER_CU = 16;

int plain_grid(uint32_t n, int block) {
    const int cus = device_cu_count();
    long long want = ((long long)n + block - 1) / block;
    long long cap = (long long)cus * PLAIN_BLOCKS_PER_CU;  // static grid-stride schedule
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    return (int)want;
}

template <typename K>
int wide_grid(K kernel, const TraceTuning& tun) {
    int per_cu = tun.blocks_per_cu;
    if (per_cu <= 0) {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, tun.block, 0) != hipSuccess || occ < 1)
            occ = 2;
        per_cu = occ;
    }
    return device_cu_count() * per_cu;
}

}  // namespace

int device_cu_count() {
    if (g_cu_count == 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
        g_cu_count = cus;
    }
    return g_cu_count;
}

hipError_t copy_bytes(void* dst, const void* src, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 16;
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    const size_t cap = (size_t)device_cu_count() * 16;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, st, (uint4*)dst, (const uint4*)src, n);
    return hipGetLastError();
}

uint64_t gather_probe_threads() { return (uint64_t)device_cu_count() * 8 * 256; }

hipError_t gather_probe(const void* buf, uint64_t records, uint32_t* out, hipStream_t st) {
    const unsigned blocks = (unsigned)(gather_probe_threads() / 256);
    hipLaunchKernelGGL(k_gather64, dim3(blocks), dim3(256), 0, st, (const uint4*)buf, (uint32_t)(records - 1),
                       (uint32_t)records, out);
    return hipGetLastError();
}

uint64_t fma_chain_threads() { return (uint64_t)device_cu_count() * 8 * 256; }

hipError_t fma_chain(float* out, uint32_t iters, hipStream_t st) {
    const unsigned blocks = (unsigned)(fma_chain_threads() / 256);
    hipLaunchKernelGGL(k_fma_chain, dim3(blocks), dim3(256), 0, st, out, iters, 0.999f, 1e-3f);
    return hipGetLastError();
}

hipError_t fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, st, p, value, n);
    return hipGetLastError();
}
hipError_t zero3_u32(uint32_t* a, size_t na, uint32_t* b, size_t nb, uint32_t* c, size_t nc, hipStream_t st) {
    const size_t n = na + nb + nc;
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero3_u32, dim3((unsigned)blocks), dim3(256), 0, st, a, na, b, nb, c, nc);
    return hipGetLastError();
}

static uint32_t plain_levels(const SceneView& s) {
    // the reference-order stack holds at most depth + 1 entries
    uint32_t need = s.depth + 2;
    return need > (uint32_t)STK_PLAIN ? need - (uint32_t)STK_PLAIN : 0u;
}
static uint32_t wide_levels(const SceneView& s, int stack) {
    // at most three pending slots per record level on the current path (of
    // the ordered records or of the occlusion records)
    uint32_t need = 3 * std::max(s.qdepth, s.odepth) + 1;
    return need > (uint32_t)stack ? need - (uint32_t)stack : 0u;
}

// the per-ray kernel instance for a stack depth
template <bool ANY, bool SEG, bool CQ>
static auto wide_kernel(int stack) -> decltype(&k_trace_wide<ANY, SEG, CQ, STK_DEFAULT>) {
    switch (stack) {
        case 8: return &k_trace_wide<ANY, SEG, CQ, 8>;
        case 24: return &k_trace_wide<ANY, SEG, CQ, 24>;
        default: return &k_trace_wide<ANY, SEG, CQ, 16>;
    }
}

// the packet kernel's stack holds at most three pending slots per record
// level on the current path plus one
static bool packet_fits(const SceneView& s) { return 3 * s.qdepth + 1 <= (uint32_t)PSTK; }
bool camera_fusable(const SceneView& s, const TraceTuning& tun) {
    return tun.scheme == SCHEME_PACKET && packet_fits(s) && s.prims64 != nullptr;
}
static bool use_compressed(const SceneView& s, const TraceTuning& tun) {
    return s.cnodes != nullptr && tun.scheme != SCHEME_RAY_EXACT;
}

// Lanes per ray of a batched occlusion launch of up to n rays (n: the count,
// or the capacity when the count is on the device).  Several lanes per ray
// (k_occlude_group) need the compressed records and a stack bound with room
// for more than one entry a step.  By launch size, from shadow-ray launches
// of direct-lighting frames (occlusion ms at 1 / 4 / 8 lanes per ray):
//   512x512 1 spp, 16K-triangle soup, 62,815 of 262,144:     0.201 / 0.060 / 0.048
//   512x512 4 spp, the same soup, 251,606 of 1,048,576:       0.266 / 0.127 / 0.106
//   512x512 16 spp, the same soup, 1,004,569 of 4,194,304:    0.518 / 0.367 / 0.358
//   512x512 1 spp, 82K-triangle sphere, 413,511 of 1,048,576: 0.311 / 0.173 / 0.161
//   512x512 4 spp, that sphere, 1,654,016 of 4,194,304:       0.573 / 0.535 / 0.559
//   1024x1024 16 spp, 1.3M-triangle sphere, 26.5M of 67.1M:   3.47 / 10.7 / 10.7
// and from launches of random shadow segments (origins in the box, uniform
// directions, lengths up to half its diagonal) in random-triangle soups,
// four triangles a leaf (ms at 1 / 4 / 8 lanes per ray):
//   200K triangles, 131,072 rays:     0.345 / 0.147 / 0.104
//   200K triangles, 1,048,576 rays:   0.718 / 0.794 / 0.567
//   200K triangles, 4,194,304 rays:   1.53 / 2.97 / 2.09
//   2M triangles, 2,097,152 rays:     1.31 / 1.63 / 1.30
//   2M triangles, 4,194,304 rays:     2.20 / 3.23 / 2.55
// With few rays per SIMD the per-ray walk is a chain of dependent record
// loads; with many, the group's idle lanes (a narrow frontier) cost more.
// Four lanes a ray hold more registers and LDS per lane and lose to eight
// wherever either wins.
constexpr uint32_t GROUP_MAX = 1u << 20;
static int group_lanes(const SceneView& s, uint32_t n, const TraceTuning& tun) {
    if (tun.lanes == 1 || !s.cnodes || !s.prims64) return 1;
    if (3u * (uint64_t)s.odepth + 8u > (uint64_t)GSTK) return 1;
    if (tun.lanes >= 2) return tun.lanes;
    return n <= GROUP_MAX ? 8 : 1;
}

template <bool SEG>
static auto group_kernel(int G) -> decltype(&k_occlude_group<8, SEG>) {
    return G == 4 ? &k_occlude_group<4, SEG> : G == 8 ? &k_occlude_group<8, SEG> : &k_occlude_group<16, SEG>;
}

size_t trace_any_block_spill(const SceneView& s, const TraceTuning& tun) {
    return (size_t)wide_levels(s, tun.stack) * tun.block * sizeof(uint2);
}

size_t trace_spill_bytes(const SceneView& s, int path, const TraceTuning& tun) {
    const int block = tun.block;
    if (path == PATH_PLAIN) {
        size_t threads = (size_t)device_cu_count() * PLAIN_BLOCKS_PER_CU * block;
        return (size_t)plain_levels(s) * threads * sizeof(uint32_t);
    }
    TraceTuning t = tun;
    // the largest persistent grid of any batched traversal kernel
    size_t threads = 0;
    const int k = t.stack;
    for (size_t g : {wide_grid(wide_kernel<false, false, false>(k), t), wide_grid(wide_kernel<true, false, false>(k), t),
                     wide_grid(wide_kernel<false, true, false>(k), t), wide_grid(wide_kernel<true, true, false>(k), t),
                     wide_grid(wide_kernel<false, false, true>(k), t), wide_grid(wide_kernel<true, false, true>(k), t),
                     wide_grid(wide_kernel<false, true, true>(k), t), wide_grid(wide_kernel<true, true, true>(k), t)})
        threads = std::max(threads, g * (size_t)block);
    return (size_t)wide_levels(s, k) * threads * sizeof(uint2);
}

template <bool ANY>
static hipError_t launch(const SceneView& s, Ray* rays, uint32_t n, Hit* hits, uint32_t* occ, uint8_t* occ8,
                         int path, const TraceTuning& tun, TraceScratch* scr, hipStream_t st,
                         const uint32_t* n_dev) {
    if (n == 0) return hipSuccess;
    const int block = tun.block;
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    if (ANY && !occ8) {
        hipError_t e = fill_u32(occ, 0u, ((size_t)n + 31) / 32, st);
        if (e != hipSuccess) return e;
    }
    uint32_t* err = scr->counter + 1;
    if (path == PATH_PLAIN) {
        if (n_dev) return hipErrorInvalidValue;
        const int grid = plain_grid(n, block);
        const uint32_t lv = plain_levels(s);
        if ((size_t)lv * grid * block * sizeof(uint32_t) > scr->spill_bytes) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_trace_plain<ANY>, dim3(grid), dim3(block), 0, st, s, rays, n, hits, occ, occ8,
                           (uint32_t*)scr->spill, lv, err);
        return hipGetLastError();
    }
    // occlusion has no packet form (pbrtx_set_tuning refuses it)
    if (ANY && tun.scheme == SCHEME_PACKET) return hipErrorNotSupported;
    if (!ANY && tun.scheme == SCHEME_PACKET && packet_fits(s) && s.prims64) {
        hipError_t e = fill_u32(scr->counter, 0u, 1, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_closest_packet<false>, dim3(wide_grid(k_closest_packet<false>, tun)), dim3(block), 0,
                           st, s, rays, n, n_dev, hits, scr->counter, err, CamParams{}, CamTable{}, CamOut{});
        return hipGetLastError();
    }
    if (!stack_supported(tun.stack)) return hipErrorInvalidValue;
    const bool cq = use_compressed(s, tun);
    if (ANY && cq && s.prims64) {
        const int G = group_lanes(s, n, tun);
        if (G > 1) {
            const uint32_t lim = (uint32_t)GSTK - 3u * s.odepth;
            const uint32_t gpb = (uint32_t)block / (uint32_t)G;
            const unsigned grid = (unsigned)(((size_t)n + gpb - 1) / gpb);
            const auto kern = group_kernel<false>(G);
            hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, s, rays, n, n_dev, occ, occ8, lim, err,
                               (const TraceSeg*)nullptr, 0u);
            return hipGetLastError();
        }
    }
    const auto kern = cq ? wide_kernel<ANY, false, true>(tun.stack) : wide_kernel<ANY, false, false>(tun.stack);
    int grid = wide_grid(kern, tun);
    const uint32_t lv = wide_levels(s, tun.stack);
    const size_t per_block = (size_t)lv * block * sizeof(uint2);
    if (per_block * grid > scr->spill_bytes) {
        // a fixed scratch: as many persistent blocks as its spill holds
        if (!scr->cap_grid || per_block > scr->spill_bytes) return hipErrorInvalidValue;
        grid = (int)(scr->spill_bytes / per_block);
    }
    uint32_t* parts = scr->counter + WIDE_PARTS_AT;
    hipError_t e = fill_u32(parts, 0u, (size_t)WIDE_PARTS * WIDE_PART_STRIDE, st);
    if (e != hipSuccess) return e;
    const uint32_t refill = (uint32_t)(tun.refill_pct < 0 ? 0 : (tun.refill_pct > 100 ? 100 : tun.refill_pct));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, s, rays, n, n_dev, hits, occ, occ8, parts,
                       (uint2*)scr->spill, lv, refill, err, (const TraceSeg*)nullptr, 0u);
    return hipGetLastError();
}

hipError_t trace_closest(const SceneView& s, Ray* rays, uint32_t n, Hit* hits, int path,
                         const TraceTuning& tun, TraceScratch* scr, hipStream_t st, const uint32_t* n_dev) {
    return launch<false>(s, rays, n, hits, nullptr, nullptr, path, tun, scr, st, n_dev);
}
hipError_t trace_any(const SceneView& s, const Ray* rays, uint32_t n, uint32_t* occ, int path,
                     const TraceTuning& tun, TraceScratch* scr, hipStream_t st, const uint32_t* n_dev) {
    return launch<true>(s, const_cast<Ray*>(rays), n, nullptr, occ, nullptr, path, tun, scr, st, n_dev);
}
hipError_t trace_any_bytes(const SceneView& s, const Ray* rays, uint32_t n, uint8_t* occ8, int path,
                           const TraceTuning& tun, TraceScratch* scr, hipStream_t st, const uint32_t* n_dev) {
    return launch<true>(s, const_cast<Ray*>(rays), n, nullptr, nullptr, occ8, path, tun, scr, st, n_dev);
}

// Scratch of the entry point that allocates nothing: the trace counter block
// (first NOALLOC_COUNTER_BYTES) and the spill stacks (the rest), one array per
// GPU for the life of the process.  Host code uses the address
// hipGetSymbolAddress returns for the current device (noalloc_scratch), never
// the array's name.
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

int noalloc_scratch(TraceScratch* out, hipError_t* err) {
    static std::mutex mu;
    static void* ptr[NOALLOC_MAX_DEVICES] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { *err = e; return -1; }
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) { *err = hipErrorInvalidDevice; return -1; }
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) { *err = e != hipSuccess ? e : hipErrorInvalidSymbol; return -1; }
        ptr[dev] = p;
    }
    out->counter = (uint32_t*)ptr[dev];
    out->spill = (uint8_t*)ptr[dev] + NOALLOC_COUNTER_BYTES;
    out->spill_bytes = NOALLOC_SCRATCH_BYTES - NOALLOC_COUNTER_BYTES;
    out->cap_grid = true;
    return 0;
}

hipError_t trace_camera(const SceneView& s, const CamParams& cam, uint32_t n, Hit* hits, const TraceTuning& tun,
                        TraceScratch* scr, hipStream_t st, const CamTable* ct, const CamOut* co) {
    if (co && ct) return hipErrorInvalidValue;  // the list form is for one camera
    if (n == 0) return hipSuccess;
    if (!camera_fusable(s, tun)) return hipErrorNotSupported;
    const int block = tun.block;
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    // (the camera kernel runs a fixed schedule: no work counter to clear)
    // Packets over pixel tiles (64 / spp pixels each, as square as a power
    // of two allows) when the launch is whole rows of the image starting at
    // a row, spp is a power of two up to 64 and the tiles divide its rows and
    // the row width; else packets of 64 consecutive samples (64 / spp pixels
    // of one row).  Which lane makes which sample's ray is all this changes.
    CamParams c = cam;
    c.tile_lw = 0xFFu;
    const uint32_t spp = cam.spp;
    if (!ct && spp >= 1 && spp <= 64 && (spp & (spp - 1)) == 0 && cam.xres > 0) {
        uint32_t sl = 0;
        while ((1u << sl) < spp) ++sl;
        const uint32_t pl = 6 - sl;           // log2 pixels per tile
        const uint32_t wl = (pl + 1) / 2;     // log2 tile width
        const uint32_t hl = pl - wl;
        const uint64_t row = (uint64_t)cam.xres * spp;
        if (cam.g0 % row == 0 && n % row == 0 && (cam.xres & ((1u << wl) - 1u)) == 0 &&
            ((n / row) & ((1u << hl) - 1u)) == 0) {
            c.tile_lw = wl;
            c.spp_l = sl;
        }
    }
    // several cameras: their own instance, so that the single-camera kernel
    // keeps its registers
    if (ct)
        hipLaunchKernelGGL((k_closest_packet<true, true>), dim3(wide_grid(k_closest_packet<true, true>, tun)),
                           dim3(block), 0, st, s, (Ray*)nullptr, n, (const uint32_t*)nullptr, hits, scr->counter,
                           scr->counter + 1, c, *ct, CamOut{});
    else
        hipLaunchKernelGGL((k_closest_packet<true, false>), dim3(wide_grid(k_closest_packet<true, false>, tun)),
                           dim3(block), 0, st, s, (Ray*)nullptr, n, (const uint32_t*)nullptr, hits, scr->counter,
                           scr->counter + 1, c, CamTable{}, co ? *co : CamOut{});
    return hipGetLastError();
}

hipError_t trace_segments(const SceneView& s, bool any, const TraceSeg* d_segs, uint32_t nseg, uint32_t total,
                          const TraceTuning& tun, TraceScratch* scr, hipStream_t st) {
    if (total == 0 || nseg == 0) return hipSuccess;
    const int block = tun.block;
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    if (!stack_supported(tun.stack)) return hipErrorInvalidValue;
    const uint32_t lv = wide_levels(s, tun.stack);
    const bool cq = use_compressed(s, tun);
    uint32_t* err = scr->counter + 1;
    if (any && cq && s.prims64) {
        const int G = group_lanes(s, total, tun);
        if (G > 1) {
            const uint32_t lim = (uint32_t)GSTK - 3u * s.odepth;
            const uint32_t gpb = (uint32_t)block / (uint32_t)G;
            const unsigned grid = (unsigned)(((size_t)total + gpb - 1) / gpb);
            hipLaunchKernelGGL(group_kernel<true>(G), dim3(grid), dim3(block), 0, st, s, (const Ray*)nullptr, total,
                               (const uint32_t*)nullptr, (uint32_t*)nullptr, (uint8_t*)nullptr, lim, err, d_segs,
                               nseg);
            return hipGetLastError();
        }
    }
    const auto kern = cq ? (any ? wide_kernel<true, true, true>(tun.stack) : wide_kernel<false, true, true>(tun.stack))
                         : (any ? wide_kernel<true, true, false>(tun.stack) : wide_kernel<false, true, false>(tun.stack));
    const int grid = wide_grid(kern, tun);
    if ((size_t)lv * grid * block * sizeof(uint2) > scr->spill_bytes) return hipErrorInvalidValue;
    uint32_t* parts = scr->counter + WIDE_PARTS_AT;
    hipError_t e = fill_u32(parts, 0u, (size_t)WIDE_PARTS * WIDE_PART_STRIDE, st);
    if (e != hipSuccess) return e;
    const uint32_t refill = (uint32_t)(tun.refill_pct < 0 ? 0 : (tun.refill_pct > 100 ? 100 : tun.refill_pct));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, s, (Ray*)nullptr, total, (const uint32_t*)nullptr,
                       (Hit*)nullptr, (uint32_t*)nullptr, (uint8_t*)nullptr, parts, (uint2*)scr->spill, lv, refill,
                       err, d_segs, nseg);
    return hipGetLastError();
}

}  // namespace px
