// bvh_build.cpp -- host-side BVH construction (BVHAccelerator::new,
// primitive/aggregates/bvh.rs:105-362) producing the reference's tree exactly,
// written as an iterative pre-order build over one index permutation with
// in-place partitions (no per-node allocation), then packed into the two
// GPU layouts (32-byte reference nodes and 64-byte child-pair nodes).
#include "bvh_build.hpp"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>

namespace px {
namespace {

struct Box {
    f3 lo, hi;
};
// f32::min / f32::max ignore a NaN operand (fmin / fmax); argument order as
// in bbox.rs:155-183 so signed zeros resolve as they do there.
inline float rmin(float a, float b) { return fminf(a, b); }
inline float rmax(float a, float b) { return fmaxf(a, b); }
inline float comp(const f3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// BBox::new: bbox.rs:36-41
inline Box box_empty() { return Box{mk3(FLT_MAX, FLT_MAX, FLT_MAX), mk3(-FLT_MAX, -FLT_MAX, -FLT_MAX)}; }
// Union<Point>: bbox.rs:155-168
inline Box box_union_p(const Box& b, const f3& p) {
    return Box{mk3(rmin(b.lo.x, p.x), rmin(b.lo.y, p.y), rmin(b.lo.z, p.z)),
               mk3(rmax(b.hi.x, p.x), rmax(b.hi.y, p.y), rmax(b.hi.z, p.z))};
}
// Union<BBox>: bbox.rs:170-183
inline Box box_union(const Box& a, const Box& b) {
    return Box{mk3(rmin(a.lo.x, b.lo.x), rmin(a.lo.y, b.lo.y), rmin(a.lo.z, b.lo.z)),
               mk3(rmax(a.hi.x, b.hi.x), rmax(a.hi.y, b.hi.y), rmax(a.hi.z, b.hi.z))};
}
// surface_area: bbox.rs:73-78
inline float surface_area(const Box& b) {
    float dx = rmax(b.hi.x - b.lo.x, 0.0f);
    float dy = rmax(b.hi.y - b.lo.y, 0.0f);
    float dz = rmax(b.hi.z - b.lo.z, 0.0f);
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}
// max_extent: bbox.rs:87-96
inline int max_extent(const Box& b) {
    f3 d = sub3(b.hi, b.lo);
    if (d.x > d.y && d.x > d.z) return 0;
    if (d.y > d.z) return 1;
    return 2;
}

struct M4 {
    float m[4][4];
};
inline M4 load_m4(const float* a) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = a[4 * i + j];
    return r;
}
// ApplyTransform<Point>: transform.rs:195-207
inline f3 xf_point(const M4& m, f3 p) {
    float x = p.x, y = p.y, z = p.z;
    float xt = m.m[0][0] * x + m.m[0][1] * y + m.m[0][2] * z + m.m[0][3];
    float yt = m.m[1][0] * x + m.m[1][1] * y + m.m[1][2] * z + m.m[1][3];
    float zt = m.m[2][0] * x + m.m[2][1] * y + m.m[2][2] * z + m.m[2][3];
    float w = m.m[3][0] * x + m.m[3][1] * y + m.m[3][2] * z + m.m[3][3];
    if (w != 1.0f) return mk3(xt / w, yt / w, zt / w);
    return mk3(xt, yt, zt);
}
// ApplyTransform<Vector>: transform.rs:210-218
inline f3 xf_vector(const M4& m, f3 v) {
    return mk3(m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z,
               m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z,
               m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z);
}

// partition_by: utils/mod.rs:112-169 over prim indices keyed by centroid[dim].
// (The reference recurses on a sub-slice with that sub-slice's midpoint; the
// loop keeps that.)
void partition_by(uint32_t* v, size_t nv, const std::vector<f3>& cen, int dim) {
    auto key = [&](uint32_t p) { return comp(cen[p], dim); };
    while (true) {
        if (nv < 3) {
            if (nv == 2 && key(v[1]) < key(v[0])) std::swap(v[0], v[1]);
            return;
        }
        float fst = key(v[0]), mid = key(v[nv / 2]), lst = key(v[nv - 1]);
        float pivot;
        if (fst < mid && mid < lst) pivot = mid;
        else if (mid < fst && fst < lst) pivot = fst;
        else pivot = lst;
        size_t ls = 0, np = 0;
        for (size_t i = 0; i < nv; ++i) {
            float bv = key(v[i]);
            if (bv < pivot) {
                std::swap(v[ls + np], v[i]);
                std::swap(v[ls + np], v[ls]);
                ls += 1;
            } else if (bv == pivot) {
                std::swap(v[ls + np], v[i]);
                np += 1;
            }
        }
        size_t pidx = std::max<size_t>(ls, 1);
        if (pidx + np <= nv / 2) {
            v += pidx;
            nv -= pidx;
        } else if (pidx >= nv / 2) {
            nv = pidx;
        } else {
            return;
        }
    }
}

template <typename Pred>
size_t stable_partition(uint32_t* a, size_t n, Pred pred, uint32_t* tmp) {
    size_t l = 0, r = 0;
    for (size_t i = 0; i < n; ++i) {
        uint32_t x = a[i];
        if (pred(x)) a[l++] = x;
        else tmp[r++] = x;
    }
    if (r) memcpy(a + l, tmp, r * sizeof(uint32_t));
    return l;
}

// bucket_for_prim: bvh.rs:118-130 (`as usize` saturates: NaN and negatives -> 0)
inline int bucket_for(float c, float cmin, float cmax) {
    float pdist = c - cmin;
    float dist = cmax - cmin;
    float f = 12.0f * (pdist / dist);
    size_t b = (f != f || f <= 0.0f) ? 0 : (size_t)f;
    if (b >= 12) b = 11;
    return (int)b;
}

inline float bits_f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

}  // namespace

// The float the GPU decodes from byte q at origin o and scale s (q * s is
// exact: an 8-bit integer times a power of two, so the device's fused
// multiply-add rounds the same sum once, as here).
static float cq_decode(float o, uint32_t q, float s) {
    const float t = (float)q * s;
    return o + t;
}

// Quantizes one axis of a record: picks the scale exponent, then for every
// slot the largest q whose decode is <= lo and the smallest whose decode is
// >= hi.  Returns false when the union box is not finite or no exponent fits.
static bool cq_axis(float ulo, float uhi, const float* lo, const float* hi, const bool* used, uint8_t* qlo,
                    uint8_t* qhi, uint32_t* ebyte) {
    if (!(std::fabs(ulo) <= FLT_MAX) || !(std::fabs(uhi) <= FLT_MAX)) return false;
    const double ext = (double)uhi - (double)ulo;
    int e = -126;
    if (ext > 0.0) {
        e = (int)std::ceil(std::log2(ext / 255.0));
        if (e < -126) e = -126;
    }
    for (;; ++e) {
        // q * 2^e must stay finite for every byte q (the device decodes with
        // one fused multiply-add, equal to multiply-then-add only then)
        if (e > 119) return false;
        const float s = std::ldexp(1.0f, e);
        if (cq_decode(ulo, 255u, s) >= uhi) break;
    }
    const float s = std::ldexp(1.0f, e);
    for (int k = 0; k < 4; ++k) {
        if (!used[k]) { qlo[k] = 255; qhi[k] = 0; continue; }
        double a = std::floor(((double)lo[k] - (double)ulo) / (double)s);
        int ql = a < 0 ? 0 : (a > 255 ? 255 : (int)a);
        while (ql > 0 && cq_decode(ulo, (uint32_t)ql, s) > lo[k]) --ql;
        double b = std::ceil(((double)hi[k] - (double)ulo) / (double)s);
        int qh = b < 0 ? 0 : (b > 255 ? 255 : (int)b);
        while (qh < 255 && cq_decode(ulo, (uint32_t)qh, s) < hi[k]) ++qh;
        if (cq_decode(ulo, (uint32_t)ql, s) > lo[k] || cq_decode(ulo, (uint32_t)qh, s) < hi[k]) return false;
        qlo[k] = (uint8_t)ql;
        qhi[k] = (uint8_t)qh;
    }
    *ebyte = (uint32_t)(e + 127);
    return true;
}

// Builds the compressed layout from the four-wide records and the leaves of
// the reference tree.  Returns false if some record cannot be represented
// (non-finite bounds); the exact layout is used then.
static bool compress_layout(const std::vector<RNode>& nodes, HostBvh* out) {
    out->cnodes.assign(out->qnodes.size(), CNode{});
    for (size_t i = 0; i < out->qnodes.size(); ++i) {
        const QNode& q = out->qnodes[i];
        CNode& c = out->cnodes[i];
        bool used[4];
        float u[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int k = 0; k < 4; ++k) {
            used[k] = q.ref[k] != QEMPTY;
            c.ref[k] = q.ref[k];
            if (!used[k]) continue;
            u[0] = std::min(u[0], q.lx[k]); u[1] = std::min(u[1], q.ly[k]); u[2] = std::min(u[2], q.lz[k]);
            u[3] = std::max(u[3], q.hx[k]); u[4] = std::max(u[4], q.hy[k]); u[5] = std::max(u[5], q.hz[k]);
        }
        uint32_t ex, ey, ez;
        if (!cq_axis(u[0], u[3], q.lx, q.hx, used, c.q + 0, c.q + 4, &ex)) return false;
        if (!cq_axis(u[1], u[4], q.ly, q.hy, used, c.q + 8, c.q + 12, &ey)) return false;
        if (!cq_axis(u[2], u[5], q.lz, q.hz, used, c.q + 16, c.q + 20, &ez)) return false;
        c.ox = u[0]; c.oy = u[1]; c.oz = u[2];
        c.emeta = ex | (ey << 8) | (ez << 16) | ((q.meta & 63u) << 24);
    }
    // prim records with their leaf's exact box
    out->prims64.assign(out->prims.size(), PrimRec64{});
    for (size_t i = 0; i < nodes.size(); ++i) {
        if ((nodes[i].meta & 3u) != 3u) continue;
        const uint32_t first = nodes[i].off, cnt = nodes[i].meta >> 2;
        for (uint32_t k = 0; k < cnt; ++k) {
            const PrimRec& p = out->prims[first + k];
            PrimRec64& r = out->prims64[first + k];
            for (int j = 0; j < 9; ++j) r.w[j] = p.w[j];
            uint32_t kind;
            memcpy(&kind, &p.w[11], 4);
            if ((kind & 0xFFu) != 0u) r.w[0] = p.w[9];  // sphere record index
            for (int j = 0; j < 3; ++j) { r.w[9 + j] = nodes[i].lo[j]; r.w[12 + j] = nodes[i].hi[j]; }
            r.w[15] = p.w[11];
        }
    }
    return true;
}

bool build_bvh(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t nfaces,
               const SphereIn* sph, uint32_t nsph, int split, uint32_t max_prims, HostBvh* out,
               std::string* err) {
    auto t0 = std::chrono::steady_clock::now();
    const uint64_t N64 = (uint64_t)nfaces + nsph;
    if (N64 == 0) { *err = "scene has no primitives"; return false; }
    if (N64 >= (1ull << 30)) { *err = "too many primitives"; return false; }
    if (split < 0 || split > 2) { *err = "unknown split method"; return false; }
    if (nfaces && (!verts || !idx)) { *err = "null mesh arrays"; return false; }
    const uint32_t N = (uint32_t)N64;
    for (uint64_t i = 0; i < 3ull * nfaces; ++i)
        if (idx[i] >= nverts) { *err = "vertex index out of range"; return false; }

    std::vector<Box> pb(N);
    std::vector<f3> pc(N);
    auto V = [&](uint32_t i) { return mk3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]); };
    // Triangle::world_bound: mesh.rs:196-205 over the refined triple [c,b,a]
    for (uint32_t f = 0; f < nfaces; ++f) {
        Box b = box_empty();
        b = box_union_p(b, V(idx[3 * f + 2]));
        b = box_union_p(b, V(idx[3 * f + 1]));
        b = box_union_p(b, V(idx[3 * f + 0]));
        pb[f] = b;
    }
    out->spheres.resize(nsph);
    for (uint32_t s = 0; s < nsph; ++s) {
        const SphereIn& in = sph[s];
        if (!(in.radius > 0.0f)) { *err = "sphere radius must be positive"; return false; }
        SphereRec& rec = out->spheres[s];
        float r = in.radius;
        // Sphere::new: sphere.rs:28-44
        rec.radius = r;
        rec.zmin = std::min(std::max(fminf(in.z0, in.z1), -r), r);
        rec.zmax = std::min(std::max(fmaxf(in.z0, in.z1), -r), r);
        rec.phimax = std::min(std::max(in.phimax_deg, 0.0f), 360.0f) * PI_F / 180.0f;
        memcpy(rec.w2o, in.o2w_inv, sizeof(rec.w2o));  // Transform::inverse swaps m, m_inv
        // world_bound: sphere.rs:111-126 + transform.rs:248-264
        M4 m = load_m4(in.o2w);
        Box ob{mk3(-r, -r, rec.zmin), mk3(r, r, rec.zmax)};
        f3 tx = xf_vector(m, mk3(ob.hi.x - ob.lo.x, 0.0f, 0.0f));
        f3 ty = xf_vector(m, mk3(0.0f, ob.hi.y - ob.lo.y, 0.0f));
        f3 tz = xf_vector(m, mk3(0.0f, 0.0f, ob.hi.z - ob.lo.z));
        f3 tp = xf_point(m, ob.lo);
        Box b{tp, tp};
        b = box_union_p(b, add3(tp, tx));
        b = box_union_p(b, add3(tp, ty));
        b = box_union_p(b, add3(tp, tz));
        b = box_union_p(b, add3(add3(tp, tx), ty));
        b = box_union_p(b, add3(add3(tp, tx), tz));
        b = box_union_p(b, add3(add3(tp, ty), tz));
        b = box_union_p(b, add3(add3(add3(tp, tx), ty), tz));
        pb[nfaces + s] = b;
    }
    for (uint32_t i = 0; i < N; ++i) pc[i] = mul3(add3(pb[i].lo, pb[i].hi), 0.5f);  // bvh.rs:77

    std::vector<uint32_t> ord(N), tmp(N);
    for (uint32_t i = 0; i < N; ++i) ord[i] = i;
    std::vector<RNode>& nodes = out->rnodes;
    nodes.clear();
    nodes.reserve(2 * (size_t)N);
    std::vector<Box> nb;
    nb.reserve(2 * (size_t)N);

    struct Task {
        uint32_t b, e;
        int64_t parent;
        uint32_t depth;
    };
    std::vector<Task> stack;
    stack.push_back(Task{0, N, -1, 0});
    uint32_t maxdepth = 0;
    std::vector<uint8_t> bk;  // bucket per position of the current range
    while (!stack.empty()) {
        Task T = stack.back();
        stack.pop_back();
        const uint32_t id = (uint32_t)nodes.size();
        nodes.push_back(RNode{});
        nb.push_back(Box{});
        if (T.parent >= 0) nodes[(size_t)T.parent].off = id;
        maxdepth = std::max(maxdepth, T.depth);
        const uint32_t b = T.b, e = T.e, n = e - b;
        // recursive_build bbox: bvh.rs:192-194 (fold in the range's order)
        Box bb = box_empty();
        for (uint32_t i = b; i < e; ++i) bb = box_union(bb, pb[ord[i]]);
        auto make_leaf = [&]() {
            nodes[id].off = b;
            nodes[id].meta = 3u | (n << 2);
            nb[id] = bb;
        };
        if (n == 1) { make_leaf(); continue; }
        Box cb = box_empty();
        for (uint32_t i = b; i < e; ++i) cb = box_union_p(cb, pc[ord[i]]);
        const int dim = max_extent(cb);
        const float cmin = comp(cb.lo, dim), cmax = comp(cb.hi, dim);
        if (cmin == cmax) { make_leaf(); continue; }
        uint32_t mid;
        bool sah_leaf = false;
        auto equal_counts = [&]() {
            partition_by(ord.data() + b, n, pc, dim);
            return b + n / 2;
        };
        if (split == 1) {
            // split_middle: bvh.rs:86-90
            float pmid = 0.5f * (cmin + cmax);
            mid = b + (uint32_t)stable_partition(
                          ord.data() + b, n, [&](uint32_t p) { return comp(pc[p], dim) < pmid; }, tmp.data());
        } else if (split == 2 || n <= 4) {
            mid = equal_counts();
        } else {
            // split_surface_area_heuristic: bvh.rs:105-186
            uint32_t cnt[12] = {0};
            Box bbk[12];
            for (int k = 0; k < 12; ++k) bbk[k] = box_empty();
            bk.resize(n);
            for (uint32_t i = 0; i < n; ++i) {
                uint32_t p = ord[b + i];
                int k = bucket_for(comp(pc[p], dim), cmin, cmax);
                bk[i] = (uint8_t)k;
                cnt[k] += 1;
                bbk[k] = box_union(bbk[k], pb[p]);
            }
            const float tsa = surface_area(bb);
            int best = 0;
            float mc = FLT_MAX;
            Box b0 = box_empty();
            uint32_t c0 = 0;
            for (int i = 0; i < 11; ++i) {
                c0 += cnt[i];
                b0 = box_union(b0, bbk[i]);  // fold over buckets 0..i in order
                Box b1 = box_empty();
                uint32_t c1 = 0;
                for (int k = i + 1; k < 12; ++k) { c1 += cnt[k]; b1 = box_union(b1, bbk[k]); }
                float b0sa = (float)c0 * surface_area(b0);
                float b1sa = (float)c1 * surface_area(b1);
                float cost = 0.125f * (b0sa + b1sa) / tsa;
                if (cost < mc) { best = i; mc = cost; }
            }
            size_t mcu = (mc != mc || mc <= 0.0f) ? 0 : (size_t)mc;
            if (max_prims < n || mcu < n) {
                // stable partition on the precomputed buckets
                uint32_t* a = ord.data() + b;
                size_t l = 0, r = 0;
                for (uint32_t i = 0; i < n; ++i) {
                    uint32_t x = a[i];
                    if ((int)bk[i] <= best) a[l++] = x;
                    else tmp[r++] = x;
                }
                if (r) memcpy(a + l, tmp.data(), r * sizeof(uint32_t));
                mid = b + (uint32_t)l;
            } else {
                sah_leaf = true;
                mid = b;
            }
        }
        if (sah_leaf) { make_leaf(); continue; }
        if (mid == b || mid == e) {
            *err = "BVH split produced an empty side";  // bvh.rs:237-238 asserts
            return false;
        }
        nodes[id].meta = (uint32_t)dim;
        stack.push_back(Task{mid, e, (int64_t)id, T.depth + 1});
        stack.push_back(Task{b, mid, -1, T.depth + 1});
    }
    // interior bounds = left ∪ right (bvh.rs:247), children follow parents
    for (size_t i = nodes.size(); i-- > 0;) {
        if ((nodes[i].meta & 3u) != 3u) nb[i] = box_union(nb[i + 1], nb[nodes[i].off]);
    }
    for (size_t i = 0; i < nodes.size(); ++i) {
        nodes[i].lo[0] = nb[i].lo.x; nodes[i].lo[1] = nb[i].lo.y; nodes[i].lo[2] = nb[i].lo.z;
        nodes[i].hi[0] = nb[i].hi.x; nodes[i].hi[1] = nb[i].hi.y; nodes[i].hi[2] = nb[i].hi.z;
    }
    out->depth = maxdepth;
    out->order = ord;
    out->nfaces = nfaces;
    out->nspheres = nsph;
    out->root_lo[0] = nb[0].lo.x; out->root_lo[1] = nb[0].lo.y; out->root_lo[2] = nb[0].lo.z;
    out->root_hi[0] = nb[0].hi.x; out->root_hi[1] = nb[0].hi.y; out->root_hi[2] = nb[0].hi.z;

    // primitive records in leaf order
    out->prims.resize(N);
    for (uint32_t k = 0; k < N; ++k) {
        uint32_t id = ord[k];
        PrimRec& r = out->prims[k];
        memset(&r, 0, sizeof(r));
        r.w[10] = bits_f(id);
        if (id < nfaces) {
            f3 p1 = V(idx[3 * id + 2]), p2 = V(idx[3 * id + 1]), p3 = V(idx[3 * id + 0]);
            f3 e1 = sub3(p2, p1), e2 = sub3(p3, p1);
            r.w[0] = p1.x; r.w[1] = p1.y; r.w[2] = p1.z;
            r.w[3] = e1.x; r.w[4] = e1.y; r.w[5] = e1.z;
            r.w[6] = e2.x; r.w[7] = e2.y; r.w[8] = e2.z;
            r.w[11] = bits_f(0u);
        } else {
            r.w[9] = bits_f(id - nfaces);
            r.w[11] = bits_f(1u);
        }
    }

    // the last record of every leaf carries bit 8 of the kind word
    for (size_t i = 0; i < nodes.size(); ++i) {
        if ((nodes[i].meta & 3u) != 3u) continue;
        uint32_t cnt = nodes[i].meta >> 2;
        PrimRec& r = out->prims[nodes[i].off + cnt - 1];
        uint32_t kind;
        memcpy(&kind, &r.w[11], 4);
        r.w[11] = bits_f(kind | 0x100u);
    }

    // four-wide layout: records in depth-first pre-order of the record tree
    auto is_leaf = [&](size_t i) { return (nodes[i].meta & 3u) == 3u; };
    auto slots_of = [&](size_t a, size_t s[4]) {
        const size_t b = a + 1, c = nodes[a].off;
        const size_t none = (size_t)-1;
        if (is_leaf(b)) { s[0] = b; s[1] = none; } else { s[0] = b + 1; s[1] = nodes[b].off; }
        if (is_leaf(c)) { s[2] = c; s[3] = none; } else { s[2] = c + 1; s[3] = nodes[c].off; }
    };
    std::vector<int32_t> qidx(nodes.size(), -1);
    std::vector<size_t> order_q;
    uint32_t qdepth = 0;
    if (!is_leaf(0)) {
        std::vector<std::pair<size_t, uint32_t>> stk;
        stk.push_back({0, 1});
        while (!stk.empty()) {
            auto [a, dq] = stk.back();
            stk.pop_back();
            qidx[a] = (int32_t)order_q.size();
            order_q.push_back(a);
            qdepth = std::max(qdepth, dq);
            size_t s[4];
            slots_of(a, s);
            for (int k = 3; k >= 0; --k)
                if (s[k] != (size_t)-1 && !is_leaf(s[k])) stk.push_back({s[k], dq + 1});
        }
    }
    out->qnodes.assign(order_q.size(), QNode{});
    for (size_t qi = 0; qi < order_q.size(); ++qi) {
        const size_t a = order_q[qi];
        QNode& q = out->qnodes[qi];
        size_t s[4];
        slots_of(a, s);
        const uint32_t ab = is_leaf(a + 1) ? 0u : (nodes[a + 1].meta & 3u);
        const uint32_t ac = is_leaf(nodes[a].off) ? 0u : (nodes[nodes[a].off].meta & 3u);
        q.meta = (nodes[a].meta & 3u) | (ab << 2) | (ac << 4);
        for (int k = 0; k < 4; ++k) {
            if (s[k] == (size_t)-1) {
                q.ref[k] = QEMPTY;
                q.lx[k] = q.ly[k] = q.lz[k] = FLT_MAX;
                q.hx[k] = q.hy[k] = q.hz[k] = -FLT_MAX;
                continue;
            }
            const RNode& c = nodes[s[k]];
            q.ref[k] = is_leaf(s[k]) ? ~(int32_t)c.off : qidx[s[k]];
            q.lx[k] = c.lo[0]; q.ly[k] = c.lo[1]; q.lz[k] = c.lo[2];
            q.hx[k] = c.hi[0]; q.hy[k] = c.hi[1]; q.hz[k] = c.hi[2];
        }
    }
    out->qroot = is_leaf(0) ? ~(int32_t)nodes[0].off : 0;
    out->qdepth = qdepth;
    if (!compress_layout(nodes, out)) {
        out->cnodes.clear();
        out->prims64.clear();
    }
    auto t1 = std::chrono::steady_clock::now();
    out->build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return true;
}

}  // namespace px
