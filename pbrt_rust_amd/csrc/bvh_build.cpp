// bvh_build.cpp -- host-side BVH construction (BVHAccelerator::new,
// primitive/aggregates/bvh.rs:105-362) producing the reference's tree exactly,
// written as an iterative pre-order build over one index permutation with
// in-place partitions (no per-node allocation), then packed into the two
// GPU layouts (32-byte reference nodes and 64-byte child-pair nodes).
#include "bvh_build.hpp"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <atomic>
#include <chrono>
#include <thread>

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

// Runs f(k, begin, end) for the chunks k = 0..T-1 of [0, n) (contiguous,
// equal but the last) on T threads, the calling thread taking chunk 0;
// T = 1 below 64K items.  Chunk boundaries depend only on (n, threads).
unsigned g_threads() {
    static const unsigned t = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    return t;
}
size_t chunk_count(size_t n, unsigned threads) {
    if (threads <= 1 || n < (1u << 16)) return 1;
    return std::min<size_t>(threads, n / 16384);
}
template <typename F>
void parallel_chunks(size_t n, unsigned threads, F f) {
    const size_t T = chunk_count(n, threads);
    if (T == 1) { f((size_t)0, (size_t)0, n); return; }
    const size_t per = (n + T - 1) / T;
    std::vector<std::thread> th;
    for (size_t k = 1; k < T; ++k)
        th.emplace_back([&f, k, per, n] { f(k, std::min(n, k * per), std::min(n, (k + 1) * per)); });
    f((size_t)0, (size_t)0, std::min(n, per));
    for (auto& t : th) t.join();
}
template <typename F>
void parallel_for(size_t n, F f) {
    parallel_chunks(n, g_threads(), [&f](size_t, size_t b, size_t e) { f(b, e); });
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
bool cq_axis(float ulo, float uhi, const float* lo, const float* hi, const bool* used, uint8_t* qlo,
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
// Quantize four-wide records into compressed ones; false when some
// record's bounds are not representable.
static bool compress_records(const std::vector<QNode>& qn, std::vector<CNode>* cn) {
    cn->assign(qn.size(), CNode{});
    std::atomic<bool> ok{true};
    parallel_for(qn.size(), [&](size_t i0, size_t i1) {
    for (size_t i = i0; i < i1 && ok.load(std::memory_order_relaxed); ++i) {
        const QNode& q = qn[i];
        CNode& c = (*cn)[i];
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
        if (!cq_axis(u[0], u[3], q.lx, q.hx, used, c.q + 0, c.q + 4, &ex) ||
            !cq_axis(u[1], u[4], q.ly, q.hy, used, c.q + 8, c.q + 12, &ey) ||
            !cq_axis(u[2], u[5], q.lz, q.hz, used, c.q + 16, c.q + 20, &ez)) {
            ok.store(false);
            break;
        }
        c.ox = u[0]; c.oy = u[1]; c.oz = u[2];
        c.emeta = ex | (ey << 8) | (ez << 16) | ((q.meta & 63u) << 24);
    }
    });
    return ok.load();
}

bool compress_layout(const std::vector<RNode>& nodes, HostBvh* out) {
    if (!compress_records(out->qnodes, &out->cnodes)) return false;
    // prim records with their leaf's exact box
    out->prims64.assign(out->prims.size(), PrimRec64{});
    parallel_for(nodes.size(), [&](size_t i0, size_t i1) {
    for (size_t i = i0; i < i1; ++i) {
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
    });
    return true;
}

// Occlusion records: a second four-wide record tree over the same leaves,
// for queries whose answer does not depend on the order in which leaves are
// met.  The ordered records take a node's grandchildren, so a node with two
// leaf children makes a record of two slots; here each record takes the cut
// of its binary subtree (at most four nodes, each a leaf or the root of a
// further record) that minimizes the summed surface area of all records
// below it -- the expected number of record tests for rays spread evenly
// over position and direction -- plus a small charge per record that breaks
// ties towards fewer records.  Dynamic programming over the binary tree:
// D[i][k-1] is the least cost of covering node i's subtree with at most k
// slots, either i whole (a leaf, or a record of its own) or split between
// its two children.  Leaves cost nothing (every cut holds each leaf once).
// Empty when the tree's root is a leaf or a record is not representable.
// A binary tree over the leaves of the reference tree for the occlusion
// records.  An occlusion query's answer depends only on the leaves: a leaf
// whose exact box passes the slab test has every enclosing box pass as well
// (the rounded slab distances are monotone in the bounds), so any tree over
// the same leaves, with each interior box the union of its leaves' boxes,
// reaches exactly the leaves the reference's tree reaches.  This one is
// built for fewer box tests: binned SAH over the leaf centroids on all three
// axes (the reference bins only the widest axis), OCC_BUCKETS buckets, down
// to single leaves.  On the 1.31M-triangle displaced icosphere scene its
// four-wide records cost 1.65 against 2.50 (summed record area over the
// root's) and a sample of shadow rays tests 6.8 records instead of 8.0.
// Buckets: on the 1024x1024 direct-lighting frame's occlusion launch, 6 /
// 12 / 32 gave 3.19-3.20 / 3.18-3.22 / 3.13-3.14 ms (paired runs).
// Leaves are the reference's nodes, copied; the layout is the reference's
// (depth first, an interior node's second child at `off`).
constexpr int OCC_BUCKETS = 32;
struct OccItem {
    float lo[3], hi[3], c[3];
    uint32_t leaf;  // the reference node
};
static void occ_build_range(std::vector<OccItem>& it, size_t b0, size_t e0, const std::vector<RNode>& ref,
                            size_t task_min, std::vector<RNode>* out, std::vector<std::pair<size_t, size_t>>* tasks) {
    struct Job { size_t b, e, parent; };  // parent: node whose off is this job's index (SIZE_MAX: none)
    std::vector<Job> st;
    st.push_back({b0, e0, SIZE_MAX});
    while (!st.empty()) {
        const Job j = st.back();
        st.pop_back();
        const size_t me = out->size();
        if (j.parent != SIZE_MAX) (*out)[j.parent].off = (uint32_t)me;
        if (j.e - j.b == 1) {
            out->push_back(ref[it[j.b].leaf]);
            continue;
        }
        if (tasks && j.e - j.b <= task_min) {
            // a subtree built on its own: a placeholder naming the task
            RNode ph{};
            ph.meta = 0xFFFFFFFCu;
            ph.off = (uint32_t)tasks->size();
            tasks->push_back({j.b, j.e});
            out->push_back(ph);
            continue;
        }
        float blo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (size_t i = j.b; i < j.e; ++i)
            for (int a = 0; a < 3; ++a) {
                blo[a] = std::min(blo[a], it[i].lo[a]); bhi[a] = std::max(bhi[a], it[i].hi[a]);
                clo[a] = std::min(clo[a], it[i].c[a]); chi[a] = std::max(chi[a], it[i].c[a]);
            }
        RNode nd{};
        for (int a = 0; a < 3; ++a) { nd.lo[a] = blo[a]; nd.hi[a] = bhi[a]; }
        nd.meta = 0;
        out->push_back(nd);
        auto bucket = [&](const OccItem& x, int a) {
            const float f = (x.c[a] - clo[a]) / (chi[a] - clo[a]) * (float)OCC_BUCKETS;
            return f >= (float)(OCC_BUCKETS - 1) ? OCC_BUCKETS - 1 : (f > 0.0f ? (int)f : 0);
        };
        double best = INFINITY;
        int ba = -1, bs = 0;
        for (int a = 0; a < 3; ++a) {
            if (!(chi[a] > clo[a])) continue;
            float lo[OCC_BUCKETS][3], hi[OCC_BUCKETS][3];
            size_t cn[OCC_BUCKETS] = {};
            for (int k = 0; k < OCC_BUCKETS; ++k)
                for (int d = 0; d < 3; ++d) { lo[k][d] = FLT_MAX; hi[k][d] = -FLT_MAX; }
            for (size_t i = j.b; i < j.e; ++i) {
                const int k = bucket(it[i], a);
                ++cn[k];
                for (int d = 0; d < 3; ++d) { lo[k][d] = std::min(lo[k][d], it[i].lo[d]); hi[k][d] = std::max(hi[k][d], it[i].hi[d]); }
            }
            auto area = [](const float* l, const float* h) {
                const double x = (double)h[0] - l[0], y = (double)h[1] - l[1], z = (double)h[2] - l[2];
                return x < 0.0 ? 0.0 : 2.0 * (x * y + y * z + z * x);
            };
            double ra[OCC_BUCKETS];
            size_t rc[OCC_BUCKETS];
            float al[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, ah[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            size_t c = 0;
            for (int k = OCC_BUCKETS - 1; k >= 0; --k) {
                for (int d = 0; d < 3; ++d) { al[d] = std::min(al[d], lo[k][d]); ah[d] = std::max(ah[d], hi[k][d]); }
                c += cn[k];
                ra[k] = area(al, ah);
                rc[k] = c;
            }
            for (int d = 0; d < 3; ++d) { al[d] = FLT_MAX; ah[d] = -FLT_MAX; }
            c = 0;
            for (int k = 0; k + 1 < OCC_BUCKETS; ++k) {
                for (int d = 0; d < 3; ++d) { al[d] = std::min(al[d], lo[k][d]); ah[d] = std::max(ah[d], hi[k][d]); }
                c += cn[k];
                if (c == 0 || rc[k + 1] == 0) continue;
                const double v = area(al, ah) * (double)c + ra[k + 1] * (double)rc[k + 1];
                if (v < best) { best = v; ba = a; bs = k; }
            }
        }
        size_t mid = (j.b + j.e) / 2;  // every centroid equal: any halves
        if (ba >= 0) {
            auto pm = std::partition(it.begin() + (std::ptrdiff_t)j.b, it.begin() + (std::ptrdiff_t)j.e,
                                     [&](const OccItem& x) { return bucket(x, ba) <= bs; });
            mid = (size_t)(pm - it.begin());
        }
        st.push_back({mid, j.e, me});  // second child, taken after the first's subtree
        st.push_back({j.b, mid, SIZE_MAX});
    }
}
static std::vector<RNode> occlusion_tree(const std::vector<RNode>& nodes) {
    std::vector<OccItem> it;
    for (size_t i = 0; i < nodes.size(); ++i)
        if ((nodes[i].meta & 3u) == 3u) {
            OccItem x;
            for (int a = 0; a < 3; ++a) {
                x.lo[a] = nodes[i].lo[a];
                x.hi[a] = nodes[i].hi[a];
                x.c[a] = 0.5f * nodes[i].lo[a] + 0.5f * nodes[i].hi[a];
            }
            x.leaf = (uint32_t)i;
            it.push_back(x);
        }
    // the top of the tree on this thread, its subtrees of at most task_min
    // leaves on all threads, then spliced in depth-first order
    const unsigned T = g_threads();
    const size_t task_min = std::max<size_t>(4096, it.size() / (8 * (size_t)T) + 1);
    std::vector<RNode> top;
    std::vector<std::pair<size_t, size_t>> tasks;
    top.reserve(2 * it.size() / task_min + 16);
    occ_build_range(it, 0, it.size(), nodes, task_min, &top, &tasks);
    std::vector<std::vector<RNode>> sub(tasks.size());
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t k; (k = next.fetch_add(1)) < tasks.size();) {
            sub[k].reserve(2 * (tasks[k].second - tasks[k].first));
            occ_build_range(it, tasks[k].first, tasks[k].second, nodes, 0, &sub[k], nullptr);
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < T && t < tasks.size(); ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    std::vector<RNode> out;
    out.reserve(2 * it.size());
    std::vector<uint32_t> pos(top.size());
    // top nodes in order; a placeholder becomes its subtree, whose interior
    // offsets move with it; a top node's second child is found through pos
    for (size_t i = 0; i < top.size(); ++i) {
        pos[i] = (uint32_t)out.size();
        if (top[i].meta == 0xFFFFFFFCu) {
            const std::vector<RNode>& s = sub[top[i].off];
            const uint32_t base = (uint32_t)out.size();
            for (RNode r : s) {
                if ((r.meta & 3u) != 3u) r.off += base;
                out.push_back(r);
            }
        } else {
            out.push_back(top[i]);
        }
    }
    for (size_t i = 0; i < top.size(); ++i)
        if (top[i].meta != 0xFFFFFFFCu && (top[i].meta & 3u) != 3u) out[pos[i]].off = pos[top[i].off];
    return out;
}

static void occlusion_records(const std::vector<RNode>& ref_nodes, HostBvh* out) {
    out->ocnodes.clear();
    out->odepth = 0;
    if (ref_nodes.empty() || (ref_nodes[0].meta & 3u) == 3u) return;
    const std::vector<RNode> nodes = occlusion_tree(ref_nodes);
    const size_t N = nodes.size();
    auto is_leaf = [&](size_t i) { return (nodes[i].meta & 3u) == 3u; };
    if (N == 0 || is_leaf(0)) return;
    auto area = [&](size_t i) {
        const double dx = (double)nodes[i].hi[0] - nodes[i].lo[0], dy = (double)nodes[i].hi[1] - nodes[i].lo[1],
                     dz = (double)nodes[i].hi[2] - nodes[i].lo[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    };
    const double ra = area(0);
    const double inv = ra > 0.0 && std::isfinite(ra) ? 1.0 / ra : 0.0;
    const double REC = 1e-6;  // per-record charge (relative to the root's area)
    std::vector<std::array<float, 4>> D(N);
    // choice bits: [2k-4, 2k-2) for k = 2..4 the left share j of a split into
    // at most k slots (0: whole); [6, 8) the left share of the record split
    std::vector<uint8_t> ch(N, 0);
    // children follow their parent in the depth-first layout: reverse index
    // order meets every child before its parent
    for (size_t ii = N; ii-- > 0;) {
        if (is_leaf(ii)) { D[ii] = {0.0f, 0.0f, 0.0f, 0.0f}; continue; }
        const size_t b = ii + 1, c = nodes[ii].off;
        float S[5];
        uint32_t js[5] = {0, 0, 0, 0, 0};
        for (int k = 2; k <= 4; ++k) {
            S[k] = INFINITY;
            for (int j = 1; j < k; ++j) {
                const float v = D[b][j - 1] + D[c][k - j - 1];
                if (v < S[k]) { S[k] = v; js[k] = (uint32_t)j; }
            }
        }
        double a = area(ii) * inv;
        if (!std::isfinite(a)) a = 1.0;
        const float R = (float)(a + REC) + S[4];
        uint8_t bits = (uint8_t)(js[4] << 6);
        D[ii][0] = R;
        for (int k = 2; k <= 4; ++k) {
            if (S[k] < R) { D[ii][k - 1] = S[k]; bits |= (uint8_t)(js[k] << (2 * k - 4)); }
            else D[ii][k - 1] = R;
        }
        ch[ii] = bits;
    }
    std::vector<std::array<float, 4>>().swap(D);
    // the slots of the record rooted at interior node a
    auto slots_of = [&](size_t a, size_t s[4]) {
        int n = 0;
        struct Item { size_t i; uint32_t k; };
        Item st[8];
        int sp = 0;
        const uint32_t j = (ch[a] >> 6) & 3u;
        st[sp++] = {nodes[a].off, 4u - j};
        st[sp++] = {a + 1, j};
        while (sp > 0) {
            const Item it = st[--sp];
            const uint32_t jj = (it.k >= 2 && !is_leaf(it.i)) ? (ch[it.i] >> (2 * it.k - 4)) & 3u : 0u;
            if (jj == 0) { s[n++] = it.i; continue; }
            st[sp++] = {nodes[it.i].off, it.k - jj};
            st[sp++] = {it.i + 1, jj};
        }
        for (int k = n; k < 4; ++k) s[k] = (size_t)-1;
    };
    // records in depth-first pre-order of the record tree
    std::vector<int32_t> rec(N, -1);
    std::vector<size_t> roots;
    {
        std::vector<std::pair<size_t, uint32_t>> stk;
        stk.push_back({0, 1});
        while (!stk.empty()) {
            auto [a, dq] = stk.back();
            stk.pop_back();
            rec[a] = (int32_t)roots.size();
            roots.push_back(a);
            out->odepth = std::max(out->odepth, dq);
            size_t s[4];
            slots_of(a, s);
            for (int k = 3; k >= 0; --k)
                if (s[k] != (size_t)-1 && !is_leaf(s[k])) stk.push_back({s[k], dq + 1});
        }
    }
    std::vector<QNode> oq(roots.size(), QNode{});
    parallel_for(roots.size(), [&](size_t r0, size_t r1) {
    for (size_t r = r0; r < r1; ++r) {
        QNode& q = oq[r];
        size_t s[4];
        slots_of(roots[r], s);
        q.meta = 0;
        for (int k = 0; k < 4; ++k) {
            if (s[k] == (size_t)-1) {
                q.ref[k] = QEMPTY;
                q.lx[k] = q.ly[k] = q.lz[k] = FLT_MAX;
                q.hx[k] = q.hy[k] = q.hz[k] = -FLT_MAX;
                continue;
            }
            const RNode& c = nodes[s[k]];
            q.ref[k] = is_leaf(s[k]) ? ~(int32_t)c.off : rec[s[k]];
            q.lx[k] = c.lo[0]; q.ly[k] = c.lo[1]; q.lz[k] = c.lo[2];
            q.hx[k] = c.hi[0]; q.hy[k] = c.hi[1]; q.hz[k] = c.hi[2];
        }
    }
    });
    if (!compress_records(oq, &out->ocnodes)) {
        out->ocnodes.clear();
        out->odepth = 0;
    }
}

namespace {

constexpr uint32_t UNSET = 0xFFFFFFFFu;

// The permutation being partitioned, with each prim's bounds and centroid
// carried along (sob[i] = bounds of prim ord[i], soc[i] its centroid), so the
// passes over a node's range read memory in order; tmp* are scratch with the
// same indexing.
struct BuildCtx {
    const Box* pb;                // prim bounds, by prim id
    const std::vector<f3>* pc;    // prim centroids, by prim id
    uint32_t* ord;
    Box* sob;
    f3* soc;
    uint32_t* tmp;
    Box* tmpb;
    f3* tmpc;
    int split;
    uint32_t max_prims;
    unsigned threads;  // for the passes over one node's range
};

// Stable partition of positions [b, b+n) by keep(i) (ord, bounds and
// centroids together); returns the size of the kept front.
template <typename Keep>
size_t stable_partition3(const BuildCtx& c, uint32_t b, uint32_t n, Keep keep);

// The same partition over T chunks: kept items in order, then the others in
// order, through the scratch arrays.
template <typename Keep>
size_t stable_partition3_par(const BuildCtx& c, uint32_t b, uint32_t n, Keep keep) {
    const size_t T = chunk_count(n, c.threads);
    if (T == 1) return stable_partition3(c, b, n, keep);
    std::vector<size_t> kept(T, 0);
    parallel_chunks(n, c.threads, [&](size_t k, size_t i0, size_t i1) {
        size_t m = 0;
        for (size_t i = i0; i < i1; ++i) m += keep((uint32_t)i) ? 1 : 0;
        kept[k] = m;
    });
    std::vector<size_t> lo(T), hi(T);
    size_t tk = 0;
    for (size_t k = 0; k < T; ++k) { lo[k] = tk; tk += kept[k]; }
    const size_t per = (n + T - 1) / T;
    for (size_t k = 0; k < T; ++k) hi[k] = tk + std::min<size_t>(n, k * per) - lo[k];
    uint32_t* a = c.ord + b;
    Box* ab = c.sob + b;
    f3* ac = c.soc + b;
    uint32_t* t = c.tmp + b;
    Box* tb = c.tmpb + b;
    f3* tc = c.tmpc + b;
    parallel_chunks(n, c.threads, [&](size_t k, size_t i0, size_t i1) {
        size_t l = lo[k], r = hi[k];
        for (size_t i = i0; i < i1; ++i) {
            const size_t d = keep((uint32_t)i) ? l++ : r++;
            t[d] = a[i]; tb[d] = ab[i]; tc[d] = ac[i];
        }
    });
    parallel_chunks(n, c.threads, [&](size_t, size_t i0, size_t i1) {
        memcpy(a + i0, t + i0, (i1 - i0) * sizeof(uint32_t));
        memcpy(ab + i0, tb + i0, (i1 - i0) * sizeof(Box));
        memcpy(ac + i0, tc + i0, (i1 - i0) * sizeof(f3));
    });
    return tk;
}

template <typename Keep>
size_t stable_partition3(const BuildCtx& c, uint32_t b, uint32_t n, Keep keep) {
    uint32_t* a = c.ord + b;
    Box* ab = c.sob + b;
    f3* ac = c.soc + b;
    uint32_t* t = c.tmp + b;
    Box* tb = c.tmpb + b;
    f3* tc = c.tmpc + b;
    size_t l = 0, r = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (keep(i)) {
            a[l] = a[i]; ab[l] = ab[i]; ac[l] = ac[i];
            ++l;
        } else {
            t[r] = a[i]; tb[r] = ab[i]; tc[r] = ac[i];
            ++r;
        }
    }
    if (r) {
        memcpy(a + l, t, r * sizeof(uint32_t));
        memcpy(ab + l, tb, r * sizeof(Box));
        memcpy(ac + l, tc, r * sizeof(f3));
    }
    return l;
}

// A pre-order run of the tree: nodes with ids local to the piece (an
// interior node's off is its second child's local id, a leaf's off the
// global position of its first prim) and their boxes (leaves only; interior
// boxes are folded after splicing).
struct Piece {
    struct Def {
        uint32_t at;     // local id the subtree's first node would have taken
        uint32_t b, e, depth;
        int64_t parent;  // local id of the parent when it is a second child, else -1
    };
    std::vector<RNode> nodes;
    std::vector<Box> nb;
    std::vector<Def> deferred;  // set-aside subtrees, in pre-order
    uint32_t maxdepth = 0;
    bool ok = true;
    std::string err;
};

// recursive_build (bvh.rs:189-258) of the subtree over ord[b, e) at `depth`,
// iteratively in pre-order.  Subtrees (other than the piece's root) of
// 4096..defer_max prims are not built but listed in out->deferred.
bool build_range(const BuildCtx& c, uint32_t b0, uint32_t e0, uint32_t depth0, uint32_t defer_max, Piece* out) {
    const std::vector<f3>& pc = *c.pc;
    uint32_t* ord = c.ord;
    const Box* sob = c.sob;
    const f3* soc = c.soc;
    std::vector<RNode>& nodes = out->nodes;
    std::vector<Box>& nb = out->nb;
    struct Task {
        uint32_t b, e;
        int64_t parent;
        uint32_t depth;
    };
    std::vector<Task> stack;
    stack.push_back(Task{b0, e0, -1, depth0});
    std::vector<uint8_t> bk;  // bucket per position of the current range
    bool root = true;
    while (!stack.empty()) {
        Task T = stack.back();
        stack.pop_back();
        const uint32_t b = T.b, e = T.e, n = e - b;
        if (!root && n >= 4096 && n <= defer_max) {
            out->deferred.push_back(Piece::Def{(uint32_t)nodes.size(), b, e, T.depth, T.parent});
            continue;
        }
        root = false;
        const uint32_t id = (uint32_t)nodes.size();
        nodes.push_back(RNode{});
        nodes.back().off = UNSET;
        nb.push_back(Box{});
        if (T.parent >= 0) nodes[(size_t)T.parent].off = id;
        out->maxdepth = std::max(out->maxdepth, T.depth);
        // recursive_build bbox: bvh.rs:192-194 (fold in the range's order)
        // min / max folds: any grouping gives the same box
        const size_t NC = chunk_count(n, c.threads);
        Box bb = box_empty();
        if (NC == 1) {
            for (uint32_t i = b; i < e; ++i) bb = box_union(bb, sob[i]);
        } else {
            std::vector<Box> part(NC, box_empty());
            parallel_chunks(n, c.threads, [&](size_t k, size_t i0, size_t i1) {
                Box x = box_empty();
                for (size_t i = b + i0; i < b + i1; ++i) x = box_union(x, sob[i]);
                part[k] = x;
            });
            for (auto& x : part) bb = box_union(bb, x);
        }
        auto make_leaf = [&]() {
            nodes[id].off = b;
            nodes[id].meta = 3u | (n << 2);
            nb[id] = bb;
        };
        if (n == 1) { make_leaf(); continue; }
        Box cb = box_empty();
        if (NC == 1) {
            for (uint32_t i = b; i < e; ++i) cb = box_union_p(cb, soc[i]);
        } else {
            std::vector<Box> part(NC, box_empty());
            parallel_chunks(n, c.threads, [&](size_t k, size_t i0, size_t i1) {
                Box x = box_empty();
                for (size_t i = b + i0; i < b + i1; ++i) x = box_union_p(x, soc[i]);
                part[k] = x;
            });
            for (auto& x : part) cb = box_union(cb, x);
        }
        const int dim = max_extent(cb);
        const float cmin = comp(cb.lo, dim), cmax = comp(cb.hi, dim);
        if (cmin == cmax) { make_leaf(); continue; }
        uint32_t mid;
        bool sah_leaf = false;
        if (c.split == 1) {
            // split_middle: bvh.rs:86-90
            float pmid = 0.5f * (cmin + cmax);
            mid = b + (uint32_t)stable_partition3_par(c, b, n, [&](uint32_t i) { return comp(soc[b + i], dim) < pmid; });
        } else if (c.split == 2 || n <= 4) {
            partition_by(ord + b, n, pc, dim);
            for (uint32_t i = b; i < e; ++i) {  // carry bounds and centroids along
                c.sob[i] = c.pb[ord[i]];
                c.soc[i] = pc[ord[i]];
            }
            mid = b + n / 2;
        } else {
            // split_surface_area_heuristic: bvh.rs:105-186
            uint32_t cnt[12] = {0};
            Box bbk[12];
            for (int k = 0; k < 12; ++k) bbk[k] = box_empty();
            bk.resize(n);
            // bucket counts and boxes per chunk, merged (exact in any grouping)
            std::vector<uint32_t> pcnt(NC * 12, 0);
            std::vector<Box> pbk(NC * 12, box_empty());
            parallel_chunks(n, c.threads, [&](size_t ck, size_t i0, size_t i1) {
                uint32_t* cn = pcnt.data() + 12 * ck;
                Box* bx = pbk.data() + 12 * ck;
                for (size_t i = i0; i < i1; ++i) {
                    int k = bucket_for(comp(soc[b + i], dim), cmin, cmax);
                    bk[i] = (uint8_t)k;
                    cn[k] += 1;
                    bx[k] = box_union(bx[k], sob[b + i]);
                }
            });
            for (size_t ck = 0; ck < NC; ++ck)
                for (int k = 0; k < 12; ++k) {
                    cnt[k] += pcnt[12 * ck + k];
                    bbk[k] = box_union(bbk[k], pbk[12 * ck + k]);
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
            if (c.max_prims < n || mcu < n) {
                // stable partition on the precomputed buckets
                mid = b + (uint32_t)stable_partition3_par(c, b, n, [&](uint32_t i) { return (int)bk[i] <= best; });
            } else {
                sah_leaf = true;
                mid = b;
            }
        }
        if (sah_leaf) { make_leaf(); continue; }
        if (mid == b || mid == e) {
            out->ok = false;
            out->err = "BVH split produced an empty side";  // bvh.rs:237-238 asserts
            return false;
        }
        nodes[id].meta = (uint32_t)dim;
        stack.push_back(Task{mid, e, (int64_t)id, T.depth + 1});
        stack.push_back(Task{b, mid, -1, T.depth + 1});
    }
    return true;
}

}  // namespace

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
    parallel_for(nfaces, [&](size_t f0, size_t f1) {
        for (size_t f = f0; f < f1; ++f) {
            Box b = box_empty();
            b = box_union_p(b, V(idx[3 * f + 2]));
            b = box_union_p(b, V(idx[3 * f + 1]));
            b = box_union_p(b, V(idx[3 * f + 0]));
            pb[f] = b;
        }
    });
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
    parallel_for(N, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; ++i) pc[i] = mul3(add3(pb[i].lo, pb[i].hi), 0.5f);  // bvh.rs:77
    });

    std::vector<uint32_t> ord(N), tmp(N);
    for (uint32_t i = 0; i < N; ++i) ord[i] = i;
    std::vector<Box> sob(pb), tmpb(N);
    std::vector<f3> soc(pc), tmpc(N);
    const unsigned hw = g_threads();
    const BuildCtx ctx{pb.data(), &pc, ord.data(), sob.data(), soc.data(), tmp.data(), tmpb.data(), tmpc.data(),
                       split, max_prims, hw};
    BuildCtx ctx1 = ctx;  // the set-aside subtrees: one thread each
    ctx1.threads = 1;

    // The top of the tree is built here, subtrees of a middling size are
    // set aside and built on worker threads (they touch disjoint ranges of
    // ord and tmp, and each is the same sequence of operations as in a
    // one-thread build), then all is spliced together in pre-order.
    const uint32_t defer_max = (N < 200000 || hw == 1) ? 0u : std::max<uint32_t>(N / (8 * hw), 50000u);
    Piece top;
    if (!build_range(ctx, 0, N, 0, defer_max, &top)) { *err = top.err; return false; }
    std::vector<Piece> pieces(top.deferred.size());
    {
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t k; (k = next.fetch_add(1)) < pieces.size();) {
                const Piece::Def& d = top.deferred[k];
                Piece& pc_ = pieces[k];
                if (!build_range(ctx1, d.b, d.e, d.depth, 0u, &pc_)) continue;
                // interior boxes of the piece (children follow parents)
                for (size_t i = pc_.nodes.size(); i-- > 0;)
                    if ((pc_.nodes[i].meta & 3u) != 3u) pc_.nb[i] = box_union(pc_.nb[i + 1], pc_.nb[pc_.nodes[i].off]);
            }
        };
        std::vector<std::thread> th;
        const unsigned nt = (unsigned)std::min<size_t>(hw, pieces.size());
        for (unsigned t = 1; t < nt; ++t) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
    }
    for (auto& pc_ : pieces)
        if (!pc_.ok) { *err = pc_.err; return false; }
    // final ids: top node i goes after the pieces deferred before it
    const size_t ntop = top.nodes.size();
    std::vector<uint32_t> top_id(ntop + 1);
    std::vector<uint32_t> piece_base(pieces.size());
    {
        uint32_t id = 0;
        size_t k = 0;
        for (size_t i = 0; i <= ntop; ++i) {
            for (; k < pieces.size() && top.deferred[k].at == i; ++k) {
                piece_base[k] = id;
                id += (uint32_t)pieces[k].nodes.size();
            }
            top_id[i] = id;
            if (i < ntop) ++id;
        }
    }
    std::vector<RNode>& nodes = out->rnodes;
    std::vector<Box> nb;
    {
        size_t total = ntop;
        for (auto& pc_ : pieces) total += pc_.nodes.size();
        nodes.assign(total, RNode{});
        nb.assign(total, Box{});
    }
    uint32_t maxdepth = top.maxdepth;
    auto set_box = [](RNode& r, const Box& x) {
        r.lo[0] = x.lo.x; r.lo[1] = x.lo.y; r.lo[2] = x.lo.z;
        r.hi[0] = x.hi.x; r.hi[1] = x.hi.y; r.hi[2] = x.hi.z;
    };
    for (size_t i = 0; i < ntop; ++i) {
        RNode r = top.nodes[i];
        if ((r.meta & 3u) != 3u && r.off != UNSET) r.off = top_id[r.off];
        nodes[top_id[i]] = r;
        nb[top_id[i]] = top.nb[i];
    }
    // the pieces (boxes complete) at their final ids
    {
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t k; (k = next.fetch_add(1)) < pieces.size();) {
                const Piece& pc_ = pieces[k];
                const uint32_t base = piece_base[k];
                for (size_t i = 0; i < pc_.nodes.size(); ++i) {
                    RNode r = pc_.nodes[i];
                    if ((r.meta & 3u) != 3u) r.off += base;
                    set_box(r, pc_.nb[i]);
                    nodes[base + i] = r;
                    nb[base + i] = pc_.nb[i];
                }
            }
        };
        std::vector<std::thread> th;
        const unsigned nt = (unsigned)std::min<size_t>(hw, pieces.size());
        for (unsigned t = 1; t < nt; ++t) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
    }
    for (size_t k = 0; k < pieces.size(); ++k) {
        // a deferred second child: its parent's second-child id
        if (top.deferred[k].parent >= 0) nodes[top_id[(size_t)top.deferred[k].parent]].off = piece_base[k];
        maxdepth = std::max(maxdepth, pieces[k].maxdepth);
    }
    pieces.clear();
    // interior bounds of the top = left ∪ right (bvh.rs:247), children first
    for (size_t i = ntop; i-- > 0;) {
        const size_t f = top_id[i];
        if ((nodes[f].meta & 3u) != 3u) nb[f] = box_union(nb[f + 1], nb[nodes[f].off]);
        set_box(nodes[f], nb[f]);
    }
    out->depth = maxdepth;
    out->order = ord;
    out->nfaces = nfaces;
    out->nspheres = nsph;
    out->root_lo[0] = nb[0].lo.x; out->root_lo[1] = nb[0].lo.y; out->root_lo[2] = nb[0].lo.z;
    out->root_hi[0] = nb[0].hi.x; out->root_hi[1] = nb[0].hi.y; out->root_hi[2] = nb[0].hi.z;

    // primitive records in leaf order
    out->prims.resize(N);
    parallel_for(N, [&](size_t k0, size_t k1) {
    for (size_t k = k0; k < k1; ++k) {
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
    });

    // the last record of every leaf carries bit 8 of the kind word
    parallel_for(nodes.size(), [&](size_t i0, size_t i1) {
    for (size_t i = i0; i < i1; ++i) {
        if ((nodes[i].meta & 3u) != 3u) continue;
        uint32_t cnt = nodes[i].meta >> 2;
        PrimRec& r = out->prims[nodes[i].off + cnt - 1];
        uint32_t kind;
        memcpy(&kind, &r.w[11], 4);
        r.w[11] = bits_f(kind | 0x100u);
    }
    });

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
    parallel_for(order_q.size(), [&](size_t q0, size_t q1) {
    for (size_t qi = q0; qi < q1; ++qi) {
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
    });
    out->qroot = is_leaf(0) ? ~(int32_t)nodes[0].off : 0;
    out->qdepth = qdepth;
    if (!compress_layout(nodes, out)) {
        out->cnodes.clear();
        out->prims64.clear();
    } else {
        occlusion_records(nodes, out);
    }
    auto t1 = std::chrono::steady_clock::now();
    out->build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return true;
}

}  // namespace px
