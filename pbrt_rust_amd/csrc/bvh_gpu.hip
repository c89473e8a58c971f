// bvh_gpu.hip -- a BVH built on the GPU (linear BVH), for scenes that must
// become resident fast (serving admission).  The reference builds its tree on
// the host (BVHAccelerator::new, bvh.rs:189-258, SAH); this builder makes a
// different tree over the same primitives, so a closest-hit query returns the
// reference's primitive, t and barycentrics except on rays where the
// reference's own answer depends on its visiting order: several primitives
// sharing the closest t (the later-visited one wins, bvh.rs:401-406 with
// mesh.rs:71), or a hit lying on its leaf box's rounded boundary, which the
// pop-time box test (bvh.rs:390-395) rejects once maxt has shrunk to that t
// and accepts before.  Occlusion bits are the reference's exactly (they do
// not depend on the tree).
//
// Pipeline, all on the device:
//   1. each primitive's box (the refined triple [c,b,a], mesh.rs:196-205;
//      spheres' boxes come from the host as build_bvh computes them) and
//      centroid (bvh.rs:77);
//   2. the centroids' bounds; 63-bit Morton codes (21 bits an axis);
//   3. radix sort of (code, prim) pairs (rocPRIM);
//   4. the binary radix tree over the sorted codes (T. Karras, "Maximizing
//      Parallelism in the Construction of BVHs, Octrees, and k-d Trees",
//      HPG 2012), duplicate codes split by their sorted position;
//   5. boxes and leaf counts bottom up (the second child to arrive at a node
//      finishes it);
//   6. depth-first positions top down, one level a launch (the reference
//      layout: first child at i + 1, second child at `off`);
//   7. the layouts the traversal kernels read: reference nodes, prim records
//      in leaf order, four-wide records (interior nodes at even depth, their
//      grandchildren as slots) and their compressed form.  Any-hit queries
//      walk the compressed records (there are no separate occlusion records).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "../../include/pbrtx_gpu.h"
#include "bvh_build.hpp"
#include "capi_internal.hpp"

namespace px {
namespace {

constexpr int GB = 256;
constexpr uint32_t LEAF_BIT = 0x80000000u;  // child reference: a leaf (sorted position)
constexpr int MAX_LEVELS = 100;              // tree depth <= 63 code bits + 32 position bits + 1

struct DBox {
    float lo[3], hi[3];
};

// 1. primitive boxes and centroids
__global__ __launch_bounds__(GB) void k_prim_boxes(const float* verts, const uint32_t* idx, uint32_t nfaces,
                                                   const DBox* sph_box, uint32_t nsph, DBox* box, float* cen) {
    const uint32_t n = nfaces + nsph;
    for (uint32_t i = blockIdx.x * GB + threadIdx.x; i < n; i += gridDim.x * GB) {
        DBox b;
        if (i < nfaces) {
            // Triangle::world_bound over [c, b, a] (mesh.rs:196-205): BBox::new
            // then unions (fmin / fmax, bbox.rs:155-168)
            b = DBox{{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
            for (int k = 2; k >= 0; --k) {
                const float* p = verts + 3 * (size_t)idx[3 * (size_t)i + k];
                for (int a = 0; a < 3; ++a) {
                    b.lo[a] = fminf(b.lo[a], p[a]);
                    b.hi[a] = fmaxf(b.hi[a], p[a]);
                }
            }
        } else {
            b = sph_box[i - nfaces];
        }
        box[i] = b;
        for (int a = 0; a < 3; ++a) cen[3 * (size_t)i + a] = (b.lo[a] + b.hi[a]) * 0.5f;  // bvh.rs:77
    }
}

// 2. centroid bounds: per-block partial min / max, then one block over them
__global__ __launch_bounds__(GB) void k_cen_bounds(const float* cen, uint32_t n, float* part) {
    __shared__ float s[6][GB];
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = blockIdx.x * GB + threadIdx.x; i < n; i += gridDim.x * GB)
        for (int a = 0; a < 3; ++a) {
            const float c = cen[3 * (size_t)i + a];
            lo[a] = fminf(lo[a], c);
            hi[a] = fmaxf(hi[a], c);
        }
    for (int a = 0; a < 3; ++a) { s[a][threadIdx.x] = lo[a]; s[3 + a][threadIdx.x] = hi[a]; }
    __syncthreads();
    for (int w = GB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int a = 0; a < 3; ++a) {
                s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + w]);
                s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}
__global__ __launch_bounds__(GB) void k_cen_bounds_final(const float* part, uint32_t nb, float* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float r[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t b = 0; b < nb; ++b)
        for (int a = 0; a < 3; ++a) {
            r[a] = fminf(r[a], part[6 * b + a]);
            r[3 + a] = fmaxf(r[3 + a], part[6 * b + 3 + a]);
        }
    for (int k = 0; k < 6; ++k) out[k] = r[k];
}

__device__ __forceinline__ uint64_t spread21(uint64_t v) {
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}
__global__ __launch_bounds__(GB) void k_morton(const float* cen, uint32_t n, const float* cb, uint64_t* key,
                                               uint32_t* val) {
    for (uint32_t i = blockIdx.x * GB + threadIdx.x; i < n; i += gridDim.x * GB) {
        uint64_t code = 0;
        for (int a = 0; a < 3; ++a) {
            const float ext = cb[3 + a] - cb[a];
            float u = ext > 0.0f ? (cen[3 * (size_t)i + a] - cb[a]) / ext : 0.0f;
            u = u == u ? fminf(fmaxf(u, 0.0f), 1.0f) : 0.0f;
            const uint64_t q = (uint64_t)fminf(u * 2097152.0f, 2097151.0f);
            code |= spread21(q) << (2 - a);
        }
        key[i] = code;
        val[i] = i;
    }
}

// 4. Karras' radix tree: internal node i (0 .. n-2) covers sorted positions
// [first, last]; its children are internal nodes or leaves (LEAF_BIT).
__device__ __forceinline__ int delta(const uint64_t* k, uint32_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= (int64_t)n) return -1;
    const uint64_t x = k[i] ^ k[j];
    if (x) return __clzll((long long)x);
    return 64 + __clz((int)((uint32_t)i ^ (uint32_t)j));  // equal codes: by position
}
__global__ __launch_bounds__(GB) void k_karras(const uint64_t* k, uint32_t n, uint32_t* child, uint32_t* parent_i,
                                               uint32_t* parent_l, uint2* range) {
    for (uint32_t ii = blockIdx.x * GB + threadIdx.x; ii + 1 < n; ii += gridDim.x * GB) {
        const int64_t i = ii;
        const int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = delta(k, n, i, i - d);
        int64_t lmax = 2;
        while (delta(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
        int64_t l = 0;
        for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
            if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
        const int64_t j = i + l * d;
        const int dnode = delta(k, n, i, j);
        int64_t s = 0;
        for (int64_t t = (l + 1) >> 1;; t = (t + 1) >> 1) {
            if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
            if (t == 1) break;
        }
        const int64_t g = i + s * d + (d < 0 ? -1 : 0);
        const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
        const uint32_t left = (lo == g) ? ((uint32_t)g | LEAF_BIT) : (uint32_t)g;
        const uint32_t right = (hi == g + 1) ? ((uint32_t)(g + 1) | LEAF_BIT) : (uint32_t)(g + 1);
        child[2 * ii] = left;
        child[2 * ii + 1] = right;
        range[ii] = make_uint2((uint32_t)lo, (uint32_t)hi);
        if (left & LEAF_BIT) parent_l[left & ~LEAF_BIT] = ii;
        else parent_i[left] = ii;
        if (right & LEAF_BIT) parent_l[right & ~LEAF_BIT] = ii;
        else parent_i[right] = ii;
    }
}

// 5. boxes and leaf counts bottom up: a leaf's thread climbs while it is the
// second child to arrive
__global__ __launch_bounds__(GB) void k_bottom_up(const uint32_t* val, const DBox* pbox, uint32_t n,
                                                  const uint32_t* child, const uint32_t* parent_i,
                                                  const uint32_t* parent_l, uint32_t* arrive, DBox* ibox,
                                                  uint32_t* nleaf) {
    for (uint32_t j = blockIdx.x * GB + threadIdx.x; j < n; j += gridDim.x * GB) {
        if (n == 1) break;
        uint32_t node = parent_l[j];
        while (true) {
            __threadfence();
            if (atomicAdd(arrive + node, 1u) == 0u) break;  // the sibling finishes this node
            __threadfence();
            DBox b{{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
            uint32_t cnt = 0;
            for (int c = 0; c < 2; ++c) {
                const uint32_t ch = child[2 * node + c];
                DBox cb;
                if (ch & LEAF_BIT) {
                    cb = pbox[val[ch & ~LEAF_BIT]];
                    cnt += 1;
                } else {
                    // written by the other climber before its atomic
                    const volatile DBox* vb = ibox + ch;
                    for (int a = 0; a < 3; ++a) { cb.lo[a] = vb->lo[a]; cb.hi[a] = vb->hi[a]; }
                    cnt += ((volatile uint32_t*)nleaf)[ch];
                }
                // Union<BBox> (bbox.rs:170-183): left first, then right
                for (int a = 0; a < 3; ++a) {
                    b.lo[a] = fminf(b.lo[a], cb.lo[a]);
                    b.hi[a] = fmaxf(b.hi[a], cb.hi[a]);
                }
            }
            ibox[node] = b;
            nleaf[node] = cnt;
            if (node == 0) break;
            node = parent_i[node];
        }
    }
}

// 6. depth-first positions, one tree level a launch: internal node x at
// position p has its first child at p + 1 and its second at p + 2 * (leaves
// under the first child); depth d for both.  Lists double-buffered, counts in
// three rotating words (level L reads count L % 3, adds to (L + 1) % 3 and
// clears (L + 2) % 3, which level L - 1 read).
__global__ __launch_bounds__(GB) void k_preorder_level(int level, const uint32_t* child, const uint32_t* nleaf,
                                                       const uint32_t* list_in, uint32_t* list_out, uint32_t* counts,
                                                       uint32_t* pos_i, uint32_t* pos_l, uint32_t* depth_i,
                                                       uint32_t* maxdepth) {
    const uint32_t cnt = counts[level % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[(level + 2) % 3] = 0u;
    for (uint32_t t = blockIdx.x * GB + threadIdx.x; t < cnt; t += gridDim.x * GB) {
        const uint32_t x = list_in[t];
        const uint32_t p = pos_i[x];
        const uint32_t L = child[2 * x], R = child[2 * x + 1];
        const uint32_t nl = (L & LEAF_BIT) ? 1u : nleaf[L];
        const uint32_t pl = p + 1, pr = p + 2 * nl;
        const uint32_t d = (uint32_t)level + 1;
        atomicMax(maxdepth, d);
        if (L & LEAF_BIT) pos_l[L & ~LEAF_BIT] = pl;
        else { pos_i[L] = pl; depth_i[L] = d; list_out[atomicAdd(counts + (level + 1) % 3, 1u)] = L; }
        if (R & LEAF_BIT) pos_l[R & ~LEAF_BIT] = pr;
        else { pos_i[R] = pr; depth_i[R] = d; list_out[atomicAdd(counts + (level + 1) % 3, 1u)] = R; }
    }
}

// 7a. reference nodes in depth-first order, record flags, prim records
__global__ __launch_bounds__(GB) void k_rnodes(uint32_t n, const uint32_t* child, const DBox* ibox, const DBox* pbox,
                                               const uint32_t* val, const uint32_t* pos_i, const uint32_t* pos_l,
                                               const uint32_t* depth_i, RNode* rn, uint32_t* is_rec) {
    const uint32_t total = 2 * n - 1;
    for (uint32_t t = blockIdx.x * GB + threadIdx.x; t < total; t += gridDim.x * GB) {
        if (t < n - 1) {
            // interior node t: the children's axis of greatest centre separation
            // picks the near child (bvh.rs:409-415 reads the node's axis)
            const uint32_t L = child[2 * t], R = child[2 * t + 1];
            const DBox bl = (L & LEAF_BIT) ? pbox[val[L & ~LEAF_BIT]] : ibox[L];
            const DBox br = (R & LEAF_BIT) ? pbox[val[R & ~LEAF_BIT]] : ibox[R];
            int axis = 0;
            float best = -1.0f;
            for (int a = 0; a < 3; ++a) {
                const float sep = fabsf((br.lo[a] + br.hi[a]) - (bl.lo[a] + bl.hi[a]));
                if (sep > best) { best = sep; axis = a; }
            }
            RNode r;
            const DBox& b = ibox[t];
            for (int a = 0; a < 3; ++a) { r.lo[a] = b.lo[a]; r.hi[a] = b.hi[a]; }
            r.off = (R & LEAF_BIT) ? pos_l[R & ~LEAF_BIT] : pos_i[R];
            r.meta = (uint32_t)axis;
            rn[pos_i[t]] = r;
            is_rec[pos_i[t]] = (depth_i[t] & 1u) == 0u ? 1u : 0u;
        } else {
            const uint32_t j = t - (n - 1);  // leaf at sorted position j
            RNode r;
            const DBox& b = pbox[val[j]];
            for (int a = 0; a < 3; ++a) { r.lo[a] = b.lo[a]; r.hi[a] = b.hi[a]; }
            r.off = j;
            r.meta = 3u | (1u << 2);
            const uint32_t p = n == 1 ? 0u : pos_l[j];
            rn[p] = r;
            is_rec[p] = 0u;
        }
    }
}

__device__ __forceinline__ float as_f(uint32_t u) { return __uint_as_float(u); }

__global__ __launch_bounds__(GB) void k_prim_recs(uint32_t n, uint32_t nfaces, const uint32_t* val,
                                                  const float* verts, const uint32_t* idx, const RNode* rn,
                                                  const uint32_t* pos_l, PrimRec* pr, PrimRec64* pr64) {
    for (uint32_t j = blockIdx.x * GB + threadIdx.x; j < n; j += gridDim.x * GB) {
        const uint32_t id = val[j];
        PrimRec r;
        for (int k = 0; k < 12; ++k) r.w[k] = 0.0f;
        r.w[10] = as_f(id);
        uint32_t kind;
        if (id < nfaces) {
            const float* p1 = verts + 3 * (size_t)idx[3 * (size_t)id + 2];
            const float* p2 = verts + 3 * (size_t)idx[3 * (size_t)id + 1];
            const float* p3 = verts + 3 * (size_t)idx[3 * (size_t)id + 0];
            for (int a = 0; a < 3; ++a) {
                r.w[a] = p1[a];
                r.w[3 + a] = p2[a] - p1[a];
                r.w[6 + a] = p3[a] - p1[a];
            }
            kind = 0u;
        } else {
            r.w[9] = as_f(id - nfaces);
            kind = 1u;
        }
        kind |= 0x100u;  // one prim a leaf: every record is its leaf's last
        r.w[11] = as_f(kind);
        pr[j] = r;
        // the 64-byte record: the leaf's exact box beside the triangle
        PrimRec64 q;
        for (int k = 0; k < 9; ++k) q.w[k] = r.w[k];
        if ((kind & 0xFFu) != 0u) q.w[0] = r.w[9];
        const RNode& leaf = rn[n == 1 ? 0u : pos_l[j]];
        for (int a = 0; a < 3; ++a) { q.w[9 + a] = leaf.lo[a]; q.w[12 + a] = leaf.hi[a]; }
        q.w[15] = r.w[11];
        pr64[j] = q;
    }
}

// 7b. four-wide records: record q at depth-first position a (an interior
// node at even depth), slots = the grandchildren through each child (a leaf
// child is a slot itself); refs: ~(prim position) for a leaf, else the
// record index of the node (qidx = exclusive scan of is_rec)
__device__ __forceinline__ bool is_leaf_node(const RNode& r) { return (r.meta & 3u) == 3u; }
__global__ __launch_bounds__(GB) void k_qnodes(uint32_t nnodes, const RNode* rn, const uint32_t* is_rec,
                                               const uint32_t* qidx, QNode* qn) {
    for (uint32_t a = blockIdx.x * GB + threadIdx.x; a < nnodes; a += gridDim.x * GB) {
        if (!is_rec[a]) continue;
        QNode q;
        const uint32_t b = a + 1, c = rn[a].off;
        uint32_t s[4];
        const uint32_t none = 0xFFFFFFFFu;
        if (is_leaf_node(rn[b])) { s[0] = b; s[1] = none; } else { s[0] = b + 1; s[1] = rn[b].off; }
        if (is_leaf_node(rn[c])) { s[2] = c; s[3] = none; } else { s[2] = c + 1; s[3] = rn[c].off; }
        const uint32_t ab = is_leaf_node(rn[b]) ? 0u : (rn[b].meta & 3u);
        const uint32_t ac = is_leaf_node(rn[c]) ? 0u : (rn[c].meta & 3u);
        q.meta = (rn[a].meta & 3u) | (ab << 2) | (ac << 4);
        q.pad[0] = q.pad[1] = q.pad[2] = 0u;
        for (int k = 0; k < 4; ++k) {
            if (s[k] == none) {
                q.ref[k] = QEMPTY;
                q.lx[k] = q.ly[k] = q.lz[k] = FLT_MAX;
                q.hx[k] = q.hy[k] = q.hz[k] = -FLT_MAX;
                continue;
            }
            const RNode& r = rn[s[k]];
            q.ref[k] = is_leaf_node(r) ? ~(int32_t)r.off : (int32_t)qidx[s[k]];
            q.lx[k] = r.lo[0]; q.ly[k] = r.lo[1]; q.lz[k] = r.lo[2];
            q.hx[k] = r.hi[0]; q.hy[k] = r.hi[1]; q.hz[k] = r.hi[2];
        }
        qn[qidx[a]] = q;
    }
}

// 7c. the compressed records (as compress_records on the host: per axis the
// smallest power-of-two scale that spans the union box in 255 steps, bytes
// rounded outward and checked against the float the traversal decodes)
__device__ __forceinline__ float cq_decode_d(float o, uint32_t q, float s) {
    const float t = (float)q * s;  // exact
    return o + t;
}
__device__ bool cq_axis_d(float ulo, float uhi, const float* lo, const float* hi, const bool* used, uint8_t* qlo,
                          uint8_t* qhi, uint32_t* ebyte) {
    if (!(fabsf(ulo) <= FLT_MAX) || !(fabsf(uhi) <= FLT_MAX)) return false;
    const double ext = (double)uhi - (double)ulo;
    int e = -126;
    if (ext > 0.0) {
        e = (int)ceil(log2(ext / 255.0));
        if (e < -126) e = -126;
    }
    for (;; ++e) {
        if (e > 119) return false;
        if (cq_decode_d(ulo, 255u, ldexpf(1.0f, e)) >= uhi) break;
    }
    const float s = ldexpf(1.0f, e);
    for (int k = 0; k < 4; ++k) {
        if (!used[k]) { qlo[k] = 255; qhi[k] = 0; continue; }
        const double a = floor(((double)lo[k] - (double)ulo) / (double)s);
        int ql = a < 0 ? 0 : (a > 255 ? 255 : (int)a);
        while (ql > 0 && cq_decode_d(ulo, (uint32_t)ql, s) > lo[k]) --ql;
        const double b = ceil(((double)hi[k] - (double)ulo) / (double)s);
        int qh = b < 0 ? 0 : (b > 255 ? 255 : (int)b);
        while (qh < 255 && cq_decode_d(ulo, (uint32_t)qh, s) < hi[k]) ++qh;
        if (cq_decode_d(ulo, (uint32_t)ql, s) > lo[k] || cq_decode_d(ulo, (uint32_t)qh, s) < hi[k]) return false;
        qlo[k] = (uint8_t)ql;
        qhi[k] = (uint8_t)qh;
    }
    *ebyte = (uint32_t)(e + 127);
    return true;
}
__global__ __launch_bounds__(GB) void k_cnodes(uint32_t nq, const QNode* qn, CNode* cn, uint32_t* bad) {
    for (uint32_t i = blockIdx.x * GB + threadIdx.x; i < nq; i += gridDim.x * GB) {
        const QNode& q = qn[i];
        CNode c;
        bool used[4];
        float u[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int k = 0; k < 4; ++k) {
            used[k] = q.ref[k] != QEMPTY;
            c.ref[k] = q.ref[k];
            if (!used[k]) continue;
            u[0] = fminf(u[0], q.lx[k]); u[1] = fminf(u[1], q.ly[k]); u[2] = fminf(u[2], q.lz[k]);
            u[3] = fmaxf(u[3], q.hx[k]); u[4] = fmaxf(u[4], q.hy[k]); u[5] = fmaxf(u[5], q.hz[k]);
        }
        uint32_t ex = 0, ey = 0, ez = 0;
        if (!cq_axis_d(u[0], u[3], q.lx, q.hx, used, c.q + 0, c.q + 4, &ex) ||
            !cq_axis_d(u[1], u[4], q.ly, q.hy, used, c.q + 8, c.q + 12, &ey) ||
            !cq_axis_d(u[2], u[5], q.lz, q.hz, used, c.q + 16, c.q + 20, &ez)) {
            atomicOr(bad, 1u);
            continue;
        }
        c.ox = u[0]; c.oy = u[1]; c.oz = u[2];
        c.emeta = ex | (ey << 8) | (ez << 16) | ((q.meta & 63u) << 24);
        c.pad[0] = c.pad[1] = 0u;
        cn[i] = c;
    }
}

uint32_t grid_for(uint64_t n) {
    const uint64_t b = (n + GB - 1) / GB;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, (uint64_t)device_cu_count() * 16));
}

template <typename T>
hipError_t dalloc(T** p, size_t count) {
    return hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
}

struct Scratch {
    std::vector<void*> ptrs;
    ~Scratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    hipError_t get(T** p, size_t count) {
        hipError_t e = dalloc(p, count);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

#define GCHK(x)                                          \
    do {                                                 \
        hipError_t e_ = (x);                             \
        if (e_ != hipSuccess) return hip_fail(e_, #x);   \
    } while (0)

}  // namespace

// The spheres' world boxes and records, as build_bvh makes them (host; a
// handful per scene).
static bool sphere_inputs(const SphereIn* sph, uint32_t nsph, std::vector<DBox>* boxes,
                          std::vector<SphereRec>* recs, std::string* err) {
    // reuse the host builder's exact per-sphere arithmetic: build a one-sphere
    // tree and take its record and root box
    boxes->resize(nsph);
    recs->resize(nsph);
    for (uint32_t s = 0; s < nsph; ++s) {
        HostBvh h;
        if (!build_bvh(nullptr, 0, nullptr, 0, sph + s, 1, 0, 1, &h, err)) return false;
        (*recs)[s] = h.spheres[0];
        for (int a = 0; a < 3; ++a) { (*boxes)[s].lo[a] = h.root_lo[a]; (*boxes)[s].hi[a] = h.root_hi[a]; }
    }
    return true;
}

int gpu_build_scene(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t nfaces, const SphereIn* sph,
                    uint32_t nsph, DeviceScene* S, double* build_ms) {
    auto t0 = std::chrono::steady_clock::now();
    memset(S, 0, sizeof(*S));
    const uint64_t N64 = (uint64_t)nfaces + nsph;
    if (N64 == 0) return fail(PBRTX_ERR_BUILD, "scene has no primitives");
    if (N64 >= (1ull << 30)) return fail(PBRTX_ERR_BUILD, "too many primitives");
    if (nfaces && (!verts || !idx)) return fail(PBRTX_ERR_ARG, "null mesh arrays");
    if (nsph && !sph) return fail(PBRTX_ERR_ARG, "null sphere array");
    for (uint64_t i = 0; i < 3ull * nfaces; ++i)
        if (idx[i] >= nverts) return fail(PBRTX_ERR_BUILD, "vertex index out of range");
    std::string err;
    std::vector<DBox> sbox;
    std::vector<SphereRec> srec;
    if (!sphere_inputs(sph, nsph, &sbox, &srec, &err)) return fail(PBRTX_ERR_BUILD, err);
    const uint32_t n = (uint32_t)N64;
    const uint32_t nnodes = 2 * n - 1;
    hipStream_t st = nullptr;
    Scratch T;
    float *d_verts, *d_cen, *d_part, *d_cb;
    uint32_t *d_idx, *d_val, *d_val2, *d_child, *d_pi, *d_pl, *d_arrive, *d_nleaf, *d_posi, *d_posl, *d_depth,
        *d_list0, *d_list1, *d_counts, *d_maxd, *d_isrec, *d_qidx, *d_bad;
    uint64_t *d_key, *d_key2;
    uint2* d_range;
    DBox *d_pbox, *d_sbox, *d_ibox;
    GCHK(T.get(&d_verts, 3 * (size_t)nverts));
    GCHK(T.get(&d_idx, 3 * (size_t)nfaces));
    GCHK(T.get(&d_sbox, nsph));
    GCHK(T.get(&d_pbox, n));
    GCHK(T.get(&d_cen, 3 * (size_t)n));
    if (nverts) GCHK(hipMemcpyAsync(d_verts, verts, 12 * (size_t)nverts, hipMemcpyHostToDevice, st));
    if (nfaces) GCHK(hipMemcpyAsync(d_idx, idx, 12 * (size_t)nfaces, hipMemcpyHostToDevice, st));
    if (nsph) GCHK(hipMemcpyAsync(d_sbox, sbox.data(), sizeof(DBox) * nsph, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_prim_boxes, dim3(grid_for(n)), dim3(GB), 0, st, d_verts, d_idx, nfaces, d_sbox, nsph, d_pbox,
                       d_cen);
    GCHK(hipGetLastError());
    const uint32_t nb = std::min<uint32_t>(grid_for(n), 1024u);
    GCHK(T.get(&d_part, 6 * (size_t)nb));
    GCHK(T.get(&d_cb, 6));
    hipLaunchKernelGGL(k_cen_bounds, dim3(nb), dim3(GB), 0, st, d_cen, n, d_part);
    hipLaunchKernelGGL(k_cen_bounds_final, dim3(1), dim3(GB), 0, st, d_part, nb, d_cb);
    GCHK(hipGetLastError());
    GCHK(T.get(&d_key, n));
    GCHK(T.get(&d_key2, n));
    GCHK(T.get(&d_val, n));
    GCHK(T.get(&d_val2, n));
    hipLaunchKernelGGL(k_morton, dim3(grid_for(n)), dim3(GB), 0, st, d_cen, n, d_cb, d_key, d_val);
    GCHK(hipGetLastError());
    {
        size_t tb = 0;
        GCHK(rocprim::radix_sort_pairs(nullptr, tb, d_key, d_key2, d_val, d_val2, n, 0, 63, st));
        void* tmp;
        GCHK(T.get((uint8_t**)&tmp, tb));
        GCHK(rocprim::radix_sort_pairs(tmp, tb, d_key, d_key2, d_val, d_val2, n, 0, 63, st));
    }
    std::swap(d_key, d_key2);
    std::swap(d_val, d_val2);
    GCHK(T.get(&d_child, 2 * (size_t)n));
    GCHK(T.get(&d_pi, n));
    GCHK(T.get(&d_pl, n));
    GCHK(T.get(&d_range, n));
    GCHK(T.get(&d_arrive, n));
    GCHK(T.get(&d_nleaf, n));
    GCHK(T.get(&d_ibox, n));
    GCHK(fill_u32(d_arrive, 0u, n, st));
    if (n > 1) {
        hipLaunchKernelGGL(k_karras, dim3(grid_for(n)), dim3(GB), 0, st, d_key, n, d_child, d_pi, d_pl, d_range);
        hipLaunchKernelGGL(k_bottom_up, dim3(grid_for(n)), dim3(GB), 0, st, d_val, d_pbox, n, d_child, d_pi, d_pl,
                           d_arrive, d_ibox, d_nleaf);
        GCHK(hipGetLastError());
    }
    GCHK(T.get(&d_posi, n));
    GCHK(T.get(&d_posl, n));
    GCHK(T.get(&d_depth, n));
    GCHK(T.get(&d_list0, n));
    GCHK(T.get(&d_list1, n));
    GCHK(T.get(&d_counts, 4));
    GCHK(T.get(&d_maxd, 1));
    {
        // level 0: the root (internal node 0) at position 0, depth 0
        const uint32_t init_counts[4] = {n > 1 ? 1u : 0u, 0u, 0u, 0u};
        const uint32_t zero = 0;
        GCHK(hipMemcpyAsync(d_counts, init_counts, sizeof(init_counts), hipMemcpyHostToDevice, st));
        GCHK(hipMemcpyAsync(d_list0, &zero, 4, hipMemcpyHostToDevice, st));
        GCHK(hipMemcpyAsync(d_posi, &zero, 4, hipMemcpyHostToDevice, st));
        GCHK(hipMemcpyAsync(d_depth, &zero, 4, hipMemcpyHostToDevice, st));
        GCHK(hipMemcpyAsync(d_maxd, &zero, 4, hipMemcpyHostToDevice, st));
        GCHK(hipStreamSynchronize(st));  // the host values above stay valid until here
        for (int level = 0; level < MAX_LEVELS && n > 1; ++level) {
            hipLaunchKernelGGL(k_preorder_level, dim3(grid_for(n / 2 + 1)), dim3(GB), 0, st, level, d_child, d_nleaf,
                               (level & 1) ? d_list1 : d_list0, (level & 1) ? d_list0 : d_list1, d_counts, d_posi,
                               d_posl, d_depth, d_maxd);
            GCHK(hipGetLastError());
            if ((level & 15) == 15) {
                uint32_t c[3];
                GCHK(hipMemcpy(c, d_counts, sizeof(c), hipMemcpyDeviceToHost));
                if (c[(level + 1) % 3] == 0) break;
            }
        }
    }
    // layouts
    RNode* d_rn;
    GCHK(dalloc(&d_rn, nnodes));
    S->d_rnodes = d_rn;
    GCHK(T.get(&d_isrec, nnodes));
    GCHK(T.get(&d_qidx, nnodes));
    hipLaunchKernelGGL(k_rnodes, dim3(grid_for(nnodes)), dim3(GB), 0, st, n, d_child, d_ibox, d_pbox, d_val, d_posi,
                       d_posl, d_depth, d_rn, d_isrec);
    GCHK(hipGetLastError());
    PrimRec* d_pr;
    PrimRec64* d_pr64;
    GCHK(dalloc(&d_pr, n));
    S->d_prims = d_pr;
    GCHK(dalloc(&d_pr64, n));
    S->d_prims64 = d_pr64;
    hipLaunchKernelGGL(k_prim_recs, dim3(grid_for(n)), dim3(GB), 0, st, n, nfaces, d_val, d_verts, d_idx, d_rn, d_posl,
                       d_pr, d_pr64);
    GCHK(hipGetLastError());
    {
        size_t tb = 0;
        GCHK(rocprim::exclusive_scan(nullptr, tb, d_isrec, d_qidx, 0u, nnodes, rocprim::plus<uint32_t>(), st));
        void* tmp;
        GCHK(T.get((uint8_t**)&tmp, tb));
        GCHK(rocprim::exclusive_scan(tmp, tb, d_isrec, d_qidx, 0u, nnodes, rocprim::plus<uint32_t>(), st));
    }
    uint32_t nq = 0, maxd = 0, last_flag = 0;
    GCHK(hipMemcpyAsync(&nq, d_qidx + (nnodes - 1), 4, hipMemcpyDeviceToHost, st));
    GCHK(hipMemcpyAsync(&last_flag, d_isrec + (nnodes - 1), 4, hipMemcpyDeviceToHost, st));
    GCHK(hipMemcpyAsync(&maxd, d_maxd, 4, hipMemcpyDeviceToHost, st));
    GCHK(hipStreamSynchronize(st));
    nq += last_flag;
    if (n > 1 && maxd + 1 >= (uint32_t)MAX_LEVELS) return fail(PBRTX_ERR_BUILD, "linear BVH deeper than expected");
    QNode* d_qn;
    CNode* d_cn;
    GCHK(dalloc(&d_qn, nq));
    S->d_qnodes = d_qn;
    GCHK(dalloc(&d_cn, nq));
    S->d_cnodes = d_cn;
    GCHK(T.get(&d_bad, 1));
    GCHK(fill_u32(d_bad, 0u, 1, st));
    if (nq) {
        hipLaunchKernelGGL(k_qnodes, dim3(grid_for(nnodes)), dim3(GB), 0, st, nnodes, d_rn, d_isrec, d_qidx, d_qn);
        hipLaunchKernelGGL(k_cnodes, dim3(grid_for(nq)), dim3(GB), 0, st, nq, d_qn, d_cn, d_bad);
        GCHK(hipGetLastError());
    }
    // ordered prim ids, spheres, root box
    uint32_t* d_order;
    GCHK(dalloc(&d_order, n));
    S->d_order = d_order;
    GCHK(hipMemcpyAsync(d_order, d_val, 4 * (size_t)n, hipMemcpyDeviceToDevice, st));
    if (nsph) {
        SphereRec* d_s;
        GCHK(dalloc(&d_s, nsph));
        S->d_spheres = d_s;
        GCHK(hipMemcpyAsync(d_s, srec.data(), sizeof(SphereRec) * nsph, hipMemcpyHostToDevice, st));
    }
    RNode root;
    uint32_t bad = 0;
    GCHK(hipMemcpyAsync(&root, d_rn, sizeof(RNode), hipMemcpyDeviceToHost, st));
    GCHK(hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, st));
    GCHK(hipStreamSynchronize(st));
    SceneView& v = S->view;
    v.rnodes = d_rn;
    v.qnodes = d_qn;
    v.prims = d_pr;
    v.spheres = nsph ? (const SphereRec*)S->d_spheres : nullptr;
    const bool cq = nq > 0 && !bad;
    v.cnodes = cq ? d_cn : nullptr;
    v.prims64 = cq ? d_pr64 : nullptr;
    v.ocnodes = v.cnodes;
    v.order = d_order;
    v.n_rnodes = nnodes;
    v.n_qnodes = nq;
    v.n_prims = n;
    v.n_spheres = nsph;
    v.qroot = n > 1 ? 0 : ~(int32_t)0;
    v.ocroot = v.qroot;
    v.depth = maxd;
    v.qdepth = n > 1 ? maxd / 2 + 1 : 0;
    v.odepth = v.qdepth;
    for (int a = 0; a < 3; ++a) { v.root_lo[a] = root.lo[a]; v.root_hi[a] = root.hi[a]; }
    S->bytes = (uint64_t)nnodes * sizeof(RNode) + (uint64_t)nq * (sizeof(QNode) + sizeof(CNode)) +
               (uint64_t)n * (sizeof(PrimRec) + sizeof(PrimRec64) + 4) + (uint64_t)nsph * sizeof(SphereRec);
    *build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PBRTX_OK;
}

}  // namespace px

extern "C" int pbrtx_scene_build_gpu(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t nfaces,
                                     const pbrtx_sphere* spheres, uint32_t nspheres, pbrtx_scene** out,
                                     double* build_ms) {
    using namespace px;
    if (!out) return fail(PBRTX_ERR_ARG, "out is null");
    *out = nullptr;
    static_assert(sizeof(pbrtx_sphere) == sizeof(SphereIn), "sphere layout");
    pbrtx_scene* s = new (std::nothrow) pbrtx_scene();
    if (!s) return fail(PBRTX_ERR_OOM, "host allocation");
    double ms = 0.0;
    int rc = gpu_build_scene(verts, nverts, idx, nfaces, reinterpret_cast<const SphereIn*>(spheres), nspheres, &s->s,
                             &ms);
    if (rc) {
        scene_free(&s->s);
        delete s;
        return rc;
    }
    if (build_ms) *build_ms = ms;
    *out = s;
    return PBRTX_OK;
}
