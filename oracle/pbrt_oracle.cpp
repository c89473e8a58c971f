// pbrt_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
//
// CPU restatement of pbrt_rust's ray-query and sampling path.  Each function
// follows the reference source operation for operation (file:line cited at the
// function) so that, compiled with -ffp-contract=off and without -ffast-math,
// it produces the same IEEE binary32 results the reference would.  Where the
// reference has no implementation (the Li integrators, the area light, the
// render loop's seeding and termination) the behaviour is the one the build
// defines (deviations D1-D11) and the device code follows the same definition.
#include "pbrt_oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// L0 math: geometry/{vector,point,normal}.rs
// ---------------------------------------------------------------------------
struct V3 {
    float x, y, z;
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
static inline V3 mk(float x, float y, float z) { return V3{x, y, z}; }
// Sub / Add component-wise: vector.rs:50-91, point.rs
static inline V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
// Mul<f32>: vector.rs:94-99
static inline V3 mul(V3 a, float f) { return mk(a.x * f, a.y * f, a.z * f); }
static inline V3 neg(V3 a) { return mk(-a.x, -a.y, -a.z); }
// Dot: vector.rs:169-173, left to right
static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Cross: vector.rs:175-182
static inline V3 cross(V3 a, V3 b) {
    return mk((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
// length_squared / length: vector.rs:41-47
static inline float length_sq(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline float length(V3 a) { return std::sqrt(length_sq(a)); }
// Div<f32> as reciprocal multiply: vector.rs:116-130
static inline V3 divs(V3 a, float f) {
    float recip = 1.0f / f;
    return mul(a, recip);
}
// Normalize: normal.rs:181-186
static inline V3 normalize(V3 a) { return divs(a, length(a)); }
// Rust f32::min / f32::max ignore a NaN operand, as fminf / fmaxf do.
static inline float rmin(float a, float b) { return std::fmin(a, b); }
static inline float rmax(float a, float b) { return std::fmax(a, b); }

// ---------------------------------------------------------------------------
// Matrix4x4 / Transform: transform/matrix4x4.rs, transform/transform.rs
// ---------------------------------------------------------------------------
struct M4 {
    float m[4][4];
};
static M4 m4_identity() {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = (i == j) ? 1.0f : 0.0f;
    return r;
}
static M4 m4_from(const float* a) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = a[4 * i + j];
    return r;
}
static void m4_store(const M4& a, float* out) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[4 * i + j] = a.m[i][j];
}
// Mul: matrix4x4.rs:167-182
static M4 m4_mul(const M4& a, const M4& b) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] +
                        a.m[i][3] * b.m[3][j];
    return r;
}
static M4 m4_transpose(const M4& a) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = a.m[j][i];
    return r;
}
// lu_decompose: matrix4x4.rs:66-118 (GE with scaled partial pivoting; the row
// swap covers columns k..3 only).
static bool m4_lu(const M4& a, M4& lu, int pivot[4]) {
    float s[4];
    for (int i = 0; i < 4; ++i) {
        float acc = 0.0f;
        for (int j = 0; j < 4; ++j) acc = rmax(std::fabs(a.m[i][j]), acc);
        s[i] = acc;
    }
    lu = a;
    for (int i = 0; i < 4; ++i) pivot[i] = i;
    for (int k = 0; k < 3; ++k) {
        float c = 0.0f;
        int p = k;
        for (int i = k; i < 4; ++i) {
            float cc = std::fabs(lu.m[i][k] / s[i]);
            if (cc > c) { c = cc; p = i; }
        }
        pivot[k] = p;
        if (c == 0.0f) return false;
        if (p != k) {
            for (int j = k; j < 4; ++j) std::swap(lu.m[k][j], lu.m[p][j]);
        }
        for (int i = k + 1; i < 4; ++i) {
            float mi = lu.m[i][k] / lu.m[k][k];
            lu.m[i][k] = mi;
            for (int j = k + 1; j < 4; ++j) lu.m[i][j] = lu.m[i][j] - mi * lu.m[k][j];
        }
    }
    if (std::fabs(lu.m[3][3]) < 1.0e-6f) return false;
    return true;
}
// solve_ax_b: matrix4x4.rs:120-142
static void m4_solve(const M4& lu, const int pivot[4], float b[4]) {
    for (int k = 0; k < 3; ++k) {
        if (pivot[k] != k) std::swap(b[pivot[k]], b[k]);
        for (int i = k + 1; i < 4; ++i) b[i] = b[i] - lu.m[i][k] * b[k];
    }
    b[3] = b[3] / lu.m[3][3];
    for (int i = 2; i >= 0; --i) {
        float sum = 0.0f;
        for (int j = i + 1; j < 4; ++j) sum = sum + lu.m[i][j] * b[j];
        b[i] = (b[i] - sum) / lu.m[i][i];
    }
}
// invert_with / inverse: matrix4x4.rs:147-160
static bool m4_inverse(const M4& a, M4& out) {
    M4 lu;
    int piv[4];
    if (!m4_lu(a, lu, piv)) return false;
    M4 rows;
    for (int c = 0; c < 4; ++c) {
        float b[4] = {0, 0, 0, 0};
        b[c] = 1.0f;
        m4_solve(lu, piv, b);
        for (int j = 0; j < 4; ++j) rows.m[c][j] = b[j];
    }
    out = m4_transpose(rows);
    return true;
}

struct Xf {
    M4 m, mi;
};
static Xf xf_identity() { return Xf{m4_identity(), m4_identity()}; }
// Transform::translate: transform.rs:40-51
static Xf xf_translate(V3 v) {
    Xf t = xf_identity();
    t.m.m[0][3] = v.x; t.m.m[1][3] = v.y; t.m.m[2][3] = v.z;
    t.mi.m[0][3] = -v.x; t.mi.m[1][3] = -v.y; t.mi.m[2][3] = -v.z;
    return t;
}
// Transform::scale: transform.rs:53-65
static Xf xf_scale(float x, float y, float z) {
    Xf t = xf_identity();
    t.m.m[0][0] = x; t.m.m[1][1] = y; t.m.m[2][2] = z;
    t.mi.m[0][0] = 1.0f / x; t.mi.m[1][1] = 1.0f / y; t.mi.m[2][2] = 1.0f / z;
    return t;
}
// Mul for Transform: transform.rs:266-272 (m = a.m*b.m, m_inv = b.m_inv*a.m_inv)
static Xf xf_mul(const Xf& a, const Xf& b) { return Xf{m4_mul(a.m, b.m), m4_mul(b.mi, a.mi)}; }
static Xf xf_inverse(const Xf& a) { return Xf{a.mi, a.m}; }
// From<Matrix4x4>: transform.rs:288-293 (inverse by LU)
static bool xf_from(const M4& m, Xf& out) {
    M4 inv;
    if (!m4_inverse(m, inv)) return false;
    out = Xf{m, inv};
    return true;
}
// ApplyTransform<Point>: transform.rs:195-207
static V3 xf_point(const M4& m, V3 p) {
    float x = p.x, y = p.y, z = p.z;
    float xt = m.m[0][0] * x + m.m[0][1] * y + m.m[0][2] * z + m.m[0][3];
    float yt = m.m[1][0] * x + m.m[1][1] * y + m.m[1][2] * z + m.m[1][3];
    float zt = m.m[2][0] * x + m.m[2][1] * y + m.m[2][2] * z + m.m[2][3];
    float w = m.m[3][0] * x + m.m[3][1] * y + m.m[3][2] * z + m.m[3][3];
    if (w != 1.0f) return mk(xt / w, yt / w, zt / w);
    return mk(xt, yt, zt);
}
// ApplyTransform<Vector>: transform.rs:210-218
static V3 xf_vector(const M4& m, V3 v) {
    float x = v.x, y = v.y, z = v.z;
    return mk(m.m[0][0] * x + m.m[0][1] * y + m.m[0][2] * z,
              m.m[1][0] * x + m.m[1][1] * y + m.m[1][2] * z,
              m.m[2][0] * x + m.m[2][1] * y + m.m[2][2] * z);
}
// swaps_handedness: transform.rs:178-185
static bool xf_swaps_handedness(const Xf& t) {
    const M4& a = t.m;
    float det = a.m[0][0] * (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1]) -
                a.m[0][1] * (a.m[1][0] * a.m[2][2] - a.m[1][2] * a.m[2][0]) +
                a.m[0][2] * (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]);
    return 0.0f > det;
}

// ---------------------------------------------------------------------------
// BBox: bbox.rs
// ---------------------------------------------------------------------------
struct BBox {
    V3 pmin, pmax;
};
// BBox::new: bbox.rs:36-41
static inline BBox bbox_empty() {
    return BBox{mk(FLT_MAX, FLT_MAX, FLT_MAX), mk(-FLT_MAX, -FLT_MAX, -FLT_MAX)};
}
// Union<Point>: bbox.rs:155-168
static inline BBox bbox_union_p(const BBox& b, V3 p) {
    return BBox{mk(rmin(b.pmin.x, p.x), rmin(b.pmin.y, p.y), rmin(b.pmin.z, p.z)),
                mk(rmax(b.pmax.x, p.x), rmax(b.pmax.y, p.y), rmax(b.pmax.z, p.z))};
}
// Union<BBox>: bbox.rs:170-183
static inline BBox bbox_union(const BBox& a, const BBox& b) {
    return BBox{mk(rmin(a.pmin.x, b.pmin.x), rmin(a.pmin.y, b.pmin.y), rmin(a.pmin.z, b.pmin.z)),
                mk(rmax(a.pmax.x, b.pmax.x), rmax(a.pmax.y, b.pmax.y), rmax(a.pmax.z, b.pmax.z))};
}
// surface_area: bbox.rs:73-78
static inline float bbox_surface_area(const BBox& b) {
    float dx = rmax(b.pmax.x - b.pmin.x, 0.0f);
    float dy = rmax(b.pmax.y - b.pmin.y, 0.0f);
    float dz = rmax(b.pmax.z - b.pmin.z, 0.0f);
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}
// max_extent: bbox.rs:87-96
static inline int bbox_max_extent(const BBox& b) {
    V3 d = sub(b.pmax, b.pmin);
    if (d.x > d.y && d.x > d.z) return 0;
    if (d.y > d.z) return 1;
    return 2;
}

struct Ray {
    V3 o, d;
    float mint, maxt;
};

// BBox::intersect: bbox.rs:185-209
static bool bbox_intersect(const BBox& b, const Ray& r, float* t0o, float* t1o) {
    float t0 = r.mint, t1 = r.maxt;
    for (int i = 0; i < 3; ++i) {
        float inv = 1.0f / r.d[i];
        float ta = (b.pmin[i] - r.o[i]) * inv;
        float tb = (b.pmax[i] - r.o[i]) * inv;
        if (ta > tb) std::swap(ta, tb);
        t0 = rmax(ta, t0);
        t1 = rmin(tb, t1);
        if (t0 > t1) return false;
    }
    if (t0o) *t0o = t0;
    if (t1o) *t1o = t1;
    return true;
}

// ---------------------------------------------------------------------------
// utils/mod.rs
// ---------------------------------------------------------------------------
// quadratic: utils/mod.rs:56-92
static bool quadratic(float a, float b, float c, float* t0, float* t1) {
    float descrim = b * b - 4.0f * a * c;
    if (descrim < 0.0f) return false;
    if (std::fabs(descrim) < 1e-6f) {
        if (a == 0.0f) return false;
        float t = -b / (2.0f * a);
        *t0 = t; *t1 = t;
        return true;
    }
    float root = std::sqrt(descrim);
    float q = (b < 0.0f) ? -0.5f * (b - root) : -0.5f * (b + root);
    if (a == 0.0f) return false;
    if (q == 0.0f) return false;  // the reference asserts q != 0 (utils/mod.rs:86)
    float r0 = q / a, r1 = c / q;
    if (r0 < r1) { *t0 = r0; *t1 = r1; } else { *t0 = r1; *t1 = r0; }
    return true;
}

// partition_by: utils/mod.rs:112-169 (median-of-three, three-way, one side recursion)
template <typename T, typename F>
static void partition_by(T* v, size_t nv, F f) {
    while (true) {
        if (nv < 3) {
            if (nv == 2 && f(v[1]) < f(v[0])) std::swap(v[0], v[1]);
            return;
        }
        auto fst = f(v[0]);
        auto mid = f(v[nv / 2]);
        auto lst = f(v[nv - 1]);
        decltype(fst) pivot;
        if (fst < mid && mid < lst) pivot = mid;
        else if (mid < fst && fst < lst) pivot = fst;
        else pivot = lst;
        size_t last_smaller = 0, num_pivots = 0;
        for (size_t i = 0; i < nv; ++i) {
            auto bv = f(v[i]);
            if (bv < pivot) {
                std::swap(v[last_smaller + num_pivots], v[i]);
                std::swap(v[last_smaller + num_pivots], v[last_smaller]);
                last_smaller += 1;
            } else if (bv == pivot) {
                std::swap(v[last_smaller + num_pivots], v[i]);
                num_pivots += 1;
            }
        }
        size_t pivot_idx = std::max<size_t>(last_smaller, 1);
        if (pivot_idx + num_pivots <= nv / 2) {
            v = v + pivot_idx;
            nv = nv - pivot_idx;
        } else if (pivot_idx >= nv / 2) {
            nv = pivot_idx;
        } else {
            return;
        }
    }
}

// get_num_subwindows_2d / get_crop_window: utils/mod.rs:171-205
static void num_subwindows_2d(size_t count, size_t aspect, size_t* nx, size_t* ny) {
    size_t x = 1, y = count;
    while ((y % 2) == 0 && 2 * aspect * x < y) { y /= 2; x *= 2; }
    *nx = x; *ny = y;
}
static void crop_window(size_t num, size_t count, float aspect, float out[4]) {
    size_t nx, ny;
    if (aspect < 1.0f) {
        size_t inva = (size_t)(1.0f / aspect);
        num_subwindows_2d(count, inva, &nx, &ny);
    } else {
        size_t x, y;
        num_subwindows_2d(count, (size_t)aspect, &x, &y);
        nx = y; ny = x;
    }
    size_t xo = num % nx, yo = num / nx;
    out[0] = (float)xo / (float)nx;
    out[1] = (float)(xo + 1) / (float)nx;
    out[2] = (float)yo / (float)ny;
    out[3] = (float)(yo + 1) / (float)ny;
}
// Lerp for f32: utils/mod.rs:15-21
static inline float lerpf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// ---------------------------------------------------------------------------
// Shapes
// ---------------------------------------------------------------------------
struct Mesh {
    std::vector<V3> p;  // world space, Mesh::new pre-transforms (mesh.rs:308)
};
struct Sphere {
    Xf o2w, w2o;
    bool ro, swaps;
    float radius, zmin, zmax, thetamin, thetamax, phimax;
};
// Sphere::new: sphere.rs:28-44
static Sphere sphere_new(const Xf& o2w, const Xf& w2o, bool ro, float rad, float z0, float z1,
                         float pm) {
    Sphere s;
    s.o2w = o2w; s.w2o = w2o; s.ro = ro; s.swaps = xf_swaps_handedness(o2w);
    s.radius = rad;
    s.zmin = std::min(std::max(rmin(z0, z1), -rad), rad);
    s.zmax = std::min(std::max(rmax(z0, z1), -rad), rad);
    s.thetamin = std::acos(s.zmin / rad);
    s.thetamax = std::acos(s.zmax / rad);
    float pmc = std::min(std::max(pm, 0.0f), 360.0f);
    s.phimax = pmc * (float)M_PI / 180.0f;  // as_radians: utils/mod.rs:37-39
    return s;
}
static const float PI_F = 3.14159265358979323846f;

// Sphere::get_intersection_point on an object-space ray: sphere.rs:46-107
static bool sphere_hit_obj(const Sphere& s, const Ray& ray, float* t_out, float* phi_out) {
    float a = length_sq(ray.d);
    float b = 2.0f * dot(ray.d, ray.o);
    float c = length_sq(ray.o) - s.radius * s.radius;
    float t0, t1;
    if (!quadratic(a, b, c, &t0, &t1)) return false;
    if (t0 > ray.maxt || t1 < ray.mint) return false;
    float thit = t0;
    if (t0 < ray.mint) {
        thit = t1;
        if (thit > ray.maxt) return false;
    }
    auto get_hit = [&](float t, V3* hit, float* angle) {
        *hit = add(ray.o, mul(ray.d, t));
        if (hit->x == 0.0f && hit->y == 0.0f) hit->x = 1e-5f * s.radius;
        float ang = std::atan2(hit->y, hit->x);
        if (ang < 0.0f) ang += 2.0f * PI_F;
        *angle = ang;
    };
    auto invalid = [&](V3 h, float ang) {
        return (h.z > -s.radius && h.z < s.zmin) || (h.z < s.radius && h.z > s.zmax) ||
               (ang > s.phimax);
    };
    V3 h;
    float ang;
    get_hit(thit, &h, &ang);
    if (invalid(h, ang)) {
        if (thit == t1) return false;
        if (t1 > ray.maxt) return false;
        thit = t1;
        get_hit(thit, &h, &ang);
        if (invalid(h, ang)) return false;
    }
    *t_out = thit;
    *phi_out = ang;
    return true;
}
static Ray ray_to_object(const Xf& w2o, const Ray& r) {
    Ray o = r;
    o.o = xf_point(w2o.m, r.o);
    o.d = xf_vector(w2o.m, r.d);
    return o;
}

// Triangle::get_intersection_point: mesh.rs:41-72
static bool tri_hit(V3 p1, V3 p2, V3 p3, const Ray& r, float* t_out, float* b1_out,
                    float* b2_out) {
    V3 e1 = sub(p2, p1);
    V3 e2 = sub(p3, p1);
    V3 s1 = cross(r.d, e2);
    float divisor = dot(s1, e1);
    if (divisor == 0.0f) return false;
    float inv = 1.0f / divisor;
    V3 s = sub(r.o, p1);
    float b1 = dot(s1, s) * inv;
    if (b1 < 0.0f || b1 > 1.0f) return false;
    V3 s2 = cross(s, e1);
    float b2 = dot(r.d, s2) * inv;
    if (b2 < 0.0f || (b1 + b2) > 1.0f) return false;
    float t = dot(e2, s2) * inv;
    if (t < r.mint || t > r.maxt) return false;
    *t_out = t; *b1_out = b1; *b2_out = b2;
    return true;
}

// A refined primitive: a triangle of a mesh (vertex triple in refine order
// [c,b,a], mesh.rs:329-332) or a sphere.
struct Prim {
    int kind;  // 0 triangle, 1 sphere
    uint32_t mesh, sphere;
    uint32_t v[3];
};

// ---------------------------------------------------------------------------
// BVH: primitive/aggregates/bvh.rs
// ---------------------------------------------------------------------------
struct Info {
    uint32_t prim;
    V3 centroid;
    BBox bounds;
};
struct PNode {  // PackedBVHNode, bvh.rs:260-272
    BBox bounds;
    bool leaf;
    uint32_t off;   // second_child_offset | prim_offset
    uint32_t meta;  // axis | num_prims
};

struct Scene {
    std::vector<Mesh> meshes;
    std::vector<Sphere> spheres;
    std::vector<Prim> prims;      // BVH input order
    std::vector<PNode> nodes;     // flattened
    std::vector<uint32_t> order;  // ordered primitives (prim indices)
};

static BBox prim_world_bound(const Scene& sc, const Prim& p) {
    if (p.kind == 0) {  // Triangle world_bound: mesh.rs:196-205
        const Mesh& m = sc.meshes[p.mesh];
        BBox b = bbox_empty();
        b = bbox_union_p(b, m.p[p.v[0]]);
        b = bbox_union_p(b, m.p[p.v[1]]);
        b = bbox_union_p(b, m.p[p.v[2]]);
        return b;
    }
    // Sphere world_bound: sphere.rs:111-126 + Transform xf(BBox) transform.rs:248-264
    const Sphere& s = sc.spheres[p.sphere];
    BBox ob{mk(-s.radius, -s.radius, s.zmin), mk(s.radius, s.radius, s.zmax)};
    const M4& m = s.o2w.m;
    V3 tx = xf_vector(m, mk(ob.pmax.x - ob.pmin.x, 0.0f, 0.0f));
    V3 ty = xf_vector(m, mk(0.0f, ob.pmax.y - ob.pmin.y, 0.0f));
    V3 tz = xf_vector(m, mk(0.0f, 0.0f, ob.pmax.z - ob.pmin.z));
    V3 tp = xf_point(m, ob.pmin);
    BBox b{tp, tp};
    b = bbox_union_p(b, add(tp, tx));
    b = bbox_union_p(b, add(tp, ty));
    b = bbox_union_p(b, add(tp, tz));
    b = bbox_union_p(b, add(add(tp, tx), ty));
    b = bbox_union_p(b, add(add(tp, tx), tz));
    b = bbox_union_p(b, add(add(tp, ty), tz));
    b = bbox_union_p(b, add(add(add(tp, tx), ty), tz));
    return b;
}

struct BuildNode {  // BVHNode, bvh.rs:21-35
    BBox bounds;
    bool leaf;
    uint32_t first, count;
    int axis;
    int c1, c2;  // indices into the builder's node pool
};

struct Builder {
    const Scene& sc;
    int split;
    uint32_t max_prims;
    std::vector<BuildNode> pool;
    std::vector<uint32_t> ordered;

    static int bucket_for(const Info& p, const BBox& cb, int dim) {
        // bvh.rs:118-130
        float pdist = p.centroid[dim] - cb.pmin[dim];
        float dist = cb.pmax[dim] - cb.pmin[dim];
        float f = 12.0f * (pdist / dist);
        size_t b = (f != f || f <= 0.0f) ? 0 : (size_t)f;  // `as usize` saturates
        if (b == 12) b = 11;
        return (int)b;
    }

    // split_equal_counts: bvh.rs:92-103
    static void equal_counts(int dim, std::vector<Info>& prims, std::vector<Info>& l,
                             std::vector<Info>& r) {
        partition_by(prims.data(), prims.size(), [dim](const Info& p) { return p.centroid[dim]; });
        size_t n = prims.size();
        l.assign(prims.begin(), prims.begin() + n / 2);
        r.assign(prims.begin() + n / 2, prims.end());
    }

    int leaf(const BBox& b, const std::vector<Info>& prims) {
        BuildNode nd{};
        nd.bounds = b;
        nd.leaf = true;
        nd.first = (uint32_t)ordered.size();
        nd.count = (uint32_t)prims.size();
        for (const Info& p : prims) ordered.push_back(p.prim);
        pool.push_back(nd);
        return (int)pool.size() - 1;
    }

    // recursive_build: bvh.rs:189-258 (prim offsets assigned in the final order
    // directly: left subtree leaves precede right subtree leaves, which is what
    // right.offset(ordered_left.len()) + append produce)
    int build(std::vector<Info>& prims) {
        BBox bbox = bbox_empty();
        for (const Info& p : prims) bbox = bbox_union(bbox, prim_world_bound(sc, sc.prims[p.prim]));
        size_t n = prims.size();
        if (n == 1) return leaf(bbox, prims);
        BBox cb = bbox_empty();
        for (const Info& p : prims) cb = bbox_union_p(cb, p.centroid);
        int dim = bbox_max_extent(cb);
        if (cb.pmin[dim] == cb.pmax[dim]) return leaf(bbox, prims);
        std::vector<Info> l, r;
        if (split == 1) {
            // split_middle: bvh.rs:86-90 (stable partition)
            float pmid = 0.5f * (cb.pmin[dim] + cb.pmax[dim]);
            for (const Info& p : prims) (p.centroid[dim] < pmid ? l : r).push_back(p);
        } else if (split == 2) {
            equal_counts(dim, prims, l, r);
        } else {
            // split_surface_area_heuristic: bvh.rs:105-186
            if (n <= 4) {
                equal_counts(dim, prims, l, r);
            } else {
                uint32_t cnt[12] = {0};
                BBox bb[12];
                for (int i = 0; i < 12; ++i) bb[i] = bbox_empty();
                for (const Info& p : prims) {
                    int b = bucket_for(p, cb, dim);
                    cnt[b] += 1;
                    bb[b] = bbox_union(bb[b], p.bounds);
                }
                float cost[11];
                for (int i = 0; i < 11; ++i) {
                    uint32_t c0 = 0, c1 = 0;
                    BBox b0 = bbox_empty(), b1 = bbox_empty();
                    for (int k = 0; k <= i; ++k) { c0 += cnt[k]; b0 = bbox_union(b0, bb[k]); }
                    for (int k = i + 1; k < 12; ++k) { c1 += cnt[k]; b1 = bbox_union(b1, bb[k]); }
                    float b0sa = (float)c0 * bbox_surface_area(b0);
                    float b1sa = (float)c1 * bbox_surface_area(b1);
                    float sc_ = 0.125f;
                    float tsa = bbox_surface_area(bbox);
                    cost[i] = sc_ * (b0sa + b1sa) / tsa;
                }
                int best = 0;
                float mc = FLT_MAX;
                for (int i = 0; i < 11; ++i)
                    if (cost[i] < mc) { best = i; mc = cost[i]; }
                size_t mcu = (mc != mc || mc <= 0.0f) ? 0 : (size_t)mc;
                if (max_prims < n || mcu < n) {
                    for (const Info& p : prims) (bucket_for(p, cb, dim) <= best ? l : r).push_back(p);
                } else {
                    return leaf(bbox, prims);
                }
            }
        }
        if (l.empty() || r.empty()) return -1;  // the reference asserts (bvh.rs:237-238)
        std::vector<Info>().swap(prims);
        int left = build(l);
        if (left < 0) return -1;
        int right = build(r);
        if (right < 0) return -1;
        BuildNode nd{};
        nd.leaf = false;
        nd.bounds = bbox_union(pool[left].bounds, pool[right].bounds);
        nd.axis = dim;
        nd.c1 = left;
        nd.c2 = right;
        pool.push_back(nd);
        return (int)pool.size() - 1;
    }

    // flatten_tree: bvh.rs:282-317
    uint32_t flatten(int ni, std::vector<PNode>& out, uint32_t offset) {
        const BuildNode& nd = pool[ni];
        if (nd.leaf) {
            out.push_back(PNode{nd.bounds, true, nd.first, nd.count});
            return offset + 1;
        }
        out.push_back(PNode{nd.bounds, false, 0, (uint32_t)nd.axis});
        uint32_t c1 = flatten(nd.c1, out, offset + 1);
        uint32_t c2 = flatten(nd.c2, out, c1);
        out[offset].off = c1;
        return c2;
    }
};

// GeometricPrimitive::intersect -> Shape::intersect (mesh.rs:212-263 / sphere.rs:135-175):
// on a hit the ray's maxt becomes t (geometric.rs:64).
static bool prim_intersect(const Scene& sc, const Prim& p, Ray& r, float* t, float* b1,
                           float* b2) {
    if (p.kind == 0) {
        const Mesh& m = sc.meshes[p.mesh];
        if (!tri_hit(m.p[p.v[0]], m.p[p.v[1]], m.p[p.v[2]], r, t, b1, b2)) return false;
        r.maxt = *t;
        return true;
    }
    const Sphere& s = sc.spheres[p.sphere];
    Ray ro = ray_to_object(s.w2o, r);
    float th, phi;
    if (!sphere_hit_obj(s, ro, &th, &phi)) return false;
    *t = th; *b1 = 0.0f; *b2 = 0.0f;
    r.maxt = th;
    return true;
}

// BVHAccelerator::intersect: bvh.rs:377-421.  nvis / ntest count the nodes
// popped and prims tested; first (may be NULL) receives the same two counts
// up to the first hit, i.e. the work of an occlusion query that stops there.
static bool bvh_intersect(const Scene& sc, Ray& r, uint32_t* prim, float* t, float* b1,
                          float* b2, uint64_t* nvis, uint64_t* ntest, uint64_t* first = nullptr) {
    if (sc.nodes.empty()) return false;
    const uint64_t v0 = *nvis, t0 = *ntest;
    V3 inv = mk(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    bool neg[3] = {inv.x < 0.0f, inv.y < 0.0f, inv.z < 0.0f};
    // the reference's Vec with capacity 64; the stack never exceeds depth + 1
    std::vector<uint32_t> todo;
    todo.reserve(64);
    todo.push_back(0);
    bool hit = false;
    while (!todo.empty()) {
        uint32_t ni = todo.back();
        todo.pop_back();
        const PNode& nd = sc.nodes[ni];
        ++*nvis;
        if (!bbox_intersect(nd.bounds, r, nullptr, nullptr)) continue;
        if (nd.leaf) {
            for (uint32_t i = 0; i < nd.meta; ++i) {
                uint32_t pi = sc.order[nd.off + i];
                ++*ntest;
                float tt, bb1, bb2;
                if (prim_intersect(sc, sc.prims[pi], r, &tt, &bb1, &bb2)) {
                    if (!hit && first) { first[0] = *nvis - v0; first[1] = *ntest - t0; }
                    hit = true; *prim = pi; *t = tt; *b1 = bb1; *b2 = bb2;
                }
            }
        } else if (neg[nd.meta]) {
            todo.push_back(ni + 1);
            todo.push_back(nd.off);
        } else {
            todo.push_back(nd.off);
            todo.push_back(ni + 1);
        }
    }
    return hit;
}

template <typename F>
static void parallel_for(uint32_t n, uint32_t nthreads, F f) {
    if (nthreads <= 1 || n < 256) {
        f(0u, n);
        return;
    }
    std::atomic<uint32_t> next{0};
    const uint32_t chunk = 256;
    std::vector<std::thread> th;
    for (uint32_t k = 0; k < nthreads; ++k)
        th.emplace_back([&]() {
            while (true) {
                uint32_t b = next.fetch_add(chunk);
                if (b >= n) break;
                f(b, std::min(n, b + chunk));
            }
        });
    for (auto& t : th) t.join();
}

// ---------------------------------------------------------------------------
// Counter-based sample generator (deviation D6).  pcg4d from M. Jarzynski and
// M. Olano, "Hash Functions for GPU Rendering", JCGT 9(3), 2020: a 4-lane
// integer hash used here as a counter-based generator keyed by (pixel,
// sample, dimension block, seed).  u01 keeps the top 24 bits, so the float is
// exact and identical on every platform.
// ---------------------------------------------------------------------------
static inline void pcg4d(uint32_t v[4]) {
    for (int i = 0; i < 4; ++i) v[i] = v[i] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
    for (int i = 0; i < 4; ++i) v[i] ^= v[i] >> 16;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
}
static inline float u01(uint32_t u) { return (float)(u >> 8) * 0x1p-24f; }
static inline void rng4(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t block,
                        float out[4]) {
    uint32_t v[4] = {pixel, sample, block, seed};
    pcg4d(v);
    for (int i = 0; i < 4; ++i) out[i] = u01(v[i]);
}

// Deterministic sine / cosine on [-pi/4, pi/4] from their Taylor series
// (terms through x^9 and x^10, truncation error < 2e-9), evaluated with basic
// operations only so every platform gives the same bits.
static inline void sincos_q(float x, float* s, float* c) {
    float x2 = x * x;
    *s = x * (1.0f + x2 * (-1.0f / 6.0f + x2 * (1.0f / 120.0f + x2 * (-1.0f / 5040.0f +
                                                                       x2 * (1.0f / 362880.0f)))));
    *c = 1.0f + x2 * (-0.5f + x2 * (1.0f / 24.0f + x2 * (-1.0f / 720.0f +
                                                         x2 * (1.0f / 40320.0f +
                                                               x2 * (-1.0f / 3628800.0f)))));
}
// Concentric square-to-disk map (P. Shirley and K. Chiu, "A Low Distortion
// Map Between Disk and Square", JGT 2(3), 1997), in the two-branch form.
static inline void concentric_disk(float u1, float u2, float* dx, float* dy) {
    float a = 2.0f * u1 - 1.0f;
    float b = 2.0f * u2 - 1.0f;
    if (a == 0.0f && b == 0.0f) { *dx = 0.0f; *dy = 0.0f; return; }
    float s, c;
    if (std::fabs(a) > std::fabs(b)) {
        sincos_q((PI_F / 4.0f) * (b / a), &s, &c);
        *dx = a * c; *dy = a * s;
    } else {
        sincos_q((PI_F / 4.0f) * (a / b), &s, &c);
        *dx = b * s; *dy = b * c;
    }
}

// ---------------------------------------------------------------------------
// Render: camera, sampler, lights, BSDFs, integrators, film
// ---------------------------------------------------------------------------
struct RGB {
    float r, g, b;
};
static inline RGB rgb(float v) { return RGB{v, v, v}; }
static inline RGB rgb_add(RGB a, RGB b) { return RGB{a.r + b.r, a.g + b.g, a.b + b.b}; }
static inline RGB rgb_mul(RGB a, RGB b) { return RGB{a.r * b.r, a.g * b.g, a.b * b.b}; }
static inline RGB rgb_muls(RGB a, float f) { return RGB{a.r * f, a.g * f, a.b * f}; }
static inline RGB rgb_divs(RGB a, float f) { return RGB{a.r / f, a.g / f, a.b / f}; }
static inline bool rgb_black(RGB a) { return a.r == 0.0f && a.g == 0.0f && a.b == 0.0f; }
// rgb_to_xyz / xyz_to_rgb: spectrum.rs:31-41
static inline void rgb_to_xyz(RGB c, float xyz[3]) {
    xyz[0] = 0.412453f * c.r + 0.357580f * c.g + 0.180423f * c.b;
    xyz[1] = 0.212671f * c.r + 0.715160f * c.g + 0.072169f * c.b;
    xyz[2] = 0.019334f * c.r + 0.119193f * c.g + 0.950227f * c.b;
}
static inline void xyz_to_rgb(const float xyz[3], float out[3]) {
    out[0] = 3.240479f * xyz[0] - 1.37150f * xyz[1] - 0.498535f * xyz[2];
    out[1] = -0.969256f * xyz[0] + 1.875991f * xyz[1] + 0.041556f * xyz[2];
    out[2] = 0.055648f * xyz[0] - 0.204043f * xyz[1] + 1.057311f * xyz[2];
}
// Spectrum::y for RGB: spectrum.rs:460-469
static inline float rgb_y(RGB c) { return 0.212671f * c.r + 0.715160f * c.g + 0.072169f * c.b; }

struct AreaTri {
    V3 p1, p2, p3, nn;
    float area;
};
struct AreaLight {
    std::vector<AreaTri> tris;
    float total;
    RGB le;
};
struct PointL {
    V3 pos;
    RGB I;
};
struct Mat {
    int kind;
    RGB r;
};

struct RenderCtx {
    Scene sc;
    const orc_render_desc* d;
    M4 r2c, c2w;
    std::vector<Mat> mats;
    std::vector<AreaLight> areas;
    std::vector<PointL> points;
    uint32_t nlights;
};

// Triangle::intersect differential geometry with the default uvs
// (mesh.rs:212-254) -> dg.nn (diff_geom.rs:52-79, no orientation flip for an
// identity-transform mesh) and dpdu.
static void tri_frame(V3 p1, V3 p2, V3 p3, V3* nn, V3* dpdu_out) {
    const float u0 = 0.0f, v0 = 0.0f, u1 = 1.0f, v1 = 0.0f, u2 = 1.0f, v2 = 1.0f;
    float du1 = u0 - u2, du2 = u1 - u2, dv1 = v0 - v2, dv2 = v1 - v2;
    V3 dp1 = sub(p1, p3), dp2 = sub(p2, p3);
    float det = du1 * dv2 - dv1 * du2;
    float inv_det = 1.0f / det;  // det == 1 for the default uvs
    V3 dpdu = mul(sub(mul(dp1, dv2), mul(dp2, dv1)), inv_det);
    V3 dpdv = mul(add(mul(dp1, -du2), mul(dp2, du1)), inv_det);
    *nn = normalize(cross(dpdu, dpdv));
    *dpdu_out = dpdu;
}

// Area light (deviation D2): one-sided diffuse emitter made of mesh triangles,
// sampled uniformly by area; triangle chosen by its area share.
static bool area_sample(const AreaLight& L, V3 p, float usel, float u1, float u2, V3* ps, V3* nl) {
    float target = usel * L.total;
    size_t k = 0;
    float acc = L.tris[0].area;
    while (k + 1 < L.tris.size() && target >= acc) { ++k; acc = acc + L.tris[k].area; }
    const AreaTri& t = L.tris[k];
    float su1 = std::sqrt(u1);
    float b1 = 1.0f - su1;
    float b2 = u2 * su1;
    float b0 = (1.0f - b1) - b2;
    *ps = add(add(mul(t.p1, b1), mul(t.p2, b2)), mul(t.p3, b0));
    *nl = t.nn;
    (void)p;
    return true;
}

struct ShadowReq {
    Ray r;
    RGB c;  // contribution if unoccluded
};

// Sample one light at hit point p with normal n; returns false when the sample
// carries nothing (pdf 0, black Li or black f: the shadow ray is not traced).
// PointLight::sample_l (point.rs:28-35, wi normalised: D5) and the area light
// (D2); VisibilityTester::segment (visibility_tester.rs:16-24); the Whitted light
// term f * Li * |wi.n| / pdf (whitted.rs:52-62).
static bool sample_light(const RenderCtx& C, uint32_t li, V3 p, float eps, V3 n, V3 wo,
                         const Mat& mat, const float u[3], ShadowReq* out) {
    RGB Li;
    V3 wi;
    float pdf;
    V3 target;
    float eps2;
    if (li < C.points.size()) {
        const PointL& pl = C.points[li];
        V3 w = sub(pl.pos, p);
        float d2 = length_sq(w);
        Li = rgb_divs(pl.I, d2);
        wi = normalize(w);
        pdf = 1.0f;
        target = pl.pos;
        eps2 = 0.0f;
    } else {
        const AreaLight& al = C.areas[li - C.points.size()];
        V3 ps, nl;
        area_sample(al, p, u[0], u[1], u[2], &ps, &nl);
        V3 w = sub(ps, p);
        float d2 = length_sq(w);
        float dist = std::sqrt(d2);
        wi = mul(w, 1.0f / dist);
        float cosl = dot(nl, neg(wi));
        if (!(cosl > 0.0f)) return false;
        Li = al.le;
        pdf = d2 / (cosl * al.total);
        target = ps;
        eps2 = 1e-3f;
    }
    if (rgb_black(Li) || pdf == 0.0f) return false;
    // BSDF::f (bsdf/mod.rs:132-149) with pbrt's flag matching (D4): only the
    // reflection lobes count when wi and wo are on the same side of ng.
    RGB f = rgb(0.0f);
    if (mat.kind == 0 && dot(wi, n) * dot(wo, n) > 0.0f) f = rgb_muls(mat.r, 1.0f / PI_F);
    if (rgb_black(f)) return false;
    float dist = length(sub(p, target));
    V3 dir = divs(sub(target, p), dist);
    out->r = Ray{p, dir, eps, (1.0f - eps2) * dist};
    out->c = rgb_divs(rgb_muls(rgb_mul(f, Li), std::fabs(dot(wi, n))), pdf);
    return true;
}

static bool trace_closest(const Scene& sc, Ray& r, uint32_t* prim, float* t, float* b1, float* b2,
                          uint64_t* st) {
    uint64_t nv = 0, nt = 0;
    bool h = bvh_intersect(sc, r, prim, t, b1, b2, &nv, &nt);
    st[0] += nv; st[1] += nt;
    return h;
}
static bool trace_any(const Scene& sc, const Ray& r0, uint64_t* st) {
    Ray r = r0;
    uint32_t p; float t, b1, b2;
    uint64_t nv = 0, nt = 0, first[2] = {0, 0};
    bool h = bvh_intersect(sc, r, &p, &t, &b1, &b2, &nv, &nt, first);
    st[5] += h ? first[0] : nv;
    st[6] += h ? first[1] : nt;
    return h;
}

struct HitGeom {
    V3 p, nn, dpdu;
    float eps;
    uint32_t mat;   // material index
    int32_t light;  // area light of an emitting face, or -1
};
// The shading geometry of a sphere hit (Sphere::intersect, sphere.rs:135-175,
// and compute_dg, helpers.rs:17-43): the object-space hit of the world ray at
// t, dp/du and dp/dv there (theta from acos, sin theta from sin, as the
// reference's libm calls), the world point o2w(p_hit), the world partials
// and dg.nn = normalize(dpdu x dpdv) flipped when reverse_orientation ^
// transform_swaps_handedness (diff_geom.rs:52-60).  dn/du, dn/dv and u, v
// feed nothing the renderer computes and are left out.
static void sphere_frame(const Sphere& s, const Ray& r, float t, V3* pw, V3* nn, V3* dpdu_w) {
    const Ray ray = ray_to_object(s.w2o, r);
    const V3 ph = add(ray.o, mul(ray.d, t));  // ray.point_at (sphere.rs:142)
    const float theta = std::acos(std::min(std::max(ph.z / s.radius, -1.0f), 1.0f));
    const float zradius = std::sqrt(ph.x * ph.x + ph.y * ph.y);
    const float inv_z = 1.0f / zradius;
    const float cos_phi = ph.x * inv_z, sin_phi = ph.y * inv_z;
    const V3 dpdu = mk(-s.phimax * ph.y, s.phimax * ph.x, 0.0f);
    const V3 dpdv = mul(mk(ph.z * cos_phi, ph.z * sin_phi, -s.radius * std::sin(theta)), s.thetamax - s.thetamin);
    *pw = xf_point(s.o2w.m, ph);
    *dpdu_w = xf_vector(s.o2w.m, dpdu);
    const V3 dpdv_w = xf_vector(s.o2w.m, dpdv);
    V3 n = normalize(cross(*dpdu_w, dpdv_w));
    if (s.ro != s.swaps) n = mul(n, -1.0f);
    *nn = n;
}
static HitGeom hit_geom(const RenderCtx& C, const Ray& r, uint32_t prim, float t) {
    const Prim& pr = C.sc.prims[prim];
    HitGeom h;
    h.eps = t * 5e-4f;  // mesh.rs:262, sphere.rs:174
    if (pr.kind == 1) {
        sphere_frame(C.sc.spheres[pr.sphere], r, t, &h.p, &h.nn, &h.dpdu);
        h.mat = C.d->sphere_mat ? C.d->sphere_mat[pr.sphere] : 0;
        h.light = -1;
        return h;
    }
    const Mesh& m = C.sc.meshes[pr.mesh];
    h.p = add(r.o, mul(r.d, t));  // Ray::point_at, ray.rs:60
    tri_frame(m.p[pr.v[0]], m.p[pr.v[1]], m.p[pr.v[2]], &h.nn, &h.dpdu);
    h.mat = C.d->face_mat[prim];
    h.light = C.d->face_light ? C.d->face_light[prim] : -1;
    return h;
}
static RGB emitted(const RenderCtx& C, const HitGeom& h, V3 wo) {
    if (h.light < 0) return rgb(0.0f);
    if (dot(h.nn, wo) > 0.0f) return C.areas[(size_t)h.light].le;
    return rgb(0.0f);
}

// Camera ray for a raster position: Camera::generate_ray_differential main ray
// (camera/mod.rs:168-195,226-236,269-270) with a static cam_to_world.
static Ray camera_ray(const RenderCtx& C, float ix, float iy) {
    V3 pc = xf_point(C.r2c, mk(ix, iy, 0.0f));
    V3 d = normalize(pc);
    V3 o = mk(0.0f, 0.0f, 0.0f);
    Ray r{xf_point(C.c2w, o), xf_vector(C.c2w, d), 0.0f, FLT_MAX};
    return r;
}

// image sample position of sample s of pixel (px,py): StratifiedSampler
// (stratified.rs:60-83, montecarlo.rs:119-128) with the counter-based jitter.
static void sample_pos(uint32_t seed, uint32_t xres, uint32_t nx, uint32_t ny, bool jitter,
                       uint32_t px, uint32_t py, uint32_t s, float* ix, float* iy) {
    float u[4];
    rng4(seed, py * xres + px, s, 0, u);
    float dx = 1.0f / (float)nx, dy = 1.0f / (float)ny;
    uint32_t sx = s % nx, sy = s / nx;
    float jx = jitter ? u[0] : 0.5f, jy = jitter ? u[1] : 0.5f;
    *ix = (((float)sx + jx) * dx) + (float)px;
    *iy = (((float)sy + jy) * dy) + (float)py;
}

// ---- LDSampler's image samples (lds.rs:47-70; sampler/utils.rs:5-32 and
// ld_shuffle_scrambled_2d :52-69), with the reference's bit tricks as they
// are (van_der_corput's last swap shifts by 2, sobol2 flips on a zero bit)
// and its RNG draws taken from the counter-based generator (D6): draw k of
// pixel p is (v0 << 32 | v1) mod (2^64 - 1) of pcg4d(p, k, LD_BLOCK, seed),
// as RNG::random_uint reduces a u64 (rng.rs:19-21).
static constexpr uint32_t LD_BLOCK = 0x4C44u;
static uint64_t ld_draw(uint32_t seed, uint32_t pixel, uint32_t k) {
    uint32_t v[4] = {pixel, k, LD_BLOCK, seed};
    pcg4d(v);
    const uint64_t u = ((uint64_t)v[0] << 32) | v[1];
    return u % 0xFFFFFFFFFFFFFFFFull;
}
static float ld_vdc(uint32_t n, uint32_t scramble) {
    n = (n << 16) | (n >> 16);
    n = ((n & 0x00ff00ffu) << 8) | ((n & 0xff00ff00u) >> 8);
    n = ((n & 0x0f0f0f0fu) << 4) | ((n & 0xf0f0f0f0u) >> 4);
    n = ((n & 0x33333333u) << 2) | ((n & 0xCCCCCCCCu) >> 2);
    n = ((n & 0x55555555u) << 2) | ((n & 0xAAAAAAAAu) >> 2);
    n ^= scramble;
    return (float)((double)((n >> 8) & 0xffffffu) / (double)(1u << 24));
}
static float ld_sobol2(uint32_t n, uint32_t scramble) {
    uint32_t s = scramble, v = 1u << 31;
    while (n != 0) {
        if ((n & 0x1u) == 0) s ^= v;
        v ^= v >> 1;
        n >>= 1;
    }
    return (float)((double)((s >> 8) & 0xFFFFFFu) / (double)(1u << 24));
}
// the spp image samples of a pixel, positions within the pixel (x, y pairs)
static void ld_pixel_points(uint32_t seed, uint32_t pixel, uint32_t n, float* pts) {
    const uint32_t sc0 = (uint32_t)ld_draw(seed, pixel, 0), sc1 = (uint32_t)ld_draw(seed, pixel, 1);
    for (uint32_t i = 0; i < n; ++i) {
        pts[2 * i] = ld_vdc(i, sc0);       // sample02 (utils.rs:35-37)
        pts[2 * i + 1] = ld_sobol2(i, sc1);
    }
    // draws 2 .. n+1 go to shuffles of one point each (no change); then the
    // points are shuffled (RNG::shuffle, rng.rs:23-32) with draws n+2 ..
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t other = i + (uint32_t)(ld_draw(seed, pixel, n + 2 + i) % (uint64_t)(n - i));
        std::swap(pts[2 * i], pts[2 * other]);
        std::swap(pts[2 * i + 1], pts[2 * other + 1]);
    }
}
static uint32_t next_pow2(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}
// radical_inverse (montecarlo.rs:8-21), in f64 as the reference computes it
static double radical_inverse(uint64_t n, uint64_t b) {
    double v = 0.0;
    uint64_t num = n;
    const double inv_base = 1.0 / (double)b;
    double aib = 1.0;
    while (num > 0) {
        const double d = (double)(num % b);
        num /= b;
        aib *= inv_base;
        v += d * aib;
    }
    return v;
}
// Lerp for f32 (utils/mod.rs:15-18): a * (1 - t) + b * t
static float lerp_f(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// HaltonSampler::get_more_samples (halton.rs:48-107) over the whole image as
// one window [0, xres) x [0, yres) (the reference's task windows depend on
// the host's core count, so the build samples the image as one window):
// sample k of the sequence, k < max(xres, yres)^2 * spp, lies at
// (lerp(0, delta, rinv(k, 3)), lerp(0, delta, rinv(k, 2))) with delta =
// max(dx, dy), and is kept when inside the image.  The frame's samples are
// grouped by the pixel they fall in, in sequence order within a pixel; every
// pixel has K slots (K = the most samples any pixel received) and slots past
// its count are empty (no ray, no film weight).
struct HaltonTable {
    uint32_t K = 0;
    std::vector<uint32_t> cnt;  // samples of each pixel
    std::vector<float> pos;     // image position of slot s of pixel p at 2 (p K + s)
};
static void halton_table(uint32_t xres, uint32_t yres, uint32_t spp, HaltonTable* T) {
    const int32_t x0 = 0, x1 = (int32_t)xres, y0 = 0, y1 = (int32_t)yres;
    const float dx = (float)(x1 - x0), dy = (float)(y1 - y0);
    const float delta = dy > dx ? dy : dx;  // dy.max(dx)
    const uint64_t side = xres > yres ? xres : yres;
    const uint64_t wanted = side * side * spp;
    std::vector<std::vector<std::pair<float, float>>> per((size_t)xres * yres);
    for (uint64_t k = 0; k < wanted; ++k) {
        const float u = (float)radical_inverse(k, 3), v = (float)radical_inverse(k, 2);
        const float ix = lerp_f((float)x0, (float)x0 + delta, u);
        const float iy = lerp_f((float)y0, (float)y0 + delta, v);
        if (ix >= (float)x1 || iy >= (float)y1) continue;
        per[(size_t)(uint32_t)iy * xres + (uint32_t)ix].push_back({ix, iy});
    }
    T->K = 0;
    for (auto& v : per) T->K = std::max<uint32_t>(T->K, (uint32_t)v.size());
    T->cnt.assign(per.size(), 0);
    T->pos.assign(per.size() * T->K * 2, 0.0f);
    for (size_t p = 0; p < per.size(); ++p) {
        T->cnt[p] = (uint32_t)per[p].size();
        for (size_t s = 0; s < per[p].size(); ++s) {
            T->pos[2 * (p * T->K + s)] = per[p][s].first;
            T->pos[2 * (p * T->K + s) + 1] = per[p][s].second;
        }
    }
}

// The sampler of a render: stratified (nx x ny, jittered), low-discrepancy
// (spp rounded up to a power of two, LDSampler::new lds.rs:17-24) or Halton
// (K slots a pixel, see HaltonTable).
struct SamplerDesc {
    int kind;  // 0 stratified, 1 low-discrepancy, 2 Halton
    uint32_t seed, xres, nx, ny, spp;
    const HaltonTable* H;
};
// false: an empty Halton slot
static bool sample_at(const SamplerDesc& S, uint32_t px, uint32_t py, uint32_t s, float* ix, float* iy) {
    if (S.kind == 0) {
        sample_pos(S.seed, S.xres, S.nx, S.ny, true, px, py, s, ix, iy);
        return true;
    }
    if (S.kind == 2) {
        const size_t p = (size_t)py * S.xres + px;
        if (s >= S.H->cnt[p]) return false;
        *ix = S.H->pos[2 * (p * S.H->K + s)];
        *iy = S.H->pos[2 * (p * S.H->K + s) + 1];
        return true;
    }
    std::vector<float> pts(2 * (size_t)S.spp);
    ld_pixel_points(S.seed, py * S.xres + px, S.spp, pts.data());
    *ix = (float)px + pts[2 * s];  // CameraSample::new(x_pos as f32 + image_samples[2i], ..) lds.rs / utils.rs:140-146
    *iy = (float)py + pts[2 * s + 1];
    return true;
}

// Direct lighting Li (deviation D1): emitted light plus, for every light,
// the average of its samples' f*Li*|cos|/pdf over unoccluded samples
// (whitted.rs:49-65 light loop; pbrt-v2 UniformSampleAllLights structure).
static RGB li_direct(const RenderCtx& C, uint32_t pixel, uint32_t s, Ray r, uint64_t* st) {
    uint32_t prim; float t, b1, b2;
    st[2] += 1;
    if (!trace_closest(C.sc, r, &prim, &t, &b1, &b2, st)) return rgb(0.0f);
    HitGeom h = hit_geom(C, r, prim, t);
    V3 wo = neg(r.d);
    RGB L = emitted(C, h, wo);
    const Mat& mat = C.mats[h.mat];
    uint32_t ns_area = C.d->light_samples;
    for (uint32_t li = 0; li < C.nlights; ++li) {
        bool is_point = li < C.points.size();
        uint32_t ns = is_point ? 1 : ns_area;
        RGB Ld = rgb(0.0f);
        for (uint32_t j = 0; j < ns; ++j) {
            float u[4] = {0, 0, 0, 0};
            if (!is_point) rng4(C.d->seed, pixel, s, 1 + (uint32_t)(li - C.points.size()) * ns_area + j, u);
            ShadowReq q;
            if (!sample_light(C, li, h.p, h.eps, h.nn, wo, mat, u, &q)) continue;
            st[3] += 1;
            if (!trace_any(C.sc, q.r, st)) Ld = rgb_add(Ld, q.c);
        }
        L = rgb_add(L, rgb_divs(Ld, (float)ns));
    }
    return L;
}

// SpecularReflection::sample_f (specular.rs:34-40) with Fresnel::noop
// (F = 1) in the BSDF's local frame: wi = (-wo.x, -wo.y, wo.z), pdf 1,
// f = F * R / |cos theta_i| (abs_cos_theta = |wi.z|, bsdf/utils.rs).
static void mirror_sample_f(V3 wo, RGB r, V3* wi, float* pdf, RGB* f) {
    *wi = mk(-wo.x, -wo.y, wo.z);
    *pdf = 1.0f;
    *f = rgb_divs(r, std::fabs(wi->z));
}

// Path Li (deviation D1): pbrt-v2 PathIntegrator with one light sampled per
// vertex (no MIS), Lambertian / mirror BSDF sampling, roulette when
// bounce > rr_start, at most max_depth bounces.
static RGB li_path(const RenderCtx& C, uint32_t pixel, uint32_t s, Ray r, uint64_t* st) {
    RGB L = rgb(0.0f), beta = rgb(1.0f);
    bool specular = false;
    uint32_t prim; float t, b1, b2;
    st[2] += 1;
    if (!trace_closest(C.sc, r, &prim, &t, &b1, &b2, st)) return L;
    for (uint32_t bounce = 0;; ++bounce) {
        HitGeom h = hit_geom(C, r, prim, t);
        V3 wo = neg(r.d);
        if (bounce == 0 || specular) L = rgb_add(L, rgb_mul(beta, emitted(C, h, wo)));
        const Mat& mat = C.mats[h.mat];
        float ua[4], ub[4];
        rng4(C.d->seed, pixel, s, 1 + 2 * bounce, ua);
        rng4(C.d->seed, pixel, s, 2 + 2 * bounce, ub);
        if (C.nlights > 0 && mat.kind == 0) {
            uint32_t ln = std::min((uint32_t)(ua[0] * (float)C.nlights), C.nlights - 1);
            ShadowReq q;
            float u3[3] = {ua[1], ua[2], ua[3]};
            if (sample_light(C, ln, h.p, h.eps, h.nn, wo, mat, u3, &q)) {
                st[3] += 1;
                if (!trace_any(C.sc, q.r, st))
                    L = rgb_add(L, rgb_mul(beta, rgb_muls(q.c, (float)C.nlights)));
            }
        }
        if (bounce == C.d->max_depth) break;
        // BSDF frame (bsdf/mod.rs:69-84): nn, tn = normalize(dpdu), sn = nn x tn
        V3 nn = h.nn;
        V3 tn = normalize(h.dpdu);
        V3 sn = cross(nn, tn);
        V3 wol = mk(dot(wo, sn), dot(wo, tn), dot(wo, nn));  // world_to_local :103-105
        V3 wil;
        RGB f;
        float pdf;
        bool spec;
        if (mat.kind == 0) {
            // cosine hemisphere sample (pbrt-v2 BxDF::Sample_f default)
            float dx, dy;
            concentric_disk(ub[0], ub[1], &dx, &dy);
            float z = std::sqrt(rmax(0.0f, (1.0f - dx * dx) - dy * dy));
            wil = mk(dx, dy, z);
            if (wol.z < 0.0f) wil.z = -wil.z;
            pdf = (wol.z * wil.z > 0.0f) ? std::fabs(wil.z) * (1.0f / PI_F) : 0.0f;
            f = rgb_muls(mat.r, 1.0f / PI_F);
            spec = false;
        } else {
            mirror_sample_f(wol, mat.r, &wil, &pdf, &f);
            spec = true;
        }
        if (rgb_black(f) || pdf == 0.0f) break;
        // local_to_world: bsdf/mod.rs:107-111
        V3 wi = mk(sn.x * wil.x + tn.x * wil.y + nn.x * wil.z,
                   sn.y * wil.x + tn.y * wil.y + nn.y * wil.z,
                   sn.z * wil.x + tn.z * wil.y + nn.z * wil.z);
        specular = spec;
        beta = rgb_mul(beta, rgb_divs(rgb_muls(f, std::fabs(dot(wi, h.nn))), pdf));
        if (bounce > C.d->rr_start) {
            float q = std::min(0.5f, rgb_y(beta));
            if (ub[2] > q) break;
            beta = rgb_divs(beta, q);
        }
        r = Ray{h.p, wi, h.eps, FLT_MAX};
        st[4] += 1;
        if (!trace_closest(C.sc, r, &prim, &t, &b1, &b2, st)) break;
    }
    return L;
}

static bool setup_render(RenderCtx& C, const orc_render_desc* d) {
    C.d = d;
    Mesh m;
    m.p.resize(d->nverts);
    for (uint32_t i = 0; i < d->nverts; ++i)
        m.p[i] = mk(d->verts[3 * i], d->verts[3 * i + 1], d->verts[3 * i + 2]);
    C.sc.meshes.push_back(std::move(m));
    C.sc.prims.resize(d->nfaces + d->nspheres);
    for (uint32_t f = 0; f < d->nfaces; ++f) {
        Prim& p = C.sc.prims[f];
        p.kind = 0; p.mesh = 0; p.sphere = 0;
        p.v[0] = d->idx[3 * f + 2]; p.v[1] = d->idx[3 * f + 1]; p.v[2] = d->idx[3 * f];
    }
    // spheres after the faces (prim ids nfaces + k), as Shape::sphere makes
    // them (main.rs:314-323, sphere.rs:28-44)
    for (uint32_t k = 0; k < d->nspheres; ++k) {
        const orc_sphere_desc& q = d->spheres[k];
        Xf o2w{m4_from(q.o2w), m4_from(q.o2w_inv)};
        C.sc.spheres.push_back(sphere_new(o2w, xf_inverse(o2w), q.reverse_orientation != 0, q.radius, q.z0, q.z1,
                                          q.phimax_deg));
        Prim& p = C.sc.prims[d->nfaces + k];
        p.kind = 1; p.mesh = 0; p.sphere = k;
        p.v[0] = p.v[1] = p.v[2] = 0;
        if (d->sphere_mat && d->sphere_mat[k] >= d->nmats) return false;
    }
    for (uint32_t i = 0; i < d->nmats; ++i)
        C.mats.push_back(Mat{(int)d->mats[4 * i], RGB{d->mats[4 * i + 1], d->mats[4 * i + 2], d->mats[4 * i + 3]}});
    for (uint32_t i = 0; i < d->n_point; ++i) {
        const float* q = d->point_lights + 6 * i;
        C.points.push_back(PointL{mk(q[0], q[1], q[2]), RGB{q[3], q[4], q[5]}});
    }
    C.areas.resize(d->n_area);
    for (uint32_t i = 0; i < d->n_area; ++i) {
        C.areas[i].le = RGB{d->area_le[3 * i], d->area_le[3 * i + 1], d->area_le[3 * i + 2]};
        C.areas[i].total = 0.0f;
    }
    if (d->face_light) {
        const Mesh& mm = C.sc.meshes[0];
        for (uint32_t f = 0; f < d->nfaces; ++f) {
            int32_t l = d->face_light[f];
            if (l < 0) continue;
            if ((uint32_t)l >= d->n_area) return false;
            const Prim& p = C.sc.prims[f];
            AreaTri t;
            t.p1 = mm.p[p.v[0]]; t.p2 = mm.p[p.v[1]]; t.p3 = mm.p[p.v[2]];
            V3 dpdu;
            tri_frame(t.p1, t.p2, t.p3, &t.nn, &dpdu);
            t.area = 0.5f * length(cross(sub(t.p2, t.p1), sub(t.p3, t.p1)));  // mesh.rs:100-103
            C.areas[(size_t)l].tris.push_back(t);
            C.areas[(size_t)l].total = C.areas[(size_t)l].total + t.area;
        }
    }
    for (auto& a : C.areas)
        if (a.tris.empty()) return false;
    C.nlights = (uint32_t)(C.points.size() + C.areas.size());
    return true;
}

}  // namespace orc

using namespace orc;

struct orc_scene {
    Scene sc;
};

// ---------------------------------------------------------------------------
// C entry points
// ---------------------------------------------------------------------------
extern "C" {

int orc_bbox_intersect(const float bmin[3], const float bmax[3], const float o[3],
                       const float d[3], float mint, float maxt, float* t0, float* t1) {
    BBox b{mk(bmin[0], bmin[1], bmin[2]), mk(bmax[0], bmax[1], bmax[2])};
    Ray r{mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]), mint, maxt};
    return bbox_intersect(b, r, t0, t1) ? 1 : 0;
}

int orc_quadratic(float a, float b, float c, float* t0, float* t1) {
    return quadratic(a, b, c, t0, t1) ? 1 : 0;
}

void orc_partition_by_i32(int32_t* v, uint32_t n) {
    partition_by(v, n, [](int32_t x) { return x; });
}
void orc_partition_by_f32(float* v, uint32_t n) {
    partition_by(v, n, [](float x) { return x; });
}

void orc_get_crop_window(uint32_t num, uint32_t count, float aspect, float out[4]) {
    crop_window(num, count, aspect, out);
}

void orc_compute_sub_window(int32_t x0, int32_t x1, int32_t y0, int32_t y1, uint32_t num,
                            uint32_t count, int32_t out[4]) {
    size_t dx = (size_t)(x1 - x0), dy = (size_t)(y1 - y0);
    float aspect = (float)dx / (float)dy;
    float w[4];
    crop_window(num, count, aspect, w);
    float psx = (float)x0, psy = (float)y0, pex = (float)x1, pey = (float)y1;
    out[0] = (int32_t)lerpf(psx, pex, w[0]);
    out[1] = (int32_t)lerpf(psx, pex, w[1]);
    out[2] = (int32_t)lerpf(psy, pey, w[2]);
    out[3] = (int32_t)lerpf(psy, pey, w[3]);
}

int orc_matrix_inverse(const float m[16], float out[16]) {
    M4 inv;
    if (!m4_inverse(m4_from(m), inv)) return -1;
    m4_store(inv, out);
    return 0;
}

void orc_stratified_1d_nojitter(float* out, uint32_t n) {
    float inv_tot = 1.0f / (float)n;
    for (uint32_t i = 0; i < n; ++i) out[i] = ((float)i + 0.5f) * inv_tot;
}
void orc_stratified_2d_nojitter(float* out, uint32_t nx, uint32_t ny) {
    float dx = 1.0f / (float)nx, dy = 1.0f / (float)ny;
    for (uint32_t y = 0; y < ny; ++y)
        for (uint32_t x = 0; x < nx; ++x) {
            uint32_t off = 2 * (y * nx + x);
            out[off] = ((float)x + 0.5f) * dx;
            out[off + 1] = ((float)y + 0.5f) * dy;
        }
}

void orc_film_extents(uint32_t xres, uint32_t yres, float xw, float yw, const float crop[4],
                      int32_t out[8]) {
    // Film::image: film.rs:77-84; get_sample_extent / get_pixel_extent
    int32_t xs = (int32_t)std::ceil((float)xres * crop[0]);
    int32_t xc = std::max((int32_t)std::ceil((float)xres * crop[1]) - xs, 1);
    int32_t ys = (int32_t)std::ceil((float)yres * crop[2]);
    int32_t yc = std::max((int32_t)std::ceil((float)yres * crop[3]) - ys, 1);
    out[0] = (int32_t)std::floor((float)xs + 0.5f - xw);
    out[1] = (int32_t)std::floor((float)xs + 0.5f + (float)xc + xw);
    out[2] = (int32_t)std::floor((float)ys + 0.5f - yw);
    out[3] = (int32_t)std::floor((float)ys + 0.5f + (float)yc + yw);
    out[4] = xs; out[5] = xs + xc; out[6] = ys; out[7] = ys + yc;
}

}  // extern "C"

// ---- Filter (filter.rs:4-124), sinc_1d (utils/mod.rs:207-217) and the
// film's weight table (film.rs:89-97).  kind 0 mean, 1 triangle, 2 gaussian
// (p0 = alpha), 3 mitchell (p0 = b, p1 = c), 4 lanczos (p0 = tau).
struct OFilter {
    int kind;
    float xw, yw, ixw, iyw;
    float alpha, exp_x, exp_y, b, c, tau;
};
static bool make_filter(int kind, float xw, float yw, float p0, float p1, OFilter* f) {
    if (kind < 0 || kind > 4 || !(xw > 0.0f) || !(yw > 0.0f)) return false;
    f->kind = kind;
    f->xw = xw; f->yw = yw;
    f->ixw = 1.0f / xw; f->iyw = 1.0f / yw;  // FilterBase::new
    f->alpha = p0;
    f->exp_x = std::exp(-p0 * xw * xw);   // Filter::gaussian: (-a * xw * xw).exp()
    f->exp_y = std::exp(-p0 * yw * yw);
    f->b = p0; f->c = p1; f->tau = p0;
    return true;
}
static float sinc_1d(float x, float tau) {
    float v = std::fabs(x);
    if (v < 1e-5f) return 1.0f;
    if (v >= 1.0f) return 0.0f;
    v = v * 3.14159265358979323846f;  // f32::consts::PI
    const float vtau = v * tau;
    const float s = std::sin(vtau) / vtau;
    return s * std::sin(v) / v;
}
static float mitchell_1d(float v, float b, float c) {
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
static float filter_eval(const OFilter& f, float x, float y) {
    switch (f.kind) {
        case 0: return 1.0f;
        case 1: {
            const float dx = (f.xw - std::fabs(x)) * f.ixw;
            const float dy = (f.yw - std::fabs(y)) * f.iyw;
            return rmax(dx, 0.0f) * rmax(dy, 0.0f);
        }
        case 2: {
            const float gx = rmax(std::exp(-f.alpha * x * x) - f.exp_x, 0.0f);
            const float gy = rmax(std::exp(-f.alpha * y * y) - f.exp_y, 0.0f);
            return gx * gy;
        }
        case 3: return mitchell_1d(x * f.ixw, f.b, f.c) * mitchell_1d(y * f.iyw, f.b, f.c);
        default: return sinc_1d(x * f.ixw, f.tau) * sinc_1d(y * f.iyw, f.tau);
    }
}
constexpr int FTD = 16;  // FILTER_TABLE_DIM (film.rs:12)
static void filter_table(const OFilter& f, float* ft) {
    for (int y = 0; y < FTD; ++y) {
        const float fy = ((float)y + 0.5f) * f.yw / (float)FTD;
        for (int x = 0; x < FTD; ++x) {
            const float fx = ((float)x + 0.5f) * f.xw / (float)FTD;
            ft[y * FTD + x] = filter_eval(f, fx, fy);
        }
    }
}
// Film::add_sample's raster extent on one axis (film.rs:197-208): pixels
// [lo, hi] of an image of n pixels; and the table offset of pixel x
// (film.rs:214-222).
static void sample_extent(float image_v, float w, int32_t n, int32_t* lo, int32_t* hi) {
    const float d = image_v - 0.5f;
    *lo = std::max(0, (int32_t)std::ceil(d - w));
    *hi = std::min(n - 1, (int32_t)std::floor(d + w));
}
static int table_offset(int32_t x, float image_v, float inv_w) {
    const float d = image_v - 0.5f;
    const float f = ((float)x - d) * inv_w * (float)FTD;
    const float a = std::floor(std::fabs(f));
    if (!(a < (float)(FTD - 1))) return a == a ? FTD - 1 : 0;  // `as usize` saturates; NaN -> 0
    return (int)a;
}

extern "C" {

uint64_t orc_ld_draw(uint32_t seed, uint32_t pixel, uint32_t k) { return ld_draw(seed, pixel, k); }
float orc_ld_vdc(uint32_t n, uint32_t scramble) { return ld_vdc(n, scramble); }
float orc_ld_sobol2(uint32_t n, uint32_t scramble) { return ld_sobol2(n, scramble); }
void orc_ld_pixel_points(uint32_t seed, uint32_t pixel, uint32_t spp, float* pts) {
    ld_pixel_points(seed, pixel, spp, pts);
}

double orc_radical_inverse(uint64_t n, uint64_t b) { return radical_inverse(n, b); }

int orc_halton_table(uint32_t xres, uint32_t yres, uint32_t spp, uint32_t* K, uint32_t* cnt, float* pos) {
    if (xres == 0 || yres == 0 || spp == 0 || !K) return -1;
    HaltonTable T;
    halton_table(xres, yres, spp, &T);
    *K = T.K;
    if (cnt) std::copy(T.cnt.begin(), T.cnt.end(), cnt);
    if (pos) std::copy(T.pos.begin(), T.pos.end(), pos);
    return 0;
}

void orc_mirror_sample_f(const float wo[3], const float r[3], float wi[3], float* pdf, float f[3]) {
    V3 w;
    RGB fr;
    mirror_sample_f(mk(wo[0], wo[1], wo[2]), RGB{r[0], r[1], r[2]}, &w, pdf, &fr);
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    f[0] = fr.r; f[1] = fr.g; f[2] = fr.b;
}

int orc_filter_eval(int kind, float xw, float yw, float p0, float p1, float x, float y, float* out) {
    OFilter f;
    if (!make_filter(kind, xw, yw, p0, p1, &f)) return -1;
    *out = filter_eval(f, x, y);
    return 0;
}
int orc_filter_table(int kind, float xw, float yw, float p0, float p1, float* out256) {
    OFilter f;
    if (!make_filter(kind, xw, yw, p0, p1, &f)) return -1;
    filter_table(f, out256);
    return 0;
}

orc_scene* orc_scene_new(void) { return new orc_scene(); }
void orc_scene_free(orc_scene* s) { delete s; }

int64_t orc_scene_add_mesh(orc_scene* s, const float* o2w16, const float* pts, uint32_t npts,
                           const uint32_t* vi, uint32_t nidx) {
    if (!s || !pts || !vi || nidx % 3 != 0) return -1;
    M4 o2w = o2w16 ? m4_from(o2w16) : m4_identity();
    Mesh m;
    m.p.resize(npts);
    for (uint32_t i = 0; i < npts; ++i)
        m.p[i] = xf_point(o2w, mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
    for (uint32_t i = 0; i < nidx; ++i)
        if (vi[i] >= npts) return -2;
    uint32_t mid = (uint32_t)s->sc.meshes.size();
    s->sc.meshes.push_back(std::move(m));
    int64_t first = (int64_t)s->sc.prims.size();
    // Mesh::refine pops index triples from the end (mesh.rs:324-335) and
    // fully_refine pops again (primitive/mod.rs:47-62): the input order is face
    // order 0..F-1 and face (a,b,c) becomes the vertex triple [c,b,a].
    for (uint32_t f = 0; f < nidx / 3; ++f) {
        Prim p;
        p.kind = 0; p.mesh = mid; p.sphere = 0;
        p.v[0] = vi[3 * f + 2]; p.v[1] = vi[3 * f + 1]; p.v[2] = vi[3 * f];
        s->sc.prims.push_back(p);
    }
    return first;
}

int64_t orc_scene_add_sphere(orc_scene* s, const float o2w_m[16], const float o2w_minv[16],
                             int ro, float radius, float z0, float z1, float phimax_deg) {
    if (!s || !(radius > 0.0f)) return -1;
    Xf o2w{m4_from(o2w_m), m4_from(o2w_minv)};
    uint32_t si = (uint32_t)s->sc.spheres.size();
    s->sc.spheres.push_back(sphere_new(o2w, xf_inverse(o2w), ro != 0, radius, z0, z1, phimax_deg));
    Prim p;
    p.kind = 1; p.mesh = 0; p.sphere = si; p.v[0] = p.v[1] = p.v[2] = 0;
    s->sc.prims.push_back(p);
    return (int64_t)s->sc.prims.size() - 1;
}

int orc_scene_build(orc_scene* s, int split, uint32_t max_prims) {
    if (!s || s->sc.prims.empty()) return -1;
    std::vector<Info> infos(s->sc.prims.size());
    for (size_t i = 0; i < infos.size(); ++i) {
        BBox b = prim_world_bound(s->sc, s->sc.prims[i]);
        infos[i].prim = (uint32_t)i;
        infos[i].bounds = b;
        infos[i].centroid = mul(add(b.pmin, b.pmax), 0.5f);  // bvh.rs:74-80
    }
    Builder B{s->sc, split, max_prims, {}, {}};
    int root = B.build(infos);
    if (root < 0) return -2;
    s->sc.nodes.clear();
    B.flatten(root, s->sc.nodes, 0);
    s->sc.order = std::move(B.ordered);
    return 0;
}

uint32_t orc_scene_num_prims(const orc_scene* s) { return (uint32_t)s->sc.prims.size(); }
uint32_t orc_scene_num_nodes(const orc_scene* s) { return (uint32_t)s->sc.nodes.size(); }

int orc_scene_nodes(const orc_scene* s, float* bounds, uint32_t* off, uint32_t* meta,
                    uint8_t* leaf, uint32_t* prim_order) {
    for (size_t i = 0; i < s->sc.nodes.size(); ++i) {
        const PNode& n = s->sc.nodes[i];
        float* b = bounds + 6 * i;
        b[0] = n.bounds.pmin.x; b[1] = n.bounds.pmin.y; b[2] = n.bounds.pmin.z;
        b[3] = n.bounds.pmax.x; b[4] = n.bounds.pmax.y; b[5] = n.bounds.pmax.z;
        off[i] = n.off; meta[i] = n.meta; leaf[i] = n.leaf ? 1 : 0;
    }
    if (prim_order)
        for (size_t i = 0; i < s->sc.order.size(); ++i) prim_order[i] = s->sc.order[i];
    return 0;
}

int orc_intersect(const orc_scene* s, uint32_t n, const float* o, const float* d, const float* mint,
                  float* maxt, uint32_t* prim, float* t, float* b1, float* b2, uint64_t* stats,
                  uint32_t nthreads) {
    if (!s) return -1;
    std::atomic<uint64_t> av{0}, at{0};
    parallel_for(n, nthreads, [&](uint32_t lo, uint32_t hi) {
        uint64_t nv = 0, nt = 0;
        for (uint32_t i = lo; i < hi; ++i) {
            Ray r{mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]), mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]),
                  mint[i], maxt[i]};
            uint32_t p = 0xFFFFFFFFu;
            float tt = 0, bb1 = 0, bb2 = 0;
            if (!bvh_intersect(s->sc, r, &p, &tt, &bb1, &bb2, &nv, &nt)) p = 0xFFFFFFFFu;
            prim[i] = p; t[i] = tt; b1[i] = bb1; b2[i] = bb2;
            maxt[i] = r.maxt;
        }
        av += nv; at += nt;
    });
    if (stats) { stats[0] = av.load(); stats[1] = at.load(); }
    return 0;
}

int orc_intersect_p(const orc_scene* s, uint32_t n, const float* o, const float* d,
                    const float* mint, const float* maxt, uint8_t* occluded, uint64_t* stats,
                    uint32_t nthreads) {
    if (!s) return -1;
    std::atomic<uint64_t> av{0}, at{0};
    parallel_for(n, nthreads, [&](uint32_t lo, uint32_t hi) {
        uint64_t nv = 0, nt = 0;
        for (uint32_t i = lo; i < hi; ++i) {
            Ray r{mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]), mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]),
                  mint[i], maxt[i]};
            uint32_t p;
            float tt, bb1, bb2;
            occluded[i] = bvh_intersect(s->sc, r, &p, &tt, &bb1, &bb2, &nv, &nt) ? 1 : 0;
        }
        av += nv; at += nt;
    });
    if (stats) { stats[0] = av.load(); stats[1] = at.load(); }
    return 0;
}

int orc_intersect_brute(const orc_scene* s, uint32_t n, const float* o, const float* d,
                        const float* mint, const float* maxt, uint32_t* prim, float* t,
                        uint32_t* ties) {
    if (!s) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        Ray r{mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]), mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]),
              mint[i], maxt[i]};
        uint32_t best = 0xFFFFFFFFu, nt = 0;
        float bt = 0.0f;
        for (uint32_t k = 0; k < s->sc.prims.size(); ++k) {
            Ray q = r;  // full [mint, maxt] for every prim: find the minimum t
            float tt, bb1, bb2;
            if (prim_intersect(s->sc, s->sc.prims[k], q, &tt, &bb1, &bb2)) {
                if (best == 0xFFFFFFFFu || tt < bt) { best = k; bt = tt; nt = 0; }
                else if (tt == bt) { nt += 1; }
            }
        }
        prim[i] = best; t[i] = bt; ties[i] = nt;
    }
    return 0;
}

int orc_hit_box_misses(const orc_scene* s, uint32_t n, const float* o, const float* d, const float* mint,
                       const float* maxt, uint32_t* out) {
    if (!s) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        Ray r{mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]), mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]), mint[i], maxt[i]};
        uint32_t cnt = 0;
        for (uint32_t k = 0; k < s->sc.prims.size(); ++k) {
            Ray q = r;
            float tt, bb1, bb2;
            if (!prim_intersect(s->sc, s->sc.prims[k], q, &tt, &bb1, &bb2)) continue;
            // the prim's own box at maxt = its hit t (bvh.rs:389-392 at pop time)
            Ray rb = r;
            rb.maxt = tt;
            if (!bbox_intersect(prim_world_bound(s->sc, s->sc.prims[k]), rb, nullptr, nullptr)) ++cnt;
        }
        out[i] = cnt;
    }
    return 0;
}

int orc_sphere_intersect(const float o2w_m[16], const float o2w_minv[16], int ro, float radius,
                         float z0, float z1, float phimax_deg, const float o[3], const float d[3],
                         float mint, float maxt, float out[16]) {
    Xf o2w{m4_from(o2w_m), m4_from(o2w_minv)};
    Sphere s = sphere_new(o2w, xf_inverse(o2w), ro != 0, radius, z0, z1, phimax_deg);
    Ray r{mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]), mint, maxt};
    Ray ray = ray_to_object(s.w2o, r);
    float thit, phi;
    if (!sphere_hit_obj(s, ray, &thit, &phi)) return 0;
    // Sphere::intersect: sphere.rs:144-174
    V3 ph = add(ray.o, mul(ray.d, thit));
    float u = phi / s.phimax;
    float theta = std::acos(std::min(std::max(ph.z / s.radius, -1.0f), 1.0f));
    float v = (theta - s.thetamin) / (s.thetamax - s.thetamin);
    float zradius = std::sqrt(ph.x * ph.x + ph.y * ph.y);
    float inv_z = 1.0f / zradius;
    float cos_phi = ph.x * inv_z, sin_phi = ph.y * inv_z;
    V3 dpdu = mk(-s.phimax * ph.y, s.phimax * ph.x, 0.0f);
    float dth = s.thetamax - s.thetamin;
    V3 dpdv = mul(mk(ph.z * cos_phi, ph.z * sin_phi, -s.radius * std::sin(theta)), dth);
    V3 d2pduu = mul(mk(ph.x, ph.y, 0.0f), -s.phimax * s.phimax);
    V3 d2pduv = mul(mk(-sin_phi, cos_phi, 0.0f), dth * ph.z * s.phimax);
    V3 d2pdvv = mul(ph, -dth * dth);
    // compute_dg: helpers.rs:17-43
    float E = dot(dpdu, dpdu), F = dot(dpdu, dpdv), G = dot(dpdv, dpdv);
    V3 N = normalize(cross(dpdu, dpdv));
    float e = dot(N, d2pduu), f = dot(N, d2pduv), g = dot(N, d2pdvv);
    float inv = 1.0f / (E * G - F * F);
    V3 dndu = add(mul(dpdu, (f * F - e * G) * inv), mul(dpdv, (e * F - f * E) * inv));
    V3 dndv = add(mul(dpdu, (g * F - f * G) * inv), mul(dpdv, (f * F - g * E) * inv));
    (void)dndu; (void)dndv;
    V3 pw = xf_point(s.o2w.m, ph);
    V3 dpduw = xf_vector(s.o2w.m, dpdu), dpdvw = xf_vector(s.o2w.m, dpdv);
    V3 nn = normalize(cross(dpduw, dpdvw));  // diff_geom.rs:52-60
    if (s.ro != s.swaps) nn = mul(nn, -1.0f);
    out[0] = thit; out[1] = thit * 5e-4f;
    out[2] = pw.x; out[3] = pw.y; out[4] = pw.z;
    out[5] = nn.x; out[6] = nn.y; out[7] = nn.z;
    out[8] = u; out[9] = v;
    out[10] = dpduw.x; out[11] = dpduw.y; out[12] = dpduw.z;
    out[13] = dpdvw.x; out[14] = dpdvw.y; out[15] = dpdvw.z;
    return 1;
}

int orc_mesh_triangle_query(const float* o2w16, const float* pts, uint32_t npts, const uint32_t* vi,
                            uint32_t nidx, uint32_t k, const float o[3], const float d[3],
                            float mint, float maxt, float* area, uint32_t v_out[3]) {
    if (nidx % 3 != 0 || k >= nidx / 3) return -1;
    M4 o2w = o2w16 ? m4_from(o2w16) : m4_identity();
    std::vector<V3> p(npts);
    for (uint32_t i = 0; i < npts; ++i) p[i] = xf_point(o2w, mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
    // refine (mesh.rs:324-335): triangle k of the refined list is face F-1-k
    uint32_t F = nidx / 3, f = F - 1 - k;
    uint32_t v[3] = {vi[3 * f + 2], vi[3 * f + 1], vi[3 * f]};
    for (int i = 0; i < 3; ++i) v_out[i] = v[i];
    V3 p1 = p[v[0]], p2 = p[v[1]], p3 = p[v[2]];
    if (area) *area = 0.5f * length(cross(sub(p2, p1), sub(p3, p1)));
    Ray r{mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]), mint, maxt};
    float t, b1, b2;
    return tri_hit(p1, p2, p3, r, &t, &b1, &b2) ? 1 : 0;
}

static bool camera_matrix(uint32_t xres, uint32_t yres, const float screen[4], const M4& proj_m,
                          const float* proj_minv, M4& r2s_m, M4& s2r_m, M4& r2c_m) {
    // Projection::new: projective.rs:48-72
    Xf cam_to_screen;
    if (proj_minv) cam_to_screen = Xf{proj_m, m4_from(proj_minv)};
    else if (!xf_from(proj_m, cam_to_screen)) return false;
    Xf s2r = xf_mul(xf_mul(xf_scale((float)xres, (float)yres, 1.0f),
                           xf_scale(1.0f / (screen[1] - screen[0]), 1.0f / (screen[2] - screen[3]), 1.0f)),
                    xf_translate(mk(-screen[0], -screen[3], 0.0f)));
    Xf r2s = xf_inverse(s2r);
    Xf r2c = xf_mul(xf_inverse(cam_to_screen), r2s);
    r2s_m = r2s.m; s2r_m = s2r.m; r2c_m = r2c.m;
    return true;
}

int orc_projection(uint32_t xres, uint32_t yres, const float proj_m[16], const float* proj_minv,
                   const float screen[4], float out[48]) {
    M4 a, b, c;
    if (!camera_matrix(xres, yres, screen, m4_from(proj_m), proj_minv, a, b, c)) return -1;
    m4_store(a, out); m4_store(b, out + 16); m4_store(c, out + 32);
    return 0;
}

// Camera::perspective: camera/mod.rs:105-121 -> Projection::new
static bool perspective_r2c(uint32_t xres, uint32_t yres, const float screen[4], float fov,
                            M4& r2c) {
    float znear = 1e-2f, zfar = 1000.0f;
    M4 p = m4_identity();
    p.m[2][2] = zfar / (zfar - znear);
    p.m[2][3] = -(zfar * znear) / (zfar - znear);
    p.m[3][2] = 1.0f;
    p.m[3][3] = 0.0f;
    Xf P;
    if (!xf_from(p, P)) return false;
    float inv_tan = 1.0f / std::tan((fov * PI_F / 180.0f) / 2.0f);
    Xf persp = xf_mul(xf_scale(inv_tan, inv_tan, 1.0f), P);
    // Projection::new with a Transform that already carries its inverse
    Xf s2r = xf_mul(xf_mul(xf_scale((float)xres, (float)yres, 1.0f),
                           xf_scale(1.0f / (screen[1] - screen[0]), 1.0f / (screen[2] - screen[3]), 1.0f)),
                    xf_translate(mk(-screen[0], -screen[3], 0.0f)));
    Xf r2s = xf_inverse(s2r);
    r2c = xf_mul(xf_inverse(persp), r2s).m;
    return true;
}

int orc_camera_raster_to_camera(uint32_t xres, uint32_t yres, const float screen[4], float fov_deg,
                                float out_m[16]) {
    M4 r2c;
    if (!perspective_r2c(xres, yres, screen, fov_deg, r2c)) return -1;
    m4_store(r2c, out_m);
    return 0;
}

int orc_camera_rays(const float r2c[16], const float c2w[16], uint32_t n, const float* image_xy,
                    float* o, float* d) {
    RenderCtx C;
    C.r2c = m4_from(r2c);
    C.c2w = m4_from(c2w);
    for (uint32_t i = 0; i < n; ++i) {
        Ray r = camera_ray(C, image_xy[2 * i], image_xy[2 * i + 1]);
        o[3 * i] = r.o.x; o[3 * i + 1] = r.o.y; o[3 * i + 2] = r.o.z;
        d[3 * i] = r.d.x; d[3 * i + 1] = r.d.y; d[3 * i + 2] = r.d.z;
    }
    return 0;
}

void orc_rng4(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t block, float out[4]) {
    rng4(seed, pixel, sample, block, out);
}

void orc_pixel_samples(uint32_t seed, uint32_t xres, uint32_t px, uint32_t py, uint32_t nx,
                       uint32_t ny, int jitter, float* xy) {
    for (uint32_t s = 0; s < nx * ny; ++s)
        sample_pos(seed, xres, nx, ny, jitter != 0, px, py, s, &xy[2 * s], &xy[2 * s + 1]);
}

// SamplerRenderer::render restated (sampler_renderer.rs:61-144,184-206) for
// rows [y0,y1): per-sample Li, then the film (Film::add_sample film.rs:192-249
// with the desc's filter and its weight table; Filter::mean(0.5, 0.5) when
// the filter widths are zero) gathered per pixel from the samples of the
// neighbour pixels that can reach it, in row-major order of those pixels and
// sample order within each -- the order the samples are added in when the
// frame's samples are added in pixel and sample order.
int orc_render(const orc_render_desc* d, uint32_t y0, uint32_t y1, uint32_t nthreads,
               float* rgb_out, float* sample_l, uint64_t* stats) {
    if (!d || y1 > d->yres || y0 >= y1 || d->nx == 0 || d->ny == 0) return -1;
    auto tb = std::chrono::steady_clock::now();
    RenderCtx C;
    if (!setup_render(C, d)) return -2;
    {
        std::vector<Info> infos(C.sc.prims.size());
        for (size_t i = 0; i < infos.size(); ++i) {
            BBox b = prim_world_bound(C.sc, C.sc.prims[i]);
            infos[i] = Info{(uint32_t)i, mul(add(b.pmin, b.pmax), 0.5f), b};
        }
        Builder B{C.sc, d->split, d->max_prims, {}, {}};
        int root = B.build(infos);
        if (root < 0) return -3;
        B.flatten(root, C.sc.nodes, 0);
        C.sc.order = std::move(B.ordered);
    }
    auto tr = std::chrono::steady_clock::now();
    if (!perspective_r2c(d->xres, d->yres, d->screen, d->fov_deg, C.r2c)) return -4;
    C.c2w = m4_from(d->cam_to_world);
    const uint32_t xres = d->xres, yres = d->yres;
    if (d->sampler_kind < 0 || d->sampler_kind > 2) return -6;
    HaltonTable H;
    if (d->sampler_kind == 2) halton_table(xres, yres, d->nx * d->ny, &H);
    const SamplerDesc S{d->sampler_kind, d->seed, xres, d->nx, d->ny,
                        d->sampler_kind == 1 ? next_pow2(d->nx * d->ny)
                                             : (d->sampler_kind == 2 ? H.K : d->nx * d->ny),
                        &H};
    const uint32_t spp = S.spp;
    OFilter F;
    if (d->filter_xw == 0.0f && d->filter_yw == 0.0f) make_filter(0, 0.5f, 0.5f, 0.0f, 0.0f, &F);
    else if (!make_filter(d->filter_kind, d->filter_xw, d->filter_yw, d->filter_p0, d->filter_p1, &F)) return -5;
    float ftab[FTD * FTD];
    filter_table(F, ftab);
    // a sample of pixel q reaches pixels within floor(w + 1/2) of q; one more
    // for the rounding of its extent
    const int32_t rx = (int32_t)std::floor(F.xw + 0.5f) + 1, ry = (int32_t)std::floor(F.yw + 0.5f) + 1;
    // rows whose samples can land in [y0,y1)
    uint32_t ya = (uint32_t)std::max(0, (int32_t)y0 - ry), yb = std::min(yres, y1 + (uint32_t)ry);
    uint32_t rows = yb - ya;
    std::vector<float> L((size_t)rows * xres * spp * 3, 0.0f);
    std::vector<uint8_t> need((size_t)rows * xres * spp, 0);
    for (uint32_t yy = ya; yy < yb; ++yy)
        for (uint32_t x = 0; x < xres; ++x)
            for (uint32_t s = 0; s < spp; ++s) {
                float ix, iy;
                if (!sample_at(S, x, yy, s, &ix, &iy)) continue;  // empty slot
                int32_t a, b;
                sample_extent(iy, F.yw, (int32_t)yres, &a, &b);
                bool in = !(b < (int32_t)y0 || a >= (int32_t)y1);
                need[((size_t)(yy - ya) * xres + x) * spp + s] = in ? 1 : 0;
            }
    std::atomic<uint64_t> st_cam{0}, st_sh{0}, st_ext{0}, st_cn{0}, st_ct{0}, st_an{0}, st_at{0};
    uint32_t nwork = rows * xres;
    parallel_for(nwork, std::max(1u, nthreads), [&](uint32_t lo, uint32_t hi) {
        uint64_t st[7] = {0, 0, 0, 0, 0, 0, 0};
        for (uint32_t w = lo; w < hi; ++w) {
            uint32_t yy = ya + w / xres, x = w % xres;
            uint32_t pixel = yy * xres + x;
            for (uint32_t s = 0; s < spp; ++s) {
                size_t k = (size_t)w * spp + s;
                if (!need[k]) continue;
                float ix, iy;
                sample_at(S, x, yy, s, &ix, &iy);
                Ray r = camera_ray(C, ix, iy);
                RGB v = d->integrator == 1 ? li_path(C, pixel, s, r, st) : li_direct(C, pixel, s, r, st);
                L[3 * k] = v.r; L[3 * k + 1] = v.g; L[3 * k + 2] = v.b;
            }
        }
        st_cam += st[2]; st_sh += st[3]; st_ext += st[4];
        st_cn += st[0]; st_ct += st[1]; st_an += st[5]; st_at += st[6];
    });
    // film: gather per pixel
    for (uint32_t y = y0; y < y1; ++y)
        for (uint32_t x = 0; x < xres; ++x) {
            float xyz[3] = {0, 0, 0};
            float wsum = 0.0f;
            for (int32_t qy = (int32_t)y - ry; qy <= (int32_t)y + ry; ++qy) {
                if (qy < 0 || qy >= (int32_t)yres) continue;
                for (int32_t qx = (int32_t)x - rx; qx <= (int32_t)x + rx; ++qx) {
                    if (qx < 0 || qx >= (int32_t)xres) continue;
                    for (uint32_t s = 0; s < spp; ++s) {
                        float ix, iy;
                        if (!sample_at(S, (uint32_t)qx, (uint32_t)qy, s, &ix, &iy)) continue;
                        int32_t xa, xb, ya_, yb_;
                        sample_extent(ix, F.xw, (int32_t)xres, &xa, &xb);
                        sample_extent(iy, F.yw, (int32_t)yres, &ya_, &yb_);
                        if ((int32_t)x < xa || (int32_t)x > xb || (int32_t)y < ya_ || (int32_t)y > yb_) continue;
                        size_t k = ((size_t)((uint32_t)qy - ya) * xres + (uint32_t)qx) * spp + s;
                        RGB v{L[3 * k], L[3 * k + 1], L[3 * k + 2]};
                        float c[3];
                        rgb_to_xyz(v, c);
                        const float wt = ftab[table_offset((int32_t)y, iy, F.iyw) * FTD +
                                              table_offset((int32_t)x, ix, F.ixw)];
                        xyz[0] = xyz[0] + wt * c[0];
                        xyz[1] = xyz[1] + wt * c[1];
                        xyz[2] = xyz[2] + wt * c[2];
                        wsum = wsum + wt;
                    }
                }
            }
            // write_image with the buffer and splat fixed (film.rs:316-354; D3)
            float out[3];
            xyz_to_rgb(xyz, out);
            if (wsum != 0.0f) {
                float inv = 1.0f / wsum;
                for (int c = 0; c < 3; ++c) out[c] = rmax(out[c] * inv, 0.0f);
            }
            float* o = rgb_out + 3 * ((size_t)(y - y0) * xres + x);
            o[0] = out[0]; o[1] = out[1]; o[2] = out[2];
        }
    if (sample_l) {
        for (uint32_t y = y0; y < y1; ++y)
            for (uint32_t x = 0; x < xres; ++x)
                for (uint32_t s = 0; s < spp; ++s) {
                    size_t k = ((size_t)(y - ya) * xres + x) * spp + s;
                    size_t o = (((size_t)(y - y0) * xres + x) * spp + s) * 3;
                    sample_l[o] = L[3 * k]; sample_l[o + 1] = L[3 * k + 1]; sample_l[o + 2] = L[3 * k + 2];
                }
    }
    auto te = std::chrono::steady_clock::now();
    if (stats) {
        stats[0] = st_cam.load(); stats[1] = st_sh.load(); stats[2] = st_ext.load();
        stats[3] = (uint64_t)std::chrono::duration_cast<std::chrono::microseconds>(tr - tb).count();
        stats[4] = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(te - tr).count();
        stats[5] = st_cn.load(); stats[6] = st_ct.load(); stats[7] = st_an.load(); stats[8] = st_at.load();
    }
    return 0;
}

}  // extern "C"
