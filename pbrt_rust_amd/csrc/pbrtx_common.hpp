// pbrtx_common.hpp -- host/device shared definitions of the HIP ray-query path.
//
// Floating-point rule for every function here: IEEE binary32 basic operations
// in the reference's operation order (the library is compiled with
// -ffp-contract=off, no fast-math, correctly rounded divide and sqrt), so the
// host and the gfx950 device produce the reference's bits.
#pragma once
#include <stdint.h>
#include <float.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PX_HD __host__ __device__ __forceinline__
#else
#define PX_HD inline
#endif

namespace px {

struct f3 {
    float x, y, z;
};
PX_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
// geometry/vector.rs:50-91 component-wise
PX_HD f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
PX_HD f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
// Mul<f32>: vector.rs:94-99
PX_HD f3 mul3(f3 a, float f) { return mk3(a.x * f, a.y * f, a.z * f); }
PX_HD f3 neg3(f3 a) { return mk3(-a.x, -a.y, -a.z); }
// Dot: vector.rs:169-173 (left to right, no fused multiply-add)
PX_HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Cross: vector.rs:175-182
PX_HD f3 cross3(f3 a, f3 b) {
    return mk3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
PX_HD float lensq3(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
// Normalize = v * (1/length): vector.rs:116-130, normal.rs:181-186
PX_HD f3 normalize3(f3 a) {
    float l = sqrtf(lensq3(a));
    float r = 1.0f / l;
    return mul3(a, r);
}

// ---- GPU-resident BVH layouts ------------------------------------------------
// Reference layout: the flattened PackedBVHNode array (bvh.rs:260-326) as a
// 32-byte record, depth-first, first child at i+1.
//   meta = split axis (0..2) for an interior node; 3 | (count << 2) for a leaf
//   off  = second_child_offset (interior) or first ordered-prim position (leaf)
struct RNode {
    float lo[3];
    float hi[3];
    uint32_t off;
    uint32_t meta;
};
static_assert(sizeof(RNode) == 32, "RNode must be 32 bytes");

// Four-wide layout: one 128-byte record (one cache line) per two levels of the
// reference tree.  For an interior node A with children B (= A+1) and C
// (= A's second child), the slots are [B0, B1, C0, C1]: the children of B and
// of C; a leaf B (or C) takes the first slot of its half and leaves the
// second EMPTY.  Boxes are stored per axis so four boxes load with six 16-byte
// loads.  refs: >= 0 another record, QEMPTY no child, otherwise ~first
// position of a leaf's prims in the ordered prim array.  meta: split axis of
// A in bits 0-1, of B in bits 2-3, of C in bits 4-5 -- the three axis signs
// that fix the reference's near-first visiting order of the four slots.
struct QNode {
    float lx[4], hx[4], ly[4], hy[4], lz[4], hz[4];
    int32_t ref[4];
    uint32_t meta;
    uint32_t pad[3];
};
static_assert(sizeof(QNode) == 128, "QNode must be 128 bytes");
static const int32_t QEMPTY = (int32_t)0x80000000;

// Primitive record in BVH leaf order, 48 bytes: triangle p1, e1 = p2-p1,
// e2 = p3-p1 (mesh.rs:45-46, precomputed on the host: the same floats) with
// (p1,p2,p3) the refined vertex triple [c,b,a] of face (a,b,c)
// (mesh.rs:329-332).  w[10] = input prim id, w[11] = kind (0 triangle,
// 1 sphere: w[9] = sphere record index).
struct PrimRec {
    float w[12];
};
static_assert(sizeof(PrimRec) == 48, "PrimRec must be 48 bytes");

// Compressed four-wide layout: the record tree of QNode with each slot box
// quantized OUTWARD to 8 bits against the union box of the record:
//   bound = o + q * 2^e   (one rounding: q * 2^e is exact)
// with the origin o = the union's low corner and one power-of-two scale per
// axis.  The host picks every q so that the decoded float box contains the
// exact slot box, so a slot test may pass where the exact one fails but never
// the reverse; the exact box of every leaf is tested when the leaf is popped
// (it travels in the PrimRec64 records), which keeps results exact.  64 bytes
// = four 16-byte loads per visit instead of eight.
//   emeta: biased exponent byte of the x, y, z scales in bits 0-23 (scale =
//   the float with that exponent field), QNode.meta in bits 24-29.
//   q: lx[4], hx[4], ly[4], hy[4], lz[4], hz[4].
struct CNode {
    float ox, oy, oz;
    uint32_t emeta;
    uint8_t q[24];
    int32_t ref[4];
    uint32_t pad[2];
};
static_assert(sizeof(CNode) == 64, "CNode must be 64 bytes");

// Primitive record of the compressed layout, 64 bytes, BVH leaf order:
// w[0..8] as PrimRec (a sphere keeps its record index in w[0]), w[9..14] the
// exact box of the record's leaf (lo xyz, hi xyz), w[15] = kind | 0x100 on
// the leaf's last record.  The input prim id is looked up in the ordered id
// array only for a hit.
struct PrimRec64 {
    float w[16];
};
static_assert(sizeof(PrimRec64) == 64, "PrimRec64 must be 64 bytes");

// Sphere (sphere.rs:16-44): world->object matrix rows, radius, z range, phi_max.
struct SphereRec {
    float w2o[16];
    float radius, zmin, zmax, phimax;
};

// Ray / hit records of the C ABI (32 and 16 bytes).
struct Ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
};
struct Hit {
    uint32_t prim;
    float t, b1, b2;
};
static const uint32_t NO_HIT = 0xFFFFFFFFu;

// ---- counter-based sample generator -----------------------------------------
// pcg4d: M. Jarzynski and M. Olano, "Hash Functions for GPU Rendering",
// JCGT 9(3), 2020.  Keyed by (pixel, sample, dimension block, seed); the top
// 24 bits give an exact float in [0,1).
PX_HD void pcg4d(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
    a = a * 1664525u + 1013904223u;
    b = b * 1664525u + 1013904223u;
    c = c * 1664525u + 1013904223u;
    d = d * 1664525u + 1013904223u;
    a += b * d; b += c * a; c += a * b; d += b * c;
    a ^= a >> 16; b ^= b >> 16; c ^= c >> 16; d ^= d >> 16;
    a += b * d; b += c * a; c += a * b; d += b * c;
}
PX_HD float u01(uint32_t u) { return (float)(u >> 8) * 0x1p-24f; }
PX_HD void rng4(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t block, float out[4]) {
    uint32_t a = pixel, b = sample, c = block, d = seed;
    pcg4d(a, b, c, d);
    out[0] = u01(a); out[1] = u01(b); out[2] = u01(c); out[3] = u01(d);
}

// ---- LDSampler's image samples (lds.rs:47-70; sampler/utils.rs:5-32 and
// ld_shuffle_scrambled_2d :52-69) --------------------------------------------
// The bit tricks as the reference has them: van_der_corput's last swap
// shifts by 2 (bits move or drop), and sobol2 flips on a zero bit.  The
// reference draws its scrambles and shuffles from a seeded RNG stream
// (RNG::random_uint, rng.rs:19-21: a u64 reduced mod 2^64 - 1); here draw k
// of a pixel comes from the counter-based generator instead.
constexpr uint32_t LD_BLOCK = 0x4C44u;
PX_HD uint64_t ld_draw(uint32_t seed, uint32_t pixel, uint32_t k) {
    uint32_t a = pixel, b = k, c = LD_BLOCK, d = seed;
    pcg4d(a, b, c, d);
    return (((uint64_t)a << 32) | b) % 0xFFFFFFFFFFFFFFFFull;
}
PX_HD float ld_vdc(uint32_t n, uint32_t scramble) {
    n = (n << 16) | (n >> 16);
    n = ((n & 0x00ff00ffu) << 8) | ((n & 0xff00ff00u) >> 8);
    n = ((n & 0x0f0f0f0fu) << 4) | ((n & 0xf0f0f0f0u) >> 4);
    n = ((n & 0x33333333u) << 2) | ((n & 0xCCCCCCCCu) >> 2);
    n = ((n & 0x55555555u) << 2) | ((n & 0xAAAAAAAAu) >> 2);
    n ^= scramble;
    return (float)((n >> 8) & 0xffffffu) * 0x1p-24f;  // k / 2^24, exact
}
PX_HD float ld_sobol2(uint32_t n, uint32_t scramble) {
    uint32_t s = scramble, v = 1u << 31;
    for (; n != 0; n >>= 1) {
        if ((n & 1u) == 0) s ^= v;
        v ^= v >> 1;
    }
    return (float)((s >> 8) & 0xffffffu) * 0x1p-24f;
}
PX_HD uint32_t next_pow2(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

static const float PI_F = 3.14159265358979323846f;

// sine / cosine on [-pi/4, pi/4] from their Taylor series through x^9 / x^10
// (truncation error below 2e-9), basic operations only: identical bits on
// host and device.
PX_HD void sincos_q(float x, float* s, float* c) {
    float x2 = x * x;
    *s = x * (1.0f + x2 * (-1.0f / 6.0f + x2 * (1.0f / 120.0f + x2 * (-1.0f / 5040.0f +
                                                                       x2 * (1.0f / 362880.0f)))));
    *c = 1.0f + x2 * (-0.5f + x2 * (1.0f / 24.0f + x2 * (-1.0f / 720.0f +
                                                         x2 * (1.0f / 40320.0f +
                                                               x2 * (-1.0f / 3628800.0f)))));
}
// Concentric square-to-disk map (Shirley and Chiu, JGT 2(3), 1997), two-branch form.
PX_HD void concentric_disk(float u1, float u2, float* dx, float* dy) {
    float a = 2.0f * u1 - 1.0f;
    float b = 2.0f * u2 - 1.0f;
    if (a == 0.0f && b == 0.0f) { *dx = 0.0f; *dy = 0.0f; return; }
    float s, c;
    if (fabsf(a) > fabsf(b)) {
        sincos_q((PI_F / 4.0f) * (b / a), &s, &c);
        *dx = a * c; *dy = a * s;
    } else {
        sincos_q((PI_F / 4.0f) * (a / b), &s, &c);
        *dx = b * s; *dy = b * c;
    }
}

}  // namespace px
