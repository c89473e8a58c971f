// camera.hpp -- camera-ray generation shared by the renderer's kernels and the
// traversal kernel that generates its camera rays itself: both evaluate the
// same operations, so the rays are bit-identical wherever they are made.
#pragma once
#include "pbrtx_common.hpp"

namespace px {

// What camera-ray generation needs (a subset of the renderer's parameters).
struct CamParams {
    float r2c[16], c2w[16];  // raster->camera (with /w), camera->world
    uint32_t xres, nx, spp, seed;
    float dx, dy;            // 1/nx, 1/ny
    uint64_t g0;             // global index of the launch's first sample
    const float2* spos;      // low-discrepancy or Halton positions within the pixel, or null
    const uint32_t* scount;  // Halton: samples in each pixel's spp slots (the rest empty), or null
    // camera-ray packets over pixel tiles (set by the packet kernel's
    // launcher): log2 of the tile width and of spp, tile_lw = 0xFF for none
    uint32_t tile_lw = 0xFFu, spp_l = 0;
};

// A served batch of several requests: batch index b belongs to segment s,
// the last whose start <= b; the segment's camera is the CamParams at
// cams + cam * stride and its first sample is g0, so b makes the ray of
// sample g0 + (b - start) of that camera.
struct CamSeg {
    uint32_t start, cam;
    uint64_t g0;
};
struct CamTable {
    const CamSeg* segs = nullptr;
    const uint8_t* cams = nullptr;
    uint32_t nsegs = 0, stride = 0;
};
__host__ __device__ __forceinline__ uint32_t cam_seg_of(const CamSeg* segs, uint32_t nsegs, uint32_t b) {
    uint32_t lo = 0, hi = nsegs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (segs[mid].start <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ApplyTransform<Point> (transform.rs:195-207) and <Vector> (:210-218)
__device__ __forceinline__ f3 xf_point(const float* m, f3 p) {
    float xt = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
    float yt = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
    float zt = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
    float w = m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15];
    if (w != 1.0f) return mk3(xt / w, yt / w, zt / w);
    return mk3(xt, yt, zt);
}
__device__ __forceinline__ f3 xf_vector(const float* m, f3 v) {
    return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
               m[8] * v.x + m[9] * v.y + m[10] * v.z);
}

// StratifiedSampler image sample of pixel (px,py), sample s (stratified.rs:60-83,
// montecarlo.rs:119-128) with the counter-based jitter.  P: any parameter
// block with seed, xres, nx, dx, dy.
// A frame sampled by LDSampler instead carries its samples' positions within
// their pixels (spos, pixel-major): ix = x + the sample's position
// (CameraSample::new(x_pos as f32 + image_samples[2i], ..), sampler/utils.rs).
template <class P>
__device__ __forceinline__ void sample_pos(const P& p, uint32_t px_, uint32_t py, uint32_t s, float* ix, float* iy) {
    if (p.spos) {
        const float2 f = p.spos[((size_t)py * p.xres + px_) * p.spp + s];
        *ix = (float)px_ + f.x;
        *iy = (float)py + f.y;
        return;
    }
    float u[4];
    rng4(p.seed, py * p.xres + px_, s, 0, u);
    uint32_t sx = s % p.nx, sy = s / p.nx;
    *ix = (((float)sx + u[0]) * p.dx) + (float)px_;
    *iy = (((float)sy + u[1]) * p.dy) + (float)py;
}

// Camera::generate_ray_differential's main ray (camera/mod.rs:168-195,
// 226-236, 269-270) with a static cam_to_world.
template <class P>
__device__ __forceinline__ void camera_ray(const P& p, float ix, float iy, Ray* r) {
    f3 pc = xf_point(p.r2c, mk3(ix, iy, 0.0f));
    f3 d = normalize3(pc);
    f3 o = xf_point(p.c2w, mk3(0.0f, 0.0f, 0.0f));
    f3 dw = xf_vector(p.c2w, d);
    *r = Ray{o.x, o.y, o.z, 0.0f, dw.x, dw.y, dw.z, FLT_MAX};
}

// The camera ray of global sample g (pixel g / spp, sample g % spp).  An
// empty Halton slot gets an empty interval (tmin 1 > tmax 0): it meets
// nothing, so its radiance is zero, and the film skips it.
template <class P>
__device__ __forceinline__ void camera_ray_of(const P& p, uint64_t g, Ray* r) {
    const uint32_t pixel = (uint32_t)(g / p.spp), s = (uint32_t)(g % p.spp);
    if (p.scount && s >= p.scount[pixel]) {
        *r = Ray{0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 1.0f, 0.0f};
        return;
    }
    float ix, iy;
    sample_pos(p, pixel % p.xres, pixel / p.xres, s, &ix, &iy);
    camera_ray(p, ix, iy, r);
}

}  // namespace px
