/*
 * pbrt_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * C entry points of the CPU oracle: a restatement, operation for operation, of
 * pbrt_rust's ray-query and sampling path (BVH build and traversal, Triangle /
 * Sphere intersection, perspective camera, stratified sampling, visibility
 * segments, direct-lighting and path Li, box-filtered film).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library,
 * and only as the checker; the product library never links or calls it.
 *
 * Conventions: every function returns 0 on success and a negative code on a
 * usage error (it never aborts); float arrays are row-major, xyz interleaved.
 */
#ifndef PBRT_ORACLE_H
#define PBRT_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- element functions (golden-vector checks) ---------------------------- */
/* BBox::intersect, bbox.rs:185-209.  returns 1 on hit (t0,t1 written), 0 on miss */
int orc_bbox_intersect(const float bmin[3], const float bmax[3], const float o[3],
                       const float d[3], float mint, float maxt, float* t0, float* t1);
/* utils::quadratic, utils/mod.rs:56-92.  returns 1 with roots, 0 for None */
int orc_quadratic(float a, float b, float c, float* t0, float* t1);
/* utils::partition_by on an int / float slice, utils/mod.rs:112-169 */
void orc_partition_by_i32(int32_t* v, uint32_t n);
void orc_partition_by_f32(float* v, uint32_t n);
/* utils::get_crop_window, utils/mod.rs:183-205 -> out[4] = tx0,tx1,ty0,ty1 */
void orc_get_crop_window(uint32_t num, uint32_t count, float aspect, float out[4]);
/* SamplerBase::compute_sub_window, sampler/base.rs:29-48 */
void orc_compute_sub_window(int32_t x0, int32_t x1, int32_t y0, int32_t y1,
                            uint32_t num, uint32_t count, int32_t out[4]);
/* Matrix4x4 LU inverse, matrix4x4.rs:66-160.  returns -1 when singular */
int orc_matrix_inverse(const float m[16], float out[16]);
/* montecarlo::stratified_sample_{1d,2d} with jitter off, montecarlo.rs:106-128 */
void orc_stratified_1d_nojitter(float* out, uint32_t n);
void orc_stratified_2d_nojitter(float* out, uint32_t nx, uint32_t ny);
/* Film::image extents (film.rs:77-122, get_sample_extent / get_pixel_extent)
   for a Filter::mean(xw,yw); out = sample x0,x1,y0,y1, pixel x0,x1,y0,y1 */
void orc_film_extents(uint32_t xres, uint32_t yres, float xw, float yw,
                      const float crop[4], int32_t out[8]);

/* ---- scene + BVH ------------------------------------------------------------ */
typedef struct orc_scene orc_scene;
orc_scene* orc_scene_new(void);
void orc_scene_free(orc_scene* s);
/* Shape::triangle_mesh + Primitive::fully_refine: o2w may be NULL (identity).
   Faces become prims in face order (primitive/mod.rs:47-62, mesh.rs:324-338).
   returns the first prim index of the mesh, or <0 */
int64_t orc_scene_add_mesh(orc_scene* s, const float* o2w16, const float* pts,
                           uint32_t npts, const uint32_t* vi, uint32_t nidx);
/* Shape::sphere (sphere.rs:28-44) with o2w / w2o given as 4x4 matrices m, m_inv */
int64_t orc_scene_add_sphere(orc_scene* s, const float o2w_m[16], const float o2w_minv[16],
                             int reverse_orientation, float radius, float z0, float z1,
                             float phimax_deg);
/* BVHAccelerator::new, bvh.rs:335-362.  split: 0 sah, 1 middle, 2 equal */
int orc_scene_build(orc_scene* s, int split, uint32_t max_prims);
uint32_t orc_scene_num_prims(const orc_scene* s);
uint32_t orc_scene_num_nodes(const orc_scene* s);
/* flattened PackedBVHNode array (bvh.rs:260-326): bounds[6n] (pmin,pmax),
   off[n] (second_child_offset or prim_offset), meta[n] (axis, or nprims),
   leaf[n] (1 = leaf) ; prim_order[num_prims] = BVH-ordered prim indices */
int orc_scene_nodes(const orc_scene* s, float* bounds, uint32_t* off, uint32_t* meta,
                    uint8_t* leaf, uint32_t* prim_order);

/* BVHAccelerator::intersect (bvh.rs:375-422) for n rays (SoA).  maxt is in/out
   (the hit t becomes the new maxt, geometric.rs:64).  prim = prim index or
   0xFFFFFFFF on a miss.  stats (may be NULL): [0] nodes visited, [1] prims tested */
int orc_intersect(const orc_scene* s, uint32_t n, const float* o, const float* d,
                  const float* mint, float* maxt, uint32_t* prim, float* t, float* b1,
                  float* b2, uint64_t* stats, uint32_t nthreads);
/* Intersectable::intersect_p default (intersection.rs:63-65): a full closest-hit
   traversal whose is_some() is the occlusion bit.  occluded[i] in {0,1} */
int orc_intersect_p(const orc_scene* s, uint32_t n, const float* o, const float* d,
                    const float* mint, const float* maxt, uint8_t* occluded,
                    uint64_t* stats, uint32_t nthreads);
/* brute force over all prims in input order, same tri / sphere tests (checker
   for the BVH; ties resolved to the later prim as the leaf loop would) */
int orc_intersect_brute(const orc_scene* s, uint32_t n, const float* o, const float* d,
                        const float* mint, const float* maxt, uint32_t* prim, float* t,
                        uint32_t* ties);
/* For each ray: how many prims it hits (full interval) whose own box the
   slab test rejects at maxt = that hit's t -- rays where the traversal's
   pop-time box test at the current maxt can cull a hit, so that its result
   depends on the order in which it meets the tree's nodes. */
int orc_hit_box_misses(const orc_scene* s, uint32_t n, const float* o, const float* d, const float* mint,
                       const float* maxt, uint32_t* out);
/* Sphere::intersect full hit record (sphere.rs:135-175 + helpers.rs:17-43):
   out[0]=t, [1]=ray_epsilon, [2..4]=p, [5..7]=nn, [8]=u, [9]=v, [10..12]=dpdu,
   [13..15]=dpdv.  returns 1 on hit */
int orc_sphere_intersect(const float o2w_m[16], const float o2w_minv[16], int ro,
                         float radius, float z0, float z1, float phimax_deg,
                         const float o[3], const float d[3], float mint, float maxt,
                         float out[16]);
/* Triangle::intersect_p on one triangle of a mesh (mesh.rs:41-72,207-210),
   triangle k of the refined list (reversed order, mesh.rs:324-338);
   area of triangle k (mesh.rs:100-103); refined vertex triple of triangle k */
int orc_mesh_triangle_query(const float* o2w16, const float* pts, uint32_t npts,
                            const uint32_t* vi, uint32_t nidx, uint32_t k,
                            const float o[3], const float d[3], float mint, float maxt,
                            float* area, uint32_t v_out[3]);

/* ---- camera ----------------------------------------------------------------- */
/* Camera::perspective + Projection::new (camera/mod.rs:105-135,
   projective.rs:48-72): out raster_to_camera m[16]. */
int orc_camera_raster_to_camera(uint32_t xres, uint32_t yres, const float screen[4],
                                float fov_deg, float out_m[16]);
/* Projection::new with an arbitrary projective transform proj given as m and
   m_inv (proj_minv NULL: the inverse comes from the LU, Transform::from):
   out[0..15] raster_to_screen.m, out[16..31] screen_to_raster.m,
   out[32..47] raster_to_camera.m */
int orc_projection(uint32_t xres, uint32_t yres, const float proj_m[16],
                   const float* proj_minv, const float screen[4], float out[48]);
/* Camera::generate_ray_differential main ray (camera/mod.rs:168-195,212-271)
   for n raster positions, cam_to_world given as m[16]. */
int orc_camera_rays(const float r2c[16], const float c2w[16], uint32_t n,
                    const float* image_xy, float* o, float* d);

/* ---- sampler + renderer ------------------------------------------------------ */
/* counter-based sample value u in [0,1): block of 4 dims */
void orc_rng4(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t block, float out[4]);
/* StratifiedSampler image sample positions for pixel (px,py), all nx*ny samples
   (stratified.rs:60-127 with the counter-based generator) -> xy[2*spp] */
void orc_pixel_samples(uint32_t seed, uint32_t xres, uint32_t px, uint32_t py,
                       uint32_t nx, uint32_t ny, int jitter, float* xy);

/* Render scene description (plain arrays; shared layout is documented in the
   product header too).  materials: 4 floats each (kind 0 = matte/Lambertian,
   1 = mirror; rgb); face_light: area-light id per face or -1; area_le: 3 floats
   per area light; point_lights: 6 floats each (pos, I). */
/* a sphere (Shape::sphere, sphere.rs:28-44): object-to-world m and m_inv
   (row-major), radius, z range, phi max in degrees, reverse orientation */
typedef struct orc_sphere_desc {
    float o2w[16], o2w_inv[16];
    float radius, z0, z1, phimax_deg;
    int32_t reverse_orientation;
} orc_sphere_desc;

typedef struct orc_render_desc {
    const float* verts; uint32_t nverts;
    const uint32_t* idx; uint32_t nfaces;
    const uint32_t* face_mat; const float* mats; uint32_t nmats;
    const int32_t* face_light; const float* area_le; uint32_t n_area;
    const float* point_lights; uint32_t n_point;
    float cam_to_world[16]; float fov_deg; float screen[4];
    uint32_t xres, yres, nx, ny, seed;
    int32_t integrator;   /* 0 = direct lighting, 1 = path */
    uint32_t light_samples; /* direct: samples per area light */
    uint32_t max_depth;     /* path */
    uint32_t rr_start;      /* path: roulette when bounce > rr_start */
    int32_t split; uint32_t max_prims;
    /* film filter (filter.rs): kind 0 mean, 1 triangle, 2 gaussian (p0 =
       alpha), 3 mitchell (p0 = b, p1 = c), 4 lanczos (p0 = tau); widths 0
       mean Filter::mean(0.5, 0.5) */
    int32_t filter_kind; float filter_xw, filter_yw, filter_p0, filter_p1;
    /* 0 StratifiedSampler (nx x ny, jittered), 1 LDSampler with nx*ny samples
       per pixel rounded up to a power of two, 2 HaltonSampler with nx*ny
       samples per pixel over the image as one window (K slots a pixel, see
       orc_halton_table) */
    int32_t sampler_kind;
    /* spheres after the faces (prim ids nfaces + k), each with a material;
       not emitters */
    const orc_sphere_desc* spheres; uint32_t nspheres;
    const uint32_t* sphere_mat;
} orc_render_desc;

/* LDSampler's pieces (sampler/utils.rs:5-32, lds.rs): van_der_corput and
   sobol2 as the reference has them, and a pixel's spp image-sample
   positions within the pixel (x, y pairs) with the counter-based draws. */
uint64_t orc_ld_draw(uint32_t seed, uint32_t pixel, uint32_t k);
float orc_ld_vdc(uint32_t n, uint32_t scramble);
float orc_ld_sobol2(uint32_t n, uint32_t scramble);
void orc_ld_pixel_points(uint32_t seed, uint32_t pixel, uint32_t spp, float* pts);

/* Renders rows [y0,y1) of the image: SamplerRenderer::render restated with the
   intended semantics (sampler_renderer.rs:61-144,184-206; deviations D1-D7).
   rgb_out: xres*(y1-y0)*3 floats (final film rgb of those rows).
   sample_l (may be NULL): per (pixel,sample) radiance, 3 floats each.
   stats (9 entries): [0] camera rays, [1] shadow rays, [2] extension rays,
   [3] build ms x1000, [4] trace+shade ns, [5] nodes visited and [6] prims
   tested by closest-hit queries (camera and extension rays), [7] nodes
   visited and [8] prims tested by occlusion queries up to their first hit
   (the work of an early-exit any-hit query).  nthreads >= 1. */
/* radical_inverse (montecarlo.rs:8-21) */
double orc_radical_inverse(uint64_t n, uint64_t b);
/* HaltonSampler (halton.rs:48-107) over the whole image: *K = the most
   samples any pixel received; cnt[xres*yres] = each pixel's samples; pos[2 *
   (pixel*K + s)] = image position of its s-th sample in sequence order (pass
   NULL to get K first). */
int orc_halton_table(uint32_t xres, uint32_t yres, uint32_t spp, uint32_t* K, uint32_t* cnt, float* pos);
/* SpecularReflection::sample_f (specular.rs:34-40) with Fresnel::noop, in the
   BSDF's local frame: wi, pdf and f for reflectance r. */
void orc_mirror_sample_f(const float wo[3], const float r[3], float wi[3], float* pdf, float f[3]);
/* Filter::evaluate (filter.rs:89-124) and the film's 16x16 weight table
   (film.rs:89-97) of a filter given as in orc_render_desc. */
int orc_filter_eval(int kind, float xw, float yw, float p0, float p1, float x, float y, float* out);
int orc_filter_table(int kind, float xw, float yw, float p0, float p1, float* out256);

int orc_render(const orc_render_desc* desc, uint32_t y0, uint32_t y1, uint32_t nthreads,
               float* rgb_out, float* sample_l, uint64_t* stats);

#ifdef __cplusplus
}
#endif
#endif
