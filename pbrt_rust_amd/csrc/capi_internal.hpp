// capi_internal.hpp -- state shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "bvh_build.hpp"
#include "trace.hpp"

namespace px {

struct DeviceScene {
    SceneView view;
    void *d_rnodes, *d_qnodes, *d_prims, *d_spheres, *d_cnodes, *d_prims64, *d_order, *d_ocnodes;
    uint64_t bytes;
};

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int stream_scratch(hipStream_t st, size_t spill, TraceScratch** out);
// Frees the scratch of a stream about to be destroyed (its work finished).
void release_stream_scratch(hipStream_t st);
// the render description's split value for a tree built on the device
constexpr int32_t SPLIT_GPU_LBVH = 3;
// Builds a linear BVH over mesh faces then spheres on the device, straight
// into S (bvh_gpu.hip); *build_ms = its wall time.
int gpu_build_scene(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t nfaces, const SphereIn* sph,
                    uint32_t nsph, DeviceScene* S, double* build_ms);
int scene_from_host(const HostBvh& h, DeviceScene* s);
void scene_free(DeviceScene* s);
int run_closest(const SceneView& v, Ray* rays, uint32_t n, Hit* hits, int path, hipStream_t st);
int run_any(const SceneView& v, const Ray* rays, uint32_t n, uint32_t* occ, int path, hipStream_t st);
extern TraceTuning g_tuning[2][3];  // [path][0 closest hit, 1 occlusion, 2 camera rays]
extern uint32_t g_tuning_gen;       // bumped by every tuning change
extern bool g_camera_auto;          // camera rays: scheme picked by samples per pixel
TraceTuning camera_tuning(int path, uint32_t spp);

}  // namespace px

// the opaque handle of the C ABI
struct pbrtx_scene {
    px::DeviceScene s;
};
