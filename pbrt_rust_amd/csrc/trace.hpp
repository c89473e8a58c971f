// trace.hpp -- device scene view and traversal launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "pbrtx_common.hpp"

namespace px {

// Device-resident scene (all pointers in HBM).
struct SceneView {
    const RNode* rnodes;
    const QNode* qnodes;
    const PrimRec* prims;
    const SphereRec* spheres;
    const CNode* cnodes;       // compressed layout, null when the scene has none
    const CNode* ocnodes;      // compressed occlusion records (order-free queries); cnodes when none
    const PrimRec64* prims64;
    const uint32_t* order;     // input prim id of each leaf-order position
    uint32_t n_rnodes, n_qnodes, n_prims, n_spheres;
    int32_t qroot;
    int32_t ocroot;   // root ref of ocnodes
    uint32_t depth;   // levels of the reference tree
    uint32_t qdepth;  // levels of the four-wide records
    uint32_t odepth;  // levels of ocnodes
    float root_lo[3], root_hi[3];
};

enum TracePath : int { PATH_PLAIN = 0, PATH_BATCHED = 1 };

// Traversal scheme of the batched path: one stack per ray, or one stack per
// wave of 64 consecutive rays (occlusion only; used while the scene's stack
// bound fits the per-wave stack, else the per-ray kernel runs).
// SCHEME_RAY reads the compressed records when the scene has them and the
// exact 128-byte records otherwise; SCHEME_RAY_EXACT always reads the exact
// ones.
enum TraceScheme : int { SCHEME_RAY = 0, SCHEME_PACKET = 1, SCHEME_RAY_EXACT = 2 };

struct TraceTuning {
    int block = 256;          // threads per block
    int blocks_per_cu = 0;    // persistent grid: blocks per CU (0 = occupancy maximum)
    int refill_pct = 50;      // refill when fewer than this % of a wave's lanes trace
    int scheme = SCHEME_RAY;
    int stack = 16;           // LDS stack entries per lane of the per-ray kernel: 8, 16 or 24
    int lanes = 0;            // occlusion, batched: lanes per ray (0 = by launch size, 1, 4, 8 or 16)
};

// Scratch owned by the caller (per stream): spill stacks and a block of
// TRACE_COUNTER_WORDS device words for work counters and the overflow flag.
constexpr uint32_t TRACE_COUNTER_WORDS = 512;
struct TraceScratch {
    void* spill = nullptr;       // spill stack entries
    size_t spill_bytes = 0;
    uint32_t* counter = nullptr; // TRACE_COUNTER_WORDS words, device
    // a fixed-size scratch: the batched path's persistent grid is cut to the
    // blocks whose spill stacks fit spill_bytes (instead of failing)
    bool cap_grid = false;
};

// The fixed scratch of the entry point that allocates nothing: one device
// array per GPU of NOALLOC_SCRATCH_BYTES, the counter block first.
constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;
constexpr size_t NOALLOC_COUNTER_BYTES = 4096;
static_assert(TRACE_COUNTER_WORDS * 4 <= NOALLOC_COUNTER_BYTES, "counter block");
constexpr int NOALLOC_MAX_DEVICES = 64;
// *out = the current device's fixed scratch (its address looked up once per
// device); 0, or -1 with *err set.
int noalloc_scratch(TraceScratch* out, hipError_t* err);
// spill bytes one persistent block of the batched occlusion kernel needs on
// this scene (0: none)
size_t trace_any_block_spill(const SceneView& s, const TraceTuning& tun);

// Launch helpers; return hipSuccess or the launch error.
// closest: hits[i] for each ray, and rays[i].tmax = t on a hit.
// any: occluded bit i of occ (words zeroed by the launcher).
// n_dev (batched path only, may be null): the ray count is read from device
// memory when the kernel starts (queues filled by an earlier kernel); n is then
// the capacity (the number of occlusion words zeroed).
hipError_t trace_closest(const SceneView& s, Ray* rays, uint32_t n, Hit* hits, int path,
                         const TraceTuning& tun, TraceScratch* scr, hipStream_t st,
                         const uint32_t* n_dev = nullptr);
hipError_t trace_any(const SceneView& s, const Ray* rays, uint32_t n, uint32_t* occ, int path,
                     const TraceTuning& tun, TraceScratch* scr, hipStream_t st,
                     const uint32_t* n_dev = nullptr);
// The same with a byte per ray (occ8[i] = 0 or 1, every byte of the
// launch's rays written; no zeroing needed): plain stores instead of one
// atomic per occluded ray.
hipError_t trace_any_bytes(const SceneView& s, const Ray* rays, uint32_t n, uint8_t* occ8, int path,
                           const TraceTuning& tun, TraceScratch* scr, hipStream_t st,
                           const uint32_t* n_dev = nullptr);
struct CamParams;
// Closest hit of the camera rays of samples [cam.g0, cam.g0 + n), generated
// inside the (packet) traversal kernel; only the hits are written.
// hipErrorNotSupported unless camera_fusable(s, tun).
// ct (served batches of several requests): the batch's segments and their
// cameras instead of cam, whose samples [g0, g0 + n) are then unused.
struct CamTable;
// co (one camera): misses end in the kernel (zero radiance to co->miss[i], no
// hit record) and the hits' sample indices are listed at co->list, their
// number counted at co->count (cleared by the caller).
struct CamOut {
    float4* miss = nullptr;
    uint32_t* list = nullptr;
    uint32_t* count = nullptr;
};
hipError_t trace_camera(const SceneView& s, const CamParams& cam, uint32_t n, Hit* hits, const TraceTuning& tun,
                        TraceScratch* scr, hipStream_t st, const CamTable* ct = nullptr, const CamOut* co = nullptr);
bool camera_fusable(const SceneView& s, const TraceTuning& tun);

// One request chunk of a served launch (see trace_segments).
struct TraceSeg {
    Ray* rays;       // the chunk's rays (tmax updated on a closest hit)
    Hit* hits;       // closest hit: one record per ray
    uint32_t* occ;   // occlusion: bit k = ray k of the chunk (words pre-zeroed)
    uint32_t start;  // first launch-wide ray index of the chunk (ascending)
    uint32_t count;
};
// Batched-path traversal over `total` rays gathered from nseg chunks
// (d_segs in device memory); closest hit (any = false) or occlusion.
hipError_t trace_segments(const SceneView& s, bool any, const TraceSeg* d_segs, uint32_t nseg, uint32_t total,
                          const TraceTuning& tun, TraceScratch* scr, hipStream_t st);
// Number of spill bytes a launch needs for this scene and tuning.
size_t trace_spill_bytes(const SceneView& s, int path, const TraceTuning& tun);
hipError_t fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t st);
// zeroes a[0..na), b[0..nb) and c[0..nc) with one launch (null ranges of length 0)
hipError_t zero3_u32(uint32_t* a, size_t na, uint32_t* b, size_t nb, uint32_t* c, size_t nc, hipStream_t st);
// streaming 16-byte-vector copy (bytes a multiple of 16)
hipError_t copy_bytes(void* dst, const void* src, size_t bytes, hipStream_t st);
// 64-byte gather probe over a buffer of `records` (a power of two, at most
// 2^32) 64-byte records, each read once in a scattered order;
// gather_probe_threads() threads, out holds one word per thread
uint64_t gather_probe_threads();
hipError_t gather_probe(const void* buf, uint64_t records, uint32_t* out, hipStream_t st);
// fma-chain probe: fma_chain_threads() threads, 8 chains of `iters` fmas
// each; out holds one float per thread
uint64_t fma_chain_threads();
hipError_t fma_chain(float* out, uint32_t iters, hipStream_t st);
int device_cu_count();

}  // namespace px
