// bvh_build.hpp -- host-side BVH construction for the HIP ray-query path.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "pbrtx_common.hpp"

namespace px {

// Sphere input (Shape::sphere, sphere.rs:28-44): object-to-world as m and m_inv.
struct SphereIn {
    float o2w[16];
    float o2w_inv[16];
    float radius, z0, z1, phimax_deg;
    int reverse_orientation;
};

struct HostBvh {
    std::vector<RNode> rnodes;      // reference layout, depth-first
    std::vector<uint32_t> order;    // BVH-ordered prim ids (input numbering)
    std::vector<PrimRec> prims;     // prim records in BVH order
    std::vector<SphereRec> spheres;
    std::vector<QNode> qnodes;      // four-wide layout
    std::vector<CNode> cnodes;      // compressed four-wide layout (empty if not representable)
    std::vector<PrimRec64> prims64; // prim records of the compressed layout
    std::vector<CNode> ocnodes;     // compressed occlusion records, root at 0 (empty: use cnodes)
    uint32_t odepth = 0;            // record levels of the occlusion records
    int32_t qroot = 0;              // root ref of the four-wide layout
    uint32_t qdepth = 0;            // record levels of the four-wide layout
    float root_lo[3] = {0, 0, 0}, root_hi[3] = {0, 0, 0};
    uint32_t depth = 0;             // deepest node level (root = 0)
    uint32_t nfaces = 0, nspheres = 0;
    double build_ms = 0.0;
};

// BVHAccelerator::new (bvh.rs:335-362) over mesh faces followed by spheres.
// split: 0 SAH, 1 middle, 2 equal counts.  Returns false with *err on a usage
// error (the reference panics there).
bool build_bvh(const float* verts, uint32_t nverts, const uint32_t* idx, uint32_t nfaces,
               const SphereIn* spheres, uint32_t nspheres, int split, uint32_t max_prims,
               HostBvh* out, std::string* err);

}  // namespace px
