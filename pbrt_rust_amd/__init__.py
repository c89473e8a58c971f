"""pbrt_rust_amd -- MI355X (gfx950) ray-query and sampling path of pbrt_rust.

The compute runs in hand-written HIP kernels inside libpbrtx_gpu.so (C ABI:
include/pbrtx_gpu.h); this package is the host-side mirror of the reference's
aggregate / renderer interfaces.
"""
from ._lib import (HIT_DTYPE, NO_HIT, RAY_DTYPE, DeviceBuffer, PbrtxError, PinnedArray, device_name, lib,
                   synchronize)
from .api import (Aggregate, BVHAccelerator, DirectLightingIntegrator, Filter, HaltonSampler, LDSampler,
                  PathIntegrator, Primitive,
                  RayBatch,
                  RayServer, RenderServer, SamplerRenderer, Sphere, StratifiedSampler, Transform, TriangleMesh,
                  encode8, reset_tuning, set_tuning, write_image)

__all__ = [
    "Aggregate", "BVHAccelerator", "DirectLightingIntegrator", "Filter", "HaltonSampler", "LDSampler", "PathIntegrator", "SamplerRenderer",
    "StratifiedSampler", "DeviceBuffer", "HIT_DTYPE", "NO_HIT", "PbrtxError", "PinnedArray", "Primitive",
    "RAY_DTYPE", "RayBatch", "RayServer", "RenderServer", "Sphere", "Transform", "TriangleMesh", "device_name", "encode8", "lib", "reset_tuning", "set_tuning",
    "synchronize", "write_image",
]
