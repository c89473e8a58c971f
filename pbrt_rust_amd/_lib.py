"""ctypes binding of libpbrtx_gpu.so (the C ABI in include/pbrtx_gpu.h).

The library is built in-tree by pbrt_rust_amd.build; importing this module
fails loudly when it is missing -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libpbrtx_gpu.so")

RAY_DTYPE = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("tmin", "<f4"),
                      ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<u4"), ("t", "<f4"), ("b1", "<f4"), ("b2", "<f4")])
NO_HIT = 0xFFFFFFFF


class PbrtxError(RuntimeError):
    pass


class SphereDesc(ctypes.Structure):
    _fields_ = [("o2w", ctypes.c_float * 16), ("o2w_inv", ctypes.c_float * 16),
                ("radius", ctypes.c_float), ("z0", ctypes.c_float), ("z1", ctypes.c_float),
                ("phimax_deg", ctypes.c_float), ("reverse_orientation", ctypes.c_int32)]


class RenderDesc(ctypes.Structure):
    """pbrtx_render_desc (include/pbrtx_gpu.h)."""
    _fields_ = [
        ("verts", ctypes.c_void_p), ("nverts", ctypes.c_uint32),
        ("idx", ctypes.c_void_p), ("nfaces", ctypes.c_uint32),
        ("face_mat", ctypes.c_void_p), ("mats", ctypes.c_void_p), ("nmats", ctypes.c_uint32),
        ("face_light", ctypes.c_void_p), ("area_le", ctypes.c_void_p), ("n_area", ctypes.c_uint32),
        ("point_lights", ctypes.c_void_p), ("n_point", ctypes.c_uint32),
        ("cam_to_world", ctypes.c_float * 16), ("fov_deg", ctypes.c_float),
        ("screen", ctypes.c_float * 4),
        ("xres", ctypes.c_uint32), ("yres", ctypes.c_uint32), ("nx", ctypes.c_uint32),
        ("ny", ctypes.c_uint32), ("seed", ctypes.c_uint32),
        ("integrator", ctypes.c_int32), ("light_samples", ctypes.c_uint32),
        ("max_depth", ctypes.c_uint32), ("rr_start", ctypes.c_uint32),
        ("split", ctypes.c_int32), ("max_prims", ctypes.c_uint32),
        ("sampler", ctypes.c_int32),
        ("spheres", ctypes.c_void_p), ("nspheres", ctypes.c_uint32),
        ("sphere_mat", ctypes.c_void_p),
    ]


class RenderStats(ctypes.Structure):
    """pbrtx_render_stats (include/pbrtx_gpu.h)."""
    _fields_ = [("camera_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
                ("extension_rays", ctypes.c_uint64), ("ms_total", ctypes.c_double),
                ("ms_trace_closest", ctypes.c_double), ("ms_trace_any", ctypes.c_double),
                ("ms_other", ctypes.c_double), ("build_ms", ctypes.c_double),
                ("launches_closest", ctypes.c_uint32), ("launches_any", ctypes.c_uint32)]


class ServerConfig(ctypes.Structure):
    """pbrtx_server_config (include/pbrtx_gpu.h)."""
    _fields_ = [("arena_bytes", ctypes.c_uint64), ("scene_budget_bytes", ctypes.c_uint64),
                ("batch_rays", ctypes.c_uint32), ("mode", ctypes.c_int32)]


class RequestStats(ctypes.Structure):
    """pbrtx_request_stats (include/pbrtx_gpu.h)."""
    _fields_ = [("n_rays", ctypes.c_uint64), ("status", ctypes.c_int32), ("launches", ctypes.c_uint32),
                ("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double),
                ("first_launch_ms", ctypes.c_double), ("done_ms", ctypes.c_double)]


class ServerCounters(ctypes.Structure):
    """pbrtx_server_counters (include/pbrtx_gpu.h)."""
    _fields_ = [("rays_done", ctypes.c_uint64), ("requests_done", ctypes.c_uint64), ("steps", ctypes.c_uint64),
                ("launches", ctypes.c_uint64), ("busy_ms", ctypes.c_double), ("elapsed_ms", ctypes.c_double),
                ("arena_bytes", ctypes.c_uint64), ("scene_bytes", ctypes.c_uint64)]


class RenderServerConfig(ctypes.Structure):
    """pbrtx_render_server_config (include/pbrtx_gpu.h)."""
    _fields_ = [("arena_bytes", ctypes.c_uint64), ("batch_samples", ctypes.c_uint32), ("mode", ctypes.c_int32)]


class RenderRequest(ctypes.Structure):
    """pbrtx_render_request (include/pbrtx_gpu.h)."""
    _fields_ = [("scene", ctypes.c_uint32), ("cam_to_world", ctypes.c_float * 16), ("fov_deg", ctypes.c_float),
                ("screen", ctypes.c_float * 4), ("xres", ctypes.c_uint32), ("yres", ctypes.c_uint32),
                ("nx", ctypes.c_uint32), ("ny", ctypes.c_uint32), ("seed", ctypes.c_uint32)]


class RenderRequestStats(ctypes.Structure):
    """pbrtx_render_request_stats (include/pbrtx_gpu.h)."""
    _fields_ = [("n_samples", ctypes.c_uint64), ("status", ctypes.c_int32), ("batches", ctypes.c_uint32),
                ("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double),
                ("first_launch_ms", ctypes.c_double), ("done_ms", ctypes.c_double)]


class RenderServerCounters(ctypes.Structure):
    """pbrtx_render_server_counters (include/pbrtx_gpu.h)."""
    _fields_ = [("samples_done", ctypes.c_uint64), ("requests_done", ctypes.c_uint64), ("steps", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("camera_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
                ("extension_rays", ctypes.c_uint64), ("busy_ms", ctypes.c_double), ("elapsed_ms", ctypes.c_double),
                ("arena_bytes", ctypes.c_uint64), ("scene_bytes", ctypes.c_uint64)]


_lib = None


def lib():
    """Load libpbrtx_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PbrtxError("libpbrtx_gpu.so is not built: run `python -m pbrt_rust_amd.build` "
                         "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    c = ctypes
    vp, u32, i32, u64 = c.c_void_p, c.c_uint32, c.c_int32, c.c_uint64
    L.pbrtx_last_error.restype = c.c_char_p
    L.pbrtx_bvh_build.argtypes = [vp, u32, vp, u32, vp, u32, i32, u32, c.POINTER(vp)]
    L.pbrtx_bvh_info.argtypes = [vp, c.POINTER(u32), c.POINTER(u32), c.POINTER(u32), c.POINTER(c.c_double)]
    L.pbrtx_bvh_export.argtypes = [vp, vp, vp, vp, vp, vp]
    L.pbrtx_bvh_free.argtypes = [vp]
    L.pbrtx_bvh_layout_info.argtypes = [vp, c.POINTER(u32), c.POINTER(u32), c.POINTER(u32)]
    L.pbrtx_bvh_export_layouts.argtypes = [vp, vp, vp, vp]
    L.pbrtx_bvh_occlusion_records.argtypes = [vp, c.POINTER(u32), c.POINTER(u32), vp]
    L.pbrtx_scene_create.argtypes = [vp, c.POINTER(vp)]
    L.pbrtx_scene_destroy.argtypes = [vp]
    L.pbrtx_scene_device_bytes.argtypes = [vp, c.POINTER(u64)]
    L.pbrtx_scene_build_gpu.argtypes = [vp, u32, vp, u32, vp, u32, c.POINTER(vp), c.POINTER(c.c_double)]
    L.pbrtx_intersect.argtypes = [vp, vp, u32, vp, i32, vp]
    L.pbrtx_intersect_p.argtypes = [vp, vp, u32, vp, i32, vp]
    L.pbrtx_intersect_p_noalloc.argtypes = [vp, vp, u32, vp, vp]
    L.pbrtx_set_tuning.argtypes = [i32, i32, i32, i32, i32, i32]
    L.pbrtx_set_stack_depth.argtypes = [i32, i32, i32]
    L.pbrtx_set_ray_lanes.argtypes = [i32]
    L.pbrtx_reset_tuning.argtypes = []
    L.pbrtx_device_copy.argtypes = [vp, vp, u64, vp]
    f32 = c.c_float
    L.pbrtx_renderer_set_filter.argtypes = [vp, i32, f32, f32, f32, f32]
    L.pbrtx_filter_eval.argtypes = [i32, f32, f32, f32, f32, f32, f32, c.POINTER(f32)]
    L.pbrtx_filter_table.argtypes = [i32, f32, f32, f32, f32, c.POINTER(f32)]
    L.pbrtx_fma_probe_threads.argtypes = []
    L.pbrtx_fma_probe_threads.restype = u64
    L.pbrtx_fma_probe.argtypes = [vp, u64, ctypes.c_uint32, ctypes.POINTER(u64), vp]
    L.pbrtx_gather_probe_threads.argtypes = []
    L.pbrtx_gather_probe_threads.restype = u64
    L.pbrtx_gather_probe.argtypes = [vp, u64, vp, u64, vp]
    L.pbrtx_device_malloc.argtypes = [c.POINTER(vp), u64]
    L.pbrtx_device_free.argtypes = [vp]
    L.pbrtx_memcpy_htod.argtypes = [vp, vp, u64]
    L.pbrtx_memcpy_dtoh.argtypes = [vp, vp, u64]
    L.pbrtx_timer_start.argtypes = [vp]
    L.pbrtx_timer_stop.argtypes = [vp, c.POINTER(c.c_float)]
    L.pbrtx_device_name.argtypes = [c.c_char_p, u32]
    L.pbrtx_host_malloc.argtypes = [c.POINTER(vp), u64]
    L.pbrtx_host_free.argtypes = [vp]
    L.pbrtx_server_create.argtypes = [c.POINTER(ServerConfig), c.POINTER(vp)]
    L.pbrtx_server_add_scene.argtypes = [vp, vp, c.POINTER(u32)]
    L.pbrtx_server_submit.argtypes = [vp, u32, i32, vp, u32, vp, c.POINTER(u64)]
    L.pbrtx_server_wait.argtypes = [vp, u64, c.POINTER(RequestStats)]
    L.pbrtx_server_stats.argtypes = [vp, c.POINTER(ServerCounters)]
    L.pbrtx_server_destroy.argtypes = [vp]
    L.pbrtx_render_server_create.argtypes = [c.POINTER(RenderServerConfig), c.POINTER(vp)]
    L.pbrtx_render_server_add_scene.argtypes = [vp, c.POINTER(RenderDesc), c.POINTER(u32)]
    L.pbrtx_render_server_submit.argtypes = [vp, c.POINTER(RenderRequest), vp, c.POINTER(u64)]
    L.pbrtx_render_server_wait.argtypes = [vp, u64, c.POINTER(RenderRequestStats)]
    L.pbrtx_render_server_stats.argtypes = [vp, c.POINTER(RenderServerCounters)]
    L.pbrtx_render_server_destroy.argtypes = [vp]
    L.pbrtx_render_server_set_mode.argtypes = [vp, i32]
    if hasattr(L, "pbrtx_renderer_create"):
        L.pbrtx_renderer_create.argtypes = [c.POINTER(RenderDesc), c.POINTER(vp)]
        L.pbrtx_renderer_render.argtypes = [vp, i32, vp, c.POINTER(RenderStats)]
        L.pbrtx_renderer_destroy.argtypes = [vp]
        L.pbrtx_renderer_sample_radiance.argtypes = [vp, vp, u64]
        L.pbrtx_renderer_write_image.argtypes = [vp, c.c_char_p]
        L.pbrtx_renderer_sample_slots.argtypes = [vp, c.POINTER(u32)]
        L.pbrtx_write_image.argtypes = [c.c_char_p, vp, u32, u32]
        L.pbrtx_image_encode8.argtypes = [vp, u64, vp]
        L.pbrtx_renderer_set_chunk.argtypes = [vp, u32]
        L.pbrtx_renderer_set_graph.argtypes = [vp, i32]
        L.pbrtx_renderer_set_octant_sort.argtypes = [vp, i32]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().pbrtx_last_error()
        raise PbrtxError("pbrtx error %d: %s" % (rc, msg.decode() if msg else ""))


class DeviceBuffer:
    """A block of HBM owned by Python (pbrtx_device_malloc / free)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().pbrtx_device_malloc(ctypes.byref(p), max(self.nbytes, 4)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(lib().pbrtx_memcpy_htod(b.ptr, a.ctypes.data, a.nbytes))
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().pbrtx_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes))

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            assert out.nbytes <= self.nbytes
            check(lib().pbrtx_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().pbrtx_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """A numpy array over page-locked host memory (pbrtx_host_malloc)."""

    def __init__(self, shape, dtype):
        dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().pbrtx_host_malloc(ctypes.byref(p), max(n, 4)))
        self.ptr = p.value
        buf = (ctypes.c_char * max(n, 4)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=np.uint8, count=n).view(dtype).reshape(shape)

    def free(self):
        if getattr(self, "ptr", None):
            self.array = None
            lib().pbrtx_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(lib().pbrtx_synchronize())


def device_name():
    buf = ctypes.create_string_buffer(256)
    check(lib().pbrtx_device_name(buf, 256))
    return buf.value.decode()
