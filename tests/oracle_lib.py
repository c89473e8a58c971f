"""ctypes loader for the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker the parity tests compare
the HIP path against.  The library is built by oracle/Makefile (build() in
__graft_entry__ runs it); when it is missing here it is built on first use.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None

F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
U32P = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
U8P = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
U64P = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class RenderDesc(ctypes.Structure):
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
        ("filter_kind", ctypes.c_int32), ("filter_xw", ctypes.c_float), ("filter_yw", ctypes.c_float),
        ("filter_p0", ctypes.c_float), ("filter_p1", ctypes.c_float),
        ("sampler_kind", ctypes.c_int32),
        ("spheres", ctypes.c_void_p), ("nspheres", ctypes.c_uint32),
        ("sphere_mat", ctypes.c_void_p),
    ]

# film filters as (kind, x width, y width, p0, p1), filter.rs:44-86
FILTER_KINDS = {"mean": 0, "triangle": 1, "gaussian": 2, "mitchell": 3, "lanczos": 4}


def filter_eval(filt, x, y):
    out = ctypes.c_float()
    assert lib().orc_filter_eval(*filt, x, y, ctypes.byref(out)) == 0
    return out.value


def filter_table(filt):
    t = np.zeros(256, np.float32)
    assert lib().orc_filter_table(*filt, t) == 0
    return t


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    fp = ctypes.POINTER(ctypes.c_float)
    L.orc_bbox_intersect.argtypes = [F32P, F32P, F32P, F32P, ctypes.c_float, ctypes.c_float, fp, fp]
    L.orc_quadratic.argtypes = [ctypes.c_float] * 3 + [fp, fp]
    L.orc_partition_by_i32.argtypes = [I32P, ctypes.c_uint32]
    L.orc_partition_by_f32.argtypes = [F32P, ctypes.c_uint32]
    L.orc_filter_eval.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_float, fp]
    L.orc_filter_table.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                   F32P]
    L.orc_hit_box_misses.argtypes = [ctypes.c_void_p, ctypes.c_uint32, F32P, F32P, F32P, F32P, U32P]
    L.orc_radical_inverse.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.orc_radical_inverse.restype = ctypes.c_double
    L.orc_halton_table.argtypes = [ctypes.c_uint32] * 3 + [ctypes.POINTER(ctypes.c_uint32), ctypes.c_void_p,
                                                          ctypes.c_void_p]
    L.orc_mirror_sample_f.argtypes = [F32P, F32P, F32P, F32P, F32P]
    L.orc_mirror_sample_f.restype = None
    L.orc_ld_draw.argtypes = [ctypes.c_uint32] * 3
    L.orc_ld_draw.restype = ctypes.c_uint64
    L.orc_ld_vdc.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.orc_ld_vdc.restype = ctypes.c_float
    L.orc_ld_sobol2.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.orc_ld_sobol2.restype = ctypes.c_float
    L.orc_ld_pixel_points.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, F32P]
    L.orc_get_crop_window.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, F32P]
    L.orc_compute_sub_window.argtypes = [ctypes.c_int32] * 4 + [ctypes.c_uint32, ctypes.c_uint32, I32P]
    L.orc_matrix_inverse.argtypes = [F32P, F32P]
    L.orc_stratified_1d_nojitter.argtypes = [F32P, ctypes.c_uint32]
    L.orc_stratified_2d_nojitter.argtypes = [F32P, ctypes.c_uint32, ctypes.c_uint32]
    L.orc_film_extents.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float, F32P, I32P]
    L.orc_scene_new.restype = ctypes.c_void_p
    L.orc_scene_free.argtypes = [ctypes.c_void_p]
    L.orc_scene_add_mesh.argtypes = [ctypes.c_void_p, ctypes.c_void_p, F32P, ctypes.c_uint32, U32P, ctypes.c_uint32]
    L.orc_scene_add_mesh.restype = ctypes.c_int64
    L.orc_scene_add_sphere.argtypes = [ctypes.c_void_p, F32P, F32P, ctypes.c_int, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_float]
    L.orc_scene_add_sphere.restype = ctypes.c_int64
    L.orc_scene_build.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32]
    L.orc_scene_num_prims.argtypes = [ctypes.c_void_p]
    L.orc_scene_num_prims.restype = ctypes.c_uint32
    L.orc_scene_num_nodes.argtypes = [ctypes.c_void_p]
    L.orc_scene_num_nodes.restype = ctypes.c_uint32
    L.orc_scene_nodes.argtypes = [ctypes.c_void_p, F32P, U32P, U32P, U8P, U32P]
    L.orc_intersect.argtypes = [ctypes.c_void_p, ctypes.c_uint32, F32P, F32P, F32P, F32P, U32P,
                                F32P, F32P, F32P, U64P, ctypes.c_uint32]
    L.orc_intersect_p.argtypes = [ctypes.c_void_p, ctypes.c_uint32, F32P, F32P, F32P, F32P, U8P,
                                  U64P, ctypes.c_uint32]
    L.orc_intersect_brute.argtypes = [ctypes.c_void_p, ctypes.c_uint32, F32P, F32P, F32P, F32P,
                                      U32P, F32P, U32P]
    L.orc_sphere_intersect.argtypes = [F32P, F32P, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_float, F32P, F32P, ctypes.c_float,
                                       ctypes.c_float, F32P]
    L.orc_mesh_triangle_query.argtypes = [ctypes.c_void_p, F32P, ctypes.c_uint32, U32P, ctypes.c_uint32,
                                          ctypes.c_uint32, F32P, F32P, ctypes.c_float, ctypes.c_float,
                                          fp, U32P]
    L.orc_camera_raster_to_camera.argtypes = [ctypes.c_uint32, ctypes.c_uint32, F32P, ctypes.c_float, F32P]
    L.orc_projection.argtypes = [ctypes.c_uint32, ctypes.c_uint32, F32P, ctypes.c_void_p, F32P, F32P]
    L.orc_camera_rays.argtypes = [F32P, F32P, ctypes.c_uint32, F32P, F32P, F32P]
    L.orc_rng4.argtypes = [ctypes.c_uint32] * 4 + [F32P]
    L.orc_pixel_samples.argtypes = [ctypes.c_uint32] * 6 + [ctypes.c_int, F32P]
    L.orc_render.argtypes = [ctypes.POINTER(RenderDesc), ctypes.c_uint32, ctypes.c_uint32,
                             ctypes.c_uint32, F32P, ctypes.c_void_p, U64P]
    _lib = L
    return L


def f32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


def u32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint32))


class OracleScene:
    """BVHAccelerator over meshes and spheres, restated on the CPU."""

    def __init__(self):
        self.L = lib()
        self.h = self.L.orc_scene_new()

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_scene_free(self.h)
            self.h = None

    def add_mesh(self, pts, idx, o2w=None):
        pts = f32(pts).reshape(-1)
        idx = u32(idx).reshape(-1)
        o2w_p = None
        if o2w is not None:
            self._o2w = f32(o2w).reshape(16)
            o2w_p = self._o2w.ctypes.data
        r = self.L.orc_scene_add_mesh(self.h, o2w_p, pts, pts.size // 3, idx, idx.size)
        assert r >= 0, r
        return r

    def add_sphere(self, m, minv, radius, z0, z1, phimax, ro=0):
        r = self.L.orc_scene_add_sphere(self.h, f32(m).reshape(16), f32(minv).reshape(16), ro,
                                        radius, z0, z1, phimax)
        assert r >= 0, r
        return r

    def build(self, split="sah", max_prims=1):
        code = {"sah": 0, "middle": 1, "equal": 2}[split]
        r = self.L.orc_scene_build(self.h, code, max_prims)
        assert r == 0, r

    def nodes(self):
        n = self.L.orc_scene_num_nodes(self.h)
        npr = self.L.orc_scene_num_prims(self.h)
        b = np.zeros(6 * n, np.float32)
        off = np.zeros(n, np.uint32)
        meta = np.zeros(n, np.uint32)
        leaf = np.zeros(n, np.uint8)
        order = np.zeros(npr, np.uint32)
        self.L.orc_scene_nodes(self.h, b, off, meta, leaf, order)
        return b.reshape(n, 6), off, meta, leaf, order

    def intersect(self, o, d, mint, maxt, nthreads=1):
        o = f32(o).reshape(-1); d = f32(d).reshape(-1)
        n = o.size // 3
        mint = f32(np.broadcast_to(mint, (n,)))
        maxt = f32(np.broadcast_to(maxt, (n,))).copy()
        prim = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32); b1 = np.zeros(n, np.float32); b2 = np.zeros(n, np.float32)
        st = np.zeros(2, np.uint64)
        self.L.orc_intersect(self.h, n, o, d, mint, maxt, prim, t, b1, b2, st, nthreads)
        return dict(prim=prim, t=t, b1=b1, b2=b2, maxt=maxt, stats=st)

    def intersect_p(self, o, d, mint, maxt, nthreads=1):
        o = f32(o).reshape(-1); d = f32(d).reshape(-1)
        n = o.size // 3
        mint = f32(np.broadcast_to(mint, (n,)))
        maxt = f32(np.broadcast_to(maxt, (n,)))
        occ = np.zeros(n, np.uint8)
        st = np.zeros(2, np.uint64)
        self.L.orc_intersect_p(self.h, n, o, d, mint, maxt, occ, st, nthreads)
        return occ, st

    def box_misses(self, o, d, mint, maxt):
        """per ray: prims it hits whose own box the slab test rejects at
        maxt = the hit's t (the order-sensitive cases of the traversal)"""
        o = f32(o).reshape(-1); d = f32(d).reshape(-1)
        n = o.size // 3
        mint = f32(np.broadcast_to(mint, (n,)))
        maxt = f32(np.broadcast_to(maxt, (n,)))
        out = np.zeros(n, np.uint32)
        self.L.orc_hit_box_misses(self.h, n, o, d, mint, maxt, out)
        return out

    def brute(self, o, d, mint, maxt):
        o = f32(o).reshape(-1); d = f32(d).reshape(-1)
        n = o.size // 3
        mint = f32(np.broadcast_to(mint, (n,)))
        maxt = f32(np.broadcast_to(maxt, (n,)))
        prim = np.zeros(n, np.uint32); t = np.zeros(n, np.float32); ties = np.zeros(n, np.uint32)
        self.L.orc_intersect_brute(self.h, n, o, d, mint, maxt, prim, t, ties)
        return prim, t, ties


def make_render_desc(scene, cam, integ, keep):
    """Fill an orc_render_desc from a scene dict (see pbrt_rust_amd.scenes)."""
    d = RenderDesc()
    arrs = dict(
        verts=f32(scene["verts"]).reshape(-1), idx=u32(scene["idx"]).reshape(-1),
        face_mat=u32(scene["face_mat"]), mats=f32(scene["mats"]).reshape(-1),
        face_light=np.ascontiguousarray(np.asarray(scene["face_light"], np.int32)),
        area_le=f32(scene["area_le"]).reshape(-1) if len(scene["area_le"]) else np.zeros(3, np.float32),
        point_lights=f32(scene["point_lights"]).reshape(-1) if len(scene["point_lights"]) else np.zeros(6, np.float32),
    )
    keep.update(arrs)
    d.verts = arrs["verts"].ctypes.data; d.nverts = arrs["verts"].size // 3
    d.idx = arrs["idx"].ctypes.data; d.nfaces = arrs["idx"].size // 3
    d.face_mat = arrs["face_mat"].ctypes.data
    d.mats = arrs["mats"].ctypes.data; d.nmats = arrs["mats"].size // 4
    d.face_light = arrs["face_light"].ctypes.data
    d.area_le = arrs["area_le"].ctypes.data; d.n_area = len(scene["area_le"])
    d.point_lights = arrs["point_lights"].ctypes.data; d.n_point = len(scene["point_lights"])
    for i, v in enumerate(np.asarray(cam["cam_to_world"], np.float32).reshape(16)):
        d.cam_to_world[i] = float(v)
    d.fov_deg = cam["fov"]
    for i, v in enumerate(cam["screen"]):
        d.screen[i] = float(v)
    d.xres, d.yres = cam["xres"], cam["yres"]
    d.nx, d.ny = integ["nx"], integ["ny"]
    d.seed = integ["seed"]
    d.integrator = 1 if integ["kind"] == "path" else 0
    d.light_samples = integ.get("light_samples", 1)
    d.max_depth = integ.get("max_depth", 5)
    d.rr_start = integ.get("rr_start", 3)
    d.split = 0
    d.max_prims = 1
    sph = scene.get("spheres")
    if sph is not None and len(sph):
        from pbrt_rust_amd.api import sphere_descs
        arr = sphere_descs(sph)
        mats = u32(scene["sphere_mat"])
        keep.update(sph_arr=arr, sph_mat=mats)
        d.spheres = ctypes.cast(arr, ctypes.c_void_p)
        d.nspheres = len(sph)
        d.sphere_mat = mats.ctypes.data
    return d


def render(scene, cam, integ, y0=0, y1=None, nthreads=1, want_samples=False, filt=None):
    """filt: (kind, xw, yw, p0, p1) of the film filter, None for
    Filter::mean(0.5, 0.5)"""
    L = lib()
    keep = {}
    d = make_render_desc(scene, cam, integ, keep)
    if filt is not None:
        d.filter_kind, d.filter_xw, d.filter_yw, d.filter_p0, d.filter_p1 = filt
    d.sampler_kind = integ.get("sampler", 0)
    if y1 is None:
        y1 = cam["yres"]
    out = np.zeros((y1 - y0) * cam["xres"] * 3, np.float32)
    st = np.zeros(9, np.uint64)
    spp = integ["nx"] * integ["ny"]
    if integ.get("sampler", 0) == 1:
        spp = 1 << (spp - 1).bit_length()  # LDSampler rounds up to a power of two
    elif integ.get("sampler", 0) == 2:
        spp = halton_table(cam["xres"], cam["yres"], spp)[0]  # K slots a pixel
    samples = np.zeros((y1 - y0) * cam["xres"] * spp * 3, np.float32) if want_samples else None
    r = L.orc_render(ctypes.byref(d), y0, y1, nthreads, out,
                     samples.ctypes.data if want_samples else None, st)
    assert r == 0, r
    res = dict(rgb=out.reshape(y1 - y0, cam["xres"], 3), stats=st)
    if want_samples:
        res["samples"] = samples.reshape(y1 - y0, cam["xres"], spp, 3)
    return res


def halton_table(xres, yres, spp):
    """(K, counts[yres, xres], positions[yres, xres, K, 2]) of the
    HaltonSampler over the whole image"""
    L = lib()
    K = ctypes.c_uint32()
    assert L.orc_halton_table(xres, yres, spp, ctypes.byref(K), None, None) == 0
    cnt = np.zeros(xres * yres, np.uint32)
    pos = np.zeros(xres * yres * K.value * 2, np.float32)
    assert L.orc_halton_table(xres, yres, spp, ctypes.byref(K), cnt.ctypes.data, pos.ctypes.data) == 0
    return K.value, cnt.reshape(yres, xres), pos.reshape(yres, xres, K.value, 2)
