"""Host-side mirror of pbrt_rust's aggregate interface over the HIP path.

Names and argument meanings follow the reference:
  Shape::triangle_mesh / Shape::sphere      src/shape/mod.rs:170-190
  Primitive::simple                          src/primitive/mod.rs:79-85
  BVHAccelerator::new(p, max_prims, "sah"|"middle"|"equal")
                                             src/primitive/aggregates/bvh.rs:335-362
  Intersectable::{intersect, intersect_p}    src/intersection.rs:61-66
  Aggregate::bvh                             src/primitive/aggregates/mod.rs:27-29
Rays are batches (RayBatch) rather than single Ray objects; intersect sets
each hit ray's tmax to the hit t, the reference's maxt side effect
(src/primitive/geometric.rs:64).
"""
import ctypes
import os
import warnings

import numpy as np

from . import _lib
from ._lib import HIT_DTYPE, NO_HIT, RAY_DTYPE, DeviceBuffer, check, lib

F32 = np.float32
F32_MAX = float(np.finfo(np.float32).max)

PATHS = {"plain": 0, "batched": 1}


def _path_code(path):
    if isinstance(path, str):
        if path not in PATHS:
            raise ValueError("unknown path %r" % path)
        return PATHS[path]
    return int(path)


class Transform:
    """Transform{m, m_inv} (src/transform/transform.rs:14-19) in float32."""

    def __init__(self, m=None, m_inv=None):
        self.m = np.eye(4, dtype=F32) if m is None else np.asarray(m, F32).reshape(4, 4)
        self.m_inv = np.eye(4, dtype=F32) if m_inv is None else np.asarray(m_inv, F32).reshape(4, 4)

    @staticmethod
    def translate(v):
        t = Transform()
        t.m[:3, 3] = np.asarray(v, F32)
        t.m_inv[:3, 3] = -np.asarray(v, F32)
        return t

    @staticmethod
    def scale(x, y, z):
        t = Transform()
        t.m[0, 0], t.m[1, 1], t.m[2, 2] = F32(x), F32(y), F32(z)
        t.m_inv[0, 0], t.m_inv[1, 1], t.m_inv[2, 2] = F32(1) / F32(x), F32(1) / F32(y), F32(1) / F32(z)
        return t

    def inverse(self):
        return Transform(self.m_inv.copy(), self.m.copy())

    @staticmethod
    def _matmul(a, b):
        # Matrix4x4 product summed left to right (matrix4x4.rs:167-182)
        r = np.zeros((4, 4), F32)
        for i in range(4):
            for j in range(4):
                s = F32(a[i, 0] * b[0, j])
                s = F32(s + F32(a[i, 1] * b[1, j]))
                s = F32(s + F32(a[i, 2] * b[2, j]))
                r[i, j] = F32(s + F32(a[i, 3] * b[3, j]))
        return r

    def __mul__(self, o):
        return Transform(self._matmul(self.m, o.m), self._matmul(o.m_inv, self.m_inv))

    def xf_points(self, p):
        """ApplyTransform<Point> (transform.rs:195-207) on an (n,3) array."""
        p = np.asarray(p, F32).reshape(-1, 3)
        m = self.m
        x, y, z = p[:, 0], p[:, 1], p[:, 2]
        rows = []
        for i in range(4):
            s = m[i, 0] * x
            s = s + m[i, 1] * y
            s = s + m[i, 2] * z
            rows.append(s + m[i, 3])
        xt, yt, zt, w = rows
        one = w == F32(1)
        out = np.stack([np.where(one, xt, xt / w), np.where(one, yt, yt / w), np.where(one, zt, zt / w)], 1)
        return out.astype(F32)


class TriangleMesh:
    """Shape::triangle_mesh (shape/mod.rs:185-190): Mesh::new pre-transforms
    the points to world space (mesh.rs:299-314)."""

    def __init__(self, o2w, w2o, reverse_orientation, vi, p):
        vi = np.asarray(vi, np.uint32).reshape(-1)
        if vi.size % 3 != 0:
            raise ValueError("vertex index count must be a multiple of 3 (mesh.rs:303)")
        o2w = o2w or Transform()
        self.p = o2w.xf_points(p)
        self.vi = vi
        self.reverse_orientation = bool(reverse_orientation)

    @property
    def nfaces(self):
        return self.vi.size // 3


class Sphere:
    """Shape::sphere (shape/mod.rs:170-173, sphere.rs:28-44)."""

    def __init__(self, o2w, w2o, reverse_orientation, radius, z0, z1, phimax):
        if not radius > 0:
            raise ValueError("sphere radius must be positive (sphere.rs:30)")
        self.o2w = o2w
        self.reverse_orientation = bool(reverse_orientation)
        self.radius, self.z0, self.z1, self.phimax = float(radius), float(z0), float(z1), float(phimax)


def sphere_descs(spheres):
    """pbrtx_sphere records (a ctypes array) of Sphere shapes"""
    sph = (_lib.SphereDesc * max(1, len(spheres)))()
    for k, s in enumerate(spheres):
        for i, v in enumerate(np.asarray(s.o2w.m, F32).reshape(16)):
            sph[k].o2w[i] = float(v)
        for i, v in enumerate(np.asarray(s.o2w.m_inv, F32).reshape(16)):
            sph[k].o2w_inv[i] = float(v)
        sph[k].radius, sph[k].z0, sph[k].z1 = s.radius, s.z0, s.z1
        sph[k].phimax_deg = s.phimax
        sph[k].reverse_orientation = int(s.reverse_orientation)
    return sph


class Primitive:
    """Primitive::simple (primitive/mod.rs:79-85): a shape in the aggregate."""

    def __init__(self, shape):
        self.shape = shape

    @staticmethod
    def simple(shape):
        return Primitive(shape)


class RayBatch:
    """n Ray records {o, d, mint, maxt} (src/ray.rs:7-38); maxt defaults to
    f32::MAX as Ray::new_with does."""

    def __init__(self, o, d, mint=0.0, maxt=F32_MAX):
        o = np.asarray(o, F32).reshape(-1, 3)
        d = np.asarray(d, F32).reshape(-1, 3)
        n = o.shape[0]
        a = np.zeros(n, RAY_DTYPE)
        a["ox"], a["oy"], a["oz"] = o[:, 0], o[:, 1], o[:, 2]
        a["dx"], a["dy"], a["dz"] = d[:, 0], d[:, 1], d[:, 2]
        a["tmin"] = np.broadcast_to(np.asarray(mint, F32), (n,))
        a["tmax"] = np.broadcast_to(np.asarray(maxt, F32), (n,))
        self.a = a

    @classmethod
    def from_records(cls, a):
        """wrap an existing RAY_DTYPE array (no copy)"""
        r = cls.__new__(cls)
        r.a = np.ascontiguousarray(a, dtype=RAY_DTYPE)
        return r

    def __len__(self):
        return self.a.shape[0]

    @property
    def maxt(self):
        return self.a["tmax"]


QNODE_DTYPE = np.dtype([("lx", "<f4", 4), ("hx", "<f4", 4), ("ly", "<f4", 4), ("hy", "<f4", 4),
                        ("lz", "<f4", 4), ("hz", "<f4", 4), ("ref", "<i4", 4), ("meta", "<u4"), ("pad", "<u4", 3)])
CNODE_DTYPE = np.dtype([("o", "<f4", 3), ("emeta", "<u4"), ("q", "u1", 24), ("ref", "<i4", 4), ("pad", "<u4", 2)])


class BVHAccelerator:
    """GPU aggregate with the reference's BVHAccelerator::new signature."""

    def __init__(self, prims, max_prims=1, split_method="sah"):
        """split_method "lbvh": the tree is built on the GPU (a linear BVH,
        pbrtx_scene_build_gpu) straight into device memory; its closest hits
        equal the reference tree's except among primitives at the same t,
        its occlusion bits equal them exactly; no host tree is kept."""
        split = {"sah": 0, "middle": 1, "equal": 2, "lbvh": 3}.get(split_method)
        if split is None:
            warnings.warn('BVH split method %s unknown. Using "SAH"' % split_method)  # bvh.rs:345-348
            split = 0
        meshes = [p.shape for p in prims if isinstance(p.shape, TriangleMesh)]
        spheres = [p.shape for p in prims if isinstance(p.shape, Sphere)]
        if len(meshes) + len(spheres) != len(prims):
            raise TypeError("only triangle meshes and spheres are supported")
        kinds = [isinstance(p.shape, Sphere) for p in prims]
        if kinds != sorted(kinds):
            raise ValueError("meshes must precede spheres in the primitive list")
        # one vertex / index array over all meshes, faces in input order
        verts, idx, base = [], [], 0
        for m in meshes:
            verts.append(m.p)
            idx.append(m.vi + base)
            base += m.p.shape[0]
        self.verts = np.ascontiguousarray(np.concatenate(verts) if verts else np.zeros((0, 3), F32))
        self.idx = np.ascontiguousarray(np.concatenate(idx) if idx else np.zeros(0, np.uint32))
        sph = sphere_descs(spheres)
        L = lib()
        h = ctypes.c_void_p()
        self.nfaces = self.idx.size // 3
        self.nspheres = len(spheres)
        self._scene = None
        self._bvh = None
        self.build_ms = None
        if split == 3:
            ms = ctypes.c_double()
            check(L.pbrtx_scene_build_gpu(self.verts.ctypes.data if self.verts.size else None,
                                          self.verts.shape[0], self.idx.ctypes.data if self.idx.size else None,
                                          self.idx.size // 3, ctypes.cast(sph, ctypes.c_void_p) if spheres else None,
                                          len(spheres), ctypes.byref(h), ctypes.byref(ms)))
            self._scene = h.value
            self.build_ms = ms.value
            return
        check(L.pbrtx_bvh_build(self.verts.ctypes.data if self.verts.size else None,
                                self.verts.shape[0], self.idx.ctypes.data if self.idx.size else None,
                                self.idx.size // 3, ctypes.cast(sph, ctypes.c_void_p) if spheres else None,
                                len(spheres), split, int(max_prims), ctypes.byref(h)))
        self._bvh = h.value

    # ---- host-side tree -------------------------------------------------------
    def info(self):
        n, d, p = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        ms = ctypes.c_double()
        check(lib().pbrtx_bvh_info(self._bvh, ctypes.byref(n), ctypes.byref(d), ctypes.byref(p), ctypes.byref(ms)))
        return dict(nodes=n.value, depth=d.value, prims=p.value, build_ms=ms.value)

    def nodes(self):
        info = self.info()
        n, npr = info["nodes"], info["prims"]
        b = np.zeros((n, 6), F32)
        off = np.zeros(n, np.uint32)
        meta = np.zeros(n, np.uint32)
        leaf = np.zeros(n, np.uint8)
        order = np.zeros(npr, np.uint32)
        check(lib().pbrtx_bvh_export(self._bvh, b.ctypes.data, off.ctypes.data, meta.ctypes.data,
                                     leaf.ctypes.data, order.ctypes.data))
        return b, off, meta, leaf, order

    def layouts(self):
        """The device layouts as numpy arrays: four-wide records (QNODE_DTYPE),
        compressed records (CNODE_DTYPE, empty if none) and the compressed
        layout's prim records (n x 16 float32)."""
        nq, nc, npr = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().pbrtx_bvh_layout_info(self._bvh, ctypes.byref(nq), ctypes.byref(nc), ctypes.byref(npr)))
        q = np.zeros(nq.value, QNODE_DTYPE)
        c = np.zeros(nc.value, CNODE_DTYPE)
        p = np.zeros((npr.value if nc.value else 0, 16), F32)
        check(lib().pbrtx_bvh_export_layouts(self._bvh, q.ctypes.data, c.ctypes.data if c.size else None,
                                             p.ctypes.data if p.size else None))
        return q, c, p

    def occlusion_records(self):
        """The compressed records any-hit queries traverse (CNODE_DTYPE, root
        first; empty when they use the compressed layout) and their depth."""
        n, d = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().pbrtx_bvh_occlusion_records(self._bvh, ctypes.byref(n), ctypes.byref(d), None))
        r = np.zeros(n.value, CNODE_DTYPE)
        if n.value:
            check(lib().pbrtx_bvh_occlusion_records(self._bvh, ctypes.byref(n), ctypes.byref(d), r.ctypes.data))
        return r, d.value

    def world_bound(self):
        if self._bvh is None:  # a device-built tree: the primitives' bounds
            lo, hi = self.verts.min(axis=0), self.verts.max(axis=0)
            return lo.astype(F32), hi.astype(F32)
        b = self.nodes()[0]
        return b[0, :3].copy(), b[0, 3:].copy()

    # ---- device scene ---------------------------------------------------------
    @property
    def scene(self):
        if self._scene is None:
            s = ctypes.c_void_p()
            check(lib().pbrtx_scene_create(self._bvh, ctypes.byref(s)))
            self._scene = s.value
        return self._scene

    def device_bytes(self):
        b = ctypes.c_uint64()
        check(lib().pbrtx_scene_device_bytes(self.scene, ctypes.byref(b)))
        return b.value

    def intersect_device(self, d_rays, n, d_hits, path="batched", stream=None):
        check(lib().pbrtx_intersect(self.scene, d_rays, n, d_hits, _path_code(path), stream))

    def intersect_p_device(self, d_rays, n, d_bits, path="batched", stream=None):
        check(lib().pbrtx_intersect_p(self.scene, d_rays, n, d_bits, _path_code(path), stream))

    def intersect(self, rays, path="batched"):
        """Intersectable::intersect for a RayBatch: returns the hit records and
        writes each hit t into rays.maxt."""
        n = len(rays)
        if n == 0:
            return np.zeros(0, HIT_DTYPE)
        dr = DeviceBuffer.from_array(rays.a)
        dh = DeviceBuffer(n * HIT_DTYPE.itemsize)
        self.intersect_device(dr.ptr, n, dh.ptr, path)
        _lib.synchronize()
        rays.a[:] = dr.download(RAY_DTYPE, n)
        return dh.download(HIT_DTYPE, n)

    def intersect_p(self, rays, path="batched"):
        """Intersectable::intersect_p for a RayBatch -> bool array."""
        n = len(rays)
        if n == 0:
            return np.zeros(0, bool)
        nw = (n + 31) // 32
        dr = DeviceBuffer.from_array(rays.a)
        db = DeviceBuffer(nw * 4)
        self.intersect_p_device(dr.ptr, n, db.ptr, path)
        _lib.synchronize()
        words = db.download(np.uint32, nw)
        bits = np.unpackbits(words.view(np.uint8), bitorder="little")[:n]
        return bits.astype(bool)

    def close(self):
        L = lib()
        if self._scene:
            L.pbrtx_scene_destroy(self._scene)
            self._scene = None
        if self._bvh:
            L.pbrtx_bvh_free(self._bvh)
            self._bvh = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RayServer:
    """Request server over device-resident BVHs (pbrtx_server_*): requests of
    rays for Intersectable::intersect ("closest") or intersect_p ("any")
    against registered scenes, admitted by an HBM budget and traced with
    continuous batching ("batched") or one plain kernel per request
    ("plain").  submit() returns a ticket; wait(ticket) blocks and returns the
    result (hit records or a bool array) and the request's timings (ms since
    the server started).  Its streams overlap only on separate hardware
    queues: in a process holding other streams too, set GPU_MAX_HW_QUEUES=8
    or more (HIP's default is 4) before the library loads."""

    MODES = {"batched": 0, "plain": 1}

    def __init__(self, arena_bytes=4 << 30, batch_rays=1 << 22, mode="batched", scene_budget_bytes=0):
        cfg = _lib.ServerConfig(int(arena_bytes), int(scene_budget_bytes), int(batch_rays), self.MODES[mode])
        h = ctypes.c_void_p()
        check(lib().pbrtx_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h.value
        self._scenes = []
        self._pending = {}

    def add_scene(self, bvh):
        sid = ctypes.c_uint32()
        check(lib().pbrtx_server_add_scene(self._h, bvh.scene, ctypes.byref(sid)))
        self._scenes.append(bvh)
        return sid.value

    def submit(self, scene_id, kind, rays, out=None):
        """rays: a RayBatch or a RAY_DTYPE array; out (optional): the result
        buffer, n HIT_DTYPE records ("closest") or ceil(n/32) uint32 words
        ("any") -- page-locked arrays (PinnedArray) let copies overlap device
        work.  Both are kept alive until wait()."""
        a = np.ascontiguousarray(rays.a if isinstance(rays, RayBatch) else rays, dtype=RAY_DTYPE)
        n = a.shape[0]
        k = KINDS[kind] if isinstance(kind, str) else int(kind)
        if out is None:
            out = np.zeros(n, HIT_DTYPE) if k == 0 else np.zeros((n + 31) // 32, np.uint32)
        elif out.nbytes < (n * HIT_DTYPE.itemsize if k == 0 else (n + 31) // 32 * 4) or not out.flags.c_contiguous:
            raise ValueError("result buffer too small or not contiguous")
        t = ctypes.c_uint64()
        check(lib().pbrtx_server_submit(self._h, int(scene_id), k, a.ctypes.data, n, out.ctypes.data,
                                        ctypes.byref(t)))
        self._pending[t.value] = (a, out, k, n)
        return t.value

    def wait(self, ticket):
        st = _lib.RequestStats()
        rc = lib().pbrtx_server_wait(self._h, int(ticket), ctypes.byref(st))
        a, out, k, n = self._pending.pop(ticket)
        check(rc)
        if k == 1:
            out = np.unpackbits(out.view(np.uint8).reshape(-1)[:((n + 31) // 32) * 4], bitorder="little")[:n].astype(bool)
        else:
            out = out.view(HIT_DTYPE).reshape(-1)[:n]
        return out, dict(n_rays=st.n_rays, launches=st.launches, submit_ms=st.submit_ms, admit_ms=st.admit_ms,
                         first_launch_ms=st.first_launch_ms, done_ms=st.done_ms,
                         latency_ms=st.done_ms - st.submit_ms)

    def counters(self):
        c = _lib.ServerCounters()
        check(lib().pbrtx_server_stats(self._h, ctypes.byref(c)))
        return {f: getattr(c, f) for f, _ in c._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().pbrtx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RenderServer:
    """Render-request server over device-resident scenes
    (pbrtx_render_server_*): requests of Renderer::render with the path
    integrator -- a camera, an image size and a stratified sampler -- against
    registered scenes, admitted by an HBM budget and rendered in progressive
    tiles batched across requests ("batched") or one at a time on the plain
    path ("plain").  submit() returns a ticket; wait(ticket) blocks and
    returns the image (yres, xres, 3) and the request's timings (ms since the
    server started)."""

    MODES = {"batched": 0, "plain": 1}

    def __init__(self, arena_bytes=8 << 30, batch_samples=1 << 22, mode="batched"):
        cfg = _lib.RenderServerConfig(int(arena_bytes), int(batch_samples), self.MODES[mode])
        h = ctypes.c_void_p()
        check(lib().pbrtx_render_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h.value
        self._pending = {}

    def set_mode(self, mode):
        """"batched" or "plain"; only with no request pending."""
        check(lib().pbrtx_render_server_set_mode(self._h, self.MODES[mode]))

    def add_scene(self, scene, integrator, split="sah"):
        """scene: a scene dict (see scenes.py); integrator: a PathIntegrator
        (its depth and roulette start apply to every request of the scene)."""
        d, keep = _render_desc(scene, None, None, integrator, split)
        sid = ctypes.c_uint32()
        check(lib().pbrtx_render_server_add_scene(self._h, ctypes.byref(d), ctypes.byref(sid)))
        return sid.value

    def submit(self, scene_id, camera, sampler, out=None):
        """camera: scenes.pinhole dict; sampler: StratifiedSampler; out
        (optional): a float32 (yres, xres, 3) buffer, page-locked arrays
        (PinnedArray) let the copy overlap device work.  Kept until wait()."""
        req = _lib.RenderRequest()
        req.scene = int(scene_id)
        _set_camera(req, camera)
        req.nx, req.ny, req.seed = sampler.xs, sampler.ys, sampler.seed
        shape = (camera["yres"], camera["xres"], 3)
        if out is None:
            out = np.zeros(shape, F32)
        elif out.dtype != F32 or out.size < np.prod(shape) or not out.flags.c_contiguous:
            raise ValueError("image buffer too small, not float32 or not contiguous")
        t = ctypes.c_uint64()
        check(lib().pbrtx_render_server_submit(self._h, ctypes.byref(req), out.ctypes.data, ctypes.byref(t)))
        self._pending[t.value] = (out, shape)
        return t.value

    def wait(self, ticket):
        st = _lib.RenderRequestStats()
        rc = lib().pbrtx_render_server_wait(self._h, int(ticket), ctypes.byref(st))
        out, shape = self._pending.pop(ticket)
        check(rc)
        img = out.reshape(-1)[:int(np.prod(shape))].reshape(shape)
        return img, dict(n_samples=st.n_samples, batches=st.batches, submit_ms=st.submit_ms, admit_ms=st.admit_ms,
                         first_launch_ms=st.first_launch_ms, done_ms=st.done_ms,
                         latency_ms=st.done_ms - st.submit_ms)

    def counters(self):
        c = _lib.RenderServerCounters()
        check(lib().pbrtx_render_server_stats(self._h, ctypes.byref(c)))
        return {f: getattr(c, f) for f, _ in c._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().pbrtx_render_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Aggregate:
    """Aggregate::bvh (aggregates/mod.rs:27-29)."""

    @staticmethod
    def bvh(prims, max_prims, sm):
        return BVHAccelerator(prims, max_prims, sm)


KINDS = {"closest": 0, "any": 1, "camera": 2}


SCHEMES = {"ray": 0, "packet": 1, "ray_exact": 2}


def set_tuning(path, kind, block=256, blocks_per_cu=0, refill_pct=0, scheme="ray", stack=None, lanes=None):
    """Launch shape of the closest-hit ("closest"), occlusion ("any") or
    renderer camera-ray ("camera") kernel of a path, and its traversal scheme: "ray" (a stack per ray, the
    compressed 64-byte records when the scene has them), "ray_exact" (a stack
    per ray on the exact 128-byte records) or "packet" (a stack per wave of 64
    consecutive rays in equal-sign groups; batched closest hit and camera rays
    only, refused for occlusion).  stack: LDS stack entries per lane
    of the per-ray kernel (8, 16 or 24; None keeps the current depth).
    lanes: batched occlusion only, lanes per ray (0 chooses by launch size,
    1, 4, 8 or 16; None keeps the current choice)."""
    k = KINDS[kind] if isinstance(kind, str) else int(kind)
    sc = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
    check(lib().pbrtx_set_tuning(_path_code(path), k, block, blocks_per_cu, refill_pct, sc))
    if stack is not None:
        check(lib().pbrtx_set_stack_depth(_path_code(path), k, int(stack)))
    if lanes is not None:
        if _path_code(path) != 1 or k != 1:
            raise ValueError("lanes per ray is a setting of batched occlusion only")
        check(lib().pbrtx_set_ray_lanes(int(lanes)))


def reset_tuning():
    """The built-in launch settings for every path and kind."""
    check(lib().pbrtx_reset_tuning())


# ---------------------------------------------------------------------------
# Renderer / integrators (src/renderer.rs:8-26, src/sampler_renderer.rs,
# src/integrator/mod.rs:148-180)
# ---------------------------------------------------------------------------
class StratifiedSampler:
    """StratifiedSampler::new(x0, x1, y0, y1, xs, ys, jitter, ...) over the
    whole image (sampler/stratified.rs:22-37); jitter values come from the
    counter-based generator keyed by (seed, pixel, sample, dimension)."""

    def __init__(self, xs, ys, seed=0):
        self.xs, self.ys, self.seed = int(xs), int(ys), int(seed)

    def maximum_sample_count(self):
        return self.xs * self.ys


class LDSampler:
    """LDSampler::new(x0, x1, y0, y1, samples_per_pixel, ...) over the whole
    image (sampler/lds.rs:17-31): samples_per_pixel rounded up to a power of
    two (with the reference's warning), image samples from the scrambled
    (0,2) sequence shuffled per pixel (sampler/utils.rs:52-69); the
    scrambles and shuffles draw from the counter-based generator keyed by
    (seed, pixel)."""
    kind = 1

    def __init__(self, samples_per_pixel, seed=0):
        n = int(samples_per_pixel)
        if n < 1:
            raise ValueError("samples_per_pixel must be positive")
        ps = 1 << (n - 1).bit_length()
        if ps != n:
            print("Warning -- LDSampler using next power of two (%d) samples per pixel" % ps)  # lds.rs:19-22
        self.xs, self.ys, self.seed = ps, 1, int(seed)

    def maximum_sample_count(self):
        return self.xs


class HaltonSampler:
    """HaltonSampler::new(x0, x1, y0, y1, samples_per_pixel, ...) over the
    whole image as one window (sampler/halton.rs:18-107): max(xres, yres)^2 *
    samples_per_pixel points of the (base 3, base 2) Halton sequence scaled
    to the square, those inside the image kept, so pixels receive varying
    numbers of samples.  The renderer gives every pixel K slots (K the most
    any pixel received); a pixel's samples fill its first slots in sequence
    order and its other slots stay empty (zero radiance, no film weight).
    seed keys the light and BSDF sample draws only."""
    kind = 2

    def __init__(self, samples_per_pixel, seed=0):
        n = int(samples_per_pixel)
        if n < 1:
            raise ValueError("samples_per_pixel must be positive")
        self.xs, self.ys, self.seed = n, 1, int(seed)

    def maximum_sample_count(self):
        return 1  # halton.rs:33


class DirectLightingIntegrator:
    """Direct lighting: emitted light plus, per light, the mean of
    light_samples light samples (shadow rays) -- pbrt-v2 DirectLighting with
    light sampling only; the reference has none (integrator/mod.rs:148-168)."""
    kind = 0

    def __init__(self, light_samples=4):
        self.light_samples = int(light_samples)


class PathIntegrator:
    """Path tracing with one light sample per vertex, Lambertian / mirror BSDF
    sampling and Russian roulette once the bounce index exceeds rr_start."""
    kind = 1

    def __init__(self, max_depth=5, rr_start=3):
        self.max_depth, self.rr_start = int(max_depth), int(rr_start)


def _render_desc(scene, camera, sampler, integrator, split="sah"):
    """pbrtx_render_desc of a scene dict (see scenes.py), a camera dict
    (scenes.pinhole, or None: scene only), a sampler and an integrator; also
    returns the arrays it points into (keep them alive)."""
    from ._lib import RenderDesc
    keep = dict(
        verts=np.ascontiguousarray(scene["verts"], F32).reshape(-1),
        idx=np.ascontiguousarray(scene["idx"], np.uint32).reshape(-1),
        face_mat=np.ascontiguousarray(scene["face_mat"], np.uint32),
        mats=np.ascontiguousarray(scene["mats"], F32).reshape(-1),
        face_light=np.ascontiguousarray(scene["face_light"], np.int32),
        area_le=np.ascontiguousarray(scene["area_le"], F32).reshape(-1),
        point_lights=np.ascontiguousarray(scene["point_lights"], F32).reshape(-1),
    )
    k = keep
    d = RenderDesc()
    d.verts, d.nverts = k["verts"].ctypes.data, k["verts"].size // 3
    d.idx, d.nfaces = k["idx"].ctypes.data, k["idx"].size // 3
    d.face_mat, d.mats, d.nmats = k["face_mat"].ctypes.data, k["mats"].ctypes.data, k["mats"].size // 4
    d.face_light = k["face_light"].ctypes.data
    d.area_le, d.n_area = (k["area_le"].ctypes.data if k["area_le"].size else None), k["area_le"].size // 3
    d.point_lights = k["point_lights"].ctypes.data if k["point_lights"].size else None
    d.n_point = k["point_lights"].size // 6
    sph = scene.get("spheres")
    if sph is not None and len(sph):
        keep["spheres"] = sphere_descs(sph)
        keep["sphere_mat"] = np.ascontiguousarray(scene["sphere_mat"], np.uint32)
        if keep["sphere_mat"].size != len(sph):
            raise ValueError("one material per sphere")
        d.spheres = ctypes.cast(keep["spheres"], ctypes.c_void_p)
        d.nspheres = len(sph)
        d.sphere_mat = keep["sphere_mat"].ctypes.data
    if camera is not None:
        _set_camera(d, camera)
    if sampler is not None:
        d.nx, d.ny, d.seed = sampler.xs, sampler.ys, sampler.seed
        d.sampler = getattr(sampler, "kind", 0)
    d.integrator = integrator.kind
    d.light_samples = getattr(integrator, "light_samples", 1)
    d.max_depth = getattr(integrator, "max_depth", 5)
    d.rr_start = getattr(integrator, "rr_start", 3)
    d.split = {"sah": 0, "middle": 1, "equal": 2, "lbvh": 3}[split]
    d.max_prims = 1
    return d, keep


def _set_camera(d, camera):
    for i, v in enumerate(np.asarray(camera["cam_to_world"], F32).reshape(16)):
        d.cam_to_world[i] = float(v)
    d.fov_deg = camera["fov"]
    for i, v in enumerate(camera["screen"]):
        d.screen[i] = float(v)
    d.xres, d.yres = camera["xres"], camera["yres"]


class Filter:
    """A film's filter (filter.rs:44-124): the same constructors and
    evaluate(); the film samples it through Film::image's 16 x 16 weight
    table (film.rs:89-97, table())."""
    KINDS = {"mean": 0, "triangle": 1, "gaussian": 2, "mitchell": 3, "lanczos": 4}

    def __init__(self, kind, xw, yw, p0=0.0, p1=0.0):
        self.kind, self.xw, self.yw, self.p0, self.p1 = int(kind), float(xw), float(yw), float(p0), float(p1)
        self.evaluate(0.0, 0.0)  # rejects an unknown kind or a width that is not positive

    @classmethod
    def mean(cls, xw, yw):
        return cls(0, xw, yw)

    @classmethod
    def triangle(cls, xw, yw):
        return cls(1, xw, yw)

    @classmethod
    def gaussian(cls, xw, yw, alpha):
        return cls(2, xw, yw, alpha)

    @classmethod
    def mitchell(cls, xw, yw, b, c):
        return cls(3, xw, yw, b, c)

    @classmethod
    def lanczos(cls, xw, yw, tau):
        return cls(4, xw, yw, tau)

    def args(self):
        return (self.kind, self.xw, self.yw, self.p0, self.p1)

    def evaluate(self, x, y):
        out = ctypes.c_float()
        check(lib().pbrtx_filter_eval(*self.args(), float(x), float(y), ctypes.byref(out)))
        return out.value

    def table(self):
        t = np.zeros(256, F32)
        check(lib().pbrtx_filter_table(*self.args(), t.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return t.reshape(16, 16)


def write_image(filename, rgb):
    """Film::write_image's file output (film.rs:15-33, 316-354) of an rgb
    image (yres x xres x 3 floats): .pfm, .ppm or .png by extension; host
    only."""
    a = np.ascontiguousarray(rgb, F32)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("rgb must be yres x xres x 3")
    check(lib().pbrtx_write_image(os.fsencode(filename), a.ctypes.data, a.shape[1], a.shape[0]))


def encode8(rgb):
    """write_img's per-channel 8-bit encoding (film.rs:22-24) of floats."""
    a = np.ascontiguousarray(rgb, F32)
    out = np.zeros(a.shape, np.uint8)
    check(lib().pbrtx_image_encode8(a.ctypes.data, a.size, out.ctypes.data))
    return out


class SamplerRenderer:
    """Renderer::render for a generated scene dict (see scenes.py), a camera
    dict (scenes.pinhole) and an integrator; the frame runs on the GPU.
    filter: the film's Filter (default Filter.mean(0.5, 0.5))."""

    def __init__(self, scene, camera, sampler, integrator, split="sah", chunk=None, filter=None):
        d, self._keep = _render_desc(scene, camera, sampler, integrator, split)
        self.desc = d
        self.xres, self.yres = d.xres, d.yres
        h = ctypes.c_void_p()
        check(lib().pbrtx_renderer_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h.value
        k = ctypes.c_uint32()
        check(lib().pbrtx_renderer_sample_slots(self._h, ctypes.byref(k)))
        self.spp = k.value  # sample slots per pixel of the frame layout
        if chunk:
            check(lib().pbrtx_renderer_set_chunk(self._h, int(chunk)))
        if filter is not None:
            self.set_filter(filter)

    def set_filter(self, f):
        """The film's filter (a Filter)."""
        check(lib().pbrtx_renderer_set_filter(self._h, *f.args()))

    def set_chunk(self, samples):
        """Samples per work chunk (whole pixels): the capacity of the ray
        queues and work buffers of one pass over the frame."""
        check(lib().pbrtx_renderer_set_chunk(self._h, int(samples)))

    def set_graph(self, enable=True):
        """Graph mode: after one ordinary frame of a path, frames are replayed
        from a captured hipGraph (per-stage times are then not measured)."""
        check(lib().pbrtx_renderer_set_graph(self._h, 1 if enable else 0))

    def set_octant_sort(self, enable=True):
        """Path tracer, batched path: octant-sort each bounce's extension-ray
        queue (stable within tiles of 4096 entries) before tracing it."""
        check(lib().pbrtx_renderer_set_octant_sort(self._h, 1 if enable else 0))

    def render(self, path="batched", want_image=True):
        from ._lib import RenderStats
        st = RenderStats()
        out = np.zeros((self.yres, self.xres, 3), F32) if want_image else None
        check(lib().pbrtx_renderer_render(self._h, _path_code(path), out.ctypes.data if want_image else None,
                                          ctypes.byref(st)))
        stats = {f: getattr(st, f) for f, _ in st._fields_}
        return out, stats

    def write_image(self, filename):
        """Film::write_image of the last frame: .pfm (floats), .ppm or .png
        (8-bit, write_img's encoding)."""
        check(lib().pbrtx_renderer_write_image(self._h, os.fsencode(filename)))

    def sample_radiance(self, count=None):
        n = self.xres * self.yres * self.spp if count is None else int(count)
        out = np.zeros(n * 3, F32)
        check(lib().pbrtx_renderer_sample_radiance(self._h, out.ctypes.data, n))
        return out.reshape(n, 3)

    def close(self):
        if getattr(self, "_h", None):
            lib().pbrtx_renderer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
