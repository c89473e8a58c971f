"""The pbrt scene-description API (the `Pbrt` state machine of main.rs:355-842:
transforms, attribute blocks, render options, graphics state) and a reader
for pbrt's scene-file syntax, feeding the GPU renderer.

    p = Pbrt()
    p.look_at(0, 1, -4,  0, 0, 0,  0, 1, 0)
    p.camera("perspective", {"fov": 45})
    p.film("image", {"xresolution": 320, "yresolution": 240})
    p.sampler("lowdiscrepancy", {"pixelsamples": 16})
    p.surface_integrator("path", {"maxdepth": 5})
    p.world_begin()
    ...shapes, materials, lights...
    renderer = p.world_end()          # a SamplerRenderer, ready to render

or `load("scene.pbrt")` / `parse(text)`.  Supported: Camera "perspective";
Film "image"; Sampler "stratified" (jittered), "lowdiscrepancy" and "halton";
PixelFilter "box", "triangle", "gaussian", "mitchell", "sinc"; SurfaceIntegrator
"directlighting" and "path"; Shape "trianglemesh" and "sphere"; Material "matte" and
"mirror"; LightSource "point"; AreaLightSource "diffuse"; the transform
directives and AttributeBegin/End, TransformBegin/End.  The reference's
make_camera, make_renderer and parse_file are unimplemented (main.rs:200-202,
802); what they would build is taken from pbrt-v2's defaults.  Transforms are
composed in float32 as transform.rs does (Matrix4x4::mul, Transform::rotate,
look_at and the LU inverse of matrix4x4.rs:66-160), each product summed in
index order, sin and cos from the C library as f32::sin/cos.  Beyond the
reference's make_shape / make_material (main.rs:288-330: "sphere" and "matte"
only): "trianglemesh" and "mirror" (the renderer's second material);
ReverseOrientation; spheres are not area lights.
"""
import ctypes
import ctypes.util
import math
import re
import warnings

import numpy as np

from . import api

F32 = np.float32


# ---- float32 transforms (transform.rs:19-174, matrix4x4.rs) ----------------
_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.sinf.argtypes = [ctypes.c_float]
_libm.sinf.restype = ctypes.c_float
_libm.cosf.argtypes = [ctypes.c_float]
_libm.cosf.restype = ctypes.c_float
_PI = F32(math.pi)


def _mul(a, b):
    """Matrix4x4::mul: r[i][j] = a[i][0]*b[0][j] + ... + a[i][3]*b[3][j],
    float32, summed left to right."""
    r = np.zeros((4, 4), F32)
    for i in range(4):
        for j in range(4):
            s = F32(a[i, 0] * b[0, j])
            for k in range(1, 4):
                s = F32(s + F32(a[i, k] * b[k, j]))
            r[i, j] = s
    return r


def _inverse(m):
    """Matrix4x4::invert (matrix4x4.rs:66-160): LU with scaled partial
    pivoting, then one solve per unit vector, in float32."""
    m = np.asarray(m, F32)
    sc = [max(F32(abs(v)) for v in row) for row in m]
    sc = [max(F32(0), v) for v in sc]
    lu = m.copy()
    piv = [0, 1, 2, 3]
    for k in range(3):
        c, p = F32(0), k
        for i in range(k, 4):
            cc = F32(abs(F32(lu[i, k] / sc[i])))
            if cc > c:
                c, p = cc, i
        piv[k] = p
        if c == 0:
            raise ValueError("singular matrix")
        if p != k:
            for j in range(k, 4):
                lu[k, j], lu[p, j] = lu[p, j], lu[k, j]
        for i in range(k + 1, 4):
            mi = F32(lu[i, k] / lu[k, k])
            lu[i, k] = mi
            for j in range(k + 1, 4):
                lu[i, j] = F32(lu[i, j] - F32(mi * lu[k, j]))
    if abs(lu[3, 3]) < F32(1e-6):
        raise ValueError("singular matrix")
    cols = []
    for e in range(4):
        r = [F32(1) if q == e else F32(0) for q in range(4)]
        for k in range(3):
            if piv[k] != k:
                r[piv[k]], r[k] = r[k], r[piv[k]]
            for i in range(k + 1, 4):
                r[i] = F32(r[i] - F32(lu[i, k] * r[k]))
        r[3] = F32(r[3] / lu[3, 3])
        for i in (2, 1, 0):
            acc = F32(0)
            for j in range(i + 1, 4):
                acc = F32(acc + F32(lu[i, j] * r[j]))
            r[i] = F32(F32(r[i] - acc) / lu[i, i])
        cols.append(r)
    return np.array(cols, F32).T.copy()


def _normalize(v):
    """Vector::normalize: v * (1 / length), length = sqrt(x*x + y*y + z*z)."""
    x, y, z = (F32(c) for c in v)
    ln = F32(np.sqrt(F32(F32(F32(x * x) + F32(y * y)) + F32(z * z))))
    r = F32(F32(1) / ln)
    return np.array([F32(r * x), F32(r * y), F32(r * z)], F32)


def _cross(a, b):
    return np.array([F32(F32(a[1] * b[2]) - F32(a[2] * b[1])), F32(F32(a[2] * b[0]) - F32(a[0] * b[2])),
                     F32(F32(a[0] * b[1]) - F32(a[1] * b[0]))], F32)


class Xf:
    """A transform with its inverse, as Transform{m, m_inv}."""

    def __init__(self, m, m_inv):
        self.m, self.m_inv = np.asarray(m, F32), np.asarray(m_inv, F32)

    @staticmethod
    def identity():
        return Xf(np.eye(4, dtype=F32), np.eye(4, dtype=F32))

    @staticmethod
    def from_matrix(m):
        m = np.asarray(m, F32).reshape(4, 4)
        return Xf(m, _inverse(m))

    def __mul__(self, o):
        return Xf(_mul(self.m, o.m), _mul(o.m_inv, self.m_inv))

    def inverse(self):
        return Xf(self.m_inv, self.m)

    @staticmethod
    def translate(x, y, z):
        m, mi = np.eye(4, dtype=F32), np.eye(4, dtype=F32)
        m[:3, 3] = (x, y, z)
        mi[:3, 3] = (-F32(x), -F32(y), -F32(z))
        return Xf(m, mi)

    @staticmethod
    def scale(x, y, z):
        m = np.diag(np.array([x, y, z, 1], F32))
        mi = np.diag(np.array([F32(1) / F32(x), F32(1) / F32(y), F32(1) / F32(z), 1], F32))
        return Xf(m, mi)

    @staticmethod
    def rotate(deg, x, y, z):
        """Transform::rotate (transform.rs:117-142)."""
        a = _normalize((x, y, z))
        rad = F32(F32(F32(deg) * _PI) / F32(180))
        s, c = F32(_libm.sinf(rad)), F32(_libm.cosf(rad))
        one = F32(1)
        m = np.eye(4, dtype=F32)
        m[0, 0] = F32(a[0] * a[0]) + F32(F32(one - F32(a[0] * a[0])) * c)
        m[0, 1] = F32(F32(a[0] * a[1]) * F32(one - c)) - F32(a[2] * s)
        m[0, 2] = F32(F32(a[0] * a[2]) * F32(one - c)) + F32(a[1] * s)
        m[1, 0] = F32(F32(a[0] * a[1]) * F32(one - c)) + F32(a[2] * s)
        m[1, 1] = F32(a[1] * a[1]) + F32(F32(one - F32(a[1] * a[1])) * c)
        m[1, 2] = F32(F32(a[1] * a[2]) * F32(one - c)) - F32(a[0] * s)
        m[2, 0] = F32(F32(a[0] * a[2]) * F32(one - c)) - F32(a[1] * s)
        m[2, 1] = F32(F32(a[1] * a[2]) * F32(one - c)) + F32(a[0] * s)
        m[2, 2] = F32(a[2] * a[2]) + F32(F32(one - F32(a[2] * a[2])) * c)
        return Xf(m, m.T.copy())

    @staticmethod
    def look_at(eye, look, up):
        """Transform::look_at (transform.rs:144-174): world to camera, with the
        camera-to-world matrix as its inverse."""
        eye, look = np.asarray(eye, F32), np.asarray(look, F32)
        d = _normalize((look - eye).astype(F32))
        left = _normalize(_cross(_normalize(up), d))
        new_up = _cross(d, left)
        c2w = np.eye(4, dtype=F32)
        c2w[:3, 0], c2w[:3, 1], c2w[:3, 2], c2w[:3, 3] = left, new_up, d, eye
        return Xf(_inverse(c2w), c2w)


def _param(params, name, default=None):
    v = params.get(name, default)
    if isinstance(v, (list, tuple, np.ndarray)) and len(v) == 1:
        return v[0]
    return v


class Pbrt:
    """The scene-description state machine (main.rs:355-842) feeding the GPU
    renderer."""

    def __init__(self):
        self.ctm = Xf.identity()
        self.named = {}
        self.stack = []
        self.camera_to_world = Xf.identity()
        self.camera_params = {"fov": 90.0}
        self.film_params = {}
        self.sampler_name, self.sampler_params = "lowdiscrepancy", {}   # main.rs:171
        self.filter_name, self.filter_params = "box", {}
        self.integrator_name, self.integrator_params = "directlighting", {}
        self.in_world = False
        self._reset_world()

    def _reset_world(self):
        self.mat_state = ("matte", {})
        self.area_light = None
        self.reverse_orientation = False
        self.spheres, self.sphere_mat = [], []
        self.verts, self.idx, self.face_mat, self.face_light = [], [], [], []
        self.mats, self.area_le, self.points = [], [], []
        self.nverts = 0

    # ---- transforms (main.rs:426-520)
    def identity(self):
        self.ctm = Xf.identity()

    def translate(self, x, y, z):
        self.ctm = self.ctm * Xf.translate(x, y, z)

    def scale(self, x, y, z):
        self.ctm = self.ctm * Xf.scale(x, y, z)

    def rotate(self, angle, x, y, z):
        self.ctm = self.ctm * Xf.rotate(angle, x, y, z)

    def look_at(self, ex, ey, ez, lx, ly, lz, ux, uy, uz):
        self.ctm = self.ctm * Xf.look_at((ex, ey, ez), (lx, ly, lz), (ux, uy, uz))

    def concat_transform(self, m16):
        """rows of the matrix, as main.rs:467-476 takes them"""
        self.ctm = self.ctm * Xf.from_matrix(m16)

    def transform(self, m16):
        self.ctm = Xf.from_matrix(m16)

    def coordinate_system(self, name):
        self.named[name] = self.ctm

    def coord_sys_transform(self, name):
        if name not in self.named:
            raise ValueError("coordinate system %r not declared" % name)
        self.ctm = self.named[name]

    # ---- render options (main.rs:521-565), before WorldBegin
    def camera(self, name, params=None):
        if name != "perspective":
            raise ValueError("camera %r: only \"perspective\"" % name)
        self.camera_params = dict(params or {})
        self.camera_to_world = self.ctm.inverse()
        self.named["camera"] = self.camera_to_world

    def film(self, name, params=None):
        if name != "image":
            raise ValueError("film %r: only \"image\"" % name)
        self.film_params = dict(params or {})

    def sampler(self, name, params=None):
        if name not in ("stratified", "lowdiscrepancy", "halton"):
            raise ValueError("sampler %r: \"stratified\", \"lowdiscrepancy\" or \"halton\"" % name)
        self.sampler_name, self.sampler_params = name, dict(params or {})

    def pixel_filter(self, name, params=None):
        if name not in ("box", "triangle", "gaussian", "mitchell", "sinc"):
            raise ValueError("filter %r not known" % name)
        self.filter_name, self.filter_params = name, dict(params or {})

    def surface_integrator(self, name, params=None):
        if name not in ("directlighting", "path"):
            raise ValueError("surface integrator %r: \"directlighting\" or \"path\"" % name)
        self.integrator_name, self.integrator_params = name, dict(params or {})

    # ---- world (main.rs:394-425, 566-713, 765-800)
    def world_begin(self):
        self.in_world = True
        self.ctm = Xf.identity()
        self.named["world"] = self.ctm
        self._reset_world()

    def attribute_begin(self):
        self.stack.append((self.ctm, self.mat_state, self.area_light, self.reverse_orientation))

    def attribute_end(self):
        if not self.stack:
            raise ValueError("unmatched AttributeEnd")
        self.ctm, self.mat_state, self.area_light, self.reverse_orientation = self.stack.pop()

    def reverse_orientation_toggle(self):
        """ReverseOrientation: the graphics state's flag flips (pbrt-v2
        GraphicsState::reverseOrientation, kept by attribute blocks)."""
        self.reverse_orientation = not self.reverse_orientation

    def transform_begin(self):
        self.stack.append((self.ctm, None, None, None))

    def transform_end(self):
        if not self.stack:
            raise ValueError("unmatched TransformEnd")
        self.ctm = self.stack.pop()[0]

    def material(self, name, params=None):
        if name not in ("matte", "mirror"):
            raise ValueError("material %r: \"matte\" or \"mirror\"" % name)
        self.mat_state = (name, dict(params or {}))

    def light_source(self, name, params=None):
        if name != "point":
            raise ValueError("light %r: only \"point\"" % name)
        p = dict(params or {})
        frm = np.asarray(p.get("from", [0, 0, 0]), F32).reshape(3)
        # make_light (main.rs:298-309): light_to_world = Translate(from) * CTM,
        # the light at its image of the origin; I times "scale"
        l2w = _mul(Xf.translate(*frm).m, self.ctm.m)
        pos = api.Transform(l2w).xf_points(np.zeros((1, 3), F32))[0]
        inten = np.broadcast_to(np.asarray(p.get("I", [1, 1, 1]), F32), (3,)) * \
            np.broadcast_to(np.asarray(p.get("scale", [1, 1, 1]), F32), (3,))
        self.points.append(list(pos) + list(inten.astype(F32)))

    def area_light_source(self, name, params=None):
        if name != "diffuse":
            raise ValueError("area light %r: only \"diffuse\"" % name)
        p = dict(params or {})
        self.area_light = np.broadcast_to(np.asarray(p.get("L", [1, 1, 1]), F32), (3,)).copy()

    def _material_row(self):
        name, p = self.mat_state
        if name == "matte":
            kd = np.broadcast_to(np.asarray(p.get("Kd", [0.5, 0.5, 0.5]), F32), (3,))
            return [0.0] + list(kd)
        kr = np.broadcast_to(np.asarray(p.get("Kr", [1, 1, 1]), F32), (3,))
        return [1.0] + list(kr)

    def _material_index(self):
        row = self._material_row()
        if row not in self.mats:
            self.mats.append(row)
        return self.mats.index(row)

    def shape(self, name, params=None):
        if name == "sphere":
            # make_shape "sphere" (main.rs:314-323): radius 1, z range
            # [-radius, radius], phimax 360 by default; object_to_world = CTM
            if self.area_light is not None:
                raise ValueError("sphere: area light spheres are not rendered (spheres are not emitters here)")
            p = dict(params or {})
            radius = float(_param(p, "radius", 1.0))
            zmin = float(_param(p, "zmin", -radius))
            zmax = float(_param(p, "zmax", radius))
            phimax = float(_param(p, "phimax", 360.0))
            o2w = api.Transform(self.ctm.m, self.ctm.m_inv)
            self.spheres.append(api.Sphere(o2w, o2w.inverse(), self.reverse_orientation, radius, zmin, zmax, phimax))
            self.sphere_mat.append(self._material_index())
            return
        if name != "trianglemesh":
            raise ValueError("shape %r: \"trianglemesh\" or \"sphere\"" % name)
        p = dict(params or {})
        P = np.asarray(p["P"], F32).reshape(-1, 3)
        ind = np.asarray(p["indices"], np.int64).reshape(-1, 3)
        if ind.size and (ind.min() < 0 or ind.max() >= len(P)):
            raise ValueError("trianglemesh: vertex index out of range")
        # Mesh::new pre-transforms the vertices to world space (mesh.rs:308)
        Pw = api.Transform(self.ctm.m).xf_points(P).astype(F32)
        mat = self._material_index()
        light = -1
        if self.area_light is not None:
            self.area_le.append(list(self.area_light))
            light = len(self.area_le) - 1
        self.verts.append(Pw)
        self.idx.append(ind + self.nverts)
        self.nverts += len(P)
        self.face_mat.append(np.full(len(ind), mat, np.uint32))
        self.face_light.append(np.full(len(ind), light, np.int32))

    def scene(self):
        """The scene dict of what the world block declared (see scenes.py)."""
        if not self.verts and not self.spheres:
            raise ValueError("the world holds no shapes")
        cat = lambda xs, shape, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(shape, dt)  # noqa: E731
        return dict(verts=cat(self.verts, (0, 3), F32), idx=cat(self.idx, (0, 3), np.uint32),
                    face_mat=cat(self.face_mat, 0, np.uint32), mats=np.asarray(self.mats, F32).reshape(-1, 4),
                    face_light=cat(self.face_light, 0, np.int32),
                    area_le=np.asarray(self.area_le, F32).reshape(-1, 3),
                    point_lights=np.asarray(self.points, F32).reshape(-1, 6),
                    spheres=list(self.spheres), sphere_mat=np.asarray(self.sphere_mat, np.uint32))

    def camera_dict(self):
        xres = int(_param(self.film_params, "xresolution", 640))
        yres = int(_param(self.film_params, "yresolution", 480))
        frame = xres / yres
        default = [-frame, frame, -1.0, 1.0] if frame > 1 else [-1.0, 1.0, -1.0 / frame, 1.0 / frame]
        screen = [float(v) for v in self.camera_params.get("screenwindow", default)]
        return dict(cam_to_world=self.camera_to_world.m.astype(F32), fov=float(_param(self.camera_params, "fov", 90.0)),
                    screen=screen, xres=xres, yres=yres)

    def sampler_obj(self):
        p, seed = self.sampler_params, int(_param(self.sampler_params, "seed", 0))
        if self.sampler_name == "lowdiscrepancy":
            return api.LDSampler(int(_param(p, "pixelsamples", 4)), seed)
        if self.sampler_name == "halton":
            return api.HaltonSampler(int(_param(p, "pixelsamples", 4)), seed)
        if not bool(_param(p, "jitter", True)):
            raise ValueError("stratified sampler: only jittered strata are rendered")
        return api.StratifiedSampler(int(_param(p, "xsamples", 2)), int(_param(p, "ysamples", 2)), seed)

    def filter_obj(self):
        p, n = self.filter_params, self.filter_name
        defaults = {"box": 0.5, "triangle": 2.0, "gaussian": 2.0, "mitchell": 2.0, "sinc": 4.0}
        xw = float(_param(p, "xwidth", defaults[n]))
        yw = float(_param(p, "ywidth", defaults[n]))
        if n == "box":
            return api.Filter.mean(xw, yw)
        if n == "triangle":
            return api.Filter.triangle(xw, yw)
        if n == "gaussian":
            return api.Filter.gaussian(xw, yw, float(_param(p, "alpha", 2.0)))
        if n == "mitchell":
            return api.Filter.mitchell(xw, yw, float(_param(p, "B", 1 / 3)), float(_param(p, "C", 1 / 3)))
        return api.Filter.lanczos(xw, yw, float(_param(p, "tau", 3.0)))

    def integrator_obj(self):
        p = self.integrator_params
        if self.integrator_name == "path":
            return api.PathIntegrator(int(_param(p, "maxdepth", 5)), int(_param(p, "rrstart", 3)))
        return api.DirectLightingIntegrator(int(_param(p, "lightsamples", 1)))

    def film_filename(self):
        """The Film's "filename" (the image write_image writes); pbrt-v2's
        default names an EXR file, which this build does not write, so the
        default here is "pbrt.png"."""
        return str(_param(self.film_params, "filename", "pbrt.png"))

    def world_end(self, chunk=None):
        """The renderer of the described scene (WorldEnd, main.rs:774-800),
        with the Film's file name as its `filename`."""
        if self.stack:
            warnings.warn("missing AttributeEnd/TransformEnd at WorldEnd")
        self.in_world = False
        r = api.SamplerRenderer(self.scene(), self.camera_dict(), self.sampler_obj(), self.integrator_obj(),
                                chunk=chunk, filter=self.filter_obj())
        r.filename = self.film_filename()
        return r


# ---- scene files -------------------------------------------------------------
_TOKEN = re.compile(r'\s*(?:(#[^\n]*)|("[^"]*")|(\[)|(\])|([^\s\[\]"#]+))')
_TYPES = {"integer", "float", "point", "normal", "vector", "color", "rgb", "string", "bool", "spectrum", "point3",
          "vector3", "normal3", "texture"}


def _tokens(text):
    pos = 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            if text[pos:].strip() == "":
                return
            raise ValueError("unexpected input at %r" % text[pos:pos + 20])
        pos = m.end()
        if m.group(1):
            continue
        yield next(g for g in m.groups()[1:] if g is not None)


def _value(tok):
    if tok.startswith('"'):
        s = tok[1:-1]
        return {"true": True, "false": False}.get(s, s)
    return float(tok)


def parse(text, pbrt=None):
    """Runs a pbrt scene description (text) on a Pbrt; returns the renderer
    WorldEnd made (None without a world block)."""
    p = pbrt or Pbrt()
    toks = list(_tokens(text))
    i = 0
    renderer = None

    def args(k):
        nonlocal i
        out = []
        while len(out) < k:
            t = toks[i]
            i += 1
            if t in "[]":
                continue
            out.append(_value(t))
        if i < len(toks) and toks[i] == "]":
            i += 1
        return out

    def plist():
        nonlocal i
        params = {}
        while i < len(toks) and toks[i].startswith('"') and len(toks[i][1:-1].split()) == 2 \
                and toks[i][1:-1].split()[0] in _TYPES:
            ty, name = toks[i][1:-1].split()
            i += 1
            vals = []
            if toks[i] == "[":
                i += 1
                while toks[i] != "]":
                    vals.append(_value(toks[i]))
                    i += 1
                i += 1
            else:
                vals.append(_value(toks[i]))
                i += 1
            if ty == "integer":
                vals = [int(v) for v in vals]
            params[name] = vals
        return params

    simple = {"Identity": (0, p.identity), "Translate": (3, p.translate), "Scale": (3, p.scale),
              "Rotate": (4, p.rotate), "LookAt": (9, p.look_at)}
    named = {"Camera": p.camera, "Film": p.film, "Sampler": p.sampler, "PixelFilter": p.pixel_filter,
             "SurfaceIntegrator": p.surface_integrator, "Material": p.material, "LightSource": p.light_source,
             "AreaLightSource": p.area_light_source, "Shape": p.shape}
    blocks = {"AttributeBegin": p.attribute_begin, "AttributeEnd": p.attribute_end,
              "ReverseOrientation": p.reverse_orientation_toggle,
              "TransformBegin": p.transform_begin, "TransformEnd": p.transform_end, "WorldBegin": p.world_begin}
    while i < len(toks):
        d = toks[i]
        i += 1
        if d in simple:
            k, fn = simple[d]
            fn(*args(k))
        elif d in ("ConcatTransform", "Transform"):
            # a file lists the matrix column by column (pbrt's format); the
            # API takes rows
            m = np.asarray(args(16), F32).reshape(4, 4).T.reshape(-1)
            (p.concat_transform if d == "ConcatTransform" else p.transform)(m)
        elif d in ("CoordinateSystem", "CoordSysTransform"):
            name = _value(toks[i])
            i += 1
            (p.coordinate_system if d == "CoordinateSystem" else p.coord_sys_transform)(name)
        elif d in named:
            name = _value(toks[i])
            i += 1
            named[d](name, plist())
        elif d in blocks:
            blocks[d]()
        elif d == "WorldEnd":
            renderer = p.world_end()
        else:
            raise ValueError("directive %r is not supported" % d)
    return renderer


def load(path, pbrt=None):
    """parse() of a scene file."""
    with open(path) as f:
        return parse(f.read(), pbrt)


def run(filenames, path="batched"):
    """Pbrt::run (main.rs:828-842): each scene file is read, its world
    rendered and the image written to the Film's file name (relative names
    relative to the working directory).  Returns the image files written."""
    written = []
    for fn in filenames:
        r = load(fn)
        if r is None:
            raise ValueError("%s: no world block" % fn)
        try:
            r.render(path, want_image=False)
            r.write_image(r.filename)
            written.append(r.filename)
        finally:
            r.close()
    return written
