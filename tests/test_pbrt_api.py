"""The scene-description API and scene-file reader (pbrt_rust_amd.pbrt_api;
main.rs:355-842): float32 transforms as transform.rs / matrix4x4.rs compute
them (pinned by the reference's own inversion test and closed forms), the
graphics and render state a file sets up, and the scene it hands the renderer
(host only; the GPU test renders a file against the oracle)."""
import numpy as np
import pytest

from pbrt_rust_amd import pbrt_api as pa

F32 = np.float32


def check_mat(a, b, tol=1e-5):
    assert np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).max() <= tol


def test_reference_inversion_test():
    # matrix4x4.rs:471-503
    m = np.array([[1, -2, 3, 0], [2, -5, 12, 0], [0, 2, -10, 0], [0, 0, 0, 1]], F32)
    check_mat(pa._inverse(np.eye(4, dtype=F32)), np.eye(4))
    check_mat(pa._mul(m, pa._inverse(m)), np.eye(4))
    check_mat(pa._mul(pa._inverse(m), m), np.eye(4))
    n = np.array([[2, 3, 1, 5], [1, 0, 3, 1], [0, 2, -3, 2], [0, 2, 3, 1]], F32)
    n_inv = np.array([[18, -35, -28, 1], [9, -18, -14, 1], [-2, 4, 3, 0], [-12, 24, 19, -1]], F32)
    check_mat(pa._inverse(n), n_inv, 1e-4)
    m2 = np.array([[-0.70710677, -0.40824828, -0.57735026, 1.0], [0.0, 0.81649655, -0.57735026, 1.0],
                   [0.70710677, -0.40824828, -0.57735026, 1.0], [0.0, 0.0, 0.0, 1.0]], F32)
    m2_inv = np.array([[-0.70710677, 0.0, 0.70710677, 0.0], [-0.40824828, 0.81649655, -0.40824828, 0.0],
                       [-0.57735026, -0.57735026, -0.57735026, 1.73205], [0, 0, 0, 1]], F32)
    check_mat(pa._inverse(m2), m2_inv)
    with pytest.raises(ValueError):
        pa._inverse(np.array([[32, 8, 11, 17], [8, 20, 17, 23], [11, 17, 14, 26], [17, 23, 26, 2]], F32))


def test_transform_closed_forms():
    t = pa.Xf.translate(1, 2, 3) * pa.Xf.scale(2, 4, 8)
    assert np.array_equal(t.m, np.array([[2, 0, 0, 1], [0, 4, 0, 2], [0, 0, 8, 3], [0, 0, 0, 1]], F32))
    assert np.array_equal(pa._mul(t.m, t.m_inv), np.eye(4, dtype=F32))
    r = pa.Xf.rotate(90, 0, 0, 1)
    check_mat(r.m[:3, 0], [0, 1, 0], 1e-7)
    check_mat(pa._mul(r.m, r.m_inv), np.eye(4), 1e-6)
    # looking down +z from the origin: camera and world axes agree
    la = pa.Xf.look_at((0, 0, 0), (0, 0, 1), (0, 1, 0))
    assert np.array_equal(la.m, np.eye(4, dtype=F32)) and np.array_equal(la.m_inv, np.eye(4, dtype=F32))
    la = pa.Xf.look_at((1, 2, -4), (0, 0, 0), (0, 1, 0))
    check_mat(pa._mul(la.m, la.m_inv), np.eye(4), 1e-6)
    c2w = la.m_inv
    d = np.array([-1, -2, 4], np.float64) / np.sqrt(21)
    check_mat(c2w[:3, 2], d, 1e-6)     # camera +z toward the target
    check_mat(c2w[:3, 3], [1, 2, -4], 0)


SCENE = """
# a quad light over a ground quad and one triangle
LookAt 0 1 -4  0 0 0  0 1 0
Camera "perspective" "float fov" [ 50 ]
Film "image" "integer xresolution" [ 48 ] "integer yresolution" 32
Sampler "stratified" "integer xsamples" 2 "integer ysamples" [2] "bool jitter" "true"
PixelFilter "gaussian" "float xwidth" 2 "float ywidth" 2 "float alpha" 2
SurfaceIntegrator "directlighting" "integer lightsamples" [ 2 ]
WorldBegin
  LightSource "point" "rgb I" [ 3 3 3 ] "point from" [ 0 4 0 ]
  AttributeBegin
    AreaLightSource "diffuse" "rgb L" [ 10 10 10 ]
    Translate 0 2.5 0
    Shape "trianglemesh" "point P" [ -0.5 0 -0.5  -0.5 0 0.5  0.5 0 0.5  0.5 0 -0.5 ]
          "integer indices" [ 0 1 2  0 2 3 ]
  AttributeEnd
  Material "matte" "rgb Kd" [ 0.8 0.8 0.8 ]
  Shape "trianglemesh" "point P" [ -5 -1 -5  -5 -1 5  5 -1 5  5 -1 -5 ] "integer indices" [ 0 1 2 0 2 3 ]
  AttributeBegin
    Material "mirror" "rgb Kr" [ 0.9 0.9 0.9 ]
    Rotate 90 0 1 0
    Shape "trianglemesh" "point P" [ 0 0 0  1 0 0  0 1 0 ] "integer indices" [ 0 1 2 ]
  AttributeEnd
WorldEnd
"""


class Recorder(pa.Pbrt):
    def world_end(self, chunk=None):  # what the file declared, without a device
        return self


def test_file_builds_the_scene():
    p = pa.parse(SCENE, Recorder())
    sc = p.scene()
    assert sc["idx"].shape == (5, 3) and sc["verts"].shape == (11, 3)
    # the light quad moved up by the Translate inside its attribute block
    assert np.array_equal(sc["verts"][:4, 1], np.full(4, 2.5, F32))
    assert list(sc["face_light"]) == [0, 0, -1, -1, -1] and sc["area_le"].tolist() == [[10, 10, 10]]
    # graphics state: default matte, then matte 0.8, then mirror in a block
    assert sc["mats"].tolist() == [[0, 0.5, 0.5, 0.5], [0, 0.800000011920929, 0.800000011920929, 0.800000011920929],
                                   [1, 0.8999999761581421, 0.8999999761581421, 0.8999999761581421]]
    assert list(sc["face_mat"]) == [0, 0, 1, 1, 2]
    # the rotated triangle: (1,0,0) -> (0,0,-1) under Rotate 90 about y
    check_mat(sc["verts"][9], [0, 0, -1], 1e-7)
    # point light at Translate(from) applied to the world origin; I as given
    assert sc["point_lights"].tolist() == [[0, 4, 0, 3, 3, 3]]
    cam = p.camera_dict()
    assert (cam["xres"], cam["yres"], cam["fov"]) == (48, 32, 50.0)
    assert cam["screen"] == [-1.5, 1.5, -1.0, 1.0]
    assert np.array_equal(cam["cam_to_world"], pa.Xf.look_at((0, 1, -4), (0, 0, 0), (0, 1, 0)).m_inv)
    s = p.sampler_obj()
    assert (s.xs, s.ys, getattr(s, "kind", 0)) == (2, 2, 0)
    f = p.filter_obj()
    assert f.args() == (2, 2.0, 2.0, 2.0, 0.0)
    assert p.integrator_obj().light_samples == 2


def test_defaults_and_low_discrepancy():
    p = pa.parse('Sampler "lowdiscrepancy" "integer pixelsamples" 6\nWorldBegin\n'
                 'Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 2]\nWorldEnd', Recorder())
    assert p.sampler_obj().xs == 8  # rounded up to a power of two
    cam = p.camera_dict()
    assert (cam["xres"], cam["yres"], cam["fov"]) == (640, 480, 90.0)
    assert p.filter_obj().args() == (0, 0.5, 0.5, 0.0, 0.0)


@pytest.mark.parametrize("text", [
    'WorldBegin AreaLightSource "diffuse" Shape "sphere" "float radius" 1 WorldEnd',
    'WorldBegin AttributeEnd',
    'Bogus 1 2 3',
    'WorldBegin Shape "trianglemesh" "point P" [0 0 0 1 0 0 0 1 0] "integer indices" [0 1 3] WorldEnd',
    'Camera "orthographic"',
])
def test_unsupported_or_malformed(text):
    with pytest.raises((ValueError, KeyError)):
        pa.parse(text, Recorder())


def test_transform_file_convention():
    # a file's Transform lists columns; the API takes rows
    p = Recorder()
    assert pa.parse("ConcatTransform [1 0 0 0  0 1 0 0  0 0 1 0  5 6 7 1]", p) is None  # no world block
    assert p.ctm.m[:3, 3].tolist() == [5, 6, 7]
    q = Recorder()
    q.concat_transform([1, 0, 0, 5, 0, 1, 0, 6, 0, 0, 1, 7, 0, 0, 0, 1])
    assert np.array_equal(p.ctm.m, q.ctm.m)


def test_sphere_shapes_and_reverse_orientation():
    # make_shape "sphere" (main.rs:314-323): defaults radius 1, z range
    # [-r, r], phimax 360; object_to_world is the CTM; ReverseOrientation is
    # kept by attribute blocks
    text = """WorldBegin
    AttributeBegin
      Translate 1 2 3
      ReverseOrientation
      Material "mirror" "rgb Kr" [ 0.5 0.5 0.5 ]
      Shape "sphere" "float radius" 2 "float zmin" -1 "float phimax" 180
    AttributeEnd
    Shape "sphere"
    WorldEnd"""
    p = pa.parse(text, Recorder())
    sc = p.scene()
    assert sc["idx"].shape == (0, 3) and len(sc["spheres"]) == 2
    a, b = sc["spheres"]
    assert (a.radius, a.z0, a.z1, a.phimax, a.reverse_orientation) == (2.0, -1.0, 2.0, 180.0, True)
    assert (b.radius, b.z0, b.z1, b.phimax, b.reverse_orientation) == (1.0, -1.0, 1.0, 360.0, False)
    assert np.array_equal(a.o2w.m[:3, 3], [1, 2, 3]) and np.array_equal(a.o2w.m_inv[:3, 3], [-1, -2, -3])
    assert np.array_equal(b.o2w.m, np.eye(4, dtype=F32))
    assert sc["mats"][sc["sphere_mat"][0]].tolist() == [1, 0.5, 0.5, 0.5]
    assert sc["mats"][sc["sphere_mat"][1]].tolist() == [0, 0.5, 0.5, 0.5]
