"""The C-ABI library loads on a machine without a GPU and exports every
function include/pbrtx_gpu.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(pbrtx_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "pbrtx_intersect" in names and "pbrtx_intersect_p" in names
    assert "pbrtx_bvh_build" in names


def test_library_exports_every_declared_symbol():
    import pbrt_rust_amd
    L = pbrt_rust_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_errors_are_codes_not_aborts():
    import pbrt_rust_amd as px
    L = px.lib()
    h = ctypes.c_void_p()
    # null output pointer, empty scene, bad split: negative codes + message
    assert L.pbrtx_bvh_build(None, 0, None, 0, None, 0, 0, 1, None) < 0
    assert L.pbrtx_bvh_build(None, 0, None, 0, None, 0, 0, 1, ctypes.byref(h)) < 0
    assert b"no primitives" in L.pbrtx_last_error()
    import numpy as np
    v = np.zeros((3, 3), np.float32)
    i = np.array([0, 1, 5], np.uint32)
    assert L.pbrtx_bvh_build(v.ctypes.data, 3, i.ctypes.data, 1, None, 0, 0, 1, ctypes.byref(h)) < 0
    assert b"out of range" in L.pbrtx_last_error()
    i = np.array([0, 1, 2], np.uint32)
    assert L.pbrtx_bvh_build(v.ctypes.data, 3, i.ctypes.data, 1, None, 0, 7, 1, ctypes.byref(h)) < 0
    assert L.pbrtx_set_tuning(0, 0, 100, 0, 50, 0) < 0
    assert L.pbrtx_set_tuning(1, 3, 256, 0, 50, 0) < 0
    assert L.pbrtx_set_tuning(1, 1, 256, 0, 101, 0) < 0
    assert L.pbrtx_set_tuning(0, 1, 256, 0, 10, 1) < 0   # packet: batched only
    # occlusion has no packet kernel; closest hit and camera rays keep theirs
    assert L.pbrtx_set_tuning(1, 1, 256, 0, 0, 1) < 0 and b"no packet scheme" in L.pbrtx_last_error()
    assert L.pbrtx_set_tuning(1, 2, 256, 16, 0, 1) == 0
    assert L.pbrtx_set_tuning(1, 0, 256, 0, 0, 1) == 0
    assert L.pbrtx_reset_tuning() == 0
    assert L.pbrtx_set_tuning(1, 1, 256, 0, 10, 3) < 0
    assert L.pbrtx_intersect(None, None, 1, None, 0, None) < 0
    assert L.pbrtx_set_stack_depth(1, 1, 12) < 0 and b"stack depth" in L.pbrtx_last_error()
    assert L.pbrtx_set_stack_depth(1, 3, 16) < 0
    assert L.pbrtx_set_ray_lanes(3) < 0 and b"lanes per ray" in L.pbrtx_last_error()
    assert L.pbrtx_set_ray_lanes(32) < 0 and L.pbrtx_set_ray_lanes(2) < 0
    assert L.pbrtx_set_ray_lanes(8) == 0 and L.pbrtx_set_ray_lanes(0) == 0
    from pbrt_rust_amd._lib import RenderServerConfig
    assert L.pbrtx_render_server_create(ctypes.byref(RenderServerConfig(1 << 30, 1 << 20, 7)), ctypes.byref(h)) < 0
    assert L.pbrtx_render_server_create(ctypes.byref(RenderServerConfig(1 << 30, 1000, 0)), ctypes.byref(h)) < 0
    assert b"batch_samples" in L.pbrtx_last_error()
    assert L.pbrtx_render_server_create(ctypes.byref(RenderServerConfig(1 << 10, 1 << 20, 0)), ctypes.byref(h)) < 0
