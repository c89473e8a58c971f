"""Build recipe for libpbrtx_gpu.so (gfx950) -- hipcc directly, in-tree.

    python -m pbrt_rust_amd.build        # or __graft_entry__.build()

Floating point: -ffp-contract=off (no fused multiply-add: the reference's
Rust arithmetic never fuses), correctly rounded f32 divide and sqrt, no
denormal flushing, no fast-math; packed-FP32 instructions disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libpbrtx_gpu.so")

SOURCES = ["bvh_build.cpp", "image_out.cpp", "trace.hip", "capi.hip", "render.hip", "serve.hip", "render_serve.hip",
           "bvh_gpu.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PBRTX_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-I" + os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function",
]
DEVICE = [
    "--offload-arch=" + ARCH, "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace(".", "_") + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + COMMON + DEVICE + ["-c", path, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    shutil.move(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
