"""The measurement probes behind the bench's measured ceilings: the streaming
device copy moves exactly the bytes asked (every size, including the tail a
thread's four-wide stride leaves), and the fma-chain probe reports the flops
it launched and writes finite values."""
import ctypes

import numpy as np
import pytest

import pbrt_rust_amd as px
from pbrt_rust_amd._lib import check

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n16", [1, 255, 4096, 1 << 20, (1 << 20) + 77, 3 * 256 * 4096 + 5])
def test_device_copy_moves_exactly_the_bytes(n16):
    src = np.random.default_rng(n16).integers(0, 2**32, size=4 * n16, dtype=np.uint32)
    a = px.DeviceBuffer.from_array(src)
    b = px.DeviceBuffer.from_array(np.zeros(4 * n16 + 8, np.uint32))  # guard words after the end
    check(px.lib().pbrtx_device_copy(b.ptr, a.ptr, 16 * n16, None))
    px.synchronize()
    out = b.download(np.uint32, 4 * n16 + 8)
    assert np.array_equal(out[:4 * n16], src)
    assert not out[4 * n16:].any()


def test_fma_probe():
    L = px.lib()
    n = int(L.pbrtx_fma_probe_threads())
    assert n > 0
    out = px.DeviceBuffer(4 * n)
    flops = ctypes.c_uint64()
    check(L.pbrtx_fma_probe(out.ptr, n, 64, ctypes.byref(flops), None))
    px.synchronize()
    assert flops.value == n * 8 * 2 * 64
    v = out.download(np.float32, n)
    assert np.isfinite(v).all()
    with pytest.raises(px.PbrtxError):
        check(L.pbrtx_fma_probe(out.ptr, n - 1, 64, ctypes.byref(flops), None))
