"""zkp_batched_invert_noalloc: batched_invert with the library's static scratch.

Called through the C ABI (libpoly.so).  Each call's output must equal the
output of batched_invert (the entry point with a caller-provided temporary
buffer) on the same input, bit for bit; two calls in a row, the second with a
smaller input, so the second runs over scratch the first left behind.  The
small case is also compared with the oracle.  Views above the stated maximum
(2^26 elements) are refused on the host, before any launch.
"""
import ctypes

import numpy as np
import pytest

import oracle as orc
from helpers import cpu2gpu, device_array, gpu2cpu
from zkpoly_amd import hip
from zkpoly_amd.field import random_elements
from zkpoly_amd.poly import PolyInvert, load_lib
from zkpoly_amd.runtime import GpuBuffer, PolyPtr

pytestmark = pytest.mark.gpu
HIP_INVALID_VALUE = 1
MAX_LEN = 1 << 26


def noalloc_fn(libs):
    lib = load_lib(libs)
    f = lib.zkp_batched_invert_noalloc
    f.argtypes = [PolyPtr, ctypes.c_void_p]
    f.restype = ctypes.c_int
    return f


def with_temp(libs, stream, x, rotate):
    """batched_invert with its own temporary buffer on a rotated copy of x."""
    n = x.shape[0]
    op = PolyInvert(libs)
    size = op.get_buffer_size(n)
    tb = hip.DeviceBuffer(size)
    buf, arr = device_array(n)
    arr.rotate(rotate)
    cpu2gpu(x, arr, buf)
    op.get_fn().f([GpuBuffer(tb.addr, size), arr], [stream])
    stream.sync()
    return gpu2cpu(arr, buf)


def without_temp(libs, stream, x, rotate):
    n = x.shape[0]
    buf, arr = device_array(n)
    arr.rotate(rotate)
    cpu2gpu(x, arr, buf)
    print(f"zkp_batched_invert_noalloc n={n} rotate={rotate}", flush=True)
    rc = noalloc_fn(libs)(PolyPtr.from_array(arr), stream.raw())
    assert rc == 0, rc
    stream.sync()
    return gpu2cpu(arr, buf)


def test_noalloc_matches_batched_invert_twice(libs, stream):
    big = random_elements((1 << 20) + 5, 901)
    big[::13] = 0
    big[-1] = 0
    small = random_elements(3001, 902)
    small[::7] = 0
    for x, rot in ((big, 12345), (small, 17)):
        got = without_temp(libs, stream, x, rot)
        want = with_temp(libs, stream, x, rot)
        assert np.array_equal(got, want), x.shape[0]
    assert np.array_equal(got, orc.batch_invert(small))


def test_noalloc_refuses_above_max(libs, stream):
    buf, arr = device_array(1)
    over = PolyPtr(arr.values, MAX_LEN + 1, 0, 0, MAX_LEN + 1)
    assert noalloc_fn(libs)(over, stream.raw()) == HIP_INVALID_VALUE
