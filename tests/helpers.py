"""Test helpers: moving (rotated / sliced) scalar arrays between host and GPU."""
from __future__ import annotations

import numpy as np

from zkpoly_amd import hip
from zkpoly_amd.runtime import ScalarArray


def device_array(n_elems: int) -> tuple[hip.DeviceBuffer, ScalarArray]:
    buf = hip.DeviceBuffer(n_elems * 32)
    return buf, ScalarArray(n_elems, buf.addr)


def cpu2gpu(host: np.ndarray, arr: ScalarArray, buf: hip.DeviceBuffer) -> None:
    """Writes logical elements host[i] into arr (honouring rotation/slice)."""
    host = np.ascontiguousarray(host, dtype=np.uint64).reshape(-1, 4)
    assert host.shape[0] == arr.len
    whole = arr.slice_info.whole_len if arr.slice_info else arr.len
    full = buf.download(np.uint64, whole * 4).reshape(whole, 4)
    idx = np.array([arr.physical_index(i) for i in range(arr.len)], dtype=np.int64)
    full[idx] = host
    buf.upload(full)


def gpu2cpu(arr: ScalarArray, buf: hip.DeviceBuffer) -> np.ndarray:
    whole = arr.slice_info.whole_len if arr.slice_info else arr.len
    full = buf.download(np.uint64, whole * 4).reshape(whole, 4)
    idx = np.array([arr.physical_index(i) for i in range(arr.len)], dtype=np.int64)
    return full[idx].copy()


def upload(host: np.ndarray) -> tuple[hip.DeviceBuffer, ScalarArray]:
    host = np.ascontiguousarray(host, dtype=np.uint64).reshape(-1, 4)
    buf, arr = device_array(host.shape[0])
    buf.upload(host)
    return buf, arr


def download(arr: ScalarArray, buf: hip.DeviceBuffer) -> np.ndarray:
    return buf.download(np.uint64, arr.len * 4).reshape(arr.len, 4)
