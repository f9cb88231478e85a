"""Minimal ctypes binding of the HIP runtime (libamdhip64) used as plumbing.

Device buffers, streams and events for the host-side mirror of the reference
runtime.  The reference's equivalents are cuda_api's CudaStream / CudaEvent /
allocators (cuda_api/src/stream.rs:10-257, cuda_api/src/mem.rs:14-34); here
they are thin wrappers: every call checks its hipError_t and raises HipError.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HIP = None

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3


class HipError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        super().__init__(f"{where}: hipError {code} ({error_string(code)})")


def _lib():
    global _HIP
    if _HIP is None:
        candidates = [
            os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"),
            "libamdhip64.so",
        ]
        last = None
        for c in candidates:
            try:
                _HIP = ctypes.CDLL(c, mode=ctypes.RTLD_GLOBAL)
                break
            except OSError as e:  # pragma: no cover - depends on the machine
                last = e
        if _HIP is None:
            raise OSError(f"cannot load libamdhip64: {last}")
        h = _HIP
        vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        h.hipFree.argtypes = [vp]
        h.hipHostMalloc.argtypes = [ctypes.POINTER(vp), sz, ctypes.c_uint]
        h.hipHostFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, sz, i]
        h.hipMemcpyAsync.argtypes = [vp, vp, sz, i, vp]
        h.hipMemset.argtypes = [vp, i, sz]
        h.hipMemsetAsync.argtypes = [vp, i, sz, vp]
        h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipStreamDestroy.argtypes = [vp]
        h.hipStreamSynchronize.argtypes = [vp]
        h.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipEventDestroy.argtypes = [vp]
        h.hipEventRecord.argtypes = [vp, vp]
        h.hipStreamWaitEvent.argtypes = [vp, vp, ctypes.c_uint]
        h.hipEventSynchronize.argtypes = [vp]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        h.hipGetDeviceCount.argtypes = [ctypes.POINTER(i)]
        h.hipSetDevice.argtypes = [i]
        h.hipGetDevice.argtypes = [ctypes.POINTER(i)]
        h.hipDeviceGetName.argtypes = [ctypes.c_char_p, i, i]
        h.hipGetErrorString.argtypes = [i]
        h.hipGetErrorString.restype = ctypes.c_char_p
        h.hipDeviceSynchronize.argtypes = []
        h.hipMemGetInfo.argtypes = [ctypes.POINTER(sz), ctypes.POINTER(sz)]
        h.hipStreamBeginCapture.argtypes = [vp, i]
        h.hipStreamEndCapture.argtypes = [vp, ctypes.POINTER(vp)]
        h.hipGraphInstantiate.argtypes = [ctypes.POINTER(vp), vp, vp, vp, sz]
        h.hipGraphLaunch.argtypes = [vp, vp]
        h.hipGraphExecDestroy.argtypes = [vp]
        h.hipGraphDestroy.argtypes = [vp]
    return _HIP


def error_string(code: int) -> str:
    try:
        return _lib().hipGetErrorString(code).decode()
    except Exception:  # pragma: no cover
        return "?"


def check(code: int, where: str = "hip") -> None:
    if code != 0:
        raise HipError(code, where)


def device_count() -> int:
    n = ctypes.c_int(0)
    try:
        rc = _lib().hipGetDeviceCount(ctypes.byref(n))
    except OSError:
        return 0
    return n.value if rc == 0 else 0


def set_device(dev: int) -> None:
    check(_lib().hipSetDevice(dev), "hipSetDevice")


def get_device() -> int:
    d = ctypes.c_int(0)
    check(_lib().hipGetDevice(ctypes.byref(d)), "hipGetDevice")
    return d.value


def device_name(dev: int) -> str:
    buf = ctypes.create_string_buffer(256)
    check(_lib().hipDeviceGetName(buf, 256, dev), "hipDeviceGetName")
    return buf.value.decode(errors="replace")


def synchronize() -> None:
    check(_lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info() -> tuple[int, int]:
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    check(_lib().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceBuffer:
    """Owning device allocation of `nbytes` bytes."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p(0)
        if self.nbytes > 0:
            check(_lib().hipMalloc(ctypes.byref(self.ptr), self.nbytes), f"hipMalloc({self.nbytes})")

    @property
    def addr(self) -> int:
        return self.ptr.value or 0

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(arr)
        if a.nbytes + offset > self.nbytes:
            raise ValueError("upload larger than buffer")
        check(
            _lib().hipMemcpy(ctypes.c_void_p(self.addr + offset), a.ctypes.data, a.nbytes, hipMemcpyHostToDevice),
            "hipMemcpy H2D",
        )

    def download(self, dtype=np.uint64, count: int | None = None, offset: int = 0) -> np.ndarray:
        itemsize = np.dtype(dtype).itemsize
        if count is None:
            count = (self.nbytes - offset) // itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes + offset > self.nbytes:
            raise ValueError("download larger than buffer")
        check(
            _lib().hipMemcpy(out.ctypes.data, ctypes.c_void_p(self.addr + offset), out.nbytes, hipMemcpyDeviceToHost),
            "hipMemcpy D2H",
        )
        return out

    def zero(self) -> None:
        check(_lib().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self) -> None:
        if self.ptr.value:
            _lib().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedBuffer:
    """Page-locked host allocation (hipHostMalloc) viewed as a numpy array."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p(0)
        check(_lib().hipHostMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1), 0), "hipHostMalloc")

    @property
    def addr(self) -> int:
        return self.ptr.value or 0

    def view(self, dtype=np.uint64) -> np.ndarray:
        n = self.nbytes // np.dtype(dtype).itemsize
        buf = (ctypes.c_char * self.nbytes).from_address(self.addr)
        return np.frombuffer(buf, dtype=dtype, count=n)

    def free(self) -> None:
        if self.ptr.value:
            _lib().hipHostFree(self.ptr)
            self.ptr = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A HIP stream at the default priority (the runtime's track stream)."""

    def __init__(self, device: int = 0):
        self.device = device
        self.handle = ctypes.c_void_p(0)
        set_device(device)
        check(_lib().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def raw(self) -> ctypes.c_void_p:
        return self.handle

    def get_device(self) -> int:
        return self.device

    def sync(self) -> None:
        check(_lib().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def wait(self, event: "Event") -> None:
        """Later work on this stream waits for `event` (under capture: a graph edge)."""
        check(_lib().hipStreamWaitEvent(self.handle, event.handle, 0), "hipStreamWaitEvent")

    def destroy(self) -> None:
        if self.handle.value:
            _lib().hipStreamDestroy(self.handle)
            self.handle = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p(0)
        check(_lib().hipEventCreate(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream: Stream | None = None) -> None:
        check(_lib().hipEventRecord(self.handle, stream.raw() if stream else None), "hipEventRecord")

    def sync(self) -> None:
        check(_lib().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        check(_lib().hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle.value:
                _lib().hipEventDestroy(self.handle)
        except Exception:
            pass


hipStreamCaptureModeThreadLocal = 1


class Graph:
    """A captured and instantiated hipGraph of the work enqueued by `fill(stream)`.

    Kernel launches and asynchronous copies may be captured here (no memsets,
    which have given wrong results on replays, and no allocations).
    """

    def __init__(self, stream: Stream, fill):
        h = _lib()
        self.graph = ctypes.c_void_p(0)
        self.exec = ctypes.c_void_p(0)
        check(h.hipStreamBeginCapture(stream.raw(), hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")
        try:
            fill(stream)
        finally:
            rc = h.hipStreamEndCapture(stream.raw(), ctypes.byref(self.graph))
        check(rc, "hipStreamEndCapture")
        check(h.hipGraphInstantiate(ctypes.byref(self.exec), self.graph, None, None, 0), "hipGraphInstantiate")

    def launch(self, stream: Stream) -> None:
        check(_lib().hipGraphLaunch(self.exec, stream.raw()), "hipGraphLaunch")

    def __del__(self):
        try:
            if self.exec.value:
                _lib().hipGraphExecDestroy(self.exec)
            if self.graph.value:
                _lib().hipGraphDestroy(self.graph)
        except Exception:
            pass
