"""CpuMemoryPool over libzkp_pinned_pool.so, mirroring memory_pool/src/cpu_pool.rs.

CpuMemoryPool(max_log_factor, base_size) hands out host blocks of
base_size * 2^k bytes carved from slabs of base_size * 2^max_log_factor bytes
of pinned host memory (hipHostMalloc); `use_mmap()` switches the backing to
file-backed mappings, as the reference's use_mmap does.  Allocation failures
and bad frees raise hip.HipError where the reference panics.
`alloc_cpu(pool, n)` gives a host ScalarArray of n field elements from the
pool (ScalarArray::alloc_cpu, runtime/src/scalar.rs).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import hip
from .runtime import Libs, ScalarArray

LIB = "libzkp_pinned_pool.so"
PINNED, MMAP = 0, 1


def _lib(libs: Libs) -> ctypes.CDLL:
    lib = libs.load_relative(LIB)
    if not getattr(lib, "_zkp_typed", False):
        P, U, ULL = ctypes.c_void_p, ctypes.c_uint, ctypes.c_ulonglong
        lib.zkp_pinned_pool_create.argtypes = [U, ULL, U, ctypes.c_char_p, ctypes.POINTER(P)]
        lib.zkp_pinned_pool_alloc.argtypes = [P, ULL, ctypes.POINTER(P)]
        lib.zkp_pinned_pool_alloc_log.argtypes = [P, U, ctypes.POINTER(P)]
        lib.zkp_pinned_pool_free.argtypes = [P, P]
        lib.zkp_pinned_pool_preallocate.argtypes = [P, ULL]
        lib.zkp_pinned_pool_shrink.argtypes = [P]
        lib.zkp_pinned_pool_clear.argtypes = [P]
        for name in ("create", "alloc", "alloc_log", "free", "preallocate", "shrink", "clear"):
            getattr(lib, "zkp_pinned_pool_" + name).restype = ctypes.c_int
        lib.zkp_pinned_pool_block_size.argtypes = [P, P]
        lib.zkp_pinned_pool_used_bytes.argtypes = [P]
        lib.zkp_pinned_pool_free_blocks.argtypes = [P, U]
        for name in ("block_size", "used_bytes", "free_blocks"):
            getattr(lib, "zkp_pinned_pool_" + name).restype = ULL
        lib.zkp_pinned_pool_destroy.argtypes = [P]
        lib.zkp_pinned_pool_destroy.restype = None
        lib._zkp_typed = True
    return lib


class CpuMemoryPool:
    def __init__(self, max_log_factor: int, base_size: int, libs: Optional[Libs] = None,
                 backing: int = PINNED, tmp_dir: Optional[str] = None):
        self.lib = _lib(libs or Libs())
        self.max_log_factor = max_log_factor
        self.base_size = base_size
        self.handle = ctypes.c_void_p(0)
        hip.check(self.lib.zkp_pinned_pool_create(max_log_factor, base_size, backing,
                                                   tmp_dir.encode() if tmp_dir else None, ctypes.byref(self.handle)),
                  "zkp_pinned_pool_create")

    def use_mmap(self, tmp_dir: Optional[str] = None) -> "CpuMemoryPool":
        """The reference's use_mmap: a new, empty pool with file-backed slabs."""
        other = CpuMemoryPool(self.max_log_factor, self.base_size, backing=MMAP, tmp_dir=tmp_dir)
        other.lib = self.lib
        self.destroy()
        return other

    def allocate(self, nbytes: int) -> int:
        out = ctypes.c_void_p(0)
        hip.check(self.lib.zkp_pinned_pool_alloc(self.handle, nbytes, ctypes.byref(out)), "zkp_pinned_pool_alloc")
        return out.value

    def allocate_log(self, log_factor: int) -> int:
        out = ctypes.c_void_p(0)
        hip.check(self.lib.zkp_pinned_pool_alloc_log(self.handle, log_factor, ctypes.byref(out)),
                  "zkp_pinned_pool_alloc_log")
        return out.value

    def free(self, ptr: int) -> None:
        hip.check(self.lib.zkp_pinned_pool_free(self.handle, ptr), "zkp_pinned_pool_free")

    def block_size(self, ptr: int) -> int:
        return int(self.lib.zkp_pinned_pool_block_size(self.handle, ptr))

    def preallocate(self, num_slabs: int) -> None:
        hip.check(self.lib.zkp_pinned_pool_preallocate(self.handle, num_slabs), "zkp_pinned_pool_preallocate")

    def shrink(self) -> None:
        hip.check(self.lib.zkp_pinned_pool_shrink(self.handle), "zkp_pinned_pool_shrink")

    def clear(self) -> None:
        hip.check(self.lib.zkp_pinned_pool_clear(self.handle), "zkp_pinned_pool_clear")

    def used_bytes(self) -> int:
        return int(self.lib.zkp_pinned_pool_used_bytes(self.handle))

    def free_blocks(self, k: int) -> int:
        return int(self.lib.zkp_pinned_pool_free_blocks(self.handle, k))

    def view(self, ptr: int, nbytes: int, dtype=np.uint64) -> np.ndarray:
        """A numpy view of nbytes at ptr (a block of this pool)."""
        buf = (ctypes.c_char * nbytes).from_address(ptr)
        return np.frombuffer(buf, dtype=dtype, count=nbytes // np.dtype(dtype).itemsize)

    def destroy(self) -> None:
        if self.handle.value:
            self.lib.zkp_pinned_pool_destroy(self.handle)
            self.handle = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def alloc_cpu(pool: CpuMemoryPool, length: int) -> ScalarArray:
    """ScalarArray::alloc_cpu: `length` field elements (32 bytes each) of host memory from the pool."""
    return ScalarArray(length, pool.allocate(length * 32), "cpu")
