"""NTT operators over libntt.so, mirroring core/src/ntt.rs.

SsipNtt (core/src/ntt.rs:165-200), SsipPrecompute (:204-238),
DistributePowers (:112-150), RecomputeNtt (:249-289) and GenPqOmegas
(:292-371) with the same argument lists and checks:
  SsipNtt:          mut_vars=[poly], vars=[twiddle (len n/2), stream]
  RecomputeNtt:     mut_vars=[poly], vars=[pq (len 2^(pq_deg-1)), omegas (len 32), stream]
  DistributePowers: mut_vars=[poly], vars=[powers, stream]
  SsipPrecompute:   f(twiddle_host_array, unit) fills [1, w, ..., w^(len-1)]
  GenPqOmegas:      f(pq_host, omegas_host, len, unit); get_pq_len(log_len)
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import hip
from .runtime import (
    FuncMeta,
    Function,
    KernelType,
    Libs,
    PolyPtr,
    RegisteredFunction,
    ScalarArray,
    unwrap_scalar_array,
    unwrap_stream,
)

LIB_NAME = "libntt.so"

# (max_threads_stage1_log, max_threads_stage2_log) for 32-byte fields
# (core/src/ntt.rs:373-378); forwarded for ABI compatibility.
STAGE_THREADS = (8, 8)


def load_lib(libs: Libs) -> ctypes.CDLL:
    lib = libs.load_relative(LIB_NAME)
    if not getattr(lib, "_zkp_typed", False):
        u32p = ctypes.c_void_p
        lib.ssip_ntt.argtypes = [PolyPtr, u32p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint]
        lib.ssip_ntt.restype = ctypes.c_int
        lib.ssip_precompute.argtypes = [u32p, ctypes.c_uint, u32p]
        lib.ssip_precompute.restype = ctypes.c_int
        lib.distribute_powers.argtypes = [PolyPtr, u32p, ctypes.c_ulonglong, ctypes.c_void_p]
        lib.distribute_powers.restype = ctypes.c_int
        lib.zkp_gen_powers.argtypes = [u32p, ctypes.c_ulonglong, u32p, ctypes.c_void_p]
        lib.zkp_gen_powers.restype = ctypes.c_int
        lib.recompute_ntt.argtypes = [PolyPtr, u32p, ctypes.c_uint, u32p, ctypes.c_uint, ctypes.c_void_p,
                                      ctypes.c_uint, ctypes.c_uint]
        lib.recompute_ntt.restype = ctypes.c_int
        lib.gen_pq_omegas.argtypes = [u32p, u32p, ctypes.c_uint, ctypes.c_uint, u32p]
        lib.gen_pq_omegas.restype = None
        lib.zkp_ntt_set_variant.argtypes = [ctypes.c_int]
        lib.zkp_ntt_set_variant.restype = ctypes.c_int
        lib.zkp_ntt_set_tuning.argtypes = [ctypes.c_uint, ctypes.c_uint]
        lib.zkp_ntt_set_tuning.restype = ctypes.c_int
        lib.zkp_ntt_kernel_count.argtypes = [ctypes.c_uint]
        lib.zkp_ntt_kernel_count.restype = ctypes.c_int
        lib._zkp_typed = True
    return lib


class SsipNtt(RegisteredFunction):
    def __init__(self, libs: Libs):
        self.c_func = load_lib(libs).ssip_ntt

    def get_fn(self) -> Function:
        c_func = self.c_func

        def f(mut_var, var, gpu_mapping=None):
            assert len(mut_var) == 1
            assert len(var) == 2
            x = unwrap_scalar_array(mut_var[0])
            assert x.len > 0 and (x.len & (x.len - 1)) == 0, "len must be a power of two"
            log_len = x.len.bit_length() - 1
            twiddle = unwrap_scalar_array(var[0])
            assert twiddle.len * 2 == x.len
            stream = unwrap_stream(var[1])
            hip.set_device(stream.get_device())
            hip.check(
                c_func(PolyPtr.from_array(x), twiddle.values, log_len, stream.raw(), *STAGE_THREADS),
                "ssip_ntt",
            )

        return Function(FuncMeta("ssip_ntt", KernelType.NttPrcompute), f)


class RecomputeNtt(RegisteredFunction):
    """The NTT without the n/2 twiddle table (core/src/ntt.rs:249-289)."""

    def __init__(self, libs: Libs):
        self.c_func = load_lib(libs).recompute_ntt

    def get_fn(self) -> Function:
        c_func = self.c_func

        def f(mut_var, var, gpu_mapping=None):
            assert len(mut_var) == 1
            assert len(var) == 3
            x = unwrap_scalar_array(mut_var[0])
            assert x.len > 0 and (x.len & (x.len - 1)) == 0, "len must be a power of two"
            log_len = x.len.bit_length() - 1
            pq = unwrap_scalar_array(var[0])
            assert pq.len > 0 and (pq.len & (pq.len - 1)) == 0, "pq len must be a power of two"
            pq_deg = pq.len.bit_length()  # trailing_zeros + 1
            omegas = unwrap_scalar_array(var[1])
            assert omegas.len == 32
            stream = unwrap_stream(var[2])
            hip.set_device(stream.get_device())
            hip.check(
                c_func(PolyPtr.from_array(x), pq.values, pq_deg, omegas.values, log_len, stream.raw(),
                       *STAGE_THREADS),
                "recompute_ntt",
            )

        return Function(FuncMeta("recompute_ntt", KernelType.NttRecompute), f)


def _get_deg(deg_stage: int, max_deg_stage: int) -> int:
    """Stage-degree split of the reference's planner (core/src/ntt.rs:305-320)."""
    rounds = 1
    while True:
        deg = deg_stage if rounds == 1 else (deg_stage - 1) // rounds + 1
        if deg <= max_deg_stage:
            return deg
        rounds += 1


class GenPqOmegas:
    """Compile-time pq / omega-ladder generation for RecomputeNtt (core/src/ntt.rs:292-371)."""

    def __init__(self, libs: Libs):
        self.c_func = load_lib(libs).gen_pq_omegas

    @staticmethod
    def get_pq_len(log_len: int) -> int:
        """core/src/ntt.rs:322-335: 2^(max(deg1, deg2) - 1) for the (8, 8) stage threads."""
        s1, s2 = STAGE_THREADS
        deg1 = _get_deg((log_len + 1) // 2, s1 + 1)
        deg2 = _get_deg(log_len // 2, (s2 + 2) // 2)
        return 1 << (max(deg1, deg2) - 1)

    def get_fn(self):
        c_func = self.c_func

        def f(pq: ScalarArray, omegas: ScalarArray, length: int, unit: np.ndarray):
            assert omegas.len == 32
            pq_deg = pq.len.bit_length()
            u = np.ascontiguousarray(unit, dtype=np.uint64)
            assert u.size == 4
            c_func(pq.values, omegas.values, pq_deg, int(length), u.ctypes.data)

        return f


class DistributePowers(RegisteredFunction):
    def __init__(self, libs: Libs):
        self.c_func = load_lib(libs).distribute_powers

    def get_fn(self) -> Function:
        c_func = self.c_func

        def f(mut_var, var, gpu_mapping=None):
            assert len(mut_var) == 1
            assert len(var) == 2
            x = unwrap_scalar_array(mut_var[0])
            powers = unwrap_scalar_array(var[0])
            stream = unwrap_stream(var[1])
            hip.set_device(stream.get_device())
            hip.check(c_func(PolyPtr.from_array(x), powers.values, powers.len, stream.raw()), "distribute_powers")

        return Function(FuncMeta("distribute_powers", KernelType.DistributePowers), f)


class SsipPrecompute:
    """Compile-time twiddle generation (not a RegisteredFunction in the reference either)."""

    def __init__(self, libs: Libs):
        self.c_func = load_lib(libs).ssip_precompute

    def get_fn(self):
        c_func = self.c_func

        def f(twiddle: ScalarArray, unit: np.ndarray):
            assert twiddle.len > 0 and (twiddle.len & (twiddle.len - 1)) == 0
            u = np.ascontiguousarray(unit, dtype=np.uint64)
            assert u.size == 4
            hip.check(c_func(twiddle.values, twiddle.len, u.ctypes.data), "ssip_precompute")

        return f


def gen_powers(libs: Libs, out_dev: int, length: int, base_dev: int, stream: hip.Stream) -> None:
    """[1, b, b^2, ...] written to device memory (no reference counterpart)."""
    hip.check(load_lib(libs).zkp_gen_powers(out_dev, length, base_dev, stream.raw()), "zkp_gen_powers")


VARIANT_FUSED_RADIX4 = 0
VARIANT_RADIX2 = 1
VARIANT_HEAD_PER_PAIR = 2  # as VARIANT_FUSED_RADIX4 with one tile pair per head workgroup


def set_variant(libs: Libs, variant: int) -> int:
    """Selects the transform kernels for later calls; returns the previous variant."""
    old = load_lib(libs).zkp_ntt_set_variant(variant)
    if old < 0:
        raise ValueError(f"unknown NTT variant {variant}")
    return old


def set_tuning(libs: Libs, log_n: int, min_col_log: int) -> None:
    """Passes of 2^log_n transforms read runs of >= 2^min_col_log adjacent elements."""
    if load_lib(libs).zkp_ntt_set_tuning(log_n, min_col_log) != 0:
        raise ValueError(f"tuning out of range: log_n={log_n} min_col_log={min_col_log}")


def kernel_count(libs: Libs, log_n: int) -> int:
    return int(load_lib(libs).zkp_ntt_kernel_count(log_n))
