"""Polynomial operators over libpoly.so, mirroring core/src/poly.rs.

Same operator names, argument lists and checks as the reference's wrappers
(core/src/poly.rs:37-183 structs, :185-212 constructors, :221-783 closures):
  PolyAdd / PolySub / PolyMul: mut_var=[res], var=[a, b, stream]
  PolyZero / PolyOneLagrange / PolyOneCoef: mut_var=[target], var=[stream]
  PolyEval:     mut_var=[temp_buf, res (Scalar)], var=[poly, x (Scalar), stream];
                get_buffer_size(len)
  KateDivision: mut_var=[temp_buf, res], var=[p, b (Scalar), stream];
                |res| = |p| = 2^k; get_buffer_size(log_p)
  PolyScan:     mut_var=[temp_buf, target], var=[p, x0 (Scalar), stream];
                target = x0, x0 p_0, x0 p_0 p_1, ...; get_buffer_size(len)
  PolyInvert:   mut_var=[temp_buf, target], var=[stream]; get_buffer_size(len)
  ScalarInv / ScalarPow(exp): mut_var=[scalar], var=[stream]
  PolyPermute:  mut_var=[res_input, res_table, temp_buf], var=[input, table, stream];
                usable = |res_input| = |res_table|; get_buffer_size(usable)
Temporary buffers are runtime.GpuBuffer objects sized by get_buffer_size
(the library's own size query: a NULL buffer returns the size, runs nothing);
a smaller buffer is refused before anything is launched.

The reference's closures also accept host-resident arrays and scalars (var
empty) and compute on the CPU; this package computes on the GPU only, so such
calls raise NotImplementedError instead of falling back to host arithmetic.
"""
from __future__ import annotations

import ctypes

from . import hip
from .runtime import (
    ConstPolyPtr,
    FuncMeta,
    Function,
    GpuBuffer,
    KernelType,
    Libs,
    PolyPtr,
    RegisteredFunction,
    Scalar,
    ScalarArray,
    unwrap,
    unwrap_scalar_array,
    unwrap_stream,
)

LIB_NAME = "libpoly.so"


def load_lib(libs: Libs) -> ctypes.CDLL:
    lib = libs.load_relative(LIB_NAME)
    if not getattr(lib, "_zkp_typed", False):
        vp, ulp, st = ctypes.c_void_p, ctypes.POINTER(ctypes.c_ulong), ctypes.c_void_p
        for name in ("poly_add", "poly_sub", "poly_mul"):
            getattr(lib, name).argtypes = [PolyPtr, ConstPolyPtr, ConstPolyPtr, st]
        for name in ("poly_one_lagrange", "poly_one_coef", "poly_zero"):
            getattr(lib, name).argtypes = [PolyPtr, st]
        lib.poly_eval.argtypes = [vp, ulp, ConstPolyPtr, vp, vp, st]
        lib.kate_division.argtypes = [vp, ulp, ctypes.c_uint, ConstPolyPtr, vp, PolyPtr, st]
        lib.scan_mul.argtypes = [vp, ulp, PolyPtr, st]
        lib.batched_invert.argtypes = [vp, ulp, PolyPtr, st]
        lib.inv_scalar.argtypes = [vp, st]
        lib.scalar_pow.argtypes = [vp, ctypes.c_ulonglong, st]
        lib.permute.argtypes = [vp, ulp, ctypes.c_ulong, ConstPolyPtr, ConstPolyPtr, PolyPtr, PolyPtr, st]
        for name in ("poly_add", "poly_sub", "poly_mul", "poly_one_lagrange", "poly_one_coef", "poly_zero",
                     "poly_eval", "kate_division", "scan_mul", "batched_invert", "inv_scalar", "scalar_pow", "permute"):
            getattr(lib, name).restype = ctypes.c_int
        lib._zkp_typed = True
    return lib


def null_ptr(length: int) -> PolyPtr:
    """PolyPtr::null(len) of core/src/poly_ptr.rs: a view with no memory, for size queries."""
    return PolyPtr(None, length, 0, 0, length)


def _query(c_func, *args) -> int:
    size = ctypes.c_ulong(0)
    hip.check(c_func(None, ctypes.byref(size), *args), c_func.__name__ + " (size query)")
    return int(size.value)


def _gpu_array(v, what: str) -> ScalarArray:
    a = unwrap_scalar_array(v)
    if a.device != "gpu":
        raise NotImplementedError(f"{what}: host-resident arrays are not computed here (GPU only)")
    return a


def _gpu_scalar(v, what: str) -> Scalar:
    s = unwrap(v, Scalar, "Scalar")
    if s.device != "gpu":
        raise NotImplementedError(f"{what}: host-resident scalars are not computed here (GPU only)")
    return s


def sub_view(a: ScalarArray, start: int, end: int) -> PolyPtr:
    """Elements [start, end) of a (rotated, sliced or plain) array as a view.
    The reference takes ScalarArray::slice here, which refuses arrays that
    are already slices (runtime/src/scalar.rs:346-365); the view is the same
    where both apply."""
    assert 0 <= start <= end <= a.len
    whole = a.slice_info.whole_len if a.slice_info is not None else a.len
    off = a.physical_index(start) if start < a.len else 0
    return PolyPtr(a.values, end - start, 0, off, whole)


def scalar_view(s: Scalar) -> ConstPolyPtr:
    """A one-element view of a device scalar."""
    return ConstPolyPtr(s.value, 1, 0, 0, 1)


class _PolyFunc(RegisteredFunction):
    SYMBOL = ""

    def __init__(self, libs: Libs):
        self.lib = load_lib(libs)
        self.c_func = getattr(self.lib, self.SYMBOL)


class _Binary(_PolyFunc):
    """res = a op b (core/src/poly.rs:682-783); unequal lengths follow poly_basic.cu."""

    TYPE = KernelType.Other

    def get_fn(self) -> Function:
        c_func, sym = self.c_func, self.SYMBOL

        def f(mut_var, var, gpu_mapping=None):
            assert len(mut_var) == 1
            assert len(var) == 3
            res = _gpu_array(mut_var[0], sym)
            a = _gpu_array(var[0], sym)
            b = _gpu_array(var[1], sym)
            stream = unwrap_stream(var[2])
            hip.set_device(stream.get_device())
            hip.check(c_func(PolyPtr.from_array(res), PolyPtr.from_array(a), PolyPtr.from_array(b), stream.raw()), sym)

        return Function(FuncMeta(sym, self.TYPE), f)


class PolyAdd(_Binary):
    SYMBOL, TYPE = "poly_add", KernelType.PolyAdd


class PolySub(_Binary):
    SYMBOL, TYPE = "poly_sub", KernelType.PolySub


class PolyMul(_Binary):
    SYMBOL, TYPE = "poly_mul", KernelType.Other  # the reference tags it Other ("not used in practice")


class _Fill(_PolyFunc):
    """target filled in place (core/src/poly.rs:468-566)."""

    TYPE = KernelType.Other

    def get_fn(self) -> Function:
        c_func, sym = self.c_func, self.SYMBOL

        def f(mut_var, var, gpu_mapping=None):
            assert len(mut_var) == 1
            target = _gpu_array(mut_var[0], sym)
            assert len(var) == 1
            stream = unwrap_stream(var[0])
            hip.set_device(stream.get_device())
            hip.check(c_func(PolyPtr.from_array(target), stream.raw()), sym)

        return Function(FuncMeta(sym, self.TYPE), f)


class PolyZero(_Fill):
    SYMBOL, TYPE = "poly_zero", KernelType.NewZero


class PolyOneLagrange(_Fill):
    SYMBOL, TYPE = "poly_one_lagrange", KernelType.NewOneLagrange


class PolyOneCoef(_Fill):
    SYMBOL, TYPE = "poly_one_coef", KernelType.NewOneCoef


class PolyEval(_PolyFunc):
    """res = sum_i poly_i x^i (core/src/poly.rs:631-681)."""

    SYMBOL = "poly_eval"

    def get_buffer_size(self, length: int) -> int:
        return _query(self.c_func, null_ptr(length), None, None, None)

    def get_fn(self) -> Function:
        c_func = self.c_func

        def f(mut_var, var, gpu_mapping=None):
            assert len(mut_var) == 2
            assert len(var) == 3
            temp = unwrap(mut_var[0], GpuBuffer, "GpuBuffer")
            res = _gpu_scalar(mut_var[1], "poly_eval")
            poly = _gpu_array(var[0], "poly_eval")
            assert temp.size >= self.get_buffer_size(poly.len), "temporary buffer too small"
            x = _gpu_scalar(var[1], "poly_eval")
            stream = unwrap_stream(var[2])
            hip.set_device(stream.get_device())
            hip.check(c_func(temp.ptr, None, PolyPtr.from_array(poly), res.value, x.value, stream.raw()), "poly_eval")

        return Function(FuncMeta("poly_eval", KernelType.EvaluatePoly), f)


class KateDivision(_PolyFunc):
    """res = (p(X) - p(b)) / (X - b), res_{n-1} = 0 (core/src/poly.rs:570-628)."""

    SYMBOL = "kate_division"

    def get_buffer_size(self, log_p: int) -> int:
        return _query(self.c_func, log_p, null_ptr(1 << log_p), None, null_ptr(1 << log_p), None)

    def get_fn(self) -> Function:
        c_func = self.c_func

        def f(mut_var, var, gpu_mapping=None):
            assert len(mut_var) == 2
            assert len(var) == 3
            temp = unwrap(mut_var[0], GpuBuffer, "GpuBuffer")
            res = _gpu_array(mut_var[1], "kate_division")
            p = _gpu_array(var[0], "kate_division")
            assert res.len == p.len
            assert p.len > 0 and (p.len & (p.len - 1)) == 0, "len must be a power of two"
            log_p = p.len.bit_length() - 1
            assert temp.size >= self.get_buffer_size(log_p), "temporary buffer too small"
            b = _gpu_scalar(var[1], "kate_division")
            stream = unwrap_stream(var[2])
            hip.set_device(stream.get_device())
            hip.check(c_func(temp.ptr, None, log_p, PolyPtr.from_array(p), b.value, PolyPtr.from_array(res),
                             stream.raw()), "kate_division")

        return Function(FuncMeta("kate_division", KernelType.KateDivision), f)


class PolyScan(_PolyFunc):
    """target_0 = x0, target_i = x0 p_0 ... p_{i-1} (core/src/poly.rs:413-466).

    As in the reference, x0 and p[0..len-1) are first copied into target
    (view-aware copies: poly_add against an empty operand), then scan_mul runs
    in place on target."""

    SYMBOL = "scan_mul"

    def get_buffer_size(self, length: int) -> int:
        return _query(self.c_func, null_ptr(length), None)

    def get_fn(self) -> Function:
        c_func, copy = self.c_func, self.lib.poly_add

        def f(mut_var, var, gpu_mapping=None):
            assert len(mut_var) == 2
            assert len(var) == 3
            temp = unwrap(mut_var[0], GpuBuffer, "GpuBuffer")
            target = _gpu_array(mut_var[1], "scan_mul")
            p = _gpu_array(var[0], "scan_mul")
            assert target.len == p.len
            assert temp.size >= self.get_buffer_size(target.len), "temporary buffer too small"
            x0 = _gpu_scalar(var[1], "scan_mul")
            stream = unwrap_stream(var[2])
            hip.set_device(stream.get_device())
            if target.len == 0:
                return
            empty = null_ptr(0)
            hip.check(copy(sub_view(target, 0, 1), scalar_view(x0), empty, stream.raw()), "scan_mul (x0 copy)")
            if target.len > 1:
                hip.check(copy(sub_view(target, 1, target.len), sub_view(p, 0, p.len - 1), empty, stream.raw()),
                          "scan_mul (shift copy)")
            hip.check(c_func(temp.ptr, None, PolyPtr.from_array(target), stream.raw()), "scan_mul")

        return Function(FuncMeta("scan_mul", KernelType.ScanMul), f)


class PolyInvert(_PolyFunc):
    """target_i = target_i^-1, zeros kept (core/src/poly.rs:364-410)."""

    SYMBOL = "batched_invert"

    def get_buffer_size(self, length: int) -> int:
        return _query(self.c_func, null_ptr(length), None)

    def get_fn(self) -> Function:
        c_func = self.c_func

        def f(mut_var, var, gpu_mapping=None):
            assert len(mut_var) == 2
            assert len(var) == 1
            temp = unwrap(mut_var[0], GpuBuffer, "GpuBuffer")
            target = _gpu_array(mut_var[1], "batched_invert")
            assert temp.size >= self.get_buffer_size(target.len), "temporary buffer too small"
            stream = unwrap_stream(var[0])
            hip.set_device(stream.get_device())
            hip.check(c_func(temp.ptr, None, PolyPtr.from_array(target), stream.raw()), "batched_invert")

        return Function(FuncMeta("batched_invert", KernelType.BatchedInvert), f)


class ScalarInv(_PolyFunc):
    """scalar = scalar^-1 on the device (core/src/poly.rs:331-361)."""

    SYMBOL = "inv_scalar"

    def get_fn(self) -> Function:
        c_func = self.c_func

        def f(mut_var, var, gpu_mapping=None):
            assert len(mut_var) == 1
            s = unwrap(mut_var[0], Scalar, "Scalar")
            if len(var) == 0:
                raise NotImplementedError("scalar_inv: host-resident scalars are not computed here (GPU only)")
            assert len(var) == 1, "var len should be 1 or 0 for ScalarInv"
            stream = unwrap_stream(var[0])
            hip.set_device(stream.get_device())
            hip.check(c_func(s.value, stream.raw()), "inv_scalar")

        return Function(FuncMeta("scalar_inv", KernelType.ScalarInvert), f)


class ScalarPow(_PolyFunc):
    """scalar = scalar^exp on the device (core/src/poly.rs:283-329)."""

    SYMBOL = "scalar_pow"

    def __init__(self, libs: Libs, exp: int):
        super().__init__(libs)
        assert 0 <= exp < (1 << 64)
        self.exp = int(exp)

    def get_fn(self) -> Function:
        c_func, exp = self.c_func, self.exp

        def f(mut_var, var, gpu_mapping=None):
            assert len(mut_var) == 1
            s = unwrap(mut_var[0], Scalar, "Scalar")
            if len(var) == 0:
                raise NotImplementedError("scalar_pow: host-resident scalars are not computed here (GPU only)")
            assert len(var) == 1, "var len should be 1 for ScalarPow on the GPU"
            stream = unwrap_stream(var[0])
            hip.set_device(stream.get_device())
            hip.check(c_func(s.value, exp, stream.raw()), "scalar_pow")

        return Function(FuncMeta("scalar_pow", KernelType.ScalarPow, exp), f)


class PolyPermute(_PolyFunc):
    """Lookup permutation (A', S') of the first `usable` rows of (input, table)
    (core/src/poly.rs:221-280): A' sorted; S' holds each run start's value and
    the table's leftovers, ascending, in the other rows."""

    SYMBOL = "permute"

    def get_buffer_size(self, usable_len: int) -> int:
        n = null_ptr(usable_len)
        return _query(self.c_func, usable_len, n, n, n, n, None)

    def get_fn(self) -> Function:
        c_func = self.c_func

        def f(mut_var, var, gpu_mapping=None):
            assert len(mut_var) == 3
            assert len(var) == 3
            res_input = _gpu_array(mut_var[0], "permute")
            res_table = _gpu_array(mut_var[1], "permute")
            temp = unwrap(mut_var[2], GpuBuffer, "GpuBuffer")
            inp = _gpu_array(var[0], "permute")
            table = _gpu_array(var[1], "permute")
            assert temp.size >= self.get_buffer_size(res_input.len), "temporary buffer too small"
            stream = unwrap_stream(var[2])
            hip.set_device(stream.get_device())
            hip.check(c_func(temp.ptr, None, res_input.len, PolyPtr.from_array(inp), PolyPtr.from_array(table),
                             PolyPtr.from_array(res_input), PolyPtr.from_array(res_table), stream.raw()), "permute")

        return Function(FuncMeta("poly_permute", KernelType.PolyPermute), f)


__all__ = ["PolyAdd", "PolySub", "PolyMul", "PolyZero", "PolyOneLagrange", "PolyOneCoef", "PolyEval",
           "KateDivision", "PolyScan", "PolyInvert", "ScalarInv", "ScalarPow", "PolyPermute", "load_lib", "null_ptr", "sub_view"]
