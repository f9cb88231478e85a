"""Host-side mirror of the reference runtime's plugin/operator interface.

Mirrors (reference paths):
  * RegisteredFunction / Function / FuncMeta / KernelType
    -- runtime/src/functions.rs:15-92
  * ScalarArray with rotate / slice semantics -- runtime/src/scalar.rs:150-445
  * PolyPtr / ConstPolyPtr #[repr(C)] -- core/src/poly_ptr.rs:4-86
  * Libs (shared-library loader rooted at $ZKPOLY_COMPILER_PROJECT_ROOT/lib)
    -- common/src/load_dynamic.rs:10-67
A Function's closure takes (mut_vars, vars, gpu_mapping) exactly like the Rust
closure; violated preconditions raise AssertionError (the Rust `assert!`s) and
failing native calls raise hip.HipError (the Rust `cuda_check!` panics).
"""
from __future__ import annotations

import ctypes
import enum
import os
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import hip

PACKAGE_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def project_root() -> str:
    return os.environ.get("ZKPOLY_COMPILER_PROJECT_ROOT", PACKAGE_ROOT)


class Libs:
    """Loads shared libraries from <project root>/lib, once each."""

    def __init__(self, root: Optional[str] = None):
        self.root = root or project_root()
        self._libs: dict[str, ctypes.CDLL] = {}

    def path_of(self, name: str) -> str:
        return os.path.join(self.root, "lib", name)

    def contains(self, name: str) -> bool:
        return name in self._libs or os.path.exists(self.path_of(name))

    def load_relative(self, name: str) -> ctypes.CDLL:
        if name not in self._libs:
            path = self.path_of(name)
            if not os.path.exists(path):
                raise FileNotFoundError(
                    f"{path} is not built; run __graft_entry__.build() (the HIP libraries have no CPU fallback)"
                )
            self._libs[name] = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        return self._libs[name]


class PolyPtr(ctypes.Structure):
    _fields_ = [
        ("ptr", ctypes.c_void_p),
        ("len", ctypes.c_size_t),
        ("rotate", ctypes.c_size_t),
        ("offset", ctypes.c_size_t),
        ("whole_len", ctypes.c_size_t),
    ]

    @classmethod
    def from_array(cls, a: "ScalarArray") -> "PolyPtr":
        if a.slice_info is None:
            offset, whole = 0, a.len
        else:
            offset, whole = a.slice_info.offset, a.slice_info.whole_len
        return cls(a.values, a.len, a.get_rotation(), offset, whole)


ConstPolyPtr = PolyPtr  # identical layout; constness is a Rust-side distinction


class KernelType(enum.Enum):
    NttPrcompute = "NttPrcompute"
    NttRecompute = "NttRecompute"
    DistributePowers = "DistributePowers"
    Msm = "Msm"
    FusedArith = "FusedArith"
    # libpoly.so operators (core/src/poly.rs); ScalarPow carries its exponent
    # in FuncMeta.extra
    PolyAdd = "PolyAdd"
    PolySub = "PolySub"
    NewZero = "NewZero"
    NewOneLagrange = "NewOneLagrange"
    NewOneCoef = "NewOneCoef"
    EvaluatePoly = "EvaluatePoly"
    KateDivision = "KateDivision"
    ScanMul = "ScanMul"
    BatchedInvert = "BatchedInvert"
    ScalarInvert = "ScalarInvert"
    ScalarPow = "ScalarPow"
    PolyPermute = "PolyPermute"
    Other = "Other"


@dataclass
class FuncMeta:
    name: str
    typ: KernelType
    extra: object = None


@dataclass
class Function:
    meta: FuncMeta
    f: Callable


class RegisteredFunction:
    def get_fn(self) -> Function:  # pragma: no cover - interface
        raise NotImplementedError


@dataclass
class ScalarSlice:
    offset: int
    whole_len: int


class ScalarArray:
    """A (possibly rotated or sliced) array of field elements at `values`.

    `values` is a raw address (device or host).  Logical element i lives at
    values[(offset + i - rotate) mod whole_len] (runtime/src/scalar.rs:432-444).
    """

    def __init__(self, length: int, values: int, device: str = "gpu", device_id: int = 0):
        self.values = int(values)
        self.len = int(length)
        self.rotate_ = 0
        self.device = device
        self.device_id = device_id
        self.slice_info: Optional[ScalarSlice] = None

    def get_rotation(self) -> int:
        return self.rotate_

    def rotate(self, shift: int) -> None:
        assert self.slice_info is None, "a slice cannot be rotated"
        r = (self.rotate_ + shift) % self.len
        self.rotate_ = r

    def slice(self, start: int, end: int) -> "ScalarArray":
        assert start <= end <= self.len
        assert self.slice_info is None, "this is already a slice"
        actual = (self.len + start - self.get_rotation()) % self.len
        s = ScalarArray(end - start, self.values, self.device, self.device_id)
        s.slice_info = ScalarSlice(actual, self.len)
        return s

    def physical_index(self, i: int) -> int:
        assert 0 <= i < self.len
        if self.slice_info is None:
            off, mod = 0, self.len
        else:
            off, mod = self.slice_info.offset, self.slice_info.whole_len
        return (i + mod + off - self.rotate_) % mod


@dataclass
class Scalar:
    """One field element at `value` (device or host address)."""

    value: int
    device: str = "gpu"


@dataclass
class GpuBuffer:
    ptr: int
    size: int
    device_id: int = 0


@dataclass
class PointArray:
    """Affine points (x, y in Fq Montgomery form, 64 bytes each) at `values`."""

    values: int
    len: int
    device: str = "cpu"


class Point:
    """One affine point held in host memory (identity = (0, 0))."""

    def __init__(self):
        self.data = np.zeros(8, dtype=np.uint64)

    def as_ptr(self) -> int:
        return self.data.ctypes.data


def unwrap(v, cls, what: str):
    assert isinstance(v, cls), f"expected {what}, got {type(v).__name__}"
    return v


def unwrap_scalar_array(v) -> ScalarArray:
    return unwrap(v, ScalarArray, "ScalarArray")


def unwrap_stream(v) -> hip.Stream:
    return unwrap(v, hip.Stream, "Stream")


def identity_mapping(x: int) -> int:
    return x
