"""MSM operators over libmsm_bn254_<w>_<t>_<dbg>.so, mirroring core/src/msm.rs.

MsmConfig mirrors common/src/msm_config.rs:3-56; MSM.get_fn mirrors the
closure of core/src/msm.rs:302-369:
  vars     = [PointArray x n_precompute (host), ScalarArray x batches (host)]
  mut_vars = [GpuBuffer x cards, Point x batches]
MSM.get_buffer_size(len) is the sizing call (core/src/msm.rs:275-299).
MSMPrecompute.get_fn() fills the extra point tables (core/src/msm.rs:138-162).
Libraries missing for a configuration are built on demand, as the reference
runs make on Makefile.msm (core/src/msm.rs:121-130).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import hip
from .runtime import (
    FuncMeta,
    Function,
    GpuBuffer,
    KernelType,
    Libs,
    Point,
    PointArray,
    RegisteredFunction,
    ScalarArray,
    unwrap,
)

CORE_LIB = "libzkp_msm.so"
SCALAR_BITS = 254


@dataclass
class MsmConfig:
    window_size: int = 16
    target_window: int = 16
    cards: List[int] = field(default_factory=lambda: [0])
    debug: bool = False
    batch_per_run: int = 1
    parts: int = 1
    stage_scalars: int = 1
    stage_points: int = 1
    bits: int = SCALAR_BITS

    def get_precompute(self) -> int:
        actual = -(-self.bits // self.window_size)
        n_windows = actual if actual < self.target_window else self.target_window
        return -(-actual // n_windows)


def get_msm_lib_name(config: MsmConfig) -> str:
    return f"libmsm_bn254_{config.window_size}_{config.target_window}_{'true' if config.debug else 'false'}.so"


def ensure_msm_lib(libs: Libs, config: MsmConfig) -> ctypes.CDLL:
    name = get_msm_lib_name(config)
    if not libs.contains(name):
        from . import build

        build.build_msm_abi(config.window_size, config.target_window, config.debug)
    lib = libs.load_relative(name)
    if not getattr(lib, "_zkp_typed", False):
        P = ctypes.c_void_p
        lib.msm.argtypes = [P, P, ctypes.c_ulonglong, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                            ctypes.c_uint, P, P, ctypes.c_uint, P, P]
        lib.msm.restype = ctypes.c_int
        lib.msm_precompute.argtypes = [ctypes.c_ulonglong, P, ctypes.c_uint]
        lib.msm_precompute.restype = ctypes.c_int
        lib._zkp_typed = True
    return lib


def core_lib(libs: Libs) -> ctypes.CDLL:
    lib = libs.load_relative(CORE_LIB)
    if not getattr(lib, "_zkp_typed", False):
        P = ctypes.c_void_p
        lib.zkp_msm_workspace_size.argtypes = [ctypes.c_ulonglong, ctypes.c_uint, ctypes.c_uint]
        lib.zkp_msm_workspace_size.restype = ctypes.c_ulonglong
        lib.zkp_msm_device.argtypes = [P, P, ctypes.c_ulonglong, ctypes.c_uint, ctypes.c_uint, P, ctypes.c_ulonglong,
                                       P, P, ctypes.c_int]
        lib.zkp_msm_device.restype = ctypes.c_int
        lib.zkp_g1_mul_generator.argtypes = [P, ctypes.c_ulonglong, P, P]
        lib.zkp_g1_mul_generator.restype = ctypes.c_int
        lib.zkp_msm_config.argtypes = [ctypes.c_uint, ctypes.c_uint, P]
        lib.zkp_msm_config.restype = ctypes.c_int
        lib.zkp_msm_device_enqueue.argtypes = [P, P, ctypes.c_ulonglong, ctypes.c_uint, ctypes.c_uint, P,
                                               ctypes.c_ulonglong, P, P]
        lib.zkp_msm_device_enqueue.restype = ctypes.c_int
        lib.zkp_msm_combine.argtypes = [ctypes.c_uint, ctypes.c_uint, P, P]
        lib.zkp_msm_combine.restype = ctypes.c_int
        lib._zkp_typed = True
    return lib


def _ptr_array(addrs) -> ctypes.Array:
    return (ctypes.c_void_p * max(1, len(addrs)))(*addrs)


class MSM(RegisteredFunction):
    def __init__(self, libs: Libs, config: MsmConfig):
        self.config = config
        self.c_func = ensure_msm_lib(libs, config).msm

    def get_buffer_size(self, length: int) -> list[int]:
        cfg = self.config
        sizes = (ctypes.c_ulong * len(cfg.cards))()
        cards = (ctypes.c_uint * len(cfg.cards))(*cfg.cards)
        hip.check(
            self.c_func(None, sizes, length, cfg.batch_per_run, cfg.parts, cfg.stage_scalars, cfg.stage_points,
                        len(cfg.cards), cards, None, 0, None, None),
            "msm (sizing)",
        )
        return list(sizes)

    def get_fn(self) -> Function:
        c_func = self.c_func
        cfg = self.config
        n_pre = cfg.get_precompute()

        def f(mut_var, var, gpu_mapping=lambda i: i):
            batches = len(var) - n_pre
            assert len(mut_var) == batches + len(cfg.cards)
            pts = [unwrap(var[i], PointArray, "PointArray") for i in range(n_pre)]
            length = pts[0].len
            scalars = []
            for i in range(n_pre, len(var)):
                a = unwrap(var[i], ScalarArray, "ScalarArray")
                assert a.get_rotation() == 0, "currently, we don't support rotate in msm"
                scalars.append(a.values)
            buffers = [unwrap(v, GpuBuffer, "GpuBuffer").ptr for v in mut_var[: len(cfg.cards)]]
            results = [unwrap(v, Point, "Point").as_ptr() for v in mut_var[len(cfg.cards):]]
            cards = [gpu_mapping(c) for c in cfg.cards]
            hip.check(
                c_func(_ptr_array(buffers), None, length, cfg.batch_per_run, cfg.parts, cfg.stage_scalars,
                       cfg.stage_points, len(cards), (ctypes.c_uint * len(cards))(*cards),
                       _ptr_array([p.values for p in pts]), batches, _ptr_array(scalars), _ptr_array(results)),
                "msm",
            )

        return Function(FuncMeta("msm", KernelType.Msm, cfg), f)


class MSMPrecompute:
    def __init__(self, libs: Libs, config: MsmConfig):
        self.c_func = ensure_msm_lib(libs, config).msm_precompute

    def get_fn(self):
        c_func = self.c_func

        def f(targets: list, length: int, max_cards: int):
            addrs = [unwrap(t, PointArray, "PointArray").values for t in targets]
            hip.check(c_func(length, _ptr_array(addrs), max_cards), "msm_precompute")

        return f


def workspace_size(libs: Libs, length: int, window: int, target: int) -> int:
    n = core_lib(libs).zkp_msm_workspace_size(length, window, target)
    if n == 0:
        raise ValueError(f"invalid MSM configuration window={window} target={target} len={length}")
    return int(n)


def msm_device(libs: Libs, d_points: int, d_scalars: int, length: int, window: int, target: int,
               workspace: hip.DeviceBuffer, stream: hip.Stream, debug: bool = False) -> np.ndarray:
    """sum_i k_i P_i for device-resident tables/scalars; returns the affine point (8 u64 words)."""
    out = np.zeros(8, dtype=np.uint64)
    hip.check(
        core_lib(libs).zkp_msm_device(d_points, d_scalars, length, window, target, workspace.addr, workspace.nbytes,
                                      stream.raw(), out.ctypes.data, 1 if debug else 0),
        "zkp_msm_device",
    )
    return out


def msm_config(libs: Libs, window: int, target: int) -> dict:
    """Derived MsmConfig values of the library (msm.cuh:51-70)."""
    out = (ctypes.c_uint * 5)()
    if not core_lib(libs).zkp_msm_config(window, target, out):
        raise ValueError(f"invalid MSM configuration window={window} target={target}")
    return dict(zip(("window", "actual_windows", "n_windows", "n_precompute", "buckets"), list(out)))


class DeviceMsm:
    """One device-resident MSM split at the host boundary: enqueue() puts the
    device work and the window sums' copy-out on a stream (graph-capturable,
    allocates nothing), result() combines the copied sums on the host."""

    def __init__(self, libs: Libs, d_points: int, d_scalars: int, length: int, window: int, target: int | None = None):
        self.lib = core_lib(libs)
        self.d_points, self.d_scalars, self.length = d_points, d_scalars, length
        self.window = window
        self.target = target or -(-SCALAR_BITS // window)
        n_windows = msm_config(libs, window, self.target)["n_windows"]
        self.workspace = hip.DeviceBuffer(workspace_size(libs, length, window, self.target))
        self.sums = hip.PinnedBuffer(n_windows * 128)

    def enqueue(self, stream: hip.Stream) -> None:
        hip.check(self.lib.zkp_msm_device_enqueue(self.d_points, self.d_scalars, self.length, self.window,
                                                  self.target, self.workspace.addr, self.workspace.nbytes,
                                                  stream.raw(), self.sums.addr), "zkp_msm_device_enqueue")

    def result(self) -> np.ndarray:
        out = np.zeros(8, dtype=np.uint64)
        hip.check(self.lib.zkp_msm_combine(self.window, self.target, self.sums.addr, out.ctypes.data),
                  "zkp_msm_combine")
        return out


def mul_generator(libs: Libs, d_scalars: int, n: int, d_out: int, stream: hip.Stream) -> None:
    hip.check(core_lib(libs).zkp_g1_mul_generator(d_scalars, n, d_out, stream.raw()), "zkp_g1_mul_generator")


def smoke_check(libs: Libs, stream: hip.Stream) -> None:
    """A 4096-point MSM through the device entry point, checked against the oracle."""
    import oracle as orc

    from .field import random_elements

    n = 4096
    ks = random_elements(n, 11)
    ss = random_elements(n, 12)
    kb = hip.DeviceBuffer(n * 32)
    kb.upload(ks)
    sb = hip.DeviceBuffer(n * 32)
    sb.upload(ss)
    pb = hip.DeviceBuffer(n * 64)
    mul_generator(libs, kb.addr, n, pb.addr, stream)
    stream.sync()
    ws = hip.DeviceBuffer(workspace_size(libs, n, 12, 22))
    got = msm_device(libs, pb.addr, sb.addr, n, 12, 22, ws, stream)
    bases = pb.download(np.uint64, n * 8).reshape(n, 8)
    want = orc.best_multiexp(ss, bases)
    if list(got) != list(want):
        raise AssertionError("smoke: GPU MSM differs from the oracle")
