"""Batched proof-step form: many NTTs of uneven sizes, each followed by a fused
elementwise kernel, and many MSMs of uneven sizes, run as one hipGraph launch.

The plain form issues, per instance, ssip_ntt (core/src/ntt.rs:165-200) and
the fused kernel (core/src/fused_kernels.rs:97-196) one call after another,
and every MSM as its own synchronous call -- what the reference runtime does
instruction by instruction.  The batched form plans all transforms together
(zkp_ntt_batch_create: one bit-reversal launch and one launch per pass round
for the whole batch), runs the fused expression over all instances with its
multi-instance entry, enqueues every MSM's device work with its window sums'
copy-out (zkp_msm_device_enqueue), and captures all of it in one hipGraph, so
a batch costs one graph launch; the MSMs' host combines (a few microseconds
each) follow the launch.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import hip
from .fused import FusedKernel
from .ntt import load_lib as load_ntt
from .runtime import Libs, PolyPtr


def _ntt_batch_api(libs: Libs) -> ctypes.CDLL:
    lib = load_ntt(libs)
    if not getattr(lib, "_zkp_batch_typed", False):
        P = ctypes.c_void_p
        lib.zkp_ntt_batch_create.argtypes = [P, P, P, ctypes.c_uint, ctypes.POINTER(ctypes.c_void_p)]
        lib.zkp_ntt_batch_create.restype = ctypes.c_int
        lib.zkp_ntt_batch_run.argtypes = [P, P]
        lib.zkp_ntt_batch_run.restype = ctypes.c_int
        lib.zkp_ntt_batch_destroy.argtypes = [P]
        lib.zkp_ntt_batch_destroy.restype = None
        lib._zkp_batch_typed = True
    return lib


class NttBatch:
    """A planned batch of in-place transforms (device work lists live on the GPU)."""

    def __init__(self, libs: Libs, polys: list[PolyPtr], twiddles: list[int], log_lens: list[int]):
        self.lib = _ntt_batch_api(libs)
        n = len(polys)
        self._xs = (PolyPtr * max(1, n))(*polys)
        self._tw = (ctypes.c_void_p * max(1, n))(*twiddles)
        self._ll = (ctypes.c_uint * max(1, n))(*log_lens)
        self.plan = ctypes.c_void_p(0)
        hip.check(self.lib.zkp_ntt_batch_create(self._xs, self._tw, self._ll, n, ctypes.byref(self.plan)),
                  "zkp_ntt_batch_create")

    def run(self, stream: hip.Stream) -> None:
        hip.check(self.lib.zkp_ntt_batch_run(self.plan, stream.raw()), "zkp_ntt_batch_run")

    def __del__(self):
        try:
            if self.plan.value:
                self.lib.zkp_ntt_batch_destroy(self.plan)
        except Exception:
            pass


@dataclass
class FusedBatch:
    """Device argument arrays for <name>_batched over instances of given lengths."""

    kernel: FusedKernel
    vars_: list  # PolyPtr, count * num_vars
    muts: list  # PolyPtr, count * num_mut_vars
    lens: list
    _bufs: list = field(default_factory=list)

    def __post_init__(self):
        va = (PolyPtr * max(1, len(self.vars_)))(*self.vars_)
        ma = (PolyPtr * max(1, len(self.muts)))(*self.muts)
        prefix = np.concatenate([[0], np.cumsum(self.lens)]).astype(np.uint64)
        self.total = int(prefix[-1])
        self.count = len(self.lens)
        dv = hip.DeviceBuffer(max(1, ctypes.sizeof(va)))
        dm = hip.DeviceBuffer(max(1, ctypes.sizeof(ma)))
        dp = hip.DeviceBuffer(prefix.nbytes)
        dv.upload(np.frombuffer(bytes(va), dtype=np.uint8))
        dm.upload(np.frombuffer(bytes(ma), dtype=np.uint8))
        dp.upload(prefix)
        local = hip.DeviceBuffer(self.kernel.local_bytes) if self.kernel.local_bytes else None
        self._bufs = [dv, dm, dp, local]

    def run(self, stream: hip.Stream) -> None:
        dv, dm, dp, local = self._bufs
        hip.check(self.kernel.c_batched(dv.addr, dm.addr, dp.addr, self.count, self.total, True, stream.raw(),
                                        local.addr if local is not None else None),
                  self.kernel.meta.name + "_batched")


def msm_branches(lengths: list[int], branches: int) -> list[list[int]]:
    """MSM indices per graph branch: largest first, each to the branch with the
    fewest points so far (ties to the lower branch), kept in index order within
    a branch."""
    loads = [0] * branches
    out = [[] for _ in range(branches)]
    for j in sorted(range(len(lengths)), key=lambda j: (-lengths[j], j)):
        b = min(range(branches), key=lambda k: (loads[k], k))
        out[b].append(j)
        loads[b] += lengths[j]
    return [sorted(x) for x in out if x]


class ProofStepBatch:
    """NTT batch + fused batch + MSMs captured as one hipGraph.

    The MSMs do not depend on the transforms, and most of an MSM below 2^18
    points is latency-bound (sort passes over few keys, the narrow reduction
    levels), so they run on `branches` parallel branches of the graph beside
    the NTT and fused kernels: the capture forks from the capturing stream
    onto side streams through an event and joins back through one event per
    branch."""

    def __init__(self, stream: hip.Stream, ntt: NttBatch, fused: FusedBatch, msms=(), branches: int = 4):
        """branches = 0 puts every MSM on the capturing stream after the fused kernel."""
        self.ntt, self.fused, self.msms = ntt, fused, list(msms)
        groups = msm_branches([m.length for m in self.msms], branches) if self.msms and branches > 0 else []
        self.branches = len(groups)  # parallel MSM branches in the graph (0: MSMs on the capturing stream)
        self._side = [hip.Stream(stream.get_device()) for _ in groups]
        self._fork = hip.Event()
        self._join = [hip.Event() for _ in groups]

        def fill(s):
            if groups:
                self._fork.record(s)
            for side, group in zip(self._side, groups):
                side.wait(self._fork)
                for j in group:
                    self.msms[j].enqueue(side)
            ntt.run(s)
            fused.run(s)
            if not groups:
                for m in self.msms:
                    m.enqueue(s)
            for side, ev in zip(self._side, self._join):
                ev.record(side)
                s.wait(ev)

        self.graph = hip.Graph(stream, fill)

    def launch(self, stream: hip.Stream) -> None:
        self.graph.launch(stream)

    def msm_results(self) -> list:
        """The MSMs' affine results once the launch has completed."""
        return [m.result() for m in self.msms]


def default_msm_window(log_n: int) -> int:
    """Window for an MSM of 2^log_n points when the autotuner has no entry."""
    return max(8, min(16, log_n - 4))


class Config4Workload:
    """`count` instances: NTT of x_i (log size uniform in [lo, hi], seeded), then
    out_i = a_i * b_i + gamma * c_i - z (seeded vectors of the same size, seeded
    scalars gamma, z), with the fused kernel `kernel` (vars [a, b, c, gamma, z]).
    `tuning` (an autotune.Tuning) is installed before the transforms are planned."""

    def __init__(self, libs: Libs, stream: hip.Stream, kernel: FusedKernel, count: int = 64, lo: int = 10,
                 hi: int = 22, seed: int = 4, tuning=None, msm_count: int = 0, msm_lo: int = 14, msm_hi: int = 18):
        from .field import mont_limbs, omega, random_elements
        from .ntt import gen_powers

        rng = np.random.Generator(np.random.PCG64(seed))
        self.log_lens = [int(x) for x in rng.integers(lo, hi + 1, size=count)]
        self.lens = [1 << k for k in self.log_lens]
        self.kernel = kernel
        self.stream = stream
        gz = random_elements(2, seed * 1000 + 999)
        self.gamma_z = gz
        self.d_gz = hip.DeviceBuffer(64)
        self.d_gz.upload(gz)
        self.tw = {}
        for k in sorted(set(self.log_lens)):
            w = hip.DeviceBuffer(32)
            w.upload(mont_limbs(omega(k)))
            t = hip.DeviceBuffer((1 << k) // 2 * 32)
            gen_powers(libs, t.addr, (1 << k) // 2, w.addr, stream)
            self.tw[k] = (t, w)
        self.host = []  # (x, a, b, c)
        self.dev = []  # DeviceBuffer x, a, b, c, out
        for i, n in enumerate(self.lens):
            arrs = [random_elements(n, seed * 100000 + 10 * i + j) for j in range(4)]
            self.host.append(arrs)
            bufs = []
            for a in arrs:
                b = hip.DeviceBuffer(n * 32)
                b.upload(a)
                bufs.append(b)
            bufs.append(hip.DeviceBuffer(n * 32))
            self.dev.append(bufs)
        stream.sync()
        self.x_ptrs = [PolyPtr(d[0].addr, n, 0, 0, n) for d, n in zip(self.dev, self.lens)]
        g_ptr = PolyPtr(self.d_gz.addr, 1, 0, 0, 1)
        z_ptr = PolyPtr(self.d_gz.addr + 32, 1, 0, 0, 1)
        self.fused_vars = []
        self.fused_muts = []
        for d, n in zip(self.dev, self.lens):
            self.fused_vars += [PolyPtr(d[1].addr, n, 0, 0, n), PolyPtr(d[2].addr, n, 0, 0, n),
                                PolyPtr(d[3].addr, n, 0, 0, n), g_ptr, z_ptr]
            self.fused_muts.append(PolyPtr(d[4].addr, n, 0, 0, n))
        # per-instance device argument arrays for the plain (one call per instance) path
        self.plain_args = []
        for i in range(count):
            va = (PolyPtr * 5)(*self.fused_vars[5 * i : 5 * i + 5])
            ma = (PolyPtr * 1)(self.fused_muts[i])
            dv, dm = hip.DeviceBuffer(ctypes.sizeof(va)), hip.DeviceBuffer(ctypes.sizeof(ma))
            dv.upload(np.frombuffer(bytes(va), dtype=np.uint8))
            dm.upload(np.frombuffer(bytes(ma), dtype=np.uint8))
            self.plain_args.append((dv, dm))
        self.ntt_lib = load_ntt(libs)
        if tuning is not None:
            tuning.apply(libs)
        self.ntt_batch = NttBatch(libs, self.x_ptrs, [self.tw[k][0].addr for k in self.log_lens], self.log_lens)
        self.fused_batch = FusedBatch(kernel, self.fused_vars, self.fused_muts, self.lens)
        self.total_elements = sum(self.lens)
        self._setup_msms(libs, rng, seed, tuning, msm_count, msm_lo, msm_hi)

    def _setup_msms(self, libs: Libs, rng, seed: int, tuning, count: int, lo: int, hi: int) -> None:
        """`count` MSMs of 2^U{lo..hi} points: prefixes of one resident table of
        bases P_i = k_i G, each with its own seeded scalars."""
        from .field import random_elements
        from .msm import DeviceMsm, mul_generator, msm_device

        self._msm_device = msm_device
        self.libs = libs
        self.total_msm_points = 0
        self.msm_log_lens = [int(x) for x in rng.integers(lo, hi + 1, size=count)]
        self.msm_lens = [1 << k for k in self.msm_log_lens]
        self.msms, self.msm_scalars = [], []
        self.msm_base_scalars = None
        if count == 0:
            return
        top = max(self.msm_lens)
        self.msm_base_scalars = random_elements(top, seed * 1000 + 77)
        kb = hip.DeviceBuffer(top * 32)
        kb.upload(self.msm_base_scalars)
        self.msm_bases = hip.DeviceBuffer(top * 64)
        mul_generator(libs, kb.addr, top, self.msm_bases.addr, self.stream)
        self._msm_scalar_bufs = []
        for j, n in enumerate(self.msm_lens):
            sc = random_elements(n, seed * 1000 + 100 + j)
            b = hip.DeviceBuffer(n * 32)
            b.upload(sc)
            self.msm_scalars.append(sc)
            self._msm_scalar_bufs.append(b)
            k = self.msm_log_lens[j]
            w = tuning.msm_window(k, default_msm_window(k)) if tuning is not None else default_msm_window(k)
            self.msms.append(DeviceMsm(libs, self.msm_bases.addr, b.addr, n, w))
        # one eager run before any capture: the first MSM on a device sets launch limits
        for m in self.msms:
            m.enqueue(self.stream)
        self.stream.sync()
        self.total_msm_points = sum(self.msm_lens)

    def reset_inputs(self) -> None:
        for d, h in zip(self.dev, self.host):
            d[0].upload(h[0])

    def run_plain(self) -> list:
        """One ssip_ntt + one fused call per instance, in order, then every MSM
        as its own synchronous call; returns the MSM results."""
        s = self.stream
        for i, k in enumerate(self.log_lens):
            hip.check(self.ntt_lib.ssip_ntt(self.x_ptrs[i], self.tw[k][0].addr, k, s.raw(), 8, 8), "ssip_ntt")
            dv, dm = self.plain_args[i]
            hip.check(self.kernel.c_func(dv.addr, dm.addr, self.lens[i], True, None, s.raw()), "fused")
        out = []
        for m in self.msms:
            out.append(self._msm_device(self.libs, m.d_points, m.d_scalars, m.length, m.window, m.target,
                                        m.workspace, s))
        return out

    def capture(self, branches: int = 4) -> ProofStepBatch:
        return ProofStepBatch(self.stream, self.ntt_batch, self.fused_batch, self.msms, branches)

    def results(self, i: int) -> tuple[np.ndarray, np.ndarray]:
        n = self.lens[i]
        x = self.dev[i][0].download(np.uint64, n * 4).reshape(n, 4)
        out = self.dev[i][4].download(np.uint64, n * 4).reshape(n, 4)
        return x, out
