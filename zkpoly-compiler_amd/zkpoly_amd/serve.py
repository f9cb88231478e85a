"""Serving form: Poisson streams of NTT / MSM proof-step requests through the
continuous-batching engine in libzkp_serve.so (see csrc/serve/serve.cpp).

Workload (70 % NTT with log2 size uniform in 14..22, 30 % MSM with log2 size
uniform in 16..21, seeded inputs, Poisson arrivals at rate lambda): this module
builds the trace and the resident state, runs it, and reports p50 / p99
latency (arrival -> result on the host) and aggregate elements/s.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import hip
from .field import mont_limbs, omega, random_elements
from .runtime import Libs

LIB = "libzkp_serve.so"


class ServeConfig(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_uint),
        ("hbm_budget", ctypes.c_ulonglong),
        ("ntt_min_log", ctypes.c_uint),
        ("ntt_max_log", ctypes.c_uint),
        ("msm_max_log", ctypes.c_uint),
        ("msm_window", ctypes.c_uint),
        ("msm_bases_host", ctypes.c_void_p),
        ("twiddles_host", ctypes.c_void_p),
    ]


class ServeRequest(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_uint),
        ("log_n", ctypes.c_uint),
        ("arrival_s", ctypes.c_double),
        ("input", ctypes.c_void_p),
        ("output", ctypes.c_void_p),
        ("start_s", ctypes.c_double),
        ("finish_s", ctypes.c_double),
        ("status", ctypes.c_int),
    ]


def _lib(libs: Libs) -> ctypes.CDLL:
    lib = libs.load_relative(LIB)
    if not getattr(lib, "_zkp_typed", False):
        P = ctypes.c_void_p
        lib.zkp_serve_create.argtypes = [ctypes.POINTER(ServeConfig), ctypes.POINTER(P)]
        lib.zkp_serve_create.restype = ctypes.c_int
        lib.zkp_serve_run.argtypes = [P, ctypes.POINTER(ServeRequest), ctypes.c_uint]
        lib.zkp_serve_run.restype = ctypes.c_int
        lib.zkp_serve_footprint.argtypes = [P, ctypes.c_uint, ctypes.c_uint]
        lib.zkp_serve_footprint.restype = ctypes.c_ulonglong
        lib.zkp_serve_destroy.argtypes = [P]
        lib.zkp_serve_destroy.restype = None
        lib._zkp_typed = True
    return lib


def host_twiddles(libs: Libs, stream: hip.Stream, k: int) -> np.ndarray:
    """[1, w, ..., w^(2^(k-1)-1)] for w = omega_{2^k}, generated on the GPU, on the host."""
    from .ntt import gen_powers

    n = 1 << (k - 1)
    w = hip.DeviceBuffer(32)
    w.upload(mont_limbs(omega(k)))
    t = hip.DeviceBuffer(n * 32)
    gen_powers(libs, t.addr, n, w.addr, stream)
    stream.sync()
    return t.download(np.uint64, n * 4).reshape(n, 4)


@dataclass
class Resident:
    ntt_min: int
    ntt_max: int
    msm_max: int
    twiddles: list
    bases: np.ndarray  # (2^msm_max, 8)
    base_scalars: np.ndarray  # k_i of bases = k_i G


def make_resident(libs: Libs, stream: hip.Stream, ntt_min=14, ntt_max=22, msm_max=21, seed=55) -> Resident:
    from .msm import mul_generator

    tw = [np.ascontiguousarray(host_twiddles(libs, stream, k)) for k in range(ntt_min, ntt_max + 1)]
    n = 1 << msm_max
    ks = random_elements(n, seed)
    kb = hip.DeviceBuffer(n * 32)
    kb.upload(ks)
    pb = hip.DeviceBuffer(n * 64)
    mul_generator(libs, kb.addr, n, pb.addr, stream)
    stream.sync()
    bases = np.ascontiguousarray(pb.download(np.uint64, n * 8).reshape(n, 8))
    return Resident(ntt_min, ntt_max, msm_max, tw, bases, ks)


class Server:
    def __init__(self, libs: Libs, res: Resident, slots: int = 4, hbm_budget: int = 8 << 30, msm_window: int = 16):
        self.lib = _lib(libs)
        self.res = res
        self._tw_ptrs = (ctypes.c_void_p * len(res.twiddles))(*[t.ctypes.data for t in res.twiddles])
        cfg = ServeConfig(slots, hbm_budget, res.ntt_min, res.ntt_max, res.msm_max, msm_window,
                          res.bases.ctypes.data, ctypes.cast(self._tw_ptrs, ctypes.c_void_p))
        self.handle = ctypes.c_void_p(0)
        hip.check(self.lib.zkp_serve_create(ctypes.byref(cfg), ctypes.byref(self.handle)), "zkp_serve_create")

    def footprint(self, kind: int, log_n: int) -> int:
        return int(self.lib.zkp_serve_footprint(self.handle, kind, log_n))

    def run(self, reqs) -> None:
        arr = (ServeRequest * len(reqs))(*reqs)
        rc = self.lib.zkp_serve_run(self.handle, arr, len(reqs))
        for i in range(len(reqs)):
            reqs[i].start_s, reqs[i].finish_s, reqs[i].status = arr[i].start_s, arr[i].finish_s, arr[i].status
        hip.check(rc, "zkp_serve_run")

    def __del__(self):
        try:
            if self.handle.value:
                self.lib.zkp_serve_destroy(self.handle)
        except Exception:
            pass


@dataclass
class Job:
    kind: int
    log_n: int
    arrival: float
    input_id: int


class Trace:
    """Poisson arrivals; each request reuses one of `variants` seeded inputs of its
    size (inputs are read-only), and gets its own pinned output buffer."""

    def __init__(self, rate: float, count: int, seed: int, variants: int = 2, ntt_range=(14, 22), msm_range=(16, 21)):
        rng = np.random.Generator(np.random.PCG64(seed))
        gaps = rng.exponential(1.0 / rate, size=count)
        arrivals = np.cumsum(gaps) - gaps[0]
        self.jobs = []
        for i in range(count):
            kind = 0 if rng.random() < 0.7 else 1
            lo, hi = ntt_range if kind == 0 else msm_range
            log_n = int(rng.integers(lo, hi + 1))
            self.jobs.append(Job(kind, log_n, float(arrivals[i]), int(rng.integers(0, variants))))
        self.inputs: dict = {}
        for j in self.jobs:
            key = (j.kind, j.log_n, j.input_id)
            if key not in self.inputs:
                n = 1 << j.log_n
                pb = hip.PinnedBuffer(n * 32)
                pb.view()[:] = random_elements(n, 7919 * seed + 1000 * j.kind + 10 * j.log_n + j.input_id).reshape(-1)
                self.inputs[key] = pb
        self.outputs = []
        for j in self.jobs:
            n = 1 << j.log_n
            self.outputs.append(hip.PinnedBuffer(n * 32 if j.kind == 0 else 64))

    def requests(self):
        out = []
        for j, o in zip(self.jobs, self.outputs):
            inp = self.inputs[(j.kind, j.log_n, j.input_id)]
            out.append(ServeRequest(j.kind, j.log_n, j.arrival, inp.addr, o.addr, 0.0, 0.0, 0))
        return out

    def output(self, i: int) -> np.ndarray:
        j = self.jobs[i]
        v = self.outputs[i].view()
        return v.reshape(-1, 4) if j.kind == 0 else v.reshape(8)

    def input(self, key) -> np.ndarray:
        return self.inputs[key].view().reshape(-1, 4)


def latency_stats(reqs, jobs) -> dict:
    lat = np.array([r.finish_s - r.arrival_s for r in reqs]) * 1e3
    wall = max(r.finish_s for r in reqs) - min(r.arrival_s for r in reqs)
    elems = sum(1 << j.log_n for j in jobs)
    return {
        "requests": len(reqs),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "max_ms": round(float(lat.max()), 3),
        "wall_s": round(wall, 4),
        "elements": elems,
        "gelem_s": round(elems / wall / 1e9, 4),
        "achieved_req_s": round(len(reqs) / wall, 2),
    }
