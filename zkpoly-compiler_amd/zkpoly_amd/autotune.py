"""Project autotuner: NTT pass depth per transform size and MSM window per size.

Candidates are timed on the device with HIP events and the fastest per size is
kept.  The choice is stored as JSON and applied before work is planned:
  * NTT: `min_col_log` per log2 size through zkp_ntt_set_tuning, i.e. how
    many butterfly stages each LDS pass takes (1: fewest, deepest passes);
    every candidate computes the same network, so results are bit-identical.
  * MSM: the window s for msm / zkp_msm_device (12..18), with one window set
    per scalar (no precomputed tables).
The reference picks its MSM configuration from an analytic cost model
(core/src/msm.rs:35-98, core/src/msm/tuning/cost_model.py:98-309) and its
NTT stage split from fixed thread limits (core/src/ntt.rs:305-335); here both
are measured on the GPU the work runs on.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field
from typing import Callable, Iterable

NTT_CANDIDATES = (1, 2, 3)
MSM_WINDOWS = tuple(range(12, 19))
DEFAULT_MIN_COL_LOG = 1
DEFAULT_MSM_WINDOW = 16


@dataclass
class Tuning:
    ntt: dict = field(default_factory=dict)  # log_n -> min_col_log
    msm: dict = field(default_factory=dict)  # log_n -> window
    timings: dict = field(default_factory=dict)  # "ntt/<log_n>/<c>" or "msm/<log_n>/<w>" -> ms
    device: str = ""

    def to_json(self) -> str:
        return json.dumps(
            {
                "ntt": {str(k): v for k, v in sorted(self.ntt.items())},
                "msm": {str(k): v for k, v in sorted(self.msm.items())},
                "timings": self.timings,
                "device": self.device,
            },
            indent=1,
            sort_keys=True,
        )

    @staticmethod
    def from_json(text: str) -> "Tuning":
        d = json.loads(text)
        return Tuning(
            ntt={int(k): int(v) for k, v in d.get("ntt", {}).items()},
            msm={int(k): int(v) for k, v in d.get("msm", {}).items()},
            timings=dict(d.get("timings", {})),
            device=str(d.get("device", "")),
        )

    def save(self, path: str) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            f.write(self.to_json())
        os.replace(tmp, path)

    def apply(self, libs) -> None:
        """Installs the NTT choices in libntt (process-wide, before planning)."""
        from .ntt import set_tuning

        for log_n, c in self.ntt.items():
            set_tuning(libs, log_n, c)

    def msm_window(self, log_n: int, default: int = DEFAULT_MSM_WINDOW) -> int:
        return self.msm.get(log_n, default)

    def merge(self, other: "Tuning") -> "Tuning":
        return Tuning({**self.ntt, **other.ntt}, {**self.msm, **other.msm}, {**self.timings, **other.timings},
                      other.device or self.device)


def load(path: str) -> Tuning | None:
    try:
        with open(path) as f:
            return Tuning.from_json(f.read())
    except (OSError, ValueError):
        return None


def default_cache_path() -> str:
    env = os.environ.get("ZKPOLY_TUNE_CACHE")
    if env:
        return env
    return os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib", "tuning.json")


def pick(times: dict) -> object:
    """Fastest candidate; ties go to the candidate listed first."""
    best, best_t = None, float("inf")
    for cand, t in times.items():
        if t < best_t:
            best, best_t = cand, t
    return best


def tune(kind: str, sizes: Iterable[int], candidates: Iterable[int], measure: Callable[[int, int], float],
         rounds: int = 2) -> Tuning:
    """Times measure(size, candidate) for every pair in `rounds` interleaved
    rounds (so clock ramp-up or a noisy moment does not favour one candidate)
    and keeps the argmin of each candidate's best time per size."""
    out = Tuning()
    cands = list(candidates)
    for size in sizes:
        times = {c: float("inf") for c in cands}
        for _ in range(max(1, rounds)):
            for c in cands:
                times[c] = min(times[c], float(measure(size, c)))
        for c in cands:
            out.timings[f"{kind}/{size}/{c}"] = round(times[c], 5)
        (out.ntt if kind == "ntt" else out.msm)[size] = pick(times)
    return out


# ---------------------------------------------------------------- device measurements

def _device_name() -> str:
    from . import hip

    try:
        return hip.device_name(hip.get_device())
    except Exception:
        return "hip"


class NttTimer:
    """measure(log_n, min_col_log): mean ms of ssip_ntt on a resident 2^log_n input."""

    def __init__(self, libs, stream, steps: int = 5, warmup: int = 3):
        self.libs, self.stream, self.steps, self.warmup = libs, stream, steps, warmup
        self._cache = {}

    def _setup(self, k: int):
        if k not in self._cache:
            from . import hip
            from .field import mont_limbs, omega, random_elements
            from .ntt import gen_powers

            n = 1 << k
            x = hip.DeviceBuffer(n * 32)
            tw = hip.DeviceBuffer(max(1, n // 2) * 32)
            b = hip.DeviceBuffer(32)
            b.upload(random_elements(1, 1000 + k))
            gen_powers(self.libs, x.addr, n, b.addr, self.stream)  # input x[i] = b^i
            w = hip.DeviceBuffer(32)
            w.upload(mont_limbs(omega(k)))
            gen_powers(self.libs, tw.addr, max(1, n // 2), w.addr, self.stream)
            self.stream.sync()
            self._cache = {k: (x, tw, b, w)}  # keep one size resident at a time
        return self._cache[k]

    def __call__(self, k: int, c: int) -> float:
        from . import hip
        from .ntt import load_lib, set_tuning
        from .runtime import PolyPtr

        x, tw, _, _ = self._setup(k)
        lib = load_lib(self.libs)
        set_tuning(self.libs, k, c)
        ptr = PolyPtr(x.addr, 1 << k, 0, 0, 1 << k)

        def run():
            hip.check(lib.ssip_ntt(ptr, tw.addr, k, self.stream.raw(), 8, 8), "ssip_ntt")

        for _ in range(self.warmup):
            run()
        a, e = hip.Event(), hip.Event()
        a.record(self.stream)
        for _ in range(self.steps):
            run()
        e.record(self.stream)
        e.sync()
        return a.elapsed_ms(e) / self.steps


class MsmTimer:
    """measure(log_n, window): mean ms of zkp_msm_device over resident bases k_i G."""

    def __init__(self, libs, stream, steps: int = 3, warmup: int = 1, seed: int = 77):
        self.libs, self.stream, self.steps, self.warmup, self.seed = libs, stream, steps, warmup, seed
        self._k = None

    def _setup(self, k: int):
        if self._k != k:
            from . import hip
            from .field import random_elements
            from .msm import mul_generator

            n = 1 << k
            ks = hip.DeviceBuffer(n * 32)
            ks.upload(random_elements(n, self.seed + k))
            self.points = hip.DeviceBuffer(n * 64)
            mul_generator(self.libs, ks.addr, n, self.points.addr, self.stream)
            self.scalars = hip.DeviceBuffer(n * 32)
            self.scalars.upload(random_elements(n, self.seed + 100 + k))
            self.stream.sync()
            self._k = k
            self._ws = None

    def __call__(self, k: int, w: int) -> float:
        from . import hip
        from .msm import msm_device, workspace_size

        self._setup(k)
        target = -(-254 // w)
        need = workspace_size(self.libs, 1 << k, w, target)
        if self._ws is None or self._ws.nbytes < need:
            self._ws = None
            self._ws = hip.DeviceBuffer(need)
        for _ in range(self.warmup):
            msm_device(self.libs, self.points.addr, self.scalars.addr, 1 << k, w, target, self._ws, self.stream)
        ev = [hip.Event() for _ in range(2)]
        ev[0].record(self.stream)
        for _ in range(self.steps):
            msm_device(self.libs, self.points.addr, self.scalars.addr, 1 << k, w, target, self._ws, self.stream)
        ev[1].record(self.stream)
        ev[1].sync()
        return ev[0].elapsed_ms(ev[1]) / self.steps


def autotune(libs, stream, ntt_sizes: Iterable[int] = (), msm_sizes: Iterable[int] = (),
             ntt_candidates=NTT_CANDIDATES, msm_windows=MSM_WINDOWS, path: str | None = None) -> Tuning:
    """Measures, applies the NTT choices, and stores the result at `path` (if given)."""
    t = Tuning(device=_device_name())
    ntt_sizes, msm_sizes = list(ntt_sizes), list(msm_sizes)
    if ntt_sizes:
        t = t.merge(tune("ntt", ntt_sizes, ntt_candidates, NttTimer(libs, stream)))
    if msm_sizes:
        t = t.merge(tune("msm", msm_sizes, msm_windows, MsmTimer(libs, stream)))
    t.apply(libs)
    if path:
        t.save(path)
    return t


__all__ = ["Tuning", "tune", "pick", "load", "autotune", "NttTimer", "MsmTimer", "default_cache_path",
           "NTT_CANDIDATES", "MSM_WINDOWS"]
