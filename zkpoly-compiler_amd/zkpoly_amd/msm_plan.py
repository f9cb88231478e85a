"""MSM configuration under a device-memory limit: the get_best_config analogue.

The reference's get_best_config (core/src/msm.rs:35-98) runs an analytic cost
model (core/src/msm/tuning/cost_model.py:98-309) over a grid of
  s      window bits,
  alpha  windows per precomputed point table (MsmConfig.target_window),
  c      points per part (2^ec), w = 2^k / c parts,
  h      batches per run (MsmConfig.batch_per_run),
keeps the candidates whose modelled peak memory is below the limit, and takes
the first one of least modelled time; MsmConfig then gets stage_scalars =
stage_points = 1 for one part and 2 otherwise.

Two planners here:
  * `reference_choice` restates that model with its A100 coefficients, so the
    choices the reference would make are reproduced (pinned against rows the
    reference's script wrote, tests/golden/msm_cost_model_rows.json);
  * `plan_msm` is the planner of this library for MI355X: the same search space,
    each candidate's device memory taken from the library's own sizing call
    (exact, including the stage buffers and the batch_per_run bucket sets),
    and its time from a model of this implementation -- per-entry sort and
    accumulation costs per window measured on MI355X, the bucket reduction per
    window and batch, and PCIe copies overlapped with compute when there is
    more than one stage.  Windows outside 8..21 are not offered (the library's
    sort paths cover 2..21; below 8 the entry count explodes).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Iterable, Optional

from .msm import MsmConfig

SCALAR_BITS = 254


@dataclass(frozen=True)
class Choice:
    alpha: int  # windows per table (target_window)
    s: int  # window bits
    c: int  # points per part
    w: int  # parts
    h: int  # batches per run

    def to_config(self, bits: int = SCALAR_BITS, cards=(0,)) -> MsmConfig:
        stages = 1 if self.w == 1 else 2
        return MsmConfig(window_size=self.s, target_window=self.alpha, cards=list(cards), batch_per_run=self.h,
                         parts=self.w, stage_scalars=stages, stage_points=stages, bits=bits)


def _ceil_div(a: int, b: int) -> int:
    return (a + b - 1) // b


# ---------------------------------------------------------------- the reference's model

REF_COEFFS = {"scatter": 1.0, "bucket_sum": 7.0, "reduction": 40.0, "amp_scatter": 1.5, "amp_sum": 1.05,
              "transfer": 0.8}
REF_SMALL_SCATTER = 2 ** 17  # below these part sizes the per-entry costs are amplified
REF_SMALL_SUM = 2 ** 20


def _ref_windows(p_bits: int) -> range:
    return range(8, 24) if p_bits < 400 else range(16, 31)


def _ref_part_logs(k: int, p_bits: int) -> range:
    if k < 18:
        return range(max(k - 4, 1), k + 1)
    if k <= 25:
        return range(18, k + 1)
    return range(22 if p_bits < 400 else 20, k + 1)


def _ref_runs(n: int) -> list:
    if n == 1:
        return [1]
    out, v = [], 1
    while v <= n:
        v *= 2
        out.append(v)
    return out


def _ref_time_and_memory(k: int, n: int, l_bits: int, p_bits: int, s: int, alpha: int, ec: int, h: int,
                         co: dict) -> tuple[float, float]:
    """(modelled time, modelled peak bytes) of one candidate, in the
    reference's units; the arithmetic follows cost_model.py's order."""
    c = 2 ** ec
    w = 2 ** (k - ec)
    windows = _ceil_div(l_bits, s)
    scalar_b = _ceil_div(l_bits, 8)
    point_b = _ceil_div(p_bits, 8)
    s_points = _ceil_div(l_bits, s * alpha) * c * (point_b * 2)
    s_scalars = c * scalar_b
    t_points = s_points * co["transfer"]
    t_scalars = s_scalars * co["transfer"]
    t_scatter = co["scatter"] * (c * windows)
    if c < REF_SMALL_SCATTER:
        t_scatter = co["scatter"] * (c * windows) * (co["amp_scatter"] ** math.log2(REF_SMALL_SCATTER / c))
    t_sum = co["bucket_sum"] * (c * windows)
    if c < REF_SMALL_SUM:
        t_sum = co["bucket_sum"] * (c * windows) * (co["amp_sum"] ** math.log2(REF_SMALL_SUM / c))
    t_red = co["reduction"] * ((2 ** s) * alpha)
    n_rows = _ceil_div(n, h)
    n_cols = _ceil_div(2 ** k, w * c)
    t_first = t_scalars + max(t_scatter, t_points)
    t_inner = max(h * t_scalars + t_points, h * (t_scatter + t_sum))
    t_last = h * t_sum + h * t_red + (h - 1) * t_scatter
    total = n_rows * n_cols * (t_first + (w - 1) * t_inner + t_last)
    s_indices = c * windows * 8
    s_buckets = (2 ** s) * alpha * (point_b * 4) / 2
    mem = s_points * 4 + s_scalars * 4 + s_indices * 4 + h * s_buckets * 4
    return total, mem


def reference_choice(k: int, n: int, l_bits: int = SCALAR_BITS, p_bits: int = SCALAR_BITS,
                     mem: int = 40 * 2 ** 30, coeffs: Optional[dict] = None) -> Optional[Choice]:
    """The choice get_best_config gets from cost_model.py for --k --n --l --p
    --mem, or None when no candidate fits (the script then fails)."""
    co = coeffs or REF_COEFFS
    best, best_t = None, None
    for s in _ref_windows(p_bits):
        for alpha in range(1, _ceil_div(l_bits, s) + 1):
            for ec in _ref_part_logs(k, p_bits):
                for h in _ref_runs(n):
                    t, m = _ref_time_and_memory(k, n, l_bits, p_bits, s, alpha, ec, h, co)
                    if m < mem and (best_t is None or t < best_t):
                        best, best_t = Choice(alpha, s, 2 ** ec, 2 ** (k - ec), h), t
    return best


# ---------------------------------------------------------------- the library's planner (MI355X)

# Per-entry cost (ns per (point, window) entry) of the sort plus the bucket
# accumulation, by window, from zkp_msm_device timings at 2^20 points on one
# MI355X (window sweep 12..18) less the modelled reduction and host combine;
# other windows take the nearest measured neighbour's cost, 5 % dearer per bit
# of distance.
MI355X_ENTRY_NS = {12: 0.117, 13: 0.121, 14: 0.123, 15: 0.107, 16: 0.100, 17: 0.108, 18: 0.124}
MI355X_REDUCE_US_PER_LEVEL = 5.0  # bucket reduction: about 2 (s - 1) dependent point operations deep (quads)
MI355X_REDUCE_US_PER_KBUCKET = 0.35  # ... plus its work, per 1024 buckets (n_windows * 2^(s-1))
MI355X_HOST_US = 60.0  # host combine (Horner over the window sums, one inversion) per batch
MI355X_PART_US = 250.0  # fixed cost of one part (about 25 launches, their tails, the partial levels of a smaller part)
MI355X_SMALL_PART = 2 ** 18  # below this part size per-entry costs grow (fewer waves per launch)
MI355X_LINK_GBS = 55.0  # host -> device copy rate from pinned memory
MI355X_MERGE_US_PER_KBUCKET = 0.15  # a later part's scratch buckets reset and merged, per 1024 buckets
MI355X_SORT_SHARE = 0.24  # share of the per-entry cost spent sorting (needs the scalars only)


def _entry_ns(s: int) -> float:
    if s in MI355X_ENTRY_NS:
        return MI355X_ENTRY_NS[s]
    near = min(MI355X_ENTRY_NS, key=lambda w: (abs(w - s), w))
    return MI355X_ENTRY_NS[near] * (1.0 + 0.05 * abs(s - near))


def model_time_us(cfg: MsmConfig, length: int, batches: int) -> float:
    """Modelled wall time (us) of msm(cfg) on `length` host points and
    `batches` host scalar arrays in pinned memory, following how the library
    runs it: per run of batch_per_run batches one pass over the parts (the
    first part of every run after the first is still resident, the pass
    direction alternates), each part sorted and accumulated once per batch;
    with more than one stage the copies of the next part and scalars overlap
    the current accumulation; then one reduction and host combine per batch."""
    windows = _ceil_div(cfg.bits, cfg.window_size)
    n_windows = min(windows, cfg.target_window)
    n_pre = cfg.get_precompute()
    parts = max(1, min(cfg.parts, length))
    c = _ceil_div(length, parts)
    amp = max(1.0, (MI355X_SMALL_PART / max(c, 1)) ** 0.25)
    buckets = n_windows * 2 ** (cfg.window_size - 1)
    first = MI355X_PART_US + c * windows * _entry_ns(cfg.window_size) * amp / 1000.0
    # the parts after a run's first accumulate into a scratch bucket set that is merged afterwards
    compute = first + (buckets / 1024 * MI355X_MERGE_US_PER_KBUCKET if parts > 1 else 0.0)
    copy_points = c * n_pre * 64 / (MI355X_LINK_GBS * 1e3)
    copy_scalars = c * 32 / (MI355X_LINK_GBS * 1e3)
    reduce = (MI355X_REDUCE_US_PER_LEVEL * 2 * (cfg.window_size - 1) + buckets / 1024 * MI355X_REDUCE_US_PER_KBUCKET
              + MI355X_HOST_US)
    bpr = max(1, cfg.batch_per_run)
    overlap = cfg.stage_points > 1 and cfg.stage_scalars > 1
    total = 0.0
    for r in range(_ceil_div(batches, bpr)):
        h = min(bpr, batches - r * bpr)
        copies = parts if r == 0 else parts - 1  # parts whose points cross the link this run
        sort = (compute - MI355X_PART_US) * MI355X_SORT_SHARE
        if overlap:
            first = copy_points if r == 0 else 0.0
            total += first + copy_scalars + parts * max(h * compute, h * copy_scalars + copy_points * copies / parts)
        elif copies:
            # each part: its scalars, then its points while the part's first sort runs
            total += parts * (copy_scalars + max(copy_points * copies / parts, sort) + (compute - sort))
            total += (h - 1) * parts * max(copy_scalars, compute)  # later batches' scalars arrive during the previous
        else:
            total += parts * (copy_scalars + compute) + (h - 1) * parts * max(copy_scalars, compute)
        total += h * reduce
    return total


def plan_msm(length: int, batches: int, mem_limit: int, libs=None, bits: int = SCALAR_BITS,
             windows: Iterable[int] = range(8, 22), memory: Optional[Callable[[MsmConfig, int], int]] = None,
             cards=(0,)) -> Optional[MsmConfig]:
    """Fastest modelled configuration whose device buffer fits mem_limit bytes.

    memory(cfg, length) -> bytes defaults to the library's sizing call (the
    msm symbol with buffers == NULL), which counts the stage buffers, the
    batch_per_run bucket sets and the sort workspace exactly."""
    if memory is None:
        from .msm import MSM
        from .runtime import Libs

        libs = libs or Libs()

        def memory(cfg: MsmConfig, n: int) -> int:
            return MSM(libs, cfg).get_buffer_size(n)[0]

    k = max(1, (length - 1).bit_length())
    best, best_t = None, None
    part_logs = range(max(0, k - 6), k + 1)
    for s in windows:
        actual = _ceil_div(bits, s)
        # no precomputation, or tables for half / a third / a quarter of the windows
        alphas = sorted({actual, _ceil_div(actual, 2), _ceil_div(actual, 3), _ceil_div(actual, 4)}, reverse=True)
        for alpha in alphas:
            for ec in part_logs:
                parts = 2 ** (k - ec)
                for h in sorted({1, min(batches, 2), min(batches, 4), batches}):
                    stages = 1 if parts == 1 else 2
                    cfg = MsmConfig(window_size=s, target_window=alpha, cards=list(cards), batch_per_run=h,
                                    parts=parts, stage_scalars=stages, stage_points=stages, bits=bits)
                    t = model_time_us(cfg, length, batches)
                    if best_t is not None and t >= best_t:
                        continue
                    if memory(cfg, length) > mem_limit:
                        continue
                    best, best_t = cfg, t
    return best


__all__ = ["Choice", "reference_choice", "plan_msm", "model_time_us", "REF_COEFFS"]
