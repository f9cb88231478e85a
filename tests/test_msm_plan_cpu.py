"""MSM configuration planners on the CPU.

* reference_choice reproduces every row the reference's cost model wrote
  (tests/golden/msm_cost_model_rows.json, made by make_msm_config_rows.py from
  core/src/msm/tuning/cost_model.py): alpha, s, c, w, h per (k, n, mem), and
  no choice where the script found no candidate under the memory limit.
* plan_msm keeps every choice within the memory limit as the library's own
  sizing call counts it, trades parts / batch_per_run / precomputation for
  memory as the limit tightens, and gives up only when nothing fits.
"""
import json
import os

import pytest

from zkpoly_amd.msm import MSM, MsmConfig
from zkpoly_amd.msm_plan import Choice, model_time_us, plan_msm, reference_choice
from zkpoly_amd.runtime import Libs

ROWS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "msm_cost_model_rows.json")))["rows"]


@pytest.mark.parametrize("row", ROWS, ids=lambda r: f"k{r['k']}-n{r['n']}-mem{r['mem'] >> 30}G")
def test_reference_choice_matches_cost_model_rows(row):
    got = reference_choice(row["k"], row["n"], 254, 254, row["mem"])
    if "error" in row:
        assert got is None
    else:
        assert got == Choice(row["alpha"], row["s"], row["c"], row["w"], row["h"])


def test_choice_maps_to_msm_config_like_get_best_config():
    # core/src/msm.rs:77-97: (precompute, window, _, parts, batch_per_run); stages 1 or 2
    cfg = Choice(alpha=17, s=15, c=1 << 18, w=4, h=2).to_config()
    assert (cfg.window_size, cfg.target_window, cfg.parts, cfg.batch_per_run) == (15, 17, 4, 2)
    assert (cfg.stage_scalars, cfg.stage_points) == (2, 2)
    one = Choice(alpha=16, s=16, c=1 << 20, w=1, h=1).to_config()
    assert (one.stage_scalars, one.stage_points) == (1, 1)


@pytest.fixture(scope="module")
def libs():
    return Libs()


def sizing(libs):
    def memory(cfg, n):
        return MSM(libs, cfg).get_buffer_size(n)[0]

    return memory


@pytest.mark.parametrize("length,batches", [(1 << 20, 1), (1 << 20, 4), (1 << 16, 3), (3 << 18, 2)])
def test_plan_fits_memory_and_tightens(libs, length, batches):
    mem = sizing(libs)
    loose = plan_msm(length, batches, 288 << 30, libs=libs)
    assert loose is not None and mem(loose, length) <= 288 << 30
    need = mem(loose, length)
    tight_limit = need // 3
    tight = plan_msm(length, batches, tight_limit, libs=libs)
    if tight is not None:
        assert mem(tight, length) <= tight_limit
        assert model_time_us(tight, length, batches) >= model_time_us(loose, length, batches)
    assert plan_msm(length, batches, 1 << 20, libs=libs) is None  # 1 MiB: nothing fits


def test_plan_prefers_parts_under_a_tight_limit(libs):
    """2^22 points in 96 MiB: the points alone of one part holding all of them
    (256 MiB) do not fit, so the plan streams parts."""
    cfg = plan_msm(1 << 22, 1, 96 << 20, libs=libs)
    assert cfg is not None and cfg.parts > 1
    assert (cfg.stage_points, cfg.stage_scalars) == (2, 2)


def test_model_counts_resident_points_across_runs():
    """With one part the points cross the link once per call whatever the
    runs; four runs of one batch differ from one run of four only by the
    three scalar copies that a shared run hides behind the accumulation."""
    one = MsmConfig(window_size=16, target_window=16, batch_per_run=1)
    four = MsmConfig(window_size=16, target_window=16, batch_per_run=4)
    scalars_copy_us = (1 << 20) * 32 / (55.0 * 1e3)
    assert model_time_us(one, 1 << 20, 4) - model_time_us(four, 1 << 20, 4) == pytest.approx(3 * scalars_copy_us)
