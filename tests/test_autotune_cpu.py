"""Autotuner selection and persistence with an injected timer (no GPU)."""
from zkpoly_amd import autotune as at


def test_pick_is_argmin_with_first_listed_on_ties():
    assert at.pick({1: 3.0, 2: 2.0, 3: 2.5}) == 2
    assert at.pick({3: 1.0, 1: 1.0}) == 3


def test_tune_keeps_argmin_per_size_and_all_timings():
    table = {(20, 1): 0.30, (20, 2): 0.25, (20, 3): 0.27, (24, 1): 3.4, (24, 2): 3.6, (24, 3): 3.7}
    t = at.tune("ntt", [20, 24], [1, 2, 3], lambda k, c: table[(k, c)])
    assert t.ntt == {20: 2, 24: 1}
    assert t.timings["ntt/24/1"] == 3.4 and len(t.timings) == 6


def test_tune_keeps_each_candidates_best_round():
    """A candidate measured slowly once (clock ramp-up) still wins on its best sweep."""
    seq = iter([9.0, 3.3, 3.4, 3.1, 3.3, 3.4])  # first sweep: c1 slow; second sweep: c1 fastest
    t = at.tune("ntt", [24], [1, 2, 3], lambda k, c: next(seq))
    assert t.ntt == {24: 1} and t.timings["ntt/24/1"] == 3.1
    m = at.tune("msm", [20], [14, 16], lambda k, w: {14: 7.0, 16: 6.5}[w])
    assert m.msm_window(20) == 16 and m.msm_window(18, default=13) == 13


def test_json_round_trip_and_merge(tmp_path):
    t = at.Tuning(ntt={24: 1, 20: 2}, msm={20: 16}, timings={"ntt/24/1": 3.4}, device="dev")
    p = str(tmp_path / "sub" / "tuning.json")
    t.save(p)
    back = at.load(p)
    assert back == t
    merged = back.merge(at.Tuning(ntt={24: 3}, device=""))
    assert merged.ntt == {24: 3, 20: 2} and merged.device == "dev"
    assert at.load(str(tmp_path / "missing.json")) is None


def test_apply_installs_choices_in_libntt():
    from zkpoly_amd.ntt import kernel_count, set_tuning
    from zkpoly_amd.runtime import Libs

    libs = Libs()
    try:
        at.Tuning(ntt={24: 1}).apply(libs)
        assert kernel_count(libs, 24) == 3  # fused head + 10 + 9 stages
        at.Tuning(ntt={24: 2}).apply(libs)
        assert kernel_count(libs, 24) == 4  # fused head + 7 + 6 + 6 stages
    finally:
        set_tuning(libs, 24, at.DEFAULT_MIN_COL_LOG)
