"""Batched proof-step form (config 4 shape): NTT + fused a*b + gamma*c - z over
instances of uneven size, as one hipGraph launch, bit-exact per instance
against the oracle and identical to the one-call-per-instance path.

The graph is replayed several times with eager work (input resets, host
synchronisation) between replays, and every replay is checked.
"""
import numpy as np
import pytest

import oracle as orc
from zkpoly_amd import fused
from zkpoly_amd.field import from_mont
from zkpoly_amd.batch import Config4Workload

pytestmark = pytest.mark.gpu
FR = orc.FR


def expected(w, i):
    x, a, b, c = w.host[i]
    k = w.log_lens[i]
    n = w.lens[i]
    g = np.broadcast_to(w.gamma_z[0], (n, 4)).copy()
    z = np.broadcast_to(w.gamma_z[1], (n, 4)).copy()
    out = orc.vec("orc_vec_sub", FR, orc.vec("orc_vec_add", FR, orc.vec("orc_vec_mul", FR, a, b),
                                             orc.vec("orc_vec_mul", FR, g, c)), z)
    return orc.ntt(x, k), out


@pytest.fixture(scope="module")
def kernel(libs):
    return fused.FusedKernel(libs, fused.FusedKernelMeta("fused_arith_abgcz", 5, 1, fused.DEMO_DIR))


def test_graph_replays_bit_exact(libs, stream, kernel):
    w = Config4Workload(libs, stream, kernel, count=24, lo=1, hi=16, seed=7)
    want = [expected(w, i) for i in range(len(w.lens))]
    graph = w.capture()
    for replay in range(3):
        w.reset_inputs()
        stream.sync()
        graph.launch(stream)
        stream.sync()
        for i in range(len(w.lens)):
            x, out = w.results(i)
            assert np.array_equal(x, want[i][0]), (replay, i, w.log_lens[i])
            assert np.array_equal(out, want[i][1]), (replay, i)


def test_graph_replays_legacy_variant(libs, stream, kernel):
    """The batch plan of the radix-2 kernels (variant 1) stays bit-exact too."""
    from zkpoly_amd.ntt import set_variant

    old = set_variant(libs, 1)
    try:
        w = Config4Workload(libs, stream, kernel, count=10, lo=1, hi=14, seed=9)
        graph = w.capture()
    finally:
        set_variant(libs, old)
    want = [expected(w, i) for i in range(len(w.lens))]
    for replay in range(2):
        w.reset_inputs()
        stream.sync()
        graph.launch(stream)
        stream.sync()
        for i in range(len(w.lens)):
            x, out = w.results(i)
            assert np.array_equal(x, want[i][0]), (replay, i, w.log_lens[i])
            assert np.array_equal(out, want[i][1]), (replay, i)


def msm_closed_form(w, j):
    n = w.msm_lens[j]
    total = from_mont(orc.to_int(orc.inner_product(w.msm_base_scalars[:n], w.msm_scalars[j])))
    return list(orc.g1_mul_canonical(orc.generator(), total))


@pytest.mark.parametrize("branches", [4, 0])
def test_graph_with_msms_replays_bit_exact(libs, stream, kernel, branches):
    """NTTs, the fused kernel and MSMs of uneven sizes (2^10..2^16 points, each
    its own window) in one graph, the MSMs on parallel branches beside the
    transforms (or all on the capturing stream); three replays with eager work
    between them, every NTT, fused output and MSM result checked, and the MSMs
    equal to the plain synchronous calls."""
    w = Config4Workload(libs, stream, kernel, count=10, lo=6, hi=15, seed=11, msm_count=6, msm_lo=10, msm_hi=16)
    print("msm sizes", w.msm_log_lens, "windows", [m.window for m in w.msms], "branches", branches, flush=True)
    want = [expected(w, i) for i in range(len(w.lens))]
    want_msm = [msm_closed_form(w, j) for j in range(len(w.msms))]
    graph = w.capture(branches)
    for replay in range(3):
        w.reset_inputs()
        for m in w.msms:
            m.sums.view()[:] = 0
        stream.sync()
        graph.launch(stream)
        stream.sync()
        for i in range(len(w.lens)):
            x, out = w.results(i)
            assert np.array_equal(x, want[i][0]), (replay, i, w.log_lens[i])
            assert np.array_equal(out, want[i][1]), (replay, i)
        got = graph.msm_results()
        for j in range(len(w.msms)):
            assert list(got[j]) == want_msm[j], (replay, j, w.msm_log_lens[j])
    plain = w.run_plain()
    stream.sync()
    assert [list(r) for r in plain] == want_msm


def test_plain_path_matches(libs, stream, kernel):
    w = Config4Workload(libs, stream, kernel, count=12, lo=10, hi=18, seed=8)
    w.run_plain()
    stream.sync()
    for i in range(len(w.lens)):
        x, out = w.results(i)
        e = expected(w, i)
        assert np.array_equal(x, e[0]) and np.array_equal(out, e[1]), i


def test_config4_full_sizes(libs, stream, kernel):
    """64 instances, log sizes 10..22 as in the workload, one graph launch."""
    w = Config4Workload(libs, stream, kernel, count=64, lo=10, hi=22, seed=4)
    graph = w.capture()
    graph.launch(stream)
    stream.sync()
    for i in range(len(w.lens)):
        x, out = w.results(i)
        e = expected(w, i)
        assert np.array_equal(x, e[0]), (i, w.log_lens[i])
        assert np.array_equal(out, e[1]), i
