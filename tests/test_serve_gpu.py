"""Serving form: continuous batching of NTT / MSM requests admitted by HBM
footprint.  Every result is bit-exact against the oracle and identical to the
same request served alone; with a tight memory budget the requests that are
in flight together never exceed it.
"""
import os

import numpy as np
import pytest

import oracle as orc
from zkpoly_amd import hip
from zkpoly_amd.field import from_mont
from zkpoly_amd.serve import Resident, Server, ServeRequest, Trace, latency_stats, make_resident

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def resident(libs, stream):
    return make_resident(libs, stream, ntt_min=10, ntt_max=15, msm_max=14, seed=3)


def expected_for(res: Resident, trace: Trace):
    want = {}
    for key in trace.inputs:
        kind, log_n, _ = key
        x = trace.input(key)
        if kind == 0:
            want[key] = orc.ntt(x, log_n, min(16, os.cpu_count() or 1))
        else:
            n = 1 << log_n
            total = from_mont(orc.to_int(orc.inner_product(res.base_scalars[:n], x)))
            want[key] = orc.g1_mul_canonical(orc.generator(), total)
    return want


def check_all(trace, reqs, want):
    for i, (j, r) in enumerate(zip(trace.jobs, reqs)):
        assert r.status == 0
        got = trace.output(i)
        assert np.array_equal(got, want[(j.kind, j.log_n, j.input_id)]), (i, j)


def test_mix_equals_alone_and_oracle(libs, resident):
    trace = Trace(rate=3000.0, count=48, seed=1, ntt_range=(10, 15), msm_range=(10, 14))
    want = expected_for(resident, trace)
    srv = Server(libs, resident, slots=4, hbm_budget=1 << 30, msm_window=12)
    reqs = trace.requests()
    srv.run(reqs)
    check_all(trace, reqs, want)
    # overlapped: at some instant more than one request was in flight
    ev = sorted([(r.start_s, 1) for r in reqs] + [(r.finish_s, -1) for r in reqs])
    live, peak = 0, 0
    for _, d in ev:
        live += d
        peak = max(peak, live)
    assert peak > 1
    # each request alone (one slot, no other work) gives the same bytes
    alone = Server(libs, resident, slots=1, hbm_budget=1 << 30, msm_window=12)
    for i in range(len(trace.jobs)):
        r = trace.requests()[i]
        r.arrival_s = 0.0
        before = trace.output(i).copy()
        trace.outputs[i].view()[:] = 0
        alone.run([r])
        assert np.array_equal(trace.output(i), before), i


@pytest.fixture(scope="module")
def resident_full(libs, stream):
    """The serving workload's resident state: twiddles for NTT 2^14..2^22, one 2^21 base table."""
    return make_resident(libs, stream, ntt_min=14, ntt_max=22, msm_max=21, seed=55)


def test_workload_sizes_mix_equals_alone_and_oracle(libs, resident_full):
    """Config 5 at its sizes: a Poisson mix of NTT 2^14..2^22 and MSM 2^16..2^21
    (70 / 30) through four slots; every result equals the oracle, and each
    request served alone afterwards gives the same bytes."""
    trace = Trace(rate=1000.0, count=32, seed=1, variants=1, ntt_range=(14, 22), msm_range=(16, 21))
    print("serve mix:", [(j.kind, j.log_n) for j in trace.jobs], flush=True)
    want = expected_for(resident_full, trace)
    srv = Server(libs, resident_full, slots=4, hbm_budget=16 << 30, msm_window=16)
    reqs = trace.requests()
    srv.run(reqs)
    check_all(trace, reqs, want)
    ev = sorted([(r.start_s, 1) for r in reqs] + [(r.finish_s, -1) for r in reqs])
    live, peak = 0, 0
    for _, d in ev:
        live += d
        peak = max(peak, live)
    assert peak > 1
    alone = Server(libs, resident_full, slots=1, hbm_budget=16 << 30, msm_window=16)
    for i in range(len(trace.jobs)):
        r = trace.requests()[i]
        r.arrival_s = 0.0
        trace.outputs[i].view()[:] = 0
        alone.run([r])
        assert r.status == 0
        j = trace.jobs[i]
        assert np.array_equal(trace.output(i), want[(j.kind, j.log_n, j.input_id)]), i


def test_admission_respects_hbm_budget(libs, resident):
    trace = Trace(rate=1e6, count=24, seed=2, ntt_range=(13, 15), msm_range=(13, 14))
    probe = Server(libs, resident, slots=1, hbm_budget=64 << 20, msm_window=12)
    fp = [probe.footprint(j.kind, j.log_n) for j in trace.jobs]
    budget = max(fp) + min(fp)  # room for the largest plus at most a small one
    srv = Server(libs, resident, slots=6, hbm_budget=budget, msm_window=12)
    reqs = trace.requests()
    srv.run(reqs)
    check_all(trace, reqs, expected_for(resident, trace))
    points = sorted({r.start_s for r in reqs})
    for t in points:
        live = sum(f for f, r in zip(fp, reqs) if r.start_s <= t < r.finish_s)
        assert live <= budget + 256 * len(reqs)


def test_rejects_requests_out_of_range(libs, resident):
    srv = Server(libs, resident, slots=1, hbm_budget=1 << 26, msm_window=12)
    out = hip.PinnedBuffer(1 << 20)
    inp = hip.PinnedBuffer(1 << 20)
    bad = [ServeRequest(0, 20, 0.0, inp.addr, out.addr, 0, 0, 0)]  # NTT size outside 10..15
    with pytest.raises(hip.HipError):
        srv.run(bad)
    bad = [ServeRequest(1, 15, 0.0, inp.addr, out.addr, 0, 0, 0)]  # more points than resident bases
    with pytest.raises(hip.HipError):
        srv.run(bad)


def test_latency_stats_shape(libs, resident):
    trace = Trace(rate=500.0, count=10, seed=3, ntt_range=(10, 12), msm_range=(10, 12))
    srv = Server(libs, resident, slots=2, hbm_budget=1 << 28, msm_window=12)
    reqs = trace.requests()
    srv.run(reqs)
    st = latency_stats(reqs, trace.jobs)
    assert st["requests"] == 10 and st["p99_ms"] >= st["p50_ms"] > 0
