"""GPU parity of libntt.so against the oracle (bit-exact, canonical Montgomery).

Mirrors core/tests/ntt.rs: test_ssip_ntt (:26-148: random omega that is not a
root of unity, GPU poly rotated by 3 and sliced out of a 2^(k+1) buffer, k up
to 20), test_distribute_zeta (:261-388: [1, zeta, zeta^2] distribution on a
zero-padded 4x extended domain), plus the workload sizes 2^20 / 2^22 / 2^24
(forward, inverse, coset g = 7) and the 4x extended coset of 2^22.
"""
import os

import numpy as np
import pytest

import oracle as orc
from helpers import cpu2gpu, device_array, download, gpu2cpu, upload
from zkpoly_amd import hip
from zkpoly_amd.field import FR_MODULUS, mont_limbs, omega, random_elements, to_mont
from zkpoly_amd.ntt import (DistributePowers, GenPqOmegas, RecomputeNtt, SsipNtt, SsipPrecompute, gen_powers,
                            kernel_count, set_tuning, set_variant)
from zkpoly_amd.runtime import ScalarArray

pytestmark = pytest.mark.gpu
THREADS = os.cpu_count() or 1


def twiddle_table(libs, w_mont, n):
    """Host twiddles through ssip_precompute, then uploaded (as the compiler does)."""
    host = np.zeros((n // 2, 4), dtype=np.uint64)
    arr = ScalarArray(n // 2, host.ctypes.data, device="cpu")
    SsipPrecompute(libs).get_fn()(arr, w_mont)
    return host


def gpu_ntt(libs, stream, data, w_mont, rotate=0, sliced=False):
    n = data.shape[0]
    tw_host = twiddle_table(libs, w_mont, n)
    tw_buf, tw = upload(tw_host)
    if sliced:
        buf, whole = device_array(2 * n)
        whole.rotate(rotate)
        x = whole.slice(n, 2 * n)
    else:
        buf, x = device_array(n)
        if rotate:
            x.rotate(rotate)
    cpu2gpu(data, x, buf)
    SsipNtt(libs).get_fn().f([x], [tw, stream], lambda i: i)
    stream.sync()
    return gpu2cpu(x, buf), tw_host


def test_ssip_precompute_matches_oracle(libs):
    for k in (1, 2, 5, 11, 16):
        w = orc.omega(k)
        got = twiddle_table(libs, w, 1 << k)
        assert np.array_equal(got, orc.twiddles(1 << (k - 1), w))


@pytest.mark.parametrize("k", list(range(1, 21)))
def test_ssip_ntt_random_omega_rotated_slice(libs, stream, k):
    """core/tests/ntt.rs:26-148 with a seeded (not OS) RNG."""
    n = 1 << k
    data = random_elements(n, 1000 + k)
    w = random_elements(1, 5000 + k)[0]  # almost surely not a root of unity
    got, _ = gpu_ntt(libs, stream, data, w, rotate=3, sliced=True)
    want = orc.best_fft(data, w, k, THREADS)
    assert np.array_equal(got, want), f"k={k}"


DEFAULT_MIN_COL_LOG = 1


@pytest.mark.parametrize("variant,min_col", [(0, 1), (0, 3), (0, 6), (1, 2), (2, 2), (3, 1), (3, 3)])
def test_variants_and_tunings(libs, stream, variant, min_col):
    """Every kernel plan gives the same best_fft network: the legacy radix-2
    plan and the fused-head radix-4 plan at several pass depths."""
    ks = (1, 2, 3, 5, 9, 12, 15, 18)
    old = set_variant(libs, variant)
    try:
        for k in ks:
            set_tuning(libs, k, min_col)
            n = 1 << k
            data = random_elements(n, 300 + k)
            w = random_elements(1, 900 + k)[0]
            got, _ = gpu_ntt(libs, stream, data, w, rotate=5 % n, sliced=True)
            want = orc.best_fft(data, w, k, THREADS)
            assert np.array_equal(got, want), f"variant={variant} min_col={min_col} k={k}"
    finally:
        set_variant(libs, old)
        for k in ks:
            set_tuning(libs, k, DEFAULT_MIN_COL_LOG)


@pytest.mark.parametrize("k", [21, 23])
def test_head_variants_at_large_sizes(libs, stream, k):
    """From 2^21 on a head workgroup walks several tile pairs (loads of the next
    pair in flight during the current stages); the walk, the one-pair head and
    the radix-2 plan all give best_fft's output for a random omega on a rotated
    slice."""
    n = 1 << k
    data = random_elements(n, 4100 + k)
    w = random_elements(1, 4200 + k)[0]
    want = orc.best_fft(data, w, k, THREADS)
    for variant in (0, 2, 1, 3):
        old = set_variant(libs, variant)
        try:
            got, _ = gpu_ntt(libs, stream, data, w, rotate=(n >> 3) + 7, sliced=True)
        finally:
            set_variant(libs, old)
        assert np.array_equal(got, want), f"variant={variant} k={k}"


def test_kernel_counts(libs):
    assert kernel_count(libs, 0) == 0
    assert kernel_count(libs, 24) == 3  # fused head + 10 + 9 stages (default min_col_log 1)
    set_tuning(libs, 24, 2)
    try:
        assert kernel_count(libs, 24) == 4  # fused head + 7 + 6 + 6 stages
    finally:
        set_tuning(libs, 24, DEFAULT_MIN_COL_LOG)
    old = set_variant(libs, 1)
    try:
        assert kernel_count(libs, 24) == 4  # bit reversal + 11 + 9 + 4 stages
    finally:
        set_variant(libs, old)


@pytest.mark.parametrize("k,rot", [(4, 2), (10, 1023), (13, 5)])
def test_ssip_ntt_rotated_true_root(libs, stream, k, rot):
    n = 1 << k
    data = random_elements(n, 77 + k)
    got, _ = gpu_ntt(libs, stream, data, orc.omega(k), rotate=rot)
    assert np.array_equal(got, orc.ntt(data, k, THREADS))


def test_round_trip_2_16(libs, stream):
    """Config 1: NTT then INTT (omega^-1 table, then n^-1 through distribute_powers)."""
    k, n = 16, 1 << 16
    data = random_elements(n, 2016)
    buf, x = upload(data)
    tw_f_buf, tw_f = upload(twiddle_table(libs, orc.omega(k), n))
    tw_i_buf, tw_i = upload(twiddle_table(libs, orc.omega(k, True), n))
    ninv_buf, ninv = upload(mont_limbs(pow(n, -1, FR_MODULUS)))
    ntt, dist = SsipNtt(libs).get_fn().f, DistributePowers(libs).get_fn().f
    ntt([x], [tw_f, stream])
    stream.sync()
    fwd = download(x, buf)
    assert np.array_equal(fwd, orc.ntt(data, k, THREADS))
    ntt([x], [tw_i, stream])
    dist([x], [ninv, stream])
    stream.sync()
    assert np.array_equal(download(x, buf), data)


def pq_omegas(libs, k, w_mont, pq_len=None):
    """Host pq / omega ladder through gen_pq_omegas, uploaded (as the compiler does)."""
    pq_len = pq_len or GenPqOmegas.get_pq_len(k)
    pq = np.zeros((pq_len, 4), dtype=np.uint64)
    om = np.zeros((32, 4), dtype=np.uint64)
    GenPqOmegas(libs).get_fn()(ScalarArray(pq_len, pq.ctypes.data, device="cpu"),
                               ScalarArray(32, om.ctypes.data, device="cpu"), 1 << k, w_mont)
    pq_buf, pq_arr = upload(pq)
    om_buf, om_arr = upload(om)
    return (pq_buf, pq_arr), (om_buf, om_arr)


@pytest.mark.parametrize("k", list(range(1, 21)))
def test_recompute_ntt_random_omega_rotated_slice(libs, stream, k):
    """core/tests/ntt.rs:151-259 (random omega, pq sized by get_pq_len) on a
    rotated slice of a 2^(k+1) buffer."""
    n = 1 << k
    data = random_elements(n, 2000 + k)
    w = random_elements(1, 6000 + k)[0]
    (pqb, pq), (omb, om) = pq_omegas(libs, k, w)
    buf, whole = device_array(2 * n)
    whole.rotate(3)
    x = whole.slice(n, 2 * n)
    cpu2gpu(data, x, buf)
    RecomputeNtt(libs).get_fn().f([x], [pq, om, stream], lambda i: i)
    stream.sync()
    assert np.array_equal(gpu2cpu(x, buf), orc.best_fft(data, w, k, THREADS)), f"k={k}"


@pytest.mark.parametrize("pq_len", [1, 2, 256])
def test_recompute_ntt_any_pq_size(libs, stream, pq_len):
    """The table size is the caller's choice: coarse pq (ladder-built small
    roots) and fine pq give the same transform."""
    k = 17
    data = random_elements(1 << k, 71)
    w = random_elements(1, 72)[0]
    (pqb, pq), (omb, om) = pq_omegas(libs, k, w, pq_len)
    buf, x = device_array(1 << k)
    cpu2gpu(data, x, buf)
    RecomputeNtt(libs).get_fn().f([x], [pq, om, stream])
    stream.sync()
    assert np.array_equal(gpu2cpu(x, buf), orc.best_fft(data, w, k, THREADS))


@pytest.mark.parametrize("k", [22, 24])
def test_recompute_matches_ssip_at_workload_sizes(libs, stream, k):
    """At 2^22 / 2^24 the recompute transform equals the table transform (which
    the workload-size test pins to the oracle) byte for byte."""
    n = 1 << k
    data = random_elements(n, 8800 + k)
    w = orc.omega(k)
    (pqb, pq), (omb, om) = pq_omegas(libs, k, w)
    got_r = _run_on_device(data, lambda x: RecomputeNtt(libs).get_fn().f([x], [pq, om, stream]), stream)
    tw_buf, tw = device_array(n // 2)
    base_buf, _ = upload(w)
    gen_powers(libs, tw.values, n // 2, base_buf.addr, stream)
    got_s = _run_on_device(data, lambda x: SsipNtt(libs).get_fn().f([x], [tw, stream]), stream)
    assert np.array_equal(got_r, got_s)
    # size-independent checks: linearity against the delta input (all ones)
    delta = np.zeros((n, 4), dtype=np.uint64)
    delta[0] = mont_limbs(1)
    ones = _run_on_device(delta, lambda x: RecomputeNtt(libs).get_fn().f([x], [pq, om, stream]), stream)
    assert np.all(ones == mont_limbs(1))


def _run_on_device(data, fn, stream):
    buf, x = device_array(data.shape[0])
    cpu2gpu(data, x, buf)
    fn(x)
    stream.sync()
    return gpu2cpu(x, buf)


def test_distribute_zeta_extended_recompute(libs, stream):
    """core/tests/ntt.rs:261-388 literally: the reference runs this one through
    RecomputeNtt with pq / omegas of the extended domain."""
    k = 12
    ek = k + 2
    data = random_elements(1 << k, 32)
    padded = np.zeros((1 << ek, 4), dtype=np.uint64)
    padded[: 1 << k] = data
    zeta = pow(7, (FR_MODULUS - 1) // 3, FR_MODULUS)
    zetas = np.stack([mont_limbs(1), mont_limbs(zeta), mont_limbs(zeta * zeta % FR_MODULUS)])
    (pqb, pq), (omb, om) = pq_omegas(libs, ek, orc.omega(ek))
    buf, x = device_array(1 << ek)
    cpu2gpu(padded, x, buf)
    z_buf, z = upload(zetas)
    DistributePowers(libs).get_fn().f([x], [z, stream])
    RecomputeNtt(libs).get_fn().f([x], [pq, om, stream])
    stream.sync()
    want = orc.ntt(orc.distribute_powers(padded, zetas), ek, THREADS)
    assert np.array_equal(gpu2cpu(x, buf), want)


def test_distribute_zeta_extended(libs, stream):
    """core/tests/ntt.rs:261-388: pad to 4n, multiply by [1, zeta, zeta^2] cyclically, NTT."""
    k = 12
    ek = k + 2
    data = random_elements(1 << k, 31)
    padded = np.zeros((1 << ek, 4), dtype=np.uint64)
    padded[: 1 << k] = data
    zeta = pow(7, (FR_MODULUS - 1) // 3, FR_MODULUS)
    zetas = np.stack([mont_limbs(1), mont_limbs(zeta), mont_limbs(zeta * zeta % FR_MODULUS)])
    buf, x = device_array(1 << ek)
    x.rotate(-2)
    cpu2gpu(padded, x, buf)
    z_buf, z = upload(zetas)
    tw_buf, tw = upload(twiddle_table(libs, orc.omega(ek), 1 << ek))
    DistributePowers(libs).get_fn().f([x], [z, stream])
    SsipNtt(libs).get_fn().f([x], [tw, stream])
    stream.sync()
    want = orc.ntt(orc.distribute_powers(padded, zetas), ek, THREADS)
    assert np.array_equal(gpu2cpu(x, buf), want)


def _device_coset_powers(libs, stream, n, g):
    pw_buf, pw = device_array(n)
    base_buf, _ = upload(mont_limbs(g))
    gen_powers(libs, pw.values, n, base_buf.addr, stream)
    stream.sync()
    return pw_buf, pw


def test_gen_powers_matches_oracle(libs, stream):
    n = 5000
    pw_buf, pw = _device_coset_powers(libs, stream, n, 7)
    got = download(pw, pw_buf)
    assert np.array_equal(got, orc.twiddles(n, to_mont(7)))


@pytest.mark.parametrize("k", [20, 22, 24])
def test_workload_sizes_forward_inverse_coset(libs, stream, k):
    """Config 2 sizes: forward, inverse (with n^-1) and coset g = 7, bit-exact."""
    n = 1 << k
    data = random_elements(n, 2000 + k)
    buf, x = upload(data)
    ntt, dist = SsipNtt(libs).get_fn().f, DistributePowers(libs).get_fn().f
    tw_buf, tw = upload(twiddle_table(libs, orc.omega(k), n))
    ntt([x], [tw, stream])
    stream.sync()
    fwd = download(x, buf)
    assert np.array_equal(fwd, orc.ntt(data, k, THREADS)), "forward"
    del tw_buf
    twi_buf, twi = upload(twiddle_table(libs, orc.omega(k, True), n))
    ninv_buf, ninv = upload(mont_limbs(pow(n, -1, FR_MODULUS)))
    ntt([x], [twi, stream])
    dist([x], [ninv, stream])
    stream.sync()
    assert np.array_equal(download(x, buf), data), "inverse round trip"
    del twi_buf
    # coset: NTT(x_i g^i)
    pw_buf, pw = _device_coset_powers(libs, stream, n, 7)
    tw_buf, tw = upload(twiddle_table(libs, orc.omega(k), n))
    dist([x], [pw, stream])
    ntt([x], [tw, stream])
    stream.sync()
    want = orc.ntt(orc.distribute_powers(data, orc.twiddles(n, to_mont(7))), k, THREADS)
    assert np.array_equal(download(x, buf), want), "coset"


def test_extended_coset_4x_of_2_22(libs, stream):
    """Config 2: 2^22 coefficients zero-padded to 2^24, coset g = 7, NTT with omega_{2^24}."""
    k, ek = 22, 24
    data = random_elements(1 << k, 4422)
    padded = np.zeros((1 << ek, 4), dtype=np.uint64)
    padded[: 1 << k] = data
    buf, x = upload(padded)
    pw_buf, pw = _device_coset_powers(libs, stream, 1 << ek, 7)
    tw_buf, tw = upload(twiddle_table(libs, orc.omega(ek), 1 << ek))
    DistributePowers(libs).get_fn().f([x], [pw, stream])
    SsipNtt(libs).get_fn().f([x], [tw, stream])
    stream.sync()
    want = orc.ntt(orc.distribute_powers(padded, orc.twiddles(1 << ek, to_mont(7))), ek, THREADS)
    assert np.array_equal(download(x, buf), want)


def test_concurrent_transforms_while_tuning_changes(libs):
    """Four threads transform their own 2^18 inputs on their own streams, three
    times each, while a fifth thread keeps changing the pass depth of 2^18
    transforms (zkp_ntt_set_tuning).  Every plan computes the same network, and
    a launch reads the setting once, so every result equals the oracle."""
    import threading

    k = 18
    n = 1 << k
    tw_buf, tw = upload(twiddle_table(libs, orc.omega(k), n))
    inputs = [random_elements(n, 1800 + j) for j in range(4)]
    want = [orc.ntt(x, k, THREADS) for x in inputs]
    streams = [hip.Stream(0) for _ in range(4)]
    bufs = [upload(x) for x in inputs]
    ntt = SsipNtt(libs).get_fn()
    done = threading.Event()
    errors, results = [], [None] * 4

    def transform(j):
        try:
            buf, x = bufs[j]
            for rep in range(3):
                buf.upload(inputs[j])
                ntt.f([x], [tw, streams[j]])
                streams[j].sync()
                results[j] = download(x, buf)
                if not np.array_equal(results[j], want[j]):
                    errors.append((j, rep))
        except Exception as e:  # reported below, on the main thread
            errors.append(e)

    def retune():
        c = 1
        while not done.is_set():
            set_tuning(libs, k, c)
            c = c % 3 + 1

    tuner = threading.Thread(target=retune)
    workers = [threading.Thread(target=transform, args=(j,)) for j in range(4)]
    print("ntt concurrent transforms: 4 threads + tuner", flush=True)
    tuner.start()
    for t in workers:
        t.start()
    for t in workers:
        t.join()
    done.set()
    tuner.join()
    set_tuning(libs, k, 1)
    assert not errors, errors
