"""Pins the oracle's best_fft and best_multiexp restatements.

NTT: O(n^2) DFT with a true root (n <= 2^7), INTT(NTT(x)) = x at the
config-1 size 2^16, NTT(delta_0) = ones, NTT(c) = (n c, 0, ...), linearity,
and gen_pq_omegas / twiddles against Python big-int powers.
MSM: brute-force double-and-add for small n, the closed form
sum_i s_i (k_i G) = (sum_i k_i s_i) G, all-zero scalars -> identity (0, 0),
all-one scalars -> sum of the bases.  A plausible slip (wrong window
extraction, dropped bucket, swapped sign) breaks one of these.
"""
import numpy as np
import pytest

import oracle as orc
from zkpoly_amd.field import (
    FR_MODULUS,
    R_MONT,
    array_to_ints,
    from_mont,
    ints_to_array,
    omega,
    random_elements,
    to_mont,
)


def mont_arr(xs):
    return ints_to_array([to_mont(x) for x in xs])


@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 7])
def test_best_fft_equals_naive_dft(log_n):
    a = random_elements(1 << log_n, 100 + log_n)
    w = orc.omega(log_n)
    assert orc.to_int(w) == to_mont(omega(log_n))
    assert np.array_equal(orc.best_fft(a, w, log_n), orc.naive_dft(a, w))


def test_naive_dft_closed_form():
    # the naive DFT itself against Python big ints
    n, log_n = 8, 3
    a = random_elements(n, 5)
    xs = [from_mont(x) for x in array_to_ints(a)]
    w = omega(log_n)
    want = [sum(xs[i] * pow(w, i * j, FR_MODULUS) for i in range(n)) % FR_MODULUS for j in range(n)]
    got = [from_mont(x) for x in array_to_ints(orc.naive_dft(a, orc.omega(log_n)))]
    assert got == want


def test_round_trip_config1():
    log_n = 16
    a = random_elements(1 << log_n, 2016)
    y = orc.ntt(a, log_n)
    assert not np.array_equal(y, a)
    assert np.array_equal(orc.intt(y, log_n), a)


def test_delta_and_constant():
    log_n = 10
    n = 1 << log_n
    delta = mont_arr([1] + [0] * (n - 1))
    assert np.array_equal(orc.ntt(delta, log_n), mont_arr([1] * n))
    c = 123456789
    const = mont_arr([c] * n)
    assert array_to_ints(orc.ntt(const, log_n)) == array_to_ints(mont_arr([n * c % FR_MODULUS] + [0] * (n - 1)))


def test_linearity():
    log_n = 9
    a = random_elements(1 << log_n, 1)
    b = random_elements(1 << log_n, 2)
    s = orc.vec("orc_vec_add", orc.FR, a, b)
    lhs = orc.ntt(s, log_n)
    rhs = orc.vec("orc_vec_add", orc.FR, orc.ntt(a, log_n), orc.ntt(b, log_n))
    assert np.array_equal(lhs, rhs)


def test_coset_is_evaluation_on_shifted_domain():
    # NTT(x_i g^i)_j = sum_i x_i (g w^j)^i : check a few outputs by Horner in big ints
    log_n = 6
    n = 1 << log_n
    a = random_elements(n, 11)
    g = 7
    powers = mont_arr([pow(g, i, FR_MODULUS) for i in range(n)])
    y = orc.ntt(orc.distribute_powers(a, powers), log_n)
    xs = [from_mont(x) for x in array_to_ints(a)]
    w = omega(log_n)
    for j in (0, 1, 17, n - 1):
        pt = g * pow(w, j, FR_MODULUS) % FR_MODULUS
        acc = 0
        for c in reversed(xs):
            acc = (acc * pt + c) % FR_MODULUS
        assert from_mont(array_to_ints(y[j : j + 1])[0]) == acc


def test_distribute_cyclic_table_closed_form():
    """A table shorter than the polynomial is applied cyclically: coefficient
    i is multiplied by powers[i % len] (the [1, zeta, zeta^2] case of
    core/tests/ntt.rs:261-388); the NTT of the result evaluates
    sum_i x_i zeta^(i mod 3) w^(i j), checked by big ints."""
    log_n = 6
    n = 1 << log_n
    a = random_elements(n, 12)
    zeta = pow(7, (FR_MODULUS - 1) // 3, FR_MODULUS)
    table = [1, zeta, zeta * zeta % FR_MODULUS]
    d = orc.distribute_powers(a, mont_arr(table))
    xs = [from_mont(x) for x in array_to_ints(a)]
    assert [from_mont(x) for x in array_to_ints(d)] == [x * table[i % 3] % FR_MODULUS for i, x in enumerate(xs)]
    y = orc.ntt(d, log_n)
    w = omega(log_n)
    for j in (0, 1, 5, n - 1):
        wj = pow(w, j, FR_MODULUS)
        want = sum(x * table[i % 3] * pow(wj, i, FR_MODULUS) for i, x in enumerate(xs)) % FR_MODULUS
        assert from_mont(array_to_ints(y[j : j + 1])[0]) == want


def test_twiddles_and_pq_omegas():
    w = omega(12)
    tw = orc.twiddles(64, to_mont(w))
    assert [from_mont(x) for x in array_to_ints(tw)] == [pow(w, i, FR_MODULUS) for i in range(64)]
    pq, om = orc.gen_pq_omegas(5, 1 << 12, to_mont(w))
    step = pow(w, (1 << 12) >> 5, FR_MODULUS)
    assert [from_mont(x) for x in array_to_ints(pq)] == [pow(step, i, FR_MODULUS) for i in range(16)]
    assert [from_mont(x) for x in array_to_ints(om)] == [pow(w, 1 << i, FR_MODULUS) for i in range(32)]


# ---------------- MSM ----------------


def bases_from_k(ks):
    g = orc.generator()
    return np.stack([orc.g1_mul_canonical(g, k) for k in ks])


@pytest.mark.parametrize("n", [1, 3, 5, 31, 40])
def test_multiexp_equals_naive(n):
    rng = np.random.default_rng(n)
    ks = [int(x) for x in rng.integers(1, 1 << 62, size=n)]
    bases = bases_from_k(ks)
    s = random_elements(n, 300 + n)
    assert list(orc.best_multiexp(s, bases, threads=4)) == list(orc.naive_msm(s, bases))


def test_multiexp_closed_form():
    n = 1 << 10
    rng = np.random.default_rng(7)
    ks = [int(x) for x in rng.integers(1, 1 << 63, size=n)]
    bases = bases_from_k(ks)
    s = random_elements(n, 77)
    ss = [from_mont(x) for x in array_to_ints(s)]
    total = sum(k * x for k, x in zip(ks, ss)) % FR_MODULUS
    want = orc.g1_mul_canonical(orc.generator(), total)
    for threads in (1, 3, 8):
        assert list(orc.best_multiexp(s, bases, threads=threads)) == list(want)


def test_multiexp_zero_and_one_scalars():
    n = 100
    ks = list(range(1, n + 1))
    bases = bases_from_k(ks)
    zeros = np.zeros((n, 4), dtype=np.uint64)
    assert list(orc.best_multiexp(zeros, bases)) == [0] * 8
    ones = mont_arr([1] * n)
    want = orc.g1_mul_canonical(orc.generator(), n * (n + 1) // 2)
    assert list(orc.best_multiexp(ones, bases)) == list(want)
