"""Pins the oracle's polynomial helpers with Python big ints.

Every check is a closed form or an identity that a plausible slip (a dropped
term, a wrong sign, an off-by-one index) breaks: eval against sum a_i x^i and
its x = 0 / x = 1 special cases; kate_division against q(X)(X - b) + p(b) =
p(X) coefficient by coefficient; batch_invert against x * x^-1 = 1 with zeros
kept; scan_mul against running products.  Values cross the boundary in
Montgomery form (x * 2^256 mod r), converted with from_mont / to_mont.
"""
import numpy as np
import pytest

import oracle as orc
from zkpoly_amd.field import FR_MODULUS, array_to_ints, from_mont, ints_to_array, mont_limbs, random_elements, to_mont

P = FR_MODULUS


def vals(arr):
    return [from_mont(x) for x in array_to_ints(arr)]


def mont_array(xs):
    return ints_to_array([to_mont(x % P) for x in xs])


@pytest.mark.parametrize("n", [1, 2, 7, 64])
def test_poly_eval_closed_form(n):
    a = random_elements(n, 100 + n)
    x = random_elements(1, 200 + n)[0]
    av, xv = vals(a), vals(x.reshape(1, 4))[0]
    want = sum(c * pow(xv, i, P) for i, c in enumerate(av)) % P
    assert from_mont(orc.to_int(orc.poly_eval(a, x))) == want
    # x = 0 gives a_0, x = 1 the coefficient sum
    assert from_mont(orc.to_int(orc.poly_eval(a, mont_limbs(0)))) == av[0]
    assert from_mont(orc.to_int(orc.poly_eval(a, mont_limbs(1)))) == sum(av) % P


@pytest.mark.parametrize("n", [2, 3, 16, 64])
def test_kate_division_is_the_quotient(n):
    p = random_elements(n, 300 + n)
    b = random_elements(1, 400 + n)[0]
    pv, bv = vals(p), vals(b.reshape(1, 4))[0]
    q = vals(orc.kate_division(p, b))
    assert q[n - 1] == 0
    qd = q[: n - 1]
    # q(X) (X - b) + p(b), coefficient by coefficient, equals p(X)
    pb = sum(c * pow(bv, i, P) for i, c in enumerate(pv)) % P
    prod = [0] * n
    for i, c in enumerate(qd):
        prod[i + 1] = (prod[i + 1] + c) % P
        prod[i] = (prod[i] - bv * c) % P
    prod[0] = (prod[0] + pb) % P
    assert prod == pv


def test_kate_division_at_zero_shifts():
    p = random_elements(32, 500)
    q = vals(orc.kate_division(p, mont_limbs(0)))
    assert q == vals(p)[1:] + [0]


def test_batch_invert_with_zeros():
    n = 120
    x = random_elements(n, 600)
    x[::17] = 0
    inv = vals(orc.batch_invert(x))
    for xi, yi in zip(vals(x), inv):
        assert (yi == 0) if xi == 0 else (xi * yi % P == 1)


def test_scan_mul_running_product():
    x = random_elements(50, 700)
    got = vals(orc.scan_mul(x))
    acc = 1
    for xi, gi in zip(vals(x), got):
        acc = acc * xi % P
        assert gi == acc


def test_mont_array_round_trip():
    xs = [0, 1, 5, P - 1]
    assert vals(mont_array(xs)) == xs
