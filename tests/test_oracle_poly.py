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


def _perm(inputs, table, usable=None):
    usable = len(inputs) if usable is None else usable
    a, s = orc.permute(mont_array(inputs), mont_array(table), usable)
    return vals(a), vals(s)


def test_permute_worked_example():
    """Worked by hand: runs of A' = [1, 2,2,2, 3,3] start at rows 0, 1, 4 and take
    1, 2, 3 out of the table {1,2,2,3,4,5}; leftovers 2, 4, 5 fill rows 2, 3, 5."""
    a, s = _perm([3, 1, 3, 2, 2, 2], [1, 2, 3, 4, 5, 2])
    assert a == [1, 2, 2, 2, 3, 3]
    assert s == [1, 2, 2, 4, 3, 5]


def _check_lookup_permutation(inp, tab, a, s):
    """What halo2's lookup argument requires of (A', S'), plus the leftover order
    that makes the pair unique: A' is the input sorted; S' a permutation of the
    table; each row has A'_i = S'_i or A'_i = A'_{i-1}; run starts hold their
    own value; the other rows hold the leftovers in ascending order."""
    n = len(a)
    assert a == sorted(inp)
    assert sorted(s) == sorted(tab)
    starts = [i == 0 or a[i] != a[i - 1] for i in range(n)]
    for i in range(n):
        assert a[i] == s[i] or (i > 0 and a[i] == a[i - 1])
        if starts[i]:
            assert s[i] == a[i]
    rest = [s[i] for i in range(n) if not starts[i]]
    assert rest == sorted(rest)


@pytest.mark.parametrize("n,spread,seed", [(1, 1, 1), (50, 8, 2), (300, 40, 3), (1000, 1000, 4), (257, 2, 5)])
def test_permute_properties_small_values(n, spread, seed):
    rng = np.random.default_rng(seed)
    tab = [int(x) for x in rng.integers(0, spread, n)]
    distinct = sorted(set(tab))
    inp = [distinct[int(i)] for i in rng.integers(0, len(distinct), n)]
    a, s = _perm(inp, tab)
    _check_lookup_permutation(inp, tab, a, s)


def test_permute_properties_field_elements_and_usable_prefix():
    """Values spread over the whole field (negatives p - x share their top bits),
    and only the first `usable` rows count."""
    rng = np.random.default_rng(9)
    n, usable = 600, 575
    pool = [int(x) for x in rng.integers(0, 1 << 62, 40)] + [P - int(x) for x in rng.integers(1, 1 << 20, 40)]
    tab = [pool[int(i)] for i in rng.integers(0, len(pool), n)]
    head = sorted(set(tab[:usable]))
    inp = [head[int(i)] for i in rng.integers(0, len(head), n)]
    a, s = _perm(inp, tab, usable)
    assert len(a) == len(s) == usable
    _check_lookup_permutation(inp[:usable], tab[:usable], a, s)


def test_permute_value_missing_from_table():
    """A missing value keeps its own row and takes nothing from the table; the
    repeated rows take the smallest leftovers."""
    a, s = _perm([7, 7, 1], [1, 2, 3])
    assert a == [1, 7, 7]
    assert s == [1, 7, 2]
