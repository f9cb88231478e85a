"""Pins the oracle's BN254 G1 arithmetic.

(1) The reference's own KATs (core/src/common/curve/tests/bn254.cu:174-203):
p1 + p2 = pr, p1 + 100*p2 = pr100, 128*p1 = p1r128, points on curve.
(2) Independent affine chord-and-tangent arithmetic in Python big ints.
"""
import json
import os
import random

import pytest

import oracle as orc
from zkpoly_amd.field import FQ_MODULUS, FR_MODULUS, R_MONT

HERE = os.path.dirname(os.path.abspath(__file__))
PTS = {k: int(v, 16) for k, v in json.load(open(os.path.join(HERE, "golden", "curve_kats.json")))["points"].items()}
Q = FQ_MODULUS
RINV = pow(R_MONT, -1, Q)


def P(name):
    return orc.g1(PTS[name + "_x"], PTS[name + "_y"])


def same(a, b):
    return list(a) == list(b)


def py_add(p, q):
    """Affine addition over canonical coordinates (None = identity)."""
    if p is None:
        return q
    if q is None:
        return p
    (x1, y1), (x2, y2) = p, q
    if x1 == x2 and (y1 + y2) % Q == 0:
        return None
    if p == q:
        lam = 3 * x1 * x1 * pow(2 * y1, -1, Q) % Q
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, Q) % Q
    x3 = (lam * lam - x1 - x2) % Q
    return (x3, (lam * (x1 - x3) - y1) % Q)


def py_mul(p, k):
    acc = None
    for bit in bin(k)[2:]:
        acc = py_add(acc, acc)
        if bit == "1":
            acc = py_add(acc, p)
    return acc


def to_py(pt):
    x, y = orc.to_int(pt[:4]), orc.to_int(pt[4:])
    if x == 0 and y == 0:
        return None
    return (x * RINV % Q, y * RINV % Q)


def test_reference_kats():
    p1, p2 = P("p1"), P("p2")
    assert orc.g1_on_curve(p1) and orc.g1_on_curve(p2)
    assert same(orc.g1_add(p1, p2), P("pr"))
    assert same(orc.g1_add(p1, orc.g1_mul_canonical(p2, 100)), P("pr100"))
    assert same(orc.g1_mul_canonical(p1, 128), P("p1r128"))


def test_kats_match_closed_form():
    p1, p2 = to_py(P("p1")), to_py(P("p2"))
    assert to_py(P("pr")) == py_add(p1, p2)
    assert to_py(P("pr100")) == py_add(p1, py_mul(p2, 100))
    assert to_py(P("p1r128")) == py_mul(p1, 128)


def test_generator_and_group_order():
    g = orc.generator()
    assert orc.g1_on_curve(g)
    # r * G = identity (0, 0); (r - 1) G = -G
    assert same(orc.g1_mul_canonical(g, FR_MODULUS), [0] * 8)
    neg = orc.g1_mul_canonical(g, FR_MODULUS - 1)
    assert to_py(neg) == (1, Q - 2)


def test_random_multiples_closed_form():
    rng = random.Random(99)
    g = orc.generator()
    gp = to_py(g)
    for _ in range(6):
        k = rng.randrange(1, 1 << 40)
        assert to_py(orc.g1_mul_canonical(g, k)) == py_mul(gp, k)
    # doubling, adding a point to itself and to its negation
    a = orc.g1_mul_canonical(g, 12345)
    assert same(orc.g1_double(a), orc.g1_add(a, a))
    na = orc.g1_mul_canonical(g, FR_MODULUS - 12345)
    assert same(orc.g1_add(a, na), [0] * 8)
    assert same(orc.g1_add(a, [0] * 8), a)
