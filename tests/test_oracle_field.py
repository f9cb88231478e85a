"""Pins the oracle's field arithmetic.

Checks against (1) the reference's own known-answer vectors
(core/src/common/mont/tests/main.cu:238-525, extracted by
tests/golden/make_kats.py) and (2) Python big-int closed forms that do not
share any code with the oracle (Montgomery(x) = x * 2^256 mod p).
"""
import json
import os
import random

import numpy as np
import pytest

import oracle as orc
from zkpoly_amd.field import FQ_MODULUS, FR_MODULUS, R_MONT

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "field_kats.json")))
R_INV_R = pow(R_MONT, -1, FR_MODULUS)


def I(x):
    return int(x, 16)


@pytest.mark.parametrize("inst", ["instance1", "instance2"])
def test_reference_kats(inst):
    k = KATS["instances"][inst]
    a, b = I(k["a"]), I(k["b"])
    # Montgomery product of the raw words (main.cu:298-303)
    assert orc.to_int(orc.mul(orc.FR, a, b)) == I(k["prod_mont"])
    # element add / sub of the raw words (main.cu:267-277)
    assert orc.to_int(orc.add(orc.FR, a, b)) == I(k["sum_mont"])
    assert orc.to_int(orc.sub(orc.FR, a, b)) == I(k["sub_mont"])
    ea = orc.from_canonical(orc.FR, a)
    # from_number(a) squared (main.cu:305-309)
    assert orc.to_int(orc.mul(orc.FR, ea, ea)) == I(k["a_square_mont"])
    # from_number(a)^b (main.cu:311-315)
    assert orc.to_int(orc.pow_(orc.FR, ea, b)) == I(k["pow_mont"])
    # from_number(a)^-1 (main.cu:317-321)
    assert orc.to_int(orc.inv(orc.FR, ea)) == I(k["a_inv_mont"])


def test_reference_kat_convert():
    c = KATS["convert"]
    x, xm = I(c["x"]), I(c["x_mont"])
    assert orc.to_int(orc.from_canonical(orc.FR, x)) == xm
    assert orc.to_int(orc.to_canonical(orc.FR, xm)) == x


def test_kats_match_closed_form():
    # the KATs themselves agree with the mathematics (guards the extraction)
    for inst in ("instance1", "instance2"):
        k = KATS["instances"][inst]
        a, b = I(k["a"]), I(k["b"])
        assert I(k["prod_mont"]) == a * b * R_INV_R % FR_MODULUS
        assert I(k["sum_mont"]) == (a + b) % FR_MODULUS
        assert I(k["sub_mont"]) == (a - b) % FR_MODULUS
        assert I(k["a_square_mont"]) == a * a * R_MONT % FR_MODULUS
        assert I(k["pow_mont"]) == pow(a, b, FR_MODULUS) * R_MONT % FR_MODULUS
        assert I(k["a_inv_mont"]) == pow(a, -1, FR_MODULUS) * R_MONT % FR_MODULUS


@pytest.mark.parametrize("fid,p", [(orc.FR, FR_MODULUS), (orc.FQ, FQ_MODULUS)])
def test_closed_forms_random(fid, p):
    rng = random.Random(1234 + fid)
    rinv = pow(R_MONT, -1, p)
    specials = [0, 1, 2, p - 1, p - 2, (p - 1) // 2, R_MONT % p, (R_MONT * R_MONT) % p]
    vals = specials + [rng.randrange(p) for _ in range(200)]
    for i in range(len(vals)):
        a, b = vals[i], vals[(i * 7 + 3) % len(vals)]
        assert orc.to_int(orc.mul(fid, a, b)) == a * b * rinv % p
        assert orc.to_int(orc.add(fid, a, b)) == (a + b) % p
        assert orc.to_int(orc.sub(fid, a, b)) == (a - b) % p
        assert orc.to_int(orc.neg(fid, a)) == (-a) % p
    for a in vals[:40]:
        if a == 0:
            assert orc.to_int(orc.inv(fid, 0)) == 0
            continue
        # inverse of a Montgomery element: (aR^-1)^-1 R = a^-1 R^2
        assert orc.to_int(orc.inv(fid, a)) == pow(a, -1, p) * R_MONT * R_MONT % p


def test_vector_ops_match_scalar_ops():
    from zkpoly_amd.field import random_elements

    a = random_elements(64, 1)
    b = random_elements(64, 2)
    m = orc.vec("orc_vec_mul", orc.FR, a, b)
    for i in range(64):
        assert np.array_equal(m[i], orc.mul(orc.FR, a[i], b[i]))


def test_random_elements_are_canonical():
    from zkpoly_amd.field import array_to_ints, random_elements

    xs = array_to_ints(random_elements(4096, 7))
    assert all(0 <= x < FR_MODULUS for x in xs)
    assert len(set(xs)) == 4096
    # deterministic per seed
    assert np.array_equal(random_elements(16, 7), random_elements(16, 7))
