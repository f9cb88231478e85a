"""GPU parity of libpoly.so against the oracle (bit-exact, Montgomery form).

Mirrors core/tests/poly.rs at its size K = 20 with its rotations: test_binary
(:29-118, a rotated 10, b rotated -10, res rotated 3), test_eval (:120-178),
test_kate (:180-246, p rotated 2, res rotated -2; the reference compares
q_0..q_{n-2}, here q_{n-1} = 0 is checked too), test_zero_one (:248-338),
test_scan (:340-413, p rotated 3, res rotated -5), test_invert (:415-471, with
zeros), test_invert_scalar / test_pow_scalar (:473-567, including zero).  Seeded
inputs replace the reference's XorShift / OsRng draws.  Beyond the reference:
ragged lengths around the 2048-element scan tile, more than 256 tiles (the
partial scan's multi-element path), sliced views, operands of unequal length
(poly_basic.cu's copy / negate tail) and empty inputs.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import cpu2gpu, device_array, gpu2cpu
from zkpoly_amd import hip
from zkpoly_amd.field import FR_MODULUS, array_to_ints, from_mont, mont_limbs, random_elements, to_mont
from zkpoly_amd.poly import (KateDivision, PolyAdd, PolyEval, PolyInvert, PolyMul, PolyOneCoef,
                             PolyOneLagrange, PolyScan, PolySub, PolyZero, ScalarInv, ScalarPow)
from zkpoly_amd.runtime import GpuBuffer, Scalar

pytestmark = pytest.mark.gpu
FR = orc.FR
K = 20
RAGGED = (1, 2, 3, 255, 2047, 2048, 2049, 4097, 3 * (1 << 19) + 17)


def put(host, rotate=0, slice_of=None):
    """Device array holding `host` logically, optionally rotated or sliced
    out of a larger buffer at [slice_of[0], slice_of[0] + n)."""
    n = host.shape[0]
    if slice_of is not None:
        start, whole_len = slice_of
        buf, whole = device_array(whole_len)
        whole.rotate(rotate)
        arr = whole.slice(start, start + n)
    else:
        buf, arr = device_array(n)
        if rotate:
            arr.rotate(rotate)
    cpu2gpu(host, arr, buf)
    return buf, arr


def empty(n, rotate=0):
    buf, arr = device_array(max(n, 1))
    arr.len = n
    if rotate and n:
        arr.rotate(rotate)
    return buf, arr


def dev_scalar(limbs):
    buf = hip.DeviceBuffer(32)
    buf.upload(np.ascontiguousarray(limbs, dtype=np.uint64).reshape(4))
    return buf, Scalar(buf.addr)


def read_scalar(buf):
    return buf.download(np.uint64, 4)


def temp_buffer(nbytes):
    buf = hip.DeviceBuffer(max(nbytes, 1))
    return buf, GpuBuffer(buf.addr, nbytes)


def run_binary(libs, stream, op, a_host, b_host, res_len, ra=0, rb=0, rr=0):
    ab, a = put(a_host, ra)
    bb, b = put(b_host, rb)
    rbuf, res = empty(res_len, rr)
    op(libs).get_fn().f([res], [a, b, stream])
    stream.sync()
    return gpu2cpu(res, rbuf)


def want_binary(name, a, b):
    """poly_basic.cu:4-54: the common part combined, the longer operand's tail
    copied (negated for a - b when b is the longer)."""
    m = min(a.shape[0], b.shape[0])
    head = orc.vec(name, FR, a[:m], b[:m])
    if a.shape[0] >= b.shape[0]:
        tail = a[m:]
    else:
        tail = b[m:] if name != "orc_vec_sub" else orc.vec("orc_vec_neg", FR, b[m:])
    return np.concatenate([head, tail.reshape(-1, 4)])


@pytest.mark.parametrize("op,name", [(PolyAdd, "orc_vec_add"), (PolySub, "orc_vec_sub"), (PolyMul, "orc_vec_mul")])
def test_binary_rotated(libs, stream, op, name):
    n = 1 << K
    a, b = random_elements(n, 11), random_elements(n, 12)
    got = run_binary(libs, stream, op, a, b, n, ra=10, rb=-10, rr=3)
    assert np.array_equal(got, orc.vec(name, FR, a, b))


@pytest.mark.parametrize("op,name", [(PolyAdd, "orc_vec_add"), (PolySub, "orc_vec_sub"), (PolyMul, "orc_vec_mul")])
@pytest.mark.parametrize("la,lb", [(1000, 1500), (1500, 1000), (1, 4097), (4097, 0), (0, 3)])
def test_binary_unequal_lengths(libs, stream, op, name, la, lb):
    a, b = random_elements(la, 20 + la), random_elements(lb, 30 + lb)
    a, b = a.reshape(-1, 4), b.reshape(-1, 4)
    ab, av = put(a) if la else empty(0)
    bb, bv = put(b) if lb else empty(0)
    n = max(la, lb)
    rbuf, res = empty(n, 7)
    op(libs).get_fn().f([res], [av, bv, stream])
    stream.sync()
    assert np.array_equal(gpu2cpu(res, rbuf), want_binary(name, a, b))


def test_binary_on_slices_and_in_place(libs, stream):
    n = 5000
    a, b = random_elements(n, 41), random_elements(n, 42)
    ab, av = put(a, rotate=123, slice_of=(300, 8192))
    bb, bv = put(b, rotate=5000, slice_of=(300, 16384))  # wraps round the end of its buffer
    PolyMul(libs).get_fn().f([av], [av, bv, stream])  # res aliases a: element-wise, so exact
    stream.sync()
    assert np.array_equal(gpu2cpu(av, ab), orc.vec("orc_vec_mul", FR, a, b))


def test_zero_one(libs, stream):
    n = 1 << K
    one = mont_limbs(1)
    for op, want0, want_rest in ((PolyZero, 0, 0), (PolyOneCoef, 1, 0), (PolyOneLagrange, 1, 1)):
        buf, arr = put(random_elements(n, 50), rotate=17)
        op(libs).get_fn().f([arr], [stream])
        stream.sync()
        got = gpu2cpu(arr, buf)
        rows = [one if v else np.zeros(4, np.uint64) for v in (want0, want_rest)]
        assert np.array_equal(got[0], rows[0]), op.__name__
        assert (got[1:] == rows[1]).all(), op.__name__


@pytest.mark.parametrize("n", [1 << K] + list(RAGGED))
def test_eval(libs, stream, n):
    a = random_elements(n, 60 + n % 1000)
    x = random_elements(1, 61)[0]
    ab, av = put(a, rotate=n // 3)
    xb, xs = dev_scalar(x)
    rb, rs = dev_scalar(np.zeros(4, np.uint64))
    op = PolyEval(libs)
    tb, tmp = temp_buffer(op.get_buffer_size(n))
    op.get_fn().f([tmp, rs], [av, xs, stream])
    stream.sync()
    assert np.array_equal(read_scalar(rb), orc.poly_eval(a, x)), n


def test_eval_special_points_and_empty(libs, stream):
    n = 10000
    a = random_elements(n, 70)
    ab, av = put(a)
    op = PolyEval(libs)
    tb, tmp = temp_buffer(op.get_buffer_size(n))
    rb, rs = dev_scalar(random_elements(1, 71)[0])
    for x in (mont_limbs(0), mont_limbs(1), mont_limbs(FR_MODULUS - 1)):
        xb, xs = dev_scalar(x)
        op.get_fn().f([tmp, rs], [av, xs, stream])
        stream.sync()
        assert np.array_equal(read_scalar(rb), orc.poly_eval(a, x))
    # an empty polynomial evaluates to zero everywhere
    eb, ev = empty(0)
    xb, xs = dev_scalar(random_elements(1, 72)[0])
    op.get_fn().f([tmp, rs], [ev, xs, stream])
    stream.sync()
    assert not read_scalar(rb).any()


@pytest.mark.parametrize("k", [K, 1, 2, 11, 12, 13, 21])
def test_kate(libs, stream, k):
    n = 1 << k
    p = random_elements(n, 80 + k)
    b = random_elements(1, 90 + k)[0]
    pb, pv = put(p, rotate=2 % n)
    bb, bs = dev_scalar(b)
    qbuf, qv = empty(n, -2 % n if n > 2 else 0)
    op = KateDivision(libs)
    tb, tmp = temp_buffer(op.get_buffer_size(k))
    op.get_fn().f([tmp, qv], [pv, bs, stream])
    stream.sync()
    assert np.array_equal(gpu2cpu(qv, qbuf), orc.kate_division(p, b)), k


def test_kate_in_place_and_at_zero(libs, stream):
    k = 14
    n = 1 << k
    p = random_elements(n, 95)
    pb, pv = put(p, rotate=1000)
    bb, bs = dev_scalar(mont_limbs(0))
    op = KateDivision(libs)
    tb, tmp = temp_buffer(op.get_buffer_size(k))
    op.get_fn().f([tmp, pv], [pv, bs, stream])  # q over p: each element is read before it is written
    stream.sync()
    got = gpu2cpu(pv, pb)
    assert np.array_equal(got[:-1], p[1:]) and not got[-1].any()  # division by X shifts down


@pytest.mark.parametrize("n", [1 << K] + list(RAGGED))
def test_scan(libs, stream, n):
    p = random_elements(n, 100 + n % 1000)
    x0 = random_elements(1, 101)[0]
    pb, pv = put(p, rotate=3 % n if n > 3 else 0)
    xb, xs = dev_scalar(x0)
    rbuf, rv = empty(n, -5 % n if n > 5 else 0)
    op = PolyScan(libs)
    tb, tmp = temp_buffer(op.get_buffer_size(n))
    op.get_fn().f([tmp, rv], [pv, xs, stream])
    stream.sync()
    got = gpu2cpu(rv, rbuf)
    # target = x0, x0 p_0, x0 p_0 p_1, ... (core/tests/poly.rs:404-412)
    want = orc.scan_mul(np.concatenate([x0.reshape(1, 4), p[: n - 1]]))
    assert np.array_equal(got, want), n


def test_scan_on_a_slice(libs, stream):
    n = 6000
    p = random_elements(n, 110)
    x0 = random_elements(1, 111)[0]
    pb, pv = put(p, rotate=3000, slice_of=(1000, 8192))  # physical start 6192: wraps
    xb, xs = dev_scalar(x0)
    rbuf, rwhole = device_array(16384)
    rwhole.rotate(7000)
    rv = rwhole.slice(2000, 2000 + n)  # physical start 11384: wraps too
    op = PolyScan(libs)
    tb, tmp = temp_buffer(op.get_buffer_size(n))
    op.get_fn().f([tmp, rv], [pv, xs, stream])
    stream.sync()
    want = orc.scan_mul(np.concatenate([x0.reshape(1, 4), p[: n - 1]]))
    assert np.array_equal(gpu2cpu(rv, rbuf), want)


@pytest.mark.parametrize("n", [1 << K] + list(RAGGED))
def test_invert(libs, stream, n):
    x = random_elements(n, 120 + n % 1000)
    x[::17] = 0  # zeros stay zero and do not disturb their neighbours
    if n > 2:
        x[-1] = 0
    buf, arr = put(x, rotate=n // 5)
    op = PolyInvert(libs)
    tb, tmp = temp_buffer(op.get_buffer_size(n))
    op.get_fn().f([tmp, arr], [stream])
    stream.sync()
    assert np.array_equal(gpu2cpu(arr, buf), orc.batch_invert(x)), n


def test_invert_all_zero_and_reused_temp(libs, stream):
    op = PolyInvert(libs)
    n = 4097
    tb, tmp = temp_buffer(op.get_buffer_size(n))
    zb, zv = put(np.zeros((n, 4), np.uint64))
    op.get_fn().f([tmp, zv], [stream])
    stream.sync()
    assert not gpu2cpu(zv, zb).any()
    # the same buffer serves a smaller call afterwards
    x = random_elements(1000, 130)
    xb, xv = put(x)
    op.get_fn().f([tmp, xv], [stream])
    stream.sync()
    assert np.array_equal(gpu2cpu(xv, xb), orc.batch_invert(x))


def test_invert_scalar(libs, stream):
    f = ScalarInv(libs).get_fn().f
    for v in (random_elements(1, 140)[0], mont_limbs(1), mont_limbs(FR_MODULUS - 1)):
        sb, s = dev_scalar(v)
        f([s], [stream])
        stream.sync()
        assert np.array_equal(read_scalar(sb), orc.inv(FR, v))
    sb, s = dev_scalar(np.zeros(4, np.uint64))
    f([s], [stream])
    stream.sync()
    assert not read_scalar(sb).any()


@pytest.mark.parametrize("exp", [0, 1, 2, 65537, 0xFFFFFFFFFFFFFFFF, 0x9E3779B97F4A7C15])
def test_pow_scalar(libs, stream, exp):
    v = random_elements(1, 150)[0]
    f = ScalarPow(libs, exp).get_fn().f
    sb, s = dev_scalar(v)
    f([s], [stream])
    stream.sync()
    want = pow(from_mont(array_to_ints(v.reshape(1, 4))[0]), exp, FR_MODULUS)
    assert np.array_equal(read_scalar(sb), mont_limbs(want))
    assert np.array_equal(read_scalar(sb), orc.pow_(FR, v, exp))
    zb, z = dev_scalar(np.zeros(4, np.uint64))
    f([z], [stream])
    stream.sync()
    assert np.array_equal(read_scalar(zb), mont_limbs(1 if exp == 0 else 0))


# ---------------------------------------------------------------- permute

R2 = orc.fe((1 << 512) % FR_MODULUS)


def mont_small(xs):
    """Montgomery form of small non-negative ints (< 2^64), vectorised: x * R^2 * R^-1."""
    limbs = np.zeros((len(xs), 4), dtype=np.uint64)
    limbs[:, 0] = np.asarray(xs, dtype=np.uint64)
    return orc.vec("orc_vec_mul", FR, limbs, np.broadcast_to(R2, limbs.shape).copy())


def run_permute(libs, stream, inp, tab, usable, views=None):
    """views: optional (rotate_in, rotate_tab, slice_res) for the GPU arrays."""
    from zkpoly_amd.poly import PolyPermute

    rot_in, rot_tab, slice_res = views or (0, 0, False)
    ib, iv = put(inp, rotate=rot_in)
    tb_, tv = put(tab, rotate=rot_tab)
    if slice_res:
        rib, riw = device_array(usable + 1000)
        riw.rotate(700)
        ri = riw.slice(300, 300 + usable)
        rtb, rtw = device_array(usable + 999)
        rt = rtw.slice(999, 999 + usable)
    else:
        rib, ri = empty(usable)
        rtb, rt = empty(usable)
    op = PolyPermute(libs)
    tmpb, tmp = temp_buffer(op.get_buffer_size(usable))
    op.get_fn().f([ri, rt, tmp], [iv, tv, stream])
    stream.sync()
    return gpu2cpu(ri, rib), gpu2cpu(rt, rtb)


def check_permute(libs, stream, inp, tab, usable, views=None):
    got_in, got_tab = run_permute(libs, stream, inp, tab, usable, views)
    want_in, want_tab = orc.permute(inp, tab, usable)
    assert np.array_equal(got_in, want_in), "permuted input"
    assert np.array_equal(got_tab, want_tab), "permuted table"


def lookup_case(n, usable, seed, spread):
    """core/tests/poly.rs:569-624: table values x mod spread, input drawn from
    the distinct values of the table's usable rows."""
    rng = np.random.default_rng(seed)
    tab = rng.integers(0, spread, n, dtype=np.uint64)
    distinct = np.unique(tab[:usable])
    inp = distinct[rng.integers(0, len(distinct), n)]
    return mont_small(inp), mont_small(tab)


def test_permute_reference_case(libs, stream):
    """core/tests/poly.rs:569-765: K = 20, 24 blinding rows + 1, table values
    below 2 * len, input sampled from the usable table's distinct values."""
    n = 1 << K
    usable = n - 25
    inp, tab = lookup_case(n, usable, 1, 2 * n)
    check_permute(libs, stream, inp, tab, usable)


@pytest.mark.parametrize("usable", [1, 2, 7, 2047, 2048, 2049, 4097, 3 * 2048 + 5, 100003, (1 << 21) + 3])
def test_permute_sizes(libs, stream, usable):
    n = usable + 3
    inp, tab = lookup_case(n, usable, usable, max(2, usable // 3))
    check_permute(libs, stream, inp, tab, usable)


def test_permute_field_elements_with_negatives(libs, stream):
    """Compressed lookup expressions look like random field elements; negatives
    p - x share their top bits: the sort compares whole 256-bit values."""
    n = 300000
    rng = np.random.default_rng(5)
    pool = np.concatenate([random_elements(4000, 6), orc.vec("orc_vec_neg", FR, mont_small(rng.integers(1, 1 << 40, 4000)))])
    tab = pool[rng.integers(0, len(pool), n)]
    distinct = np.unique(tab, axis=0)
    inp = distinct[rng.integers(0, len(distinct), n)]
    check_permute(libs, stream, inp, tab, n)


def test_permute_views_and_degenerate_values(libs, stream):
    n = 50000
    inp, tab = lookup_case(n, n, 7, 1000)
    check_permute(libs, stream, inp, tab, n, views=(123, n - 5, True))
    # every value equal
    same = mont_small(np.full(n, 42))
    check_permute(libs, stream, same, same, n)
    # all distinct: the permutation is the sorted table twice
    tab = mont_small(np.random.default_rng(8).permutation(n))
    got_in, got_tab = run_permute(libs, stream, tab[::-1].copy(), tab, n)
    assert np.array_equal(got_in, got_tab) and np.array_equal(got_in, orc.permute(tab, tab, n)[0])


def test_permute_value_missing_from_table(libs, stream):
    """Beyond the reference (which asserts on the device): a missing value keeps
    its row and the other rows take the smallest leftovers, as the oracle says."""
    n = 5000
    inp, tab = lookup_case(n, n, 9, 300)
    inp[::97] = mont_small([10 ** 6])  # not in the table
    check_permute(libs, stream, inp, tab, n)


def test_permute_reuses_temp_for_a_smaller_call(libs, stream):
    from zkpoly_amd.poly import PolyPermute

    op = PolyPermute(libs)
    big, small = 70000, 3000
    tmpb, tmp = temp_buffer(op.get_buffer_size(big))
    for usable, seed in ((big, 11), (small, 12)):
        inp, tab = lookup_case(usable, usable, seed, usable // 2)
        ib, iv = put(inp)
        tb_, tv = put(tab)
        rib, ri = empty(usable)
        rtb, rt = empty(usable)
        op.get_fn().f([ri, rt, tmp], [iv, tv, stream])
        stream.sync()
        want_in, want_tab = orc.permute(inp, tab, usable)
        assert np.array_equal(gpu2cpu(ri, rib), want_in)
        assert np.array_equal(gpu2cpu(rt, rtb), want_tab)


# ---------------------------------------------------------------- full size (2^24)


def gpu_eval(libs, stream, arr, x_limbs):
    op = PolyEval(libs)
    tb, tmp = temp_buffer(op.get_buffer_size(arr.len))
    xb, xs = dev_scalar(x_limbs)
    rb, rs = dev_scalar(np.zeros(4, np.uint64))
    op.get_fn().f([tmp, rs], [arr, xs, stream])
    stream.sync()
    return array_to_ints(read_scalar(rb).reshape(1, 4))[0]


def test_full_size_properties(libs, stream):
    """2^24 elements, checked through properties that do not depend on the size
    (the oracle pins every entry point bit-exactly up to 2^21 above): after one
    inversion x_i * y_i = 1 on 4096 sampled rows and every zero is still zero,
    and inverting again gives the input back; Kate's quotient q satisfies
    p(x) = q(x) (x - b) + p(b) at a random x, with the GPU's own eval; the
    prefix products end in the product of all elements."""
    k = 24
    n = 1 << k
    x = random_elements(n, 900)
    x[::1000] = 0
    buf, arr = put(x, rotate=12345)
    op = PolyInvert(libs)
    tb, tmp = temp_buffer(op.get_buffer_size(n))
    op.get_fn().f([tmp, arr], [stream])
    stream.sync()
    y = gpu2cpu(arr, buf)
    zero = ~x.any(axis=1)
    assert not y[zero].any(), "zeros must stay zero"
    rows = np.random.default_rng(904).choice(np.flatnonzero(~zero), 4096, replace=False)
    prod = orc.vec("orc_vec_mul", FR, np.ascontiguousarray(x[rows]), np.ascontiguousarray(y[rows]))
    assert np.array_equal(prod, np.broadcast_to(mont_limbs(1), prod.shape)), "x * x^-1 != 1"
    op.get_fn().f([tmp, arr], [stream])
    stream.sync()
    assert np.array_equal(gpu2cpu(arr, buf), x)
    del tb, tmp

    # Kate: q = (p - p(b)) / (X - b) in place over p
    b = random_elements(1, 901)[0]
    z = random_elements(1, 902)[0]
    p_at_z = from_mont(gpu_eval(libs, stream, arr, z))
    p_at_b = from_mont(gpu_eval(libs, stream, arr, b))
    kd = KateDivision(libs)
    tb, tmp = temp_buffer(kd.get_buffer_size(k))
    bb, bs = dev_scalar(b)
    kd.get_fn().f([tmp, arr], [arr, bs, stream])
    stream.sync()
    q_at_z = from_mont(gpu_eval(libs, stream, arr, z))
    zv, bv = from_mont(array_to_ints(z.reshape(1, 4))[0]), from_mont(array_to_ints(b.reshape(1, 4))[0])
    assert p_at_z == (q_at_z * (zv - bv) + p_at_b) % FR_MODULUS
    del tb, tmp, buf, arr

    # scan_mul: the last prefix product is the product of all (checked on a
    # 2^16 slice of the output against big ints)
    m = 1 << 16
    y = random_elements(n, 903)
    ybuf, yarr = put(y)
    sc = PolyScan(libs)
    tb, tmp = temp_buffer(sc.get_buffer_size(n))
    rbuf, rarr = empty(n)
    one_b, one_s = dev_scalar(mont_limbs(1))
    sc.get_fn().f([tmp, rarr], [yarr, one_s, stream])
    stream.sync()
    out = gpu2cpu(rarr, rbuf)
    # target_i = y_0 ... y_{i-1}: a window of consecutive outputs differs by the elements between them
    lo = n - m
    prod = from_mont(array_to_ints(out[lo].reshape(1, 4))[0])
    for v in array_to_ints(y[lo:n - 1]):
        prod = prod * from_mont(v) % FR_MODULUS
    assert prod == from_mont(array_to_ints(out[n - 1].reshape(1, 4))[0])


def test_permute_full_size_properties(libs, stream):
    """2^24 rows with values below 2^20 (the oracle pins permute bit-exactly up
    to 2^21 + 3 rows above).  These properties fix the result completely:
    A' is the input in ascending canonical order; at every run start of A' the
    table column holds the run's value; the other rows hold the table entries
    left over after the run starts, ascending, in row order."""
    n = 1 << 24
    inp, tab = lookup_case(n, n, 905, 1 << 20)
    got_in, got_tab = run_permute(libs, stream, inp, tab, n)
    one = np.zeros((n, 4), dtype=np.uint64)
    one[:, 0] = 1

    def small(a):  # Montgomery -> canonical; the values fit in limb 0
        c = orc.vec("orc_vec_mul", FR, np.ascontiguousarray(a), one)
        assert not c[:, 1:].any()
        return c[:, 0]

    a, s, i_c, t_c = small(got_in), small(got_tab), small(inp), small(tab)
    assert np.array_equal(a, np.sort(i_c)), "A' is the sorted input"
    start = np.ones(n, dtype=bool)
    start[1:] = a[1:] != a[:-1]
    assert np.array_equal(s[start], a[start]), "run starts take their own value"
    used = np.bincount(a[start].astype(np.int64), minlength=1 << 20)
    have = np.bincount(t_c.astype(np.int64), minlength=1 << 20)
    assert (have >= used).all()
    left = np.repeat(np.arange(1 << 20, dtype=np.uint64), have - used)
    assert np.array_equal(s[~start], left), "leftovers ascending in row order"
