"""Fused elementwise kernels: codegen checks (CPU) and GPU parity with the oracle.

The reference has no fused-kernel unit test (fused kernels are exercised only
through full compiler runs); the expected values here come from the oracle's
elementwise field ops, which are pinned against the reference's field KATs.
Semantics checked: kernel_gen.rs:457-566 of core/src/fused_kernels (Pp / Ss / Sp
ops, Lagrange vs coefficient form, Neg, Pow), in-place outputs and the
vars / mut_vars ordering of common/src/arith.rs:411-452.
"""
import json

import numpy as np
import pytest

import oracle as orc
from helpers import cpu2gpu, device_array, download, gpu2cpu, upload
from zkpoly_amd import fused, hip
from zkpoly_amd.field import mont_limbs, random_elements
from zkpoly_amd.runtime import GpuBuffer, Libs, Scalar, ScalarArray

FR = orc.FR


# ---------------------------------------------------------------- CPU: codegen


def test_codegen_arity_and_symbols():
    src = fused.codegen(Libs(), fused.graph_ab_gamma_c_minus_z(), "k1")
    assert 'extern "C" hipError_t k1(const ConstPolyPtr* vars, const PolyPtr* mut_vars' in src
    assert "k1_batched" in src and "out2[0] = 5; out2[1] = 1;" in src
    src = fused.codegen(Libs(), fused.graph_inplace_mixed(), "k2")
    # x is in place: vars = [y, s], mut_vars = [x', d, t, q]
    assert "out2[0] = 2; out2[1] = 4;" in src


def test_codegen_marks_mutable_vars_the_kernel_reads():
    """Only the in-place output's slot is read; the others are written in full."""
    src = fused.codegen(Libs(), fused.graph_inplace_mixed(), "k2")
    assert 'void k2_mut_reads(unsigned char* out) { out[0] = 1; out[1] = 0; out[2] = 0; out[3] = 0;' in src
    src = fused.codegen(Libs(), fused.graph_ab_gamma_c_minus_z(), "k3")
    assert 'void k3_mut_reads(unsigned char* out) { out[0] = 0;' in src


def test_codegen_coef_rule_text():
    src = fused.codegen(Libs(), fused.graph_coef_rules(), "k3")
    assert src.count("(idx == 0 && is_first)") >= 3 * 2  # single + batched kernels


@pytest.mark.parametrize(
    "mutate,msg",
    [
        (lambda g: g["g"][0].append({"op": {"Arith": {"Bin": [{"Pp": "Div"}, 0, 1]}}}) or g["outputs"].__setitem__(0, len(g["g"][0])) or g["g"][0].append({"op": {"Output": {"outer_id": 99, "typ": "ScalarArray", "store_node": len(g["g"][0]) - 1, "in_node": None}}}), "division"),
        (lambda g: g.__setitem__("poly_repr", "Weird"), "poly_repr"),
        (lambda g: g["outputs"].append(0), "Output"),
        (lambda g: g["g"][0][5]["op"]["Arith"]["Bin"].__setitem__(1, 999), "out of range"),
    ],
)
def test_codegen_rejects_invalid_graphs(mutate, msg):
    g = json.loads(fused.graph_ab_gamma_c_minus_z())
    mutate(g)
    with pytest.raises(fused.CodegenError) as e:
        fused.codegen(Libs(), json.dumps(g), "bad")
    assert msg.lower() in str(e.value).lower()


def test_codegen_rejects_bad_name():
    with pytest.raises(fused.CodegenError):
        fused.codegen(Libs(), fused.graph_ab_gamma_c_minus_z(), "1bad-name")


def test_codegen_accepts_flat_vertex_list():
    g = json.loads(fused.graph_ab_gamma_c_minus_z())
    g["g"] = g["g"][0]
    assert "k4_kernel" in fused.codegen(Libs(), json.dumps(g), "k4")


# ---------------------------------------------------------------- GPU


def meta(name, nv, nm):
    return fused.FusedKernelMeta(name, nv, nm, fused.DEMO_DIR)


def call(kernel, outputs, inputs, stream):
    m = kernel.meta
    arg = hip.DeviceBuffer(fused.arg_buffer_bytes(m))
    local = hip.DeviceBuffer(256)
    kernel.get_fn().f([GpuBuffer(arg.addr, arg.nbytes), GpuBuffer(local.addr, 256)] + outputs, [stream] + inputs)
    stream.sync()


def scalar_dev(x_mont):
    b = hip.DeviceBuffer(32)
    b.upload(np.asarray(x_mont, dtype=np.uint64))
    return b, Scalar(b.addr)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 1000, 1 << 16])
def test_ab_gamma_c_minus_z(libs, stream, n):
    a, b, c = (random_elements(n, s) for s in (1, 2, 3))
    gz = random_elements(2, 4)
    bufs = [upload(x) for x in (a, b, c)]
    gb, gs = scalar_dev(gz[0])
    zb, zs = scalar_dev(gz[1])
    ob, oa = device_array(n)
    k = fused.FusedKernel(libs, meta("fused_arith_abgcz", 5, 1))
    call(k, [oa], [bufs[0][1], bufs[1][1], bufs[2][1], gs, zs], stream)
    want = orc.vec("orc_vec_mul", FR, a, b)
    want = orc.vec("orc_vec_add", FR, want, orc.vec("orc_vec_mul", FR, np.broadcast_to(gz[0], (n, 4)).copy(), c))
    want = orc.vec("orc_vec_sub", FR, want, np.broadcast_to(gz[1], (n, 4)).copy())
    assert np.array_equal(download(oa, ob), want)


@pytest.mark.gpu
def test_rotated_and_sliced_operands(libs, stream):
    n = 4096
    a, b, c = (random_elements(n, s) for s in (11, 12, 13))
    gz = random_elements(2, 14)
    ab, aa = device_array(n)
    aa.rotate(5)
    cpu2gpu(a, aa, ab)
    bb, whole = device_array(2 * n)
    whole.rotate(-7)
    bs = whole.slice(n // 2, n // 2 + n)
    cpu2gpu(b, bs, bb)
    cb, ca = upload(c)
    gb, gs = scalar_dev(gz[0])
    zb, zs = scalar_dev(gz[1])
    ob, oa = device_array(n)
    oa.rotate(3)
    k = fused.FusedKernel(libs, meta("fused_arith_abgcz", 5, 1))
    call(k, [oa], [aa, bs, ca, gs, zs], stream)
    want = orc.vec("orc_vec_mul", FR, a, b)
    want = orc.vec("orc_vec_add", FR, want, orc.vec("orc_vec_mul", FR, np.broadcast_to(gz[0], (n, 4)).copy(), c))
    want = orc.vec("orc_vec_sub", FR, want, np.broadcast_to(gz[1], (n, 4)).copy())
    assert np.array_equal(gpu2cpu(oa, ob), want)


@pytest.mark.gpu
def test_coefficient_form_rules(libs, stream):
    n = 777
    p = random_elements(n, 21)
    s = random_elements(1, 22)[0]
    pb, pa = upload(p)
    sb, ss = scalar_dev(s)
    outs = [device_array(n) for _ in range(4)]
    k = fused.FusedKernel(libs, meta("fused_arith_coef_rules", 2, 4))
    call(k, [o[1] for o in outs], [pa, ss], stream)
    got = [download(o[1], o[0]) for o in outs]
    sv = np.broadcast_to(s, (n, 4)).copy()
    neg_p = orc.vec("orc_vec_neg", FR, p)
    # s + p only on the constant coefficient
    w0 = p.copy()
    w0[0] = orc.add(FR, s, p[0])
    # s - p: constant coefficient s - p0, the rest -p
    w1 = neg_p.copy()
    w1[0] = orc.sub(FR, s, p[0])
    # p - s: constant coefficient p0 - s, the rest p
    w2 = p.copy()
    w2[0] = orc.sub(FR, p[0], s)
    w3 = orc.vec("orc_vec_mul", FR, sv, p)
    for i, w in enumerate((w0, w1, w2, w3)):
        assert np.array_equal(got[i], w), i


@pytest.mark.gpu
def test_inplace_scalar_output_pow_neg(libs, stream):
    n = 5000
    x = random_elements(n, 31)
    y = random_elements(n, 32)
    s = random_elements(1, 33)[0]
    xb, xa = upload(x)
    yb, ya = upload(y)
    sb, ss = scalar_dev(s)
    db, da = device_array(n)
    tb, ts = scalar_dev(np.zeros(4, dtype=np.uint64))
    qb, qa = device_array(n)
    k = fused.FusedKernel(libs, meta("fused_arith_inplace_mixed", 2, 4))
    # mut_vars: [x (in place), d, t (scalar), q]; vars: [y, s]
    call(k, [xa, da, ts, qa], [ya, ss], stream)
    xy = orc.vec("orc_vec_add", FR, x, y)
    cube = orc.vec_pow(FR, xy, 3)
    want_x = orc.vec("orc_vec_mul", FR, np.broadcast_to(s, (n, 4)).copy(), cube)
    assert np.array_equal(download(xa, xb), want_x)
    assert np.array_equal(download(da, db), orc.vec("orc_vec_neg", FR, orc.vec("orc_vec_sub", FR, x, y)))
    t = orc.add(FR, orc.mul(FR, s, s), s)
    assert list(tb.download(np.uint64, 4)) == list(t)
    assert np.array_equal(download(qa, qb), orc.vec("orc_vec_mul", FR, np.broadcast_to(t, (n, 4)).copy(), y))


@pytest.mark.gpu
def test_batched_entry_matches_single(libs, stream):
    """<name>_batched over instances of uneven length equals one call per instance."""
    import ctypes

    from zkpoly_amd.runtime import PolyPtr

    lens = [1, 17, 1024, 3000, 0, 65536]
    k = fused.FusedKernel(libs, meta("fused_arith_abgcz", 5, 1))
    keep, vars_, muts, wants = [], [], [], []
    for i, n in enumerate(lens):
        m = max(n, 1)
        a, b, c = (random_elements(m, 100 * i + j) for j in range(3))
        gz = random_elements(2, 100 * i + 7)
        bufs = [upload(v) for v in (a, b, c)]
        g_ = scalar_dev(gz[0])
        z_ = scalar_dev(gz[1])
        ob, oa = device_array(m)
        keep += bufs + [g_, z_, (ob, oa)]
        vars_ += [PolyPtr.from_array(bb[1]) for bb in bufs]
        vars_ += [PolyPtr(g_[0].addr, 1, 0, 0, 1), PolyPtr(z_[0].addr, 1, 0, 0, 1)]
        for bb in bufs:
            bb[1].len = n
        muts.append(PolyPtr(ob.addr, n, 0, 0, max(n, 1)))
        vars_[-5].len = vars_[-4].len = vars_[-3].len = n
        w = orc.vec("orc_vec_mul", FR, a, b)
        w = orc.vec("orc_vec_add", FR, w, orc.vec("orc_vec_mul", FR, np.broadcast_to(gz[0], (m, 4)).copy(), c))
        w = orc.vec("orc_vec_sub", FR, w, np.broadcast_to(gz[1], (m, 4)).copy())
        wants.append((ob, w[:n]))
    prefix = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    va = (PolyPtr * len(vars_))(*vars_)
    ma = (PolyPtr * len(muts))(*muts)
    dv, dm, dp = hip.DeviceBuffer(ctypes.sizeof(va)), hip.DeviceBuffer(ctypes.sizeof(ma)), hip.DeviceBuffer(prefix.nbytes)
    dv.upload(np.frombuffer(bytes(va), dtype=np.uint8))
    dm.upload(np.frombuffer(bytes(ma), dtype=np.uint8))
    dp.upload(prefix)
    hip.check(k.c_batched(dv.addr, dm.addr, dp.addr, len(lens), int(prefix[-1]), True, stream.raw(), None), "batched")
    stream.sync()
    for (ob, w), n in zip(wants, lens):
        if n:
            assert np.array_equal(ob.download(np.uint64, n * 4).reshape(n, 4), w)


# ---------------------------------------------------------------- pipelined form


class HostPoly:
    """A pinned host polynomial, optionally rotated (logical i at physical (i - r) mod n)."""

    def __init__(self, data, rotate=0):
        n = data.shape[0]
        self.buf = hip.PinnedBuffer(n * 32)
        self.arr = ScalarArray(n, self.buf.addr, device="cpu")
        self.arr.rotate(rotate)
        view = self.buf.view().reshape(n, 4)
        idx = (np.arange(n) + n - self.arr.get_rotation()) % n  # ScalarArray.physical_index, vectorised
        assert idx[n - 1] == self.arr.physical_index(n - 1)
        view[idx] = data
        self.idx = idx

    def logical(self):
        return self.buf.view().reshape(-1, 4)[self.idx].copy()


class HostScalar:
    def __init__(self, value):
        self.buf = hip.PinnedBuffer(32)
        self.buf.view()[:] = value
        self.s = Scalar(self.buf.addr, device="cpu")

    def value(self):
        return self.buf.view().copy()


def dev_buffers(count, nbytes):
    bufs = [hip.DeviceBuffer(nbytes) for _ in range(count)]
    return bufs, [GpuBuffer(b.addr, nbytes) for b in bufs]


def pipelined(name, nv, nm, parts, scalars, mut_scalars):
    m = fused.FusedKernelMeta(name, nv, nm, fused.DEMO_DIR, fused.PipelinedMeta(parts, scalars, mut_scalars))
    return fused.PipelinedFusedKernel(Libs(), m)


def run_pipelined(k, parts, scalars, polys, mut_scalars, mut_polys, n):
    """Lays out the reference's argument lists (core/src/fused_kernels.rs:219-227)."""
    chunk_bytes = (n // parts) * 32
    arg_b, (arg,) = dev_buffers(1, fused.pipelined_arg_buffer_bytes(k.meta))
    loc_b, (loc,) = dev_buffers(1, 256)
    ms_b, ms_g = dev_buffers(len(mut_scalars), 32)
    mp_b, mp_g = dev_buffers(3 * len(mut_polys), chunk_bytes)
    s_b, s_g = dev_buffers(len(scalars), 32)
    p_b, p_g = dev_buffers(2 * len(polys), chunk_bytes)
    mut_var = [arg, loc] + [s.s for s in mut_scalars] + [p.arr for p in mut_polys] + ms_g + mp_g
    var = [s.s for s in scalars] + [p.arr for p in polys] + s_g + p_g
    k.get_fn().f(mut_var, var, lambda i: i)


@pytest.mark.gpu
@pytest.mark.parametrize("n,parts", [(1 << 20, 8), (4 * 3000, 4), (1 << 16, 32)])
def test_pipelined_ab_gamma_c_minus_z(stream, n, parts):
    a, b, c = (random_elements(n, 300 + i) for i in range(3))
    gamma, z = random_elements(1, 310)[0], random_elements(1, 311)[0]
    ha, hb, hc = HostPoly(a, 1000), HostPoly(b), HostPoly(c, n - 7)
    out = HostPoly(np.zeros((n, 4), np.uint64), 5)
    k = pipelined("fused_arith_pipe_abgcz", 5, 1, parts, 2, 0)
    run_pipelined(k, parts, [HostScalar(gamma), HostScalar(z)], [ha, hb, hc], [], [out], n)
    g = np.broadcast_to(gamma, (n, 4)).copy()
    zz = np.broadcast_to(z, (n, 4)).copy()
    want = orc.vec("orc_vec_sub", FR, orc.vec("orc_vec_add", FR, orc.vec("orc_vec_mul", FR, a, b),
                                              orc.vec("orc_vec_mul", FR, g, c)), zz)
    assert np.array_equal(out.logical(), want)


@pytest.mark.gpu
def test_pipelined_mutable_scalar_and_in_place(libs, stream):
    n, parts = 6 * 4096, 6
    x, y = random_elements(n, 320), random_elements(n, 321)
    s = random_elements(1, 322)[0]
    hx, hy = HostPoly(x, 77), HostPoly(y, 4096 * 5 + 3)
    hq = HostPoly(np.zeros((n, 4), np.uint64))
    ht = HostScalar(np.zeros(4, np.uint64))
    k = pipelined("fused_arith_pipe_mixed", 2, 3, parts, 1, 1)
    run_pipelined(k, parts, [HostScalar(s)], [hy], [ht], [hx, hq], n)
    t = orc.add(FR, orc.mul(FR, s, s), s)
    assert np.array_equal(ht.value(), t)
    tv = np.broadcast_to(t, (n, 4)).copy()
    assert np.array_equal(hx.logical(), orc.vec("orc_vec_mul", FR, tv, orc.vec("orc_vec_add", FR, x, y)))
    assert np.array_equal(hq.logical(), orc.vec("orc_vec_neg", FR, orc.vec("orc_vec_sub", FR, x, y)))


@pytest.mark.gpu
def test_pipelined_coefficient_rules_see_element_zero_once(stream):
    """Only the first chunk carries is_first, so the constant-coefficient rules
    touch global element 0 and nothing at the start of later chunks."""
    n, parts = 4 * 1024, 4
    p = random_elements(n, 330)
    s = random_elements(1, 331)[0]
    outs = [HostPoly(np.zeros((n, 4), np.uint64), r) for r in (0, 1, 1024, n - 1)]
    k = pipelined("fused_arith_pipe_coef", 2, 4, parts, 1, 0)
    run_pipelined(k, parts, [HostScalar(s)], [HostPoly(p, 9)], [], outs, n)
    sv = np.broadcast_to(s, (n, 4)).copy()
    neg_p = orc.vec("orc_vec_neg", FR, p)
    w0 = p.copy()
    w0[0] = orc.add(FR, s, p[0])
    w1 = neg_p.copy()
    w1[0] = orc.sub(FR, s, p[0])
    w2 = p.copy()
    w2[0] = orc.sub(FR, p[0], s)
    w3 = orc.vec("orc_vec_mul", FR, sv, p)
    for i, w in enumerate((w0, w1, w2, w3)):
        assert np.array_equal(outs[i].logical(), w), i


class PageablePoly:
    """A host polynomial in ordinary (pageable) numpy memory: the driver stages
    every chunk through device buffers, write-only outputs included."""

    def __init__(self, data, rotate=0):
        n = data.shape[0]
        self.buf = np.zeros((n, 4), dtype=np.uint64)
        self.arr = ScalarArray(n, self.buf.ctypes.data, device="cpu")
        self.arr.rotate(rotate)
        self.idx = (np.arange(n) + n - self.arr.get_rotation()) % n
        self.buf[self.idx] = data

    def logical(self):
        return self.buf[self.idx].copy()


class PageableScalar:
    """A host scalar in pageable numpy memory."""

    def __init__(self, value):
        self.value = np.array(value, dtype=np.uint64)
        self.s = Scalar(self.value.ctypes.data, device="cpu")


@pytest.mark.gpu
def test_pipelined_pageable_host_memory(stream):
    n, parts = 8 * 2048, 8
    a, b, c = (random_elements(n, 340 + i) for i in range(3))
    gamma, z = random_elements(1, 350)[0], random_elements(1, 351)[0]
    out = PageablePoly(np.zeros((n, 4), np.uint64), 333)
    k = pipelined("fused_arith_pipe_abgcz", 5, 1, parts, 2, 0)
    run_pipelined(k, parts, [PageableScalar(gamma), PageableScalar(z)],
                  [PageablePoly(a, 5), PageablePoly(b), PageablePoly(c, 77)], [], [out], n)
    g = np.broadcast_to(gamma, (n, 4)).copy()
    zz = np.broadcast_to(z, (n, 4)).copy()
    want = orc.vec("orc_vec_sub", FR, orc.vec("orc_vec_add", FR, orc.vec("orc_vec_mul", FR, a, b),
                                              orc.vec("orc_vec_mul", FR, g, c)), zz)
    assert np.array_equal(out.logical(), want)


@pytest.mark.gpu
def test_pipelined_mixed_memory(stream):
    """Staged chunks (a pageable in-place polynomial) next to a pinned write-only
    output, which the kernel writes directly in each chunk."""
    n, parts = 6 * 4096, 6
    x, y = random_elements(n, 360), random_elements(n, 361)
    s = random_elements(1, 362)[0]
    hx, hy = PageablePoly(x, 4095), HostPoly(y, 12)
    hq = HostPoly(np.zeros((n, 4), np.uint64), n - 1)
    ht = HostScalar(np.zeros(4, np.uint64))
    k = pipelined("fused_arith_pipe_mixed", 2, 3, parts, 1, 1)
    run_pipelined(k, parts, [PageableScalar(s)], [hy], [ht], [hx, hq], n)
    t = orc.add(FR, orc.mul(FR, s, s), s)
    assert np.array_equal(ht.value(), t)
    tv = np.broadcast_to(t, (n, 4)).copy()
    assert np.array_equal(hx.logical(), orc.vec("orc_vec_mul", FR, tv, orc.vec("orc_vec_add", FR, x, y)))
    assert np.array_equal(hq.logical(), orc.vec("orc_vec_neg", FR, orc.vec("orc_vec_sub", FR, x, y)))
