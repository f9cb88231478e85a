"""CPU checks of the drop-in boundary (no GPU needed, no kernels launched).

* every function include/*.h declares is exported by the library its
  "library:" section names;
* PolyPtr has the reference's 40-byte #[repr(C)] layout (core/src/poly_ptr.rs:4-11);
* host-side validation returns hipErrorInvalidValue before touching a device
  (the reference's Rust asserts, core/src/ntt.rs:173-179, become error codes);
* ScalarArray rotation / slice index maps follow runtime/src/scalar.rs:351-444;
* MsmConfig and library naming follow common/src/msm_config.rs:14-24 and
  core/src/msm.rs:100-111; the MSM sizing call works without a device.
"""
import ctypes
import os
import re
import subprocess

import pytest

from zkpoly_amd import build
from zkpoly_amd.msm import MSM, MsmConfig, get_msm_lib_name
from zkpoly_amd.runtime import Libs, PolyPtr, ScalarArray

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
HIP_INVALID_VALUE = 1

DECL = re.compile(r"^\s*(?:hipError_t|unsigned long long|int|void)\s+(\w+)\s*\(", re.M)


def header_sections():
    """(header, library, [function names]) for every library: section."""
    out = []
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        parts = re.split(r"/\*\s*library:\s*(\S+)\s*\*/", text)
        for i in range(1, len(parts), 2):
            lib = parts[i].replace("<w>_<t>_<dbg>", "16_16_false")
            names = DECL.findall(parts[i + 1])
            out.append((h, lib, names))
    return out


def exported(path):
    r = subprocess.run(["nm", "-D", "--defined-only", path], stdout=subprocess.PIPE, text=True, check=True)
    return {line.split()[-1] for line in r.stdout.splitlines() if line.strip()}


def test_headers_declare_something():
    secs = header_sections()
    assert {s[1] for s in secs} >= {"libntt.so", "libpoly.so", "libzkp_msm.so", "libmsm_bn254_16_16_false.so",
                                    "libzkp_fused_codegen.so", "libzkp_fused_pipeline.so", "libzkp_serve.so"}
    assert all(s[2] for s in secs)


@pytest.mark.parametrize("header,lib,names", header_sections())
def test_library_exports_every_declared_symbol(header, lib, names):
    path = os.path.join(build.LIB, lib)
    assert os.path.exists(path), f"{lib} not built"
    ctypes.CDLL(path)  # loads without a GPU
    missing = sorted(set(names) - exported(path))
    assert not missing, f"{lib} lacks {missing} declared in {header}"


def test_polyptr_layout():
    assert ctypes.sizeof(PolyPtr) == 40
    assert [PolyPtr.__dict__[f].offset for f in ("ptr", "len", "rotate", "offset", "whole_len")] == [0, 8, 16, 24, 32]


def _ntt():
    from zkpoly_amd.ntt import load_lib

    return load_lib(Libs())


def test_ssip_ntt_rejects_bad_shapes_without_launch():
    lib = _ntt()
    # len not 2^log_len
    assert lib.ssip_ntt(PolyPtr(0x1000, 12, 0, 0, 12), 0x2000, 3, None, 8, 8) == HIP_INVALID_VALUE
    assert lib.ssip_ntt(PolyPtr(0x1000, 16, 0, 0, 16), 0x2000, 3, None, 8, 8) == HIP_INVALID_VALUE
    # null buffers
    assert lib.ssip_ntt(PolyPtr(0, 16, 0, 0, 16), 0x2000, 4, None, 8, 8) == HIP_INVALID_VALUE
    assert lib.ssip_ntt(PolyPtr(0x1000, 16, 0, 0, 16), 0, 4, None, 8, 8) == HIP_INVALID_VALUE
    # rotation on a slice, slice longer than its array, offset out of range
    assert lib.ssip_ntt(PolyPtr(0x1000, 16, 3, 5, 32), 0x2000, 4, None, 8, 8) == HIP_INVALID_VALUE
    assert lib.ssip_ntt(PolyPtr(0x1000, 16, 0, 0, 8), 0x2000, 4, None, 8, 8) == HIP_INVALID_VALUE
    assert lib.ssip_ntt(PolyPtr(0x1000, 16, 0, 32, 32), 0x2000, 4, None, 8, 8) == HIP_INVALID_VALUE
    # a length-1 transform is the identity and launches nothing
    assert lib.ssip_ntt(PolyPtr(0x1000, 1, 0, 0, 1), 0x2000, 0, None, 8, 8) == 0


def test_recompute_ntt_rejects_bad_args_without_launch():
    lib = _ntt()
    ok = PolyPtr(0x1000, 16, 0, 0, 16)
    assert lib.recompute_ntt(PolyPtr(0x1000, 12, 0, 0, 12), 0x2000, 2, 0x3000, 3, None, 8, 8) == HIP_INVALID_VALUE
    assert lib.recompute_ntt(ok, 0, 2, 0x3000, 4, None, 8, 8) == HIP_INVALID_VALUE
    assert lib.recompute_ntt(ok, 0x2000, 2, 0, 4, None, 8, 8) == HIP_INVALID_VALUE
    assert lib.recompute_ntt(ok, 0x2000, 0, 0x3000, 4, None, 8, 8) == HIP_INVALID_VALUE
    # the omega ladder has 32 entries, so transforms stop at 2^32
    big = 1 << 33
    assert lib.recompute_ntt(PolyPtr(0x1000, big, 0, 0, big), 0x2000, 2, 0x3000, 33, None, 8, 8) == HIP_INVALID_VALUE
    assert lib.recompute_ntt(PolyPtr(0x1000, 1, 0, 0, 1), 0x2000, 1, 0x3000, 0, None, 8, 8) == 0


def test_gen_pq_omegas_host_matches_oracle():
    """gen_pq_omegas is host code: checked here against the oracle's restatement
    (itself pinned to big-int powers in test_oracle_ntt_msm)."""
    import numpy as np

    import oracle as orc
    from zkpoly_amd.field import random_elements
    from zkpoly_amd.ntt import GenPqOmegas

    gen = GenPqOmegas(Libs()).get_fn()
    for log_len, seed in ((1, 1), (5, 2), (12, 3), (18, 4), (24, 5)):
        unit = random_elements(1, 40 + seed)[0]  # not a root of unity: exact powers still agree
        pq_len = GenPqOmegas.get_pq_len(log_len)
        pq = np.zeros((pq_len, 4), dtype=np.uint64)
        om = np.zeros((32, 4), dtype=np.uint64)
        gen(ScalarArray(pq_len, pq.ctypes.data, device="cpu"), ScalarArray(32, om.ctypes.data, device="cpu"),
            1 << log_len, unit)
        want_pq, want_om = orc.gen_pq_omegas(pq_len.bit_length(), 1 << log_len, unit)
        assert np.array_equal(pq, want_pq), log_len
        assert np.array_equal(om, want_om), log_len


def test_get_pq_len_follows_the_reference_stage_split():
    """core/src/ntt.rs:322-335 with stage threads (8, 8): worked by hand for
    2^24 (stage-1 degree 12 -> 2 rounds of 6; stage-2 degree 12 -> 3 rounds of
    4: 2^(6-1)) and 2^18 (9 fits one round: 2^8); never above 256 entries."""
    from zkpoly_amd.ntt import GenPqOmegas

    assert GenPqOmegas.get_pq_len(24) == 32
    assert GenPqOmegas.get_pq_len(18) == 256
    assert GenPqOmegas.get_pq_len(1) == 1
    assert max(GenPqOmegas.get_pq_len(k) for k in range(1, 33)) == 256


def test_distribute_powers_rejects_bad_args():
    lib = _ntt()
    assert lib.distribute_powers(PolyPtr(0x1000, 16, 0, 0, 16), 0x2000, 0, None) == HIP_INVALID_VALUE
    assert lib.distribute_powers(PolyPtr(0x1000, 16, 0, 0, 16), 0, 3, None) == HIP_INVALID_VALUE
    assert lib.distribute_powers(PolyPtr(0x1000, 0, 0, 0, 0), 0x2000, 3, None) == 0


def test_scalar_array_index_maps():
    # rotate(3) then slice(n, 2n) of a 2n array (core/tests/ntt.rs:82-87)
    n = 16
    a = ScalarArray(2 * n, 0)
    a.rotate(3)
    s = a.slice(n, 2 * n)
    assert s.slice_info.offset == (2 * n + n - 3) % (2 * n)
    assert [s.physical_index(i) for i in range(n)] == [(n - 3 + i) % (2 * n) for i in range(n)]
    p = PolyPtr.from_array(s)
    assert (p.len, p.rotate, p.offset, p.whole_len) == (n, 0, n - 3, 2 * n)
    # rotate(-2) of an unsliced array: logical i at (i + 2) mod n, PolyPtr offset = len - rotate
    b = ScalarArray(n, 0)
    b.rotate(-2)
    assert b.get_rotation() == n - 2
    assert [b.physical_index(i) for i in range(n)] == [(i + 2) % n for i in range(n)]
    q = PolyPtr.from_array(b)
    assert (q.len - q.rotate) % n == 2


def test_msm_config_and_names():
    assert MsmConfig().get_precompute() == 1
    assert MsmConfig(window_size=16, target_window=8).get_precompute() == 2
    assert MsmConfig(window_size=13, target_window=5).get_precompute() == 4
    assert get_msm_lib_name(MsmConfig(window_size=16, target_window=16)) == "libmsm_bn254_16_16_false.so"
    assert get_msm_lib_name(MsmConfig(window_size=12, target_window=4, debug=True)) == "libmsm_bn254_12_4_true.so"


def test_msm_sizing_call_without_device():
    op = MSM(Libs(), MsmConfig(window_size=16, target_window=16))
    small, big = op.get_buffer_size(1 << 10)[0], op.get_buffer_size(1 << 20)[0]
    assert 0 < small < big
    # two parts need less device memory than one
    op2 = MSM(Libs(), MsmConfig(window_size=16, target_window=16, parts=2))
    assert op2.get_buffer_size(1 << 20)[0] < big


def test_msm_config_query():
    from zkpoly_amd.msm import core_lib

    lib = core_lib(Libs())
    out = (ctypes.c_uint * 5)()
    assert lib.zkp_msm_config(16, 16, out) == 1
    assert list(out) == [16, 16, 16, 1, 1 << 15]
    assert lib.zkp_msm_config(1, 16, out) == 0
    assert lib.zkp_msm_workspace_size(1 << 20, 16, 16) > 0


def test_msm_config_extra_carry_window_keeps_reference_tables():
    """For s = 2, 3 the top signed digit of r - 1 carries, so one window more
    than ceil(254 / s) is needed.  With a target covering the reference's
    windows it joins them (one table, as the reference's MsmConfig says);
    where it would need a table the reference does not count, the config is
    rejected instead of reading past the caller's tables."""
    from zkpoly_amd.msm import core_lib

    lib = core_lib(Libs())
    out = (ctypes.c_uint * 5)()
    assert lib.zkp_msm_config(2, 127, out) == 1
    assert list(out) == [2, 128, 128, 1, 2]
    assert lib.zkp_msm_config(3, 85, out) == 1
    assert list(out) == [3, 86, 86, 1, 4]
    assert lib.zkp_msm_config(2, 64, out) == 1
    assert list(out)[2:4] == [64, 2]  # ceil(127 / 64) = 2 tables in the reference too
    for s, target in ((2, 1), (3, 1), (3, 5), (3, 17)):
        assert lib.zkp_msm_config(s, target, out) == 0, (s, target)
    # the windows in use are unaffected
    for s in range(4, 25):
        assert lib.zkp_msm_config(s, -(-254 // s), out) == 1
        assert out[1] == -(-254 // s) and out[3] == 1


def _poly():
    from zkpoly_amd.poly import load_lib

    return load_lib(Libs())


def _a256(x):
    return (x + 255) // 256 * 256


@pytest.mark.parametrize("n", [1, 2047, 2048, 2049, 1 << 20, (1 << 24) + 5])
def test_poly_buffer_size_queries(n):
    """The NULL-buffer size protocol of core/src/poly.rs:265-280, 364-378,
    413-427, 570-586, 631-647: the size comes back and nothing is launched.
    Scans keep one 32-byte product (64-byte Horner pair for eval / kate) per
    2048-element tile, twice; batched_invert keeps the tile products and the
    products before and after each tile, and the inverse of the total."""
    from zkpoly_amd.poly import KateDivision, PolyEval, PolyInvert, PolyScan

    libs = Libs()
    tiles = -(-n // 2048)
    assert PolyScan(libs).get_buffer_size(n) == 2 * _a256(tiles * 32)
    assert PolyEval(libs).get_buffer_size(n) == _a256(tiles * 64) + 256
    assert PolyInvert(libs).get_buffer_size(n) == 3 * _a256(tiles * 32) + 256
    if n & (n - 1) == 0:
        k = n.bit_length() - 1
        assert KateDivision(libs).get_buffer_size(k) == 2 * _a256(tiles * 64)


def test_poly_entry_points_reject_bad_args_without_launch():
    lib = _poly()
    size = ctypes.c_ulong(0)
    a16, a8 = PolyPtr(0x1000, 16, 0, 0, 16), PolyPtr(0x2000, 8, 0, 0, 8)
    # the result must hold max(|a|, |b|) elements
    assert lib.poly_add(a8, a16, a8, None) == HIP_INVALID_VALUE
    assert lib.poly_sub(a8, a8, a16, None) == HIP_INVALID_VALUE
    assert lib.poly_mul(PolyPtr(0, 16, 0, 0, 16), a16, a8, None) == HIP_INVALID_VALUE
    # bad views: rotation on a slice, slice longer than its array
    assert lib.poly_zero(PolyPtr(0x1000, 16, 3, 5, 32), None) == HIP_INVALID_VALUE
    assert lib.poly_one_coef(PolyPtr(0x1000, 16, 0, 0, 8), None) == HIP_INVALID_VALUE
    assert lib.poly_one_lagrange(PolyPtr(0, 16, 0, 0, 16), None) == HIP_INVALID_VALUE
    # empty operands do nothing
    assert lib.poly_add(PolyPtr(0, 0, 0, 0, 0), PolyPtr(0, 0, 0, 0, 0), PolyPtr(0, 0, 0, 0, 0), None) == 0
    assert lib.poly_zero(PolyPtr(0, 0, 0, 0, 0), None) == 0
    # size queries need somewhere to write the size
    assert lib.scan_mul(None, None, a16, None) == HIP_INVALID_VALUE
    assert lib.batched_invert(None, None, a16, None) == HIP_INVALID_VALUE
    # kate_division: both views must hold 2^log_p elements, b non-null
    assert lib.kate_division(0x9000, None, 4, a8, 0x3000, a16, None) == HIP_INVALID_VALUE
    assert lib.kate_division(0x9000, None, 4, a16, 0x3000, a8, None) == HIP_INVALID_VALUE
    assert lib.kate_division(0x9000, None, 4, a16, None, a16, None) == HIP_INVALID_VALUE
    assert lib.kate_division(0x9000, None, 41, a16, 0x3000, a16, None) == HIP_INVALID_VALUE
    # poly_eval needs res and x
    assert lib.poly_eval(0x9000, None, a16, None, 0x3000, None) == HIP_INVALID_VALUE
    assert lib.poly_eval(0x9000, None, a16, 0x3000, None, None) == HIP_INVALID_VALUE
    assert lib.poly_eval(None, ctypes.byref(size), a16, None, None, None) == 0 and size.value > 0
    # empty scans / inversions launch nothing
    assert lib.scan_mul(0x9000, None, PolyPtr(0, 0, 0, 0, 0), None) == 0
    assert lib.batched_invert(0x9000, None, PolyPtr(0, 0, 0, 0, 0), None) == 0
    assert lib.inv_scalar(None, None) == HIP_INVALID_VALUE
    assert lib.scalar_pow(None, 5, None) == HIP_INVALID_VALUE


def test_poly_operators_refuse_host_arrays():
    """The reference computes host-resident operands on the CPU; this package
    has no CPU compute path and says so."""
    import numpy as np

    from zkpoly_amd.poly import PolyZero, ScalarInv
    from zkpoly_amd.runtime import Scalar

    host = np.zeros((4, 4), dtype=np.uint64)
    with pytest.raises(NotImplementedError):
        PolyZero(Libs()).get_fn().f([ScalarArray(4, host.ctypes.data, device="cpu")], [])
    with pytest.raises(NotImplementedError):
        ScalarInv(Libs()).get_fn().f([Scalar(host.ctypes.data, device="cpu")], [])


@pytest.mark.parametrize("n", [1, 2047, 2048, 2049, (1 << 20) - 25])
def test_permute_buffer_size_and_rejects(n):
    """Sizing as core/src/poly.rs:264-280 calls it: two key buffers for each of
    input and table, the slot and rank flags, the count scan's two per-tile
    arrays and the merge splits (one per 2048-key tile of each array)."""
    from zkpoly_amd.poly import PolyPermute

    tiles = -(-n // 2048)
    want = 4 * _a256(n * 32) + 2 * _a256(n * 4) + 2 * _a256(tiles * 8) + _a256(2 * tiles * 4)
    assert PolyPermute(Libs()).get_buffer_size(n) == want
    lib = _poly()
    whole, part = PolyPtr(0x1000, 64, 0, 0, 64), PolyPtr(0x2000, 60, 0, 0, 60)
    # input and table of one length; results exactly `usable` long; usable <= len
    assert lib.permute(0x9000, None, 60, whole, PolyPtr(0x3000, 63, 0, 0, 63), part, part, None) == HIP_INVALID_VALUE
    assert lib.permute(0x9000, None, 60, whole, whole, whole, part, None) == HIP_INVALID_VALUE
    assert lib.permute(0x9000, None, 64, part, part, whole, whole, None) == HIP_INVALID_VALUE
    assert lib.permute(0x9000, None, 60, whole, whole, PolyPtr(0, 60, 0, 0, 60), part, None) == HIP_INVALID_VALUE
    assert lib.permute(0x9000, None, 1 << 32, whole, whole, part, part, None) == HIP_INVALID_VALUE
    assert lib.permute(None, None, 60, whole, whole, part, part, None) == HIP_INVALID_VALUE
    empty = PolyPtr(0, 0, 0, 0, 0)
    assert lib.permute(0x9000, None, 0, whole, whole, empty, empty, None) == 0


def test_fused_pipelined_rejects_before_any_work():
    """The reference's asserts (core/src/fused_kernels.rs:202, 229-242, 262,
    318-321, 335) become hipErrorInvalidValue returned before a stream exists."""
    from zkpoly_amd.fused import _pipeline_lib

    f = _pipeline_lib(Libs()).zkp_fused_pipelined
    host = (PolyPtr * 1)(PolyPtr(0x1000, 64, 3, 0, 64))
    dev = (ctypes.c_void_p * 3)(0x2000, 0x3000, 0x4000)
    none = (ctypes.c_void_p * 1)()

    def call(host_p=host, len_=64, parts=4, argsz=3 * 40, kernel=0x5000, arg=0x6000):
        return f(kernel, 0, none, none, 0, none, none, 0, None, none, 1, host_p, dev, None, len_, parts, arg, argsz,
                 None, 0)

    assert call(parts=3) == HIP_INVALID_VALUE  # divide_parts must exceed 3
    assert call(len_=66) == HIP_INVALID_VALUE  # len a multiple of divide_parts
    assert call(argsz=3 * 40 - 1) == HIP_INVALID_VALUE  # three sets of one mut var
    assert call(kernel=0) == HIP_INVALID_VALUE
    assert call(arg=0) == HIP_INVALID_VALUE
    sliced = (PolyPtr * 1)(PolyPtr(0x1000, 64, 0, 8, 128))
    assert call(host_p=sliced) == HIP_INVALID_VALUE  # "pipelined fused kernel doesn't support slice"
    short = (PolyPtr * 1)(PolyPtr(0x1000, 32, 0, 0, 32))
    assert call(host_p=short) == HIP_INVALID_VALUE


def test_poly_operators_refuse_small_temporary_buffers():
    """A temporary buffer below the size query is refused on the host, before
    any launch (the library itself trusts the size, as the reference's does)."""
    from zkpoly_amd.poly import PolyEval, PolyInvert
    from zkpoly_amd.runtime import GpuBuffer, Scalar

    arr = ScalarArray(4096, 0x1000)
    inv = PolyInvert(Libs())
    with pytest.raises(AssertionError, match="too small"):
        inv.get_fn().f([GpuBuffer(0x2000, inv.get_buffer_size(4096) - 1), arr], [None])
    ev = PolyEval(Libs())
    with pytest.raises(AssertionError, match="too small"):
        ev.get_fn().f([GpuBuffer(0x2000, 16), Scalar(0x3000)], [arr, Scalar(0x4000), None])


def test_reference_abi_sizing_grows_with_stages_and_sets(libs):
    """The sizing call counts stage_points point buffers, stage_scalars scalar
    buffers and batch_per_run bucket sets."""
    n = 1 << 16
    base = MSM(libs, MsmConfig(window_size=16, target_window=16)).get_buffer_size(n)[0]
    two_sets = MSM(libs, MsmConfig(window_size=16, target_window=16, batch_per_run=2)).get_buffer_size(n)[0]
    staged = MSM(libs, MsmConfig(window_size=16, target_window=16, stage_points=2, stage_scalars=2)).get_buffer_size(n)[0]
    bucket_set = 16 * (1 << 15) * 128
    assert two_sets - base == bucket_set
    assert staged - base == n * 64 + n * 32
