"""ctypes wrapper of the CPU oracle (liboracle.so, built from zkp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the bench's
CPU-baseline leg, and only as the checker.  The product never imports it.

All vectors are numpy uint64 arrays of shape (n, 4) holding Montgomery-form
limbs; points are (n, 8) arrays {x limbs, y limbs} in Fq Montgomery form.
Which reference function each entry restates is listed in zkp_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
FR, FQ = 0, 1

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P, U64, I, U32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32
        for name in ("orc_mul", "orc_add", "orc_sub"):
            getattr(L, name).argtypes = [I, P, P, P]
        for name in ("orc_neg", "orc_inv", "orc_from_canonical", "orc_to_canonical"):
            getattr(L, name).argtypes = [I, P, P]
        L.orc_pow.argtypes = [I, P, P, P, I]
        for name in ("orc_vec_mul", "orc_vec_add", "orc_vec_sub"):
            getattr(L, name).argtypes = [I, P, P, P, U64]
        L.orc_vec_neg.argtypes = [I, P, P, U64]
        L.orc_vec_pow.argtypes = [I, P, P, U64, U64]
        L.orc_best_fft.argtypes = [P, P, U32, I]
        L.orc_naive_dft.argtypes = [P, P, U64, P]
        L.orc_small_field_dft.argtypes = [P, P, U64, U64, U64]
        L.orc_omega.argtypes = [P, U32, I]
        L.orc_distribute_powers.argtypes = [P, U64, P, U64]
        L.orc_twiddles.argtypes = [P, U64, P]
        L.orc_gen_pq_omegas.argtypes = [P, P, U32, U64, P]
        L.orc_g1_add.argtypes = [P, P, P]
        L.orc_g1_double.argtypes = [P, P]
        L.orc_g1_mul_canonical.argtypes = [P, P, P]
        L.orc_g1_mul_scalar.argtypes = [P, P, P]
        L.orc_g1_on_curve.argtypes = [P]
        L.orc_g1_on_curve.restype = I
        L.orc_best_multiexp.argtypes = [P, P, P, U64, I]
        L.orc_naive_msm.argtypes = [P, P, P, U64]
        L.orc_inner_product.argtypes = [P, P, P, U64]
        L.orc_poly_eval.argtypes = [P, P, U64, P]
        L.orc_kate_division.argtypes = [P, P, U64, P]
        L.orc_batch_invert.argtypes = [P, U64]
        L.orc_scan_mul.argtypes = [P, U64]
        L.orc_permute.argtypes = [P, P, P, P, U64]
        _lib = L
    return _lib


def _c(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def fe(x) -> np.ndarray:
    """4-limb uint64 array from a Python int (no Montgomery conversion)."""
    if isinstance(x, np.ndarray):
        return _c(x).reshape(4)
    return np.array([(x >> (64 * i)) & ((1 << 64) - 1) for i in range(4)], dtype=np.uint64)


def to_int(a) -> int:
    a = _c(np.asarray(a)).reshape(-1)
    return sum(int(a[i]) << (64 * i) for i in range(4))


def binop(name: str, fid: int, a, b) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint64)
    a, b = fe(a), fe(b)
    getattr(lib(), name)(fid, _p(out), _p(a), _p(b))
    return out


def mul(fid, a, b):
    return binop("orc_mul", fid, a, b)


def add(fid, a, b):
    return binop("orc_add", fid, a, b)


def sub(fid, a, b):
    return binop("orc_sub", fid, a, b)


def unop(name: str, fid: int, a) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint64)
    a = fe(a)
    getattr(lib(), name)(fid, _p(out), _p(a))
    return out


def neg(fid, a):
    return unop("orc_neg", fid, a)


def inv(fid, a):
    return unop("orc_inv", fid, a)


def from_canonical(fid, a):
    return unop("orc_from_canonical", fid, a)


def to_canonical(fid, a):
    return unop("orc_to_canonical", fid, a)


def pow_(fid, a, e: int) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint64)
    ew = np.array([(e >> (64 * i)) & ((1 << 64) - 1) for i in range(4)], dtype=np.uint64)
    a = fe(a)
    lib().orc_pow(fid, _p(out), _p(a), _p(ew), 4)
    return out


def vec(name: str, fid: int, *arrs) -> np.ndarray:
    arrs = [_c(a).reshape(-1, 4) for a in arrs]
    n = arrs[0].shape[0]
    out = np.zeros((n, 4), dtype=np.uint64)
    getattr(lib(), name)(fid, _p(out), *[_p(a) for a in arrs], n)
    return out


def vec_pow(fid: int, a: np.ndarray, e: int) -> np.ndarray:
    a = _c(a).reshape(-1, 4)
    out = np.zeros_like(a)
    lib().orc_vec_pow(fid, _p(out), _p(a), e, a.shape[0])
    return out


def omega(k: int, inverse: bool = False) -> np.ndarray:
    """Montgomery-form omega_{2^k} (compiler/src/utils.rs:51-77)."""
    out = np.zeros(4, dtype=np.uint64)
    lib().orc_omega(_p(out), k, 1 if inverse else 0)
    return out


def best_fft(a: np.ndarray, omega_mont, log_n: int, threads: int | None = None) -> np.ndarray:
    """Returns best_fft(a, omega) (core/tests/msm_halo2.rs:221-324)."""
    x = _c(a).reshape(-1, 4).copy()
    assert x.shape[0] == 1 << log_n
    w = fe(omega_mont)
    lib().orc_best_fft(_p(x), _p(w), log_n, threads or (os.cpu_count() or 1))
    return x


def naive_dft(a: np.ndarray, omega_mont) -> np.ndarray:
    x = _c(a).reshape(-1, 4)
    out = np.zeros_like(x)
    w = fe(omega_mont)
    lib().orc_naive_dft(_p(out), _p(x), x.shape[0], _p(w))
    return out


def small_field_dft(a: np.ndarray, root: int, p: int) -> np.ndarray:
    x = np.ascontiguousarray(a, dtype=np.uint64)
    out = np.zeros_like(x)
    lib().orc_small_field_dft(_p(out), _p(x), x.shape[0], root, p)
    return out


def distribute_powers(a: np.ndarray, powers: np.ndarray) -> np.ndarray:
    x = _c(a).reshape(-1, 4).copy()
    pw = _c(powers).reshape(-1, 4)
    lib().orc_distribute_powers(_p(x), x.shape[0], _p(pw), pw.shape[0])
    return x


def twiddles(length: int, omega_mont) -> np.ndarray:
    out = np.zeros((length, 4), dtype=np.uint64)
    w = fe(omega_mont)
    lib().orc_twiddles(_p(out), length, _p(w))
    return out


def gen_pq_omegas(pq_deg: int, length: int, unit) -> tuple[np.ndarray, np.ndarray]:
    pq = np.zeros(((1 << pq_deg) >> 1, 4), dtype=np.uint64)
    om = np.zeros((32, 4), dtype=np.uint64)
    u = fe(unit)
    lib().orc_gen_pq_omegas(_p(pq), _p(om), pq_deg, length, _p(u))
    return pq, om


def ntt(a: np.ndarray, log_n: int, threads: int | None = None) -> np.ndarray:
    return best_fft(a, omega(log_n), log_n, threads)


def intt(a: np.ndarray, log_n: int, threads: int | None = None) -> np.ndarray:
    """n^-1 * best_fft(a, omega^-1) (compiler/src/transit/type2/intt_mending.rs:27-59)."""
    y = best_fft(a, omega(log_n, True), log_n, threads)
    n_inv = inv(FR, from_canonical(FR, 1 << log_n))
    return vec("orc_vec_mul", FR, y, np.broadcast_to(n_inv, y.shape).copy())


# ---------------- G1 ----------------


def g1(x_mont, y_mont) -> np.ndarray:
    return np.concatenate([fe(x_mont), fe(y_mont)])


def g1_add(p, q) -> np.ndarray:
    out = np.zeros(8, dtype=np.uint64)
    p, q = _c(p).reshape(8), _c(q).reshape(8)
    lib().orc_g1_add(_p(out), _p(p), _p(q))
    return out


def g1_double(p) -> np.ndarray:
    out = np.zeros(8, dtype=np.uint64)
    p = _c(p).reshape(8)
    lib().orc_g1_double(_p(out), _p(p))
    return out


def g1_mul_canonical(p, k: int) -> np.ndarray:
    out = np.zeros(8, dtype=np.uint64)
    p = _c(p).reshape(8)
    kk = fe(k)
    lib().orc_g1_mul_canonical(_p(out), _p(p), _p(kk))
    return out


def g1_on_curve(p) -> bool:
    p = _c(p).reshape(8)
    return bool(lib().orc_g1_on_curve(_p(p)))


def generator() -> np.ndarray:
    return g1(from_canonical(FQ, 1), from_canonical(FQ, 2))


def best_multiexp(scalars: np.ndarray, bases: np.ndarray, threads: int | None = None) -> np.ndarray:
    """best_multiexp (core/tests/msm_halo2.rs:182-209), affine result."""
    s = _c(scalars).reshape(-1, 4)
    b = _c(bases).reshape(-1, 8)
    assert s.shape[0] == b.shape[0]
    out = np.zeros(8, dtype=np.uint64)
    lib().orc_best_multiexp(_p(out), _p(s), _p(b), s.shape[0], threads or (os.cpu_count() or 1))
    return out


def naive_msm(scalars: np.ndarray, bases: np.ndarray) -> np.ndarray:
    s = _c(scalars).reshape(-1, 4)
    b = _c(bases).reshape(-1, 8)
    out = np.zeros(8, dtype=np.uint64)
    lib().orc_naive_msm(_p(out), _p(s), _p(b), s.shape[0])
    return out


def inner_product(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = _c(a).reshape(-1, 4)
    b = _c(b).reshape(-1, 4)
    out = np.zeros(4, dtype=np.uint64)
    lib().orc_inner_product(_p(out), _p(a), _p(b), a.shape[0])
    return out


def poly_eval(a: np.ndarray, x) -> np.ndarray:
    """sum_i a_i x^i (Horner from the top, core/tests/poly.rs:173-176)."""
    a = _c(a).reshape(-1, 4)
    out = np.zeros(4, dtype=np.uint64)
    lib().orc_poly_eval(_p(out), _p(a), a.shape[0], _p(fe(x)))
    return out


def kate_division(a: np.ndarray, b) -> np.ndarray:
    """halo2_proofs::arithmetic::kate_division; n outputs with q[n-1] = 0."""
    a = _c(a).reshape(-1, 4)
    q = np.zeros_like(a)
    lib().orc_kate_division(_p(q), _p(a), a.shape[0], _p(fe(b)))
    return q


def batch_invert(a: np.ndarray) -> np.ndarray:
    """Element inverses, zero kept (ff batch_invert)."""
    x = _c(a).reshape(-1, 4).copy()
    lib().orc_batch_invert(_p(x), x.shape[0])
    return x


def scan_mul(a: np.ndarray) -> np.ndarray:
    """Inclusive prefix products."""
    x = _c(a).reshape(-1, 4).copy()
    lib().orc_scan_mul(_p(x), x.shape[0])
    return x


def permute(inp: np.ndarray, table: np.ndarray, usable: int) -> tuple[np.ndarray, np.ndarray]:
    """halo2 lookup permutation of the first `usable` rows (core/tests/poly.rs:635-690)."""
    a = _c(inp).reshape(-1, 4)
    t = _c(table).reshape(-1, 4)
    assert a.shape[0] >= usable and t.shape[0] >= usable
    out_in = np.zeros((usable, 4), dtype=np.uint64)
    out_tab = np.zeros((usable, 4), dtype=np.uint64)
    lib().orc_permute(_p(out_in), _p(out_tab), _p(a), _p(t), usable)
    return out_in, out_tab
