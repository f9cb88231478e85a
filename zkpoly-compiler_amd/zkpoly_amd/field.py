"""BN254 constants and host-side element encoding.

Elements travel as 4 x u64 little-endian limbs in Montgomery form (x * 2^256
mod p), byte-identical to the reference's 8 x u32 layout
(core/src/common/mont/src/field.cuh:637-644).  Constants: Fr modulus
core/src/common/mont/src/bn254_fr.cuh:21, Fq modulus bn254_fq.cuh:13, curve
y^2 = x^3 + 3 core/src/common/curve/src/bn254.cuh:11-18, root-of-unity choice
compiler/src/utils.rs:51-77.

Synthetic inputs: numpy's PCG64 generator (stable across platforms and numpy
versions) seeded per workload; each element is four 64-bit draws, masked to
254 bits and redrawn while >= p, and the value is used directly as the
Montgomery representation -- a uniform canonical representative is a uniform
field element in Montgomery form.
"""
from __future__ import annotations

import numpy as np

FR_MODULUS = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
FQ_MODULUS = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
R_MONT = 1 << 256
FR_TWO_ADICITY = 28
FR_GENERATOR = 7
# 7^((r-1)/2^28), canonical
FR_ROOT_OF_UNITY = 0x03DDB9F5166D18B798865EA93DD31F743215CF6DD39329C8D34F1ED960C37C9C
G1_GENERATOR = (1, 2)
G1_B = 3

_M64 = (1 << 64) - 1


def to_limbs(x: int) -> np.ndarray:
    return np.array([(x >> (64 * i)) & _M64 for i in range(4)], dtype=np.uint64)


def from_limbs(a) -> int:
    a = np.asarray(a, dtype=np.uint64).reshape(-1)
    return sum(int(a[i]) << (64 * i) for i in range(4))


def to_mont(x: int, p: int = FR_MODULUS) -> int:
    return (x % p) * R_MONT % p


def from_mont(x: int, p: int = FR_MODULUS) -> int:
    return x * pow(R_MONT, -1, p) % p


def mont_limbs(x: int, p: int = FR_MODULUS) -> np.ndarray:
    return to_limbs(to_mont(x, p))


def array_to_ints(arr: np.ndarray) -> list[int]:
    a = np.asarray(arr, dtype=np.uint64).reshape(-1, 4)
    return [int(r[0]) | (int(r[1]) << 64) | (int(r[2]) << 128) | (int(r[3]) << 192) for r in a]


def ints_to_array(xs) -> np.ndarray:
    out = np.empty((len(xs), 4), dtype=np.uint64)
    for i, x in enumerate(xs):
        out[i] = to_limbs(x)
    return out


def omega(k: int, inverse: bool = False) -> int:
    """Canonical primitive 2^k-th root of unity ROOT_OF_UNITY^(2^(28-k))."""
    if not 0 <= k <= FR_TWO_ADICITY:
        raise ValueError("k out of range")
    w = pow(FR_ROOT_OF_UNITY, 1 << (FR_TWO_ADICITY - k), FR_MODULUS)
    return pow(w, -1, FR_MODULUS) if inverse else w


def _lt_modulus(rows: np.ndarray, p: int) -> np.ndarray:
    """Row-wise rows < p for (n, 4) uint64 limb rows."""
    pl = to_limbs(p)
    lt = np.zeros(rows.shape[0], dtype=bool)
    eq = np.ones(rows.shape[0], dtype=bool)
    for i in (3, 2, 1, 0):
        lt |= eq & (rows[:, i] < pl[i])
        eq &= rows[:, i] == pl[i]
    return lt


def random_elements(n: int, seed: int, p: int = FR_MODULUS) -> np.ndarray:
    """n uniform elements of F_p as (n, 4) uint64 Montgomery-form limbs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = rng.integers(0, 1 << 64, size=(n, 4), dtype=np.uint64, endpoint=False)
    out[:, 3] &= np.uint64((1 << 62) - 1)
    bad = ~_lt_modulus(out, p)
    while bad.any():
        k = int(bad.sum())
        fresh = rng.integers(0, 1 << 64, size=(k, 4), dtype=np.uint64, endpoint=False)
        fresh[:, 3] &= np.uint64((1 << 62) - 1)
        out[bad] = fresh
        bad = ~_lt_modulus(out, p)
    return out
