"""The MSM's bucket arithmetic on the CPU: Fq29 / G1X29 (csrc/field/fq29.hpp,
csrc/curve/xyzz29.hpp -- Fq in nine 29-bit limbs, Montgomery radix 2^261,
lazily reduced) run through a host driver (fq29_host_check.cpp, built here
with hipcc as host code) and checked with Python integers:

* products of reduced, raw (below 32p), lazy-sum and negated operands equal
  a b mod p, and the exact zero test holds for 2p;
* madd chains over random multiples of G (with negated points, and the
  accumulator's own value or its negation, the doubling and cancelling
  cases) equal the affine sum; so do the general sum of a result with itself,
  its double, and a sum after the packed store/load round trip; every
  intermediate keeps the bound invariant (X < 8p, Y < 4p, ZZ, ZZZ < 2p).

The kernels run the same header code on the device; their results are
checked against the oracle by the GPU MSM tests.
"""
import os
import random
import shutil
import subprocess

import pytest

P = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
R = 1 << 256
RINV = pow(R, -1, P)
G = (1, 2)
HERE = os.path.dirname(os.path.abspath(__file__))


def to_m(x):
    return x * R % P


def from_m(x):
    return x * RINV % P


def hx(x):
    return f"{x:064x}"


def aff_add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    (x1, y1), (x2, y2) = a, b
    if x1 == x2:
        if (y1 + y2) % P == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, -1, P) % P
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, P) % P
    x3 = (lam * lam - x1 - x2) % P
    return x3, (lam * (x1 - x3) - y1) % P


def aff_neg(a):
    return None if a is None else (a[0], (-a[1]) % P)


def mul(k, a):
    r = None
    while k:
        if k & 1:
            r = aff_add(r, a)
        a = aff_add(a, a)
        k >>= 1
    return r


def xyzz_affine(fields):
    X, Y, ZZ, ZZZ = (from_m(int(f, 16)) for f in fields)
    if ZZ == 0:
        return None
    return X * pow(ZZ, -1, P) % P, Y * pow(ZZZ, -1, P) % P


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    hipcc = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    if not os.path.exists(hipcc):
        hipcc = shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc is not installed")
    exe = str(tmp_path_factory.mktemp("fq29") / "fq29_host_check")
    src = os.path.join(HERE, "fq29_host_check.cpp")
    subprocess.run([hipcc, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", src, "-o", exe],
                   check=True, capture_output=True)
    return exe


def run(exe, text):
    r = subprocess.run([exe], input=text, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.split("\n")


def test_products_and_lazy_forms(driver):
    rng = random.Random(29)
    cases = [(rng.randrange(P), rng.randrange(P)) for _ in range(3000)]
    cases += [(0, 0), (1, P - 1), (P - 1, P - 1), (to_m(1), to_m(1))]
    out = run(driver, "".join(f"mul {hx(a)} {hx(b)}\n" for a, b in cases))
    for (am, bm), line in zip(cases, out):
        f = line.split()
        a, b = from_m(am), from_m(bm)
        assert from_m(int(f[0], 16)) == a * b % P
        assert from_m(int(f[1], 16)) == a * b % P
        assert from_m(int(f[2], 16)) == (2 * a + 2 * b) ** 2 % P
        assert from_m(int(f[3], 16)) == (a - 2 * b) % P
        assert from_m(int(f[4], 16)) == -a * b % P
        assert f[5] == "1"
        assert f[6] == ("1" if a * b % P == 0 else "0")
        assert all(int(x, 16) < P for x in f[:5])  # canonical on the way out


def test_madd_chains_and_point_forms(driver):
    rng = random.Random(261)
    pts = [mul(rng.randrange(1, P), G) for _ in range(24)]
    chains, text = [], []
    for _ in range(60):
        steps, acc = [], None
        for _ in range(rng.randrange(1, 24)):
            if acc is not None and rng.random() < 0.15:
                neg = rng.random() < 0.5
                steps.append(f"ACC {int(neg)}")
                acc = aff_add(acc, aff_neg(acc) if neg else acc)
            else:
                a = rng.choice(pts)
                neg = rng.random() < 0.5
                steps.append(f"{hx(to_m(a[0]))} {hx(to_m(a[1]))} {int(neg)}")
                acc = aff_add(acc, aff_neg(a) if neg else a)
        chains.append(acc)
        text.append(f"chain {len(steps)}\n" + "".join(s + "\n" for s in steps))
    out = run(driver, "".join(text))
    for k, acc in enumerate(chains):
        res, twice, dbl, back, ok = out[5 * k: 5 * k + 5]
        assert ok == "1", "an intermediate broke the bound invariant"
        for line, want in ((res, acc), (twice, aff_add(acc, acc)), (dbl, aff_add(acc, acc)),
                           (back, mul(3, acc) if acc else None)):
            f = line.split()
            assert f[4] == "1"
            assert xyzz_affine(f[:4]) == want
