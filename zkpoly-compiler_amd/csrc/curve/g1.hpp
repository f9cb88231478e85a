// BN254 G1 (y^2 = x^3 + 3 over Fq) in affine and XYZZ coordinates, host+device.
//
// Affine points are {x, y} in Fq Montgomery form (64 bytes); the identity is
// (0, 0), as in the reference (core/src/common/curve/src/curve_xyzz.cuh:204-210,
// core/src/common/curve/src/bn254.cuh:11-18).  XYZZ points {X, Y, ZZ, ZZZ}
// represent x = X/ZZ, y = Y/ZZZ with ZZ^3 = ZZZ^2; the identity has ZZ = 0.
// Formulas (Explicit-Formulas Database, "xyzz" coordinates for short
// Weierstrass curves, a = 0): madd-2008-s, add-2008-s, dbl-2008-s-1,
// mdbl-2008-s-1 -- the same set the reference uses (curve_xyzz.cuh:233-419),
// with the equal-point and inverse-point cases handled explicitly so the sum
// is exact for any inputs.
#pragma once
#include "../field/bn254.hpp"

namespace zkp {

struct G1Affine {
  Fq x, y;

  ZKP_HD static G1Affine identity() { return G1Affine{Fq::zero(), Fq::zero()}; }
  ZKP_HD bool is_identity() const { return x.is_zero() && y.is_zero(); }
  ZKP_HD G1Affine neg() const { return G1Affine{x, y.neg()}; }

  __device__ __forceinline__ static G1Affine load(const u32* p) {
    return G1Affine{Fq::load(p), Fq::load(p + 8)};
  }
  __device__ __forceinline__ void store(u32* p) const {
    x.store(p);
    y.store(p + 8);
  }
};

struct G1XYZZ {
  Fq x, y, zz, zzz;

  ZKP_HD static G1XYZZ identity() { return G1XYZZ{Fq::zero(), Fq::zero(), Fq::zero(), Fq::zero()}; }
  ZKP_HD bool is_identity() const { return zz.is_zero(); }

  ZKP_HD static G1XYZZ from_affine(const G1Affine& a) {
    if (a.is_identity()) return identity();
    return G1XYZZ{a.x, a.y, Fq::one(), Fq::one()};
  }

  __device__ __forceinline__ static G1XYZZ load(const u32* p) {
    return G1XYZZ{Fq::load(p), Fq::load(p + 8), Fq::load(p + 16), Fq::load(p + 24)};
  }
  __device__ __forceinline__ void store(u32* p) const {
    x.store(p);
    y.store(p + 8);
    zz.store(p + 16);
    zzz.store(p + 24);
  }

  ZKP_HD G1XYZZ neg() const { return G1XYZZ{x, y.neg(), zz, zzz}; }

  // dbl-2008-s-1 (a = 0)
  ZKP_HD G1XYZZ dbl() const {
    if (is_identity()) return *this;
    const Fq u = y.dbl();
    const Fq v = u.sqr();
    const Fq w = u * v;
    const Fq s = x * v;
    const Fq x2 = x.sqr();
    const Fq m = x2.dbl() + x2;
    G1XYZZ r;
    r.x = m.sqr() - s.dbl();
    r.y = m * (s - r.x) - w * y;
    r.zz = v * zz;
    r.zzz = w * zzz;
    return r;
  }

  // mdbl-2008-s-1: 2 * a for an affine a (a != identity)
  ZKP_HD static G1XYZZ dbl_affine(const G1Affine& a) {
    const Fq u = a.y.dbl();
    const Fq v = u.sqr();
    const Fq w = u * v;
    const Fq s = a.x * v;
    const Fq x2 = a.x.sqr();
    const Fq m = x2.dbl() + x2;
    G1XYZZ r;
    r.x = m.sqr() - s.dbl();
    r.y = m * (s - r.x) - w * a.y;
    r.zz = v;
    r.zzz = w;
    return r;
  }

  // madd-2008-s: this + a (a affine)
  ZKP_HD G1XYZZ add_affine(const G1Affine& a) const {
    if (a.is_identity()) return *this;
    if (is_identity()) return from_affine(a);
    const Fq u2 = a.x * zz;
    const Fq s2 = a.y * zzz;
    const Fq p = u2 - x;
    const Fq rr = s2 - y;
    if (p.is_zero()) {
      if (rr.is_zero()) return dbl_affine(a);
      return identity();
    }
    const Fq pp = p.sqr();
    const Fq ppp = p * pp;
    const Fq q = x * pp;
    G1XYZZ r;
    r.x = rr.sqr() - ppp - q.dbl();
    r.y = rr * (q - r.x) - y * ppp;
    r.zz = zz * pp;
    r.zzz = zzz * ppp;
    return r;
  }

  // add-2008-s: this + o
  ZKP_HD G1XYZZ add(const G1XYZZ& o) const {
    if (o.is_identity()) return *this;
    if (is_identity()) return o;
    const Fq u1 = x * o.zz;
    const Fq u2 = o.x * zz;
    const Fq s1 = y * o.zzz;
    const Fq s2 = o.y * zzz;
    const Fq p = u2 - u1;
    const Fq rr = s2 - s1;
    if (p.is_zero()) {
      if (rr.is_zero()) return dbl();
      return identity();
    }
    const Fq pp = p.sqr();
    const Fq ppp = p * pp;
    const Fq q = u1 * pp;
    G1XYZZ r;
    r.x = rr.sqr() - ppp - q.dbl();
    r.y = rr * (q - r.x) - s1 * ppp;
    r.zz = zz * o.zz * pp;
    r.zzz = zzz * o.zzz * ppp;
    return r;
  }

  // affine normalisation with one inversion: 1/ZZ = (ZZ * (1/ZZZ))^2
  ZKP_HD G1Affine to_affine() const {
    if (is_identity()) return G1Affine::identity();
    const Fq zzz_inv = zzz.inv();
    const Fq t = zz * zzz_inv;
    const Fq zz_inv = t.sqr();
    return G1Affine{x * zz_inv, y * zzz_inv};
  }
};

}  // namespace zkp
