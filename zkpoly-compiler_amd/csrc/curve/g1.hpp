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
// is exact for any inputs.  Independent products of a formula are grouped
// (Fq::muln), which interleaves their instruction chains on the device.
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

  // dbl-2008-s-1 (a = 0): U = 2Y, V = U^2, W = U V, S = X V, M = 3 X^2,
  // X3 = M^2 - 2S, Y3 = M (S - X3) - W Y, ZZ3 = V ZZ, ZZZ3 = W ZZZ
  ZKP_HD G1XYZZ dbl() const {
    if (is_identity()) return *this;
    const Fq u = y.dbl();
    Fq p1[2];
    Fq::muln<2>({u, x}, {u, x}, p1);  // V, X^2
    const Fq v = p1[0];
    const Fq m = p1[1].dbl() + p1[1];
    Fq p2[4];
    Fq::muln<4>({u, x, m, v}, {v, v, m, zz}, p2);  // W, S, M^2, ZZ3
    const Fq w = p2[0], s = p2[1];
    G1XYZZ r;
    r.x = p2[2] - s.dbl();
    r.zz = p2[3];
    Fq p3[3];
    Fq::muln<3>({m, w, w}, {s - r.x, y, zzz}, p3);
    r.y = p3[0] - p3[1];
    r.zzz = p3[2];
    return r;
  }

  // mdbl-2008-s-1: 2 * a for an affine a (a != identity)
  ZKP_HD static G1XYZZ dbl_affine(const G1Affine& a) {
    const Fq u = a.y.dbl();
    Fq p1[2];
    Fq::muln<2>({u, a.x}, {u, a.x}, p1);  // V, X^2
    const Fq v = p1[0];
    const Fq m = p1[1].dbl() + p1[1];
    Fq p2[3];
    Fq::muln<3>({u, a.x, m}, {v, v, m}, p2);  // W, S, M^2
    const Fq w = p2[0], s = p2[1];
    G1XYZZ r;
    r.x = p2[2] - s.dbl();
    Fq p3[2];
    Fq::muln<2>({m, w}, {s - r.x, a.y}, p3);
    r.y = p3[0] - p3[1];
    r.zz = v;
    r.zzz = w;
    return r;
  }

  // madd-2008-s: this + a (a affine)
  // U2 = x ZZ, S2 = y ZZZ, P = U2 - X, R = S2 - Y, PP = P^2, PPP = P PP,
  // Q = X PP, X3 = R^2 - PPP - 2Q, Y3 = R (Q - X3) - Y PPP, ZZ3 = ZZ PP,
  // ZZZ3 = ZZZ PPP
  ZKP_HD G1XYZZ add_affine(const G1Affine& a) const {
    if (a.is_identity()) return *this;
    if (is_identity()) return from_affine(a);
    Fq p1[2];
    Fq::muln<2>({a.x, a.y}, {zz, zzz}, p1);  // U2, S2
    const Fq p = p1[0] - x;
    const Fq rr = p1[1] - y;
    if (p.is_zero()) {
      if (rr.is_zero()) return dbl_affine(a);
      return identity();
    }
    Fq p2[2];
    Fq::muln<2>({p, rr}, {p, rr}, p2);  // PP, R^2
    const Fq pp = p2[0];
    Fq p3[3];
    Fq::muln<3>({p, x, zz}, {pp, pp, pp}, p3);  // PPP, Q, ZZ3
    const Fq ppp = p3[0], q = p3[1];
    G1XYZZ r;
    r.x = p2[1] - ppp - q.dbl();
    r.zz = p3[2];
    Fq p4[3];
    Fq::muln<3>({rr, y, zzz}, {q - r.x, ppp, ppp}, p4);
    r.y = p4[0] - p4[1];
    r.zzz = p4[2];
    return r;
  }

  // add-2008-s: this + o
  // U1 = X1 ZZ2, U2 = X2 ZZ1, S1 = Y1 ZZZ2, S2 = Y2 ZZZ1, P = U2 - U1,
  // R = S2 - S1, PP = P^2, PPP = P PP, Q = U1 PP, X3 = R^2 - PPP - 2Q,
  // Y3 = R (Q - X3) - S1 PPP, ZZ3 = ZZ1 ZZ2 PP, ZZZ3 = ZZZ1 ZZZ2 PPP
  ZKP_HD G1XYZZ add(const G1XYZZ& o) const {
    if (o.is_identity()) return *this;
    if (is_identity()) return o;
    Fq p1[4];
    Fq::muln<4>({x, o.x, y, o.y}, {o.zz, zz, o.zzz, zzz}, p1);  // U1, U2, S1, S2
    const Fq u1 = p1[0], s1 = p1[2];
    const Fq p = p1[1] - u1;
    const Fq rr = p1[3] - s1;
    if (p.is_zero()) {
      if (rr.is_zero()) return dbl();
      return identity();
    }
    Fq p2[4];
    Fq::muln<4>({p, rr, zz, zzz}, {p, rr, o.zz, o.zzz}, p2);  // PP, R^2, ZZ1 ZZ2, ZZZ1 ZZZ2
    const Fq pp = p2[0];
    Fq p3[3];
    Fq::muln<3>({p, u1, p2[2]}, {pp, pp, pp}, p3);  // PPP, Q, ZZ3
    const Fq ppp = p3[0], q = p3[1];
    G1XYZZ r;
    r.x = p2[1] - ppp - q.dbl();
    r.zz = p3[2];
    Fq p4[3];
    Fq::muln<3>({rr, s1, p2[3]}, {q - r.x, ppp, ppp}, p4);
    r.y = p4[0] - p4[1];
    r.zzz = p4[2];
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
