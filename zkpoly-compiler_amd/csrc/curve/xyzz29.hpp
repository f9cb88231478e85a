// BN254 G1 XYZZ points over Fq29 (nine 29-bit limbs, radix 2^261): the
// buckets of the MSM (host + device).
//
// Same formulas as G1XYZZ in g1.hpp (EFD "xyzz", a = 0: madd-2008-s,
// add-2008-s, dbl-2008-s-1, with the equal and opposite point cases handled
// exactly), with Fq29's lazy reduction.  Invariant of every point passed
// between operations: X < 8p, Y < 4p, ZZ < 2p, ZZZ < 2p, limbs normalised; the
// identity has every limb zero (ZZ = 0 is the test).  Each formula below keeps
// it, with these bounds (A B <= 168 for every product):
//  madd:  U2 = x ZZ (32 x 2), S2 = y ZZZ; P = U2 + 8p - X < 10p,
//         R = S2 + 4p - Y < 6p; PP = P^2 (100), PPP = P PP, Q = X PP, R^2 (36);
//         X3 = R^2 + 6p - PPP - 2Q < 8p; Y3 = R (Q + 8p - X3) (60) + 2p - Y PPP < 4p
//  add:   U1, U2, S1, S2 < 2p; P, R < 4p; X3 < 8p; Q - X3 < 10p (40); Y3 < 4p
//  dbl:   U = 2Y < 8p; V = U^2 (64), X^2 (64); M = 3 X^2 < 6p; X3 = M^2 + 4p - 2S < 6p;
//         Y3 = M (S + 8p - X3) (60) + 2p - W Y < 4p
// Affine inputs come from HBM in the form of bn254.hpp and are read as
// x 2^261 below 32p (Fq29::load_mont256); they meet only ZZ and ZZZ in
// products (32 x 2), and are brought below 2p (Fq29::reduce_2p) where they
// become a point's X and Y.
// In HBM a point is four packed elements (128 bytes); X is brought below 2p
// first so that every coordinate is below 2^256.
#pragma once
#include "../field/fq29.hpp"

namespace zkp {

struct Affine29 {
  Fq29 x, y;  // below 32p
  bool inf;

  // An affine point of G1Affine's layout (x, y in the form of bn254.hpp, (0, 0) the identity).
  __device__ __forceinline__ static Affine29 load(const u32* p, bool negate) {
    u32 wx[8], wy[8];
    Fq29::load_words(p, wx);
    Fq29::load_words(p + 8, wy);
    u32 any = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) any |= wx[i] | wy[i];
    Affine29 a;
    a.inf = any == 0;
    a.x = Fq29::from_words<5>(wx);
    a.y = Fq29::from_words<5>(wy);
    // -y = 32 p - y (below 32 p; y = 0 only for points of order 2, which G1 has none of)
    if (negate) a.y = Fq29::zero().sub<32>(a.y);
    return a;
  }
};

struct G1X29 {
  Fq29 x, y, zz, zzz;

  ZKP_HD static G1X29 identity() { return G1X29{Fq29::zero(), Fq29::zero(), Fq29::zero(), Fq29::zero()}; }
  ZKP_HD bool is_identity() const { return zz.all_zero(); }

  // (reduce_2p, not a product by one: a run's first point takes this path in
  // a few lanes of the wave while the others wait)
  ZKP_HD static G1X29 from_affine(const Affine29& a) {
    const Fq29 one = Fq29::one();
    return G1X29{a.x.reduce_2p(), a.y.reduce_2p(), one, one};
  }

  __device__ __forceinline__ static G1X29 load(const u32* p) {
    return G1X29{Fq29::load(p), Fq29::load(p + 8), Fq29::load(p + 16), Fq29::load(p + 24)};
  }
  __device__ __forceinline__ void store(u32* p) const {
    x.reduce_2p().store(p);
    y.store(p + 8);
    zz.store(p + 16);
    zzz.store(p + 24);
  }
  // The four coordinates as canonical x 2^256: G1XYZZ's layout and form.
  __device__ __forceinline__ void store_mont256(u32* p) const {
    const Fq29* c[4] = {&x, &y, &zz, &zzz};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      u32 w[8];
      c[k]->to_mont256(w);
      uint4* q = reinterpret_cast<uint4*>(p + 8 * k);
      q[0] = make_uint4(w[0], w[1], w[2], w[3]);
      q[1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
  }

  // dbl-2008-s-1
  ZKP_HD G1X29 dbl() const {
    if (is_identity()) return *this;
    const Fq29 u = y + y;
    const Fq29 v = u * u;
    const Fq29 x2 = x * x;
    const Fq29 m = (x2 + x2 + x2).norm();
    const Fq29 w = u * v;
    const Fq29 s = x * v;
    const Fq29 m2 = m * m;
    G1X29 r;
    r.zz = v * zz;
    r.x = m2.sub<4>(s + s);
    const Fq29 a = m * s.sub<8>(r.x);
    const Fq29 b = w * y;
    r.y = a.sub<2>(b);
    r.zzz = w * zzz;
    return r;
  }

  // madd-2008-s: this + a
  ZKP_HD G1X29 add_affine(const Affine29& a) const {
    if (a.inf) return *this;
    if (is_identity()) return from_affine(a);
    const Fq29 u2 = a.x * zz;
    const Fq29 s2 = a.y * zzz;
    const Fq29 p = u2.sub<8>(x);
    const Fq29 rr = s2.sub<4>(y);
    if (p.is_zero_modp()) {
      if (rr.is_zero_modp()) return from_affine(a).dbl();
      return identity();
    }
    const Fq29 pp = p * p;
    const Fq29 r2 = rr * rr;
    const Fq29 ppp = p * pp;
    const Fq29 q = x * pp;
    G1X29 r;
    r.zz = zz * pp;
    r.x = r2.sub<6>(ppp + q + q);
    const Fq29 t = rr * q.sub<8>(r.x);
    const Fq29 b = y * ppp;
    r.y = t.sub<2>(b);
    r.zzz = zzz * ppp;
    return r;
  }

  // add-2008-s: this + o
  ZKP_HD G1X29 add(const G1X29& o) const {
    if (o.is_identity()) return *this;
    if (is_identity()) return o;
    const Fq29 u1 = x * o.zz;
    const Fq29 u2 = o.x * zz;
    const Fq29 s1 = y * o.zzz;
    const Fq29 s2 = o.y * zzz;
    const Fq29 p = u2.sub<2>(u1);
    const Fq29 rr = s2.sub<2>(s1);
    if (p.is_zero_modp()) {
      if (rr.is_zero_modp()) return dbl();
      return identity();
    }
    const Fq29 pp = p * p;
    const Fq29 r2 = rr * rr;
    const Fq29 zz12 = zz * o.zz;
    const Fq29 zzz12 = zzz * o.zzz;
    const Fq29 ppp = p * pp;
    const Fq29 q = u1 * pp;
    G1X29 r;
    r.zz = zz12 * pp;
    r.x = r2.sub<6>(ppp + q + q);
    const Fq29 t = rr * q.sub<8>(r.x);
    const Fq29 b = s1 * ppp;
    r.y = t.sub<2>(b);
    r.zzz = zzz12 * ppp;
    return r;
  }
};

// ---- quad-cooperative XYZZ arithmetic (device) ----
//
// Where few point operations are left (the last levels of the bucket
// reduction) each one would run on a single lane with its products one after
// another.  Here the four lanes of a quad (q = lane & 3) hold the same
// operands and split every phase of independent products among themselves:
// lane q computes product q and DPP quad permutes hand every product to all
// four lanes, so add-2008-s takes four product latencies (4, 4, 3, 3
// products) instead of 14 and dbl-2008-s-1 three (2, 4, 3) instead of 9.
// Every branch depends only on the shared operands, so the four lanes of a
// quad always run together.
template <int SRC>
__device__ __forceinline__ Fq29 quad29_from(const Fq29& x) {
  Fq29 r;
#pragma unroll
  for (int i = 0; i < 9; i++) r.l[i] = (u32)__builtin_amdgcn_mov_dpp((int)x.l[i], SRC * 0x55, 0xf, 0xf, false);
  return r;
}

// lane q of the quad takes operand q.  The choice is a bit-select on masks of
// q's two bits (v_bfi_b32), kept opaque: written as a select over the operands
// the optimiser folds it into a load of operand[q] from a stack copy.
__device__ __forceinline__ u32 quad29_bfi(u32 m, u32 x1, u32 x0) {
  u32 r;
  asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(x1), "v"(x0));
  return r;
}

__device__ __forceinline__ Fq29 quad29_pick(u32 q, const Fq29& a, const Fq29& b, const Fq29& c, const Fq29& d) {
  const u32 m1 = 0u - (q & 1u), m2 = 0u - ((q >> 1) & 1u);
  Fq29 r;
#pragma unroll
  for (int i = 0; i < 9; i++)
    r.l[i] = quad29_bfi(m2, quad29_bfi(m1, d.l[i], c.l[i]), quad29_bfi(m1, b.l[i], a.l[i]));
  return r;
}

// out[k] = l[k] * r[k] for k < N (N <= 4) on every lane of the quad
template <int N>
__device__ __forceinline__ void quad29_mul(u32 q, const Fq29 (&l)[N], const Fq29 (&r)[N], Fq29 (&out)[N]) {
  const Fq29 a = quad29_pick(q, l[0], l[N > 1 ? 1 : 0], l[N > 2 ? 2 : 0], l[N > 3 ? 3 : 0]);
  const Fq29 b = quad29_pick(q, r[0], r[N > 1 ? 1 : 0], r[N > 2 ? 2 : 0], r[N > 3 ? 3 : 0]);
  const Fq29 m = a * b;
  out[0] = quad29_from<0>(m);
  if (N > 1) out[N > 1 ? 1 : 0] = quad29_from<1>(m);
  if (N > 2) out[N > 2 ? 2 : 0] = quad29_from<2>(m);
  if (N > 3) out[N > 3 ? 3 : 0] = quad29_from<3>(m);
}

__device__ __forceinline__ G1X29 quad29_dbl(const G1X29& a, u32 q) {
  if (a.is_identity()) return a;
  const Fq29 u = a.y + a.y;
  Fq29 p1[2];
  quad29_mul<2>(q, {u, a.x}, {u, a.x}, p1);  // V, X^2
  const Fq29 v = p1[0];
  const Fq29 m = (p1[1] + p1[1] + p1[1]).norm();
  Fq29 p2[4];
  quad29_mul<4>(q, {u, a.x, m, v}, {v, v, m, a.zz}, p2);  // W, S, M^2, ZZ3
  const Fq29 w = p2[0], s = p2[1];
  G1X29 r;
  r.x = p2[2].sub<4>(s + s);
  r.zz = p2[3];
  Fq29 p3[3];
  quad29_mul<3>(q, {m, w, w}, {s.sub<8>(r.x), a.y, a.zzz}, p3);
  r.y = p3[0].sub<2>(p3[1]);
  r.zzz = p3[2];
  return r;
}

__device__ __forceinline__ G1X29 quad29_add(const G1X29& a, const G1X29& o, u32 q) {
  if (o.is_identity()) return a;
  if (a.is_identity()) return o;
  Fq29 p1[4];
  quad29_mul<4>(q, {a.x, o.x, a.y, o.y}, {o.zz, a.zz, o.zzz, a.zzz}, p1);  // U1, U2, S1, S2
  const Fq29 u1 = p1[0], s1 = p1[2];
  const Fq29 p = p1[1].sub<2>(u1);
  const Fq29 rr = p1[3].sub<2>(s1);
  if (p.is_zero_modp()) {
    if (rr.is_zero_modp()) return quad29_dbl(a, q);
    return G1X29::identity();
  }
  Fq29 p2[4];
  quad29_mul<4>(q, {p, rr, a.zz, a.zzz}, {p, rr, o.zz, o.zzz}, p2);  // PP, R^2, ZZ1 ZZ2, ZZZ1 ZZZ2
  const Fq29 pp = p2[0];
  Fq29 p3[3];
  quad29_mul<3>(q, {p, u1, p2[2]}, {pp, pp, pp}, p3);  // PPP, Q, ZZ3
  const Fq29 ppp = p3[0], qq = p3[1];
  G1X29 r;
  r.x = p2[1].sub<6>(ppp + qq + qq);
  r.zz = p3[2];
  Fq29 p4[3];
  quad29_mul<3>(q, {rr, s1, p2[3]}, {qq.sub<8>(r.x), ppp, ppp}, p4);
  r.y = p4[0].sub<2>(p4[1]);
  r.zzz = p4[2];
  return r;
}

}  // namespace zkp
