// BN254 prime fields (Fr scalar field, Fq base field) for gfx950 and host.
//
// Representation: 8 x u32 little-endian limbs holding the canonical Montgomery
// form x*R mod p (R = 2^256), always fully reduced (< p).  This is byte-identical
// to halo2curves' 4 x u64 Montgomery limbs, which is what the reference runtime
// hands to its kernels (core/src/ntt.rs:186, core/src/msm.rs:325; limb layout
// core/src/common/mont/src/field.cuh:637-644).
//
// Multiplication (device): product scanning on 32-bit limbs with interleaved
// Montgomery reduction; every limb product is one v_mad_u64_u32 whose carry-out
// is folded into a third accumulator word.  Host: CIOS on 64-bit limbs with
// 128-bit intermediates.  Both return the canonical representative, so they
// agree bit for bit.  The top limb of both moduli is below 2^30, so the
// running value never exceeds 2p < 2^255.
// Semantics follow the reference's mont::Element (field.cuh:907-1204):
// canonical output, zero-preserving negation, Fermat inversion.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ZKP_HD __host__ __device__ __forceinline__

namespace zkp {

typedef uint32_t u32;
typedef uint64_t u64;

// 32-bit add / subtract with carry through clang's __builtin_addc and
// __builtin_subc: the AMDGPU backend lowers a chain of them to one v_add_co /
// v_addc (v_sub_co / v_subb) per limb, where the same chain written with
// 64-bit sums and shifts costs three to five instructions per limb.
ZKP_HD u32 adc(u32 a, u32 b, u32& carry) { return __builtin_addc(a, b, carry, &carry); }
ZKP_HD u32 sbb(u32 a, u32 b, u32& borrow) { return __builtin_subc(a, b, borrow, &borrow); }

struct FrParams {
  // r = 0x30644e72e131a029b85045b68181585d2833e84879b9709143e1f593f0000001
  static constexpr u32 P[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                               0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  static constexpr u32 INV = 0xefffffffu;  // -r^-1 mod 2^32
  static constexpr u32 ONE[8] = {0x4ffffffbu, 0xac96341cu, 0x9f60cd29u, 0x36fc7695u,
                                 0x7879462eu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
  static constexpr u32 R2[8] = {0xae216da7u, 0x1bb8e645u, 0xe35c59e3u, 0x53fe3ab1u,
                                0x53bb8085u, 0x8c49833du, 0x7f4e44a5u, 0x0216d0b1u};
  static constexpr u32 R3[8] = {0xb4bf0040u, 0x5e94d8e1u, 0x1cfbb6b8u, 0x2a489cbeu,
                                0xa19fcfedu, 0x893cc664u, 0x7fcc657cu, 0x0cf8594bu};  // 2^768 mod r
};

struct FqParams {
  // q = 0x30644e72e131a029b85045b68181585d97816a916871ca8d3c208c16d87cfd47
  static constexpr u32 P[8] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                               0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  static constexpr u32 INV = 0xe4866389u;  // -q^-1 mod 2^32
  static constexpr u32 ONE[8] = {0xc58f0d9du, 0xd35d438du, 0xf5c70b3du, 0x0a78eb28u,
                                 0x7879462cu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
  static constexpr u32 R2[8] = {0x538afa89u, 0xf32cfc5bu, 0xd44501fbu, 0xb5e71911u,
                                0x0a417ff6u, 0x47ab1effu, 0xcab8351fu, 0x06d89f71u};
  static constexpr u32 R3[8] = {0xda1530dfu, 0xb1cd6dafu, 0xa7283db6u, 0x62f210e6u,
                                0x0ada0afbu, 0xef7f0b0cu, 0x2d592544u, 0x20fd6e90u};  // 2^768 mod q
};

template <class Params>
struct Fp {
  static constexpr int LIMBS = 8;
  u32 v[8];

  ZKP_HD static Fp zero() {
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = 0;
    return r;
  }
  ZKP_HD static Fp one() {
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = Params::ONE[i];
    return r;
  }
  ZKP_HD static Fp from_raw(const u32* p) {
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = p[i];
    return r;
  }

  // 2 x 16-byte vector loads / stores (elements are 32-byte aligned in HBM).
  __device__ __forceinline__ static Fp load(const u32* p) {
    const uint4* q = reinterpret_cast<const uint4*>(p);
    uint4 a = q[0], b = q[1];
    Fp r;
    r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
    r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
    return r;
  }
  __device__ __forceinline__ void store(u32* p) const {
    uint4* q = reinterpret_cast<uint4*>(p);
    q[0] = make_uint4(v[0], v[1], v[2], v[3]);
    q[1] = make_uint4(v[4], v[5], v[6], v[7]);
  }

  ZKP_HD bool is_zero() const {
    u32 acc = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) acc |= v[i];
    return acc == 0;
  }
  ZKP_HD bool operator==(const Fp& o) const {
    u32 acc = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) acc |= v[i] ^ o.v[i];
    return acc == 0;
  }
  ZKP_HD bool operator!=(const Fp& o) const { return !(*this == o); }

  // r = x - p if x >= p else x (x < 2p).
  ZKP_HD static Fp reduce_once(const Fp& x) {
    Fp d;
    u32 borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.v[i] = sbb(x.v[i], Params::P[i], borrow);
    // borrow set <=> x < p
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = borrow ? x.v[i] : d.v[i];
    return r;
  }

  ZKP_HD Fp operator+(const Fp& o) const {
    Fp s;
    u32 c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s.v[i] = adc(v[i], o.v[i], c);
    // 2p < 2^255: no carry out of limb 7
    return reduce_once(s);
  }

  ZKP_HD Fp operator-(const Fp& o) const {
    Fp d;
    u32 borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.v[i] = sbb(v[i], o.v[i], borrow);
    // add p back if the subtraction wrapped
    const u32 mask = 0u - borrow;
    u32 c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.v[i] = adc(d.v[i], Params::P[i] & mask, c);
    return d;
  }

  ZKP_HD Fp neg() const {
    // zero stays zero (reference field.cuh:1066-1079)
    Fp d;
    u32 borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.v[i] = sbb(Params::P[i], v[i], borrow);
    u32 nz = is_zero() ? 0u : 0xffffffffu;
#pragma unroll
    for (int i = 0; i < 8; i++) d.v[i] &= nz;
    return d;
  }

  ZKP_HD Fp dbl() const { return *this + *this; }

  // ---- lazy forms for butterflies: values kept in [0, 4p) (4p < 2^256) ----
  // limb i of 2p
  static constexpr u32 p2_limb(int i) { return (Params::P[i] << 1) | (i ? Params::P[i - 1] >> 31 : 0u); }

  // x - 2p if x >= 2p, for x < 4p: result in [0, 2p).
  ZKP_HD static Fp reduce_2p(const Fp& x) {
    Fp d;
    u32 borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.v[i] = sbb(x.v[i], p2_limb(i), borrow);
    Fp r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = borrow ? x.v[i] : d.v[i];
    return r;
  }
  // [0, 4p) -> canonical.
  ZKP_HD static Fp reduce_4p(const Fp& x) { return reduce_once(reduce_2p(x)); }
  // a + b without reduction (the caller keeps the sum below 2^256).
  ZKP_HD static Fp add_raw(const Fp& a, const Fp& b) {
    Fp s;
    u32 c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s.v[i] = adc(a.v[i], b.v[i], c);
    return s;
  }
  // a + 2p - b for a, b < 2p: result in (0, 4p).
  ZKP_HD static Fp sub_2p(const Fp& a, const Fp& b) {
    Fp d;
    u32 c = 0, borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d.v[i] = adc(a.v[i], sbb(p2_limb(i), b.v[i], borrow), c);  // 2p - b >= 0
    return d;
  }
  // Montgomery product without the final subtraction: for a < 4p and b < p
  // the result a*b*R^-1 (mod p) lies in [0, 2p).
  ZKP_HD static Fp mul_lazy(const Fp& a, const Fp& b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return mul_ps_raw(a, b);
#else
    return mul_cios64_raw(a, b);
#endif
  }
  // Two independent lazy products (device: interleaved chains).
  ZKP_HD static void mul_lazy2(const Fp& a0, const Fp& b0, const Fp& a1, const Fp& b1, Fp& r0, Fp& r1) {
#if defined(__HIP_DEVICE_COMPILE__)
    const Fp a[2] = {a0, a1}, b[2] = {b0, b1};
    Fp r[2];
    mul_ps_raw_n<2>(a, b, r);
    r0 = r[0];
    r1 = r[1];
#else
    r0 = mul_cios64_raw(a0, b0);
    r1 = mul_cios64_raw(a1, b1);
#endif
  }
  // N independent canonical products r[n] = a[n] * b[n].
  template <int N>
  ZKP_HD static void muln(const Fp (&a)[N], const Fp (&b)[N], Fp (&r)[N]) {
#if defined(__HIP_DEVICE_COMPILE__)
    mul_ps_raw_n<N>(a, b, r);
#pragma unroll
    for (int n = 0; n < N; n++) r[n] = reduce_once(r[n]);
#else
    for (int n = 0; n < N; n++) r[n] = a[n] * b[n];
#endif
  }
  // Two / three independent canonical products.
  ZKP_HD static void mul2(const Fp& a0, const Fp& b0, const Fp& a1, const Fp& b1, Fp& r0, Fp& r1) {
    mul_lazy2(a0, b0, a1, b1, r0, r1);
    r0 = reduce_once(r0);
    r1 = reduce_once(r1);
  }
  ZKP_HD static void mul3(const Fp& a0, const Fp& b0, const Fp& a1, const Fp& b1, const Fp& a2, const Fp& b2, Fp& r0,
                          Fp& r1, Fp& r2) {
#if defined(__HIP_DEVICE_COMPILE__)
    const Fp a[3] = {a0, a1, a2}, b[3] = {b0, b1, b2};
    Fp r[3];
    mul_ps_raw_n<3>(a, b, r);
    r0 = reduce_once(r[0]);
    r1 = reduce_once(r[1]);
    r2 = reduce_once(r[2]);
#else
    r0 = a0 * b0;
    r1 = a1 * b1;
    r2 = a2 * b2;
#endif
  }

  // Montgomery product a*b*R^-1 mod p.
  //
  // Device: product scanning (column-wise) with interleaved reduction.  Each
  // 32x32 limb product is one v_mad_u64_u32 into a 64-bit column accumulator
  // whose carry-out (VCC) is folded into a third word by one v_addc/v_cndmask,
  // i.e. 2 VALU ops per limb product and no register-pair shuffling (128 limb
  // products + 8 m-digit multiplies per call).
  // Host: CIOS with the same interleaved reduction (identical results; both
  // return the canonical representative).
  ZKP_HD Fp operator*(const Fp& b) const {
#if defined(__HIP_DEVICE_COMPILE__)
    return mul_ps(*this, b);
#else
    return mul_cios64(*this, b);
#endif
  }

#if !defined(__HIP_DEVICE_COMPILE__)
  // -p^-1 mod 2^64 by Newton iteration on the low limb.
  static constexpr u64 inv64() {
    const u64 p0 = (u64)Params::P[0] | ((u64)Params::P[1] << 32);
    u64 x = 1;
    for (int i = 0; i < 6; i++) x *= 2 - p0 * x;  // x = p0^-1 mod 2^64
    return (u64)0 - x;
  }
  // Host CIOS on 4 x 64-bit limbs (same memory layout as 8 x 32).
  static Fp mul_cios64(const Fp& a, const Fp& b) { return reduce_once(mul_cios64_raw(a, b)); }
  static Fp mul_cios64_raw(const Fp& a, const Fp& b) {
    typedef unsigned __int128 u128;
    u64 A[4], B[4], P[4], t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
      A[i] = (u64)a.v[2 * i] | ((u64)a.v[2 * i + 1] << 32);
      B[i] = (u64)b.v[2 * i] | ((u64)b.v[2 * i + 1] << 32);
      P[i] = (u64)Params::P[2 * i] | ((u64)Params::P[2 * i + 1] << 32);
    }
    constexpr u64 INV64 = inv64();
    for (int i = 0; i < 4; i++) {
      u128 c = 0;
      for (int j = 0; j < 4; j++) {
        c += (u128)A[j] * B[i] + t[j];
        t[j] = (u64)c;
        c >>= 64;
      }
      u128 s = (u128)t[4] + c;
      t[4] = (u64)s;
      t[5] = (u64)(s >> 64);
      const u64 m = t[0] * INV64;
      c = ((u128)m * P[0] + t[0]) >> 64;
      for (int j = 1; j < 4; j++) {
        c += (u128)m * P[j] + t[j];
        t[j - 1] = (u64)c;
        c >>= 64;
      }
      s = (u128)t[4] + c;
      t[3] = (u64)s;
      t[4] = t[5] + (u64)(s >> 64);
    }
    Fp r;
    for (int i = 0; i < 4; i++) {
      r.v[2 * i] = (u32)t[i];
      r.v[2 * i + 1] = (u32)(t[i] >> 32);
    }
    return r;
  }
#endif

  ZKP_HD static Fp mul_cios(const Fp& a, const Fp& b) {
    u32 t[8];
#pragma unroll
    for (int i = 0; i < 8; i++) t[i] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const u32 bi = b.v[i];
      u64 A = (u64)a.v[0] * bi + t[0];
      const u32 t0 = (u32)A;
      const u32 m = t0 * Params::INV;
      u64 C = (u64)m * Params::P[0] + t0;
#pragma unroll
      for (int j = 1; j < 8; j++) {
        A = (u64)a.v[j] * bi + t[j] + (A >> 32);
        C = (u64)m * Params::P[j] + (u32)A + (C >> 32);
        t[j - 1] = (u32)C;
      }
      t[7] = (u32)(C >> 32) + (u32)(A >> 32);
    }
    return reduce_once(from_raw(t));
  }

#if defined(__HIP_DEVICE_COMPILE__)
  // acc(64) += a*b; c2 += carry-out of the 64-bit accumulate.
  __device__ __forceinline__ static void mac(u64& acc, u32& c2, u32 a, u32 b) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_addc_co_u32_e32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "+v"(c2)
        : "v"(a), "v"(b)
        : "vcc");
  }
  // first product of a column: c2 = carry-out (starts the column's top word)
  __device__ __forceinline__ static void mac_first(u64& acc, u32& c2, u32 a, u32 b) {
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_cndmask_b32_e64 %1, 0, 1, vcc"
        : "+v"(acc), "=v"(c2)
        : "v"(a), "v"(b)
        : "vcc");
  }
  __device__ __forceinline__ static Fp mul_ps(const Fp& A, const Fp& B) { return reduce_once(mul_ps_raw(A, B)); }
  __device__ __forceinline__ static Fp mul_ps_raw(const Fp& A, const Fp& B) {
    const u32* a = A.v;
    const u32* b = B.v;
    u32 m[8];
    u32 t[8];
    u64 acc = 0;
    u32 c2 = 0;
#pragma unroll
    for (int k = 0; k < 15; k++) {
      bool first = true;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int j = k - i;
        if (j < 0 || j > 7) continue;
        if (first) {
          mac_first(acc, c2, a[i], b[j]);
          first = false;
        } else {
          mac(acc, c2, a[i], b[j]);
        }
      }
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int j = k - i;
        if (i >= k || j < 1 || j > 7) continue;
        mac(acc, c2, m[i], Params::P[j]);
      }
      if (k < 8) {
        m[k] = (u32)acc * Params::INV;
        mac(acc, c2, m[k], Params::P[0]);
      } else {
        t[k - 8] = (u32)acc;
      }
      acc = (acc >> 32) | ((u64)c2 << 32);
    }
    t[7] = (u32)acc;
    return from_raw(t);
  }

  // N independent products r[n] = a[n] * b[n] * R^-1 in [0, 2p), computed in
  // lockstep: the limb products of the N chains alternate, so each chain's
  // dependent multiply-adds issue between the others' and a lone wave is not
  // held by the latency of one chain.
  template <int N>
  __device__ __forceinline__ static void mul_ps_raw_n(const Fp (&A)[N], const Fp (&B)[N], Fp (&R)[N]) {
    u32 m[N][8];
    u32 t[N][8];
    u64 acc[N];
    u32 c2[N];
#pragma unroll
    for (int n = 0; n < N; n++) {
      acc[n] = 0;
      c2[n] = 0;
    }
#pragma unroll
    for (int k = 0; k < 15; k++) {
      bool first = true;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int j = k - i;
        if (j < 0 || j > 7) continue;
#pragma unroll
        for (int n = 0; n < N; n++) {
          if (first)
            mac_first(acc[n], c2[n], A[n].v[i], B[n].v[j]);
          else
            mac(acc[n], c2[n], A[n].v[i], B[n].v[j]);
        }
        first = false;
      }
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int j = k - i;
        if (i >= k || j < 1 || j > 7) continue;
#pragma unroll
        for (int n = 0; n < N; n++) mac(acc[n], c2[n], m[n][i], Params::P[j]);
      }
#pragma unroll
      for (int n = 0; n < N; n++) {
        if (k < 8) {
          m[n][k] = (u32)acc[n] * Params::INV;
        } else {
          t[n][k - 8] = (u32)acc[n];
        }
      }
      if (k < 8) {
#pragma unroll
        for (int n = 0; n < N; n++) mac(acc[n], c2[n], m[n][k], Params::P[0]);
      }
#pragma unroll
      for (int n = 0; n < N; n++) acc[n] = (acc[n] >> 32) | ((u64)c2[n] << 32);
    }
#pragma unroll
    for (int n = 0; n < N; n++) {
      t[n][7] = (u32)acc[n];
      R[n] = from_raw(t[n]);
    }
  }
#endif

  ZKP_HD Fp sqr() const { return (*this) * (*this); }

  // this^e for a 64-bit exponent (square-and-multiply, MSB first).
  ZKP_HD Fp pow(u64 e) const {
    Fp r = one();
    bool started = false;
    for (int bit = 63; bit >= 0; bit--) {
      if (started) r = r.sqr();
      if ((e >> bit) & 1) {
        r = started ? r * (*this) : *this;
        started = true;
      }
    }
    return r;
  }

  // this^e for an exponent of nwords little-endian u32 words.
  ZKP_HD Fp pow_words(const u32* e, int nwords) const {
    Fp r = one();
    for (int w = nwords - 1; w >= 0; w--) {
      for (int bit = 31; bit >= 0; bit--) {
        r = r.sqr();
        if ((e[w] >> bit) & 1) r = r * (*this);
      }
    }
    return r;
  }

  // Inverse by the binary extended Euclidean algorithm, for a lone thread
  // inverting one value (variable time: a few hundred 256-bit shifts and
  // subtractions in place of Fermat's ~380 serial products).  The integer
  // inverse of the Montgomery word aR is a^-1 R^-1; one product with R^3 gives
  // a^-1 R.  Invariants: x1 * w = u and x2 * w = v (mod p), with w the word.
  // Zero maps to zero.  Any 256-bit word is first reduced below p, so the
  // loops always end (gcd(u, p) = 1 for 0 < u < p).
  ZKP_HD Fp inv_binary() const {
    const Fp p = from_raw(Params::P);
    Fp u = *this;
    while (!u.less_than(p)) u.sub_words(p);  // at most 5 times: 2^256 < 6p
    if (u.is_zero()) return zero();
    Fp v = p, x1 = zero(), x2 = zero();
    x1.v[0] = 1;
    while (!u.is_one_word() && !v.is_one_word()) {
      while (!(u.v[0] & 1)) {
        u.shr1();
        x1.half_mod();
      }
      while (!(v.v[0] & 1)) {
        v.shr1();
        x2.half_mod();
      }
      if (!u.less_than(v)) {
        u.sub_words(v);
        x1 = x1 - x2;
      } else {
        v.sub_words(u);
        x2 = x2 - x1;
      }
    }
    return (u.is_one_word() ? x1 : x2) * from_raw(Params::R3);
  }
  ZKP_HD bool is_one_word() const {
    u32 acc = v[0] ^ 1u;
#pragma unroll
    for (int i = 1; i < 8; i++) acc |= v[i];
    return acc == 0;
  }
  ZKP_HD bool less_than(const Fp& o) const {
    u32 borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) (void)sbb(v[i], o.v[i], borrow);
    return borrow != 0;
  }
  ZKP_HD void sub_words(const Fp& o) {
    u32 borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = sbb(v[i], o.v[i], borrow);
  }
  ZKP_HD void shr1() {
#pragma unroll
    for (int i = 0; i < 7; i++) v[i] = (v[i] >> 1) | (v[i + 1] << 31);
    v[7] >>= 1;
  }
  // x / 2 mod p for x < p: (x + p) / 2 when x is odd (x + p < 2p < 2^256)
  ZKP_HD void half_mod() {
    if (v[0] & 1) {
      u32 carry = 0;
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = adc(v[i], Params::P[i], carry);
    }
    shr1();
  }

  // Fermat inverse x^(p-2); inverse of zero is zero.
  ZKP_HD Fp inv() const {
    u32 e[8];
    u64 borrow = 2;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      u64 t = (u64)Params::P[i] - borrow;
      e[i] = (u32)t;
      borrow = (t >> 63) & 1;
    }
    return pow_words(e, 8);
  }

  // canonical integer -> Montgomery form (x * R^2 * R^-1)
  ZKP_HD static Fp from_canonical(const Fp& x) { return x * from_raw(Params::R2); }
  // Montgomery form -> canonical integer (x * 1 * R^-1)
  ZKP_HD Fp to_canonical() const {
    Fp o = zero();
    o.v[0] = 1;
    return (*this) * o;
  }
};

typedef Fp<FrParams> Fr;
typedef Fp<FqParams> Fq;

}  // namespace zkp
