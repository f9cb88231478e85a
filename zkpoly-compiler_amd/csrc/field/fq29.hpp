// BN254 Fq in nine 29-bit limbs with Montgomery radix R' = 2^261: the
// arithmetic of the MSM's buckets (host + device).  The limb layout and the
// bounds are written for any modulus below 2^254 (F29<Params>); the NTT keeps
// the 8 x 32-bit product of bn254.hpp (see the note on F29 below).
//
// Why 29 bits: a 29 x 29-bit limb product is below 2^58, so one column of the
// product-scanning multiplication (at most nine a_i b_j and nine m_i p_j
// terms, plus the carry) stays below 2^64 in a single 64-bit accumulator.
// Every limb product is then one v_mad_u64_u32 whose carry-out is never
// needed, where the 32-bit-limb product of bn254.hpp spends a second
// instruction per limb product folding that carry: 228 VALU instructions per
// product here against 308 there (gfx950 ISA of both, counted), 81 + 81 limb
// products against 64 + 64.
//
// Values are kept lazily reduced, and every bound is a multiple of p:
//  * a product of a < A p and b < B p with A B <= 168 (< 2^261 / p) is below
//    2p, since (a b + m p) / R' < a b / R' + p; its limbs are normalised
//    (< 2^29).  Operand limbs may be up to 2^30 (limb products up to 2^60:
//    9 2^60 + 9 2^58 + carry < 2^64).
//  * a + b is limb-wise, with no carry and no reduction.
//  * sub<K>(a, b) = a + K p - b for b < K p, limb-wise against a copy of K p
//    whose lower limbs are raised by 2^31 (and the next lowered by 4), so no
//    limb goes negative, then normalised.
//  * zero tests are exact for any value below 16 p; reduce_2p brings any
//    normalised value under 2p, canonical() under p.
//
// Conversions: an element x stored in the form of bn254.hpp (x 2^256 mod p,
// canonical, 8 x 32-bit words) read with a 5-bit shift is 32 (x 2^256) =
// x 2^261, this form, below 32 p; a product with K256 = 2^256 mod p maps this
// form back to x 2^256.  Buckets in HBM hold this form packed into 8 x 32-bit
// words (any value below 2^256).
#pragma once
#include "bn254.hpp"

namespace zkp {

// The base field q (the MSM's buckets).
struct Fq29Params {
  static constexpr u32 P[9] = {0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u,
                               0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
  static constexpr u32 INV = 0x04866389u;  // -p^-1 mod 2^29
  static constexpr u32 ONE[9] = {0x157ccc21u, 0x141c2758u, 0x185230d3u, 0x014c0419u, 0x0aa36fb9u,
                                 0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};  // 2^261 mod p
  static constexpr u32 K256[9] = {0x058f0d9du, 0x1aea1c6eu, 0x11c2cf74u, 0x11d651ebu, 0x1462c0a7u,
                                  0x11b7bc3cu, 0x1cbd99bau, 0x183340fbu, 0x000e0a77u};  // 2^256 mod p
};

// The NTT's twiddle product was tried through this form too (x read as is, the
// twiddle 5 bits up, the result packed back): 267 instructions against 295, but
// 171 v_mad_u64_u32 against 136, and the 2^24 transform ran 2.21 ms against
// 2.14 ms -- the multiply-adds, not the instruction count, set its pace.
template <class Params>
struct F29 {
  u32 l[9];

  static constexpr u32 MASK = (1u << 29) - 1;
  static constexpr u32 INV = Params::INV;
  static constexpr u32 P_TOP = Params::P[8];  // p >> 232

  ZKP_HD static F29 zero() {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = 0;
    return r;
  }
  ZKP_HD static F29 konst(const u32 (&c)[9]) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = c[i];
    return r;
  }
  ZKP_HD static F29 one() { return konst(Params::ONE); }

  ZKP_HD bool all_zero() const {
    u32 acc = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) acc |= l[i];
    return acc == 0;
  }

  // (w << SHIFT) for a 256-bit little-endian word array w, SHIFT <= 5.
  template <int SHIFT>
  ZKP_HD static F29 from_words(const u32 (&w)[8]) {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      const int s = 29 * i - SHIFT;  // first source bit of limb i
      u32 v;
      if (s < 0) {
        v = w[0] << (-s);
      } else {
        const int word = s >> 5, off = s & 31;
        const u32 lo = w[word];
        const u32 hi = word + 1 < 8 ? w[word + 1] : 0u;
        v = off ? (lo >> off) | (hi << (32 - off)) : lo;
      }
      r.l[i] = i < 8 ? v & MASK : v;
    }
    return r;
  }

  // Packs normalised limbs of a value below 2^256.
  ZKP_HD void to_words(u32 (&w)[8]) const {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int b = 32 * k, i = b / 29, o = b % 29;
      u64 v = (u64)l[i] >> o;
      int have = 29 - o;
      if (i + 1 < 9) v |= (u64)l[i + 1] << have;
      have += 29;
      if (have < 32 && i + 2 < 9) v |= (u64)l[i + 2] << have;
      w[k] = (u32)v;
    }
  }

  __device__ __forceinline__ static void load_words(const u32* p, u32 (&w)[8]) {
    const uint4* q = reinterpret_cast<const uint4*>(p);
    const uint4 a = q[0], b = q[1];
    w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
    w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
  }
  // an element of this form packed in HBM
  __device__ __forceinline__ static F29 load(const u32* p) {
    u32 w[8];
    load_words(p, w);
    return from_words<0>(w);
  }
  // an element in the form of bn254.hpp (canonical x 2^256), as x 2^261 below 32 p
  __device__ __forceinline__ static F29 load_mont256(const u32* p) {
    u32 w[8];
    load_words(p, w);
    return from_words<5>(w);
  }
  __device__ __forceinline__ void store(u32* p) const {
    u32 w[8];
    to_words(w);
    uint4* q = reinterpret_cast<uint4*>(p);
    q[0] = make_uint4(w[0], w[1], w[2], w[3]);
    q[1] = make_uint4(w[4], w[5], w[6], w[7]);
  }

  // Montgomery product a b 2^-261 (bounds above).
  ZKP_HD F29 operator*(const F29& b) const {
    u32 m[9];
    F29 r;
    u64 acc = 0;
#pragma unroll
    for (int k = 0; k < 17; k++) {
#pragma unroll
      for (int i = 0; i < 9; i++) {
        const int j = k - i;
        if (j >= 0 && j <= 8) acc += (u64)l[i] * b.l[j];
      }
#pragma unroll
      for (int i = 0; i < 9; i++) {
        const int j = k - i;
        if (i < k && j >= 1 && j <= 8) acc += (u64)m[i] * Params::P[j];
      }
      if (k < 9) {
        m[k] = ((u32)acc * INV) & MASK;
        acc += (u64)m[k] * Params::P[0];  // clears the low 29 bits
      } else {
        r.l[k - 9] = (u32)acc & MASK;
      }
      acc >>= 29;
    }
    r.l[8] = (u32)acc;
    return r;
  }

  ZKP_HD F29 operator+(const F29& b) const {
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = l[i] + b.l[i];
    return r;
  }

  // Carries every limb above 29 bits into the next (value unchanged).
  ZKP_HD F29 norm() const {
    F29 r;
    u32 c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const u32 t = l[i] + c;
      r.l[i] = t & MASK;
      c = t >> 29;
    }
    r.l[8] = l[8] + c;
    return r;
  }

  // K p with every lower limb raised by 2^31 and the next lowered by 4.
  template <int K>
  struct Lifted {
    u32 l[9];
    constexpr Lifted() : l{} {
      u64 c = 0;
      for (int i = 0; i < 9; i++) {
        c += (u64)K * Params::P[i];
        l[i] = i < 8 ? (u32)(c & MASK) : (u32)c;
        c >>= 29;
      }
      for (int i = 0; i < 8; i++) {
        l[i] += 1u << 31;
        l[i + 1] -= 4u;
      }
    }
  };

  // this + K p - b, normalised; b < K p with limbs below 2^31 - 4, limbs of this below 2^30.
  template <int K>
  ZKP_HD F29 sub(const F29& b) const {
    constexpr Lifted<K> kp{};
    F29 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.l[i] = l[i] + kp.l[i] - b.l[i];
    return r.norm();
  }

  // x == 0 mod p for normalised limbs and a value below 16 p: if x = q p then
  // the top limb divided by p >> 232 is exactly q.  The low limb settles all
  // but a 2^-29 share of the nonzero values; the rest compare every limb.
  ZKP_HD bool is_zero_modp() const {
    const u32 q = l[8] / P_TOP;
    if (((u32)((u64)q * Params::P[0]) ^ l[0]) & MASK) return false;
    u64 c = 0;
    u32 diff = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      c += (u64)q * Params::P[i];
      diff |= (i < 8 ? ((u32)c & MASK) : (u32)c) ^ l[i];
      c >>= 29;
    }
    return diff == 0;
  }

  // A normalised value below 2^261 brought below 2p: q' = top / (p_top + 1)
  // never exceeds q = floor(x / p), and is at least q - 1 since
  // top > q p_top - 1 and (q + 1) / (p_top + 1) < 1 for every such x.
  ZKP_HD F29 reduce_2p() const {
    const u32 q = l[8] / (P_TOP + 1);
    F29 r;
    int64_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      c += (int64_t)l[i] - (int64_t)((u64)q * Params::P[i]);
      r.l[i] = i < 8 ? (u32)c & MASK : (u32)c;
      c >>= 29;  // arithmetic: the borrow
    }
    return r;
  }

  // The canonical value (below p) of a normalised value.
  ZKP_HD F29 canonical() const {
    const F29 r = reduce_2p();
    F29 t;
    int64_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      c += (int64_t)r.l[i] - (int64_t)Params::P[i];
      t.l[i] = i < 8 ? (u32)c & MASK : (u32)c;
      if (i < 8) c >>= 29;
    }
    return c < 0 ? r : t;
  }

  // x 2^261 (this form, below 16 p: the product by K256 stays below 2p) -> canonical x 2^256,
  // packed: the form of bn254.hpp.
  ZKP_HD void to_mont256(u32 (&w)[8]) const {
    ((*this) * konst(Params::K256)).canonical().to_words(w);
  }
};

using Fq29 = F29<Fq29Params>;

}  // namespace zkp
