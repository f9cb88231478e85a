// libpoly.so -- BN254 Fr polynomial kernels for gfx950.
//
// C ABI of the reference's libpoly.so (core/src/poly/src/poly.h:4-28;
// cudaError_t -> hipError_t, cudaStream_t -> hipStream_t), bound by the
// closures of core/src/poly.rs.  Semantics:
//  * poly_add / poly_sub / poly_mul: element i < min(|a|, |b|) is a op b; the
//    rest of the longer operand is copied (for a - b with |a| < |b|: negated)
//    up to max(|a|, |b|) (poly_basic.cu:4-54, poly_basic.cuh:4-52);
//  * poly_zero / poly_one_coef / poly_one_lagrange: 0..0 / 1,0..0 / 1..1;
//  * inv_scalar (zero stays zero) / scalar_pow on one device element
//    (poly_basic.cu:65-73);
//  * scan_mul: in-place inclusive prefix product (scan_mul.cuh:7-27);
//  * batched_invert: x_i^-1 for x_i != 0, zero stays zero
//    (batched_invert.cuh:49-91), by Montgomery's trick over tiles (a tile
//    fold, one workgroup for the carries and the one inversion, a rescan);
//  * poly_eval: sum_i a_i x^i into a device element (poly_eval.cuh:8-71);
//  * kate_division: q_i = sum_{j>i} p_j b^(j-i-1), i.e. the quotient of
//    (p(X) - p(b)) / (X - b), with q_{n-1} = 0 (kate_division.cuh:55-83).
// Field arithmetic is exact, so any association order gives the reference's
// bytes; the work is organised for the GPU rather than after cub.
//
// Scans and folds run over a monoid in three launches: every workgroup folds a
// tile of TILE = 2048 elements (8 per thread, then an ordered LDS scan over the
// 256 thread folds); one workgroup scans the tile folds into per-tile carries;
// every workgroup rescans its tile seeded with its carry and emits results.
// Monoids: products (scan_mul, batched_invert) and the Horner pair
//   (S, M): (S_l, M_l) o (S_u, M_u) = (S_l + M_l S_u, M_l M_u),
// whose fold over elements (a_i, x) in index order is (sum a_i x^i, x^n):
// poly_eval is its total and kate_division its reverse exclusive scan over
// (p_j, b).  A reverse scan runs forward over mirrored indices with the
// combine's operands swapped.
//
// Temporary buffers follow the reference's protocol: with temp_buf == NULL
// the size is written to *temp_buf_size and nothing runs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "scan.hpp"

namespace zkp {
namespace poly {

// ---------------------------------------------------------------- elementwise

// op: 0 add, 1 sub, 2 mul.  a is the longer operand; `neg` (sub with |a| < |b|
// swapped) means the result is b - a on the common part and -a on the tail.
__global__ void __launch_bounds__(TH)
    k_binary(PolyView r, PolyView a, PolyView b, uint64_t len, uint64_t upper, int op, bool neg) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < upper; i += stride) {
    const Fr x = ld(a, i);
    Fr y;
    if (i >= len) {
      y = (op == 1 && neg) ? x.neg() : x;
    } else {
      const Fr z = ld(b, i);
      y = op == 0 ? x + z : op == 2 ? x * z : (neg ? z - x : x - z);
    }
    st(r, i, y);
  }
}

// mode: 0 all zero, 1 one then zeros, 2 all one
__global__ void __launch_bounds__(TH) k_fill(PolyView r, int mode) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < r.len; i += stride)
    st(r, i, (mode == 2 || (mode == 1 && i == 0)) ? Fr::one() : Fr::zero());
}

__global__ void k_inv_scalar(uint32_t* x) {
  if (threadIdx.x == 0 && blockIdx.x == 0) Fr::load(x).inv_binary().store(x);
}

__global__ void k_pow_scalar(uint32_t* x, uint64_t e) {
  if (threadIdx.x == 0 && blockIdx.x == 0) Fr::load(x).pow(e).store(x);
}

// ---------------------------------------------------------------- monoids

struct ProdOp {
  struct T {
    Fr v;
  };
  __device__ static T id() { return T{Fr::one()}; }
  __device__ static T comb(const T& l, const T& u) { return T{l.v * u.v}; }
};

struct HornerOp {
  struct T {
    Fr s, m;
  };
  __device__ static T id() { return T{Fr::zero(), Fr::one()}; }
  __device__ static T comb(const T& l, const T& u) {
    Fr ms, mm;
    Fr::mul2(l.m, u.s, l.m, u.m, ms, mm);
    return T{l.s + ms, mm};
  }
};

// Reverse order: the same monoid with the operands of the combine swapped.
template <class Op>
struct RevOp {
  using T = typename Op::T;
  __device__ static T id() { return Op::id(); }
  __device__ static T comb(const T& l, const T& u) { return Op::comb(u, l); }
};


// ---------------------------------------------------------------- element sources and sinks

struct SrcProd {  // x_k
  PolyView v;
  __device__ ProdOp::T operator()(uint64_t k) const { return ProdOp::T{ld(v, k)}; }
};
struct SrcNonZero {  // x_k, or 1 for x_k = 0
  PolyView v;
  __device__ ProdOp::T operator()(uint64_t k) const {
    const Fr x = ld(v, k);
    return ProdOp::T{x.is_zero() ? Fr::one() : x};
  }
};
// (sum_t v_{i0+t} x^t, x^cnt) over t < cnt <= PER: Horner from the top (cnt - 1
// products) and x^cnt by squaring, instead of two products per element.
__device__ __forceinline__ HornerOp::T horner_run(const PolyView& v, uint64_t i0, uint32_t cnt, const Fr& x) {
  if (cnt == 0) return HornerOp::id();
  Fr e[PER];
#pragma unroll
  for (int t = 0; t < PER; t++)
    if ((uint32_t)t < cnt) e[t] = ld(v, i0 + t);
  Fr s = e[0];
#pragma unroll
  for (int t = PER - 1; t >= 0; t--) {
    if ((uint32_t)t + 1 == cnt) s = e[t];
    else if ((uint32_t)t + 1 < cnt) s = s * x + e[t];
  }
  Fr m = x, base = x;  // x^cnt, cnt >= 1
  bool started = false;
  for (uint32_t c = cnt; c; c >>= 1) {
    if (c & 1) {
      m = started ? m * base : base;
      started = true;
    }
    if (c >> 1) base = base.sqr();
  }
  return HornerOp::T{s, m};
}

struct SrcHorner {  // (a_k, x)
  PolyView v;
  const uint32_t* x;
  __device__ HornerOp::T operator()(uint64_t k) const { return HornerOp::T{ld(v, k), Fr::load(x)}; }
  // elements k0 .. k0+cnt-1 in order: a contiguous run of a
  __device__ HornerOp::T fold(uint64_t k0, uint64_t n) const {
    const uint32_t cnt = k0 < n ? (uint32_t)(n - k0 < (uint64_t)PER ? n - k0 : PER) : 0;
    return horner_run(v, k0, cnt, Fr::load(x));
  }
};
struct SrcHornerRev {  // (p_{n-1-k}, b)
  PolyView v;
  const uint32_t* b;
  __device__ HornerOp::T operator()(uint64_t k) const { return HornerOp::T{ld(v, v.len - 1 - k), Fr::load(b)}; }
  // under RevOp the run k0 .. k0+cnt-1 folds as the Horner run of
  // p_{n-k0-cnt} .. p_{n-1-k0} in ascending index order
  __device__ HornerOp::T fold(uint64_t k0, uint64_t n) const {
    const uint32_t cnt = k0 < n ? (uint32_t)(n - k0 < (uint64_t)PER ? n - k0 : PER) : 0;
    return horner_run(v, n - k0 - cnt, cnt, Fr::load(b));
  }
};
// scan_mul's rescan: every thread loads its 8 elements once, folds them for
// the block scan, then runs the prefix product from its carry-in.
__global__ void __launch_bounds__(TH) k_scan_mul_tile(PolyView v, uint64_t n, const ProdOp::T* __restrict__ carry) {
  __shared__ ProdOp::T sh[TH];
  const uint64_t k0 = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * PER;
  Fr x[PER];
  Fr t = Fr::one();
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const uint64_t k = k0 + j;
    x[j] = k < n ? ld(v, k) : Fr::one();
    t = j ? t * x[j] : x[j];
  }
  ProdOp::T tot;
  Fr acc = carry[blockIdx.x].v * block_exclusive<ProdOp>(ProdOp::T{t}, sh, &tot).v;
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const uint64_t k = k0 + j;
    if (k < n) {
      acc = acc * x[j];
      st(v, k, acc);
    }
  }
}

// Kate's rescan: thread t of tile b gets the fold of everything above its run
// (the tile carry and the threads before it in mirrored order), whose S part is
// q at the run's top index; below that, q_{i-1} = p_i + b q_i, one product per
// element (the generic rescan would also carry the pair's b^count part).
__global__ void __launch_bounds__(TH) k_kate_tile(PolyView p, PolyView q, const uint32_t* b, uint64_t n,
                                                  const HornerOp::T* __restrict__ carry) {
  using Op = RevOp<HornerOp>;
  __shared__ HornerOp::T sh[TH];
  const SrcHornerRev src{p, b};
  const uint64_t k0 = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * PER;
  HornerOp::T tot;
  const HornerOp::T excl = block_exclusive<Op>(src.fold(k0, n), sh, &tot);
  Fr s = Op::comb(carry[blockIdx.x], excl).s;
  const Fr bb = Fr::load(b);
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const uint64_t k = k0 + j;
    if (k < n) {
      const uint64_t i = n - 1 - k;
      const Fr pi = ld(p, i);  // read before q_i is written: p and q may be one array
      st(q, i, s);
      if (j + 1 < PER) s = pi + bb * s;
    }
  }
}

// ---------------------------------------------------------------- batched inversion
//
// Montgomery's trick over tiles, zeros read as one: after the tile fold of the
// non-zero elements, one workgroup forms for every tile the products of the
// tiles before (pre) and after (suf) it and the inverse of the whole product;
// then every thread of a tile rescans its 8 elements, so that
//   x_i^-1 = (product of non-zero x_j, j < i) * (product of non-zero x_j, j > i) * total^-1.

// One workgroup: pre[b], suf[b] and *inv = (product of all tile folds)^-1.
__global__ void __launch_bounds__(TH) k_invert_carries(const ProdOp::T* __restrict__ partial, uint64_t nb,
                                                       Fr* __restrict__ pre, Fr* __restrict__ suf, Fr* inv) {
  __shared__ ProdOp::T sh[TH];
  const uint64_t per = (nb + TH - 1) / TH;
  const uint64_t lo = (uint64_t)threadIdx.x * per;
  const uint64_t hi = lo + per < nb ? lo + per : nb;
  Fr v = Fr::one();
  for (uint64_t j = lo; j < hi; j++) v = v * partial[j].v;
  ProdOp::T tot;
  Fr acc = block_exclusive<ProdOp>(ProdOp::T{v}, sh, &tot).v;
  for (uint64_t j = lo; j < hi; j++) {
    pre[j] = acc;
    acc = acc * partial[j].v;
  }
  // products commute: scanning the threads in mirrored order gives each the
  // product of the later segments
  acc = block_exclusive<ProdOp>(ProdOp::T{v}, sh, &tot, TH - 1 - threadIdx.x).v;
  for (uint64_t j = hi; j-- > lo;) {
    suf[j] = acc;
    acc = acc * partial[j].v;
  }
  if (threadIdx.x == 0) *inv = tot.v.inv_binary();
}

// Tile rescan: every thread holds its 8 elements (zeros as one), takes the
// products of the threads before and after it within the tile, and forms each
// inverse from its left and right products: 4 products per element.
__global__ void __launch_bounds__(TH) k_invert_tile(PolyView v, uint64_t n, const Fr* __restrict__ pre,
                                                    const Fr* __restrict__ suf, const Fr* __restrict__ inv) {
  __shared__ ProdOp::T sh[TH];
  const uint64_t k0 = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * PER;
  Fr x[PER];
  Fr t = Fr::one();
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const uint64_t k = k0 + j;
    x[j] = k < n ? ld(v, k) : Fr::one();
    if (x[j].is_zero()) x[j] = Fr::one();
    t = j ? t * x[j] : x[j];
  }
  ProdOp::T tot;
  const Fr before = block_exclusive<ProdOp>(ProdOp::T{t}, sh, &tot).v;
  const Fr after = block_exclusive<ProdOp>(ProdOp::T{t}, sh, &tot, TH - 1 - threadIdx.x).v;
  Fr left, right;
  Fr::mul2(pre[blockIdx.x], before, after, suf[blockIdx.x], left, right);
  right = right * (*inv);
  // x[j] <- product of everything after element j
#pragma unroll
  for (int j = PER - 1; j >= 0; j--) {
    const Fr e = x[j];
    x[j] = right;
    if (j) right = right * e;
  }
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const uint64_t k = k0 + j;
    if (k < n) {
      const Fr y = ld(v, k);  // read again: x[j] now holds the right product
      const bool zero = y.is_zero();
      st(v, k, zero ? Fr::zero() : left * x[j]);
      if (j + 1 < PER && !zero) left = left * y;
    }
  }
}

__global__ void k_store_s(const HornerOp::T* total, uint32_t* res) {
  if (threadIdx.x == 0 && blockIdx.x == 0) total->s.store(res);
}


}  // namespace poly
}  // namespace zkp

using namespace zkp;
using namespace zkp::poly;

// ---------------------------------------------------------------- C ABI

static hipError_t binary(PolyPtr r, ConstPolyPtr a, ConstPolyPtr b, hipStream_t stream, int op) {
  PolyView vr, va, vb;
  if (!make_view(r, &vr) || !make_view(a, &va) || !make_view(b, &vb)) return hipErrorInvalidValue;
  bool neg = false;
  if (va.len < vb.len) {
    const PolyView t = va;
    va = vb;
    vb = t;
    neg = true;  // only sub uses it
  }
  const uint64_t len = vb.len, upper = va.len;
  if (upper == 0) return hipSuccess;
  if (vr.len < upper || vr.ptr == nullptr || va.ptr == nullptr || (len && vb.ptr == nullptr))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_binary, dim3(grid_for(upper)), dim3(TH), 0, stream, vr, va, vb, len, upper, op,
                     op == 1 && neg);
  return hipGetLastError();
}

static hipError_t fill(PolyPtr r, hipStream_t stream, int mode) {
  PolyView v;
  if (!make_view(r, &v)) return hipErrorInvalidValue;
  if (v.len == 0) return hipSuccess;
  if (v.ptr == nullptr) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(v.len)), dim3(TH), 0, stream, v, mode);
  return hipGetLastError();
}

// ---------------------------------------------------------------- batched inversion launch

static size_t invert_part_bytes(uint64_t n) { return align256(tiles_of(n ? n : 1) * sizeof(Fr)); }

// partial tile folds, pre, suf (one Fr per tile each), then the inverse of the total
static size_t invert_temp_bytes(uint64_t n) { return 3 * invert_part_bytes(n) + align256(sizeof(Fr)); }

static_assert(sizeof(ProdOp::T) == sizeof(Fr), "tile folds are stored as Fr");

// The three launches of batched_invert over a temporary region t of at least
// invert_temp_bytes(v.len).  Every byte the kernels read is written by an
// earlier launch of the same call, so t needs no clearing.
static hipError_t invert_launch(const PolyView& v, char* t, hipStream_t stream) {
  const uint64_t n = v.len;
  if (n == 0) return hipSuccess;
  if (v.ptr == nullptr) return hipErrorInvalidValue;
  const uint64_t nb = tiles_of(n);
  if (nb > 0x7fffffffull) return hipErrorInvalidValue;
  const size_t part = invert_part_bytes(n);
  auto* partial = reinterpret_cast<ProdOp::T*>(t);
  Fr* pre = reinterpret_cast<Fr*>(t + part);
  Fr* suf = reinterpret_cast<Fr*>(t + 2 * part);
  Fr* inv = reinterpret_cast<Fr*>(t + 3 * part);
  hipLaunchKernelGGL((k_tile_fold<ProdOp, SrcNonZero>), dim3((unsigned)nb), dim3(TH), 0, stream, SrcNonZero{v}, n,
                     partial);
  hipLaunchKernelGGL(k_invert_carries, dim3(1), dim3(TH), 0, stream, partial, nb, pre, suf, inv);
  hipLaunchKernelGGL(k_invert_tile, dim3((unsigned)nb), dim3(TH), 0, stream, v, n, pre, suf, inv);
  return hipGetLastError();
}

// Scratch of zkp_batched_invert_noalloc: room for the temporaries of the
// largest accepted polynomial, ZKP_INVERT_NOALLOC_MAX_LEN elements (2^26:
// 3 x 1 MiB of tile products + one element).  Every GPU the library runs on
// holds its own copy until the process ends.  In host code the name `scratch`
// is only a stand-in for the symbol; its device address per GPU comes from
// hipGetSymbolAddress and is the only form ever passed to a kernel.
#define ZKP_INVERT_NOALLOC_MAX_LEN (1ull << 26)
#define ZKP_INVERT_SCRATCH_BYTES (3 * (ZKP_INVERT_NOALLOC_MAX_LEN / TILE) * 32 + 256)
static __device__ __align__(256) unsigned char scratch[ZKP_INVERT_SCRATCH_BYTES];

static_assert(ZKP_INVERT_NOALLOC_MAX_LEN % TILE == 0, "the maximum is a whole number of tiles");

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[kMaxDevices];  // device address of `scratch` per GPU, once resolved

// The current device's address of `scratch`, resolved on the first call on
// that device (with that device current) and kept for the process lifetime.
static hipError_t invert_scratch(char** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (ptr == nullptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = static_cast<char*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

extern "C" {

// r = a + b (poly.h:4)
hipError_t poly_add(PolyPtr r, ConstPolyPtr a, ConstPolyPtr b, hipStream_t stream) { return binary(r, a, b, stream, 0); }

// r = a - b (poly.h:6)
hipError_t poly_sub(PolyPtr r, ConstPolyPtr a, ConstPolyPtr b, hipStream_t stream) { return binary(r, a, b, stream, 1); }

// r = a * b elementwise (poly.h:8)
hipError_t poly_mul(PolyPtr r, ConstPolyPtr a, ConstPolyPtr b, hipStream_t stream) { return binary(r, a, b, stream, 2); }

// every element one (poly.h:10)
hipError_t poly_one_lagrange(PolyPtr target, hipStream_t stream) { return fill(target, stream, 2); }

// 1, 0, ..., 0 (poly.h:12)
hipError_t poly_one_coef(PolyPtr target, hipStream_t stream) { return fill(target, stream, 1); }

// every element zero (poly.h:14)
hipError_t poly_zero(PolyPtr target, hipStream_t stream) { return fill(target, stream, 0); }

// *res = sum_i poly_i x^i; res and x are device elements (poly.h:16)
hipError_t poly_eval(void* temp_buf, unsigned long* temp_buf_size, ConstPolyPtr poly, unsigned int* res,
                     const unsigned int* x, hipStream_t stream) {
  PolyView v;
  if (!make_view(poly, &v)) return hipErrorInvalidValue;
  const uint64_t nb = tiles_of(v.len ? v.len : 1);
  const size_t part = align256(nb * sizeof(HornerOp::T));
  if (temp_buf == nullptr) {
    if (temp_buf_size == nullptr) return hipErrorInvalidValue;
    *temp_buf_size = part + align256(sizeof(HornerOp::T));
    return hipSuccess;
  }
  if (res == nullptr || x == nullptr) return hipErrorInvalidValue;
  if (v.len == 0) {
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(TH), 0, stream, PolyView{res, 1, 0, 1}, 0);
    return hipGetLastError();
  }
  if (v.ptr == nullptr) return hipErrorInvalidValue;
  auto* partial = reinterpret_cast<HornerOp::T*>(temp_buf);
  auto* total = reinterpret_cast<HornerOp::T*>(reinterpret_cast<char*>(temp_buf) + part);
  const SrcHorner src{v, x};
  hipLaunchKernelGGL((k_tile_fold<HornerOp, SrcHorner>), dim3((unsigned)nb), dim3(TH), 0, stream, src, v.len,
                     partial);
  hipLaunchKernelGGL((k_partials_scan<HornerOp>), dim3(1), dim3(TH), 0, stream, partial, nb,
                     (HornerOp::T*)nullptr, total);
  hipLaunchKernelGGL(k_store_s, dim3(1), dim3(1), 0, stream, total, res);
  return hipGetLastError();
}

// q = (p(X) - p(b)) / (X - b) for p of 2^log_p coefficients, q_{n-1} = 0
// (poly.h:18)
hipError_t kate_division(void* temp_buf, unsigned long* temp_buf_size, unsigned int log_p, ConstPolyPtr p,
                         const unsigned int* b, PolyPtr q, hipStream_t stream) {
  if (log_p > 40) return hipErrorInvalidValue;
  const uint64_t n = 1ull << log_p;
  const uint64_t nb = tiles_of(n);
  const size_t part = align256(nb * sizeof(HornerOp::T));
  if (temp_buf == nullptr) {
    if (temp_buf_size == nullptr) return hipErrorInvalidValue;
    *temp_buf_size = 2 * part;
    return hipSuccess;
  }
  PolyView vp, vq;
  if (!make_view(p, &vp) || !make_view(q, &vq) || vp.len != n || vq.len != n) return hipErrorInvalidValue;
  if (vp.ptr == nullptr || vq.ptr == nullptr || b == nullptr) return hipErrorInvalidValue;
  using Op = RevOp<HornerOp>;
  auto* partial = reinterpret_cast<HornerOp::T*>(temp_buf);
  auto* carry = reinterpret_cast<HornerOp::T*>(reinterpret_cast<char*>(temp_buf) + part);
  if (nb > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_tile_fold<Op, SrcHornerRev>), dim3((unsigned)nb), dim3(TH), 0, stream, SrcHornerRev{vp, b}, n,
                     partial);
  hipLaunchKernelGGL((k_partials_scan<Op>), dim3(1), dim3(TH), 0, stream, partial, nb, carry, (HornerOp::T*)nullptr);
  hipLaunchKernelGGL(k_kate_tile, dim3((unsigned)nb), dim3(TH), 0, stream, vp, vq, b, n, carry);
  return hipGetLastError();
}

// target_i = target_0 * ... * target_i in place (poly.h:20)
hipError_t scan_mul(void* temp_buffer, unsigned long* buffer_size, PolyPtr target, hipStream_t stream) {
  PolyView v;
  if (!make_view(target, &v)) return hipErrorInvalidValue;
  const uint64_t nb = tiles_of(v.len ? v.len : 1);
  const size_t part = align256(nb * sizeof(ProdOp::T));
  if (temp_buffer == nullptr) {
    if (buffer_size == nullptr) return hipErrorInvalidValue;
    *buffer_size = 2 * part;
    return hipSuccess;
  }
  if (v.len == 0) return hipSuccess;
  if (v.ptr == nullptr) return hipErrorInvalidValue;
  auto* partial = reinterpret_cast<ProdOp::T*>(temp_buffer);
  auto* carry = reinterpret_cast<ProdOp::T*>(reinterpret_cast<char*>(temp_buffer) + part);
  if (nb > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_tile_fold<ProdOp, SrcProd>), dim3((unsigned)nb), dim3(TH), 0, stream, SrcProd{v}, v.len,
                     partial);
  hipLaunchKernelGGL((k_partials_scan<ProdOp>), dim3(1), dim3(TH), 0, stream, partial, nb, carry,
                     (ProdOp::T*)nullptr);
  hipLaunchKernelGGL(k_scan_mul_tile, dim3((unsigned)nb), dim3(TH), 0, stream, v, v.len, carry);
  return hipGetLastError();
}

// poly_i = poly_i^-1 in place, zeros kept (poly.h:22)
hipError_t batched_invert(void* temp_buffer, unsigned long* buffer_size, PolyPtr poly, hipStream_t stream) {
  PolyView v;
  if (!make_view(poly, &v)) return hipErrorInvalidValue;
  if (temp_buffer == nullptr) {
    if (buffer_size == nullptr) return hipErrorInvalidValue;
    *buffer_size = invert_temp_bytes(v.len);
    return hipSuccess;
  }
  return invert_launch(v, reinterpret_cast<char*>(temp_buffer), stream);
}

// batched_invert without a temporary buffer: the scratch is the library's
// static device array (one per GPU), so nothing is allocated on any call.
hipError_t zkp_batched_invert_noalloc(PolyPtr poly, hipStream_t stream) {
  PolyView v;
  if (!make_view(poly, &v)) return hipErrorInvalidValue;
  if (v.len > ZKP_INVERT_NOALLOC_MAX_LEN || invert_temp_bytes(v.len) > ZKP_INVERT_SCRATCH_BYTES)
    return hipErrorInvalidValue;
  char* t = nullptr;
  const hipError_t e = invert_scratch(&t);
  if (e != hipSuccess) return e;
  return invert_launch(v, t, stream);
}

// *target = *target^-1, zero stays zero (poly.h:24)
hipError_t inv_scalar(unsigned int* target, hipStream_t stream) {
  if (target == nullptr) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_inv_scalar, dim3(1), dim3(1), 0, stream, target);
  return hipGetLastError();
}

// *target = *target^exp (poly.h:26)
hipError_t scalar_pow(unsigned int* target, unsigned long long exp, hipStream_t stream) {
  if (target == nullptr) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pow_scalar, dim3(1), dim3(1), 0, stream, target, (uint64_t)exp);
  return hipGetLastError();
}

}  // extern "C"
