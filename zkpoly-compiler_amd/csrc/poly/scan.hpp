// Shared machinery of libpoly.so: the element helpers and the three-launch
// monoid scan (tile fold, one-workgroup scan of the tile folds, seeded tile
// rescan) used by the scans, folds and the permutation.
//
// A monoid Op supplies T, id() and comb(l, u) (l precedes u).  A source
// src(k) yields element k as an Op::T; a sink(k, exclusive, element, inclusive)
// receives the folds of elements [0, k) and [0, k] and element k.  Every workgroup folds TILE =
// TH * PER consecutive elements (PER per thread, then an ordered LDS scan over
// the TH thread folds).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../common/poly_ptr.hpp"
#include "../field/bn254.hpp"

namespace zkp {
namespace poly {

constexpr int TH = 256;
constexpr int PER = 8;
constexpr int TILE = TH * PER;

__device__ __forceinline__ Fr ld(const PolyView& v, uint64_t i) { return Fr::load(v.elem(i)); }
__device__ __forceinline__ void st(const PolyView& v, uint64_t i, const Fr& x) { x.store(v.elem(i)); }

// ---------------------------------------------------------------- scan machinery

// A source may fold a thread's run itself (src.fold(k0, n)), when it knows a
// cheaper way than element-by-element combines.
template <class Src, class = void>
struct has_fold : std::false_type {};
template <class Src>
struct has_fold<Src, std::void_t<decltype(&Src::fold)>> : std::true_type {};

// Fold of elements [k0, min(k0 + PER, n)) in order.
template <class Op, class Src>
__device__ __forceinline__ typename Op::T thread_fold(const Src& src, uint64_t k0, uint64_t n) {
  if constexpr (has_fold<Src>::value) {
    return src.fold(k0, n);
  } else {
    typename Op::T acc = Op::id();
    for (int j = 0; j < PER; j++) {
      const uint64_t k = k0 + j;
      if (k >= n) break;
      acc = Op::comb(acc, src(k));
    }
    return acc;
  }
}

// Exclusive scan of one value per thread in the order of `pos` (a permutation
// of the thread indices; thread order by default), Hillis-Steele in LDS;
// *total = fold of all of them.  Every thread of the block calls it.
template <class Op>
__device__ __forceinline__ typename Op::T block_exclusive(const typename Op::T& v, typename Op::T* sh,
                                                          typename Op::T* total, uint32_t pos) {
  using T = typename Op::T;
  sh[pos] = v;
  __syncthreads();
  for (uint32_t off = 1; off < TH; off <<= 1) {
    T cur = sh[pos];
    if (pos >= off) cur = Op::comb(sh[pos - off], cur);
    __syncthreads();
    sh[pos] = cur;
    __syncthreads();
  }
  const T excl = pos ? sh[pos - 1] : Op::id();
  *total = sh[TH - 1];
  __syncthreads();  // sh is reused by the caller
  return excl;
}

template <class Op>
__device__ __forceinline__ typename Op::T block_exclusive(const typename Op::T& v, typename Op::T* sh,
                                                          typename Op::T* total) {
  return block_exclusive<Op>(v, sh, total, threadIdx.x);
}

// Fold of one value per thread in thread order; the result is valid in thread
// 0.  An in-order tree: at stride `off` the first TH / (2 off) threads combine
// the adjacent ranges [p, p + off) and [p + off, p + 2 off), p = 2 off t, so the
// active threads stay packed into the fewest waves (about 9 wave-level combines
// for 256 threads, against 32 for the exclusive scan).
template <class Op>
__device__ __forceinline__ typename Op::T block_fold(const typename Op::T& v, typename Op::T* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t off = 1; off < (uint32_t)TH; off <<= 1) {
    const uint32_t p = threadIdx.x * 2 * off;
    if (p < (uint32_t)TH) sh[p] = Op::comb(sh[p], sh[p + off]);
    __syncthreads();
  }
  return sh[0];
}

// partial[b] = fold of tile b.
template <class Op, class Src>
__global__ void __launch_bounds__(TH) k_tile_fold(Src src, uint64_t n, typename Op::T* __restrict__ partial) {
  __shared__ typename Op::T sh[TH];
  const uint64_t k0 = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * PER;
  const typename Op::T tot = block_fold<Op>(thread_fold<Op>(src, k0, n), sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// One workgroup: carry[b] = fold of partial[0..b) (when carry is given) and
// *total = fold of all nb partials (when total is given).
template <class Op>
__global__ void __launch_bounds__(TH) k_partials_scan(const typename Op::T* __restrict__ partial, uint64_t nb,
                                                      typename Op::T* __restrict__ carry,
                                                      typename Op::T* __restrict__ total) {
  using T = typename Op::T;
  __shared__ T sh[TH];
  const uint64_t per = (nb + TH - 1) / TH;
  const uint64_t lo = (uint64_t)threadIdx.x * per;
  const uint64_t hi = lo + per < nb ? lo + per : nb;
  T v = Op::id();
  for (uint64_t j = lo; j < hi; j++) v = Op::comb(v, partial[j]);
  T tot;
  T acc = block_exclusive<Op>(v, sh, &tot);
  if (carry) {
    for (uint64_t j = lo; j < hi; j++) {
      carry[j] = acc;
      acc = Op::comb(acc, partial[j]);
    }
  }
  if (total && threadIdx.x == 0) *total = tot;
}

// Rescan of every tile seeded with its carry; sink(k, exclusive, element, inclusive)
// receives the fold of elements [0, k) and element k.
template <class Op, class Src, class Sink>
__global__ void __launch_bounds__(TH)
    k_tile_scan(Src src, Sink sink, uint64_t n, const typename Op::T* __restrict__ carry) {
  using T = typename Op::T;
  __shared__ T sh[TH];
  const uint64_t k0 = (uint64_t)blockIdx.x * TILE + (uint64_t)threadIdx.x * PER;
  T tot;
  const T excl = block_exclusive<Op>(thread_fold<Op>(src, k0, n), sh, &tot);
  T acc = Op::comb(carry[blockIdx.x], excl);
  for (int j = 0; j < PER; j++) {
    const uint64_t k = k0 + j;
    if (k >= n) break;
    const T e = src(k);
    const T next = Op::comb(acc, e);
    sink(k, acc, e, next);
    acc = next;
  }
}

static inline uint64_t tiles_of(uint64_t n) { return (n + TILE - 1) / TILE; }
static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Runs the three phases of a scan of n > 0 elements; partial / carry hold
// tiles_of(n) entries each; total (optional) receives the fold of everything.
template <class Op, class Src, class Sink>
static hipError_t run_scan(const Src& src, const Sink& sink, uint64_t n, typename Op::T* partial,
                           typename Op::T* carry, typename Op::T* total, hipStream_t stream) {
  const uint64_t nb = tiles_of(n);
  if (nb > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_tile_fold<Op, Src>), dim3((unsigned)nb), dim3(TH), 0, stream, src, n, partial);
  hipLaunchKernelGGL((k_partials_scan<Op>), dim3(1), dim3(TH), 0, stream, partial, nb, carry, total);
  hipLaunchKernelGGL((k_tile_scan<Op, Src, Sink>), dim3((unsigned)nb), dim3(TH), 0, stream, src, sink, n, carry);
  return hipGetLastError();
}

static inline unsigned grid_for(uint64_t n) {
  uint64_t g = (n + TH - 1) / TH;
  if (g > 65536) g = 65536;
  return g ? (unsigned)g : 1u;
}

}  // namespace poly
}  // namespace zkp
