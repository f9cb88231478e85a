// Device helpers included by generated fused kernels (see codegen.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../common/poly_ptr.hpp"
#include "../field/bn254.hpp"

namespace zkp {
namespace fused {

constexpr int THREADS = 256;

// PolyPtr -> index map (iter.cuh:309-327: a rotation becomes offset = len - rotate).
template <class P>
__device__ __forceinline__ PolyView view(const P& p) {
  PolyView v;
  v.ptr = const_cast<uint32_t*>(p.ptr);
  v.len = p.len;
  v.whole_len = p.whole_len ? p.whole_len : p.len;
  v.offset = p.rotate ? p.len - p.rotate : p.offset;
  return v;
}

// grid-stride launches: enough waves to fill 256 CUs several times over
inline unsigned grid_for(uint64_t n) {
  uint64_t g = (n + THREADS - 1) / THREADS;
  const uint64_t cap = 256 * 64;
  return (unsigned)(g < cap ? (g ? g : 1) : cap);
}

// largest k < n with prefix[k] <= g (prefix has n + 1 non-decreasing entries)
__device__ __forceinline__ unsigned instance_of(const unsigned long long* prefix, unsigned n, uint64_t g) {
  unsigned lo = 0, hi = n;  // invariant: prefix[lo] <= g < prefix[hi]
  while (hi - lo > 1) {
    const unsigned mid = (lo + hi) >> 1;
    if (prefix[mid] <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

}  // namespace fused
}  // namespace zkp
