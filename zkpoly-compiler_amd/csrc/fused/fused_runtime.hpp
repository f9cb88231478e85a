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

// grid of a kernel that keeps per-thread spill slots: at most `cap` workgroups
inline unsigned grid_capped(uint64_t n, unsigned cap) {
  const unsigned g = grid_for(n);
  return g < cap ? g : cap;
}

// Spill slots of generated kernels: slot s of a thread is two 16-byte planes,
// each plane lane-contiguous over the `threads` threads of the kernel's capped
// grid, so a wave's spill or reload is one coalesced 1 KiB access per plane.
// `base` is the slot area already offset by the thread's global index.
__device__ __forceinline__ void spill_store(uint4* base, unsigned s, uint64_t threads, const Fr& x) {
  base[(2ull * s) * threads] = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
  base[(2ull * s + 1) * threads] = make_uint4(x.v[4], x.v[5], x.v[6], x.v[7]);
}
__device__ __forceinline__ Fr spill_load(const uint4* base, unsigned s, uint64_t threads) {
  const uint4 a = base[(2ull * s) * threads], b = base[(2ull * s + 1) * threads];
  Fr r;
  r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
  r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
  return r;
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
