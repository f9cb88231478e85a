// Polynomial view passed across the C ABI, and its device-side index map.
//
// PolyPtr / ConstPolyPtr are the 40-byte #[repr(C)] structs the reference
// runtime builds from a ScalarArray (core/src/common/iter/src/iter.cuh:11-26,
// core/src/poly_ptr.rs:4-86).  Logical element i of the view lives at
//   ptr[(offset + i) % whole_len]
// and a non-zero rotation (allowed only on an unsliced array) is folded into
// offset = len - rotate (iter.cuh:309-327).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

extern "C" {
typedef struct PolyPtr {
  uint32_t* ptr;
  size_t len;
  size_t rotate;
  size_t offset;
  size_t whole_len;
} PolyPtr;

typedef struct ConstPolyPtr {
  const uint32_t* ptr;
  size_t len;
  size_t rotate;
  size_t offset;
  size_t whole_len;
} ConstPolyPtr;
}

namespace zkp {

// Normalised view: rotation folded in, whole_len never zero.  Physical index of
// logical i is offset + i, minus whole_len once if it wraps (offset < whole_len
// and i < len <= whole_len, so one conditional subtraction is exact).
struct PolyView {
  uint32_t* ptr;
  uint64_t len;
  uint64_t offset;
  uint64_t whole_len;

  __host__ __device__ __forceinline__ uint64_t phys(uint64_t i) const {
    uint64_t p = offset + i;
    return p >= whole_len ? p - whole_len : p;
  }
  __host__ __device__ __forceinline__ uint32_t* elem(uint64_t i) const {
    return ptr + phys(i) * 8;
  }
};

// Validates a PolyPtr the way the reference's SliceIterator asserts do
// (iter.cuh:50-55, 311-314) and returns the normalised view.
inline bool make_view(const uint32_t* ptr, size_t len, size_t rotate, size_t offset,
                      size_t whole_len, PolyView* out) {
  if (whole_len == 0) whole_len = len;
  if (len > whole_len) return false;
  if (rotate != 0) {
    if (len != whole_len || offset != 0 || rotate >= len) return false;
    offset = len - rotate;
  }
  if (len > 0 && offset >= whole_len) return false;
  out->ptr = const_cast<uint32_t*>(ptr);
  out->len = len;
  out->offset = offset;
  out->whole_len = whole_len;
  return true;
}

inline bool make_view(const PolyPtr& p, PolyView* out) {
  return make_view(p.ptr, p.len, p.rotate, p.offset, p.whole_len, out);
}
inline bool make_view(const ConstPolyPtr& p, PolyView* out) {
  return make_view(p.ptr, p.len, p.rotate, p.offset, p.whole_len, out);
}

}  // namespace zkp
