/* Types shared by the zkpoly-compiler_amd C ABI.
 *
 * PolyPtr / ConstPolyPtr: the 40-byte #[repr(C)] view the reference runtime
 * passes for a ScalarArray (core/src/poly_ptr.rs:4-20, mirrored from
 * core/src/common/iter/src/iter.cuh:11-26).  Logical element i lives at
 * ptr[(offset + i) % whole_len] (8 x u32 words per element, Montgomery form);
 * a non-zero rotate (only on an unsliced array) means offset = len - rotate
 * (iter.cuh:309-327).
 *
 * Errors are hipError_t (0 = hipSuccess), the HIP counterpart of the
 * cudaError_t every reference entry point returns (core/src/common/error/src/check.cuh:3-10);
 * precondition violations return hipErrorInvalidValue (1).
 */
#ifndef ZKPOLY_TYPES_H
#define ZKPOLY_TYPES_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

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

#ifdef __cplusplus
}
#endif

#endif /* ZKPOLY_TYPES_H */
