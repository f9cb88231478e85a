/* libpoly.so -- BN254 Fr polynomial kernels (the reference's
 * core/src/poly/src/poly.h:4-28, bound by core/src/poly.rs).
 *
 * All polynomial pointers are device views; asynchronous on `stream`.  Entry
 * points with a temporary buffer follow the reference's protocol: called with
 * temp buffer NULL they write the bytes they need to *size and run nothing;
 * called with a buffer of at least that size they run.  Nothing is allocated.
 * Bad views or pointers return hipErrorInvalidValue without a launch.
 */
#ifndef ZKPOLY_POLY_H
#define ZKPOLY_POLY_H

#include "zkpoly_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* library: libpoly.so */

/* r_i = a_i + b_i for i < min(|a|, |b|), then the longer operand's elements up
 * to max(|a|, |b|) (r holds at least that many).  Replaces poly_add,
 * core/src/poly/src/poly.h:4 (poly_basic.cu:4-18). */
hipError_t poly_add(PolyPtr r, ConstPolyPtr a, ConstPolyPtr b, hipStream_t stream);

/* r_i = a_i - b_i on the common part; then a_i, or -b_i when b is the longer.
 * Replaces poly_sub, poly.h:6 (poly_basic.cu:20-36). */
hipError_t poly_sub(PolyPtr r, ConstPolyPtr a, ConstPolyPtr b, hipStream_t stream);

/* r_i = a_i * b_i on the common part, then the longer operand's elements.
 * Replaces poly_mul, poly.h:8 (poly_basic.cu:38-52). */
hipError_t poly_mul(PolyPtr r, ConstPolyPtr a, ConstPolyPtr b, hipStream_t stream);

/* Every element 1.  Replaces poly_one_lagrange, poly.h:10. */
hipError_t poly_one_lagrange(PolyPtr target, hipStream_t stream);

/* 1, 0, ..., 0.  Replaces poly_one_coef, poly.h:12. */
hipError_t poly_one_coef(PolyPtr target, hipStream_t stream);

/* Every element 0.  Replaces poly_zero, poly.h:14. */
hipError_t poly_zero(PolyPtr target, hipStream_t stream);

/* *res = sum_i poly_i * x^i (res, x: device elements).  Replaces poly_eval,
 * poly.h:16 (poly_eval.cuh:8-71). */
hipError_t poly_eval(void* temp_buf, unsigned long* temp_buf_size, ConstPolyPtr poly, unsigned int* res,
                     const unsigned int* x, hipStream_t stream);

/* q_i = sum_{j>i} p_j b^(j-i-1) for p, q of 2^log_p elements (the quotient of
 * (p(X) - p(b)) / (X - b); q_{n-1} = 0).  Replaces kate_division, poly.h:18
 * (kate_division.cuh:55-83). */
hipError_t kate_division(void* temp_buf, unsigned long* temp_buf_size, unsigned int log_p, ConstPolyPtr p,
                         const unsigned int* b, PolyPtr q, hipStream_t stream);

/* In-place inclusive prefix product.  Replaces scan_mul, poly.h:20
 * (scan_mul.cuh:7-27). */
hipError_t scan_mul(void* temp_buffer, unsigned long* buffer_size, PolyPtr target, hipStream_t stream);

/* In-place inverses; zero elements stay zero.  Replaces batched_invert,
 * poly.h:22 (batched_invert.cuh:49-91). */
hipError_t batched_invert(void* temp_buffer, unsigned long* buffer_size, PolyPtr poly, hipStream_t stream);

/* batched_invert without a temporary buffer: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  The temporaries live in a static
 * device array of the library (about 3 MiB, one copy per GPU, held until the
 * process ends), so poly.len is limited to 2^26 elements; longer views return
 * hipErrorInvalidValue without a launch.  Nothing is allocated.  Calls that
 * overlap on one GPU share that array and race: serialise them (one stream,
 * or events).  Same results as batched_invert. */
hipError_t zkp_batched_invert_noalloc(PolyPtr poly, hipStream_t stream);

/* *target = *target^-1 on the device; zero stays zero.  Replaces inv_scalar,
 * poly.h:24. */
hipError_t inv_scalar(unsigned int* target, hipStream_t stream);

/* *target = *target^exp on the device.  Replaces scalar_pow, poly.h:26. */
hipError_t scalar_pow(unsigned int* target, unsigned long long exp, hipStream_t stream);

/* Lookup permutation of the first `usable` rows: res_input = input sorted by
 * canonical value; res_table holds, at each row that starts a run of equal
 * values, that value (one copy leaves the table multiset), and in the other
 * rows the table's leftover values ascending.  input and table have equal
 * lengths >= usable; res_input and res_table hold usable elements.  A value
 * missing from the table keeps its own row (the reference asserts instead).
 * Replaces permute, poly.h:28 (permute.cuh:84-221). */
hipError_t permute(void* temp_buffer, unsigned long* buffer_size, unsigned long usable, ConstPolyPtr input,
                   ConstPolyPtr table, PolyPtr res_input, PolyPtr res_table, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* ZKPOLY_POLY_H */
