/* Fused elementwise kernels: the code generator and the generated kernel ABI.
 *
 * Generated kernels (one libfused_kernels.so per program, symbol = the
 * FusedKernelMeta name) have the reference's signature,
 * core/src/fused_kernels/kernel_gen.rs:286-290 (bound by
 * core/src/fused_kernels.rs:40-52,97-196):
 *
 *   hipError_t <name>(const ConstPolyPtr* vars, const PolyPtr* mut_vars,
 *                     unsigned long long len, bool is_first,
 *                     void* local_buffer, hipStream_t stream);
 *
 * vars / mut_vars are DEVICE arrays; scalars are 1-element views.
 * local_buffer holds at least <name>_local_buffer_bytes() bytes (the
 * analogue of FusedOp::get_temp_buffer_size, kernel_gen.rs:136-146: the
 * graph's scalar table and register spill slots; 0 for small graphs, which
 * accept any buffer).  Each library also exports <name>_batched (instance k
 * reads vars[k*NV..], mut_vars[k*NM..] and elements [prefix[k],
 * prefix[k+1]); the same local_buffer, last), <name>_arity, <name>_mut_reads
 * (one flag per mut_var: 1 when the kernel reads it) and
 * <name>_local_buffer_bytes.
 */
#ifndef ZKPOLY_FUSED_H
#define ZKPOLY_FUSED_H

#include <stdbool.h>

#include "zkpoly_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef hipError_t (*zkp_fused_fn)(const ConstPolyPtr* vars, const PolyPtr* mut_vars, unsigned long long len,
                                   bool is_first, void* local_buffer, hipStream_t stream);
typedef hipError_t (*zkp_fused_batched_fn)(const ConstPolyPtr* vars, const PolyPtr* mut_vars,
                                           const unsigned long long* prefix, unsigned int instances,
                                           unsigned long long total, bool is_first, hipStream_t stream,
                                           void* local_buffer);
typedef unsigned long long (*zkp_fused_local_bytes_fn)(void);

/* library: libzkp_fused_codegen.so */

/* HIP source of one kernel from the serde JSON of common::arith::ArithGraph
 * (common/src/arith.rs:324-331).  Replaces FusedOp::new + FusedOp::gen,
 * core/src/fused_kernels/kernel_gen.rs:47-191 (called from
 * compiler/src/transit/type3/lowering/kernel_gen.rs:96-158).  0 on success,
 * -1 with a message in err. */
int zkp_fused_codegen(const char* graph_json, const char* name, char* out, size_t out_cap, size_t* out_len,
                      char* err, size_t err_cap);

/* Options of the kernels generated afterwards (process-wide, generator only):
 * field elements held in registers per element before values spill to
 * local_buffer (default 16, at least 3) and the waves per SIMD the compiler
 * must keep (default 2; 0 leaves it to the compiler).  No counterpart in the
 * reference (its register limit is derived inside FusedOp::new,
 * kernel_gen.rs:88-106). */
void zkp_fused_set_options(unsigned int reg_budget, unsigned int waves_per_eu);

/* Generates every kernel into out_dir/<name>.hip and compiles out_dir/libfused_kernels.so
 * (the reference's xmake build of the generated sources, core/src/fused_kernels.rs:76-95). */
int zkp_fused_build(const char* const* graph_jsons, const char* const* names, int count, const char* out_dir,
                    const char* hipcc, const char* include_dir, const char* const* extra_flags, int n_extra,
                    char* err, size_t err_cap);

/* library: libzkp_fused_pipeline.so */

/* Streams a generated kernel over host-resident polynomials in divide_parts
 * (> 3) chunks: three streams (copy in, compute, copy out), two device
 * buffers per read-only polynomial and three per mutable one
 * (dev_polys[j * num_polys + i], dev_mut_polys[j * num_mut_polys + i]), each
 * of len / divide_parts elements; scalars in one-element device buffers,
 * mutable scalars copied back at the end.  mut_poly_read (NULL: all) marks
 * the mutable polynomials the kernel reads; the others are not copied in
 * (<name>_mut_reads of a generated library gives the flags per mut_var).
 * Argument arrays are laid out
 * [scalars..., polys...] per buffer set in arg_buffer, which holds
 * 2 * (num_scalars + num_polys) + 3 * (num_mut_scalars + num_mut_polys)
 * PolyPtr.  Host polynomials are unsliced (rotation allowed).  When every
 * host polynomial is mapped pinned memory (hipHostMalloc) the kernel runs once
 * over the host arrays directly and the chunk buffers go unused.  Synchronous.
 * Replaces PipelinedFusedKernel::get_fn, core/src/fused_kernels.rs:198-516
 * (host code in the reference's runtime). */
hipError_t zkp_fused_pipelined(zkp_fused_fn kernel, unsigned int num_scalars, const unsigned int* const* host_scalars,
                               unsigned int* const* dev_scalars, unsigned int num_mut_scalars,
                               unsigned int* const* host_mut_scalars, unsigned int* const* dev_mut_scalars,
                               unsigned int num_polys, const ConstPolyPtr* host_polys, unsigned int* const* dev_polys,
                               unsigned int num_mut_polys, const PolyPtr* host_mut_polys,
                               unsigned int* const* dev_mut_polys, const unsigned char* mut_poly_read,
                               unsigned long long len, unsigned int divide_parts,
                               void* arg_buffer, size_t arg_buffer_size, void* local_buffer, int device);

#ifdef __cplusplus
}
#endif

#endif /* ZKPOLY_FUSED_H */
