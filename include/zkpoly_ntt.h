/* libntt.so -- BN254 Fr NTT entry points (the reference's core/src/ntt/src/ntt.h:6-17).
 *
 * All data pointers are device pointers unless stated; asynchronous entry
 * points run on `stream` and allocate nothing.
 */
#ifndef ZKPOLY_NTT_H
#define ZKPOLY_NTT_H

#include "zkpoly_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* library: libntt.so */

/* In-place natural-order forward NTT of x (x.len = 2^log_len) with the twiddle
 * table [1, w, ..., w^(len/2 - 1)].  Replaces ssip_ntt, core/src/ntt/src/ntt.h:6
 * (bound by core/src/ntt.rs:24-37,165-200).  The two max_threads hints are
 * accepted and ignored. */
hipError_t ssip_ntt(PolyPtr x, const unsigned int* twiddle, unsigned int log_len, hipStream_t stream,
                    const unsigned int max_threads_stage1_log, const unsigned int max_threads_stage2_log);

/* twiddle[i] = unit^i for i < len, into HOST memory; synchronous.  Replaces
 * ssip_precompute, core/src/ntt/src/ntt.h:9 (bound by core/src/ntt.rs:39-50,204-238). */
hipError_t ssip_precompute(unsigned int* twiddle, unsigned int len, const unsigned int* unit);

/* x[i] *= powers[i mod power_num].  Replaces distribute_powers,
 * core/src/ntt/src/ntt.h:17 (bound by core/src/ntt.rs:69-80,112-150). */
hipError_t distribute_powers(PolyPtr poly, const unsigned int* powers, unsigned long long power_num,
                             hipStream_t stream);

/* In-place forward NTT of x whose twiddles are recomputed from pq
 * (2^(pq_deg - 1) entries) and the 32-entry ladder omegas (device arrays from
 * gen_pq_omegas); no n/2 table.  Bit-identical to ssip_ntt with the table of
 * omegas[0].  Replaces recompute_ntt, core/src/ntt/src/ntt.h:12 (bound by
 * core/src/ntt.rs:249-289).  log_len <= 32. */
hipError_t recompute_ntt(PolyPtr x, const unsigned int* pq, unsigned int pq_deg, const unsigned int* omegas,
                         unsigned int log_len, hipStream_t stream, const unsigned int max_threads_stage1_log,
                         const unsigned int max_threads_stage2_log);

/* Host arrays: pq[i] = (unit^(len >> pq_deg))^i for i < 2^(pq_deg - 1), and
 * omegas[i] = unit^(2^i) for i < 32; synchronous, no GPU work.  Replaces
 * gen_pq_omegas, core/src/ntt/src/ntt.h:15 (bound by core/src/ntt.rs:337-371). */
void gen_pq_omegas(unsigned int* pq, unsigned int* omegas, unsigned int pq_deg, unsigned int len,
                   unsigned int* unit);

/* ---- extensions (no counterpart in the reference ABI) ---- */

/* out[i] = base^i for i < len (device), asynchronous: coset power tables and
 * twiddle tables built directly in HBM. */
hipError_t zkp_gen_powers(unsigned int* out, unsigned long long len, const unsigned int* base, hipStream_t stream);

/* ssip_ntt that also records events[0] before the first kernel and
 * events[i + 1] after kernel i; kinds[i] = 0 (bit reversal, fused with the first stages by default) or 1 (butterfly
 * pass); *num_kernels = kernels launched.  Measurement hook. */
hipError_t zkp_ntt_with_events(PolyPtr x, const unsigned int* twiddle, unsigned int log_len, hipStream_t stream,
                               hipEvent_t* events, int max_events, int* kinds, int* num_kernels);

/* Plans `count` independent in-place transforms (device xs[i], twiddles[i],
 * length 2^log_lens[i]) to run as one bit-reversal launch plus one launch per
 * pass round for the whole batch; allocates the device work lists. */
hipError_t zkp_ntt_batch_create(const PolyPtr* xs, const unsigned int* const* twiddles, const unsigned int* log_lens,
                                unsigned int count, void** plan_out);

/* Launches a planned batch on `stream` (kernel launches only: capturable in a hipGraph). */
hipError_t zkp_ntt_batch_run(void* plan, hipStream_t stream);

/* Frees a plan after the work using it has completed. */
void zkp_ntt_batch_destroy(void* plan);

/* Transform variant for later calls and batch plans in this process: 0 = bit
 * reversal fused with the first stages (persistent workgroups) + radix-4 passes
 * (default), 1 = plain bit reversal + radix-2 passes, 2 = as 0 with one tile
 * pair per head workgroup (batch plans treat 2 as 0).  Returns the previous
 * variant, -1 if unknown. */
int zkp_ntt_set_variant(int variant);

/* Autotuner hook: passes of 2^log_n transforms read runs of >= 2^min_col_log
 * adjacent elements (smaller: fewer, deeper passes).  0, or -1 if out of range. */
int zkp_ntt_set_tuning(unsigned int log_n, unsigned int min_col_log);

/* Kernels one transform of 2^log_n launches with the current settings. */
int zkp_ntt_kernel_count(unsigned int log_n);

/* dst[i] = src[i] for n16 16-byte words: plain copy used to calibrate HBM counters. */
hipError_t zkp_copy_u4(void* dst, const void* src, unsigned long long n16, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* ZKPOLY_NTT_H */
