/* MSM entry points.
 *
 * libmsm_bn254_<window>_<target_windows>_<debug>.so (one per MsmConfig, named
 * as core/src/msm.rs:100-111 builds it) exports msm / msm_precompute of the
 * reference's core/src/msm/src/msm.h:4-11.  They forward to libzkp_msm.so,
 * which also exports the zkp_* entry points below.
 */
#ifndef ZKPOLY_MSM_H
#define ZKPOLY_MSM_H

#include "zkpoly_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* library: libmsm_bn254_<w>_<t>_<dbg>.so */

/* sum_i k_i P_i for `batches` scalar arrays over host point tables, results
 * (affine {x, y}, Fq Montgomery, identity (0,0)) to host.  buffers == NULL:
 * fill buffer_sizes[card] with the device bytes each card needs.  Synchronous.
 * Replaces msm, core/src/msm/src/msm.h:4-9 (bound by core/src/msm.rs:113-128,302-369). */
hipError_t msm(void* const* buffers, unsigned long* buffer_sizes, unsigned long long len, unsigned int batch_per_run,
               unsigned int parts, unsigned int stage_scalers, unsigned int stage_points, unsigned int num_cards,
               const unsigned int* cards, unsigned int const* const* h_points, unsigned int batches,
               unsigned int const* const* h_scaler_batch, unsigned int* const* h_result);

/* h_points[j][i] = 2^(j * n_windows * window) * h_points[0][i] for
 * 1 <= j < n_precompute (host, affine).  Replaces msm_precompute,
 * core/src/msm/src/msm.h:11 (bound by core/src/msm.rs:130-162). */
hipError_t msm_precompute(unsigned long long len, unsigned int* const* h_points, unsigned int max_cards);

/* library: libzkp_msm.so */

/* msm / msm_precompute for a run-time MsmConfig (what the per-config libraries call). */
hipError_t zkp_msm_host(unsigned int window, unsigned int target_windows, int debug_flag, void* const* buffers,
                        unsigned long* buffer_sizes, unsigned long long len, unsigned int batch_per_run,
                        unsigned int parts, unsigned int stage_scalers, unsigned int stage_points,
                        unsigned int num_cards, const unsigned int* cards, unsigned int const* const* h_points,
                        unsigned int batches, unsigned int const* const* h_scaler_batch,
                        unsigned int* const* h_result);
hipError_t zkp_msm_precompute_host(unsigned int window, unsigned int target_windows, unsigned long long len,
                                   unsigned int* const* h_points, unsigned int max_cards);

/* ---- extensions (no counterpart in the reference ABI) ---- */

/* Device workspace bytes for zkp_msm_device (0 if the configuration is invalid). */
unsigned long long zkp_msm_workspace_size(unsigned long long len, unsigned int window, unsigned int target_windows);

/* MSM over device-resident point tables (n_precompute tables of len affine
 * points, table-major) and device scalars; synchronizes `stream` and writes
 * the affine result to host h_result[16]. */
hipError_t zkp_msm_device(const unsigned int* d_points, const unsigned int* d_scalars, unsigned long long len,
                          unsigned int window, unsigned int target_windows, void* d_workspace,
                          unsigned long long workspace_bytes, hipStream_t stream, unsigned int* h_result, int debug);

/* zkp_msm_device without its host step, for batching MSMs into one stream or
 * hipGraph: enqueues the device work on `stream` and, if h_sums is not NULL,
 * an asynchronous copy of the n_windows window sums (128-byte XYZZ points)
 * into pinned host memory.  Allocates nothing and never waits.  Run one MSM
 * on the device before capturing this call into a graph (the first one sets
 * the scatter kernels' LDS limits). */
hipError_t zkp_msm_device_enqueue(const unsigned int* d_points, const unsigned int* d_scalars, unsigned long long len,
                                  unsigned int window, unsigned int target_windows, void* d_workspace,
                                  unsigned long long workspace_bytes, hipStream_t stream, void* h_sums);

/* The affine result (h_result[16]) from window sums zkp_msm_device_enqueue copied out. */
hipError_t zkp_msm_combine(unsigned int window, unsigned int target_windows, const void* h_sums,
                           unsigned int* h_result);

/* d_out[i] = k_i * G, G = (1, 2), for Fr Montgomery-form scalars (device, asynchronous). */
hipError_t zkp_g1_mul_generator(const unsigned int* d_scalars, unsigned long long n, unsigned int* d_out,
                                hipStream_t stream);

/* {window, actual_windows, n_windows, n_precompute, buckets} of a configuration; 0 if invalid. */
int zkp_msm_config(unsigned int window, unsigned int target_windows, unsigned int* out5);

#ifdef __cplusplus
}
#endif

#endif /* ZKPOLY_MSM_H */
