/* Continuous-batching server for NTT / MSM proof-step requests on one GPU.
 *
 * No counterpart in the reference ABI: it is the proof-step analogue of the
 * reference's multi-request scheduler (scheduler/src/scheduler.rs:215-310,
 * scheduler/src/scheduler/core.rs:197-383), admitting work by HBM footprint
 * instead of host memory.
 */
#ifndef ZKPOLY_SERVE_H
#define ZKPOLY_SERVE_H

#include "zkpoly_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ZkpServeConfig {
  unsigned int slots;                        /* worker threads, one HIP stream each */
  unsigned long long hbm_budget;             /* bytes of request memory */
  unsigned int ntt_min_log, ntt_max_log;     /* accepted NTT sizes; twiddles resident */
  unsigned int msm_max_log;                  /* resident MSM base table: 2^msm_max_log points */
  unsigned int msm_window;
  const unsigned int* msm_bases_host;        /* affine points, host */
  const unsigned int* const* twiddles_host;  /* [log - ntt_min_log] -> 2^(log-1) elements, host */
} ZkpServeConfig;

typedef struct ZkpServeRequest {
  unsigned int kind;   /* 0: forward NTT of 2^log_n elements, 1: MSM of 2^log_n scalars */
  unsigned int log_n;
  double arrival_s;    /* release time, seconds after zkp_serve_run starts */
  const unsigned int* input;  /* host */
  unsigned int* output;       /* host: 2^log_n elements (NTT) or one affine point (MSM) */
  double start_s;      /* filled: admission time */
  double finish_s;     /* filled: result on the host */
  int status;          /* filled: hipError_t */
} ZkpServeRequest;

/* library: libzkp_serve.so */

hipError_t zkp_serve_create(const ZkpServeConfig* cfg, void** engine);
unsigned long long zkp_serve_footprint(void* engine, unsigned int kind, unsigned int log_n);
hipError_t zkp_serve_run(void* engine, ZkpServeRequest* reqs, unsigned int count);
void zkp_serve_destroy(void* engine);

#ifdef __cplusplus
}
#endif

#endif /* ZKPOLY_SERVE_H */
