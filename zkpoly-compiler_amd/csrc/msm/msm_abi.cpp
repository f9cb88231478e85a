// libmsm_bn254_<window>_<target_windows>_<debug>.so: the reference's MSM C ABI
// (core/src/msm/src/msm.h:4-11) bound to one compiled-in MsmConfig, the way the
// reference builds one library per configuration (core/src/msm.rs:100-111,
// Makefile.msm).  All work happens in libzkp_msm.so.
#include <hip/hip_runtime_api.h>

#ifndef ZKP_MSM_WINDOW
#error "ZKP_MSM_WINDOW must be defined"
#endif
#ifndef ZKP_MSM_TARGET
#error "ZKP_MSM_TARGET must be defined"
#endif
#ifndef ZKP_MSM_DEBUG
#define ZKP_MSM_DEBUG 0
#endif

extern "C" {
hipError_t zkp_msm_host(unsigned int window, unsigned int target_windows, int debug_flag, void* const* buffers,
                        unsigned long* buffer_sizes, unsigned long long len, unsigned int batch_per_run,
                        unsigned int parts, unsigned int stage_scalers, unsigned int stage_points,
                        unsigned int num_cards, const unsigned int* cards, unsigned int const* const* h_points,
                        unsigned int batches, unsigned int const* const* h_scaler_batch,
                        unsigned int* const* h_result);
hipError_t zkp_msm_precompute_host(unsigned int window, unsigned int target_windows, unsigned long long len,
                                   unsigned int* const* h_points, unsigned int max_cards);

hipError_t msm(void* const* buffers, unsigned long* buffer_sizes, unsigned long long len, unsigned int batch_per_run,
               unsigned int parts, unsigned int stage_scalers, unsigned int stage_points, unsigned int num_cards,
               const unsigned int* cards, unsigned int const* const* h_points, unsigned int batches,
               unsigned int const* const* h_scaler_batch, unsigned int* const* h_result) {
  return zkp_msm_host(ZKP_MSM_WINDOW, ZKP_MSM_TARGET, ZKP_MSM_DEBUG, buffers, buffer_sizes, len, batch_per_run, parts,
                      stage_scalers, stage_points, num_cards, cards, h_points, batches, h_scaler_batch, h_result);
}

hipError_t msm_precompute(unsigned long long len, unsigned int* const* h_points, unsigned int max_cards) {
  return zkp_msm_precompute_host(ZKP_MSM_WINDOW, ZKP_MSM_TARGET, len, h_points, max_cards);
}
}
