// Issue-rate microbenchmark for the integer and fp64 instructions that bound
// 256-bit modular arithmetic on gfx950: v_mad_u64_u32, v_mul_lo_u32,
// v_mul_hi_u32, 64-bit add (v_add_co/v_addc), fp64 FMA, and the Montgomery
// product of the BN254 scalar field built from them.
//
// Each kernel runs 8 independent dependency chains per lane so the issue rate,
// not latency, is measured.  Prints one JSON object per kernel with lane-ops/s.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include "../field/bn254.hpp"

using namespace zkp;

#define CHECK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_),       \
              __FILE__, __LINE__);                                            \
      exit(1);                                                                \
    }                                                                         \
  } while (0)

constexpr int CHAINS = 8;

__global__ void k_mad64(u64* out, u32 y, int iters) {
  u64 acc[CHAINS];
#pragma unroll
  for (int k = 0; k < CHAINS; k++) acc[k] = threadIdx.x * 977u + k;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < CHAINS; k++) acc[k] = (u64)(u32)acc[k] * y + acc[k];
  }
  u64 s = 0;
#pragma unroll
  for (int k = 0; k < CHAINS; k++) s ^= acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_mullo(u64* out, u32 y, int iters) {
  u32 acc[CHAINS];
#pragma unroll
  for (int k = 0; k < CHAINS; k++) acc[k] = threadIdx.x * 977u + 2 * k + 1;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < CHAINS; k++) acc[k] = acc[k] * y;
  }
  u32 s = 0;
#pragma unroll
  for (int k = 0; k < CHAINS; k++) s ^= acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_mulhi(u64* out, u32 y, int iters) {
  u32 acc[CHAINS];
#pragma unroll
  for (int k = 0; k < CHAINS; k++) acc[k] = threadIdx.x * 977u + 2 * k + 1 + 0x80000000u;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < CHAINS; k++) acc[k] = __umulhi(acc[k], y) | 0x80000000u;
  }
  u32 s = 0;
#pragma unroll
  for (int k = 0; k < CHAINS; k++) s ^= acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_add64(u64* out, u32 y, int iters) {
  u64 acc[CHAINS];
#pragma unroll
  for (int k = 0; k < CHAINS; k++) acc[k] = threadIdx.x * 0x100000001ull + k;
  const u64 yy = ((u64)y << 32) | y;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < CHAINS; k++) acc[k] = acc[k] + yy;
  }
  u64 s = 0;
#pragma unroll
  for (int k = 0; k < CHAINS; k++) s ^= acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_fma64(u64* out, u32 y, int iters) {
  double acc[CHAINS];
  const double m = 1.0 + 1.0 / (double)(y | 1);
  const double a = 0.5 / (double)(y | 1);
#pragma unroll
  for (int k = 0; k < CHAINS; k++) acc[k] = threadIdx.x * 1e-3 + k;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < CHAINS; k++) acc[k] = fma(acc[k], m, a);
  }
  double s = 0;
#pragma unroll
  for (int k = 0; k < CHAINS; k++) s += acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = (u64)s;
}

__global__ void k_frmul(u64* out, u32 y, int iters) {
  constexpr int C = 4;
  Fr acc[C];
  Fr m = Fr::one();
  m.v[0] ^= y;
#pragma unroll
  for (int k = 0; k < C; k++) {
    acc[k] = Fr::one();
    acc[k].v[1] ^= threadIdx.x + k;
  }
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < C; k++) acc[k] = acc[k] * m;
  }
  u64 s = 0;
#pragma unroll
  for (int k = 0; k < C; k++) s ^= acc[k].v[0] ^ acc[k].v[7];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// One dependent chain in one wave: latency of a Montgomery product.
__global__ void k_frmul_latency(u64* out, u32 y, int iters) {
  Fr acc = Fr::one();
  Fr m = Fr::one();
  m.v[0] ^= y;
  acc.v[1] ^= threadIdx.x;
  for (int it = 0; it < iters; it++) acc = acc * m;
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc.v[0] ^ acc.v[7];
}

typedef void (*kern_t)(u64*, u32, int);

static void run(const char* name, kern_t k, int iters, double ops_per_iter, int blocks,
                int threads, u64* dout) {
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  hipLaunchKernelGGL(k, dim3(blocks), dim3(threads), 0, 0, dout, 0x9e3779b9u, iters);
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  float best = 1e30f;
  for (int rep = 0; rep < 5; rep++) {
    CHECK(hipEventRecord(a, 0));
    hipLaunchKernelGGL(k, dim3(blocks), dim3(threads), 0, 0, dout, 0x9e3779b9u, iters);
    CHECK(hipEventRecord(b, 0));
    CHECK(hipEventSynchronize(b));
    float ms;
    CHECK(hipEventElapsedTime(&ms, a, b));
    if (ms < best) best = ms;
  }
  double lane_ops = (double)blocks * threads * iters * ops_per_iter;
  printf("{\"kernel\": \"%s\", \"ms\": %.4f, \"lane_ops_per_s\": %.4e}\n", name, best,
         lane_ops / (best * 1e-3));
  fflush(stdout);
  CHECK(hipEventDestroy(a));
  CHECK(hipEventDestroy(b));
}

int main(int argc, char** argv) {
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  printf("{\"device\": \"%s\", \"gcn_arch\": \"%s\", \"cus\": %d, \"clock_khz\": %d, "
         "\"lds_per_block\": %zu, \"l2\": %d}\n",
         prop.name, prop.gcnArchName, prop.multiProcessorCount, prop.clockRate,
         prop.sharedMemPerBlock, prop.l2CacheSize);
  int blocks = prop.multiProcessorCount * 16;
  int threads = 256;
  u64* dout;
  CHECK(hipMalloc(&dout, (size_t)blocks * threads * sizeof(u64)));
  run("v_mad_u64_u32", k_mad64, 4096, CHAINS, blocks, threads, dout);
  run("v_mul_lo_u32", k_mullo, 4096, CHAINS, blocks, threads, dout);
  run("v_mul_hi_u32+or", k_mulhi, 4096, CHAINS, blocks, threads, dout);
  run("add_u64", k_add64, 4096, CHAINS, blocks, threads, dout);
  run("v_fma_f64", k_fma64, 4096, CHAINS, blocks, threads, dout);
  run("fr_mont_mul", k_frmul, 256, 4, blocks, threads, dout);
  {
    // latency: a single wave, one chain
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    const int iters = 20000;
    hipLaunchKernelGGL(k_frmul_latency, dim3(1), dim3(64), 0, 0, dout, 0x9e3779b9u, iters);
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(a, 0));
    hipLaunchKernelGGL(k_frmul_latency, dim3(1), dim3(64), 0, 0, dout, 0x9e3779b9u, iters);
    CHECK(hipEventRecord(b, 0));
    CHECK(hipEventSynchronize(b));
    float ms;
    CHECK(hipEventElapsedTime(&ms, a, b));
    printf("{\"kernel\": \"fr_mont_mul_latency_one_wave\", \"ns_per_mul\": %.2f}\n", ms * 1e6 / iters);
  }
  CHECK(hipFree(dout));
  return 0;
}
