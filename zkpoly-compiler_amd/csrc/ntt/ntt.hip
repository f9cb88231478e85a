// libntt.so -- BN254 Fr number-theoretic transforms for gfx950.
//
// C ABI identical to the reference's core/src/ntt/src/ntt.h:6-17 (cudaError_t
// -> hipError_t, cudaStream_t -> hipStream_t), so the reference runtime's
// closures (core/src/ntt.rs:165-200, 112-150) bind these symbols unchanged.
//
// Transform: in-place forward DFT X[j] = sum_i a[i] w^(ij) with natural order
// in and out (matches halo2's best_fft, core/tests/ntt.rs:58,139-143).  The
// network is exactly best_fft's radix-2 decimation-in-time network
// (core/tests/msm_halo2.rs:221-324): a bit-reversal permutation followed by
// log2(n) butterfly stages b' = a - b*w, a' = a + b*w with stage-s twiddle
// w = tw[(i mod 2^s) * n / 2^(s+1)].  Because the butterflies and twiddles are
// the same, the output equals best_fft even for an omega that is not a root of
// unity (the reference's own tests use a random omega, core/tests/ntt.rs:53).
//
// HBM layout: the polynomial is the caller's array of 32-byte elements (8 x u32
// Montgomery limbs), addressed through the PolyPtr index map; the twiddle table
// is the caller's [1, w, ..., w^(n/2-1)] (what ssip_precompute writes, the same
// table the reference's compiler stores as a constant).
//
// Kernels:
//  * bitrev_kernel: in-place swap of i <-> rev(i).
//  * dit_pass<TILE_LOG>: a workgroup loads a tile of 2^TILE_LOG elements --
//    2^d butterfly-connected rows times 2^c adjacent columns, so every global
//    access is a run of >= 4 consecutive 32-byte elements -- into LDS (two
//    16-byte planes), runs d consecutive stages there, and writes it back.
//    A transform of 2^24 is bitrev + 3 passes (11 + 9 + 4 stages).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../common/poly_ptr.hpp"
#include "../field/bn254.hpp"

namespace zkp {
namespace ntt {

constexpr int THREADS = 256;
constexpr int TILE_LOG = 11;          // 2048 elements = 64 KiB of LDS per workgroup
constexpr int MIN_COL_LOG_LATER = 2;  // later passes read runs of >= 4 elements (128 B)

__device__ __forceinline__ Fr load_elem(const PolyView& v, uint64_t i) { return Fr::load(v.elem(i)); }
__device__ __forceinline__ void store_elem(const PolyView& v, uint64_t i, const Fr& x) { x.store(v.elem(i)); }

__global__ void __launch_bounds__(THREADS) bitrev_kernel(PolyView v, uint32_t log_n) {
  const uint64_t n = 1ull << log_n;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t j = __brevll(i) >> (64 - log_n);
    if (i < j) {
      Fr a = load_elem(v, i);
      Fr b = load_elem(v, j);
      store_elem(v, i, b);
      store_elem(v, j, a);
    }
  }
}

// One pass of DIT stages [s0, s0 + d) over tile `tile` of 2^d rows x 2^c_log columns.
//   s0 > 0 : columns are consecutive low indices, rows stride 2^s0
//            idx(c, k) = lo_block * C + c + (k << s0) + (hi << (s0 + d))
//   s0 == 0: columns are consecutive contiguous groups of 2^d elements
//            idx(c, k) = (tile * C + c) << d | k
__device__ __forceinline__ void dit_tile(uint4* __restrict__ lds0, uint4* __restrict__ lds1, const PolyView& v,
                                         const uint32_t* __restrict__ tw, uint32_t log_n, uint32_t s0, uint32_t d,
                                         uint32_t c_log, uint64_t tile) {
  const uint32_t C = 1u << c_log;
  const uint32_t R = 1u << d;
  const uint32_t T = R << c_log;

  uint64_t base, cstride, kstride;
  if (s0 == 0) {
    base = (tile << c_log) << d;
    cstride = R;
    kstride = 1;
  } else {
    const uint64_t lo_blocks = (1ull << s0) >> c_log;
    const uint64_t hi = tile / lo_blocks;
    const uint64_t lo_block = tile % lo_blocks;
    base = lo_block * C + (hi << (s0 + d));
    cstride = 1;
    kstride = 1ull << s0;
  }

  // load: consecutive threads take consecutive global addresses
  for (uint32_t e = threadIdx.x; e < T; e += THREADS) {
    uint32_t c, k;
    if (cstride == 1) {
      c = e & (C - 1);
      k = e >> c_log;
    } else {
      k = e & (R - 1);
      c = e >> d;
    }
    const uint64_t idx = base + c * cstride + (uint64_t)k * kstride;
    const uint4* src = reinterpret_cast<const uint4*>(v.elem(idx));
    const uint32_t pos = (k << c_log) | c;
    lds0[pos] = src[0];
    lds1[pos] = src[1];
  }
  __syncthreads();

  for (uint32_t m = 0; m < d; m++) {
    const uint32_t s = s0 + m;
    const uint32_t half = 1u << m;
    const uint64_t jmask = (1ull << s) - 1;
    const uint32_t tw_shift = log_n - 1 - s;
    for (uint32_t b = threadIdx.x; b < (T >> 1); b += THREADS) {
      const uint32_t c = b & (C - 1);
      const uint32_t q = b >> c_log;
      const uint32_t ka = ((q >> m) << (m + 1)) | (q & (half - 1));
      const uint32_t kb = ka + half;
      const uint32_t pa = (ka << c_log) | c;
      const uint32_t pb = (kb << c_log) | c;
      const uint64_t idx_a = base + c * cstride + (uint64_t)ka * kstride;
      const uint64_t j = idx_a & jmask;
      const Fr w = Fr::load(tw + ((j << tw_shift) << 3));
      Fr x, y;
      {
        uint4 a0 = lds0[pa], a1 = lds1[pa], b0 = lds0[pb], b1 = lds1[pb];
        x.v[0] = a0.x; x.v[1] = a0.y; x.v[2] = a0.z; x.v[3] = a0.w;
        x.v[4] = a1.x; x.v[5] = a1.y; x.v[6] = a1.z; x.v[7] = a1.w;
        y.v[0] = b0.x; y.v[1] = b0.y; y.v[2] = b0.z; y.v[3] = b0.w;
        y.v[4] = b1.x; y.v[5] = b1.y; y.v[6] = b1.z; y.v[7] = b1.w;
      }
      const Fr t = y * w;
      const Fr u = x + t;
      const Fr l = x - t;
      lds0[pa] = make_uint4(u.v[0], u.v[1], u.v[2], u.v[3]);
      lds1[pa] = make_uint4(u.v[4], u.v[5], u.v[6], u.v[7]);
      lds0[pb] = make_uint4(l.v[0], l.v[1], l.v[2], l.v[3]);
      lds1[pb] = make_uint4(l.v[4], l.v[5], l.v[6], l.v[7]);
    }
    __syncthreads();
  }

  for (uint32_t e = threadIdx.x; e < T; e += THREADS) {
    uint32_t c, k;
    if (cstride == 1) {
      c = e & (C - 1);
      k = e >> c_log;
    } else {
      k = e & (R - 1);
      c = e >> d;
    }
    const uint64_t idx = base + c * cstride + (uint64_t)k * kstride;
    uint4* dst = reinterpret_cast<uint4*>(v.elem(idx));
    const uint32_t pos = (k << c_log) | c;
    dst[0] = lds0[pos];
    dst[1] = lds1[pos];
  }
}

__global__ void __launch_bounds__(THREADS)
    dit_pass(PolyView v, const uint32_t* __restrict__ tw, uint32_t log_n, uint32_t s0, uint32_t d,
             uint32_t c_log) {
  __shared__ uint4 lds0[1 << TILE_LOG];
  __shared__ uint4 lds1[1 << TILE_LOG];
  dit_tile(lds0, lds1, v, tw, log_n, s0, d, c_log, blockIdx.x);
}

// ---------------------------------------------------------------- batched (many instances per launch)

// One transform of a batch.
struct BatchInstance {
  PolyView v;
  const uint32_t* tw;
  uint32_t log_n;
  uint32_t pad;
};

// Pass parameters of one instance in one round (d == 0: no pass this round).
struct BatchPass {
  uint32_t s0, d, c_log, pad;
};

// largest k < n with prefix[k] <= g (prefix: n + 1 non-decreasing entries)
__device__ __forceinline__ uint32_t find_instance(const uint64_t* prefix, uint32_t n, uint64_t g) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Bit reversal of every instance; prefix[k] = sum of earlier instance lengths.
__global__ void __launch_bounds__(THREADS)
    bitrev_batched(const BatchInstance* __restrict__ inst, const uint64_t* __restrict__ prefix, uint32_t count) {
  const uint64_t total = prefix[count];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
    const uint32_t k = find_instance(prefix, count, g);
    const uint64_t i = g - prefix[k];
    const uint32_t log_n = inst[k].log_n;
    if (log_n == 0) continue;
    const uint64_t j = __brevll(i) >> (64 - log_n);
    if (i < j) {
      const PolyView v = inst[k].v;
      Fr a = load_elem(v, i);
      Fr b = load_elem(v, j);
      store_elem(v, i, b);
      store_elem(v, j, a);
    }
  }
}

// One pass round over all instances: block b runs tile (b - tile_prefix[k]) of
// instance k, with tile_prefix[k] = sum of earlier instances' tile counts.
__global__ void __launch_bounds__(THREADS)
    dit_pass_batched(const BatchInstance* __restrict__ inst, const BatchPass* __restrict__ pass,
                     const uint64_t* __restrict__ tile_prefix, uint32_t count) {
  __shared__ uint4 lds0[1 << TILE_LOG];
  __shared__ uint4 lds1[1 << TILE_LOG];
  const uint64_t b = blockIdx.x;
  if (b >= tile_prefix[count]) return;
  const uint32_t k = find_instance(tile_prefix, count, b);
  const BatchPass p = pass[k];
  dit_tile(lds0, lds1, inst[k].v, inst[k].tw, inst[k].log_n, p.s0, p.d, p.c_log, b - tile_prefix[k]);
}

// out[i] = base^i for i < len; each thread starts its chunk with one pow.
__global__ void __launch_bounds__(THREADS)
    gen_powers_kernel(uint32_t* __restrict__ out, uint64_t len, const uint32_t* __restrict__ base_ptr,
                      uint32_t chunk) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t start = t * chunk;
  if (start >= len) return;
  const Fr base = Fr::load(base_ptr);
  Fr x = base.pow(start);
  const uint64_t end = start + chunk < len ? start + chunk : len;
  for (uint64_t i = start; i < end; i++) {
    x.store(out + i * 8);
    x = x * base;
  }
}

// x[i] *= powers[i mod m]  (core/src/ntt/src/distribute_powers.cuh:4-19)
__global__ void __launch_bounds__(THREADS)
    distribute_powers_kernel(PolyView v, const uint32_t* __restrict__ powers, uint64_t m, uint64_t mask,
                             bool pow2) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.len; i += stride) {
    const uint64_t pi = pow2 ? (i & mask) : (i % m);
    const Fr p = Fr::load(powers + pi * 8);
    store_elem(v, i, load_elem(v, i) * p);
  }
}

struct PassPlan {
  uint32_t s0, d, c_log;
  uint64_t tiles;
};

// Splits log_n DIT stages into LDS passes.  Pass 0 takes up to TILE_LOG stages
// over contiguous groups; later passes take <= TILE_LOG - MIN_COL_LOG_LATER
// stages each, balanced, with as many adjacent columns as the tile holds.
static int plan_passes(uint32_t log_n, PassPlan* out) {
  int np = 0;
  uint32_t d0 = log_n < (uint32_t)TILE_LOG ? log_n : (uint32_t)TILE_LOG;
  uint32_t c0 = TILE_LOG - d0;
  if (c0 > log_n - d0) c0 = log_n - d0;
  out[np++] = PassPlan{0, d0, c0, (1ull << log_n) >> (d0 + c0)};
  uint32_t s0 = d0;
  const uint32_t max_d = TILE_LOG - MIN_COL_LOG_LATER;
  while (s0 < log_n) {
    const uint32_t rem = log_n - s0;
    const uint32_t parts = (rem + max_d - 1) / max_d;
    const uint32_t d = (rem + parts - 1) / parts;
    uint32_t c = TILE_LOG - d;
    if (c > s0) c = s0;
    out[np++] = PassPlan{s0, d, c, (1ull << log_n) >> (d + c)};
    s0 += d;
  }
  return np;
}

static inline unsigned grid_for(uint64_t work, unsigned threads, unsigned cap) {
  uint64_t g = (work + threads - 1) / threads;
  if (g > cap) g = cap;
  if (g == 0) g = 1;
  return (unsigned)g;
}

__global__ void __launch_bounds__(THREADS)
    copy_u4_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Validates, then launches bitrev + the planned passes on `stream`.  When
// `events` is given, events[0] is recorded first and events[i + 1] after
// kernel i (bounded by max_events).
static hipError_t launch_ntt(PolyPtr x, const unsigned int* twiddle, unsigned int log_len, hipStream_t stream,
                             hipEvent_t* events, int max_events, int* kinds, int* num_kernels) {
  PolyView v;
  if (!make_view(x, &v)) return hipErrorInvalidValue;
  if (log_len > 40 || v.len != (1ull << log_len)) return hipErrorInvalidValue;
  if (log_len == 0) return hipSuccess;
  if (v.ptr == nullptr || twiddle == nullptr) return hipErrorInvalidValue;
  int ev = 0, nk = 0;
  hipError_t err;
  if (events && ev < max_events) {
    err = hipEventRecord(events[ev++], stream);
    if (err != hipSuccess) return err;
  }
  hipLaunchKernelGGL(bitrev_kernel, dim3(grid_for(v.len, THREADS, 65536)), dim3(THREADS), 0, stream, v,
                     (uint32_t)log_len);
  err = hipGetLastError();
  if (err != hipSuccess) return err;
  if (kinds && nk < max_events) kinds[nk] = 0;
  nk++;
  if (events && ev < max_events) {
    err = hipEventRecord(events[ev++], stream);
    if (err != hipSuccess) return err;
  }
  PassPlan plan[48];
  const int np = plan_passes(log_len, plan);
  for (int p = 0; p < np; p++) {
    hipLaunchKernelGGL(dit_pass, dim3((unsigned)plan[p].tiles), dim3(THREADS), 0, stream, v, twiddle,
                       (uint32_t)log_len, plan[p].s0, plan[p].d, plan[p].c_log);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    if (kinds && nk < max_events) kinds[nk] = 1;
    nk++;
    if (events && ev < max_events) {
      err = hipEventRecord(events[ev++], stream);
      if (err != hipSuccess) return err;
    }
  }
  if (num_kernels) *num_kernels = nk;
  return hipSuccess;
}

// Device work lists of a batch of transforms (built once, replayed by
// zkp_ntt_batch_run, which only launches kernels and so can be captured in a
// hipGraph).
struct BatchPlan {
  uint32_t count = 0;
  BatchInstance* d_inst = nullptr;
  uint64_t* d_elem_prefix = nullptr;
  uint64_t total_elems = 0;
  std::vector<BatchPass*> d_pass;
  std::vector<uint64_t*> d_tile_prefix;
  std::vector<uint64_t> total_tiles;

  ~BatchPlan() {
    if (d_inst) (void)hipFree(d_inst);
    if (d_elem_prefix) (void)hipFree(d_elem_prefix);
    for (auto p : d_pass) (void)hipFree(p);
    for (auto p : d_tile_prefix) (void)hipFree(p);
  }
};

}  // namespace ntt
}  // namespace zkp

using namespace zkp;
using namespace zkp::ntt;

extern "C" {

// In-place natural-order forward NTT of x (len = 2^log_len) with twiddle table
// [1, w, ..., w^(len/2-1)] on the device.  Replaces ssip_ntt of
// core/src/ntt/src/ntt.h:6 / ntt.cu:7-9.  The two thread-shape hints tune the
// reference's CUDA kernels and are accepted but not needed here.
hipError_t ssip_ntt(PolyPtr x, const unsigned int* twiddle, unsigned int log_len, hipStream_t stream,
                    const unsigned int max_threads_stage1_log, const unsigned int max_threads_stage2_log) {
  (void)max_threads_stage1_log;
  (void)max_threads_stage2_log;
  return zkp::ntt::launch_ntt(x, twiddle, log_len, stream, nullptr, 0, nullptr, nullptr);
}

// ssip_ntt that also records events[0] before the first kernel and
// events[i + 1] after kernel i (i < *num_kernels), for per-kernel timing of the
// same launches.  kinds[i] = 0 for the bit-reversal, 1 for a butterfly pass.
// No counterpart in the reference ABI (measurement hook).
hipError_t zkp_ntt_with_events(PolyPtr x, const unsigned int* twiddle, unsigned int log_len, hipStream_t stream,
                               hipEvent_t* events, int max_events, int* kinds, int* num_kernels) {
  int nk = 0;
  hipError_t err = zkp::ntt::launch_ntt(x, twiddle, log_len, stream, events, max_events, kinds, &nk);
  if (num_kernels) *num_kernels = nk;
  return err;
}

// dst[i] = src[i] for n 16-byte words: a plain streaming copy used to calibrate
// the HBM traffic counters (FETCH_SIZE / WRITE_SIZE) against known bytes.
hipError_t zkp_copy_u4(void* dst, const void* src, unsigned long long n16, hipStream_t stream) {
  if (n16 == 0) return hipSuccess;
  if (dst == nullptr || src == nullptr) return hipErrorInvalidValue;
  hipLaunchKernelGGL(zkp::ntt::copy_u4_kernel, dim3(zkp::ntt::grid_for(n16, zkp::ntt::THREADS, 65536)),
                     dim3(zkp::ntt::THREADS), 0, stream, reinterpret_cast<uint4*>(dst),
                     reinterpret_cast<const uint4*>(src), (uint64_t)n16);
  return hipGetLastError();
}

// twiddle[i] = unit^i for i < len, written to HOST memory; synchronous.
// Replaces ssip_precompute of core/src/ntt/src/ntt.h:9 / ntt.cu:11-14 (the
// reference computes the same prefix product with a device scan,
// precompute.cuh:8-86).  The second argument is the table length (the Rust
// wrapper passes twiddle.len, core/src/ntt.rs:221-226).
hipError_t ssip_precompute(unsigned int* twiddle, unsigned int len, const unsigned int* unit) {
  if (len == 0) return hipSuccess;
  if (twiddle == nullptr || unit == nullptr) return hipErrorInvalidValue;
  uint32_t *d_out = nullptr, *d_unit = nullptr;
  hipStream_t s = nullptr;
  hipError_t err = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (err != hipSuccess) return err;
  err = hipMalloc(&d_out, (size_t)len * 32);
  if (err == hipSuccess) err = hipMalloc(&d_unit, 32);
  if (err == hipSuccess) err = hipMemcpyAsync(d_unit, unit, 32, hipMemcpyHostToDevice, s);
  if (err == hipSuccess) {
    const uint32_t chunk = 256;
    const uint64_t threads = ((uint64_t)len + chunk - 1) / chunk;
    hipLaunchKernelGGL(gen_powers_kernel, dim3(grid_for(threads, THREADS, 1u << 30)), dim3(THREADS), 0, s,
                       d_out, (uint64_t)len, d_unit, chunk);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipMemcpyAsync(twiddle, d_out, (size_t)len * 32, hipMemcpyDeviceToHost, s);
  if (err == hipSuccess) err = hipStreamSynchronize(s);
  if (d_out) (void)hipFree(d_out);
  if (d_unit) (void)hipFree(d_unit);
  (void)hipStreamDestroy(s);
  return err;
}

// x[i] *= powers[i mod power_num] on the device.  Replaces distribute_powers of
// core/src/ntt/src/ntt.h:17 / distribute_powers.cuh:13-19.
hipError_t distribute_powers(PolyPtr poly, const unsigned int* powers, unsigned long long power_num,
                             hipStream_t stream) {
  PolyView v;
  if (!make_view(poly, &v)) return hipErrorInvalidValue;
  if (v.len == 0) return hipSuccess;
  if (v.ptr == nullptr || powers == nullptr || power_num == 0) return hipErrorInvalidValue;
  const bool pow2 = (power_num & (power_num - 1)) == 0;
  hipLaunchKernelGGL(distribute_powers_kernel, dim3(grid_for(v.len, THREADS, 8192)), dim3(THREADS), 0,
                     stream, v, powers, (uint64_t)power_num, (uint64_t)(power_num - 1), pow2);
  return hipGetLastError();
}

// out[i] = base^i for i < len, device memory, asynchronous on stream.  No
// counterpart in the reference ABI: used to build coset power tables and
// twiddle tables directly in HBM.
hipError_t zkp_gen_powers(unsigned int* out, unsigned long long len, const unsigned int* base,
                          hipStream_t stream) {
  if (len == 0) return hipSuccess;
  if (out == nullptr || base == nullptr) return hipErrorInvalidValue;
  const uint32_t chunk = 256;
  const uint64_t threads = (len + chunk - 1) / chunk;
  hipLaunchKernelGGL(gen_powers_kernel, dim3(grid_for(threads, THREADS, 1u << 30)), dim3(THREADS), 0, stream,
                     out, (uint64_t)len, base, chunk);
  return hipGetLastError();
}

}  // extern "C"

extern "C" {

// Plans `count` independent in-place transforms (xs[i], twiddles[i] device
// pointers, length 2^log_lens[i]) to run together: one bit-reversal launch and
// one launch per pass round for the whole batch.  Allocates the device work
// lists and returns an opaque plan.  No counterpart in the reference ABI.
hipError_t zkp_ntt_batch_create(const PolyPtr* xs, const unsigned int* const* twiddles, const unsigned int* log_lens,
                                unsigned int count, void** plan_out) {
  if (plan_out == nullptr || (count > 0 && (xs == nullptr || twiddles == nullptr || log_lens == nullptr)))
    return hipErrorInvalidValue;
  BatchPlan* plan = new BatchPlan();
  plan->count = count;
  std::vector<BatchInstance> inst(count);
  std::vector<uint64_t> elem_prefix(count + 1, 0);
  std::vector<std::vector<PassPlan>> plans(count);
  size_t rounds = 0;
  for (unsigned i = 0; i < count; i++) {
    PolyView v;
    if (!make_view(xs[i], &v) || log_lens[i] > 40 || v.len != (1ull << log_lens[i]) ||
        (log_lens[i] > 0 && (v.ptr == nullptr || twiddles[i] == nullptr))) {
      delete plan;
      return hipErrorInvalidValue;
    }
    inst[i] = BatchInstance{v, twiddles[i], log_lens[i], 0};
    elem_prefix[i + 1] = elem_prefix[i] + (log_lens[i] > 0 ? v.len : 0);
    if (log_lens[i] > 0) {
      PassPlan pp[48];
      const int np = plan_passes(log_lens[i], pp);
      plans[i].assign(pp, pp + np);
      if ((size_t)np > rounds) rounds = np;
    }
  }
  plan->total_elems = elem_prefix[count];
  hipError_t err = hipMalloc(&plan->d_inst, sizeof(BatchInstance) * (count ? count : 1));
  if (err == hipSuccess) err = hipMalloc(&plan->d_elem_prefix, sizeof(uint64_t) * (count + 1));
  if (err == hipSuccess && count)
    err = hipMemcpy(plan->d_inst, inst.data(), sizeof(BatchInstance) * count, hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = hipMemcpy(plan->d_elem_prefix, elem_prefix.data(), sizeof(uint64_t) * (count + 1), hipMemcpyHostToDevice);
  for (size_t r = 0; r < rounds && err == hipSuccess; r++) {
    std::vector<BatchPass> passes(count, BatchPass{0, 0, 0, 0});
    std::vector<uint64_t> tile_prefix(count + 1, 0);
    for (unsigned i = 0; i < count; i++) {
      uint64_t tiles = 0;
      if (r < plans[i].size()) {
        const PassPlan& q = plans[i][r];
        passes[i] = BatchPass{q.s0, q.d, q.c_log, 0};
        tiles = q.tiles;
      }
      tile_prefix[i + 1] = tile_prefix[i] + tiles;
    }
    BatchPass* dp = nullptr;
    uint64_t* dt = nullptr;
    err = hipMalloc(&dp, sizeof(BatchPass) * (count ? count : 1));
    if (err == hipSuccess) err = hipMalloc(&dt, sizeof(uint64_t) * (count + 1));
    if (dp) plan->d_pass.push_back(dp);
    if (dt) plan->d_tile_prefix.push_back(dt);
    if (err == hipSuccess && count)
      err = hipMemcpy(dp, passes.data(), sizeof(BatchPass) * count, hipMemcpyHostToDevice);
    if (err == hipSuccess)
      err = hipMemcpy(dt, tile_prefix.data(), sizeof(uint64_t) * (count + 1), hipMemcpyHostToDevice);
    plan->total_tiles.push_back(tile_prefix[count]);
  }
  if (err != hipSuccess) {
    delete plan;
    return err;
  }
  *plan_out = plan;
  return hipSuccess;
}

// Launches the planned batch on `stream` (kernel launches only).
hipError_t zkp_ntt_batch_run(void* plan_ptr, hipStream_t stream) {
  BatchPlan* plan = reinterpret_cast<BatchPlan*>(plan_ptr);
  if (plan == nullptr) return hipErrorInvalidValue;
  if (plan->total_elems == 0) return hipSuccess;
  hipLaunchKernelGGL(bitrev_batched, dim3(grid_for(plan->total_elems, THREADS, 65536)), dim3(THREADS), 0, stream,
                     plan->d_inst, plan->d_elem_prefix, plan->count);
  hipError_t err = hipGetLastError();
  for (size_t r = 0; r < plan->d_pass.size() && err == hipSuccess; r++) {
    if (plan->total_tiles[r] == 0) continue;
    hipLaunchKernelGGL(dit_pass_batched, dim3((unsigned)plan->total_tiles[r]), dim3(THREADS), 0, stream, plan->d_inst,
                       plan->d_pass[r], plan->d_tile_prefix[r], plan->count);
    err = hipGetLastError();
  }
  return err;
}

// Frees a plan (after the work using it has completed).
void zkp_ntt_batch_destroy(void* plan_ptr) { delete reinterpret_cast<BatchPlan*>(plan_ptr); }

}  // extern "C"
