// libzkp_msm.so -- Pippenger MSM over BN254 G1 (kernels + host orchestration).
//
// The reference's per-configuration libraries libmsm_bn254_<w>_<t>_<dbg>.so
// (core/src/msm.rs:100-111) are thin wrappers (msm_abi.cpp) that bind its C
// ABI -- msm / msm_precompute of core/src/msm/src/msm.h:4-11, cudaError_t ->
// hipError_t -- to zkp_msm_host / zkp_msm_precompute_host below with their
// compiled-in MsmConfig.  The zkp_msm_* entry points take the window at run
// time.
//
// Algorithm (result sum_i k_i P_i, affine; identity (0, 0)):
//  1. digits: scalar -> canonical integer -> ceil(254/s) signed s-bit digits in
//     [-2^(s-1), 2^(s-1)) (the reference's signed_digit, msm_impl.cuh:22-33);
//     logical window w uses point table w / n_windows and physical window
//     w % n_windows (precomputed tables 2^(j*n_windows*s) P, msm_impl.cuh:358-371).
//  2. sort of the (window, |digit|) keys with values (point index | sign << 31):
//     a two-level partition sort in LDS (coarse bins, then the low bucket
//     bits), no device atomics; windows above 20 bits use a device-atomic
//     counting sort instead.
//  3. bucket accumulation, load balanced: every thread owns 32 consecutive
//     sorted entries and adds each run of equal keys into an XYZZ sum (mixed
//     additions); runs that cross a chunk edge become (key, partial) pairs
//     that the next level reduces the same way, so a bucket holding any share
//     of the points (e.g. the scalar-1 bucket) costs log levels, not a serial walk.
//     Buckets and partials are G1X29 (xyzz29.hpp): Fq in nine 29-bit limbs,
//     radix 2^261, lazily reduced -- one instruction per limb product.
//  4. bucket reduction per window, S_w = sum_b b * B_b, by halving levels
//     (X, Y) -> (2 (X_2i + X_2i+1), X_2i+1 + Y_2i + Y_2i+1): X' and Y' run on
//     different threads, two point operations deep per level, since a chain of
//     point additions is latency bound; the last 128 items of a window are
//     halved inside one workgroup (LDS), all levels in one launch, and the
//     window sums leave in G1XYZZ's form (8 x 32-bit limbs, radix 2^256).
//  5. host: Horner over the window sums (s doublings each) and one inversion
//     to affine -- the serial tail runs where a modular product is ~30x lower
//     latency than one GPU wave (measured 738 ns per dependent product).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <mutex>
#include <thread>
#include <vector>

#include "../curve/g1.hpp"
#include "../curve/xyzz29.hpp"

namespace zkp {
namespace msm {

constexpr int THREADS = 256;
constexpr uint32_t CHUNK = 32;   // sorted entries per thread at accumulation level 0
constexpr uint32_t CHUNK_L = 8;  // (key, partial) entries per thread at the later levels
constexpr uint32_t RED_TAIL = 128;  // items per window finished inside one workgroup
constexpr uint32_t SCALAR_BITS = 254;
constexpr int MAX_LEVELS = 24;

#define ZKP_CHECK(x)                          \
  do {                                        \
    hipError_t e_ = (x);                      \
    if (e_ != hipSuccess) return e_;          \
  } while (0)

struct Config {
  uint32_t s;               // window bits
  uint32_t actual_windows;  // ceil(254 / s)
  uint32_t n_windows;       // physical windows per table
  uint32_t n_precompute;    // point tables
  uint32_t buckets;         // 2^(s-1)
};

static inline uint32_t div_ceil32(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// MsmConfig semantics of core/src/msm/src/msm.cuh:51-70 / common/src/msm_config.rs:14-24.
// Top s-bit window of r - 1 (the largest canonical scalar) starting at bit `bit`.
static uint32_t top_window_of_r_minus_1(uint32_t bit, uint32_t s) {
  uint32_t v[8];
  for (int i = 0; i < 8; i++) v[i] = FrParams::P[i];
  v[0] -= 1;  // r is odd
  const uint32_t limb = bit >> 5, off = bit & 31;
  uint64_t two = limb < 8 ? v[limb] : 0;
  if (limb + 1 < 8) two |= (uint64_t)v[limb + 1] << 32;
  return (uint32_t)((two >> off) & ((1ull << s) - 1));
}

static bool make_config(uint32_t s, uint32_t target, Config* c) {
  if (s < 2 || s > 24 || target == 0) return false;
  c->s = s;
  const uint32_t ref_windows = div_ceil32(SCALAR_BITS, s);  // the reference's count
  const uint32_t ref_n_windows = ref_windows < target ? ref_windows : target;
  const uint32_t ref_precompute = div_ceil32(ref_windows, ref_n_windows);
  c->actual_windows = ref_windows;
  // the signed recoding must not carry out of the top window: top digit + 1 <
  // 2^(s-1); where it can (s = 2, 3) one more window is needed
  if (top_window_of_r_minus_1((c->actual_windows - 1) * s, s) + 1 >= (1u << (s - 1))) c->actual_windows += 1;
  c->n_windows = c->actual_windows < target ? c->actual_windows : target;
  // a target covering the reference's windows covers the extra one too, so the
  // caller's table count (the reference's MsmConfig) stays right
  if (target >= ref_windows) c->n_windows = c->actual_windows;
  c->n_precompute = div_ceil32(c->actual_windows, c->n_windows);
  if (c->n_precompute != ref_precompute) return false;  // would need a table the caller does not have
  c->buckets = 1u << (s - 1);
  return true;
}

// ---------------------------------------------------------------- digits

// Signed s-bit digit of window w; `carry` flows from lower windows.
__device__ __forceinline__ int next_digit(const Fr& c, uint32_t w, uint32_t s, int& carry) {
  const uint32_t bit = w * s;
  const uint32_t limb = bit >> 5, off = bit & 31;
  uint64_t two = 0;
  if (limb < 8) two = c.v[limb];
  if (limb + 1 < 8) two |= (uint64_t)c.v[limb + 1] << 32;
  int d = (int)((two >> off) & ((1ull << s) - 1)) + carry;
  if (d >= (1 << (s - 1))) {
    d -= 1 << s;
    carry = 1;
  } else {
    carry = 0;
  }
  return d;
}

// Wave-aggregated atomic increment of counter[bin] for the active lanes;
// returns each lane's previous value plus its rank among lanes with the same
// bin.  Up to 4 rounds serve the bin of the lowest active lane with a single
// atomic for all lanes sharing it; what is left (all-distinct bins) falls back
// to one atomic per lane.  This removes the serialisation on the few bins of a
// narrow top window (e.g. bits 252..253 for s = 12, 14, 18) and on a bucket that
// receives a fixed share of all scalars.  Must be called by the whole wave.
__device__ __forceinline__ uint32_t aggregated_increment(uint32_t* counter, uint32_t bin, bool active) {
  const uint32_t lane = __lane_id();
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t result = 0;
  for (int round = 0; round < 4; round++) {
    const uint64_t act = __ballot(active);
    if (act == 0) return result;
    const int leader = __ffsll((long long)act) - 1;
    const uint32_t lbin = __shfl(bin, leader);
    const uint64_t same = __ballot(active && bin == lbin);
    if (round > 0 && __popcll(same) < 2) break;
    uint32_t start = 0;
    if ((int)lane == leader) start = atomicAdd(&counter[lbin], (uint32_t)__popcll(same));
    start = __shfl(start, leader);
    if (active && bin == lbin) {
      result = start + (uint32_t)__popcll(same & below);
      active = false;
    }
  }
  if (active) result = atomicAdd(&counter[bin], 1u);
  return result;
}

// All lanes of a wave run the same number of iterations (inactive tails keep
// participating in the wave-wide ballots).
__global__ void __launch_bounds__(THREADS)
    k_histogram(const uint32_t* __restrict__ scalars, uint64_t n, Config cfg, uint32_t* __restrict__ hist) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t rounds = (n + stride - 1) / stride;
  for (uint64_t r = 0; r < rounds; r++) {
    const uint64_t i = r * stride + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const Fr c = valid ? Fr::load(scalars + i * 8).to_canonical() : Fr::zero();
    int carry = 0;
    for (uint32_t w = 0; w < cfg.actual_windows; w++) {
      const int d = next_digit(c, w, cfg.s, carry);
      const uint32_t pw = w % cfg.n_windows;
      const uint32_t bin = pw * cfg.buckets + (uint32_t)(d < 0 ? -d : d) - 1;
      aggregated_increment(hist, bin, valid && d != 0);
    }
  }
}

__global__ void __launch_bounds__(THREADS)
    k_scatter(const uint32_t* __restrict__ scalars, uint64_t n, Config cfg, uint32_t* __restrict__ cursor,
              uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t rounds = (n + stride - 1) / stride;
  for (uint64_t r = 0; r < rounds; r++) {
    const uint64_t i = r * stride + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const Fr c = valid ? Fr::load(scalars + i * 8).to_canonical() : Fr::zero();
    int carry = 0;
    for (uint32_t w = 0; w < cfg.actual_windows; w++) {
      const int d = next_digit(c, w, cfg.s, carry);
      const uint32_t pw = w % cfg.n_windows;
      const uint32_t table = w / cfg.n_windows;
      const uint32_t key = pw * cfg.buckets + (uint32_t)(d < 0 ? -d : d) - 1;
      const bool active = valid && d != 0;
      const uint32_t pos = aggregated_increment(cursor, key, active);
      if (active) {
        keys[pos] = key;
        vals[pos] = (uint32_t)(table * n + i) | (d < 0 ? 0x80000000u : 0u);
      }
    }
  }
}

// ---------------------------------------------------------------- scan (exclusive, in place)

constexpr int SCAN_BLOCK = 1024;  // elements per block: 256 threads x 4

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
  // inclusive scan inside each wave by shuffles, then the waves' totals through LDS
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (uint32_t w = 0; w < THREADS / 64; w++) {
    const uint32_t t = sh[w];
    before += w < wave ? t : 0u;
    all += t;
  }
  *total = all;
  __syncthreads();  // sh is free again for the caller
  return before + x - v;
}

__global__ void __launch_bounds__(THREADS) k_scan_blocks(uint32_t* data, uint32_t n, uint32_t* block_sums) {
  __shared__ uint32_t sh[THREADS];
  const uint32_t base = blockIdx.x * SCAN_BLOCK + threadIdx.x * 4;
  uint32_t v[4];
  for (int k = 0; k < 4; k++) v[k] = base + k < n ? data[base + k] : 0;
  const uint32_t local = v[0] + v[1] + v[2] + v[3];
  uint32_t total;
  uint32_t pre = block_exclusive_scan(local, sh, &total);
  for (int k = 0; k < 4; k++) {
    if (base + k < n) data[base + k] = pre;
    pre += v[k];
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// single block: exclusive scan of up to 256 * 16 block sums
__global__ void __launch_bounds__(THREADS) k_scan_sums(uint32_t* sums, uint32_t nb) {
  __shared__ uint32_t sh[THREADS];
  const uint32_t per = (nb + THREADS - 1) / THREADS;
  const uint32_t base = threadIdx.x * per;
  uint32_t local = 0;
  for (uint32_t k = 0; k < per; k++)
    if (base + k < nb) local += sums[base + k];
  uint32_t total;
  uint32_t pre = block_exclusive_scan(local, sh, &total);
  for (uint32_t k = 0; k < per; k++) {
    if (base + k < nb) {
      const uint32_t x = sums[base + k];
      sums[base + k] = pre;
      pre += x;
    }
  }
}

__global__ void __launch_bounds__(THREADS) k_scan_add(uint32_t* data, uint32_t n, const uint32_t* block_sums,
                                                      uint32_t* cursor) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const uint32_t v = data[i] + block_sums[i / SCAN_BLOCK];
    data[i] = v;
    if (cursor) cursor[i] = v;
  }
}

// ---------------------------------------------------------------- partition sort
//
// Sorts the (key, value) entries by key without device atomics, in two levels.
// A key is pw * buckets + |d| - 1; its coarse bin is key >> low_bits (the
// physical window and the top bits of the bucket).
//  1. k_part_count: every tile of `tp` points counts its entries per coarse
//     bin in LDS and writes the counts to matrix[bin * tiles + tile]; an
//     exclusive scan of the matrix gives every (bin, tile) run its place.
//  2. k_part_scatter: every tile ranks its entries within their bins in LDS,
//     stages them in bin order and writes each run contiguously.
//  3. k_seg_*: every coarse bin's segment is sorted by the low bits in slices
//     (LDS counting, one device atomic per nonzero bucket per slice of a
//     multi-slice segment).  The order inside a bucket is not kept; bucket
//     sums do not depend on it.

struct SortPlan {
  uint32_t low_bits;  // bits of the bucket sorted by the second level (<= 12)
  uint32_t nc;        // coarse bins: n_windows << (s - 1 - low_bits)
  uint32_t tp;        // points per tile
  uint32_t tiles;
  uint32_t lds_part;  // dynamic LDS of k_part_scatter (bytes)
};

// Counts the tile's nonzero digits per coarse bin into cnt (LDS, zeroed by the caller).
__device__ __forceinline__ void part_count_tile(const uint32_t* __restrict__ scalars, uint64_t n, const Config& cfg,
                                                uint32_t low_bits, uint64_t first, uint32_t tp, uint32_t* cnt) {
  for (uint32_t p = threadIdx.x; p < tp; p += THREADS) {
    const uint64_t i = first + p;
    if (i >= n) break;
    const Fr c = Fr::load(scalars + i * 8).to_canonical();
    int carry = 0;
    for (uint32_t w = 0; w < cfg.actual_windows; w++) {
      const int d = next_digit(c, w, cfg.s, carry);
      if (d == 0) continue;
      const uint32_t key = (w % cfg.n_windows) * cfg.buckets + (uint32_t)(d < 0 ? -d : d) - 1;
      atomicAdd(&cnt[key >> low_bits], 1u);
    }
  }
}

__global__ void __launch_bounds__(THREADS)
    k_part_count(const uint32_t* __restrict__ scalars, uint64_t n, Config cfg, SortPlan sp,
                 uint32_t* __restrict__ matrix) {
  extern __shared__ uint32_t dyn[];
  uint32_t* cnt = dyn;  // nc
  for (uint32_t b = threadIdx.x; b < sp.nc; b += THREADS) cnt[b] = 0;
  __syncthreads();
  part_count_tile(scalars, n, cfg, sp.low_bits, (uint64_t)blockIdx.x * sp.tp, sp.tp, cnt);
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < sp.nc; b += THREADS) matrix[(uint64_t)b * sp.tiles + blockIdx.x] = cnt[b];
}

// Exclusive scan of a[0..m) in LDS by the whole block; sh is THREADS words of scratch.
__device__ __forceinline__ void lds_exclusive_scan(uint32_t* a, uint32_t m, uint32_t* sh) {
  const uint32_t per = (m + THREADS - 1) / THREADS;
  const uint32_t base = threadIdx.x * per;
  uint32_t local = 0;
  for (uint32_t k = 0; k < per; k++)
    if (base + k < m) local += a[base + k];
  uint32_t total;
  uint32_t pre = block_exclusive_scan(local, sh, &total);
  for (uint32_t k = 0; k < per; k++) {
    if (base + k < m) {
      const uint32_t x = a[base + k];
      a[base + k] = pre;
      pre += x;
    }
  }
  __syncthreads();
}

// Points per thread and windows per point whose entries k_part_scatter keeps
// in registers, with each entry's rank in its bin from the counting atomic:
// one digit pass over the tile instead of a counting pass and a placing pass.
// Other tile shapes take the two passes.
constexpr uint32_t PART_PPT = 2, PART_WMAX = 22;

__global__ void __launch_bounds__(THREADS)
    k_part_scatter(const uint32_t* __restrict__ scalars, uint64_t n, Config cfg, SortPlan sp,
                   const uint32_t* __restrict__ offsets, uint32_t* __restrict__ keys_out,
                   uint32_t* __restrict__ vals_out) {
  extern __shared__ uint32_t dyn[];
  __shared__ uint32_t sh[THREADS];
  uint32_t* base = dyn;               // nc: entries of each bin, then its first staging slot
  uint32_t* cur = dyn + sp.nc;        // nc: entries of each bin / staging cursor
  uint32_t* skey = dyn + 2 * sp.nc;   // tp * actual_windows
  uint32_t* sval = skey + (size_t)sp.tp * cfg.actual_windows;
  const uint64_t first = (uint64_t)blockIdx.x * sp.tp;
  for (uint32_t b = threadIdx.x; b < sp.nc; b += THREADS) base[b] = 0;
  __syncthreads();
  const bool regs = sp.tp == PART_PPT * THREADS && cfg.actual_windows <= PART_WMAX;  // uniform
  uint32_t ek[PART_PPT][PART_WMAX], er[PART_PPT][PART_WMAX];  // key | sign << 31 (~0: none), rank
  if (regs) {
#pragma unroll
    for (uint32_t p = 0; p < PART_PPT; p++) {
      const uint64_t i = first + p * THREADS + threadIdx.x;
      const bool valid = i < n;
      const Fr c = valid ? Fr::load(scalars + i * 8).to_canonical() : Fr::zero();
      int carry = 0;
#pragma unroll
      for (uint32_t w = 0; w < PART_WMAX; w++) {
        ek[p][w] = ~0u;
        er[p][w] = 0;
        if (w < cfg.actual_windows) {
          const int d = next_digit(c, w, cfg.s, carry);
          if (valid && d != 0) {
            const uint32_t key = (w % cfg.n_windows) * cfg.buckets + (uint32_t)(d < 0 ? -d : d) - 1;
            ek[p][w] = key | (d < 0 ? 0x80000000u : 0u);
            er[p][w] = atomicAdd(&base[key >> sp.low_bits], 1u);
          }
        }
      }
    }
  } else {
    part_count_tile(scalars, n, cfg, sp.low_bits, first, sp.tp, base);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < sp.nc; b += THREADS) cur[b] = base[b];
  lds_exclusive_scan(base, sp.nc, sh);
  uint32_t staged;
  if (regs) {
#pragma unroll
    for (uint32_t p = 0; p < PART_PPT; p++) {
      const uint32_t i = (uint32_t)(first + p * THREADS + threadIdx.x);
#pragma unroll
      for (uint32_t w = 0; w < PART_WMAX; w++) {
        if (ek[p][w] == ~0u) continue;
        const uint32_t key = ek[p][w] & 0x7fffffffu;
        const uint32_t slot = base[key >> sp.low_bits] + er[p][w];
        skey[slot] = key;
        sval[slot] = (uint32_t)((w / cfg.n_windows) * n + i) | (ek[p][w] & 0x80000000u);
      }
    }
    staged = base[sp.nc - 1] + cur[sp.nc - 1];
  } else {
    for (uint32_t b = threadIdx.x; b < sp.nc; b += THREADS) cur[b] = base[b];
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < sp.tp; p += THREADS) {
      const uint64_t i = first + p;
      if (i >= n) break;
      const Fr c = Fr::load(scalars + i * 8).to_canonical();
      int carry = 0;
      for (uint32_t w = 0; w < cfg.actual_windows; w++) {
        const int d = next_digit(c, w, cfg.s, carry);
        if (d == 0) continue;
        const uint32_t key = (w % cfg.n_windows) * cfg.buckets + (uint32_t)(d < 0 ? -d : d) - 1;
        const uint32_t slot = atomicAdd(&cur[key >> sp.low_bits], 1u);
        skey[slot] = key;
        sval[slot] = (uint32_t)((w / cfg.n_windows) * n + i) | (d < 0 ? 0x80000000u : 0u);
      }
    }
    __syncthreads();
    staged = cur[sp.nc - 1];  // the last bin's cursor ends at the tile's entry count
  }
  __syncthreads();
  for (uint32_t slot = threadIdx.x; slot < staged; slot += THREADS) {
    const uint32_t key = skey[slot];
    const uint32_t b = key >> sp.low_bits;
    const uint32_t dst = offsets[(uint64_t)b * sp.tiles + blockIdx.x] + (slot - base[b]);
    keys_out[dst] = key;
    vals_out[dst] = sval[slot];
  }
}

// Second level: every coarse bin's segment is sorted by the low bits in slices
// of at most SEG_SLICE entries, so a segment holding most entries of a narrow
// window (s = 12, 13, 14, 18 and the small windows) is spread over many
// workgroups.  Bucket counts of a one-slice segment are stored, those of a
// longer one added up with one device atomic per nonzero bucket per slice.
constexpr uint32_t SEG_SLICE = 8192;  // entries per workgroup, staged in LDS at once (64 KB)
constexpr uint32_t SEG_PER_THREAD = SEG_SLICE / THREADS;

// largest k < n with prefix[k] <= g (prefix: n + 1 non-decreasing entries)
__device__ __forceinline__ uint32_t find_segment(const uint32_t* prefix, uint32_t n, uint32_t g) {
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

__device__ __forceinline__ uint32_t seg_begin(const uint32_t* offsets, const SortPlan& sp, uint32_t cb) {
  return offsets[(uint64_t)cb * sp.tiles];
}

// slice_prefix[cb] = slices of the segments before cb (nc + 1 entries); one workgroup.
__global__ void __launch_bounds__(THREADS)
    k_seg_plan(const uint32_t* __restrict__ offsets, SortPlan sp, uint32_t* __restrict__ slice_prefix) {
  __shared__ uint32_t sh[THREADS];
  extern __shared__ uint32_t cnt[];  // nc + 1
  for (uint32_t cb = threadIdx.x; cb <= sp.nc; cb += THREADS) {
    uint32_t k = 0;
    if (cb < sp.nc) {
      const uint32_t size = seg_begin(offsets, sp, cb + 1) - seg_begin(offsets, sp, cb);
      k = (size + SEG_SLICE - 1) / SEG_SLICE;
    }
    cnt[cb] = k;
  }
  __syncthreads();
  lds_exclusive_scan(cnt, sp.nc + 1, sh);
  for (uint32_t cb = threadIdx.x; cb <= sp.nc; cb += THREADS) slice_prefix[cb] = cnt[cb];
}

struct SegSlice {
  uint32_t cb, lo, hi;  // segment and entry range [lo, hi)
  bool single;          // the segment is one slice
};

__device__ __forceinline__ bool seg_slice(const uint32_t* offsets, const uint32_t* slice_prefix, const SortPlan& sp,
                                          SegSlice* out) {
  const uint32_t b = blockIdx.x;
  if (b >= slice_prefix[sp.nc]) return false;
  const uint32_t cb = find_segment(slice_prefix, sp.nc, b);
  const uint32_t start = seg_begin(offsets, sp, cb), end = seg_begin(offsets, sp, cb + 1);
  const uint32_t j = b - slice_prefix[cb];
  out->cb = cb;
  out->lo = start + j * SEG_SLICE;
  out->hi = out->lo + SEG_SLICE < end ? out->lo + SEG_SLICE : end;
  out->single = slice_prefix[cb + 1] - slice_prefix[cb] == 1;
  return true;
}

// ghist[cb * nb + bin] (zeroed beforehand) = entries of the segment in each bucket.
__global__ void __launch_bounds__(THREADS)
    k_seg_hist(const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ offsets,
               const uint32_t* __restrict__ slice_prefix, SortPlan sp, uint32_t* __restrict__ ghist) {
  extern __shared__ uint32_t h[];  // nb
  SegSlice sl;
  if (!seg_slice(offsets, slice_prefix, sp, &sl)) return;  // uniform per block
  const uint32_t nb = 1u << sp.low_bits, mask = nb - 1;
  for (uint32_t b = threadIdx.x; b < nb; b += THREADS) h[b] = 0;
  __syncthreads();
  for (uint32_t e = sl.lo + threadIdx.x; e < sl.hi; e += THREADS) atomicAdd(&h[keys_in[e] & mask], 1u);
  __syncthreads();
  uint32_t* g = ghist + (uint64_t)sl.cb * nb;
  for (uint32_t b = threadIdx.x; b < nb; b += THREADS) {
    const uint32_t c = h[b];
    if (sl.single)
      g[b] = c;
    else if (c)
      atomicAdd(&g[b], c);
  }
}

// Per segment: ghist becomes the scatter cursor, the first entry of each bucket.
__global__ void __launch_bounds__(THREADS)
    k_seg_scan(const uint32_t* __restrict__ offsets, SortPlan sp, uint32_t* __restrict__ ghist) {
  extern __shared__ uint32_t h[];  // nb
  __shared__ uint32_t sh[THREADS];
  const uint32_t nb = 1u << sp.low_bits;
  const uint32_t cb = blockIdx.x;
  const uint32_t start = seg_begin(offsets, sp, cb);
  uint32_t* g = ghist + (uint64_t)cb * nb;
  for (uint32_t b = threadIdx.x; b < nb; b += THREADS) h[b] = g[b];
  __syncthreads();
  lds_exclusive_scan(h, nb, sh);
  for (uint32_t b = threadIdx.x; b < nb; b += THREADS) g[b] = start + h[b];
}

// Every slice takes its run of each bucket from the cursor (one atomic per
// nonzero bucket; a one-slice segment owns the whole bucket) and moves its
// entries in one round: each thread holds its entries in registers with their
// rank inside their bucket (the value the counting atomic returned), stages
// them in bucket order in LDS, and the slice is written out so that each
// bucket's entries land as one contiguous run (coalesced stores instead of one
// scattered store per entry).  The order inside a bucket is not kept.
__global__ void __launch_bounds__(THREADS)
    k_seg_scatter(const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
                  const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ slice_prefix, SortPlan sp,
                  uint32_t* __restrict__ cursor, uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out) {
  extern __shared__ uint32_t h[];  // cnt, lbase, gcur (nb each), then the staged keys and values
  __shared__ uint32_t sh[THREADS];
  SegSlice sl;
  if (!seg_slice(offsets, slice_prefix, sp, &sl)) return;  // uniform per block
  const uint32_t nb = 1u << sp.low_bits, mask = nb - 1;
  uint32_t* cnt = h;
  uint32_t* lbase = h + nb;
  uint32_t* gcur = h + 2 * nb;
  uint32_t* skey = h + 3 * nb;
  uint32_t* sval = skey + SEG_SLICE;
  for (uint32_t b = threadIdx.x; b < nb; b += THREADS) cnt[b] = 0;
  __syncthreads();
  uint32_t key[SEG_PER_THREAD], val[SEG_PER_THREAD], rank[SEG_PER_THREAD];
#pragma unroll
  for (uint32_t k = 0; k < SEG_PER_THREAD; k++) {
    const uint32_t e = sl.lo + k * THREADS + threadIdx.x;
    if (e < sl.hi) {
      key[k] = keys_in[e];
      val[k] = vals_in[e];
      rank[k] = atomicAdd(&cnt[key[k] & mask], 1u);
    }
  }
  __syncthreads();
  uint32_t* cur = cursor + (uint64_t)sl.cb * nb;
  for (uint32_t b = threadIdx.x; b < nb; b += THREADS) {
    const uint32_t c = cnt[b];
    gcur[b] = sl.single ? cur[b] : (c ? atomicAdd(&cur[b], c) : 0u);
    lbase[b] = c;
  }
  __syncthreads();
  lds_exclusive_scan(lbase, nb, sh);
#pragma unroll
  for (uint32_t k = 0; k < SEG_PER_THREAD; k++) {
    const uint32_t e = sl.lo + k * THREADS + threadIdx.x;
    if (e < sl.hi) {
      const uint32_t slot = lbase[key[k] & mask] + rank[k];
      skey[slot] = key[k];
      sval[slot] = val[k];
    }
  }
  __syncthreads();
  const uint32_t n = sl.hi - sl.lo;
  for (uint32_t slot = threadIdx.x; slot < n; slot += THREADS) {
    const uint32_t k = skey[slot];
    const uint32_t b = k & mask;
    const uint32_t dst = gcur[b] + (slot - lbase[b]);
    keys_out[dst] = k;
    vals_out[dst] = sval[slot];
  }
}

// Dynamic LDS of k_seg_scatter for 2^low_bits buckets per segment.
static inline uint32_t seg_scatter_lds(uint32_t low_bits) { return (3u * (1u << low_bits) + 2u * SEG_SLICE) * 4u; }

// ---------------------------------------------------------------- accumulation

// Closes a bucket.  Every bucket is completed at exactly one level (its whole
// run, or all its partials, inside one chunk), so in the first part of an MSM
// (fresh buckets hold the identity) the sum is stored, in later parts added.
// Identity sums are skipped: slots without a partial carry a key too.
__device__ __forceinline__ void bucket_close(u32* buckets, uint32_t key, const G1X29& sum, bool fresh) {
  if (sum.is_identity()) return;
  u32* p = buckets + (size_t)key * 32;
  if (fresh) {
    sum.store(p);
    return;
  }
  const G1X29 cur = G1X29::load(p);
  (cur.is_identity() ? sum : cur.add(sum)).store(p);
}

__device__ __forceinline__ Affine29 fetch_point(const u32* __restrict__ points, uint32_t v) {
  return Affine29::load(points + (size_t)(v & 0x7fffffffu) * 16, (v & 0x80000000u) != 0);
}

// One accumulation level.  Thread t owns sorted entries [t*CH, ...) and sums
// each run of equal keys, one entry per loop step in every lane (a wave takes
// CH steps whatever the run lengths).  A run closed on both sides goes to its bucket;
// the chunk's first run (if it continues from the previous chunk) and last run
// (if it continues into the next) are emitted as (key, partial) at slots 2t and
// 2t + 1 of the next level, identity where there is nothing to carry, so every
// key's pieces stay contiguous.
template <bool FROM_POINTS, uint32_t CH>
__device__ __forceinline__ void accumulate_chunk(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                 const u32* __restrict__ src, uint32_t* counts, uint32_t level, bool fresh,
                                                 u32* buckets, uint32_t* __restrict__ next_keys,
                                                 u32* __restrict__ next_pts) {
  const uint32_t N = counts[level];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nchunks = (N + CH - 1) / CH;
  if (t == 0) counts[level + 1] = 2 * nchunks;
  if (t >= nchunks) return;
  const uint32_t start = t * CH;
  const uint32_t end = start + CH < N ? start + CH : N;
  const uint32_t key_first = keys[start], key_last = keys[end - 1];
  const bool open_left = start > 0 && keys[start - 1] == key_first;
  const bool open_right = end < N && keys[end] == key_last;
  bool a_done = false, b_done = false;
  G1X29 acc = G1X29::identity();
  uint32_t run_start = start;
  for (uint32_t i = start; i < end; i++) {
    const uint32_t key = keys[i];
    if (FROM_POINTS)
      acc = acc.add_affine(fetch_point(src, vals[i]));
    else
      acc = acc.add(G1X29::load(src + (size_t)i * 32));
    const bool run_end = (i + 1 == end) || keys[i + 1] != key;
    if (run_end) {
      const bool first = run_start == start, last = i + 1 == end;
      if (first && open_left) {
        if (next_pts) acc.store(next_pts + (size_t)(2 * t) * 32);
        a_done = true;
      } else if (last && open_right) {
        if (next_pts) acc.store(next_pts + (size_t)(2 * t + 1) * 32);
        b_done = true;
      } else {
        bucket_close(buckets, key, acc, fresh);
      }
      acc = G1X29::identity();
      run_start = i + 1;
    }
  }
  if (next_keys) {
    next_keys[2 * t] = key_first;
    next_keys[2 * t + 1] = key_last;
    if (!a_done) G1X29::identity().store(next_pts + (size_t)(2 * t) * 32);
    if (!b_done) G1X29::identity().store(next_pts + (size_t)(2 * t + 1) * 32);
  }
}

// Level 0: entries are (key, point index | sign << 31).  Three waves per SIMD
// (168 VGPRs, a few values spilled) run it faster than two without spills: the
// dependent loads of an entry (key, point index, point) wait less.
__global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(3)))
    k_accumulate0(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                  const u32* __restrict__ points, uint32_t* counts, bool fresh, u32* buckets,
                  uint32_t* __restrict__ next_keys, u32* __restrict__ next_pts) {
  accumulate_chunk<true, CHUNK>(keys, vals, points, counts, 0, fresh, buckets, next_keys, next_pts);
}

// Level >= 1: entries are (key, XYZZ partial).  With run_flag given, the
// level runs only when *run_flag is set (k_close_runs found a long run).
__global__ void __launch_bounds__(THREADS)
    k_accumulate(const uint32_t* __restrict__ keys, const u32* __restrict__ pts, uint32_t* counts, uint32_t level,
                 bool fresh, u32* buckets, uint32_t* __restrict__ next_keys, u32* __restrict__ next_pts,
                 const uint32_t* __restrict__ run_flag) {
  if (run_flag && *run_flag == 0u) return;
  accumulate_chunk<false, CHUNK_L>(keys, nullptr, pts, counts, level, fresh, buckets, next_keys, next_pts);
}

// After the first chunked level the open pieces of a bucket form short runs
// of equal keys (for scalars spread over the buckets, a bucket that crosses
// chunk edges leaves two to four slots), which fixed chunk edges would keep
// deferring level after level.  Here the thread at the first slot of every
// run of at most RUN_MAX slots adds the run, closes the bucket and marks the
// slots consumed (ZZ = 0, the identity), so the bucket is completed once.  A
// longer run (a bucket holding a large share of the points) raises *run_flag
// and stays for the chunked levels, which see the consumed slots as identity.
constexpr uint32_t RUN_MAX = 16;

__global__ void __launch_bounds__(THREADS)
    k_close_runs(const uint32_t* __restrict__ keys, u32* __restrict__ pts, const uint32_t* __restrict__ counts,
                 uint32_t level, bool fresh, u32* buckets, uint32_t* __restrict__ run_flag) {
  const uint32_t N = counts[level];
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const uint32_t key = keys[s];
  if (s > 0 && keys[s - 1] == key) return;  // not the first slot of its run
  uint32_t e = s + 1;
  while (e < N && e - s <= RUN_MAX && keys[e] == key) e++;
  if (e - s > RUN_MAX) {
    *run_flag = 1u;
    return;
  }
  G1X29 acc = G1X29::identity();
  for (uint32_t i = s; i < e; i++) {
    u32* p = pts + (size_t)i * 32;
    acc = acc.add(G1X29::load(p));
    Fq::zero().store(p + 16);
  }
  bucket_close(buckets, key, acc, fresh);
}

__global__ void __launch_bounds__(THREADS) k_zero_u32(uint32_t* p, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

__global__ void __launch_bounds__(THREADS) k_fill_identity(u32* pts, uint64_t count) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint4 z = make_uint4(0, 0, 0, 0);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count * 8; i += stride)
    reinterpret_cast<uint4*>(pts)[i] = z;
}

// dst[0, n16) = src[0, n16) in 16-byte words, src being pinned host memory
// mapped into the device's address space: the host-to-device copies of the
// drop-in msm's points and scalars.  A grid of PULL_BLOCKS workgroups with four
// loads in flight per lane keeps the PCIe link full (57 GB/s measured on
// MI355X for 64 MiB, the link's rate) on a small share of the CUs, so the
// compute stream's sort runs beside it.  DMA copies reach the same rate on some
// copy engines and about half of it on others, and which engine serves a stream
// depends on what the process did before (a one-part 2^20 call took 3.8 ms or
// 4.9 ms that way); the kernel's rate does not.
constexpr uint32_t PULL_BLOCKS = 64;
constexpr int PULL_UNROLL = 4;
__global__ void __launch_bounds__(THREADS) k_pull_host(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                       uint64_t n16) {
  const uint64_t stride = (uint64_t)gridDim.x * THREADS * PULL_UNROLL;
  for (uint64_t base = (uint64_t)blockIdx.x * THREADS * PULL_UNROLL + threadIdx.x; base < n16; base += stride) {
    uint4 v[PULL_UNROLL];
#pragma unroll
    for (int u = 0; u < PULL_UNROLL; u++) {
      const uint64_t i = base + (uint64_t)u * THREADS;
      if (i < n16) v[u] = src[i];
    }
#pragma unroll
    for (int u = 0; u < PULL_UNROLL; u++) {
      const uint64_t i = base + (uint64_t)u * THREADS;
      if (i < n16) dst[i] = v[u];
    }
  }
}

// dst[i] += src[i] over H buckets (XYZZ): a later part's buckets, accumulated
// into a scratch set of their own, join the batch's buckets in one pass with
// no divergence, instead of every run of that part reading and adding its
// bucket inside the accumulation loop.
__global__ void __launch_bounds__(THREADS) k_merge_buckets(u32* __restrict__ dst, const u32* __restrict__ src,
                                                           uint32_t H) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= H) return;
  const G1X29 b = G1X29::load(src + (size_t)i * 32);
  if (b.is_identity()) return;
  u32* d = dst + (size_t)i * 32;
  const G1X29 a = G1X29::load(d);
  (a.is_identity() ? b : a.add(b)).store(d);
}

__global__ void k_set_count(uint32_t* counts, const uint32_t* offsets, uint32_t h) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counts[0] = offsets[h];
}

// ---------------------------------------------------------------- bucket reduction
//
// Per window: S = sum_j j X_j + sum_j Y_j over j < M (M a power of two).  A
// halving level maps the pair (2i, 2i + 1) to
//   X'_i = 2 (X_2i + X_2i+1),   Y'_i = X_2i+1 + Y_2i + Y_2i+1,
// which keeps S = sum_i i X'_i + sum_i Y'_i over M / 2 items; at M = 1,
// S = Y_0.  Level 0 has X = Y = buckets, bucket b at j = b - 1, which makes
// S = sum_b b B_b.  X' and Y' are computed by different threads (different
// waves), so a level is two point operations deep.

// The first two levels at once, one thread per block of four buckets
// B_0..B_3 (items 4i..4i+3 of a window): X'' = 4T with T = B_0 + .. + B_3, and
// Y'' = B_0 + 2B_1 + 3B_2 + 4B_3 = T + s_1 + s_2 + s_3 with the suffix sums
// s_r = B_r + .. + B_3, so that S = sum_i i X''_i + sum_i Y''_i over M / 4
// items.  Six additions and two doublings (102 products) per four buckets,
// where two halving levels take 51 products per pair, 153 per four buckets.
__global__ void __launch_bounds__(THREADS)
    k_reduce_block4(const u32* __restrict__ B, uint32_t M, uint32_t n_windows, u32* __restrict__ Xo,
                    u32* __restrict__ Yo) {
  const uint32_t blocks = (M >> 2) * n_windows;
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= blocks) return;
  const size_t j0 = (size_t)gid * 4;  // windows are consecutive runs of M items: block gid covers 4 gid..4 gid+3
  const G1X29 b3 = G1X29::load(B + (j0 + 3) * 32);
  const G1X29 s2 = G1X29::load(B + (j0 + 2) * 32).add(b3);
  const G1X29 u = s2.add(b3);  // s_2 + s_3
  const G1X29 s1 = G1X29::load(B + (j0 + 1) * 32).add(s2);
  const G1X29 t = G1X29::load(B + j0 * 32).add(s1);
  t.add(s1).add(u).store(Yo + (size_t)gid * 32);
  t.dbl().dbl().store(Xo + (size_t)gid * 32);
}

// One level over all windows: threads [0, Mo * n_windows) write X', the next
// Mo * n_windows threads write Y' (Mo = M / 2).
__global__ void __launch_bounds__(THREADS)
    k_reduce_level(const u32* __restrict__ X, const u32* __restrict__ Y, uint32_t M, uint32_t n_windows,
                   bool first, u32* __restrict__ Xo, u32* __restrict__ Yo) {
  const uint32_t Mo = M >> 1;
  const uint32_t items = Mo * n_windows;
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= 2 * items) return;
  const bool yside = gid >= items;
  const uint32_t it = yside ? gid - items : gid;
  const size_t j0 = (size_t)(it / Mo) * M + 2 * (it % Mo);
  const G1X29 x1 = G1X29::load(X + (j0 + 1) * 32);
  if (!yside) {
    G1X29::load(X + j0 * 32).add(x1).dbl().store(Xo + (size_t)it * 32);
  } else {
    const u32* src = first ? X : Y;
    const G1X29 y1 = first ? x1 : G1X29::load(Y + (j0 + 1) * 32);
    G1X29::load(src + j0 * 32).add(y1).add(x1).store(Yo + (size_t)it * 32);
  }
}

// The same level with a quad of lanes per point operation (quad29_add /
// quad29_dbl): for the narrow levels, where the operations are few and each
// one's latency, not the throughput, sets the time.
__global__ void __launch_bounds__(THREADS)
    k_reduce_level_quad(const u32* __restrict__ X, const u32* __restrict__ Y, uint32_t M, uint32_t n_windows,
                        bool first, u32* __restrict__ Xo, u32* __restrict__ Yo) {
  const uint32_t Mo = M >> 1;
  const uint32_t items = Mo * n_windows;
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t op = gid >> 2, q = gid & 3;
  if (op >= 2 * items) return;  // whole quads: 2 * items * 4 threads are launched in multiples of 4
  const bool yside = op >= items;
  const uint32_t it = yside ? op - items : op;
  const size_t j0 = (size_t)(it / Mo) * M + 2 * (it % Mo);
  const G1X29 x1 = G1X29::load(X + (j0 + 1) * 32);
  G1X29 r;
  if (!yside) {
    r = quad29_dbl(quad29_add(G1X29::load(X + j0 * 32), x1, q), q);
  } else {
    const u32* src = first ? X : Y;
    const G1X29 y1 = first ? x1 : G1X29::load(Y + (j0 + 1) * 32);
    r = quad29_add(quad29_add(G1X29::load(src + j0 * 32), y1, q), x1, q);
  }
  if (q == 0) r.store((yside ? Yo : Xo) + (size_t)it * 32);
}

// The last levels of window blockIdx.x (M <= RED_TAIL items) in LDS, all in
// one launch, a quad of lanes per point operation: QUAD_TAIL_THREADS threads,
// the operations writing X' on the lower half of them; out[w] = S_w.  At
// 512 threads the kernel may keep 256 VGPRs, which the quad's operand copies
// need (at 1024 it spilled).
constexpr uint32_t QUAD_TAIL_THREADS = 4 * RED_TAIL;
__global__ void __launch_bounds__(QUAD_TAIL_THREADS)
    k_reduce_tail_quad(const u32* __restrict__ X, const u32* __restrict__ Y, uint32_t M, bool first,
                       u32* __restrict__ out) {
  __shared__ G1X29 sx[RED_TAIL], sy[RED_TAIL];
  const size_t base = (size_t)blockIdx.x * M;
  for (uint32_t j = threadIdx.x; j < M; j += QUAD_TAIL_THREADS) {
    sx[j] = G1X29::load(X + (base + j) * 32);
    sy[j] = first ? sx[j] : G1X29::load(Y + (base + j) * 32);
  }
  __syncthreads();
  const uint32_t op = threadIdx.x >> 2, q = threadIdx.x & 3;
  const bool yside = op >= RED_TAIL / 2;
  const uint32_t i = yside ? op - RED_TAIL / 2 : op;
  for (uint32_t P = M >> 1; P >= 1; P >>= 1) {
    const bool act = i < P;  // the same for the four lanes of a quad
    G1X29 r;
    if (act) {
      const G1X29 x1 = sx[2 * i + 1];
      r = yside ? quad29_add(quad29_add(sy[2 * i], sy[2 * i + 1], q), x1, q)
                : quad29_dbl(quad29_add(sx[2 * i], x1, q), q);
    }
    __syncthreads();
    if (act && q == 0) {
      if (yside)
        sy[i] = r;
      else
        sx[i] = r;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) sy[0].store_mont256(out + (size_t)blockIdx.x * 32);  // the host reads G1XYZZ
}

// ---------------------------------------------------------------- precompute

// table j (1 <= j < n_precompute) of point i = 2^(j * shift) * P_i, affine
__global__ void __launch_bounds__(THREADS)
    k_precompute(u32* points, uint64_t len, uint32_t n_precompute, uint32_t shift) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= len) return;
  G1XYZZ p = G1XYZZ::from_affine(G1Affine::load(points + gid * 16));
  for (uint32_t j = 1; j < n_precompute; j++) {
    for (uint32_t k = 0; k < shift; k++) p = p.dbl();
    p.to_affine().store(points + (gid + j * len) * 16);
  }
}

// out[i] = k_i * G for G = (1, 2), k_i given as Fr Montgomery-form elements;
// double-and-add over the canonical bits, affine output.
__global__ void __launch_bounds__(THREADS)
    k_mul_generator(const u32* __restrict__ scalars, uint64_t n, u32* __restrict__ out) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n) return;
  const Fr k = Fr::load(scalars + gid * 8).to_canonical();
  Fq one = Fq::one();
  const G1Affine g{one, one + one};
  G1XYZZ acc = G1XYZZ::identity();
  for (int bit = 253; bit >= 0; bit--) {
    acc = acc.dbl();
    if ((k.v[bit >> 5] >> (bit & 31)) & 1) acc = acc.add_affine(g);
  }
  acc.to_affine().store(out + gid * 16);
}

// ---------------------------------------------------------------- host orchestration

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Device workspace for one MSM of `len` points (per table) and config cfg.
struct Layout {
  size_t hist, block_sums, cursor, keys0, vals0, counts, lvl_keys[2], lvl_pts[2], buckets, red[4], sums, total;
  size_t matrix, keys_tmp, vals_tmp, slices;  // partition sort (when `part`)
  uint32_t max_slices;
  bool part;
  SortPlan sp;
  uint64_t matrix_n;  // nc * tiles + 1
  uint32_t nb_matrix;
  uint32_t H, nb_scan;
  uint64_t n_entries_max;
  uint32_t level_max[MAX_LEVELS];
  int levels;  // accumulation levels after level 0
  uint32_t red_m[MAX_LEVELS];
  int red_levels;  // one-level launches before the in-workgroup tail
  size_t bucket_bytes;  // one bucket set (n_windows * 2^(s-1) XYZZ points)
  uint32_t sets;        // bucket sets from `buckets` on (one per batch of a run)
};

// Partition-sort parameters for `len` points; false when the window is too wide
// for the LDS tables (then the atomic counting sort runs instead).
constexpr uint32_t PART_LDS_MAX = 96 * 1024;
static bool make_sort_plan(const Config& cfg, uint64_t len, SortPlan* sp) {
  const uint32_t bits = cfg.s - 1;  // bucket bits
  uint32_t cb = bits < 5 ? bits : 5;
  uint32_t low = bits - cb;
  if (low > 12) {
    low = 12;
    cb = bits - 12;
  }
  if (cb > 7) return false;  // windows above 20 bits
  sp->low_bits = low;
  sp->nc = cfg.n_windows << cb;
  uint32_t tp = 1024;
  while (tp > 64 && (2ull * sp->nc + 2ull * tp * cfg.actual_windows) * 4 > PART_LDS_MAX) tp >>= 1;
  if ((2ull * sp->nc + 2ull * tp * cfg.actual_windows) * 4 > PART_LDS_MAX) return false;
  sp->tp = tp;
  sp->tiles = (uint32_t)((len + tp - 1) / tp);
  sp->lds_part = (uint32_t)((2ull * sp->nc + 2ull * tp * cfg.actual_windows) * 4);
  return sp->tiles > 0;
}

static bool make_layout(const Config& cfg, uint64_t len, Layout* L, uint32_t sets = 1) {
  memset(L, 0, sizeof(*L));
  if (sets == 0) return false;
  L->sets = sets;
  L->H = cfg.n_windows * cfg.buckets;
  L->nb_scan = div_ceil32((uint64_t)L->H + 1, SCAN_BLOCK);
  if (L->nb_scan > THREADS * 1024) return false;
  L->part = len > 0 && make_sort_plan(cfg, len, &L->sp);
  if (L->part) {
    L->matrix_n = (uint64_t)L->sp.nc * L->sp.tiles + 1;
    if (L->matrix_n >= 0xffffffffull) return false;
    L->nb_matrix = div_ceil32(L->matrix_n, SCAN_BLOCK);
    if (L->nb_matrix > THREADS * 1024) L->part = false;
  }
  L->n_entries_max = len * cfg.actual_windows;
  if (len * cfg.n_precompute >= 0x80000000ull || L->n_entries_max >= 0xffffffffull) return false;
  // accumulation level sizes (upper bounds)
  uint64_t m = L->n_entries_max;
  int lv = 0;
  uint64_t max_even = 0, max_odd = 0;
  uint32_t chunk = CHUNK;  // level 0, then CHUNK_L
  while (m > chunk) {
    m = 2 * ((m + chunk - 1) / chunk);
    chunk = CHUNK_L;
    if (lv >= MAX_LEVELS - 1) return false;
    L->level_max[lv] = (uint32_t)m;
    if (lv % 2 == 0) max_even = m > max_even ? m : max_even;
    else max_odd = m > max_odd ? m : max_odd;
    lv++;
  }
  L->levels = lv;
  if (max_even < 2) max_even = 2;
  if (max_odd < 2) max_odd = 2;
  // reduction level sizes (halving down to RED_TAIL items per window)
  uint32_t M = cfg.buckets;
  int rl = 0;
  uint64_t red_max0 = 0, red_max1 = 0;
  while (M > RED_TAIL) {
    if (rl >= MAX_LEVELS) return false;
    M >>= 1;
    L->red_m[rl] = M;
    if (rl % 2 == 0) red_max0 = M > red_max0 ? M : red_max0;
    else red_max1 = M > red_max1 ? M : red_max1;
    rl++;
  }
  L->red_levels = rl;
  size_t off = 0;
  L->hist = off; off = align256(off + ((size_t)L->H + 1) * 4);
  L->cursor = off; off = align256(off + ((size_t)L->H + 1) * 4);
  L->block_sums = off; off = align256(off + (size_t)(L->nb_scan > L->nb_matrix ? L->nb_scan : L->nb_matrix) * 4);
  if (L->part) {
    L->matrix = off; off = align256(off + (size_t)L->matrix_n * 4);
    L->keys_tmp = off; off = align256(off + L->n_entries_max * 4);
    L->vals_tmp = off; off = align256(off + L->n_entries_max * 4);
    L->slices = off; off = align256(off + ((size_t)L->sp.nc + 1) * 4);
    L->max_slices = (uint32_t)((L->n_entries_max + SEG_SLICE - 1) / SEG_SLICE) + L->sp.nc;
  }
  L->counts = off; off = align256(off + MAX_LEVELS * 4 * 2);
  L->keys0 = off; off = align256(off + L->n_entries_max * 4);
  L->vals0 = off; off = align256(off + L->n_entries_max * 4);
  L->lvl_keys[0] = off; off = align256(off + max_even * 4);
  L->lvl_pts[0] = off; off = align256(off + max_even * 128);
  L->lvl_keys[1] = off; off = align256(off + max_odd * 4);
  L->lvl_pts[1] = off; off = align256(off + max_odd * 128);
  L->bucket_bytes = align256((size_t)L->H * 128);
  L->buckets = off; off += L->bucket_bytes * sets;
  // reduction ping-pong: X and Y for even / odd levels
  L->red[0] = off; off = align256(off + (size_t)cfg.n_windows * red_max0 * 128);
  L->red[1] = off; off = align256(off + (size_t)cfg.n_windows * red_max0 * 128);
  L->red[2] = off; off = align256(off + (size_t)cfg.n_windows * (red_max1 ? red_max1 : 1) * 128);
  L->red[3] = off; off = align256(off + (size_t)cfg.n_windows * (red_max1 ? red_max1 : 1) * 128);
  L->sums = off; off = align256(off + (size_t)cfg.n_windows * 128);
  L->total = off;
  return true;
}

static inline unsigned grid_for(uint64_t work, unsigned cap) {
  uint64_t g = (work + THREADS - 1) / THREADS;
  if (g > cap) g = cap;
  return g == 0 ? 1u : (unsigned)g;
}

// The partition and segment scatters take more than 64 KB of dynamic LDS; the
// permission is set once per device, at each kernel's largest request, on the
// first MSM there (outside any graph capture a caller may start later).
static std::mutex g_attr_mu;
static bool g_attr_done[64];

static hipError_t ensure_lds_limits() {
  int dev = 0;
  ZKP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_attr_mu);
  if (g_attr_done[dev]) return hipSuccess;
  ZKP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_part_scatter),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)PART_LDS_MAX));
  ZKP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_seg_scatter),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)seg_scatter_lds(12)));
  g_attr_done[dev] = true;
  return hipSuccess;
}

struct PhaseTimer {
  bool on;
  hipStream_t s;
  std::vector<hipEvent_t> ev;
  std::vector<const char*> names;
  void mark(const char* name) {
    if (!on) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s);
    ev.push_back(e);
    names.push_back(name);
  }
  void report() {
    if (!on || ev.size() < 2) return;
    (void)hipEventSynchronize(ev.back());
    for (size_t i = 1; i < ev.size(); i++) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, ev[i - 1], ev[i]);
      fprintf(stderr, "[msm] %-12s %.4f ms\n", names[i], ms);
    }
    for (auto e : ev) (void)hipEventDestroy(e);
    ev.clear();
    names.clear();
  }
};

static inline u32* bucket_set(const Layout& L, char* ws, uint32_t set) {
  return reinterpret_cast<u32*>(ws + L.buckets + (size_t)set * L.bucket_bytes);
}

// Zeroes bucket set `set` (fill kernel, safe under graph capture).
static hipError_t reset_buckets(const Config& cfg, const Layout& L, char* ws, uint32_t set, hipStream_t stream) {
  hipLaunchKernelGGL(k_fill_identity, dim3(grid_for((uint64_t)L.H * 8, 16384)), dim3(THREADS), 0, stream,
                     bucket_set(L, ws, set), (uint64_t)L.H);
  return hipGetLastError();
}

// First half of a part: the signed digits of its scalars sorted by bucket into
// keys0 / vals0 (point index and sign), with the entry counts in `counts`.
// Reads only the scalars, so it can run while the part's points are still
// crossing the link.
static hipError_t sort_part(const Config& cfg, const Layout& L, char* ws, const u32* d_scalars, uint64_t len,
                            hipStream_t stream, PhaseTimer& pt) {
  ZKP_CHECK(ensure_lds_limits());
  uint32_t* hist = reinterpret_cast<uint32_t*>(ws + L.hist);
  uint32_t* cursor = reinterpret_cast<uint32_t*>(ws + L.cursor);
  uint32_t* bsums = reinterpret_cast<uint32_t*>(ws + L.block_sums);
  uint32_t* counts = reinterpret_cast<uint32_t*>(ws + L.counts);
  uint32_t* keys0 = reinterpret_cast<uint32_t*>(ws + L.keys0);
  uint32_t* vals0 = reinterpret_cast<uint32_t*>(ws + L.vals0);
  if (L.part) {
    // partition sort (no device atomics); `len` may be below the planned maximum
    SortPlan sp = L.sp;
    sp.tiles = (uint32_t)((len + sp.tp - 1) / sp.tp);
    const uint64_t mn = (uint64_t)sp.nc * sp.tiles + 1;
    uint32_t* matrix = reinterpret_cast<uint32_t*>(ws + L.matrix);
    uint32_t* ktmp = reinterpret_cast<uint32_t*>(ws + L.keys_tmp);
    uint32_t* vtmp = reinterpret_cast<uint32_t*>(ws + L.vals_tmp);
    pt.mark("start");
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(THREADS), 0, stream, matrix + (mn - 1), (uint64_t)1);
    hipLaunchKernelGGL(k_part_count, dim3(sp.tiles), dim3(THREADS), sp.nc * 4, stream, d_scalars, len, cfg, sp, matrix);
    ZKP_CHECK(hipGetLastError());
    pt.mark("count");
    const uint32_t nbm = div_ceil32(mn, SCAN_BLOCK);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nbm), dim3(THREADS), 0, stream, matrix, (uint32_t)mn, bsums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(THREADS), 0, stream, bsums, nbm);
    hipLaunchKernelGGL(k_scan_add, dim3(div_ceil32(mn, THREADS)), dim3(THREADS), 0, stream, matrix, (uint32_t)mn, bsums,
                       nullptr);
    hipLaunchKernelGGL(k_set_count, dim3(1), dim3(64), 0, stream, counts, matrix, (uint32_t)(mn - 1));
    ZKP_CHECK(hipGetLastError());
    pt.mark("scan");
    hipLaunchKernelGGL(k_part_scatter, dim3(sp.tiles), dim3(THREADS), sp.lds_part, stream, d_scalars, len, cfg, sp,
                       matrix, ktmp, vtmp);
    ZKP_CHECK(hipGetLastError());
    pt.mark("partition");
    const uint32_t nbk = 1u << sp.low_bits;
    uint32_t* slices = reinterpret_cast<uint32_t*>(ws + L.slices);
    hipLaunchKernelGGL(k_seg_plan, dim3(1), dim3(THREADS), (sp.nc + 1) * 4, stream, matrix, sp, slices);
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(L.H, 4096)), dim3(THREADS), 0, stream, cursor, (uint64_t)L.H);
    hipLaunchKernelGGL(k_seg_hist, dim3(L.max_slices), dim3(THREADS), nbk * 4, stream, ktmp, matrix, slices, sp, cursor);
    hipLaunchKernelGGL(k_seg_scan, dim3(sp.nc), dim3(THREADS), nbk * 4, stream, matrix, sp, cursor);
    const uint32_t lds_seg = seg_scatter_lds(sp.low_bits);
    hipLaunchKernelGGL(k_seg_scatter, dim3(L.max_slices), dim3(THREADS), lds_seg, stream, ktmp, vtmp, matrix, slices,
                       sp, cursor, keys0, vals0);
    ZKP_CHECK(hipGetLastError());
    pt.mark("segsort");
  } else {
  const uint32_t h1 = L.H + 1;
  hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(h1, 4096)), dim3(THREADS), 0, stream, hist, (uint64_t)h1);
  ZKP_CHECK(hipGetLastError());
  pt.mark("start");
  hipLaunchKernelGGL(k_histogram, dim3(grid_for(len, 8192)), dim3(THREADS), 0, stream, d_scalars, len, cfg, hist);
  ZKP_CHECK(hipGetLastError());
  pt.mark("histogram");
  hipLaunchKernelGGL(k_scan_blocks, dim3(L.nb_scan), dim3(THREADS), 0, stream, hist, h1, bsums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(THREADS), 0, stream, bsums, L.nb_scan);
  hipLaunchKernelGGL(k_scan_add, dim3(div_ceil32(h1, THREADS)), dim3(THREADS), 0, stream, hist, h1, bsums, cursor);
  hipLaunchKernelGGL(k_set_count, dim3(1), dim3(64), 0, stream, counts, hist, L.H);
  ZKP_CHECK(hipGetLastError());
  pt.mark("scan");
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(len, 8192)), dim3(THREADS), 0, stream, d_scalars, len, cfg, cursor,
                     keys0, vals0);
  ZKP_CHECK(hipGetLastError());
  pt.mark("scatter");
  }
  return hipSuccess;
}

// Second half: adds the sorted entries' points (n_precompute tables laid out
// table-major at d_points) into bucket set `set`.
static hipError_t accumulate_sorted(const Config& cfg, const Layout& L, char* ws, uint32_t set, const u32* d_points,
                                    bool fresh, hipStream_t stream, PhaseTimer& pt) {
  uint32_t* counts = reinterpret_cast<uint32_t*>(ws + L.counts);
  uint32_t* keys0 = reinterpret_cast<uint32_t*>(ws + L.keys0);
  uint32_t* vals0 = reinterpret_cast<uint32_t*>(ws + L.vals0);
  u32* buckets = bucket_set(L, ws, set);
  const uint64_t nchunks0 = (L.n_entries_max + CHUNK - 1) / CHUNK;
  // with a single chunk every run is closed: no next level
  uint32_t* nk = L.levels > 0 ? reinterpret_cast<uint32_t*>(ws + L.lvl_keys[0]) : nullptr;
  u32* np = L.levels > 0 ? reinterpret_cast<u32*>(ws + L.lvl_pts[0]) : nullptr;
  hipLaunchKernelGGL(k_accumulate0, dim3(grid_for(nchunks0, 1u << 30)), dim3(THREADS), 0, stream, keys0, vals0,
                     d_points, counts, fresh, buckets, nk, np);
  ZKP_CHECK(hipGetLastError());
  pt.mark("accumulate0");
  // level 1 chunked, then the short runs of level 2 closed directly; the later
  // chunked levels only run when a long run remains.  A top window of few
  // bits puts its digits in a handful of buckets (runs far above RUN_MAX), so
  // then every level runs and the run pass is left out.
  uint32_t* run_flag = counts + MAX_LEVELS;
  const uint32_t top_bits = SCALAR_BITS - (cfg.actual_windows - 1) * cfg.s;
  const bool close_runs = top_bits >= 10;
  for (int lv = 0; lv < L.levels; lv++) {
    const int cur = lv % 2, nxt = 1 - cur;
    const bool last = lv == L.levels - 1;
    const uint64_t nch = (L.level_max[lv] + CHUNK_L - 1) / CHUNK_L;
    hipLaunchKernelGGL(k_accumulate, dim3(grid_for(nch, 1u << 30)), dim3(THREADS), 0, stream,
                       reinterpret_cast<const uint32_t*>(ws + L.lvl_keys[cur]),
                       reinterpret_cast<const u32*>(ws + L.lvl_pts[cur]), counts, (uint32_t)(lv + 1), fresh, buckets,
                       last ? nullptr : reinterpret_cast<uint32_t*>(ws + L.lvl_keys[nxt]),
                       last ? nullptr : reinterpret_cast<u32*>(ws + L.lvl_pts[nxt]),
                       lv == 0 || !close_runs ? nullptr : run_flag);
    ZKP_CHECK(hipGetLastError());
    if (lv == 0 && !last && close_runs) {
      hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(THREADS), 0, stream, run_flag, (uint64_t)1);
      hipLaunchKernelGGL(k_close_runs, dim3(grid_for(L.level_max[1], 1u << 30)), dim3(THREADS), 0, stream,
                         reinterpret_cast<const uint32_t*>(ws + L.lvl_keys[nxt]),
                         reinterpret_cast<u32*>(ws + L.lvl_pts[nxt]), counts, 2u, fresh, buckets, run_flag);
      ZKP_CHECK(hipGetLastError());
    }
  }
  pt.mark("levels");
  return hipSuccess;
}

// Adds sum_i k_i P_i for one part (len points, n_precompute tables laid out
// table-major at d_points) into bucket set `set`.
static hipError_t accumulate_part(const Config& cfg, const Layout& L, char* ws, uint32_t set, const u32* d_points,
                                  const u32* d_scalars, uint64_t len, bool fresh, hipStream_t stream, PhaseTimer& pt) {
  ZKP_CHECK(sort_part(cfg, L, ws, d_scalars, len, stream, pt));
  return accumulate_sorted(cfg, L, ws, set, d_points, fresh, stream, pt);
}

// Levels with at most this many point operations run a quad of lanes per
// operation (their time is the operations' latency, not their count).
constexpr uint64_t QUAD_LEVEL_OPS = 65536;

// Reduction of bucket set `set`; the n_windows window sums end at *out_dev (XYZZ).
static hipError_t reduce_buckets(const Config& cfg, const Layout& L, char* ws, uint32_t set, hipStream_t stream,
                                 const u32** out_dev, PhaseTimer& pt) {
  const u32* X = bucket_set(L, ws, set);
  const u32* Y = X;
  uint32_t M = cfg.buckets;
  if (M < 2 || (M & (M - 1))) return hipErrorInvalidValue;  // 2^(s-1) buckets, s >= 2
  int rl0 = 0;
  if (L.red_levels >= 2) {
    // levels 0 and 1 in one pass, into level 1's buffers
    u32* Xo = reinterpret_cast<u32*>(ws + L.red[2]);
    u32* Yo = reinterpret_cast<u32*>(ws + L.red[3]);
    const uint32_t blocks = (M >> 2) * cfg.n_windows;
    hipLaunchKernelGGL(k_reduce_block4, dim3(div_ceil32(blocks, THREADS)), dim3(THREADS), 0, stream, X, M,
                       cfg.n_windows, Xo, Yo);
    ZKP_CHECK(hipGetLastError());
    X = Xo;
    Y = Yo;
    M >>= 2;
    rl0 = 2;
  }
  for (int rl = rl0; rl < L.red_levels; rl++) {
    u32* Xo = reinterpret_cast<u32*>(ws + L.red[(rl % 2) * 2]);
    u32* Yo = reinterpret_cast<u32*>(ws + L.red[(rl % 2) * 2 + 1]);
    const uint32_t Mo = L.red_m[rl];
    const uint64_t ops = 2ull * Mo * cfg.n_windows;
    if (ops <= QUAD_LEVEL_OPS)
      hipLaunchKernelGGL(k_reduce_level_quad, dim3(div_ceil32(4 * ops, THREADS)), dim3(THREADS), 0, stream, X, Y, M,
                         cfg.n_windows, rl == 0, Xo, Yo);
    else
      hipLaunchKernelGGL(k_reduce_level, dim3(div_ceil32(ops, THREADS)), dim3(THREADS), 0, stream, X, Y, M,
                         cfg.n_windows, rl == 0, Xo, Yo);
    ZKP_CHECK(hipGetLastError());
    X = Xo;
    Y = Yo;
    M = Mo;
  }
  u32* sums = reinterpret_cast<u32*>(ws + L.sums);
  hipLaunchKernelGGL(k_reduce_tail_quad, dim3(cfg.n_windows), dim3(QUAD_TAIL_THREADS), 0, stream, X, Y, M,
                     L.red_levels == 0, sums);
  ZKP_CHECK(hipGetLastError());
  *out_dev = sums;
  pt.mark("reduce");
  return hipSuccess;
}

// Host Horner over the window sums: sum_w 2^(s w) S_w.
static G1XYZZ combine_windows(const Config& cfg, const std::vector<G1XYZZ>& sums) {
  G1XYZZ acc = G1XYZZ::identity();
  for (int w = (int)cfg.n_windows - 1; w >= 0; w--) {
    if (!acc.is_identity())
      for (uint32_t k = 0; k < cfg.s; k++) acc = acc.dbl();
    acc = acc.add(sums[w]);
  }
  return acc;
}

static hipError_t fetch_window_sums(const Config& cfg, const u32* d_sums, hipStream_t stream,
                                    std::vector<G1XYZZ>& out) {
  out.resize(cfg.n_windows);
  ZKP_CHECK(hipMemcpyAsync(out.data(), d_sums, sizeof(G1XYZZ) * cfg.n_windows, hipMemcpyDeviceToHost, stream));
  return hipStreamSynchronize(stream);
}

// Full device-resident MSM of one scalar batch: result (XYZZ, host).
static hipError_t msm_device(const Config& cfg, const u32* d_points, const u32* d_scalars, uint64_t len, char* ws,
                             const Layout& L, hipStream_t stream, G1XYZZ* result, bool debug) {
  PhaseTimer pt{debug, stream, {}, {}};
  ZKP_CHECK(reset_buckets(cfg, L, ws, 0, stream));
  if (len > 0) ZKP_CHECK(accumulate_part(cfg, L, ws, 0, d_points, d_scalars, len, true, stream, pt));
  const u32* d_sums = nullptr;
  ZKP_CHECK(reduce_buckets(cfg, L, ws, 0, stream, &d_sums, pt));
  std::vector<G1XYZZ> sums;
  ZKP_CHECK(fetch_window_sums(cfg, d_sums, stream, sums));
  auto t0 = std::chrono::steady_clock::now();
  *result = combine_windows(cfg, sums);
  if (debug) {
    pt.report();
    auto t1 = std::chrono::steady_clock::now();
    fprintf(stderr, "[msm] host combine %.4f ms\n", std::chrono::duration<double, std::milli>(t1 - t0).count());
  }
  return hipSuccess;
}

// ---------------------------------------------------------------- host-streamed MSM (the drop-in msm)
//
// Plan of one call: each card takes a contiguous range of points, split into
// `parts` pieces that are streamed through its buffer; batches run in groups
// of batch_per_run that share one pass over the parts, each with its own
// bucket set.  Copies go on two copy streams (points, scalars) into
// stage_points / stage_scalars rotating device buffers, ordered against the
// compute stream by events, so the next part's points and the next batch's
// scalars cross the link while the current accumulation runs.  Runs
// alternate the direction of their pass over the parts, so the part a run
// ends with is still resident when the next run starts with it (with one part
// the points cross the link once per call).
struct HostPlan {
  uint64_t part_len;      // points per part (the last may be shorter)
  uint32_t scratch_set;   // bucket set for the parts after a run's first (UINT32_MAX: one part)
  uint32_t stage_points;  // rotating device buffers for point parts
  uint32_t stage_scalars; // rotating device buffers for scalar parts
  uint32_t batch_per_run; // bucket sets
  size_t pts_bytes;       // one point stage: part_len * n_precompute affine points
  size_t sc_bytes;        // one scalar stage: part_len Fr elements
  Layout L;

  size_t total() const { return stage_points * pts_bytes + stage_scalars * sc_bytes + L.total; }
};

static bool make_host_plan(const Config& cfg, uint64_t card_len, uint32_t parts, uint32_t batch_per_run,
                           uint32_t stage_scalars, uint32_t stage_points, HostPlan* P) {
  if (parts == 0 || batch_per_run == 0 || stage_scalars == 0 || stage_points == 0) return false;
  if (batch_per_run > 1024 || stage_scalars > 64 || stage_points > 64) return false;
  P->part_len = (card_len + parts - 1) / parts;
  P->stage_points = stage_points;
  P->stage_scalars = stage_scalars;
  P->batch_per_run = batch_per_run;
  const bool several = P->part_len < card_len;
  P->scratch_set = several ? batch_per_run : UINT32_MAX;
  if (!make_layout(cfg, P->part_len ? P->part_len : 1, &P->L, batch_per_run + (several ? 1 : 0))) return false;
  P->pts_bytes = align256((size_t)P->part_len * cfg.n_precompute * 64);
  P->sc_bytes = align256((size_t)P->part_len * 32);
  return true;
}

// One card's streams (compute, point copies, scalar copies), the events that
// order them, and pinned slots for the window sums.  Made once per device and
// kept for the process's lifetime in a free list (stream, event and pinned
// allocations cost more than a small MSM); a call takes one set for itself,
// grows it if its plan needs more stages, and gives it back drained.
struct CardRes {
  hipStream_t s = nullptr, cp = nullptr, cs = nullptr;
  std::vector<hipEvent_t> pt_ready, pt_free, sc_ready, sc_free;
  G1XYZZ* h_sums = nullptr;  // batch_per_run x n_windows window sums (pinned)
  size_t sums_cap = 0;

  hipError_t ensure(const Config& cfg, const HostPlan& P) {
    if (s == nullptr) ZKP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    if (cp == nullptr) ZKP_CHECK(hipStreamCreateWithFlags(&cp, hipStreamNonBlocking));
    if (cs == nullptr) ZKP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    auto grow = [](std::vector<hipEvent_t>& v, uint32_t n) -> hipError_t {
      while (v.size() < n) {
        hipEvent_t e;
        ZKP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        v.push_back(e);
      }
      return hipSuccess;
    };
    ZKP_CHECK(grow(pt_ready, P.stage_points));
    ZKP_CHECK(grow(pt_free, P.stage_points));
    ZKP_CHECK(grow(sc_ready, P.stage_scalars));
    ZKP_CHECK(grow(sc_free, P.stage_scalars));
    const size_t need = sizeof(G1XYZZ) * cfg.n_windows * P.batch_per_run;
    if (sums_cap < need) {
      if (h_sums) ZKP_CHECK(hipHostFree(h_sums));
      h_sums = nullptr;
      sums_cap = 0;
      ZKP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_sums), need, hipHostMallocDefault));
      sums_cap = need;
    }
    // every stage starts free
    for (uint32_t i = 0; i < P.stage_points; i++) ZKP_CHECK(hipEventRecord(pt_free[i], s));
    for (uint32_t i = 0; i < P.stage_scalars; i++) ZKP_CHECK(hipEventRecord(sc_free[i], s));
    return hipSuccess;
  }
  void drain() {
    for (hipStream_t q : {s, cp, cs})
      if (q) (void)hipStreamSynchronize(q);
  }
};

static std::mutex g_res_mu;
static std::vector<CardRes*>* g_res_free = new std::vector<CardRes*>[64];  // per device; never freed

struct ResLease {
  int dev;
  CardRes* r = nullptr;
  explicit ResLease(int d) : dev(d) {
    std::lock_guard<std::mutex> lock(g_res_mu);
    auto& v = g_res_free[dev];
    if (!v.empty()) {
      r = v.back();
      v.pop_back();
    }
    if (r == nullptr) r = new CardRes;
  }
  ~ResLease() {
    r->drain();  // nothing of this call may still run when the set is reused
    std::lock_guard<std::mutex> lock(g_res_mu);
    g_res_free[dev].push_back(r);
  }
};

// Host-to-device copy of `bytes` on stream s.  With `pull`, pinned host memory
// that the device addresses at the host's own virtual address (hipHostMalloc'd
// or registered memory) is pulled by k_pull_host when both ends are 16-byte
// aligned; any other source (pageable memory), and every copy without `pull`,
// goes through hipMemcpyAsync.
static hipError_t copy_in(void* dst, const void* src, size_t bytes, hipStream_t s, bool pull) {
  if (bytes == 0) return hipSuccess;
  if (!pull) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  hipPointerAttribute_t a{};
  const hipError_t q = hipPointerGetAttributes(&a, src);
  if (q != hipSuccess) (void)hipGetLastError();  // an unregistered pointer may be reported as an error
  const bool pinned = q == hipSuccess && a.type == hipMemoryTypeHost && a.devicePointer != nullptr &&
                      a.devicePointer == a.hostPointer;
  if (pinned && bytes % 16 == 0 && (uintptr_t)src % 16 == 0 && (uintptr_t)dst % 16 == 0) {
    hipLaunchKernelGGL(k_pull_host, dim3(PULL_BLOCKS), dim3(THREADS), 0, s, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), (uint64_t)(bytes / 16));
    return hipGetLastError();
  }
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
}

// One card's share: points [c_off, c_off + c_len) of every table, all
// batches; adds the card's partial result of batch b into totals[b].
//
// Per part, each batch's scalars are copied first; the part's sort reads only
// them, so it runs while the part's points are still on the link; the
// accumulation waits for the points.  The next part's points are requested
// as soon as this part's first scalars are on their way, into another point
// stage, so they cross the link during this part's sorts and accumulations.
static hipError_t msm_card(const Config& cfg, const HostPlan& P, unsigned card, char* buf, uint64_t c_off,
                           uint64_t c_len, const u32* const* h_points, unsigned batches, const u32* const* h_scalars,
                           G1XYZZ* totals, bool debug) {
  ZKP_CHECK(hipSetDevice((int)card));
  int dev = 0;
  ZKP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  ResLease lease(dev);
  CardRes& R = *lease.r;
  ZKP_CHECK(R.ensure(cfg, P));
  const Layout& L = P.L;
  char* ws = buf + P.stage_points * P.pts_bytes + P.stage_scalars * P.sc_bytes;
  auto pts = [&](uint32_t st) { return reinterpret_cast<u32*>(buf + st * P.pts_bytes); };
  auto scs = [&](uint32_t st) { return reinterpret_cast<u32*>(buf + P.stage_points * P.pts_bytes + st * P.sc_bytes); };
  const uint32_t nparts = c_len ? (uint32_t)((c_len + P.part_len - 1) / P.part_len) : 0;
  // The copy kernel needs room on the CUs.  With one part the points cross the
  // link beside the sort only; with several, each part's copy runs beside the
  // previous part's accumulation, whose waves hold every CU's registers, and
  // the kernel would wait for them (2^20 points, window 16, four parts: 4.6 ms
  // with the kernel, 3.9 ms with DMA), so those copies go to the DMA engines.
  const bool pull = nparts == 1;
  auto part_off = [&](uint32_t p) { return c_off + (uint64_t)p * P.part_len; };
  auto part_len = [&](uint32_t p) {
    const uint64_t o = part_off(p), end = c_off + c_len;
    return end - o < P.part_len ? end - o : P.part_len;
  };
  // which part each point stage holds (-1: none); the copy has been enqueued
  std::vector<int64_t> held(P.stage_points, -1);
  std::vector<uint64_t> last_use(P.stage_points, 0);
  uint64_t clock = 0;
  uint32_t sc_next = 0;
  // stage holding part p, enqueueing its copy into the least recently used
  // stage other than `keep` when it is not resident; `after` (may be null) is
  // an event the copy waits for first (the scalars that should go first)
  auto points_of = [&](uint32_t p, int64_t keep, hipEvent_t after, uint32_t* out) -> hipError_t {
    for (uint32_t st = 0; st < P.stage_points; st++)
      if (held[st] == (int64_t)p) {
        *out = st;
        last_use[st] = ++clock;
        return hipSuccess;
      }
    uint32_t st = P.stage_points;
    for (uint32_t k = 0; k < P.stage_points; k++)
      if ((int64_t)k != keep && (st == P.stage_points || last_use[k] < last_use[st])) st = k;
    if (st == P.stage_points) st = 0;  // one stage: reuse it after its last reader
    ZKP_CHECK(hipStreamWaitEvent(R.cp, R.pt_free[st], 0));
    if (after != nullptr) ZKP_CHECK(hipStreamWaitEvent(R.cp, after, 0));
    const uint64_t off = part_off(p), n = part_len(p);
    for (uint32_t t = 0; t < cfg.n_precompute; t++)
      ZKP_CHECK(copy_in(pts(st) + (size_t)t * n * 16, h_points[t] + off * 16, n * 64, R.cp, pull));
    ZKP_CHECK(hipEventRecord(R.pt_ready[st], R.cp));
    held[st] = p;
    last_use[st] = ++clock;
    *out = st;
    return hipSuccess;
  };
  for (unsigned r0 = 0, run = 0; r0 < batches; r0 += P.batch_per_run, run++) {
    const uint32_t nb = batches - r0 < P.batch_per_run ? batches - r0 : P.batch_per_run;
    PhaseTimer pt{debug, R.s, {}, {}};
    for (uint32_t j = 0; j < nb; j++) ZKP_CHECK(reset_buckets(cfg, L, ws, j, R.s));
    const bool forward = (run & 1) == 0;
    for (uint32_t idx = 0; idx < nparts; idx++) {
      const uint32_t p = forward ? idx : nparts - 1 - idx;
      const uint64_t off = part_off(p), n = part_len(p);
      int64_t st = -1;
      for (uint32_t j = 0; j < nb; j++) {
        const uint32_t ss = sc_next++ % P.stage_scalars;
        ZKP_CHECK(hipStreamWaitEvent(R.cs, R.sc_free[ss], 0));
        ZKP_CHECK(copy_in(scs(ss), h_scalars[r0 + j] + off * 8, n * 32, R.cs, pull));
        ZKP_CHECK(hipEventRecord(R.sc_ready[ss], R.cs));
        if (j == 0) {
          // the part's points after its first scalars (resident parts are not copied again), then the next
          // part's points into another stage
          uint32_t s0 = 0;
          ZKP_CHECK(points_of(p, -1, R.sc_ready[ss], &s0));
          st = s0;
          if (idx + 1 < nparts && P.stage_points > 1) {
            uint32_t s1 = 0;
            ZKP_CHECK(points_of(forward ? p + 1 : p - 1, st, nullptr, &s1));
          }
        }
        ZKP_CHECK(hipStreamWaitEvent(R.s, R.sc_ready[ss], 0));
        ZKP_CHECK(sort_part(cfg, L, ws, scs(ss), n, R.s, pt));
        ZKP_CHECK(hipEventRecord(R.sc_free[ss], R.s));
        if (j == 0) ZKP_CHECK(hipStreamWaitEvent(R.s, R.pt_ready[st], 0));
        if (idx == 0 || P.scratch_set == UINT32_MAX) {
          ZKP_CHECK(accumulate_sorted(cfg, L, ws, j, pts((uint32_t)st), idx == 0, R.s, pt));
        } else {
          ZKP_CHECK(reset_buckets(cfg, L, ws, P.scratch_set, R.s));
          ZKP_CHECK(accumulate_sorted(cfg, L, ws, P.scratch_set, pts((uint32_t)st), true, R.s, pt));
          hipLaunchKernelGGL(k_merge_buckets, dim3(div_ceil32(L.H, THREADS)), dim3(THREADS), 0, R.s,
                             bucket_set(L, ws, j), bucket_set(L, ws, P.scratch_set), L.H);
          ZKP_CHECK(hipGetLastError());
        }
      }
      ZKP_CHECK(hipEventRecord(R.pt_free[st], R.s));
    }
    for (uint32_t j = 0; j < nb; j++) {
      const u32* d_sums = nullptr;
      ZKP_CHECK(reduce_buckets(cfg, L, ws, j, R.s, &d_sums, pt));
      ZKP_CHECK(hipMemcpyAsync(R.h_sums + (size_t)j * cfg.n_windows, d_sums, sizeof(G1XYZZ) * cfg.n_windows,
                               hipMemcpyDeviceToHost, R.s));
    }
    ZKP_CHECK(hipStreamSynchronize(R.s));
    pt.report();
    for (uint32_t j = 0; j < nb; j++) {
      std::vector<G1XYZZ> sums(R.h_sums + (size_t)j * cfg.n_windows, R.h_sums + (size_t)(j + 1) * cfg.n_windows);
      totals[r0 + j] = totals[r0 + j].add(combine_windows(cfg, sums));
    }
  }
  return hipSuccess;
}

}  // namespace msm
}  // namespace zkp

using namespace zkp;
using namespace zkp::msm;

extern "C" {

// Multi-scalar multiplication of host points / host scalars, results to host
// (affine), for MsmConfig (window, target_windows).  Behind msm of
// core/src/msm/src/msm.h:4-9 / msm.cu:9-41, with the same calling convention:
// with buffers == NULL only buffer_sizes[card] (bytes of device memory per
// card) is filled.  Points are split over the cards in contiguous ranges, one
// host thread per card (MultiGPUMSM, msm_impl.cuh:830-909); each card streams
// its range in `parts` pieces with batch_per_run batches per pass, through
// stage_points / stage_scalers rotating buffers on copy streams (MSM::run,
// msm_impl.cuh:374-603); the cards' results are summed on the host.
// Synchronous.  h_points holds n_precompute tables (msm_precompute),
// h_scaler_batch `batches` scalar arrays (Fr Montgomery, no rotation),
// h_result `batches` affine points.  Host arrays in pinned memory (as the
// runtime's CPU pool gives) let the copies overlap the accumulation.
hipError_t zkp_msm_host(unsigned int window, unsigned int target_windows, int debug_flag, void* const* buffers,
                        unsigned long* buffer_sizes, unsigned long long len, unsigned int batch_per_run,
                        unsigned int parts, unsigned int stage_scalers, unsigned int stage_points,
                        unsigned int num_cards, const unsigned int* cards, unsigned int const* const* h_points,
                        unsigned int batches, unsigned int const* const* h_scaler_batch,
                        unsigned int* const* h_result) {
  Config cfg;
  if (!make_config(window, target_windows, &cfg)) return hipErrorInvalidValue;
  if (num_cards == 0 || cards == nullptr) return hipErrorInvalidValue;
  const uint64_t card_len = (len + num_cards - 1) / num_cards;
  HostPlan P;
  if (!make_host_plan(cfg, card_len, parts, batch_per_run, stage_scalers, stage_points, &P))
    return hipErrorInvalidValue;
  if (buffers == nullptr) {
    if (buffer_sizes == nullptr) return hipErrorInvalidValue;
    for (unsigned c = 0; c < num_cards; c++) buffer_sizes[c] = P.total();
    return hipSuccess;
  }
  if (batches > 0 && (h_points == nullptr || h_scaler_batch == nullptr || h_result == nullptr))
    return hipErrorInvalidValue;
  const bool debug = debug_flag != 0;
  std::vector<std::vector<G1XYZZ>> partial(num_cards, std::vector<G1XYZZ>(batches, G1XYZZ::identity()));
  std::vector<hipError_t> errs(num_cards, hipSuccess);
  std::vector<std::thread> threads;
  for (unsigned c = 0; c < num_cards; c++) {
    const uint64_t c_off = c * card_len;
    if (c_off >= len) break;
    const uint64_t c_len = len - c_off < card_len ? len - c_off : card_len;
    auto job = [&, c, c_off, c_len]() {
      errs[c] = msm_card(cfg, P, cards[c], reinterpret_cast<char*>(buffers[c]), c_off, c_len, h_points, batches,
                         h_scaler_batch, partial[c].data(), debug);
    };
    if (num_cards == 1) job();
    else threads.emplace_back(job);
  }
  for (auto& t : threads) t.join();
  for (unsigned c = 0; c < num_cards; c++)
    if (errs[c] != hipSuccess) return errs[c];
  for (unsigned b = 0; b < batches; b++) {
    G1XYZZ total = G1XYZZ::identity();
    for (unsigned c = 0; c < num_cards; c++) total = total.add(partial[c][b]);
    const G1Affine a = total.to_affine();
    memcpy(h_result[b], a.x.v, 32);
    memcpy(h_result[b] + 8, a.y.v, 32);
  }
  return hipSuccess;
}

// Fills tables 1..n_precompute-1: h_points[j][i] = 2^(j * n_windows * s) * h_points[0][i]
// (affine), for MsmConfig (window, target_windows).  Behind msm_precompute of
// core/src/msm/src/msm.h:11 / msm_impl.cuh:776-828.  Synchronous; uses at most max_cards devices.
hipError_t zkp_msm_precompute_host(unsigned int window, unsigned int target_windows, unsigned long long len,
                                   unsigned int* const* h_points, unsigned int max_cards) {
  Config cfg;
  if (!make_config(window, target_windows, &cfg)) return hipErrorInvalidValue;
  if (cfg.n_precompute == 1 || len == 0) return hipSuccess;
  if (h_points == nullptr || max_cards == 0) return hipErrorInvalidValue;
  int devices = 0;
  ZKP_CHECK(hipGetDeviceCount(&devices));
  if (devices < 1) return hipErrorNoDevice;
  if ((unsigned)devices > max_cards) devices = (int)max_cards;
  const uint64_t per_dev = (len + devices - 1) / devices;
  for (int d = 0; d < devices; d++) {
    const uint64_t off = (uint64_t)d * per_dev;
    if (off >= len) break;
    const uint64_t n = len - off < per_dev ? len - off : per_dev;
    ZKP_CHECK(hipSetDevice(d));
    hipStream_t s;
    ZKP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const uint64_t part = n < (1ull << 20) ? n : (1ull << 20);
    u32* dp = nullptr;
    hipError_t err = hipMalloc(&dp, (size_t)part * 64 * cfg.n_precompute);
    for (uint64_t p = 0; p < n && err == hipSuccess; p += part) {
      const uint64_t cur = n - p < part ? n - p : part;
      err = hipMemcpyAsync(dp, h_points[0] + (off + p) * 16, cur * 64, hipMemcpyHostToDevice, s);
      if (err != hipSuccess) break;
      hipLaunchKernelGGL(k_precompute, dim3(div_ceil32(cur, THREADS)), dim3(THREADS), 0, s, dp, cur,
                         cfg.n_precompute, cfg.n_windows * cfg.s);
      err = hipGetLastError();
      for (uint32_t j = 1; j < cfg.n_precompute && err == hipSuccess; j++)
        err = hipMemcpyAsync(h_points[j] + (off + p) * 16, dp + (size_t)j * cur * 16, cur * 64,
                             hipMemcpyDeviceToHost, s);
      if (err == hipSuccess) err = hipStreamSynchronize(s);
    }
    if (dp) (void)hipFree(dp);
    (void)hipStreamDestroy(s);
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

// ---- extension entry points (no counterpart in the reference ABI) ----

// Device workspace bytes for zkp_msm_device with `len` points per table.
unsigned long long zkp_msm_workspace_size(unsigned long long len, unsigned int window, unsigned int target_windows) {
  Config cfg;
  if (!make_config(window, target_windows, &cfg)) return 0;
  Layout L;
  if (!make_layout(cfg, len ? len : 1, &L)) return 0;
  return L.total;
}

// MSM with points (n_precompute tables of len affine points, table-major) and
// scalars already in device memory.  Runs on `stream`, then synchronizes it and
// writes the affine result to host memory h_result[16].
hipError_t zkp_msm_device(const unsigned int* d_points, const unsigned int* d_scalars, unsigned long long len,
                          unsigned int window, unsigned int target_windows, void* d_workspace,
                          unsigned long long workspace_bytes, hipStream_t stream, unsigned int* h_result, int debug) {
  Config cfg;
  if (!make_config(window, target_windows, &cfg)) return hipErrorInvalidValue;
  Layout L;
  if (!make_layout(cfg, len ? len : 1, &L)) return hipErrorInvalidValue;
  if (d_workspace == nullptr || workspace_bytes < L.total || h_result == nullptr) return hipErrorInvalidValue;
  if (len > 0 && (d_points == nullptr || d_scalars == nullptr)) return hipErrorInvalidValue;
  G1XYZZ r;
  ZKP_CHECK(msm_device(cfg, d_points, d_scalars, len, reinterpret_cast<char*>(d_workspace), L, stream, &r,
                       debug != 0));
  const G1Affine a = r.to_affine();
  memcpy(h_result, a.x.v, 32);
  memcpy(h_result + 8, a.y.v, 32);
  return hipSuccess;
}

// zkp_msm_device split at the host boundary, for callers that batch MSMs into
// one stream or hipGraph: enqueues the whole device part (digits, sort,
// accumulation, bucket reduction) on `stream` and, when h_sums is not NULL,
// an asynchronous copy of the n_windows window sums (XYZZ, 128 bytes each)
// into h_sums (pinned host memory, n_windows * 128 bytes).  Nothing is
// allocated and nothing waits; zkp_msm_combine turns the sums into the
// result once the stream has passed the copy.  Launch permissions are set on
// the device's first MSM: run one MSM (or this call) outside a capture first.
hipError_t zkp_msm_device_enqueue(const unsigned int* d_points, const unsigned int* d_scalars, unsigned long long len,
                                  unsigned int window, unsigned int target_windows, void* d_workspace,
                                  unsigned long long workspace_bytes, hipStream_t stream, void* h_sums) {
  Config cfg;
  if (!make_config(window, target_windows, &cfg)) return hipErrorInvalidValue;
  Layout L;
  if (!make_layout(cfg, len ? len : 1, &L)) return hipErrorInvalidValue;
  if (d_workspace == nullptr || workspace_bytes < L.total) return hipErrorInvalidValue;
  if (len > 0 && (d_points == nullptr || d_scalars == nullptr)) return hipErrorInvalidValue;
  char* ws = reinterpret_cast<char*>(d_workspace);
  PhaseTimer pt{false, stream, {}, {}};
  ZKP_CHECK(reset_buckets(cfg, L, ws, 0, stream));
  if (len > 0) ZKP_CHECK(accumulate_part(cfg, L, ws, 0, d_points, d_scalars, len, true, stream, pt));
  const u32* d_sums = nullptr;
  ZKP_CHECK(reduce_buckets(cfg, L, ws, 0, stream, &d_sums, pt));
  if (h_sums != nullptr)
    ZKP_CHECK(hipMemcpyAsync(h_sums, d_sums, sizeof(G1XYZZ) * cfg.n_windows, hipMemcpyDeviceToHost, stream));
  return hipSuccess;
}

// Result of an MSM from its window sums (as zkp_msm_device_enqueue copied
// them): Horner over the windows on the host, affine into h_result[16].
hipError_t zkp_msm_combine(unsigned int window, unsigned int target_windows, const void* h_sums,
                           unsigned int* h_result) {
  Config cfg;
  if (!make_config(window, target_windows, &cfg)) return hipErrorInvalidValue;
  if (h_sums == nullptr || h_result == nullptr) return hipErrorInvalidValue;
  const G1XYZZ* p = reinterpret_cast<const G1XYZZ*>(h_sums);
  std::vector<G1XYZZ> sums(p, p + cfg.n_windows);
  const G1Affine a = combine_windows(cfg, sums).to_affine();
  memcpy(h_result, a.x.v, 32);
  memcpy(h_result + 8, a.y.v, 32);
  return hipSuccess;
}

// out[i] = k_i * G (affine, Fq Montgomery) for Fr Montgomery scalars k_i on the
// device; asynchronous on stream.
hipError_t zkp_g1_mul_generator(const unsigned int* d_scalars, unsigned long long n, unsigned int* d_out,
                                hipStream_t stream) {
  if (n == 0) return hipSuccess;
  if (d_scalars == nullptr || d_out == nullptr) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mul_generator, dim3(div_ceil32(n, THREADS)), dim3(THREADS), 0, stream, d_scalars, (uint64_t)n,
                     d_out);
  return hipGetLastError();
}

// Derived configuration for (window, target_windows): {window,
// actual_windows, n_windows, n_precompute, buckets}; returns 0 if invalid.
int zkp_msm_config(unsigned int window, unsigned int target_windows, unsigned int* out5) {
  Config c;
  if (!make_config(window, target_windows, &c)) return 0;
  out5[0] = c.s;
  out5[1] = c.actual_windows;
  out5[2] = c.n_windows;
  out5[3] = c.n_precompute;
  out5[4] = c.buckets;
  return 1;
}

}  // extern "C"
