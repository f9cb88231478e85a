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
// Kernels (default plan):
//  * bitrev_head_stream: the bit-reversal permutation fused with DIT stages
//    0..4; a workgroup swaps a pair of 32 x 32 tiles through LDS and runs the
//    five stages on them before writing back, walking its share of the tile
//    pairs with the next pair's loads in flight during the current stages.
//  * dit_pass_r4: a workgroup loads a tile of 2^TILE_LOG elements -- 2^d
//    butterfly-connected rows times 2^c adjacent columns, so every global
//    access is a run of >= 2^c consecutive 32-byte elements -- into LDS (two
//    16-byte planes), runs d consecutive stages there two at a time (radix-4
//    register blocks), and writes it back.
//    Between stages and passes values stay in [0, 4p) (lazy reduction); the
//    last pass stores canonical values.
//    A transform of 2^24 is the head + 3 passes (7 + 6 + 6 stages), or
//    + 2 passes (10 + 9) with min_col_log = 1.
// The first version (bitrev_kernel + radix-2 dit_pass, 11 + 9 + 4 stages) is
// kept selectable through zkp_ntt_set_variant for comparison.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <cstring>
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
  uint32_t head_t;  // stages fused into the bit reversal (0: none)
  uint32_t head_g;  // the head's 2^head_g workgroups walk the tile pairs
};

// Pass parameters of one instance in one round (d == 0: no pass this round);
// pad != 0 marks the instance's last pass (canonical output).
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

// ---------------------------------------------------------------- fused bit reversal + first stages
//
// Index i = (a, M, b): a = top t bits, M = middle L - 2t bits, b = low t bits.
// The bit-reversed array is y[(a, M, b)] = x[(rev b, rev M, rev a)], so the
// 2^t x 2^t tile of x at middle M' = rev M, read transposed, is the tile of y
// at middle M.  A block owns the pair {M, rev M}: it loads both x tiles
// (rows of 2^t contiguous elements), runs DIT stages 0..t-1 (which only mix the
// low t bits, i.e. stay inside a y row) through the transposed LDS index map,
// and writes both y tiles back in place.  Rows are padded by one element so
// the transposed accesses spread over the LDS banks.
constexpr uint32_t HEAD_T_MAX = 5;
constexpr uint32_t HEAD_W = 1u << HEAD_T_MAX;
constexpr uint32_t HEAD_PITCH = HEAD_W + 1;

__device__ __forceinline__ uint32_t rev_bits(uint32_t x, uint32_t bits) { return bits ? __brev(x) >> (32 - bits) : 0; }

__device__ __forceinline__ Fr lds_get(const uint4* l0, const uint4* l1, uint32_t p) {
  const uint4 a = l0[p], b = l1[p];
  Fr r;
  r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
  r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
  return r;
}
__device__ __forceinline__ void lds_put(uint4* l0, uint4* l1, uint32_t p, const Fr& x) {
  l0[p] = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
  l1[p] = make_uint4(x.v[4], x.v[5], x.v[6], x.v[7]);
}

constexpr uint32_t HEAD_BUF = HEAD_W * HEAD_PITCH;  // one 32 x 33 tile plane

// Two tile buffers per plane (x at M at offset 0, x at rev M at HEAD_BUF) and
// the head's twiddles r[i] = w^(i * 2^(log_n - t)), i < 2^(t - 1).  The second
// plane starts 128 bytes into the LDS bank cycle relative to the first, so the
// two lanes that move the two halves of one element hit different banks.
struct HeadLds {
  uint4 p0[2 * HEAD_BUF];
  uint4 pad[8];
  uint4 p1[2 * HEAD_BUF];
  uint4 r0[HEAD_W / 2], r1[HEAD_W / 2];
};
static_assert((2 * HEAD_BUF * 16 + 128) % 256 == 128, "planes half a bank cycle apart");

// The head moves tiles between HBM and LDS as 16-byte halves of elements:
// half h of a tile is plane (h & 1) of element h >> 1, and consecutive lanes
// take consecutive halves, so one wave instruction reads or writes 1 KiB of a
// tile row (32 elements, contiguous in memory) and every 128-byte line it
// touches is filled by that one instruction.
__device__ __forceinline__ uint4* head_plane(HeadLds& L, uint32_t plane) { return plane ? L.p1 : L.p0; }
__device__ __forceinline__ const uint4* head_plane(const HeadLds& L, uint32_t plane) { return plane ? L.p1 : L.p0; }

// Per-thread staging of the head's global loads: half h = j * NT +
// threadIdx.x of the tile at M (a) and of the tile at rev M (b), j < 2048 / NT.
struct HeadStaged {
  uint4 a, b;
};

// Load j of the thread for a workgroup of NT threads.  The loads are
// unconditional: the half index is clamped for t < 5, and x at M is read twice
// when M == rev M.  Callers keep the results in named values so the staging
// stays in registers.
template <int NT>
__device__ __forceinline__ HeadStaged head_fetch(const PolyView& v, uint32_t t, uint32_t shift_a, uint64_t M,
                                                 uint64_t Mr, uint32_t j) {
  const uint32_t W = 1u << t;
  const uint32_t h = (threadIdx.x + j * NT) & (2 * W * W - 1);
  const uint32_t e = h >> 1, plane = h & 1;
  const uint32_t a = e >> t, b = e & (W - 1);
  const uint4* s = reinterpret_cast<const uint4*>(v.elem(((uint64_t)a << shift_a) | (M << t) | b));
  const uint4* q = reinterpret_cast<const uint4*>(v.elem(((uint64_t)a << shift_a) | (Mr << t) | b));
  return HeadStaged{s[plane], q[plane]};
}

// Staged load j into the two LDS tile buffers (x at M at offset 0, x at rev M
// at HEAD_BUF); the caller puts a barrier after the last.
template <int NT>
__device__ __forceinline__ void head_place(HeadLds& L, uint32_t t, bool pair, uint32_t j, const HeadStaged& g) {
  const uint32_t W = 1u << t;
  const uint32_t h = threadIdx.x + j * NT;
  if (h < 2 * W * W) {
    const uint32_t e = h >> 1;
    const uint32_t p = (e >> t) * HEAD_PITCH + (e & (W - 1));
    uint4* dst = head_plane(L, h & 1);
    dst[p] = g.a;
    if (pair) dst[HEAD_BUF + p] = g.b;
  }
}

// DIT stages 0..t-1 on the placed tile(s), through the transposed index map;
// ends with a barrier.  LAZY keeps values in [0, 4p) (Harvey butterflies, as in
// the passes), for callers whose later passes accept such inputs; otherwise
// every stage output is canonical.
template <bool LAZY, int NT>
__device__ __forceinline__ void head_stages(HeadLds& L, uint32_t t, bool pair) {
  constexpr int U = 512 / NT;   // quads per thread per round (radix-4 stages)
  constexpr int R2 = 1024 / NT;  // butterflies per thread per round (radix-2 stage)
  const uint32_t W = 1u << t;
  // tile 0 = y at M (in the buffer of x at Mr), tile 1 = y at Mr (buffer of x at M)
  const uint32_t y0 = pair ? HEAD_BUF : 0;
  const uint32_t per_tile = W * (W >> 1);
  const uint32_t total = (pair ? 2 : 1) * per_tile;
  // w^0 is 1 in every table built as powers; the product by it is skipped
  // only when the table holds exactly 1, so any other table is still honoured.
  // The skip happens at stage 0 only, whose inputs are the caller's canonical
  // values, so the lazy bounds below hold for the unmultiplied operand too.
  const bool unit0 = lds_get(L.r0, L.r1, 0) == Fr::one();
  // stages two at a time: quads of rows {k0, k0+h, k0+2h, k0+3h} (bits m, m+1
  // of k0 clear) of one column, two quads per thread
  uint32_t m = 0;
  for (; m + 1 < t; m += 2) {
    const uint32_t h = 1u << m;
    const uint32_t per_tile_q = W * (W >> 2);
    const uint32_t totq = (pair ? 2 : 1) * per_tile_q;
    const bool skip1 = m == 0 && unit0;
    for (uint32_t u0 = threadIdx.x; u0 < totq; u0 += U * NT) {
      Fr x[U][4];
      uint32_t pos[U][4], jw[U];
      bool live[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const uint32_t uq = u0 + u * NT;
        live[u] = uq < totq;
        const uint32_t uu = live[u] ? uq : u0;
        const uint32_t tile = uu >= per_tile_q ? 1 : 0;
        const uint32_t r = uu - tile * per_tile_q;
        // consecutive lanes take consecutive LDS columns ra (the W independent
        // transforms of a tile are interchangeable), so a row's accesses hit
        // distinct banks
        const uint32_t ra = r & (W - 1);
        const uint32_t q = r >> t;
        const uint32_t k0 = ((q >> m) << (m + 2)) | (q & (h - 1));
        const uint32_t off = tile ? 0 : y0;
        jw[u] = k0 & (h - 1);
#pragma unroll
        for (int r4 = 0; r4 < 4; r4++) {
          pos[u][r4] = off + rev_bits(k0 + r4 * h, t) * HEAD_PITCH + ra;
          x[u][r4] = lds_get(L.p0, L.p1, pos[u][r4]);
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        Fr t1 = x[u][1], t3 = x[u][3];
        const Fr w2 = lds_get(L.r0, L.r1, jw[u] << (t - 2 - m));
        const Fr w3 = lds_get(L.r0, L.r1, (jw[u] + h) << (t - 2 - m));
        if (LAZY) {
          if (!skip1) {
            const Fr w1 = lds_get(L.r0, L.r1, jw[u] << (t - 1 - m));
            t1 = Fr::mul_lazy(x[u][1], w1);
            t3 = Fr::mul_lazy(x[u][3], w1);
          }
          const Fr x0 = Fr::reduce_2p(x[u][0]), x2 = Fr::reduce_2p(x[u][2]);
          const Fr a0 = Fr::add_raw(x0, t1), a1 = Fr::sub_2p(x0, t1);
          const Fr a2 = Fr::add_raw(x2, t3), a3 = Fr::sub_2p(x2, t3);
          // at m = 0 the second stage's first twiddle is w^0 too: a reduction instead of a product
          Fr t2;
          if (skip1)
            t2 = Fr::reduce_2p(a2);
          else
            t2 = Fr::mul_lazy(a2, w2);
          const Fr t4 = Fr::mul_lazy(a3, w3);
          const Fr a0n = Fr::reduce_2p(a0), a1n = Fr::reduce_2p(a1);
          x[u][0] = Fr::add_raw(a0n, t2);
          x[u][2] = Fr::sub_2p(a0n, t2);
          x[u][1] = Fr::add_raw(a1n, t4);
          x[u][3] = Fr::sub_2p(a1n, t4);
        } else {
          if (!skip1) {
            const Fr w1 = lds_get(L.r0, L.r1, jw[u] << (t - 1 - m));
            t1 = t1 * w1;
            t3 = t3 * w1;
          }
          const Fr a0 = x[u][0] + t1, a1 = x[u][0] - t1;
          const Fr a2 = x[u][2] + t3, a3 = x[u][2] - t3;
          Fr t2 = a2;
          if (!skip1) t2 = a2 * w2;
          const Fr t4 = a3 * w3;
          x[u][0] = a0 + t2;
          x[u][2] = a0 - t2;
          x[u][1] = a1 + t4;
          x[u][3] = a1 - t4;
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (!live[u]) continue;
#pragma unroll
        for (int r4 = 0; r4 < 4; r4++) lds_put(L.p0, L.p1, pos[u][r4], x[u][r4]);
      }
    }
    __syncthreads();
  }
  if (m < t) {  // odd last stage, radix-2
    const uint32_t half = 1u << m;
    const bool skip = m == 0 && unit0;
    for (uint32_t u0 = threadIdx.x; u0 < total; u0 += R2 * NT) {
      Fr x[R2], y[R2];
      uint32_t pa[R2], pb[R2], wi[R2];
      bool live[R2];
#pragma unroll
      for (int j = 0; j < R2; j++) {
        const uint32_t u = u0 + j * NT;
        live[j] = u < total;
        const uint32_t uu = live[j] ? u : u0;
        const uint32_t tile = uu >= per_tile ? 1 : 0;
        const uint32_t r = uu - tile * per_tile;
        const uint32_t ra = r & (W - 1);  // consecutive lanes, consecutive LDS columns
        const uint32_t q = r >> t;
        const uint32_t blo = ((q >> m) << (m + 1)) | (q & (half - 1));
        const uint32_t off = tile ? 0 : y0;
        pa[j] = off + rev_bits(blo, t) * HEAD_PITCH + ra;
        pb[j] = off + rev_bits(blo + half, t) * HEAD_PITCH + ra;
        wi[j] = (blo & (half - 1)) << (t - 1 - m);
        x[j] = lds_get(L.p0, L.p1, pa[j]);
        y[j] = lds_get(L.p0, L.p1, pb[j]);
      }
      if (!skip) {
#pragma unroll
        for (int j = 0; j < R2; j++) {
          const Fr w = lds_get(L.r0, L.r1, wi[j]);
          y[j] = LAZY ? Fr::mul_lazy(y[j], w) : y[j] * w;
        }
      }
#pragma unroll
      for (int j = 0; j < R2; j++) {
        const Fr tt = y[j];
        if (LAZY) {
          const Fr xr = Fr::reduce_2p(x[j]);
          y[j] = Fr::sub_2p(xr, tt);
          x[j] = Fr::add_raw(xr, tt);
        } else {
          y[j] = x[j] - tt;
          x[j] = x[j] + tt;
        }
      }
#pragma unroll
      for (int j = 0; j < R2; j++) {
        if (!live[j]) continue;
        lds_put(L.p0, L.p1, pa[j], x[j]);
        lds_put(L.p0, L.p1, pb[j], y[j]);
      }
    }
    __syncthreads();
  }
}

// Both y tiles back in place (y at M from the buffer at y0, y at Mr from 0),
// as 16-byte halves on consecutive lanes.
template <int NT>
__device__ __forceinline__ void head_store(const HeadLds& L, const PolyView& v, uint32_t t, uint32_t shift_a,
                                           uint64_t M, uint64_t Mr, bool pair) {
  const uint32_t W = 1u << t;
  const uint32_t y0 = pair ? HEAD_BUF : 0;
  for (uint32_t h = threadIdx.x; h < 2 * W * W; h += NT) {
    const uint32_t e = h >> 1, plane = h & 1;
    const uint32_t a = e >> t, b = e & (W - 1);
    const uint32_t p = rev_bits(b, t) * HEAD_PITCH + rev_bits(a, t);
    const uint4* src = head_plane(L, plane);
    reinterpret_cast<uint4*>(v.elem(((uint64_t)a << shift_a) | (M << t) | b))[plane] = src[y0 + p];
    if (pair) reinterpret_cast<uint4*>(v.elem(((uint64_t)a << shift_a) | (Mr << t) | b))[plane] = src[p];
  }
}

// One tile pair {M, rev M} per workgroup.  The caller has written L.r
// (visible after the first barrier below).
// (256 threads.)
template <bool LAZY>
__device__ __forceinline__ void head_tile(HeadLds& L, const PolyView& v, uint32_t log_n, uint32_t t, uint64_t M) {
  const uint32_t mbits = log_n - 2 * t;
  const uint64_t Mr = mbits ? (__brevll(M) >> (64 - mbits)) : 0;
  if (M > Mr) return;  // the pair is handled by block Mr (uniform per block)
  const bool pair = M != Mr;
  const uint32_t shift_a = log_n - t;
  static_assert(2 * HEAD_W * HEAD_W == 8 * THREADS, "eight staged loads per thread");
  HeadStaged g[8];
#pragma unroll
  for (uint32_t j = 0; j < 8; j++) g[j] = head_fetch<THREADS>(v, t, shift_a, M, Mr, j);
#pragma unroll
  for (uint32_t j = 0; j < 8; j++) head_place<THREADS>(L, t, pair, j, g[j]);
  __syncthreads();
  head_stages<LAZY, THREADS>(L, t, pair);
  head_store<THREADS>(L, v, t, shift_a, M, Mr, pair);
}

// L.r[i] = tw[i << (log_n - t)] for i < 2^(t - 1)
__device__ __forceinline__ void head_twiddles_from_table(HeadLds& L, const uint32_t* __restrict__ tw, uint32_t log_n,
                                                         uint32_t t) {
  if (threadIdx.x < ((1u << t) >> 1))
    lds_put(L.r0, L.r1, threadIdx.x, Fr::load(tw + (((uint64_t)threadIdx.x << (log_n - t)) << 3)));
}

// Persistent head: 2^g workgroups (g <= mbits = log_n - 2t, mbits - g even)
// walk the tile pairs.  Walk index k = H : L (h = (mbits - g) / 2 bits each) of
// workgroup b is the tile row
//   M = H << (h + g) | b << h | L,   rev M = rev(L) << (h + g) | rev(b) << h | rev(H),
// so the workgroup's own bits sit in the middle.  The pair {M, rev M} is taken
// by this workgroup when H < rev(L) (M is the smaller of the two); when H ==
// rev(L), rev M is the same walk index in workgroup rev(b), and the pair goes to
// the smaller of b, rev(b) for even H and to the larger for odd H.  Every pair
// is thus taken exactly once, and each workgroup gets the same number of pairs
// up to 2^(h-1).  The loads of the next pair are issued before the stages of
// the current one and land in registers while they run (barriers wait only for
// LDS), so the tile traffic overlaps the butterflies.  The stages are lazy: a
// radix-4 pass always follows the head and takes values in [0, 4p).
// Workgroup b's walk (b < 2^g), HEAD_THREADS threads; the caller has written L.r.
// (512 threads measured no faster: the staged loads of the next pair do not fit
// the 128 VGPRs of four waves per SIMD and spill.)
constexpr int HEAD_THREADS = 256;
__device__ __forceinline__ void head_walk(HeadLds& L, const PolyView& v, uint32_t log_n, uint32_t t, uint32_t g,
                                          uint64_t b) {
  const uint32_t mbits = log_n - 2 * t;
  const uint32_t h = (mbits - g) >> 1;
  const uint64_t K = 1ull << (2 * h);
  const uint64_t lmask = (1ull << h) - 1;
  const uint64_t rb = g ? (__brevll(b) >> (64 - g)) : 0;
  const uint32_t shift_a = log_n - t;
  auto rev_h = [&](uint64_t x) -> uint64_t { return h ? (__brevll(x) >> (64 - h)) : 0; };
  auto item = [&](uint64_t k) -> uint64_t { return ((k >> h) << (h + g)) | (b << h) | (k & lmask); };
  auto rev = [&](uint64_t M) -> uint64_t { return mbits ? (__brevll(M) >> (64 - mbits)) : 0; };
  auto owned = [&](uint64_t k) -> bool {
    const uint64_t H = k >> h, rl = rev_h(k & lmask);
    if (H != rl) return H < rl;
    return b == rb || ((H & 1) ? b > rb : b < rb);
  };
  auto next = [&](uint64_t k) -> uint64_t {  // first owned k' >= k, or K
    while (k < K && !owned(k)) k++;
    return k;
  };
  uint64_t k = next(0);
  if (k >= K) return;
  static_assert(2 * HEAD_W * HEAD_W == 8 * HEAD_THREADS, "eight staged loads per thread");
  HeadStaged stg[8];
  auto fetch = [&](uint64_t M) {
    const uint64_t Mr = rev(M);
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) stg[j] = head_fetch<HEAD_THREADS>(v, t, shift_a, M, Mr, j);
  };
  fetch(item(k));
  for (;;) {
    const uint64_t M = item(k), Mr = rev(M);
    const bool pair = M != Mr;
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) head_place<HEAD_THREADS>(L, t, pair, j, stg[j]);
    __syncthreads();
    const uint64_t kn = next(k + 1);
    if (kn < K) fetch(item(kn));
    head_stages<true, HEAD_THREADS>(L, t, pair);
    head_store<HEAD_THREADS>(L, v, t, shift_a, M, Mr, pair);
    if (kn >= K) break;
    __syncthreads();  // the store phase has read the tiles before they are replaced
    k = kn;
  }
}

__global__ void __launch_bounds__(HEAD_THREADS)
    bitrev_head_stream(PolyView v, const uint32_t* __restrict__ tw, uint32_t log_n, uint32_t t, uint32_t g) {
  __shared__ HeadLds L;
  head_twiddles_from_table(L, tw, log_n, t);
  head_walk(L, v, log_n, t, g, blockIdx.x);
}

// Workgroup bits of bitrev_head_stream for mbits = log_n - 2t: at most
// HEAD_STREAM_G_MAX (1024 workgroups, two rounds of the two per CU that the
// 68 KB of LDS allows on 256 CUs), with mbits - g even.
constexpr uint32_t HEAD_STREAM_G_MAX = 10;
static inline uint32_t head_stream_g(uint32_t mbits) {
  uint32_t g = mbits < HEAD_STREAM_G_MAX ? mbits : HEAD_STREAM_G_MAX;
  if ((mbits - g) & 1) g--;
  return g;
}

__global__ void __launch_bounds__(THREADS)
    bitrev_head(PolyView v, const uint32_t* __restrict__ tw, uint32_t log_n, uint32_t t) {
  __shared__ HeadLds L;
  head_twiddles_from_table(L, tw, log_n, t);
  head_tile<true>(L, v, log_n, t, blockIdx.x);
}

// Fused bit reversal + first head_t stages of every instance; instance k owns
// blocks [head_prefix[k], head_prefix[k + 1]) (2^head_g of them walking its tile
// pairs as in bitrev_head_stream, or none when head_t == 0).
__global__ void __launch_bounds__(HEAD_THREADS)
    bitrev_head_batched(const BatchInstance* __restrict__ inst, const uint64_t* __restrict__ head_prefix,
                        uint32_t count) {
  __shared__ HeadLds L;
  const uint64_t b = blockIdx.x;
  if (b >= head_prefix[count]) return;
  const uint32_t k = find_instance(head_prefix, count, b);
  const BatchInstance I = inst[k];
  head_twiddles_from_table(L, I.tw, I.log_n, I.head_t);
  head_walk(L, I.v, I.log_n, I.head_t, I.head_g, b - head_prefix[k]);
}

// ---------------------------------------------------------------- radix-4 pass
//
// Same tile geometry as dit_tile (2^d rows x 2^c_log columns); the stages run
// two at a time (pass_stages).
__device__ __forceinline__ uint64_t tile_index(uint64_t base, uint64_t cstride, uint64_t kstride, uint32_t c,
                                               uint32_t k) {
  return base + c * cstride + (uint64_t)k * kstride;
}

// Global geometry of one pass tile (tile_index(base, cstride, kstride, c, k)).
struct TileGeom {
  uint64_t base, cstride, kstride;
};

__device__ __forceinline__ TileGeom tile_geom(uint32_t s0, uint32_t d, uint32_t c_log, uint64_t tile) {
  if (s0 == 0) return TileGeom{(tile << c_log) << d, 1ull << d, 1};
  // 2^(s0 - c_log) column blocks per row block (c_log <= s0)
  const uint32_t lo_bits = s0 - c_log;
  const uint64_t hi = tile >> lo_bits;
  const uint64_t lo_block = tile & ((1ull << lo_bits) - 1);
  return TileGeom{lo_block << c_log | (hi << (s0 + d)), 1, 1ull << s0};
}

// Element e of a tile: its LDS position (row k, column c -> k * C + c) and its
// global index.  Consecutive e walk the contiguous direction in memory.
__device__ __forceinline__ uint64_t tile_elem(const TileGeom& g, uint32_t d, uint32_t c_log, uint32_t e,
                                              uint32_t& pos) {
  uint32_t c, k;
  if (g.cstride == 1) {
    c = e & ((1u << c_log) - 1);
    k = e >> c_log;
  } else {
    k = e & ((1u << d) - 1);
    c = e >> d;
  }
  pos = (k << c_log) | c;
  return tile_index(g.base, g.cstride, g.kstride, c, k);
}

// Stages [s0, s0 + d) on the tile in LDS; ends with a barrier.  Butterflies are
// lazy (Harvey): values stay in [0, 4p).  Stages go two at a time: a thread
// owns quads of rows {k0, k0+h, k0+2h, k0+3h} (h = 2^m, bits m and m+1 of k0
// clear), which stages m and m+1 connect only among themselves; stage m uses
// one twiddle for both of its butterflies (k0 and k0+2h agree modulo 2^(s0+m)),
// stage m+1 two.  Per LDS round trip a thread does 2 quads = 8 butterflies with
// 4 independent products per stage; an odd last stage runs radix-2.
template <int NT, int U>
__device__ __forceinline__ void pass_stages(uint4* __restrict__ lds0, uint4* __restrict__ lds1,
                                            const uint32_t* __restrict__ tw, uint32_t log_n, uint32_t s0, uint32_t d,
                                            uint32_t c_log, const TileGeom& geo) {
  const uint32_t C = 1u << c_log;
  const uint32_t T = (1u << d) << c_log;
  uint32_t m = 0;
  for (; m + 1 < d; m += 2) {
    const uint32_t s = s0 + m;
    const uint32_t h = 1u << m;
    const uint64_t mask1 = (1ull << s) - 1, mask2 = (1ull << (s + 1)) - 1;
    const uint32_t sh1 = log_n - 1 - s, sh2 = log_n - 2 - s;
    const uint32_t quads = T >> 2;
    for (uint32_t g0 = threadIdx.x; g0 < quads; g0 += U * NT) {
      Fr x[U][4], w1[U], w2[U], w3[U];
      uint32_t pos[U][4];
      bool live[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const uint32_t g = g0 + u * NT;
        live[u] = g < quads;
        const uint32_t gg = live[u] ? g : g0;
        const uint32_t c = gg & (C - 1);
        const uint32_t qr = gg >> c_log;
        const uint32_t k0 = ((qr >> m) << (m + 2)) | (qr & (h - 1));
        const uint64_t i0 = tile_index(geo.base, geo.cstride, geo.kstride, c, k0);
        const uint64_t i1 = tile_index(geo.base, geo.cstride, geo.kstride, c, k0 + h);
        w1[u] = Fr::load(tw + (((i0 & mask1) << sh1) << 3));
        w2[u] = Fr::load(tw + (((i0 & mask2) << sh2) << 3));
        w3[u] = Fr::load(tw + (((i1 & mask2) << sh2) << 3));
#pragma unroll
        for (int r = 0; r < 4; r++) {
          pos[u][r] = ((k0 + r * h) << c_log) | c;
          x[u][r] = lds_get(lds0, lds1, pos[u][r]);
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        Fr t1, t3, t2, t4;
        Fr::mul_lazy2(x[u][1], w1[u], x[u][3], w1[u], t1, t3);
        const Fr x0 = Fr::reduce_2p(x[u][0]), x2 = Fr::reduce_2p(x[u][2]);
        const Fr a0 = Fr::add_raw(x0, t1), a1 = Fr::sub_2p(x0, t1);
        const Fr a2 = Fr::add_raw(x2, t3), a3 = Fr::sub_2p(x2, t3);
        Fr::mul_lazy2(a2, w2[u], a3, w3[u], t2, t4);
        const Fr a0n = Fr::reduce_2p(a0), a1n = Fr::reduce_2p(a1);
        x[u][0] = Fr::add_raw(a0n, t2);
        x[u][2] = Fr::sub_2p(a0n, t2);
        x[u][1] = Fr::add_raw(a1n, t4);
        x[u][3] = Fr::sub_2p(a1n, t4);
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (!live[u]) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) lds_put(lds0, lds1, pos[u][r], x[u][r]);
      }
    }
    __syncthreads();
  }
  if (m < d) {  // odd last stage, radix-2
    const uint32_t s = s0 + m;
    const uint32_t half = 1u << m;
    const uint64_t jmask = (1ull << s) - 1;
    const uint32_t tw_shift = log_n - 1 - s;
    for (uint32_t b = threadIdx.x; b < (T >> 1); b += NT) {
      const uint32_t c = b & (C - 1);
      const uint32_t q = b >> c_log;
      const uint32_t ka = ((q >> m) << (m + 1)) | (q & (half - 1));
      const uint32_t pa = (ka << c_log) | c;
      const uint32_t pb = ((ka + half) << c_log) | c;
      const Fr w = Fr::load(tw + (((tile_index(geo.base, geo.cstride, geo.kstride, c, ka) & jmask) << tw_shift) << 3));
      const Fr x = Fr::reduce_2p(lds_get(lds0, lds1, pa));
      const Fr t = Fr::mul_lazy(lds_get(lds0, lds1, pb), w);
      lds_put(lds0, lds1, pa, Fr::add_raw(x, t));
      lds_put(lds0, lds1, pb, Fr::sub_2p(x, t));
    }
    __syncthreads();
  }
}

// The tile back to memory; the transform's last pass (final) stores
// canonical values.
template <int NT>
__device__ __forceinline__ void pass_store(const uint4* __restrict__ lds0, const uint4* __restrict__ lds1,
                                           const PolyView& v, const TileGeom& geo, uint32_t d, uint32_t c_log,
                                           bool final_pass) {
  const uint32_t T = 1u << (d + c_log);
  for (uint32_t e = threadIdx.x; e < T; e += NT) {
    uint32_t pos;
    uint4* dst = reinterpret_cast<uint4*>(v.elem(tile_elem(geo, d, c_log, e, pos)));
    if (final_pass) {
      const Fr y = Fr::reduce_4p(lds_get(lds0, lds1, pos));
      dst[0] = make_uint4(y.v[0], y.v[1], y.v[2], y.v[3]);
      dst[1] = make_uint4(y.v[4], y.v[5], y.v[6], y.v[7]);
    } else {
      dst[0] = lds0[pos];
      dst[1] = lds1[pos];
    }
  }
}

// One tile per workgroup of NT threads (U quads per thread per round trip):
// load, stages, store.
template <int NT, int U>
__device__ __forceinline__ void dit_tile_r4(uint4* __restrict__ lds0, uint4* __restrict__ lds1, const PolyView& v,
                                            const uint32_t* __restrict__ tw, uint32_t log_n, uint32_t s0, uint32_t d,
                                            uint32_t c_log, uint64_t tile, bool final_pass) {
  const uint32_t T = 1u << (d + c_log);
  const TileGeom geo = tile_geom(s0, d, c_log, tile);
  for (uint32_t e = threadIdx.x; e < T; e += NT) {
    uint32_t pos;
    const uint4* src = reinterpret_cast<const uint4*>(v.elem(tile_elem(geo, d, c_log, e, pos)));
    lds0[pos] = src[0];
    lds1[pos] = src[1];
  }
  __syncthreads();
  pass_stages<NT, U>(lds0, lds1, tw, log_n, s0, d, c_log, geo);
  pass_store<NT>(lds0, lds1, v, geo, d, c_log, final_pass);
}

// Passes run 512 threads per workgroup, one quad per thread per round trip:
// with two workgroups per CU (64 KB of LDS each) that is four waves per SIMD in
// at most 128 VGPRs, whose issue covers the waits (barriers, tile loads) that
// two waves left exposed (2^24: 2.34 -> 2.18 ms).
constexpr int PASS_THREADS = 512;
__global__ void __launch_bounds__(PASS_THREADS)
    dit_pass_r4(PolyView v, const uint32_t* __restrict__ tw, uint32_t log_n, uint32_t s0, uint32_t d, uint32_t c_log,
                uint32_t final_pass) {
  __shared__ uint4 lds0[1 << TILE_LOG];
  __shared__ uint4 lds1[1 << TILE_LOG];
  dit_tile_r4<PASS_THREADS, 1>(lds0, lds1, v, tw, log_n, s0, d, c_log, blockIdx.x, final_pass != 0);
}

// The earlier shape: 256 threads, two quads each per round trip (variant 3).
__global__ void __launch_bounds__(THREADS)
    dit_pass_r4_narrow(PolyView v, const uint32_t* __restrict__ tw, uint32_t log_n, uint32_t s0, uint32_t d,
                       uint32_t c_log, uint32_t final_pass) {
  __shared__ uint4 lds0[1 << TILE_LOG];
  __shared__ uint4 lds1[1 << TILE_LOG];
  dit_tile_r4<THREADS, 2>(lds0, lds1, v, tw, log_n, s0, d, c_log, blockIdx.x, final_pass != 0);
}

__global__ void __launch_bounds__(PASS_THREADS)
    dit_pass_r4_batched(const BatchInstance* __restrict__ inst, const BatchPass* __restrict__ pass,
                        const uint64_t* __restrict__ tile_prefix, uint32_t count) {
  __shared__ uint4 lds0[1 << TILE_LOG];
  __shared__ uint4 lds1[1 << TILE_LOG];
  const uint64_t b = blockIdx.x;
  if (b >= tile_prefix[count]) return;
  const uint32_t k = find_instance(tile_prefix, count, b);
  const BatchPass p = pass[k];
  dit_tile_r4<PASS_THREADS, 1>(lds0, lds1, inst[k].v, inst[k].tw, inst[k].log_n, p.s0, p.d, p.c_log,
                               b - tile_prefix[k], p.pad != 0);
}

// ---------------------------------------------------------------- recompute NTT
//
// recompute_ntt (core/src/ntt/src/recompute_ntt.cuh:361-438) runs the same
// transform without the n/2-entry twiddle table: the caller passes
// pq[i] = w^(i * 2^(log_n - pq_deg)) for i < 2^(pq_deg - 1) and the ladder
// omegas[i] = w^(2^i), i < 32 (gen_pq_omegas).  Every twiddle is formed here as
// an exact product of those powers, so it equals the table entry w^e for the
// same exponent e and the result is bit-identical to ssip_ntt / best_fft.
//
// Twiddle of stage s = s0 + m at global index i (i mod 2^s = lo + kp * 2^s0,
// lo < 2^s0 the tile column, kp < 2^m the row bits):
//   w^((i mod 2^s) * 2^(L-1-s)) = colbase(lo, m) * small(kp << (d - 1 - m))
//   colbase(lo, m) = w^(lo * 2^(L-1-s0-m)),  small(j) = w^(j * 2^(L-d)).
// A workgroup walks consecutive tiles, so lo grows by C from one tile to the
// next and colbase is advanced with one product per (column, stage).
constexpr int RTILE_LOG = 10;          // 32 KiB of tile + 16 KiB of stage twiddles
constexpr uint32_t RMAX_C_LOG = 5;     // <= 32 columns
constexpr uint32_t RMAX_D = RTILE_LOG - 1;
constexpr uint32_t RMAX_COLBASE = 160;  // max C * d over plans with c <= 5, c + d <= 10
constexpr uint32_t RBLOCKS = 1024;

// w^(e * 2^sh) as the product of omegas[b + sh] over the set bits b of e.
__device__ __forceinline__ Fr pow_ladder(const uint32_t* __restrict__ omegas, uint64_t e, uint32_t sh) {
  Fr r = Fr::one();
  bool first = true;
  while (e) {
    const uint32_t b = __ffsll((unsigned long long)e) - 1;
    e &= e - 1;
    const Fr f = Fr::load(omegas + (b + sh) * 8);
    r = first ? f : r * f;
    first = false;
  }
  return r;
}

// r[j] = w^(j * 2^(log_n - D)) for j < 2^(D - 1): copied out of pq when it is
// fine enough, else built by doubling from the ladder.  Block-cooperative (any
// block size); ends with a barrier.
__device__ void build_small_roots(uint4* r0, uint4* r1, const uint32_t* __restrict__ pq, uint32_t pq_deg,
                                  const uint32_t* __restrict__ omegas, uint32_t log_n, uint32_t D) {
  const uint32_t cnt = (1u << D) >> 1;
  const uint32_t nt = blockDim.x;
  if (pq_deg >= D && pq_deg <= log_n) {
    for (uint32_t j = threadIdx.x; j < cnt; j += nt) lds_put(r0, r1, j, Fr::load(pq + ((uint64_t)j << (pq_deg - D)) * 8));
  } else {
    if (threadIdx.x == 0) lds_put(r0, r1, 0, Fr::one());
    for (uint32_t i = 0; (1u << i) < cnt; i++) {
      __syncthreads();
      const Fr g = Fr::load(omegas + (log_n - D + i) * 8);
      for (uint32_t j = threadIdx.x; j < (1u << i); j += nt) lds_put(r0, r1, j + (1u << i), lds_get(r0, r1, j) * g);
    }
  }
  __syncthreads();
}

// The persistent head walk (bitrev_head_stream) with the head's twiddles built
// from pq / the omega ladder.
__global__ void __launch_bounds__(HEAD_THREADS)
    rbitrev_head_stream(PolyView v, const uint32_t* __restrict__ pq, uint32_t pq_deg,
                        const uint32_t* __restrict__ omegas, uint32_t log_n, uint32_t t, uint32_t g) {
  __shared__ HeadLds L;
  build_small_roots(L.r0, L.r1, pq, pq_deg, omegas, log_n, t);
  head_walk(L, v, log_n, t, g, blockIdx.x);
}

struct RPassLds {
  uint4 x0[1 << RTILE_LOG], x1[1 << RTILE_LOG];              // tile
  uint4 w0[1 << (RTILE_LOG - 1)], w1[1 << (RTILE_LOG - 1)];  // twiddles of the current stage
  uint4 r0[1 << (RMAX_D - 1)], r1[1 << (RMAX_D - 1)];        // small roots
  uint4 b0[RMAX_COLBASE], b1[RMAX_COLBASE];                  // colbase of the current tile, [c * d + m]
  uint4 z0[RMAX_COLBASE], z1[RMAX_COLBASE];                  // colbase at lo_block 0
};

// Stages [s0, s0 + d) over tiles [blockIdx.x * per_block, ...) of 2^d rows x
// 2^c_log columns, tile geometry as dit_tile.
// 512 threads (one butterfly each per stage at the 1024-element tile); lazy
// butterflies, values in [0, 4p) between stages and passes, canonical stores
// in the transform's last pass (final_pass).
constexpr int RPASS_THREADS = 512;
__global__ void __launch_bounds__(RPASS_THREADS)
    rdit_pass(PolyView v, const uint32_t* __restrict__ pq, uint32_t pq_deg, const uint32_t* __restrict__ omegas,
              uint32_t log_n, uint32_t s0, uint32_t d, uint32_t c_log, uint64_t tiles, uint32_t per_block,
              uint32_t final_pass) {
  __shared__ RPassLds S;
  const uint32_t C = 1u << c_log;
  const uint32_t R = 1u << d;
  const uint32_t T = R << c_log;
  const uint64_t first = (uint64_t)blockIdx.x * per_block;
  if (first >= tiles) return;
  const uint64_t last = first + per_block < tiles ? first + per_block : tiles;
  const uint64_t lo_blocks = s0 ? ((1ull << s0) >> c_log) : 1;

  if (s0) {
    const uint64_t lob = first % lo_blocks;
    for (uint32_t e = threadIdx.x; e < C * d; e += RPASS_THREADS) {
      const uint32_t c = e / d, m = e - c * d;
      const uint32_t sh = log_n - 1 - s0 - m;
      lds_put(S.z0, S.z1, e, pow_ladder(omegas, c, sh));
      lds_put(S.b0, S.b1, e, pow_ladder(omegas, lob * C + c, sh));
    }
  }
  build_small_roots(S.r0, S.r1, pq, pq_deg, omegas, log_n, d);  // ends with a barrier

  for (uint64_t tile = first; tile < last; tile++) {
    uint64_t base, cstride, kstride;
    if (s0 == 0) {
      base = (tile << c_log) << d;
      cstride = R;
      kstride = 1;
    } else {
      const uint64_t hi = tile / lo_blocks;
      const uint64_t lo_block = tile % lo_blocks;
      base = lo_block * C + (hi << (s0 + d));
      cstride = 1;
      kstride = 1ull << s0;
      if (tile != first) {  // lo grows by C, or wraps to 0 with the next hi
        const bool wrap = lo_block == 0;
        for (uint32_t e = threadIdx.x; e < C * d; e += RPASS_THREADS) {
          const uint32_t m = e % d;
          const Fr nb = wrap ? lds_get(S.z0, S.z1, e)
                             : lds_get(S.b0, S.b1, e) * Fr::load(omegas + (c_log + log_n - 1 - s0 - m) * 8);
          lds_put(S.b0, S.b1, e, nb);
        }
      }
    }
    for (uint32_t e = threadIdx.x; e < T; e += RPASS_THREADS) {
      uint32_t c, k;
      if (cstride == 1) {
        c = e & (C - 1);
        k = e >> c_log;
      } else {
        k = e & (R - 1);
        c = e >> d;
      }
      const uint4* src = reinterpret_cast<const uint4*>(v.elem(tile_index(base, cstride, kstride, c, k)));
      const uint32_t pos = (k << c_log) | c;
      S.x0[pos] = src[0];
      S.x1[pos] = src[1];
    }
    __syncthreads();
    for (uint32_t m = 0; m < d; m++) {
      const uint32_t h = 1u << m;
      for (uint32_t e = threadIdx.x; e < (C << m); e += RPASS_THREADS) {
        const uint32_t c = e >> m, kp = e & (h - 1);
        Fr w = lds_get(S.r0, S.r1, kp << (d - 1 - m));
        if (s0) w = lds_get(S.b0, S.b1, c * d + m) * w;
        lds_put(S.w0, S.w1, e, w);
      }
      __syncthreads();
      for (uint32_t b = threadIdx.x; b < (T >> 1); b += RPASS_THREADS) {
        const uint32_t c = b & (C - 1);
        const uint32_t q = b >> c_log;
        const uint32_t ka = ((q >> m) << (m + 1)) | (q & (h - 1));
        const uint32_t pa = (ka << c_log) | c;
        const uint32_t pb = ((ka + h) << c_log) | c;
        const Fr w = lds_get(S.w0, S.w1, (c << m) | (ka & (h - 1)));
        const Fr x = Fr::reduce_2p(lds_get(S.x0, S.x1, pa));
        const Fr t = Fr::mul_lazy(lds_get(S.x0, S.x1, pb), w);
        lds_put(S.x0, S.x1, pa, Fr::add_raw(x, t));
        lds_put(S.x0, S.x1, pb, Fr::sub_2p(x, t));
      }
      __syncthreads();
    }
    for (uint32_t e = threadIdx.x; e < T; e += RPASS_THREADS) {
      uint32_t c, k;
      if (cstride == 1) {
        c = e & (C - 1);
        k = e >> c_log;
      } else {
        k = e & (R - 1);
        c = e >> d;
      }
      uint4* dst = reinterpret_cast<uint4*>(v.elem(tile_index(base, cstride, kstride, c, k)));
      const uint32_t pos = (k << c_log) | c;
      if (final_pass) {
        const Fr y = Fr::reduce_4p(lds_get(S.x0, S.x1, pos));
        dst[0] = make_uint4(y.v[0], y.v[1], y.v[2], y.v[3]);
        dst[1] = make_uint4(y.v[4], y.v[5], y.v[6], y.v[7]);
      } else {
        dst[0] = S.x0[pos];
        dst[1] = S.x1[pos];
      }
    }
    __syncthreads();
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

// v2 plan: a fused bit-reversal + first t stages (t = min(5, log_n / 2)), then
// radix-4 passes of at most TILE_LOG - min_col_log stages each, balanced, with
// as many adjacent columns as the tile holds.
struct PlanV2 {
  uint32_t head_t;  // 0: plain bit reversal, no fused stages
  int np;
  PassPlan pass[48];
};

// Tuning state, read by every launch and written by the autotuner hooks
// (zkp_ntt_set_variant / zkp_ntt_set_tuning).  Both are atomics and a launch
// reads each once, so a setter that races with launches on other threads
// changes which plan the next launch picks but never mixes two plans.
//
// per log_n minimum column exponent for v2 passes (runs of >= 2^c elements).
// The default 1 (fewest, deepest passes) measured fastest at 2^20..2^24 with
// 512-thread passes.
static std::atomic<uint32_t> g_min_col_log[41] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                                  1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
// 0: persistent head + radix-4 passes; 1: plain bit reversal + radix-2 passes;
// 2: as 0 with one tile pair per head workgroup; 3: as 0 with 256-thread
// passes (batch plans treat 2 and 3 as 0)
static std::atomic<int> g_variant{0};

static void plan_v2(uint32_t log_n, PlanV2* p) {
  uint32_t t = log_n / 2;
  if (t > HEAD_T_MAX) t = HEAD_T_MAX;
  p->head_t = t;
  p->np = 0;
  uint32_t s0 = t;
  uint32_t min_c = g_min_col_log[log_n < 41 ? log_n : 40].load(std::memory_order_relaxed);
  if (min_c > TILE_LOG - 1) min_c = TILE_LOG - 1;
  const uint32_t max_d = TILE_LOG - min_c;
  while (s0 < log_n) {
    const uint32_t rem = log_n - s0;
    const uint32_t parts = (rem + max_d - 1) / max_d;
    const uint32_t d = (rem + parts - 1) / parts;
    uint32_t c = TILE_LOG - d;
    if (s0 == 0) {
      if (c > log_n - d) c = log_n - d;
    } else if (c > s0) {
      c = s0;
    }
    p->pass[p->np++] = PassPlan{s0, d, c, (1ull << log_n) >> (d + c)};
    s0 += d;
  }
}

// Recompute plan: the fused head (t = min(5, log_n / 2)), then balanced passes
// of <= RMAX_D stages over tiles of <= 2^RTILE_LOG elements with <= 2^RMAX_C_LOG
// columns.
static void plan_recompute(uint32_t log_n, PlanV2* p) {
  uint32_t t = log_n / 2;
  if (t > HEAD_T_MAX) t = HEAD_T_MAX;
  p->head_t = t;
  p->np = 0;
  uint32_t s0 = t;
  while (s0 < log_n) {
    const uint32_t rem = log_n - s0;
    const uint32_t parts = (rem + RMAX_D - 1) / RMAX_D;
    const uint32_t d = (rem + parts - 1) / parts;
    uint32_t c = RTILE_LOG - d;
    if (c > RMAX_C_LOG) c = RMAX_C_LOG;
    if (s0 == 0) {
      if (c > log_n - d) c = log_n - d;
    } else if (c > s0) {
      c = s0;
    }
    p->pass[p->np++] = PassPlan{s0, d, c, (1ull << log_n) >> (d + c)};
    s0 += d;
  }
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
  auto mark = [&](int kind) -> hipError_t {
    if (kinds && nk < max_events) kinds[nk] = kind;
    nk++;
    if (events && ev < max_events) return hipEventRecord(events[ev++], stream);
    return hipSuccess;
  };
  if (events && ev < max_events) {
    err = hipEventRecord(events[ev++], stream);
    if (err != hipSuccess) return err;
  }
  const int variant = g_variant.load(std::memory_order_relaxed);
  if (variant == 1) {
    hipLaunchKernelGGL(bitrev_kernel, dim3(grid_for(v.len, THREADS, 65536)), dim3(THREADS), 0, stream, v,
                       (uint32_t)log_len);
    if ((err = hipGetLastError()) != hipSuccess || (err = mark(0)) != hipSuccess) return err;
    PassPlan plan[48];
    const int np = plan_passes(log_len, plan);
    for (int p = 0; p < np; p++) {
      hipLaunchKernelGGL(dit_pass, dim3((unsigned)plan[p].tiles), dim3(THREADS), 0, stream, v, twiddle,
                         (uint32_t)log_len, plan[p].s0, plan[p].d, plan[p].c_log);
      if ((err = hipGetLastError()) != hipSuccess || (err = mark(1)) != hipSuccess) return err;
    }
  } else {
    PlanV2 plan;
    plan_v2(log_len, &plan);
    if (plan.head_t == 0) {
      hipLaunchKernelGGL(bitrev_kernel, dim3(grid_for(v.len, THREADS, 65536)), dim3(THREADS), 0, stream, v,
                         (uint32_t)log_len);
    } else {
      if (variant == 2) {
        hipLaunchKernelGGL(bitrev_head, dim3((unsigned)(1ull << (log_len - 2 * plan.head_t))), dim3(THREADS), 0,
                           stream, v, twiddle, (uint32_t)log_len, plan.head_t);
      } else {
        const uint32_t g = head_stream_g((uint32_t)log_len - 2 * plan.head_t);
        hipLaunchKernelGGL(bitrev_head_stream, dim3(1u << g), dim3(HEAD_THREADS), 0, stream, v, twiddle,
                           (uint32_t)log_len, plan.head_t, g);
      }
    }
    if ((err = hipGetLastError()) != hipSuccess || (err = mark(0)) != hipSuccess) return err;
    for (int p = 0; p < plan.np; p++) {
      const PassPlan& pp = plan.pass[p];
      if (variant == 3)
        hipLaunchKernelGGL(dit_pass_r4_narrow, dim3((unsigned)pp.tiles), dim3(THREADS), 0, stream, v, twiddle,
                           (uint32_t)log_len, pp.s0, pp.d, pp.c_log, (uint32_t)(p + 1 == plan.np));
      else
        hipLaunchKernelGGL(dit_pass_r4, dim3((unsigned)pp.tiles), dim3(PASS_THREADS), 0, stream, v, twiddle,
                           (uint32_t)log_len, pp.s0, pp.d, pp.c_log, (uint32_t)(p + 1 == plan.np));
      if ((err = hipGetLastError()) != hipSuccess || (err = mark(1)) != hipSuccess) return err;
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
  int variant = 0;
  BatchInstance* d_inst = nullptr;
  uint64_t* d_elem_prefix = nullptr;  // variant 1: elements; variant 0: head blocks
  uint64_t total_elems = 0;
  uint64_t total_head_blocks = 0;
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

// Transform variant for all later calls and batch plans in this process:
// 0 = fused bit reversal + radix-4 passes (default), 1 = separate bit reversal
// + radix-2 passes.  Returns the previous variant, or -1 for an unknown one.
// Safe to call while other threads launch transforms: each launch reads the
// variant once.  No counterpart in the reference ABI (tuning hook).
int zkp_ntt_set_variant(int variant) {
  if (variant < 0 || variant > 3) return -1;
  return zkp::ntt::g_variant.exchange(variant);
}

// Tuning of transforms of 2^log_n (variant 0): every pass reads runs of at
// least 2^min_col_log adjacent elements, so a smaller value means fewer,
// deeper passes.  Returns 0, or -1 when out of range.  No counterpart in the
// reference ABI (autotuner hook).
int zkp_ntt_set_tuning(unsigned int log_n, unsigned int min_col_log) {
  if (log_n > 40 || min_col_log > (unsigned)zkp::ntt::TILE_LOG - 1) return -1;
  zkp::ntt::g_min_col_log[log_n].store(min_col_log, std::memory_order_relaxed);
  return 0;
}

// Number of kernels one transform of 2^log_n launches with the current
// variant and tuning (0 for log_n == 0 or > 40).
int zkp_ntt_kernel_count(unsigned int log_n) {
  if (log_n == 0 || log_n > 40) return 0;
  if (zkp::ntt::g_variant.load() == 1) {
    zkp::ntt::PassPlan plan[48];
    return 1 + zkp::ntt::plan_passes(log_n, plan);
  }
  zkp::ntt::PlanV2 p;
  zkp::ntt::plan_v2(log_n, &p);
  return 1 + p.np;
}

// In-place natural-order forward NTT of x (len = 2^log_len) whose twiddles are
// recomputed from pq (2^(pq_deg - 1) entries) and the 32-entry ladder omegas,
// both device arrays written by gen_pq_omegas.  Replaces recompute_ntt of
// core/src/ntt/src/ntt.h:12 / recompute_ntt.cuh:361-438 (bound by
// core/src/ntt.rs:249-289).  Allocates nothing; asynchronous on stream.  The
// thread-shape hints are accepted but not needed here.
hipError_t recompute_ntt(PolyPtr x, const unsigned int* pq, unsigned int pq_deg, const unsigned int* omegas,
                         unsigned int log_len, hipStream_t stream, const unsigned int max_threads_stage1_log,
                         const unsigned int max_threads_stage2_log) {
  (void)max_threads_stage1_log;
  (void)max_threads_stage2_log;
  using namespace zkp::ntt;
  PolyView v;
  if (!make_view(x, &v)) return hipErrorInvalidValue;
  if (log_len > 32 || v.len != (1ull << log_len)) return hipErrorInvalidValue;
  if (log_len == 0) return hipSuccess;
  if (v.ptr == nullptr || pq == nullptr || omegas == nullptr || pq_deg == 0) return hipErrorInvalidValue;
  PlanV2 plan;
  plan_recompute(log_len, &plan);
  if (plan.head_t == 0) {
    hipLaunchKernelGGL(bitrev_kernel, dim3(grid_for(v.len, THREADS, 65536)), dim3(THREADS), 0, stream, v,
                       (uint32_t)log_len);
  } else {
    const uint32_t g = head_stream_g((uint32_t)log_len - 2 * plan.head_t);
    hipLaunchKernelGGL(rbitrev_head_stream, dim3(1u << g), dim3(HEAD_THREADS), 0, stream, v, pq, (uint32_t)pq_deg,
                       omegas, (uint32_t)log_len, plan.head_t, g);
  }
  hipError_t err = hipGetLastError();
  for (int p = 0; p < plan.np && err == hipSuccess; p++) {
    const PassPlan& q = plan.pass[p];
    const uint64_t blocks = q.tiles < RBLOCKS ? q.tiles : RBLOCKS;
    const uint32_t per_block = (uint32_t)((q.tiles + blocks - 1) / blocks);
    const unsigned grid = (unsigned)((q.tiles + per_block - 1) / per_block);
    hipLaunchKernelGGL(rdit_pass, dim3(grid), dim3(RPASS_THREADS), 0, stream, v, pq, (uint32_t)pq_deg, omegas,
                       (uint32_t)log_len, q.s0, q.d, q.c_log, q.tiles, per_block, (uint32_t)(p + 1 == plan.np));
    err = hipGetLastError();
  }
  return err;
}

// pq[i] = (unit^(len >> pq_deg))^i for i < 2^(pq_deg - 1) (pq[0] = 1 is always
// written) and omegas[i] = unit^(2^i) for i < 32; host arrays, synchronous.
// Replaces gen_pq_omegas of core/src/ntt/src/ntt.h:15 / precompute.cuh:88-112
// (bound by core/src/ntt.rs:337-371).
void gen_pq_omegas(unsigned int* pq, unsigned int* omegas, unsigned int pq_deg, unsigned int len,
                   unsigned int* unit) {
  using zkp::Fr;
  auto put = [](unsigned int* dst, const Fr& x) { memcpy(dst, x.v, sizeof(x.v)); };
  const Fr u = Fr::from_raw(unit);
  const uint64_t cnt = pq_deg < 64 ? (1ull << pq_deg) >> 1 : 0;
  put(pq, Fr::one());
  if (cnt > 1) {
    const Fr step = u.pow(pq_deg < 32 ? (uint64_t)(len >> pq_deg) : 0);
    Fr cur = step;
    for (uint64_t i = 1; i < cnt; i++) {
      put(pq + i * 8, cur);
      cur = cur * step;
    }
  }
  Fr cur = u;
  for (int i = 0; i < 32; i++) {
    put(omegas + i * 8, cur);
    cur = cur.sqr();
  }
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
  plan->variant = g_variant.load();
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
    inst[i] = BatchInstance{v, twiddles[i], log_lens[i], 0, 0};
    if (log_lens[i] == 0) {
      elem_prefix[i + 1] = elem_prefix[i];
      continue;
    }
    if (plan->variant == 1) {
      elem_prefix[i + 1] = elem_prefix[i] + v.len;
      PassPlan pp[48];
      const int np = plan_passes(log_lens[i], pp);
      plans[i].assign(pp, pp + np);
    } else {
      PlanV2 p2;
      plan_v2(log_lens[i], &p2);
      inst[i].head_t = p2.head_t;
      inst[i].head_g = p2.head_t ? head_stream_g(log_lens[i] - 2 * p2.head_t) : 0;
      // head_t == 0 only for log_n == 1, whose bit reversal is the identity
      elem_prefix[i + 1] = elem_prefix[i] + (p2.head_t ? (1ull << inst[i].head_g) : 0);
      plans[i].assign(p2.pass, p2.pass + p2.np);
    }
    if (plans[i].size() > rounds) rounds = plans[i].size();
  }
  if (plan->variant == 1)
    plan->total_elems = elem_prefix[count];
  else
    plan->total_head_blocks = elem_prefix[count];
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
        passes[i] = BatchPass{q.s0, q.d, q.c_log, (uint32_t)(r + 1 == plans[i].size())};
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
  hipError_t err = hipSuccess;
  if (plan->variant == 1) {
    if (plan->total_elems == 0) return hipSuccess;
    hipLaunchKernelGGL(bitrev_batched, dim3(grid_for(plan->total_elems, THREADS, 65536)), dim3(THREADS), 0, stream,
                       plan->d_inst, plan->d_elem_prefix, plan->count);
    err = hipGetLastError();
  } else if (plan->total_head_blocks > 0) {
    hipLaunchKernelGGL(bitrev_head_batched, dim3((unsigned)plan->total_head_blocks), dim3(HEAD_THREADS), 0, stream,
                       plan->d_inst, plan->d_elem_prefix, plan->count);
    err = hipGetLastError();
  }
  for (size_t r = 0; r < plan->d_pass.size() && err == hipSuccess; r++) {
    if (plan->total_tiles[r] == 0) continue;
    if (plan->variant == 1)
      hipLaunchKernelGGL(dit_pass_batched, dim3((unsigned)plan->total_tiles[r]), dim3(THREADS), 0, stream,
                         plan->d_inst, plan->d_pass[r], plan->d_tile_prefix[r], plan->count);
    else
      hipLaunchKernelGGL(dit_pass_r4_batched, dim3((unsigned)plan->total_tiles[r]), dim3(PASS_THREADS), 0, stream,
                         plan->d_inst, plan->d_pass[r], plan->d_tile_prefix[r], plan->count);
    err = hipGetLastError();
  }
  return err;
}

// Frees a plan (after the work using it has completed).
void zkp_ntt_batch_destroy(void* plan_ptr) { delete reinterpret_cast<BatchPlan*>(plan_ptr); }

}  // extern "C"
