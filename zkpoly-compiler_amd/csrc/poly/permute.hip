// permute -- the lookup-argument permutation of libpoly.so for gfx950.
//
// C ABI of the reference's permute (core/src/poly/src/poly.h:28,
// permute.cuh:84-221), bound by PolyPermute (core/src/poly.rs:221-280).  For
// the first `usable` rows of input A and table S:
//   A' = A sorted ascending by canonical value;
//   S'_i = A'_i where row i starts a run of equal values in A' (one copy of the
//   value leaves the table multiset);
//   the other rows take the values left in the table, ascending, in row order
// (halo2's permute_expression_pair, as core/tests/poly.rs:635-690 states it).
// Every input value should occur in the table.  Where one does not, the
// reference stops with a device assert; here that row keeps its own value and
// takes no copy out, and the other rows take the smallest leftovers.  Every
// access stays inside its buffers either way.
//
// Steps, all asynchronous on `stream`, allocating nothing:
//  1. canonical keys of A and S into two key buffers (one launch for both);
//  2. sort, the same launches serving both arrays: every workgroup sorts a tile
//     of 2048 keys in LDS (4 per thread by a sorting network, then merges of
//     runs 4 .. 1024); merge passes then double the sorted runs through HBM:
//     one thread per output tile finds the tile's merge-path split, then each
//     workgroup merges its 2048 outputs staged in LDS.  Keys are whole 256-bit
//     values compared by a borrow chain, so the cost does not depend on how
//     the values are spread (small lookup tables, random-looking compressed
//     expressions and negatives alike);
//  3. every run start of A' finds its lower bound in S' and marks that table
//     slot used; A' is written out in Montgomery form;
//  4. one scan over rows of (run starts, unused slots) counts (tile folds, then
//     a rescan by wave ballots) compacts the unused slots in order into the
//     leftover list, writes S' at run starts, and gives every other row its
//     rank among the non-starts;
//  5. the row of rank r takes leftover r.
#include "scan.hpp"

namespace zkp {
namespace poly {
namespace perm {

constexpr int STILE = 2048;        // keys per LDS tile: 64 KiB
constexpr int TTH = 512;           // threads of a tile-sort workgroup
constexpr int TPER = STILE / TTH;  // keys per thread of the tile sort (a 4-key network, then merges)
constexpr int MTH = 1024;          // threads of a merge-pass workgroup (the same 2048-key tile)
constexpr int MPER = STILE / MTH;  // outputs per thread
constexpr uint32_t START = 0xffffffffu;

// a < b as 256-bit integers: the borrow out of a - b.
__device__ __forceinline__ bool key_lt(const Fr& a, const Fr& b) {
  u32 borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) (void)sbb(a.v[i], b.v[i], borrow);
  return borrow != 0;
}

__device__ __forceinline__ void cas(Fr& a, Fr& b) {
  if (key_lt(b, a)) {
    const Fr t = a;
    a = b;
    b = t;
  }
}

// LDS holds keys as 16-byte chunks; chunk c sits at c ^ ((c >> 4) & 15), so a
// wave reading key 8t + j in lane t spreads over all chunk slots of a row.
__device__ __forceinline__ uint32_t swz(uint32_t c) { return c ^ ((c >> 4) & 15u); }

__device__ __forceinline__ Fr lds_get(const uint4* s, uint32_t k) {
  const uint4 a = s[swz(2 * k)], b = s[swz(2 * k + 1)];
  Fr r;
  r.v[0] = a.x, r.v[1] = a.y, r.v[2] = a.z, r.v[3] = a.w;
  r.v[4] = b.x, r.v[5] = b.y, r.v[6] = b.z, r.v[7] = b.w;
  return r;
}

__device__ __forceinline__ void lds_put(uint4* s, uint32_t k, const Fr& x) {
  s[swz(2 * k)] = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
  s[swz(2 * k + 1)] = make_uint4(x.v[4], x.v[5], x.v[6], x.v[7]);
}

// Count of A's elements among the first d outputs of merge(A, B), ties going
// to A.  Reads A[mid] with mid < la and B[d - 1 - mid] with d - 1 - mid < lb.
template <class GA, class GB>
__device__ __forceinline__ uint32_t merge_path(const GA& A, uint32_t la, const GB& B, uint32_t lb, uint32_t d) {
  uint32_t lo = d > lb ? d - lb : 0, hi = d < la ? d : la;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (!key_lt(B(d - 1 - mid), A(mid)))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// m (<= N) outputs of merge(A, B) from split (i, j), ties going to A.
template <int N, class GA, class GB>
__device__ __forceinline__ void merge_run(const GA& A, uint32_t la, const GB& B, uint32_t lb, uint32_t i, uint32_t j,
                                          uint32_t m, Fr (&out)[N]) {
  Fr a = i < la ? A(i) : Fr::zero();
  Fr b = j < lb ? B(j) : Fr::zero();
#pragma unroll
  for (int q = 0; q < N; q++) {
    if ((uint32_t)q < m) {
      const bool take_a = i < la && (j >= lb || !key_lt(b, a));
      if (take_a) {
        out[q] = a;
        if (++i < la) a = A(i);
      } else {
        out[q] = b;
        if (++j < lb) b = B(j);
      }
    }
  }
}

__device__ __forceinline__ Fr* pick(Fr* a0, Fr* a1) { return blockIdx.y ? a1 : a0; }

// 1. keys[k] = canonical value of view element k, k < n (blockIdx.y: input / table).
__global__ void __launch_bounds__(TH) k_to_keys(PolyView a, PolyView t, Fr* ka, Fr* kt, uint32_t n) {
  const PolyView v = blockIdx.y ? t : a;
  Fr* out = pick(ka, kt);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    ld(v, k).to_canonical().store(out[k].v);
}

// 2a. Sorts keys[b*STILE, min((b+1)*STILE, n)) in place; the tile's missing
// tail is padded with 2^256 - 1, above every canonical value.  Each of the TTH
// threads sorts TPER keys in registers, then runs of TPER .. STILE / 2 are
// merged pairwise in LDS (short serial chains and many waves: the merges wait
// on LDS latency, not on VALU).
__global__ void __launch_bounds__(TTH) k_tile_sort(Fr* k0, Fr* k1, uint32_t n) {
  __shared__ uint4 s[2 * STILE];
  Fr* keys = pick(k0, k1);
  const uint32_t base = blockIdx.x * STILE;
  const uint32_t cnt = n - base < (uint32_t)STILE ? n - base : (uint32_t)STILE;
  const uint4* src = reinterpret_cast<const uint4*>(keys + base);
  for (uint32_t c = threadIdx.x; c < 2u * STILE; c += TTH)
    s[swz(c)] = (c >> 1) < cnt ? src[c] : make_uint4(~0u, ~0u, ~0u, ~0u);
  __syncthreads();
  const uint32_t d0 = threadIdx.x * TPER;
  {
    static_assert(TPER == 4, "the register network below sorts four keys");
    Fr x[TPER];
#pragma unroll
    for (int q = 0; q < TPER; q++) x[q] = lds_get(s, d0 + q);
    cas(x[0], x[1]), cas(x[2], x[3]), cas(x[0], x[2]), cas(x[1], x[3]), cas(x[1], x[2]);
#pragma unroll
    for (int q = 0; q < TPER; q++) lds_put(s, d0 + q, x[q]);  // the thread's own slots
  }
  __syncthreads();
  for (uint32_t L = TPER; L < (uint32_t)STILE; L <<= 1) {
    // the thread's outputs [d0, d0 + TPER) lie in one pair of runs (2L > TPER)
    const uint32_t pb = d0 & ~(2 * L - 1);
    const uint32_t d = d0 - pb;
    const auto A = [&](uint32_t i) { return lds_get(s, pb + i); };
    const auto B = [&](uint32_t j) { return lds_get(s, pb + L + j); };
    const uint32_t i = merge_path(A, L, B, L, d);
    Fr out[TPER];
    merge_run<TPER>(A, L, B, L, i, d - i, TPER, out);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPER; q++) lds_put(s, d0 + q, out[q]);
    __syncthreads();
  }
  uint4* dst = reinterpret_cast<uint4*>(keys + base);
  for (uint32_t c = threadIdx.x; c < 2u * cnt; c += TTH) dst[c] = s[swz(c)];
}

// The pair of runs of L (the last run may be short, or missing) holding
// output position o: base pb, run lengths la and lb; B starts at pb + la
// (la == L whenever lb > 0).
struct Pair {
  uint64_t pb;
  uint32_t la, lb;
};

__device__ __forceinline__ Pair pair_of(uint64_t o, uint32_t n, uint32_t L) {
  Pair p;
  p.pb = o / (2ull * L) * (2ull * L);
  const uint64_t rest = n - p.pb;
  p.la = (uint32_t)(rest < L ? rest : L);
  p.lb = (uint32_t)(rest > L ? (rest - L < L ? rest - L : L) : 0);
  return p;
}

// 2b. split[y][b] = A's share of the outputs before b*STILE within its pair,
// b < tiles: every search runs in its own thread, so no workgroup of the merge
// waits on a chain of dependent loads.
__global__ void __launch_bounds__(TH) k_merge_split(const Fr* in0, const Fr* in1, uint32_t n, uint32_t L,
                                                    uint32_t tiles, uint32_t* split) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= tiles) return;
  const Fr* in = blockIdx.y ? in1 : in0;
  const uint64_t o = (uint64_t)b * STILE;
  const Pair p = pair_of(o, n, L);
  const Fr* A = in + p.pb;
  const Fr* B = A + p.la;
  const auto GA = [&](uint32_t i) { return Fr::load(A[i].v); };
  const auto GB = [&](uint32_t j) { return Fr::load(B[j].v); };
  split[blockIdx.y * tiles + b] = merge_path(GA, p.la, GB, p.lb, (uint32_t)(o - p.pb));
}

// 2c. Runs of L sorted keys merged pairwise into runs of 2L: workgroup b
// writes outputs [b*STILE, min((b+1)*STILE, n)) from the splits of 2b.
__global__ void __launch_bounds__(MTH) k_merge_pass(const Fr* in0, Fr* out0, const Fr* in1, Fr* out1, uint32_t n,
                                                     uint32_t L, const uint32_t* split) {
  __shared__ uint4 s[2 * STILE];
  const Fr* in = blockIdx.y ? in1 : in0;
  Fr* out = pick(out0, out1);
  const uint32_t tiles = gridDim.x;
  const uint64_t o0 = (uint64_t)blockIdx.x * STILE;
  const uint64_t o1 = o0 + STILE < n ? o0 + STILE : n;
  const Pair p = pair_of(o0, n, L);
  const uint64_t pb = p.pb;
  const uint32_t la = p.la, lb = p.lb;
  const Fr* A = in + pb;
  const Fr* B = A + la;
  const uint32_t d0 = (uint32_t)(o0 - pb), d1 = (uint32_t)(o1 - pb);
  // the end of the tile is the end of its pair, or the start of the next tile
  // inside the same pair
  const uint32_t i0 = split[blockIdx.y * tiles + blockIdx.x];
  const uint32_t i1 = d1 == la + lb ? la : split[blockIdx.y * tiles + blockIdx.x + 1];
  const uint32_t j0 = d0 - i0, j1 = d1 - i1;
  const uint32_t na = i1 - i0, nb = j1 - j0;  // na + nb = o1 - o0 <= STILE
  const uint32_t tot = na + nb;
  for (uint32_t c = threadIdx.x; c < 2 * tot; c += MTH) {
    const uint32_t q = c >> 1;
    const uint4* src = q < na ? reinterpret_cast<const uint4*>(A + i0 + q)
                              : reinterpret_cast<const uint4*>(B + j0 + (q - na));
    s[swz(c)] = src[c & 1];
  }
  __syncthreads();
  const uint32_t d = threadIdx.x * MPER;
  const uint32_t m = d < tot ? (tot - d < (uint32_t)MPER ? tot - d : (uint32_t)MPER) : 0;
  Fr res[MPER];
  if (m) {
    const auto LA = [&](uint32_t i) { return lds_get(s, i); };
    const auto LB = [&](uint32_t j) { return lds_get(s, na + j); };
    const uint32_t i = merge_path(LA, na, LB, nb, d);
    merge_run<MPER>(LA, na, LB, nb, i, d - i, m, res);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MPER; q++)
    if ((uint32_t)q < m) lds_put(s, d + q, res[q]);
  __syncthreads();
  uint4* dst = reinterpret_cast<uint4*>(out + o0);
  for (uint32_t c = threadIdx.x; c < 2 * tot; c += MTH) dst[c] = s[swz(c)];
}

__global__ void __launch_bounds__(TH) k_zero(uint32_t* x, uint32_t n) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) x[k] = 0;
}

// 3. A' out in Montgomery form; start[k] = 1 where row k starts a run; each run
// start marks the lower bound of its value in S' when the value is there.
__global__ void __launch_bounds__(TH) k_lookup(const Fr* sa, const Fr* ss, uint32_t n, uint32_t* used,
                                               uint32_t* start, PolyView res_input) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const Fr a = Fr::load(sa[k].v);
    st(res_input, k, Fr::from_canonical(a));
    const bool first = k == 0 || a != Fr::load(sa[k - 1].v);
    start[k] = first;
    if (first) {
      uint32_t lo = 0, hi = n;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (key_lt(Fr::load(ss[mid].v), a))
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo < n && Fr::load(ss[lo].v) == a) used[lo] = 1;  // distinct run values: distinct slots
    }
  }
}

// 4. Counts (run starts, unused table slots) before row k.
struct CountOp {
  struct T {
    uint32_t f, u;
  };
  __device__ static T id() { return T{0, 0}; }
  __device__ static T comb(const T& l, const T& r) { return T{l.f + r.f, l.u + r.u}; }
};

struct SrcCounts {
  const uint32_t* start;
  const uint32_t* used;
  __device__ CountOp::T operator()(uint64_t k) const { return CountOp::T{start[k], 1u - used[k]}; }
};

// Unused slot k goes to leftover[#unused before k]; a run start writes S'_k =
// A'_k and rank[k] = START; any other row keeps its rank among non-starts.
struct SinkCounts {
  const Fr* sa;
  const Fr* ss;
  Fr* leftover;
  uint32_t* rank;  // aliases `start`: element k is read before it is written
  PolyView res_table;
  __device__ void operator()(uint64_t k, const CountOp::T& excl, const CountOp::T& e, const CountOp::T&) const {
    if (e.u) Fr::load(ss[k].v).store(leftover[excl.u].v);
    if (e.f) {
      st(res_table, k, Fr::from_canonical(Fr::load(sa[k].v)));
      rank[k] = START;
    } else {
      rank[k] = (uint32_t)k - excl.f;
    }
  }
};

// 4b. The rescan of a tile for 0/1 flags, one element per lane: each row of TH
// consecutive elements gets its exclusive counts from wave ballots (the set
// lanes below a lane, by mbcnt) and the totals of the waves before it (LDS);
// rows follow one another from the tile's carry.  Loads and stores are
// coalesced, with no per-thread serial chain.
__global__ void __launch_bounds__(TH) k_count_tile(const uint32_t* __restrict__ start,
                                                   const uint32_t* __restrict__ used, uint64_t n,
                                                   const CountOp::T* __restrict__ carry, SinkCounts sink) {
  constexpr uint32_t WAVES = TH / 64;
  __shared__ uint32_t wf[WAVES], wu[WAVES];
  const uint32_t wave = threadIdx.x >> 6;
  CountOp::T run = carry[blockIdx.x];
  for (uint32_t r = 0; r < (uint32_t)PER; r++) {
    const uint64_t k = (uint64_t)blockIdx.x * TILE + (uint64_t)r * TH + threadIdx.x;
    const bool in = k < n;
    const uint32_t f = in ? start[k] : 0u;
    const uint32_t u = in ? 1u - used[k] : 0u;
    const unsigned long long mf = __ballot(f != 0), mu = __ballot(u != 0);
    const uint32_t bf = __builtin_amdgcn_mbcnt_hi((uint32_t)(mf >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mf, 0u));
    const uint32_t bu = __builtin_amdgcn_mbcnt_hi((uint32_t)(mu >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mu, 0u));
    if ((threadIdx.x & 63) == 0) {
      wf[wave] = (uint32_t)__popcll(mf);
      wu[wave] = (uint32_t)__popcll(mu);
    }
    __syncthreads();
    CountOp::T excl = run;
    uint32_t rf = 0, ru = 0;
#pragma unroll
    for (uint32_t w = 0; w < WAVES; w++) {
      if (w < wave) excl.f += wf[w], excl.u += wu[w];
      rf += wf[w], ru += wu[w];
    }
    __syncthreads();  // wf / wu are rewritten by the next row
    excl.f += bf;
    excl.u += bu;
    if (in) sink(k, excl, CountOp::T{f, u}, CountOp::T{excl.f + f, excl.u + u});
    run.f += rf;
    run.u += ru;
  }
}

// 5. The non-start row of rank r takes leftover r (r < #non-starts <= #unused).
__global__ void __launch_bounds__(TH) k_fill_back(const uint32_t* rank, const Fr* leftover, uint32_t n,
                                                  PolyView res_table) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const uint32_t r = rank[k];
    if (r != START) st(res_table, k, Fr::from_canonical(Fr::load(leftover[r].v)));
  }
}

struct Layout {
  size_t keys, flags, part, split, total;
};

static Layout layout(uint64_t n) {
  const uint64_t m = n ? n : 1;
  Layout l;
  l.keys = align256(m * sizeof(Fr));
  l.flags = align256(m * sizeof(uint32_t));
  l.part = align256(tiles_of(m) * sizeof(CountOp::T));
  l.split = align256(2 * ((m + STILE - 1) / STILE) * sizeof(uint32_t));
  l.total = 4 * l.keys + 2 * l.flags + 2 * l.part + l.split;
  return l;
}

}  // namespace perm
}  // namespace poly
}  // namespace zkp

using namespace zkp;
using namespace zkp::poly;
using namespace zkp::poly::perm;

extern "C" {

// (A', S') = lookup permutation of the first `usable` rows of (input, table)
// (poly.h:28)
hipError_t permute(void* temp_buffer, unsigned long* buffer_size, unsigned long usable, ConstPolyPtr input,
                   ConstPolyPtr table, PolyPtr res_input, PolyPtr res_table, hipStream_t stream) {
  if (usable >= 0xffffffffull) return hipErrorInvalidValue;  // rows are counted in 32 bits
  const uint32_t n = (uint32_t)usable;
  const Layout lay = layout(n);
  if (temp_buffer == nullptr) {
    if (buffer_size == nullptr) return hipErrorInvalidValue;
    *buffer_size = lay.total;
    return hipSuccess;
  }
  PolyView vi, vt, ri, rt;
  if (!make_view(input, &vi) || !make_view(table, &vt) || !make_view(res_input, &ri) || !make_view(res_table, &rt))
    return hipErrorInvalidValue;
  // the reference's asserts (permute.cuh:122-125)
  if (vi.len != vt.len || ri.len != usable || rt.len != usable || usable > vi.len) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  if (!vi.ptr || !vt.ptr || !ri.ptr || !rt.ptr) return hipErrorInvalidValue;

  char* t = reinterpret_cast<char*>(temp_buffer);
  Fr* ka[2] = {reinterpret_cast<Fr*>(t), reinterpret_cast<Fr*>(t + lay.keys)};
  Fr* ks[2] = {reinterpret_cast<Fr*>(t + 2 * lay.keys), reinterpret_cast<Fr*>(t + 3 * lay.keys)};
  uint32_t* used = reinterpret_cast<uint32_t*>(t + 4 * lay.keys);
  uint32_t* start = reinterpret_cast<uint32_t*>(t + 4 * lay.keys + lay.flags);
  auto* partial = reinterpret_cast<CountOp::T*>(t + 4 * lay.keys + 2 * lay.flags);
  auto* carry = reinterpret_cast<CountOp::T*>(t + 4 * lay.keys + 2 * lay.flags + lay.part);
  uint32_t* split = reinterpret_cast<uint32_t*>(t + 4 * lay.keys + 2 * lay.flags + 2 * lay.part);

  const unsigned g = grid_for(n);
  hipLaunchKernelGGL(k_to_keys, dim3(g, 2), dim3(TH), 0, stream, vi, vt, ka[0], ks[0], n);
  const unsigned tiles = (n + STILE - 1) / STILE;
  hipLaunchKernelGGL(k_tile_sort, dim3(tiles, 2), dim3(TTH), 0, stream, ka[0], ks[0], n);
  int cur = 0;
  for (uint64_t L = STILE; L < n; L <<= 1) {
    hipLaunchKernelGGL(k_merge_split, dim3((tiles + TH - 1) / TH, 2), dim3(TH), 0, stream, ka[cur], ks[cur], n,
                       (uint32_t)L, tiles, split);
    hipLaunchKernelGGL(k_merge_pass, dim3(tiles, 2), dim3(MTH), 0, stream, ka[cur], ka[cur ^ 1], ks[cur],
                       ks[cur ^ 1], n, (uint32_t)L, split);
    cur ^= 1;
  }
  const Fr* sa = ka[cur];
  const Fr* ss = ks[cur];
  Fr* leftover = ks[cur ^ 1];
  hipLaunchKernelGGL(k_zero, dim3(g), dim3(TH), 0, stream, used, n);
  hipLaunchKernelGGL(k_lookup, dim3(g), dim3(TH), 0, stream, sa, ss, n, used, start, ri);
  const uint64_t cnb = tiles_of(n);
  hipLaunchKernelGGL((k_tile_fold<CountOp, SrcCounts>), dim3((unsigned)cnb), dim3(TH), 0, stream,
                     SrcCounts{start, used}, (uint64_t)n, partial);
  hipLaunchKernelGGL((k_partials_scan<CountOp>), dim3(1), dim3(TH), 0, stream, partial, cnb, carry,
                     (CountOp::T*)nullptr);
  hipLaunchKernelGGL(k_count_tile, dim3((unsigned)cnb), dim3(TH), 0, stream, start, used, (uint64_t)n, carry,
                     SinkCounts{sa, ss, leftover, start, rt});
  hipLaunchKernelGGL(k_fill_back, dim3(g), dim3(TH), 0, stream, start, leftover, n, rt);
  return hipGetLastError();
}

}  // extern "C"
