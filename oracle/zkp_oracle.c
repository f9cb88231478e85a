/*
 * zkp_oracle.c -- CPU oracle for the BN254 prover hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
 * file: only tests/, the smoke check and the bench's CPU-baseline leg use it,
 * and only as the checker.  It is deliberately independent of the device code:
 * 4 x u64 limbs with 128-bit products (the device uses 8 x u32 CIOS), Jacobian
 * curve coordinates (the device uses XYZZ).
 *
 * What it restates (paths relative to the reference repository):
 *  - halo2curves 0.6.0 field semantics (not vendored; restated from the
 *    mathematics): canonical Montgomery form x*2^256 mod p, 4 x u64 LE limbs,
 *    byte-identical to the reference's 8 x u32 layout
 *    (core/src/common/mont/src/field.cuh:637-644).
 *  - best_fft + recursive_butterfly_arithmetic: core/tests/msm_halo2.rs:221-324
 *    (bit-reversal, twiddles = [1, w, ..., w^(n/2-1)] by running product,
 *    recursive radix-2 DIT with a thread split at the top levels).
 *  - multiexp_serial / best_multiexp: core/tests/msm_halo2.rs:63-209
 *    (window c = ceil(ln n), segments = NUM_BITS/c + 1, bucket summation by
 *    parts, per-thread chunking n/threads, sum of chunk results).
 *  - omega choice ROOT_OF_UNITY^(2^(S-k)): compiler/src/utils.rs:51-77.
 *  - INTT = NTT(w^-1) * n^-1: compiler/src/transit/type2/intt_mending.rs:27-59.
 *  - distribute_powers x[i] *= powers[i mod m]:
 *    core/src/ntt/src/distribute_powers.cuh:4-19.
 *  - SliceIterator index map (offset + i) % whole_len with rotate folded into
 *    offset = len - rotate: core/src/common/iter/src/iter.cuh:309-327.
 *  - curve y^2 = x^3 + 3 over Fq, identity affine (0,0):
 *    core/src/common/curve/src/bn254.cuh:11-18, curve_xyzz.cuh:204-210.
 *  - polynomial helpers the reference's poly tests compare against
 *    (core/tests/poly.rs): eval_polynomial by Horner from the top (:173-176),
 *    halo2_proofs::arithmetic::kate_division (q_i = a_{i+1} + b q_{i+1} from
 *    the top, n - 1 outputs; :240-244), ff's batch_invert (zeros skipped;
 *    :464-470), the running product of scan_mul (:408-411) and the lookup
 *    permutation of test_poly_permute (:635-690).
 *
 * How each piece is pinned (tests/test_oracle_*.py): field ops against the
 * reference's own known-answer vectors (core/src/common/mont/tests/main.cu)
 * and Python big-int closed forms; curve ops against the reference KATs
 * (core/src/common/curve/tests/bn254.cu:174-203); best_fft against an O(n^2)
 * DFT and INTT/NTT round trips; multiexp against double-and-add and the
 * closed form (sum k_i s_i) G for bases k_i G.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef uint32_t u32;
typedef unsigned __int128 u128;

typedef struct {
  u64 v[4];
} fe;

typedef struct {
  u64 p[4];
  u64 inv; /* -p^-1 mod 2^64 */
  u64 r2[4];
  u64 one[4];
  u32 num_bits;
} fparams;

static const fparams FR = {
    {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
    0xc2e1f593efffffffULL,
    {0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL, 0x0216d0b17f4e44a5ULL},
    {0xac96341c4ffffffbULL, 0x36fc76959f60cd29ULL, 0x666ea36f7879462eULL, 0x0e0a77c19a07df2fULL},
    254};
static const fparams FQ = {
    {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL},
    0x87d20782e4866389ULL,
    {0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL, 0x47ab1eff0a417ff6ULL, 0x06d89f71cab8351fULL},
    {0xd35d438dc58f0d9dULL, 0x0a78eb28f5c70b3dULL, 0x666ea36f7879462cULL, 0x0e0a77c19a07df2fULL},
    254};

static const fparams* field_of(int fid) { return fid == 0 ? &FR : &FQ; }

/* ---------------- 256-bit field arithmetic (4 x u64) ---------------- */

static inline int geq_p(const u64 a[4], const u64 p[4]) {
  for (int i = 3; i >= 0; i--) {
    if (a[i] != p[i]) return a[i] > p[i];
  }
  return 1;
}

static inline void sub_p(u64 a[4], const u64 p[4]) {
  u64 borrow = 0;
  for (int i = 0; i < 4; i++) {
    u128 t = (u128)a[i] - p[i] - borrow;
    a[i] = (u64)t;
    borrow = (u64)(t >> 64) & 1;
  }
}

static inline void f_add(const fparams* F, fe* o, const fe* a, const fe* b) {
  u64 s[4];
  u64 c = 0;
  for (int i = 0; i < 4; i++) {
    u128 t = (u128)a->v[i] + b->v[i] + c;
    s[i] = (u64)t;
    c = (u64)(t >> 64);
  }
  if (c || geq_p(s, F->p)) sub_p(s, F->p);
  memcpy(o->v, s, 32);
}

static inline void f_sub(const fparams* F, fe* o, const fe* a, const fe* b) {
  u64 d[4];
  u64 borrow = 0;
  for (int i = 0; i < 4; i++) {
    u128 t = (u128)a->v[i] - b->v[i] - borrow;
    d[i] = (u64)t;
    borrow = (u64)(t >> 64) & 1;
  }
  if (borrow) {
    u64 c = 0;
    for (int i = 0; i < 4; i++) {
      u128 t = (u128)d[i] + F->p[i] + c;
      d[i] = (u64)t;
      c = (u64)(t >> 64);
    }
  }
  memcpy(o->v, d, 32);
}

static inline int f_is_zero(const fe* a) { return (a->v[0] | a->v[1] | a->v[2] | a->v[3]) == 0; }

static inline void f_neg(const fparams* F, fe* o, const fe* a) {
  if (f_is_zero(a)) {
    memset(o, 0, sizeof(fe));
    return;
  }
  fe z;
  memcpy(z.v, F->p, 32);
  f_sub(F, o, &z, a);
}

/* Montgomery product a*b*2^-256 mod p: separated-operand CIOS on 64-bit words. */
static inline void f_mul(const fparams* F, fe* o, const fe* a, const fe* b) {
  u64 t[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; i++) {
    u64 c = 0;
    for (int j = 0; j < 4; j++) {
      u128 x = (u128)a->v[j] * b->v[i] + t[j] + c;
      t[j] = (u64)x;
      c = (u64)(x >> 64);
    }
    u128 x = (u128)t[4] + c;
    t[4] = (u64)x;
    t[5] = (u64)(x >> 64);
    u64 m = t[0] * F->inv;
    x = (u128)m * F->p[0] + t[0];
    c = (u64)(x >> 64);
    for (int j = 1; j < 4; j++) {
      x = (u128)m * F->p[j] + t[j] + c;
      t[j - 1] = (u64)x;
      c = (u64)(x >> 64);
    }
    x = (u128)t[4] + c;
    t[3] = (u64)x;
    t[4] = t[5] + (u64)(x >> 64);
  }
  u64 r[4] = {t[0], t[1], t[2], t[3]};
  if (t[4] || geq_p(r, F->p)) sub_p(r, F->p);
  memcpy(o->v, r, 32);
}

static inline void f_one(const fparams* F, fe* o) { memcpy(o->v, F->one, 32); }

/* a^e, e given as nwords little-endian u64 words; MSB-first square-and-multiply. */
static void f_pow(const fparams* F, fe* o, const fe* a, const u64* e, int nwords) {
  fe r, base = *a;
  f_one(F, &r);
  for (int w = nwords - 1; w >= 0; w--) {
    for (int bit = 63; bit >= 0; bit--) {
      f_mul(F, &r, &r, &r);
      if ((e[w] >> bit) & 1) f_mul(F, &r, &r, &base);
    }
  }
  *o = r;
}

static void f_inv(const fparams* F, fe* o, const fe* a) {
  u64 e[4];
  memcpy(e, F->p, 32);
  /* p - 2; the low word of both moduli is > 2 */
  e[0] -= 2;
  f_pow(F, o, a, e, 4);
}

static void f_from_canonical(const fparams* F, fe* o, const fe* x) {
  fe r2;
  memcpy(r2.v, F->r2, 32);
  f_mul(F, o, x, &r2);
}

static void f_to_canonical(const fparams* F, fe* o, const fe* x) {
  fe one = {{1, 0, 0, 0}};
  f_mul(F, o, x, &one);
}

/* ---------------- exported scalar field ops (fid 0 = Fr, 1 = Fq) ---------------- */

void orc_mul(int fid, u64* o, const u64* a, const u64* b) { f_mul(field_of(fid), (fe*)o, (const fe*)a, (const fe*)b); }
void orc_add(int fid, u64* o, const u64* a, const u64* b) { f_add(field_of(fid), (fe*)o, (const fe*)a, (const fe*)b); }
void orc_sub(int fid, u64* o, const u64* a, const u64* b) { f_sub(field_of(fid), (fe*)o, (const fe*)a, (const fe*)b); }
void orc_neg(int fid, u64* o, const u64* a) { f_neg(field_of(fid), (fe*)o, (const fe*)a); }
void orc_inv(int fid, u64* o, const u64* a) { f_inv(field_of(fid), (fe*)o, (const fe*)a); }
void orc_pow(int fid, u64* o, const u64* a, const u64* e, int nwords) { f_pow(field_of(fid), (fe*)o, (const fe*)a, e, nwords); }
void orc_from_canonical(int fid, u64* o, const u64* x) { f_from_canonical(field_of(fid), (fe*)o, (const fe*)x); }
void orc_to_canonical(int fid, u64* o, const u64* x) { f_to_canonical(field_of(fid), (fe*)o, (const fe*)x); }

/* Elementwise vector ops over n elements (used for fused-kernel parity checks). */
void orc_vec_mul(int fid, u64* o, const u64* a, const u64* b, u64 n) {
  const fparams* F = field_of(fid);
  for (u64 i = 0; i < n; i++) f_mul(F, (fe*)o + i, (const fe*)a + i, (const fe*)b + i);
}
void orc_vec_add(int fid, u64* o, const u64* a, const u64* b, u64 n) {
  const fparams* F = field_of(fid);
  for (u64 i = 0; i < n; i++) f_add(F, (fe*)o + i, (const fe*)a + i, (const fe*)b + i);
}
void orc_vec_sub(int fid, u64* o, const u64* a, const u64* b, u64 n) {
  const fparams* F = field_of(fid);
  for (u64 i = 0; i < n; i++) f_sub(F, (fe*)o + i, (const fe*)a + i, (const fe*)b + i);
}
void orc_vec_neg(int fid, u64* o, const u64* a, u64 n) {
  const fparams* F = field_of(fid);
  for (u64 i = 0; i < n; i++) f_neg(F, (fe*)o + i, (const fe*)a + i);
}
void orc_vec_pow(int fid, u64* o, const u64* a, u64 e, u64 n) {
  const fparams* F = field_of(fid);
  for (u64 i = 0; i < n; i++) f_pow(F, (fe*)o + i, (const fe*)a + i, &e, 1);
}

/* ---------------- thread helper ---------------- */

typedef void* (*job_fn)(void* arg);

static void run_parallel(job_fn fn, void* args, size_t arg_size, int njobs) {
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)njobs);
  for (int i = 1; i < njobs; i++)
    pthread_create(&th[i], NULL, fn, (char*)args + (size_t)i * arg_size);
  fn(args);
  for (int i = 1; i < njobs; i++) pthread_join(th[i], NULL);
  free(th);
}

/* ---------------- best_fft (core/tests/msm_halo2.rs:221-324) ---------------- */

static u64 bitreverse(u64 n, u32 l) {
  u64 r = 0;
  for (u32 i = 0; i < l; i++) {
    r = (r << 1) | (n & 1);
    n >>= 1;
  }
  return r;
}

typedef struct {
  fe* a;
  u64 n;
  u64 twiddle_chunk;
  const fe* tw;
  int depth; /* remaining levels at which the two halves run on separate threads */
} rba_args;

static void rba(rba_args* x);

static void* rba_thread(void* p) {
  rba((rba_args*)p);
  return NULL;
}

/* recursive_butterfly_arithmetic: halves first (joined), then the twiddled combine. */
static void rba(rba_args* x) {
  const fparams* F = &FR;
  fe* a = x->a;
  u64 n = x->n;
  if (n == 2) {
    fe t = a[1];
    a[1] = a[0];
    f_add(F, &a[0], &a[0], &t);
    f_sub(F, &a[1], &a[1], &t);
    return;
  }
  rba_args l = {a, n / 2, x->twiddle_chunk * 2, x->tw, x->depth - 1};
  rba_args r = {a + n / 2, n / 2, x->twiddle_chunk * 2, x->tw, x->depth - 1};
  if (x->depth > 0) {
    pthread_t th;
    pthread_create(&th, NULL, rba_thread, &l);
    rba(&r);
    pthread_join(th, NULL);
  } else {
    rba(&l);
    rba(&r);
  }
  fe* left = a;
  fe* right = a + n / 2;
  /* twiddle one */
  fe t = right[0];
  right[0] = left[0];
  f_add(F, &left[0], &left[0], &t);
  f_sub(F, &right[0], &right[0], &t);
  for (u64 i = 1; i < n / 2; i++) {
    fe u;
    f_mul(F, &u, &right[i], &x->tw[i * x->twiddle_chunk]);
    right[i] = left[i];
    f_add(F, &left[i], &left[i], &u);
    f_sub(F, &right[i], &right[i], &u);
  }
}

static u32 log2_floor(u64 v) {
  u32 p = 0;
  while (((u64)1 << (p + 1)) <= v) p++;
  return p;
}

/* In-place forward DFT X[j] = sum a[i] w^(ij) of a power-of-two Fr vector
 * (Montgomery form), natural order in and out, exactly best_fft's network. */
void orc_best_fft(u64* a_raw, const u64* omega_raw, u32 log_n, int threads) {
  fe* a = (fe*)a_raw;
  const fparams* F = &FR;
  u64 n = (u64)1 << log_n;
  if (threads < 1) threads = 1;
  for (u64 k = 0; k < n; k++) {
    u64 rk = bitreverse(k, log_n);
    if (k < rk) {
      fe t = a[k];
      a[k] = a[rk];
      a[rk] = t;
    }
  }
  if (n == 1) return;
  u64 half = n / 2;
  fe* tw = (fe*)malloc(sizeof(fe) * half);
  fe w;
  f_one(F, &w);
  for (u64 i = 0; i < half; i++) {
    tw[i] = w;
    f_mul(F, &w, &w, (const fe*)omega_raw);
  }
  /* both branches of the reference compute the same butterflies; the thread
   * split only changes who runs them */
  rba_args top = {a, n, 1, tw, (int)log2_floor((u64)threads)};
  rba(&top);
  free(tw);
}

/* The same transform over a 64-bit-word prime field < 2^32 (the reference's
 * small-field NTT test, core/src/ntt/tests/small_field.cuh:4-5), used only to
 * check that the butterfly network equals the O(n^2) DFT for a true root. */
void orc_small_field_dft(u64* out, const u64* in, u64 n, u64 root, u64 p) {
  for (u64 j = 0; j < n; j++) {
    u64 wj = 1, acc = 0, e = (j % n);
    /* wj = root^j */
    u64 b = root % p;
    wj = 1;
    while (e) {
      if (e & 1) wj = (u64)((u128)wj * b % p);
      b = (u64)((u128)b * b % p);
      e >>= 1;
    }
    u64 x = 1;
    for (u64 i = 0; i < n; i++) {
      acc = (u64)(((u128)in[i] * x + acc) % p);
      x = (u64)((u128)x * wj % p);
    }
    out[j] = acc;
  }
}

/* O(n^2) DFT over Fr, for pinning best_fft on small sizes. */
void orc_naive_dft(u64* out_raw, const u64* in_raw, u64 n, const u64* omega_raw) {
  const fparams* F = &FR;
  const fe* in = (const fe*)in_raw;
  fe* out = (fe*)out_raw;
  const fe* omega = (const fe*)omega_raw;
  fe wj;
  f_one(F, &wj);
  for (u64 j = 0; j < n; j++) {
    fe acc = {{0, 0, 0, 0}}, x;
    f_one(F, &x);
    for (u64 i = 0; i < n; i++) {
      fe t;
      f_mul(F, &t, &in[i], &x);
      f_add(F, &acc, &acc, &t);
      f_mul(F, &x, &x, &wj);
    }
    out[j] = acc;
    f_mul(F, &wj, &wj, omega);
  }
}

/* omega_k = ROOT_OF_UNITY^(2^(28-k)) in Montgomery form, optionally inverted
 * (compiler/src/utils.rs:51-77). */
void orc_omega(u64* out, u32 k, int inv) {
  const fparams* F = &FR;
  /* 7^((r-1)/2^28), Montgomery form */
  fe w = {{0x9632c7c5b639feb8ULL, 0x985ce3400d0ff299ULL, 0xb2dd880001b0ecd8ULL, 0x1d69070d6d98ce29ULL}};
  for (u32 i = k; i < 28; i++) f_mul(F, &w, &w, &w);
  if (inv) f_inv(F, &w, &w);
  memcpy(out, w.v, 32);
}

/* x[i] *= powers[i mod m] (core/src/ntt/src/distribute_powers.cuh:4-19) */
void orc_distribute_powers(u64* x_raw, u64 n, const u64* powers_raw, u64 m) {
  const fparams* F = &FR;
  fe* x = (fe*)x_raw;
  const fe* pw = (const fe*)powers_raw;
  for (u64 i = 0; i < n; i++) f_mul(F, &x[i], &x[i], &pw[i % m]);
}

/* sum_i a_i x^i by Horner from the top (core/tests/poly.rs:173-176) */
void orc_poly_eval(u64* out, const u64* a_raw, u64 n, const u64* x_raw) {
  const fparams* F = &FR;
  const fe* a = (const fe*)a_raw;
  fe acc;
  memset(&acc, 0, sizeof(acc));
  for (u64 i = n; i-- > 0;) {
    f_mul(F, &acc, &acc, (const fe*)x_raw);
    f_add(F, &acc, &acc, &a[i]);
  }
  memcpy(out, acc.v, 32);
}

/* halo2 kate_division: b' = -b; from the top, lead = a_j - tmp, q_{j-1} = lead,
 * tmp = lead * b'; n - 1 outputs q_0..q_{n-2}; q_{n-1} is written as 0 */
void orc_kate_division(u64* q_raw, const u64* a_raw, u64 n, const u64* b_raw) {
  const fparams* F = &FR;
  const fe* a = (const fe*)a_raw;
  fe* q = (fe*)q_raw;
  fe nb, tmp, lead;
  f_neg(F, &nb, (const fe*)b_raw);
  memset(&tmp, 0, sizeof(tmp));
  if (n == 0) return;
  memset(&q[n - 1], 0, sizeof(fe));
  for (u64 j = n - 1; j >= 1; j--) {
    f_sub(F, &lead, &a[j], &tmp);
    q[j - 1] = lead;
    f_mul(F, &tmp, &lead, &nb);
  }
}

/* every element inverted by Fermat; zero stays zero (ff batch_invert semantics) */
void orc_batch_invert(u64* x_raw, u64 n) {
  const fparams* F = &FR;
  fe* x = (fe*)x_raw;
  for (u64 i = 0; i < n; i++)
    if (!f_is_zero(&x[i])) f_inv(F, &x[i], &x[i]);
}

/* x_i = x_0 * ... * x_i in place */
void orc_scan_mul(u64* x_raw, u64 n) {
  const fparams* F = &FR;
  fe* x = (fe*)x_raw;
  for (u64 i = 1; i < n; i++) f_mul(F, &x[i], &x[i - 1], &x[i]);
}

/* halo2 permute_expression_pair as the reference's test restates it
 * (core/tests/poly.rs:635-690): A' = first `usable` inputs sorted by canonical
 * value; a row that starts a run of equal values in A' takes that value in S'
 * and one copy of it leaves the table multiset; the remaining table values,
 * ascending, fill the other rows in order.  An input value missing from the
 * table (the reference stops there) keeps its own value and takes no copy out;
 * the other rows then take the smallest leftovers. */
static int cmp_canon(const void* pa, const void* pb) {
  const u64* a = (const u64*)pa;
  const u64* b = (const u64*)pb;
  for (int i = 3; i >= 0; i--)
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  return 0;
}

void orc_permute(u64* out_in_raw, u64* out_tab_raw, const u64* in_raw, const u64* tab_raw, u64 usable) {
  const fparams* F = &FR;
  fe* sa = (fe*)malloc((usable ? usable : 1) * sizeof(fe));
  fe* st = (fe*)malloc((usable ? usable : 1) * sizeof(fe));
  unsigned char* used = (unsigned char*)calloc(usable ? usable : 1, 1);
  fe* out_in = (fe*)out_in_raw;
  fe* out_tab = (fe*)out_tab_raw;
  for (u64 i = 0; i < usable; i++) {
    f_to_canonical(F, &sa[i], (const fe*)in_raw + i);
    f_to_canonical(F, &st[i], (const fe*)tab_raw + i);
  }
  qsort(sa, usable, sizeof(fe), cmp_canon);
  qsort(st, usable, sizeof(fe), cmp_canon);
  /* run starts: match one table copy each (both lists ascending) */
  u64 t = 0;
  for (u64 i = 0; i < usable; i++) {
    if (i > 0 && cmp_canon(&sa[i], &sa[i - 1]) == 0) continue;
    while (t < usable && cmp_canon(&st[t], &sa[i]) < 0) t++;
    if (t < usable && cmp_canon(&st[t], &sa[i]) == 0) used[t++] = 1;
    f_from_canonical(F, &out_tab[i], &sa[i]);
  }
  /* leftovers, ascending, into the repeated rows in row order */
  u64 next = 0;
  for (u64 i = 0; i < usable; i++) {
    if (i == 0 || cmp_canon(&sa[i], &sa[i - 1]) != 0) continue;
    while (used[next]) next++;
    f_from_canonical(F, &out_tab[i], &st[next++]);
  }
  for (u64 i = 0; i < usable; i++) f_from_canonical(F, &out_in[i], &sa[i]);
  free(sa);
  free(st);
  free(used);
}

/* twiddle table [1, w, ..., w^(len-1)] by running product (ssip_precompute,
 * core/src/ntt/src/precompute.cuh:8-86 computes the same prefix product). */
void orc_twiddles(u64* out_raw, u64 len, const u64* omega_raw) {
  const fparams* F = &FR;
  fe* out = (fe*)out_raw;
  fe w;
  f_one(F, &w);
  for (u64 i = 0; i < len; i++) {
    out[i] = w;
    f_mul(F, &w, &w, (const fe*)omega_raw);
  }
}

/* gen_pq_omegas (core/src/ntt/src/precompute.cuh:88-112):
 * pq[i] = (unit^(len >> pq_deg))^i for i < 2^(pq_deg-1); omegas[i] = unit^(2^i), i < 32 */
void orc_gen_pq_omegas(u64* pq_raw, u64* omegas_raw, u32 pq_deg, u64 len, const u64* unit_raw) {
  const fparams* F = &FR;
  fe* pq = (fe*)pq_raw;
  fe* om = (fe*)omegas_raw;
  const fe* unit = (const fe*)unit_raw;
  u64 cnt = ((u64)1 << pq_deg) >> 1;
  memset(pq, 0, sizeof(fe) * cnt);
  f_one(F, &pq[0]);
  u64 e = len >> pq_deg;
  fe tw;
  f_pow(F, &tw, unit, &e, 1);
  if (pq_deg > 1) {
    pq[1] = tw;
    fe last = tw;
    for (u64 i = 2; i < cnt; i++) {
      f_mul(F, &last, &last, &tw);
      pq[i] = last;
    }
  }
  om[0] = *unit;
  fe last = *unit;
  for (u32 i = 1; i < 32; i++) {
    f_mul(F, &last, &last, &last);
    om[i] = last;
  }
}

/* ---------------- BN254 G1 (Jacobian, a = 0, b = 3) ---------------- */

typedef struct {
  fe x, y;
} g1a; /* affine; identity = (0,0) */
typedef struct {
  fe x, y, z;
} g1j; /* Jacobian; identity has z = 0 */

static void j_identity(g1j* p) { memset(p, 0, sizeof(g1j)); }
static int j_is_identity(const g1j* p) { return f_is_zero(&p->z); }
static int a_is_identity(const g1a* p) { return f_is_zero(&p->x) && f_is_zero(&p->y); }

static void j_from_affine(g1j* o, const g1a* a) {
  if (a_is_identity(a)) {
    j_identity(o);
    return;
  }
  o->x = a->x;
  o->y = a->y;
  f_one(&FQ, &o->z);
}

/* dbl-2009-l */
static void j_double(g1j* o, const g1j* p) {
  const fparams* F = &FQ;
  if (j_is_identity(p) || f_is_zero(&p->y)) {
    j_identity(o);
    return;
  }
  fe A, B, C, D, E, Fv, t, X3, Y3, Z3;
  f_mul(F, &A, &p->x, &p->x);
  f_mul(F, &B, &p->y, &p->y);
  f_mul(F, &C, &B, &B);
  f_add(F, &t, &p->x, &B);
  f_mul(F, &t, &t, &t);
  f_sub(F, &t, &t, &A);
  f_sub(F, &t, &t, &C);
  f_add(F, &D, &t, &t);
  f_add(F, &E, &A, &A);
  f_add(F, &E, &E, &A);
  f_mul(F, &Fv, &E, &E);
  f_sub(F, &X3, &Fv, &D);
  f_sub(F, &X3, &X3, &D);
  f_sub(F, &t, &D, &X3);
  f_mul(F, &Y3, &E, &t);
  fe C8;
  f_add(F, &C8, &C, &C);
  f_add(F, &C8, &C8, &C8);
  f_add(F, &C8, &C8, &C8);
  f_sub(F, &Y3, &Y3, &C8);
  f_mul(F, &Z3, &p->y, &p->z);
  f_add(F, &Z3, &Z3, &Z3);
  o->x = X3;
  o->y = Y3;
  o->z = Z3;
}

/* add-2007-bl with the equal / opposite / identity cases */
static void j_add(g1j* o, const g1j* p, const g1j* q) {
  const fparams* F = &FQ;
  if (j_is_identity(p)) {
    *o = *q;
    return;
  }
  if (j_is_identity(q)) {
    *o = *p;
    return;
  }
  fe Z1Z1, Z2Z2, U1, U2, S1, S2, H, I, J, r, V, t;
  f_mul(F, &Z1Z1, &p->z, &p->z);
  f_mul(F, &Z2Z2, &q->z, &q->z);
  f_mul(F, &U1, &p->x, &Z2Z2);
  f_mul(F, &U2, &q->x, &Z1Z1);
  f_mul(F, &S1, &p->y, &q->z);
  f_mul(F, &S1, &S1, &Z2Z2);
  f_mul(F, &S2, &q->y, &p->z);
  f_mul(F, &S2, &S2, &Z1Z1);
  f_sub(F, &H, &U2, &U1);
  f_sub(F, &r, &S2, &S1);
  if (f_is_zero(&H)) {
    if (f_is_zero(&r)) {
      j_double(o, p);
    } else {
      j_identity(o);
    }
    return;
  }
  f_add(F, &r, &r, &r);
  f_add(F, &I, &H, &H);
  f_mul(F, &I, &I, &I);
  f_mul(F, &J, &H, &I);
  f_mul(F, &V, &U1, &I);
  g1j res;
  f_mul(F, &res.x, &r, &r);
  f_sub(F, &res.x, &res.x, &J);
  f_sub(F, &res.x, &res.x, &V);
  f_sub(F, &res.x, &res.x, &V);
  f_sub(F, &t, &V, &res.x);
  f_mul(F, &res.y, &r, &t);
  f_mul(F, &t, &S1, &J);
  f_add(F, &t, &t, &t);
  f_sub(F, &res.y, &res.y, &t);
  f_add(F, &t, &p->z, &q->z);
  f_mul(F, &t, &t, &t);
  f_sub(F, &t, &t, &Z1Z1);
  f_sub(F, &t, &t, &Z2Z2);
  f_mul(F, &res.z, &t, &H);
  *o = res;
}

/* madd-2007-bl: Jacobian + affine */
static void j_add_affine(g1j* o, const g1j* p, const g1a* q) {
  const fparams* F = &FQ;
  if (a_is_identity(q)) {
    *o = *p;
    return;
  }
  if (j_is_identity(p)) {
    j_from_affine(o, q);
    return;
  }
  fe Z1Z1, U2, S2, H, HH, I, J, r, V, t;
  f_mul(F, &Z1Z1, &p->z, &p->z);
  f_mul(F, &U2, &q->x, &Z1Z1);
  f_mul(F, &S2, &q->y, &p->z);
  f_mul(F, &S2, &S2, &Z1Z1);
  f_sub(F, &H, &U2, &p->x);
  f_sub(F, &r, &S2, &p->y);
  if (f_is_zero(&H)) {
    if (f_is_zero(&r)) {
      g1j qj;
      j_from_affine(&qj, q);
      j_double(o, &qj);
    } else {
      j_identity(o);
    }
    return;
  }
  f_mul(F, &HH, &H, &H);
  f_add(F, &I, &HH, &HH);
  f_add(F, &I, &I, &I);
  f_mul(F, &J, &H, &I);
  f_add(F, &r, &r, &r);
  f_mul(F, &V, &p->x, &I);
  g1j res;
  f_mul(F, &res.x, &r, &r);
  f_sub(F, &res.x, &res.x, &J);
  f_sub(F, &res.x, &res.x, &V);
  f_sub(F, &res.x, &res.x, &V);
  f_sub(F, &t, &V, &res.x);
  f_mul(F, &res.y, &r, &t);
  f_mul(F, &t, &p->y, &J);
  f_add(F, &t, &t, &t);
  f_sub(F, &res.y, &res.y, &t);
  f_add(F, &t, &p->z, &H);
  f_mul(F, &t, &t, &t);
  f_sub(F, &t, &t, &Z1Z1);
  f_sub(F, &res.z, &t, &HH);
  *o = res;
}

static void j_to_affine(g1a* o, const g1j* p) {
  const fparams* F = &FQ;
  if (j_is_identity(p)) {
    memset(o, 0, sizeof(g1a));
    return;
  }
  fe zi, zi2, zi3;
  f_inv(F, &zi, &p->z);
  f_mul(F, &zi2, &zi, &zi);
  f_mul(F, &zi3, &zi2, &zi);
  f_mul(F, &o->x, &p->x, &zi2);
  f_mul(F, &o->y, &p->y, &zi3);
}

/* canonical little-endian 256-bit scalar (Fr "to_repr") */
static void scalar_repr(u64 out[4], const fe* s_mont) {
  fe c;
  f_to_canonical(&FR, &c, s_mont);
  memcpy(out, c.v, 32);
}

/* k * P by double-and-add over the canonical bits of k (any 256-bit integer). */
static void j_mul_u256(g1j* o, const g1a* p, const u64 k[4]) {
  g1j acc;
  j_identity(&acc);
  for (int w = 3; w >= 0; w--) {
    for (int b = 63; b >= 0; b--) {
      j_double(&acc, &acc);
      if ((k[w] >> b) & 1) j_add_affine(&acc, &acc, p);
    }
  }
  *o = acc;
}

/* exported curve helpers; points are affine {x, y} Fq Montgomery, 8 u64 words */
void orc_g1_add(u64* o, const u64* p, const u64* q) {
  g1j a, b, c;
  j_from_affine(&a, (const g1a*)p);
  j_from_affine(&b, (const g1a*)q);
  j_add(&c, &a, &b);
  j_to_affine((g1a*)o, &c);
}
void orc_g1_double(u64* o, const u64* p) {
  g1j a, c;
  j_from_affine(&a, (const g1a*)p);
  j_double(&c, &a);
  j_to_affine((g1a*)o, &c);
}
/* k given in canonical (non-Montgomery) little-endian form */
void orc_g1_mul_canonical(u64* o, const u64* p, const u64* k) {
  g1j c;
  j_mul_u256(&c, (const g1a*)p, k);
  j_to_affine((g1a*)o, &c);
}
/* k given as an Fr element in Montgomery form */
void orc_g1_mul_scalar(u64* o, const u64* p, const u64* k_mont) {
  u64 k[4];
  scalar_repr(k, (const fe*)k_mont);
  orc_g1_mul_canonical(o, p, k);
}
int orc_g1_on_curve(const u64* p) {
  const g1a* a = (const g1a*)p;
  if (a_is_identity(a)) return 1;
  const fparams* F = &FQ;
  fe y2, x3, b = {{3, 0, 0, 0}};
  f_from_canonical(F, &b, &b);
  f_mul(F, &y2, &a->y, &a->y);
  f_mul(F, &x3, &a->x, &a->x);
  f_mul(F, &x3, &x3, &a->x);
  f_add(F, &x3, &x3, &b);
  return memcmp(&x3, &y2, sizeof(fe)) == 0;
}

/* ---------------- multiexp (core/tests/msm_halo2.rs:63-209) ---------------- */

static u64 get_at(u64 segment, u64 c, const u64 repr[4]) {
  u64 skip_bits = segment * c;
  u64 skip_bytes = skip_bits / 8;
  if (skip_bytes >= (254 + 7) / 8) return 0;
  const uint8_t* bytes = (const uint8_t*)repr;
  uint8_t v[8] = {0};
  for (u64 i = 0; i < 8 && skip_bytes + i < 32; i++) v[i] = bytes[skip_bytes + i];
  u64 tmp;
  memcpy(&tmp, v, 8);
  tmp >>= skip_bits - skip_bytes * 8;
  tmp %= ((u64)1 << c);
  return tmp;
}

typedef enum { B_NONE = 0, B_AFFINE = 1, B_PROJ = 2 } bucket_kind;
typedef struct {
  int kind;
  g1a a;
  g1j p;
} bucket;

static void multiexp_serial(const fe* coeffs, const g1a* bases, u64 n, g1j* acc) {
  u64* reprs = (u64*)malloc(sizeof(u64) * 4 * (n ? n : 1));
  for (u64 i = 0; i < n; i++) scalar_repr(reprs + 4 * i, &coeffs[i]);
  u64 c;
  if (n < 4)
    c = 1;
  else if (n < 32)
    c = 3;
  else
    c = (u64)ceil(log((double)(u32)n));
  u64 segments = 254 / c + 1;
  u64 nb = ((u64)1 << c) - 1;
  bucket* buckets = (bucket*)malloc(sizeof(bucket) * nb);
  for (u64 seg = segments; seg-- > 0;) {
    for (u64 i = 0; i < c; i++) j_double(acc, acc);
    for (u64 b = 0; b < nb; b++) buckets[b].kind = B_NONE;
    for (u64 i = 0; i < n; i++) {
      u64 v = get_at(seg, c, reprs + 4 * i);
      if (v == 0) continue;
      bucket* bk = &buckets[v - 1];
      if (bk->kind == B_NONE) {
        bk->kind = B_AFFINE;
        bk->a = bases[i];
      } else if (bk->kind == B_AFFINE) {
        g1j t;
        j_from_affine(&t, &bk->a);
        j_add_affine(&bk->p, &t, &bases[i]);
        bk->kind = B_PROJ;
      } else {
        j_add_affine(&bk->p, &bk->p, &bases[i]);
      }
    }
    /* summation by parts */
    g1j running;
    j_identity(&running);
    for (u64 b = nb; b-- > 0;) {
      if (buckets[b].kind == B_AFFINE)
        j_add_affine(&running, &running, &buckets[b].a);
      else if (buckets[b].kind == B_PROJ)
        j_add(&running, &buckets[b].p, &running);
      j_add(acc, acc, &running);
    }
  }
  free(buckets);
  free(reprs);
}

typedef struct {
  const fe* coeffs;
  const g1a* bases;
  u64 n;
  g1j result;
} mexp_job;

static void* mexp_worker(void* p) {
  mexp_job* j = (mexp_job*)p;
  j_identity(&j->result);
  multiexp_serial(j->coeffs, j->bases, j->n, &j->result);
  return NULL;
}

/* best_multiexp: result affine {x, y} (Fq Montgomery), identity (0,0). */
void orc_best_multiexp(u64* out, const u64* coeffs_raw, const u64* bases_raw, u64 n, int threads) {
  const fe* coeffs = (const fe*)coeffs_raw;
  const g1a* bases = (const g1a*)bases_raw;
  if (threads < 1) threads = 1;
  g1j total;
  j_identity(&total);
  if (n > (u64)threads) {
    u64 chunk = n / (u64)threads;
    int num_chunks = (int)((n + chunk - 1) / chunk);
    mexp_job* jobs = (mexp_job*)calloc((size_t)num_chunks, sizeof(mexp_job));
    for (int i = 0; i < num_chunks; i++) {
      u64 off = (u64)i * chunk;
      jobs[i].coeffs = coeffs + off;
      jobs[i].bases = bases + off;
      jobs[i].n = (off + chunk <= n) ? chunk : n - off;
    }
    run_parallel(mexp_worker, jobs, sizeof(mexp_job), num_chunks);
    for (int i = 0; i < num_chunks; i++) j_add(&total, &total, &jobs[i].result);
    free(jobs);
  } else {
    multiexp_serial(coeffs, bases, n, &total);
  }
  j_to_affine((g1a*)out, &total);
}

/* Brute-force sum_i k_i P_i by per-point double-and-add (small n only). */
void orc_naive_msm(u64* out, const u64* coeffs_raw, const u64* bases_raw, u64 n) {
  const fe* coeffs = (const fe*)coeffs_raw;
  const g1a* bases = (const g1a*)bases_raw;
  g1j total;
  j_identity(&total);
  for (u64 i = 0; i < n; i++) {
    u64 k[4];
    scalar_repr(k, &coeffs[i]);
    g1j t;
    j_mul_u256(&t, &bases[i], k);
    j_add(&total, &total, &t);
  }
  j_to_affine((g1a*)out, &total);
}

/* sum_i a_i * b_i over Fr (Montgomery in, Montgomery out) */
void orc_inner_product(u64* out, const u64* a_raw, const u64* b_raw, u64 n) {
  const fe* a = (const fe*)a_raw;
  const fe* b = (const fe*)b_raw;
  fe acc = {{0, 0, 0, 0}};
  for (u64 i = 0; i < n; i++) {
    fe t;
    f_mul(&FR, &t, &a[i], &b[i]);
    f_add(&FR, &acc, &acc, &t);
  }
  memcpy(out, acc.v, 32);
}
