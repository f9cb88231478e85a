// Host driver for tests/test_fq29_cpu.py: runs Fq29 / G1X29 (the MSM's bucket
// arithmetic, nine 29-bit limbs, radix 2^261) on the CPU over inputs given on
// stdin and prints the results in the 2^256 Montgomery form of bn254.hpp, for
// the test to check with Python integers.  Built with hipcc as host code.
//
// Input lines (field elements as 64 hex digits, canonical x 2^256 mod p):
//   mul A B         -> A B 2^-256 mod p through Fq29, and the lazy-sum forms
//   chain N         then N lines "X Y NEG": madd of each affine point (X, Y)
//                   (negated if NEG) into an accumulator starting at the
//                   identity; a point may also be "ACC NEG": the accumulator's
//                   own value (equal / opposite point cases)
//                   -> the XYZZ result, then the general sum of the result with
//                   itself, its double, and the result after a store/load
//                   round trip, each as "X Y ZZ ZZZ" and an invariant flag
#include <cstdio>
#include <cstring>
#include <string>

#include "../zkpoly-compiler_amd/csrc/curve/g1.hpp"
#include "../zkpoly-compiler_amd/csrc/curve/xyzz29.hpp"

using namespace zkp;

static bool read_hex(const char* s, u32 (&w)[8]) {
  if (strlen(s) != 64) return false;
  for (int k = 0; k < 8; k++) {
    unsigned v = 0;
    if (sscanf(s + 8 * (7 - k), "%8x", &v) != 1) return false;
    w[k] = v;
  }
  return true;
}

static void print_words(const u32 (&w)[8]) {
  for (int k = 7; k >= 0; k--) printf("%08x", w[k]);
}

static void print29(const Fq29& a) {
  u32 w[8];
  a.to_mont256(w);
  print_words(w);
}

// X < 8p, Y < 4p, ZZ, ZZZ < 2p, limbs below 2^29
static bool below(const Fq29& v, int k) {
  for (int i = 0; i < 8; i++)
    if (v.l[i] > Fq29::MASK) return false;
  int64_t c = 0;
  for (int i = 0; i < 9; i++) {
    c += (int64_t)v.l[i] - (int64_t)k * Fq29Params::P[i];
    if (i < 8) c >>= 29;
  }
  return c < 0;
}
static bool invariant(const G1X29& p) { return below(p.x, 8) && below(p.y, 4) && below(p.zz, 2) && below(p.zzz, 2); }

static void print_point(const G1X29& p) {
  print29(p.x);
  printf(" ");
  print29(p.y);
  printf(" ");
  print29(p.zz);
  printf(" ");
  print29(p.zzz);
  printf(" %d\n", invariant(p) ? 1 : 0);
}

static Affine29 affine(const u32 (&x)[8], const u32 (&y)[8], bool neg) {
  Affine29 a;
  u32 any = 0;
  for (int i = 0; i < 8; i++) any |= x[i] | y[i];
  a.inf = any == 0;
  a.x = Fq29::from_words<5>(x);
  a.y = Fq29::from_words<5>(y);
  if (neg) a.y = Fq29::zero().sub<32>(a.y);
  return a;
}

// the accumulator as an affine point of the 2^256 form (x = X / ZZ, y = Y / ZZZ) for the ACC input
static void acc_affine(const G1X29& p, u32 (&x)[8], u32 (&y)[8]) {
  u32 w[4][8];
  p.x.to_mont256(w[0]);
  p.y.to_mont256(w[1]);
  p.zz.to_mont256(w[2]);
  p.zzz.to_mont256(w[3]);
  const G1XYZZ q{Fq::from_raw(w[0]), Fq::from_raw(w[1]), Fq::from_raw(w[2]), Fq::from_raw(w[3])};
  const G1Affine a = q.to_affine();
  for (int i = 0; i < 8; i++) {
    x[i] = a.x.v[i];
    y[i] = a.y.v[i];
  }
}

int main() {
  char line[512];
  while (fgets(line, sizeof line, stdin)) {
    char op[16], s1[80], s2[80];
    if (sscanf(line, "%15s", op) != 1) continue;
    if (!strcmp(op, "mul")) {
      u32 a[8], b[8];
      if (sscanf(line, "%*s %79s %79s", s1, s2) != 2 || !read_hex(s1, a) || !read_hex(s2, b)) return 2;
      const Fq29 ra = Fq29::from_words<5>(a), rb = Fq29::from_words<5>(b);  // below 32p
      const Fq29 x = ra.reduce_2p(), y = rb.reduce_2p();                     // below 2p
      print29(x * y);                                                          // a b
      printf(" ");
      print29(ra * y);  // a raw operand (below 32p) times one below 2p
      printf(" ");
      const Fq29 s4 = ((x + x) + (y + y)).norm();  // 2a + 2b, below 8p
      print29(s4 * s4);
      printf(" ");
      print29(x.sub<4>(y + y));  // a - 2b
      printf(" ");
      print29(Fq29::zero().sub<32>(ra) * y);  // -a b
      printf(" %d %d\n", x.sub<2>(x).is_zero_modp() ? 1 : 0, (x * y).is_zero_modp() ? 1 : 0);
    } else if (!strcmp(op, "chain")) {
      int n = 0;
      if (sscanf(line, "%*s %d", &n) != 1) return 2;
      G1X29 acc = G1X29::identity();
      bool ok = true;
      for (int t = 0; t < n; t++) {
        if (!fgets(line, sizeof line, stdin)) return 2;
        int neg = 0;
        u32 x[8], y[8];
        if (!strncmp(line, "ACC", 3)) {
          if (sscanf(line, "%*s %d", &neg) != 1) return 2;
          if (acc.is_identity()) continue;
          acc_affine(acc, x, y);
        } else if (sscanf(line, "%79s %79s %d", s1, s2, &neg) != 3 || !read_hex(s1, x) || !read_hex(s2, y)) {
          return 2;
        }
        acc = acc.add_affine(affine(x, y, neg != 0));
        ok = ok && invariant(acc);
      }
      print_point(acc);
      print_point(acc.add(acc));
      print_point(acc.dbl());
      // through the packed HBM layout (X brought below 2p first)
      u32 w[4][8];
      acc.x.reduce_2p().to_words(w[0]);
      acc.y.to_words(w[1]);
      acc.zz.to_words(w[2]);
      acc.zzz.to_words(w[3]);
      const G1X29 back{Fq29::from_words<0>(w[0]), Fq29::from_words<0>(w[1]), Fq29::from_words<0>(w[2]),
                       Fq29::from_words<0>(w[3])};
      print_point(back.add(acc.dbl()));
      printf("%d\n", ok ? 1 : 0);
    } else {
      return 2;
    }
    fflush(stdout);
  }
  return 0;
}
