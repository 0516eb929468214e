/*
 * The shared-reciprocal division of the device code (rdiv/rdivd in
 * fargo3d_amd/csrc/hip/f3d_device_body.h):
 *     r = RN(1/d), q0 = RN(a r), rem = a - q0 d (one FMA),
 *     q = (rem == 0) ? q0 : RN(q0 + rem r) (one FMA)
 * must equal RN(a/d).
 *
 *   recip_division_check exhaustive P0 P1
 *     every pair of significands a, d in [1, 2) of the binary formats of
 *     precision P0..P1, in exact integer arithmetic (no hardware FP): prints
 *     "P pairs mismatches q0_wrong" per precision
 *   recip_division_check binary64 N SEED
 *     N random operand pairs plus N pairs whose quotient lies next to a
 *     rounding midpoint, in binary64 with the C library's fma; prints
 *     "samples mismatches"
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
typedef struct {
  i128 m;
  int e;
} dy; /* m * 2^e */
static int P;

static int bitlen(i128 x) {
  int n = 0;
  if (x < 0) x = -x;
  while (x) {
    n++;
    x >>= 1;
  }
  return n;
}
/* N/D * 2^e rounded to P significant bits, ties to even (D > 0) */
static dy rn_rat(i128 N, i128 D, int e) {
  dy z = {0, 0};
  if (N == 0) return z;
  int s = N < 0;
  if (s) N = -N;
  int k = 0;
  while (N / D >= ((i128)1 << P)) {
    D <<= 1;
    k++;
  }
  while (N / D < ((i128)1 << (P - 1))) {
    N <<= 1;
    k--;
  }
  i128 q = N / D, rem = N - q * D;
  if (2 * rem > D || (2 * rem == D && (q & 1))) q++;
  if (q == ((i128)1 << P)) {
    q >>= 1;
    k++;
  }
  z.m = s ? -q : q;
  z.e = e + k;
  return z;
}
static dy rn(dy x) { return (x.m == 0 || bitlen(x.m) <= P) ? x : rn_rat(x.m, 1, x.e); }
static dy mul(dy a, dy b) {
  dy z = {a.m * b.m, a.e + b.e};
  return z;
}
static dy add(dy a, dy b) {
  if (a.m == 0) return b;
  if (b.m == 0) return a;
  if (a.e > b.e) {
    dy t = a;
    a = b;
    b = t;
  }
  dy z = {a.m + (b.m << (b.e - a.e)), a.e};
  return z;
}
static int same(dy a, dy b) {
  if (a.m == 0 || b.m == 0) return a.m == b.m;
  while (!(a.m & 1)) {
    a.m >>= 1;
    a.e++;
  }
  while (!(b.m & 1)) {
    b.m >>= 1;
    b.e++;
  }
  return a.m == b.m && a.e == b.e;
}

static int exhaustive(int p0, int p1) {
  for (P = p0; P <= p1; P++) {
    long pairs = 0, bad = 0, q0bad = 0;
    const i128 lo = (i128)1 << (P - 1), hi = (i128)1 << P;
    for (i128 D = lo; D < hi; D++) {
      const dy d = {D, 1 - P};
      const dy r = rn_rat(lo, D, 0); /* 1/d = 2^(P-1)/D */
      for (i128 A = lo; A < hi; A++) {
        const dy a = {A, 1 - P};
        const dy ref = rn_rat(A, D, 0);
        const dy q0 = rn(mul(a, r));
        dy qd = mul(d, q0);
        qd.m = -qd.m;
        const dy rem = rn(add(a, qd));
        const dy q = rem.m == 0 ? q0 : rn(add(q0, mul(rem, r)));
        pairs++;
        bad += !same(q, ref);
        q0bad += !same(q0, ref);
      }
    }
    printf("%d %ld %ld %ld\n", P, pairs, bad, q0bad);
  }
  return 0;
}

static uint64_t rs;
static uint64_t xr(void) {
  rs ^= rs << 13;
  rs ^= rs >> 7;
  rs ^= rs << 17;
  return rs;
}
static double rnd(int emin, int emax) {
  uint64_t m = xr() & ((1ULL << 52) - 1);
  int e = emin + (int)(xr() % (uint64_t)(emax - emin + 1));
  uint64_t bits = ((uint64_t)(e + 1023) << 52) | m;
  double x;
  memcpy(&x, &bits, 8);
  return (xr() & 1) ? -x : x;
}
static double rdiv(double a, double d) {
  const double r = 1.0 / d, q0 = a * r, rem = fma(-q0, d, a);
  return rem == 0.0 ? q0 : fma(rem, r, q0);
}
static int binary64(long n, uint64_t seed) {
  rs = seed ? seed : 88172645463325252ULL;
  long bad = 0, samples = 0;
  for (long t = 0; t < n; t++) {
    const double a = rnd(-40, 40), d = rnd(-40, 40);
    samples++;
    bad += rdiv(a, d) != a / d;
    /* a quotient next to a midpoint: a = RN(d * m), m halfway between two
       neighbours, then the neighbours of that a */
    const double q = fabs(rnd(-20, 20)), dd = fabs(rnd(-20, 20));
    const double mid = q + 0.5 * (nextafter(q, INFINITY) - q);
    double aa = dd * mid;
    for (int k = -2; k <= 2; k++) {
      double x = aa;
      for (int u = 0; u < abs(k); u++) x = nextafter(x, k > 0 ? INFINITY : -INFINITY);
      samples++;
      bad += rdiv(x, dd) != x / dd;
    }
  }
  printf("%ld %ld\n", samples, bad);
  return 0;
}

int main(int argc, char** argv) {
  if (argc == 4 && strcmp(argv[1], "exhaustive") == 0) return exhaustive(atoi(argv[2]), atoi(argv[3]));
  if (argc == 4 && strcmp(argv[1], "binary64") == 0) return binary64(atol(argv[2]), strtoull(argv[3], 0, 10));
  fprintf(stderr, "usage: %s exhaustive P0 P1 | binary64 N SEED\n", argv[0]);
  return 2;
}
