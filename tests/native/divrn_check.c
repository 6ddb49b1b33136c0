/* The FP64 triangular solves divide by a diagonal entry d through its
 * correctly rounded reciprocal (hpgl_amd/csrc/divrn.h): q = RN(a y),
 * r = a - d q (one fused multiply-add, exact), RN(q + r y).  This program
 * restates that sequence on the host and compares it with the division
 * a / d bit for bit on random dividends and divisors over 2^-40 .. 2^40,
 * divisors with all-ones low significand bits, and quotients next to
 * powers of two.  Prints "ok <n>" or the first mismatches. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t state = 0x9e3779b97f4a7c15ull;
static uint64_t next(void) {
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return state;
}
static double from_bits(uint64_t b) {
    double d;
    memcpy(&d, &b, 8);
    return d;
}
static uint64_t to_bits(double d) {
    uint64_t b;
    memcpy(&b, &d, 8);
    return b;
}
static double random_double(int emin, int emax) {
    const uint64_t m = next() & ((1ull << 52) - 1);
    const int e = emin + (int)(next() % (uint64_t)(emax - emin + 1));
    return from_bits(((uint64_t)(e + 1023) << 52) | m);
}
static double div_rn(double a, double d, double dinv) {
    const double q = a * dinv;
    const double r = fma(-q, d, a);
    const double c = fma(r, dinv, q);
    const double m = fabs(c);
    if (m > 0x1p-960 && m < 0x1p960) return c;
    if (a == 0.0) return a;
    return a / d;
}

int main(int argc, char **argv) {
    const long n = argc > 1 ? atol(argv[1]) : 20000000;
    long bad = 0;
    for (long i = 0; i < n; ++i) {
        double a = random_double(-40, 40);
        if (next() & 1) a = -a;
        double d = random_double(-40, 40);
        const int kind = (int)(i % 8);
        if (kind == 1) d = from_bits(to_bits(d) | (((1ull << 52) - 1) >> (next() % 12)));
        if (kind == 2) a = d * from_bits(to_bits(1.0) + (next() % 5) - 2); /* quotients next to 1 */
        if (kind == 3) a = ldexp(d, (int)(next() % 20) - 10);             /* exact powers of two */
        if (kind == 4) a = (next() & 1) ? 0.0 : -0.0;
        const double dinv = 1.0 / d;
        const double got = div_rn(a, d, dinv), want = a / d;
        if (to_bits(got) != to_bits(want)) {
            if (bad < 5) printf("mismatch a=%a d=%a got %a want %a\n", a, d, got, want);
            ++bad;
        }
    }
    if (bad) {
        printf("bad %ld of %ld\n", bad, n);
        return 1;
    }
    printf("ok %ld\n", n);
    return 0;
}
