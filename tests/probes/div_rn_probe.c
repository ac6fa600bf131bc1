/* Test probe (built by tests/test_div_rn_cpu.py with gcc): div_rn of
 * disco_hd.h -- one reciprocal per divisor, then q = a y, r = fma(-b, q, a),
 * q + r y -- against the correctly rounded quotient a / b, on random pairs
 * spread over 120 binades and on quotients close to a rounding boundary
 * (a = b k (1 +- 2^-52) for integers k).  Prints the number of pairs and of
 * mismatches. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "disco_hd.h"

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint64_t next(void) {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
}
static double rnd(int emin, int emax) {
    uint64_t m = next() & ((1ull << 52) - 1);
    int e = emin + (int)(next() % (uint64_t)(emax - emin + 1));
    uint64_t bits = ((uint64_t)(e + 1023) << 52) | m;
    if (next() & 1) bits |= 1ull << 63;
    double d;
    memcpy(&d, &bits, 8);
    return d;
}

int main(int argc, char** argv) {
    long n = argc > 1 ? atol(argv[1]) : 1000000;
    long bad = 0;
    for (long i = 0; i < n; ++i) {
        double b = rnd(-60, 60), a;
        if (i % 3 == 0)
            a = b * (double)(next() % 4096 + 1) * (1.0 + ((double)(next() % 3) - 1.0) * 0x1p-52);
        else
            a = rnd(-60, 60);
        volatile double y = 1.0 / b;
        double q = div_rn(a, b, y), ex = a / b;
        if (memcmp(&q, &ex, 8) != 0) {
            if (bad < 5) printf("mismatch a=%a b=%a: %a vs %a\n", a, b, q, ex);
            ++bad;
        }
    }
    printf("pairs %ld mismatches %ld\n", n, bad);
    return bad != 0;
}
