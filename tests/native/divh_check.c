/* Checks the division used by the kernels for divisors fixed on the host
 * (divh in pllc_dev.h): with y = RN(1/h), q = RN(a*y), r = a - q*h (exact by
 * FMA) and RN(q + r*y) must equal the IEEE quotient RN(a/h).  Prints the
 * number of mismatches over n samples for the grid spacings of the test and
 * benchmark grids and for random divisors. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static uint64_t st = 0x9E3779B97F4A7C15ULL;
static uint64_t next(void)
{
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
}
static double rnd(void)
{
    const uint64_t m = next() & ((1ULL << 52) - 1);
    const int e = (int)(next() % 81) - 40;
    const double v = ldexp(1.0 + (double)m / 4503599627370496.0, e);
    return (next() & 1) ? -v : v;
}
static double divh(double a, double h, double y)
{
    const double q = a * y;
    const double r = fma(-q, h, a);
    return fma(r, y, q);
}
int main(int argc, char** argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    const double hs[] = {4.0 / 24, 1.0 / 12, 1.0 / 10, 4.0 / 64, 1.0 / 32, 4.0 / 256, 1.0 / 128, 4.0 / 512,
                         1.0 / 256, 4.0 / 96, 1.0 / 24, 1.0 / 50, 4.0 / 192, 1.0 / 48, 1.0 / 64, 1.0, 3.0};
    long bad = 0;
    for (unsigned i = 0; i < sizeof(hs) / sizeof(hs[0]); i++) {
        const double h = hs[i], y = 1.0 / h;
        for (long t = 0; t < n; t++) {
            const double a = rnd();
            if (divh(a, h, y) != a / h) bad++;
            if (divh(a, 2.0 * h, 0.5 * y) != a / (2.0 * h)) bad++;
        }
    }
    for (long t = 0; t < 4 * n; t++) {
        const double h = fabs(rnd()), a = rnd();
        if (divh(a, h, 1.0 / h) != a / h) bad++;
    }
    printf("%ld\n", bad);
    return 0;
}
