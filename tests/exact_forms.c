/* Bitwise checks, on the host's IEEE-754 double arithmetic (round to nearest even, fused
 * multiply-add correctly rounded), of two instruction forms the stage kernel uses in place
 * of the reference's expressions (shallow-water_amd/csrc/swe_face.h):
 *
 *  1. div_numerator: the last step of the split division with the remainder taken negated,
 *         fma(-fma(d, q, -n), r, q),  q = n * r,
 *     against the form with the copysign of the numerator,
 *         copysign(fma(fma(-d, q, n), r, q), n),
 *     for any reciprocal estimate r > 0 within a few ulps of 1/d (the kernel's r is
 *     rcp + two Newton steps), numerators of both signs including +-0 and exact quotients.
 *  2. half_minmod_split: fma(0.5, max(min(a, b), min(max(a, b), 0)), +0) against
 *     0.5 * minmod(a, b) of omp_kernels.hpp:45-48 for slope operands that are differences
 *     of values satisfying the kernel's range guard (0 or magnitude in [2^-240, 2^190]).
 *
 * Usage: exact_forms <cases>   prints "div <mismatches> median <mismatches>".
 */
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
static double unit(void) { return 1.0 + (double)(next() >> 11) / 9007199254740992.0; }
static uint64_t bits(double x) {
    uint64_t u;
    memcpy(&u, &x, sizeof u);
    return u;
}

/* a value satisfying the range guard: 0, -0, or a magnitude in [2^-240, 2^190] */
static double guarded(void) {
    const uint64_t b = next();
    switch (b % 8) {
        case 0: return 0.0;
        case 1: return -0.0;
        case 2: return ldexp((b & 256) ? -1.0 : 1.0, (int)((b >> 9) % 8) - 2); /* ties */
        default: {
            const double x = ldexp(unit(), (int)((b >> 9) % 430) - 240);
            return (b & 256) ? -x : x;
        }
    }
}

static double minmod_ref(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    const double m = fmin(fabs(a), fabs(b));
    return (a > 0.0 ? 1.0 : -1.0) * m;
}

int main(int argc, char **argv) {
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    long bad_div = 0, bad_med = 0;
    for (long i = 0; i < n; ++i) {
        /* 1. division */
        const double d = ldexp(unit(), (int)(next() % 400) - 200);
        double num;
        switch (i % 8) {
            case 0: num = 0.0; break;
            case 1: num = -0.0; break;
            case 2: num = d * ldexp(1.0, (int)(next() % 40) - 20); break; /* exact quotient */
            case 3: num = -d; break;
            default: num = ldexp(unit(), (int)(next() % 400) - 200) * ((next() & 1) ? -1 : 1);
        }
        double r = 1.0 / d;
        const int p = (int)(next() % 7) - 3;
        for (int k = 0; k < abs(p); ++k) r = nextafter(r, p > 0 ? INFINITY : 0.0);
        const double q = num * r;
        const double with_copysign = copysign(fma(fma(-d, q, num), r, q), num);
        const double negated = fma(-fma(d, q, -num), r, q);
        if (bits(with_copysign) != bits(negated)) ++bad_div;

        /* 2. median slope on differences of guarded values */
        const double x0 = guarded(), x1 = guarded(), x2 = guarded();
        double a = x1 - x0, b = x2 - x1;
        if (i % 5 == 0) {
            a = x0;
            b = x1;
        }
        const double ref = 0.5 * minmod_ref(a, b);
        const double med = fma(0.5, fmax(fmin(a, b), fmin(fmax(a, b), 0.0)), 0.0);
        if (bits(ref) != bits(med)) ++bad_med;
    }
    printf("div %ld median %ld\n", bad_div, bad_med);
    return (bad_div || bad_med) ? 1 : 0;
}
