/* Checks myokit_amd/csrc/mk_math.h against the C library: prints the
 * largest error in ulps per function over random arguments, and the number
 * of wrong special cases. Built and run by tests/test_oracle.py. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "mk_math.h"

static double ulps(double a, double b)
{
    double u;
    if (a == b || (isnan(a) && isnan(b))) return 0;
    u = nextafter(fabs(b), INFINITY) - fabs(b);
    return fabs(a - b) / u;
}

static double uni(void) { return rand() / (double)RAND_MAX; }

int main(void)
{
    double e = 0, l = 0, pi = 0, p = 0, s = 0, c = 0, a = 0, sp = 0;
    int i;
    srand(7);
    for (i = 0; i < 500000; ++i) {
        double x = (uni() * 2 - 1) * 740, y, t;
        double r = ulps(mk_exp(x), exp(x));
        if (r > e) e = r;
        y = exp((uni() * 2 - 1) * 700);
        r = ulps(mk_log(y), log(y));
        if (r > l) l = r;
        y = uni() * 4;
        r = ulps(mk_pow(y, 3.0), pow(y, 3.0));
        if (r > pi) pi = r;
        t = (uni() * 2 - 1) * 8;
        r = ulps(mk_pow(y, t), pow(y, t));
        if (r > p) p = r;
        t = (uni() * 2 - 1) * 100;
        if (fabs(sin(t)) > 1e-3) { r = ulps(mk_sin(t), sin(t)); if (r > s) s = r; }
        if (fabs(cos(t)) > 1e-3) { r = ulps(mk_cos(t), cos(t)); if (r > c) c = r; }
        r = ulps(mk_atan(t), atan(t));
        if (r > a) a = r;
    }
    /* special values */
    sp += !(mk_exp(1000) == INFINITY);
    sp += !(mk_exp(-1000) == 0);
    sp += !isnan(mk_exp(NAN));
    sp += !(mk_log(0) == -INFINITY);
    sp += !isnan(mk_log(-1));
    sp += !(mk_log(INFINITY) == INFINITY);
    sp += !(mk_pow(-2, 3) == -8);
    sp += !isnan(mk_pow(-2, 0.5));
    sp += !(mk_pow(0, 2) == 0);
    sp += !(mk_pow(5, 0) == 1);
    sp += !(mk_exp(0) == 1);
    sp += !(mk_log(1) == 0);
    printf("exp %g\nlog %g\npow_int %g\npow %g\nsin %g\ncos %g\natan %g\n"
           "special %g\n", e, l, pi, p, s, c, a, sp);
    return 0;
}
