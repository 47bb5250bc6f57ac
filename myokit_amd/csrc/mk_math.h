/*
 * mk_math.h -- portable double-precision elementary functions.
 *
 * One implementation of exp, log, pow, sin, cos, tan, asin, acos, atan and
 * log10, written only with IEEE-754 +, -, *, /, fused multiply-add (all
 * correctly rounded on every conforming CPU and GPU: v_fma_f64 on gfx950,
 * the FMA3 instruction or the C library's fma on the CPU), floor, rint,
 * fabs and sqrt, and compiled with floating-point contraction disabled, so
 * that a fused multiply-add appears exactly where one is written. Because
 * every operation is correctly rounded and the operation order is fixed,
 * the HIP kernel (gfx950, built by hiprtc) and the CPU oracle (gcc) produce
 * bit-identical results. That is how the double-precision engine can match
 * its CPU oracle to the last bit.
 *
 * Methods (textbook, no vendor library constants):
 *   exp  : Cody & Waite (1980) range reduction x = k*ln2 + r, with ln2 split
 *          into a 32-bit head and a tail so that k*LN2_HI is exact, then the
 *          Taylor polynomial of exp(r) to degree 13 by Horner's rule in
 *          fused multiply-adds (|r| <= ln2/2, truncation error < 2^-57),
 *          then scaling by 2^k with one rounding (exact unless the result
 *          is subnormal).
 *   log  : x = m * 2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
 *          log(m) = 2*atanh(s) = 2s(1 + s^2/3 + s^4/5 + ... + s^22/23).
 *   pow  : integer exponents |y| <= 64 by binary powering, otherwise
 *          exp(y * log(x)) with the C99 special cases.
 *   sin/cos: Cody & Waite reduction by pi/2 (three-part split), Taylor
 *          polynomials on [-pi/4, pi/4].
 *   atan : reflection for |x| > 1, two argument halvings
 *          atan(t) = 2 atan(t / (1 + sqrt(1 + t^2))), Taylor series.
 * Split constants were derived with 200-bit arithmetic from ln2 and pi;
 * Taylor coefficients are the doubles nearest 1/k!.
 *
 * Accuracy (checked against the C library in the tests): exp and log within
 * 1 ulp; pow within a few ulp for moderate |y log x|; the trigonometric
 * functions within a few ulp for |x| < 1e5.
 */
#ifndef MK_MATH_H
#define MK_MATH_H

#if defined(__HIPCC_RTC__) || defined(MK_MATH_DEVICE)
  /* Device build (hiprtc, or hipcc with MK_MATH_DEVICE defined, as in
   * exported kernels): no system headers. */
  typedef unsigned long long mk_u64;
  #define MK_FN static __device__ __forceinline__
  #define MK_BITS(d) __builtin_bit_cast(mk_u64, (d))
  #define MK_DBL(u) __builtin_bit_cast(double, (mk_u64)(u))
#else
  /* Host build (the CPU oracle). */
  #include <stdint.h>
  #include <string.h>
  typedef uint64_t mk_u64;
  #define MK_FN static inline
  static inline mk_u64 mk_bits_(double d) { mk_u64 u; memcpy(&u, &d, 8); return u; }
  static inline double mk_dbl_(mk_u64 u) { double d; memcpy(&d, &u, 8); return d; }
  #define MK_BITS(d) mk_bits_(d)
  #define MK_DBL(u) mk_dbl_((mk_u64)(u))
#endif

#define MK_INF __builtin_inf()
#define MK_NAN __builtin_nan("")
#define MK_FLOOR(x) __builtin_floor(x)
#define MK_RINT(x) __builtin_rint(x)
#define MK_FMA(a, b, c) __builtin_fma((a), (b), (c))
/* min(max(x, lo), hi) with NaN to lo: v_max_f64 / v_min_f64 on the GPU,
 * compares on the CPU (where fmax and fmin are library calls) */
#if defined(__HIPCC_RTC__) || defined(MK_MATH_DEVICE)
  #define MK_CLAMP(x, lo, hi) __builtin_fmin(__builtin_fmax((x), (lo)), (hi))
#else
  static inline double mk_clamp_(double x, double lo, double hi)
  {
      x = x >= lo ? x : lo;
      return x <= hi ? x : hi;
  }
  #define MK_CLAMP(x, lo, hi) mk_clamp_((x), (lo), (hi))
#endif
#define MK_FABS(x) __builtin_fabs(x)
#define MK_SQRT(x) __builtin_sqrt(x)

#define MK_LN2_HI 0x1.62e42fee00000p-1
#define MK_LN2_LO 0x1.a39ef35793c76p-33
#define MK_LOG2E 0x1.71547652b82fep+0
#define MK_SQRT2 0x1.6a09e667f3bcdp+0
#define MK_LN10 0x1.26bb1bbb55516p+1
#define MK_PIO2 0x1.921fb54442d18p+0
#define MK_PI 0x1.921fb54442d18p+1
#define MK_TWO_OVER_PI 0x1.45f306dc9c883p-1
#define MK_PIO2_1 0x1.921fb54400000p+0
#define MK_PIO2_2 0x1.0b4611a600000p-34
#define MK_PIO2_3 0x1.3198a2e037073p-69

/* 2^n as a double, for -1022 <= n <= 1023 (exact). */
MK_FN double mk_pow2i(int n)
{
    return MK_DBL(((mk_u64)(n + 1023)) << 52);
}

/* p * 2^k rounded once (-1076 <= k <= 1024, p in [0.5, 2)): IEEE scaleB,
 * v_ldexp_f64 on the GPU; on the CPU (where ldexp is a library call) as
 * 2^k1 * 2^k2 with k1 + k2 = k, both normal powers, so that only the
 * second product rounds, into the subnormals or to inf as scaleB does. */
#if defined(__HIPCC_RTC__) || defined(MK_MATH_DEVICE)
  #define MK_SCALE2(p, k) __builtin_ldexp((p), (k))
#else
  static inline double mk_scale2_(double p, int k)
  {
      const int k1 = k >> 1;
      return (p * mk_pow2i(k1)) * mk_pow2i(k - k1);
  }
  #define MK_SCALE2(p, k) mk_scale2_((p), (k))
#endif

MK_FN double mk_exp(double x)
{
    /* One straight path for every input (no early returns, which a GPU
     * runs as divergent branches in each of a cell's many exps). x is
     * clamped to [-746, 710] (NaN to -746), where 2^k stays within
     * 2^-1076 .. 2^1024: below -745.1332 (-1075 ln 2) the result rounds
     * to 0 and above 709.7827 (ln of the largest double) it overflows to
     * inf by itself, as exp's do (2^k applied with one rounding,
     * MK_SCALE2). A NaN argument is returned at the end. */
    double xc, k, hi, r, p, y;
    xc = MK_CLAMP(x, -746.0, 710.0);
    k = MK_RINT(xc * MK_LOG2E);
    hi = MK_FMA(-k, MK_LN2_HI, xc);     /* exact: k * LN2_HI is exact */
    r = MK_FMA(-k, MK_LN2_LO, hi);
    p = 0x1.6124613a86d09p-33;                  /* 1/13! */
    p = MK_FMA(p, r, 0x1.1eed8eff8d898p-29);    /* 1/12! */
    p = MK_FMA(p, r, 0x1.ae64567f544e4p-26);    /* 1/11! */
    p = MK_FMA(p, r, 0x1.27e4fb7789f5cp-22);    /* 1/10! */
    p = MK_FMA(p, r, 0x1.71de3a556c734p-19);    /* 1/9!  */
    p = MK_FMA(p, r, 0x1.a01a01a01a01ap-16);    /* 1/8!  */
    p = MK_FMA(p, r, 0x1.a01a01a01a01ap-13);    /* 1/7!  */
    p = MK_FMA(p, r, 0x1.6c16c16c16c17p-10);    /* 1/6!  */
    p = MK_FMA(p, r, 0x1.1111111111111p-7);     /* 1/5!  */
    p = MK_FMA(p, r, 0x1.5555555555555p-5);     /* 1/4!  */
    p = MK_FMA(p, r, 0x1.5555555555555p-3);     /* 1/3!  */
    p = MK_FMA(p, r, 0x1.0000000000000p-1);     /* 1/2!  */
    p = MK_FMA(p, r, 1.0);
    p = MK_FMA(p, r, 1.0);
    y = MK_SCALE2(p, (int)k);
    return x != x ? x : y;
}

MK_FN double mk_log(double x)
{
    mk_u64 b;
    int e = 0;
    double m, f, s, z, R, s2, de;
    if (x != x) return x;
    if (x < 0.0) return MK_NAN;
    if (x == 0.0) return -MK_INF;
    if (x == MK_INF) return x;
    if (x < 0x1p-1022) { x = x * 0x1p54; e = -54; }
    b = MK_BITS(x);
    e += (int)((b >> 52) & 0x7ff) - 1023;
    m = MK_DBL((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > MK_SQRT2) { m = m * 0.5; e += 1; }
    /* With f = m - 1 (exact) and s = f / (2 + f):
     *   log(1 + f) = 2s + s*R,  R = 2s^2/3 + 2s^4/5 + ... + 2s^22/23,
     * and since 2s = f - hfsq + s*hfsq (hfsq = f^2/2) this is
     *   log(1 + f) = f - (hfsq - s*(hfsq + R)),
     * which keeps the large term f exact. */
    f = m - 1.0;
    s = f / (2.0 + f);
    z = s * s;
    R = 0x1.642c8590b2164p-4;           /* 2/23 */
    R = R * z + 0x1.8618618618618p-4;   /* 2/21 */
    R = R * z + 0x1.af286bca1af28p-4;   /* 2/19 */
    R = R * z + 0x1.e1e1e1e1e1e1ep-4;   /* 2/17 */
    R = R * z + 0x1.1111111111111p-3;   /* 2/15 */
    R = R * z + 0x1.3b13b13b13b14p-3;   /* 2/13 */
    R = R * z + 0x1.745d1745d1746p-3;   /* 2/11 */
    R = R * z + 0x1.c71c71c71c71cp-3;   /* 2/9  */
    R = R * z + 0x1.2492492492492p-2;   /* 2/7  */
    R = R * z + 0x1.999999999999ap-2;   /* 2/5  */
    R = R * z + 0x1.5555555555555p-1;   /* 2/3  */
    R = R * z;
    s2 = 0.5 * f * f;                   /* hfsq */
    de = (double)e;
    return de * MK_LN2_HI
        - ((s2 - (s * (s2 + R) + de * MK_LN2_LO)) - f);
}

MK_FN double mk_log10(double x)
{
    return mk_log(x) / MK_LN10;
}

/* x^n for integer n by binary powering (fixed operation order). */
MK_FN double mk_powi(double x, int n)
{
    double r = 1.0, b = x;
    int neg = n < 0;
    unsigned int u = (unsigned int)(neg ? -n : n);
    while (u) {
        if (u & 1u) r = r * b;
        u >>= 1;
        if (u) b = b * b;
    }
    return neg ? 1.0 / r : r;
}

MK_FN double mk_pow(double x, double y)
{
    double ay, l;
    int yint, yodd;
    if (y == 0.0) return 1.0;
    if (x == 1.0) return 1.0;
    if (x != x || y != y) return x + y;
    ay = MK_FABS(y);
    yint = (MK_FLOOR(y) == y);
    if (yint && ay <= 64.0) return mk_powi(x, (int)y);
    yodd = yint && ay < 0x1p53 && (MK_FLOOR(y * 0.5) * 2.0 != y);
    if (x == 0.0) {
        if (y > 0.0) return yodd ? x : 0.0;
        return yodd ? 1.0 / x : MK_INF;
    }
    if (x < 0.0) {
        if (!yint) return MK_NAN;
        l = mk_exp(y * mk_log(-x));
        return yodd ? -l : l;
    }
    return mk_exp(y * mk_log(x));
}

/* sin(r) and cos(r) for |r| <= pi/4 */
MK_FN double mk_sin_kernel(double r)
{
    double z = r * r, p;
    p = 0x1.952c77030ad4ap-49;
    p = p * z - 0x1.ae7f3e733b81fp-41;
    p = p * z + 0x1.6124613a86d09p-33;
    p = p * z - 0x1.ae64567f544e4p-26;
    p = p * z + 0x1.71de3a556c734p-19;
    p = p * z - 0x1.a01a01a01a01ap-13;
    p = p * z + 0x1.1111111111111p-7;
    p = p * z - 0x1.5555555555555p-3;
    return r + (r * z) * p;
}

MK_FN double mk_cos_kernel(double r)
{
    double z = r * r, p;
    p = -0x1.6827863b97d97p-53;
    p = p * z + 0x1.ae7f3e733b81fp-45;
    p = p * z - 0x1.93974a8c07c9dp-37;
    p = p * z + 0x1.1eed8eff8d898p-29;
    p = p * z - 0x1.27e4fb7789f5cp-22;
    p = p * z + 0x1.a01a01a01a01ap-16;
    p = p * z - 0x1.6c16c16c16c17p-10;
    p = p * z + 0x1.5555555555555p-5;
    p = p * z - 0x1.0000000000000p-1;
    return 1.0 + z * p;
}

/* Reduces x to r in [-pi/4, pi/4]; returns the quadrant (0..3). */
MK_FN int mk_reduce_pio2(double x, double* r)
{
    double n = MK_FLOOR(x * MK_TWO_OVER_PI + 0.5);
    double q;
    *r = ((x - n * MK_PIO2_1) - n * MK_PIO2_2) - n * MK_PIO2_3;
    q = n - 4.0 * MK_FLOOR(n * 0.25);
    return (int)q;
}

MK_FN double mk_sin(double x)
{
    double r;
    int q;
    if (x != x || MK_FABS(x) == MK_INF) return MK_NAN;
    q = mk_reduce_pio2(x, &r);
    if (q == 0) return mk_sin_kernel(r);
    if (q == 1) return mk_cos_kernel(r);
    if (q == 2) return -mk_sin_kernel(r);
    return -mk_cos_kernel(r);
}

MK_FN double mk_cos(double x)
{
    double r;
    int q;
    if (x != x || MK_FABS(x) == MK_INF) return MK_NAN;
    q = mk_reduce_pio2(x, &r);
    if (q == 0) return mk_cos_kernel(r);
    if (q == 1) return -mk_sin_kernel(r);
    if (q == 2) return -mk_cos_kernel(r);
    return mk_sin_kernel(r);
}

MK_FN double mk_tan(double x)
{
    return mk_sin(x) / mk_cos(x);
}

MK_FN double mk_atan(double x)
{
    double ax, t, z, p, res;
    int inv;
    if (x != x) return x;
    ax = MK_FABS(x);
    if (ax == MK_INF) return x > 0 ? MK_PIO2 : -MK_PIO2;
    inv = ax > 1.0;
    t = inv ? 1.0 / ax : ax;
    t = t / (1.0 + MK_SQRT(1.0 + t * t));
    t = t / (1.0 + MK_SQRT(1.0 + t * t));
    z = t * t;
    /* atan(t) = t - t^3/3 + t^5/5 - ... - t^35/35, |t| <= 0.2 */
    p = -1.0 / 35.0;
    p = p * z + 1.0 / 33.0;
    p = p * z - 1.0 / 31.0;
    p = p * z + 1.0 / 29.0;
    p = p * z - 1.0 / 27.0;
    p = p * z + 1.0 / 25.0;
    p = p * z - 1.0 / 23.0;
    p = p * z + 1.0 / 21.0;
    p = p * z - 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z - 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z - 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z - 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z - 1.0 / 3.0;
    res = 4.0 * (t + (t * z) * p);
    if (inv) res = MK_PIO2 - res;
    return x < 0 ? -res : res;
}

MK_FN double mk_asin(double x)
{
    double ax = MK_FABS(x);
    if (x != x || ax > 1.0) return MK_NAN;
    if (ax == 1.0) return x > 0 ? MK_PIO2 : -MK_PIO2;
    return mk_atan(x / MK_SQRT((1.0 - x) * (1.0 + x)));
}

MK_FN double mk_acos(double x)
{
    if (x != x || MK_FABS(x) > 1.0) return MK_NAN;
    if (x == -1.0) return MK_PI;
    return 2.0 * mk_atan(MK_SQRT((1.0 - x) / (1.0 + x)));
}

#endif /* MK_MATH_H */
