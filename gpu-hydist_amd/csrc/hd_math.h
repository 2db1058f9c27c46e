/*
 * hd_math.h -- double-precision elementary functions shared by the HIP kernels
 * and the host code (case-constant evaluation, CPU checker), so that the GPU
 * and the CPU produce the same bits for every transcendental the ADI step uses.
 *
 * The reference kernels call the CUDA device math library (pow, log, cos, sqrt;
 * e.g. support_funcs.cu:54, Sediment_Transport.cu:91, support_funcs.cu:370).
 * Those results cannot be reproduced bit-for-bit outside that library, so this
 * project defines its own functions from textbook methods:
 *   log  : x = 2^k m, m in [sqrt(2)/2, sqrt(2)); log(m) = 2 atanh(s) with
 *          s = f/(2+f), f = m-1, atanh summed as its Taylor series in s^2
 *          (coefficients 2/(2n+1)), k ln2 added with a two-part ln2
 *          (Cody & Waite, "Software Manual for the Elementary Functions", 1980).
 *   exp  : x = k ln2 + r, |r| <= ln2/2 (Cody-Waite two-part ln2), exp(r) by its
 *          Taylor series, then scaled by 2^k.
 *   pow  : exp(y log x) with log carried as a double-double (hi + lo) and the
 *          product y*log formed exactly with fma (Dekker's two-product).
 *   sin/cos : x = k pi/2 + r with a three-part pi/2 (Cody-Waite), Taylor series
 *          on |r| <= pi/4, quadrant selection.
 * sqrt and division are the IEEE correctly rounded operations of both targets.
 *
 * Every function uses only +, -, *, /, fma, comparisons and bit casts, so with
 * floating-point contraction disabled (-ffp-contract=off on both compilers) the
 * host and the device evaluate exactly the same operation sequence.
 */
#ifndef HD_MATH_H
#define HD_MATH_H

#include <stdint.h>
#include <math.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define HD_FN __host__ __device__ static inline
#else
#define HD_FN static inline
#endif

HD_FN uint64_t hd_bits(double x) { uint64_t b; __builtin_memcpy(&b, &x, 8); return b; }
HD_FN double hd_from_bits(uint64_t b) { double x; __builtin_memcpy(&x, &b, 8); return x; }

/* 2^k for k in the normal exponent range [-1022, 1023]. */
HD_FN double hd_exp2i(int k) { return hd_from_bits((uint64_t)(k + 1023) << 52); }

/* x * 2^k without intermediate overflow/underflow for |k| <= 2098. */
HD_FN double hd_ldexp(double x, int k)
{
    if (k > 1023) { x *= hd_exp2i(1023); k -= 1023; if (k > 1023) { x *= hd_exp2i(1023); k -= 1023; if (k > 1023) k = 1023; } }
    else if (k < -1022) { x *= hd_exp2i(-1022 + 53); k += 1022 - 53; if (k < -1022) { x *= hd_exp2i(-1022 + 53); k += 1022 - 53; if (k < -1022) k = -1022; } }
    return x * hd_exp2i(k);
}

/* ln 2 split so that k * HD_LN2_HI is exact for |k| < 2^20. */
#define HD_LN2_HI 0x1.62e42ff000000p-1
#define HD_LN2_LO -0x1.718432a1b0e26p-35
#define HD_LOG2E 0x1.71547652b82fep+0
#define HD_SQRT2 0x1.6a09e667f3bcdp+0

/* log(1+f) for f = m - 1, m in [sqrt(2)/2, sqrt(2)); returns f - s*(f - R). */
HD_FN double hd_log1p_reduced(double f)
{
    double s = f / (2.0 + f);
    double z = s * s;
    /* R = sum_{n>=1} 2 z^n / (2n+1), Horner form, 12 terms (z <= 0.0295). */
    double r = 2.0 / 25.0;
    r = 2.0 / 23.0 + z * r;
    r = 2.0 / 21.0 + z * r;
    r = 2.0 / 19.0 + z * r;
    r = 2.0 / 17.0 + z * r;
    r = 2.0 / 15.0 + z * r;
    r = 2.0 / 13.0 + z * r;
    r = 2.0 / 11.0 + z * r;
    r = 2.0 / 9.0 + z * r;
    r = 2.0 / 7.0 + z * r;
    r = 2.0 / 5.0 + z * r;
    r = 2.0 / 3.0 + z * r;
    r = z * r;
    /* 2 atanh(s) = 2s + s R and 2s = f - s f. */
    return f - s * (f - r);
}

/* Split a positive finite x into 2^k * m with m in [sqrt(2)/2, sqrt(2)). */
HD_FN double hd_reduce_log_arg(double x, int* kout)
{
    int k = 0;
    uint64_t b = hd_bits(x);
    if ((b >> 52) == 0) { /* subnormal */
        x *= 0x1p54;
        k = -54;
        b = hd_bits(x);
    }
    k += (int)(b >> 52) - 1023;
    double m = hd_from_bits((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > HD_SQRT2) { m *= 0.5; k += 1; }
    *kout = k;
    return m;
}

HD_FN double hd_log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    int k;
    double m = hd_reduce_log_arg(x, &k);
    double lp = hd_log1p_reduced(m - 1.0);
    double dk = (double)k;
    return dk * HD_LN2_HI + (dk * HD_LN2_LO + lp);
}

/* exp(x) * (1 + xlo): exp of the double-double x + xlo, |xlo| << ulp(x). */
HD_FN double hd_exp_dd(double x, double xlo)
{
    if (x != x) return x;
    if (x > 709.782712893384) return __builtin_inf();
    if (x < -745.2) return 0.0;
    double kd = x * HD_LOG2E;
    kd = (kd >= 0.0) ? (double)(int64_t)(kd + 0.5) : -(double)(int64_t)(0.5 - kd);
    int k = (int)kd;
    double r = (x - kd * HD_LN2_HI) - kd * HD_LN2_LO;
    /* exp(r) - 1 - r = r^2 (1/2 + r/6 + ...), Taylor to degree 14 (|r| <= 0.347). */
    double p = 1.0 / 87178291200.0;          /* 1/14! */
    p = 1.0 / 6227020800.0 + r * p;          /* 1/13! */
    p = 1.0 / 479001600.0 + r * p;           /* 1/12! */
    p = 1.0 / 39916800.0 + r * p;            /* 1/11! */
    p = 1.0 / 3628800.0 + r * p;             /* 1/10! */
    p = 1.0 / 362880.0 + r * p;              /* 1/9! */
    p = 1.0 / 40320.0 + r * p;               /* 1/8! */
    p = 1.0 / 5040.0 + r * p;                /* 1/7! */
    p = 1.0 / 720.0 + r * p;                 /* 1/6! */
    p = 1.0 / 120.0 + r * p;                 /* 1/5! */
    p = 1.0 / 24.0 + r * p;                  /* 1/4! */
    p = 1.0 / 6.0 + r * p;                   /* 1/3! */
    p = 0.5 + r * p;                         /* 1/2! */
    double em1 = r + (r * r) * p;            /* exp(r) - 1 */
    double e = 1.0 + (em1 + (1.0 + em1) * xlo);
    return hd_ldexp(e, k);
}

HD_FN double hd_exp(double x) { return hd_exp_dd(x, 0.0); }

/* pow(x, y) for the domains the ADI/sediment formulas use (x >= 0, or x < 0
 * with integral y).  Special values follow C99 Annex F. */
HD_FN double hd_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (x == 1.0) return 1.0;
    if (x != x || y != y) return x + y;
    if (y == 1.0) return x;
    if (y == 2.0) return x * x;
    double ay = y < 0.0 ? -y : y;
    int yint = 0, yodd = 0;
    if (ay < 0x1p53) {
        double yi = (double)(int64_t)ay;
        if (yi == ay) { yint = 1; yodd = ((int64_t)ay) & 1; }
    } else {
        yint = 1;
    }
    double sign = 1.0;
    if (x < 0.0) {
        if (!yint) return __builtin_nan("");
        if (yodd) sign = -1.0;
        x = -x;
    }
    if (x == 0.0) return (y > 0.0) ? sign * 0.0 : sign * __builtin_inf();
    if (x == __builtin_inf()) return (y > 0.0) ? sign * __builtin_inf() : sign * 0.0;
    if (ay == __builtin_inf()) {
        if (x < 1.0) return (y > 0.0) ? 0.0 : __builtin_inf();
        return (y > 0.0) ? __builtin_inf() : 0.0;
    }
    int k;
    double m = hd_reduce_log_arg(x, &k);
    double lp = hd_log1p_reduced(m - 1.0);
    double dk = (double)k;
    /* log x as hi + lo */
    double a = dk * HD_LN2_HI;               /* exact */
    double lhi = a + lp;
    double llo = ((a - lhi) + lp) + dk * HD_LN2_LO;
    /* y * log x as phi + plo (two-product with fma) */
    double phi = y * lhi;
    double plo = fma(y, lhi, -phi) + y * llo;
    return sign * hd_exp_dd(phi, plo);
}

/* pi/2 split into three 33-bit parts: k*P1, k*P2 exact for |k| < 2^20. */
#define HD_PIO2_1 0x1.921fb54400000p+0
#define HD_PIO2_2 0x1.0b4611a600000p-34
#define HD_PIO2_3 0x1.3198a2e037073p-69
#define HD_2OPI 0x1.45f306dc9c883p-1

HD_FN double hd_sin_poly(double r)
{
    /* sin r = r + r z (-1/3! + z/5! - z^2/7! + ... + z^9/21!), z = r^2 */
    double z = r * r;
    double p = 1.0 / 51090942171709440000.0;   /* 1/21! */
    p = -1.0 / 121645100408832000.0 + z * p;   /* -1/19! */
    p = 1.0 / 355687428096000.0 + z * p;       /* 1/17! */
    p = -1.0 / 1307674368000.0 + z * p;        /* -1/15! */
    p = 1.0 / 6227020800.0 + z * p;            /* 1/13! */
    p = -1.0 / 39916800.0 + z * p;             /* -1/11! */
    p = 1.0 / 362880.0 + z * p;                /* 1/9! */
    p = -1.0 / 5040.0 + z * p;                 /* -1/7! */
    p = 1.0 / 120.0 + z * p;                   /* 1/5! */
    p = -1.0 / 6.0 + z * p;                    /* -1/3! */
    return r + (r * z) * p;
}

HD_FN double hd_cos_poly(double r)
{
    double z = r * r;
    double p = 1.0 / 2432902008176640000.0;    /* 1/20! */
    p = -1.0 / 6402373705728000.0 + z * p;     /* -1/18! */
    p = 1.0 / 20922789888000.0 + z * p;        /* 1/16! */
    p = -1.0 / 87178291200.0 + z * p;          /* -1/14! */
    p = 1.0 / 479001600.0 + z * p;             /* 1/12! */
    p = -1.0 / 3628800.0 + z * p;              /* -1/10! */
    p = 1.0 / 40320.0 + z * p;                 /* 1/8! */
    p = -1.0 / 720.0 + z * p;                  /* -1/6! */
    p = 1.0 / 24.0 + z * p;                    /* 1/4! */
    double hz = 0.5 * z;
    double w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + (z * z) * p);
}

/* Reduce x to r in [-pi/4, pi/4] and quadrant q; valid for |x| < 2^19 * pi. */
HD_FN double hd_trig_reduce(double x, int* q)
{
    double kd = x * HD_2OPI;
    kd = (kd >= 0.0) ? (double)(int64_t)(kd + 0.5) : -(double)(int64_t)(0.5 - kd);
    *q = (int)(((int64_t)kd) & 3);
    return ((x - kd * HD_PIO2_1) - kd * HD_PIO2_2) - kd * HD_PIO2_3;
}

HD_FN double hd_cos(double x)
{
    if (x != x || x == __builtin_inf() || x == -__builtin_inf()) return x - x;
    int q;
    double r = hd_trig_reduce(x, &q);
    switch (q) {
    case 0: return hd_cos_poly(r);
    case 1: return -hd_sin_poly(r);
    case 2: return -hd_cos_poly(r);
    default: return hd_sin_poly(r);
    }
}

HD_FN double hd_sin(double x)
{
    if (x != x || x == __builtin_inf() || x == -__builtin_inf()) return x - x;
    int q;
    double r = hd_trig_reduce(x, &q);
    switch (q) {
    case 0: return hd_sin_poly(r);
    case 1: return hd_cos_poly(r);
    case 2: return -hd_sin_poly(r);
    default: return -hd_cos_poly(r);
    }
}

/* round-half-to-even conversion (the reference's __double2int_rn,
 * support_funcs.cu:202-203) for |x| < 2^31. */
HD_FN int hd_rint_int(double x)
{
    double t = (x >= 0.0) ? x + 0x1p52 - 0x1p52 : x - 0x1p52 + 0x1p52;
    return (int)t;
}

#endif /* HD_MATH_H */
