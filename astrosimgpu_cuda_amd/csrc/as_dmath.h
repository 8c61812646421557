/* as_dmath.h -- deterministic double-precision exp, log and cos.
 *
 * The simulator's FP64 path must give the same bits on the host and on a
 * gfx950 GPU. Vendor math libraries (glibc on the host, the ROCm device
 * library on the GPU) each round their results a little differently, and in
 * a spiking network a one-ulp difference in one exp() eventually moves a
 * spike. This header therefore supplies the three transcendentals the
 * per-step arithmetic needs, written only with IEEE-754 +, -, *, /, fma (each
 * correctly rounded on both sides), exact bit manipulation and floor(), so
 * that any build with floating-point contraction disabled evaluates them to
 * identical bits on every target.
 *
 * Besides + - * /, the methods use fused multiply-add, which IEEE-754 also
 * rounds exactly once on every conforming target (the device's v_fma_f64,
 * the host's FMA instruction or C library fma), so results stay identical.
 *
 * Methods (textbook, from the mathematics):
 *   exp : Cody-Waite reduction x = k*ln2 + r, |r| <= ln2/2, with ln2 split
 *         into a 32-bit head and a tail (the head step exact inside one
 *         fused multiply-add); exp(r) by its Taylor series to degree 13
 *         (truncation < 2^-57); 2^k applied with one rounding.
 *   log : x = 2^k * m with m in [sqrt(2)/2, sqrt(2)], f = m - 1,
 *         s = f/(2+f); log(m) = 2*atanh(s) = 2s + s*R(s^2) with the atanh
 *         series to s^23 (truncation < 2^-60), evaluated in the
 *         f - (f^2/2 - s*(f^2/2 + R)) arrangement that keeps the leading
 *         term exact.
 *   cos : Cody-Waite reduction by pi/2 with a three-part constant (33+33+53
 *         bits), then the Taylor series of cos or sin on |r| <= pi/4
 *         (truncation < 2^-57). Full accuracy for |x| < 2^20; the simulator
 *         only calls it on [0, 2*pi).
 * The split constants were computed with 80-digit decimal arithmetic.
 * Accuracy against glibc is checked by tests/test_dmath.py (max error a
 * small number of ulps); equality of host and device bits by the GPU tests.
 *
 * The header is plain C99 as well as HIP C++, so the host oracle and the
 * device kernels include the very same definitions.
 */
#ifndef AS_DMATH_H
#define AS_DMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define AS_DM_FN static inline __host__ __device__
#else
#define AS_DM_FN static inline
#endif

AS_DM_FN uint64_t as_dm_bits(double x) {
    uint64_t u;
    __builtin_memcpy(&u, &x, sizeof u);
    return u;
}

AS_DM_FN double as_dm_double(uint64_t u) {
    double x;
    __builtin_memcpy(&x, &u, sizeof x);
    return x;
}

/* 2^k as a double for -1022 <= k <= 1023. */
AS_DM_FN double as_dm_pow2(int k) {
    return as_dm_double((uint64_t)(k + 1023) << 52);
}

#define AS_DM_LN2_HI 0x1.62e42fee00000p-1  /* ln2 truncated to 32 bits */
#define AS_DM_LN2_LO 0x1.a39ef35793c76p-33 /* ln2 - LN2_HI, rounded */
#define AS_DM_INV_LN2 0x1.71547652b82fep+0
#define AS_DM_PIO2_1 0x1.921fb54400000p+0  /* pi/2, first 33 bits */
#define AS_DM_PIO2_2 0x1.0b4611a600000p-34 /* next 33 bits */
#define AS_DM_PIO2_3 0x1.3198a2e037073p-69 /* remainder, rounded */
#define AS_DM_TWO_OVER_PI 0x1.45f306dc9c883p-1
#define AS_DM_SQRT2 0x1.6a09e667f3bcdp+0

/* exp on [-745.2, 709.78]: the reduction, the polynomial and the scaling. */
AS_DM_FN double as_dm_exp_core(double x) {
    const double kd = __builtin_floor(x * AS_DM_INV_LN2 + 0.5); /* k = round(x / ln2) */
    const int k = (int)kd;
    /* r = x - k*ln2, |r| <= ln2/2 (+ rounding): the head step is exact */
    const double r = __builtin_fma(-kd, AS_DM_LN2_LO, __builtin_fma(-kd, AS_DM_LN2_HI, x));
    /* exp(r) = 1 + r + r^2 * q(r), q = sum_{n>=2} r^(n-2)/n!, Horner with
     * fused multiply-adds (one rounding per step on every IEEE target) */
    double q = 1.0 / 6227020800.0;                 /* 1/13! */
    q = __builtin_fma(r, q, 1.0 / 479001600.0);     /* 1/12! */
    q = __builtin_fma(r, q, 1.0 / 39916800.0);      /* 1/11! */
    q = __builtin_fma(r, q, 1.0 / 3628800.0);       /* 1/10! */
    q = __builtin_fma(r, q, 1.0 / 362880.0);        /* 1/9!  */
    q = __builtin_fma(r, q, 1.0 / 40320.0);         /* 1/8!  */
    q = __builtin_fma(r, q, 1.0 / 5040.0);          /* 1/7!  */
    q = __builtin_fma(r, q, 1.0 / 720.0);           /* 1/6!  */
    q = __builtin_fma(r, q, 1.0 / 120.0);           /* 1/5!  */
    q = __builtin_fma(r, q, 1.0 / 24.0);            /* 1/4!  */
    q = __builtin_fma(r, q, 1.0 / 6.0);             /* 1/3!  */
    q = __builtin_fma(r, q, 0.5);                   /* 1/2!  */
    const double p = 1.0 + __builtin_fma(r * r, q, r);
    /* p * 2^k rounded once. k lies in [-1075, 1024] and p in [0.7, 1.5], so
     * p * 2^(k/2) is normal and exact, and the second scaling is the one
     * rounding; the device's ldexp instruction rounds p * 2^k once as well,
     * so both give the same bits (subnormal and overflowing results
     * included). */
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_ldexp(p, k);
#else
    const int k1 = k / 2;
    return (p * as_dm_pow2(k1)) * as_dm_pow2(k - k1);
#endif
}

AS_DM_FN double as_exp(double x) {
    if (!(x >= -745.2 && x <= 709.782712893384)) { /* NaN, underflow or overflow */
        if (x != x) return x;
        return x > 0.0 ? as_dm_double(0x7ff0000000000000ULL) : 0.0;
    }
    return as_dm_exp_core(x);
}

/* as_exp(x), bit for bit, for x that is not NaN and not above 709.78: the
 * range test becomes one clamp, because below -745.2 both this and as_exp
 * give +0 (exp(-745.2) rounds to zero). */
AS_DM_FN double as_exp_upto709(double x) {
    return as_dm_exp_core(__builtin_fmax(x, -745.2));
}

AS_DM_FN double as_log(double x) {
    if (x != x) return x;
    if (x < 0.0) return as_dm_double(0x7ff8000000000000ULL);
    if (x == 0.0) return as_dm_double(0xfff0000000000000ULL);
    uint64_t u = as_dm_bits(x);
    if (u >= 0x7ff0000000000000ULL) return x; /* +inf */
    int k = 0;
    if (u < 0x0010000000000000ULL) { /* subnormal: scale by 2^54 */
        x *= 0x1p54;
        u = as_dm_bits(x);
        k = -54;
    }
    k += (int)(u >> 52) - 1023;
    double m = as_dm_double((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > AS_DM_SQRT2) {
        m *= 0.5;
        k += 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    /* R = sum_{n=1..11} 2 z^n / (2n+1) */
    double R = 2.0 / 23.0;
    R = __builtin_fma(z, R, 2.0 / 21.0);
    R = __builtin_fma(z, R, 2.0 / 19.0);
    R = __builtin_fma(z, R, 2.0 / 17.0);
    R = __builtin_fma(z, R, 2.0 / 15.0);
    R = __builtin_fma(z, R, 2.0 / 13.0);
    R = __builtin_fma(z, R, 2.0 / 11.0);
    R = __builtin_fma(z, R, 2.0 / 9.0);
    R = __builtin_fma(z, R, 2.0 / 7.0);
    R = __builtin_fma(z, R, 2.0 / 5.0);
    R = __builtin_fma(z, R, 2.0 / 3.0);
    R = z * R;
    const double hfsq = 0.5 * f * f;
    const double kd = (double)k;
    return kd * AS_DM_LN2_HI - ((hfsq - (s * (hfsq + R) + kd * AS_DM_LN2_LO)) - f);
}

/* cos(r) on |r| <= pi/4 (+ a few ulps). */
AS_DM_FN double as_dm_cos_kernel(double r) {
    const double z = r * r;
    double c = 1.0 / 20922789888000.0;           /* +1/16! */
    c = __builtin_fma(z, c, -1.0 / 87178291200.0); /* -1/14! */
    c = __builtin_fma(z, c, 1.0 / 479001600.0);    /* +1/12! */
    c = __builtin_fma(z, c, -1.0 / 3628800.0);     /* -1/10! */
    c = __builtin_fma(z, c, 1.0 / 40320.0);        /* +1/8!  */
    c = __builtin_fma(z, c, -1.0 / 720.0);         /* -1/6!  */
    c = __builtin_fma(z, c, 1.0 / 24.0);           /* +1/4!  */
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + z * (z * c));
}

/* sin(r) on |r| <= pi/4 (+ a few ulps). */
AS_DM_FN double as_dm_sin_kernel(double r) {
    const double z = r * r;
    double s = 1.0 / 355687428096000.0;          /* +1/17! */
    s = __builtin_fma(z, s, -1.0 / 1307674368000.0); /* -1/15! */
    s = __builtin_fma(z, s, 1.0 / 6227020800.0);   /* +1/13! */
    s = __builtin_fma(z, s, -1.0 / 39916800.0);    /* -1/11! */
    s = __builtin_fma(z, s, 1.0 / 362880.0);       /* +1/9!  */
    s = __builtin_fma(z, s, -1.0 / 5040.0);        /* -1/7!  */
    s = __builtin_fma(z, s, 1.0 / 120.0);          /* +1/5!  */
    s = __builtin_fma(z, s, -1.0 / 6.0);           /* -1/3!  */
    return __builtin_fma(r * z, s, r);
}

AS_DM_FN double as_cos(double x) {
    if (x != x) return x;
    const double ax = x < 0.0 ? -x : x;
    if (!(ax < 0x1p30)) return as_dm_double(0x7ff8000000000000ULL); /* outside supported range */
    const double nd = __builtin_floor(x * AS_DM_TWO_OVER_PI + 0.5);
    const double r = ((x - nd * AS_DM_PIO2_1) - nd * AS_DM_PIO2_2) - nd * AS_DM_PIO2_3;
    const int n = ((int)nd) & 3;
    if (n == 0) return as_dm_cos_kernel(r);
    if (n == 1) return -as_dm_sin_kernel(r);
    if (n == 2) return -as_dm_cos_kernel(r);
    return as_dm_sin_kernel(r);
}

#endif /* AS_DMATH_H */
