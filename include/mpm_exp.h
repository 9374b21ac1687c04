/*
 * mpm_exp.h -- a portable, bit-reproducible double-precision exp().
 *
 * The USL volume update v <- v * exp(dt * (exx + eyy)) (reference:
 * libmpm/process_usl.c:961-970) calls libm's exp.  glibc on the host and the
 * device math library may round that call differently, and a one-ulp change
 * of a particle volume can flip the density cut-off of the stress update.
 * This function is written with IEEE +, -, *, / only, in a fixed order, so
 * that any compiler that does not contract multiplies into FMAs (both the
 * HIP kernels and the CPU parity build are compiled with
 * -ffp-contract=off) returns the same bits on the CPU and on the GPU.
 *
 * Method: Cody-Waite reduction x = k*ln2 + r with |r| <= ln2/2, a Taylor
 * series for expm1(r) summed in Horner form, and a compensated final
 * addition 1 + r + q (Fast2Sum of 1 + r, then the low part added to q) so
 * that the result is correctly rounded except when exp(x) lies within
 * ~1e-6 ulp of a rounding boundary.  Arguments with |x| <= 2^-8 (every
 * volumetric increment of a stable explicit step) take a short series.
 * Scaling by 2^k is done by building the power of two from its bits; results
 * in the subnormal range are scaled in two exact-then-rounding steps.
 */
#ifndef MPM_EXP_H
#define MPM_EXP_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define MPM_HD __host__ __device__
#else
#define MPM_HD
#endif

static inline MPM_HD double mpm_pow2i(int k)
{
    /* 2^k for -1022 <= k <= 1023 */
    uint64_t bits = (uint64_t)(k + 1023) << 52;
    double r;
    memcpy(&r, &bits, sizeof r);
    return r;
}

static inline MPM_HD double mpm_exp(double x)
{
    const double inv_ln2 = 1.4426950408889634074;      /* 1/ln2 */
    const double ln2_hi = 6.93147180369123816490e-01;  /* 0x3fe62e42fee00000 */
    const double ln2_lo = 1.90821492927058770002e-10;  /* 0x3dea39ef35793c76 */
    const double shifter = 6755399441055744.0;         /* 1.5 * 2^52 */

    if (x != x) {
        return x + x;                   /* NaN */
    }
    if (x > 709.782712893384) {
        return 1.0 / 0.0 * (x > 0.0);   /* +inf (also for x = +inf) */
    }
    if (x < -745.1332191019412) {
        return 0.0;
    }

    double r;
    double rl = 0.0;                    /* low part of the reduced argument */
    int k;
    double q;
    if (x <= 0.00390625 && x >= -0.00390625) {
        /* |x| <= 2^-8: q = expm1(x) - x to ~2^-60 relative */
        k = 0;
        r = x;
        q = r * r * (0.5 + r * (1.0 / 6.0 + r * (1.0 / 24.0 + r * (1.0 / 120.0
            + r * (1.0 / 720.0 + r * (1.0 / 5040.0))))));
    } else {
        double kd = (x * inv_ln2 + shifter) - shifter;  /* round to nearest */
        k = (int)kd;
        /* k*ln2_hi is exact (ln2_hi has 11 trailing zero bits, |k| <= 1075)
         * and so is x - k*ln2_hi (Sterbenz); rl keeps the rounding error of
         * the second subtraction. */
        double rh = x - kd * ln2_hi;
        double kl = kd * ln2_lo;
        r = rh - kl;
        rl = (rh - r) - kl;
        /* |r| <= 0.3466: Taylor to degree 14, truncation < 2^-66 */
        q = r * r * (0.5 + r * (1.0 / 6.0 + r * (1.0 / 24.0 + r * (1.0 / 120.0
            + r * (1.0 / 720.0 + r * (1.0 / 5040.0 + r * (1.0 / 40320.0
            + r * (1.0 / 362880.0 + r * (1.0 / 3628800.0 + r * (1.0 / 39916800.0
            + r * (1.0 / 479001600.0 + r * (1.0 / 6227020800.0
            + r * (1.0 / 87178291200.0)))))))))))));
    }

    /* 1 + r + q (+ rl * exp(r)) with the rounding error of 1 + r recovered
     * exactly */
    double s = 1.0 + r;
    double e = (1.0 - s) + r;           /* exact: |1| >= |r| */
    double y = s + (e + (q + (rl + rl * r)));

    if (k == 0) {
        return y;
    }
    if (k > 1023) {
        return (y * mpm_pow2i(1023)) * mpm_pow2i(k - 1023);
    }
    if (k < -1021) {
        /* land in the normal range exactly, then round once */
        return (y * mpm_pow2i(k + 108)) * mpm_pow2i(-108);
    }
    return y * mpm_pow2i(k);
}

#endif /* MPM_EXP_H */
