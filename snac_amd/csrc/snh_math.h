/*
 * snh_math.h -- scalar arithmetic of the explicit step, shared by the gfx950
 * kernels and host-side checks.  Every expression keeps the operand order of
 * the reference line it cites, so that with FMA contraction off the device
 * produces the reference's IEEE-754 double results.
 */
#ifndef SNH_MATH_H
#define SNH_MATH_H

#ifdef __HIPCC__
#define SNH_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define SNH_HD static inline
#endif

#include <limits.h>

/* x86-64 meaning of the reference's abs() applied to a double
 * (plastic_BI/Constitutive.c:290,409,413,483): cvttsd2si truncation, which
 * gives INT_MIN for NaN and out-of-range inputs, then abs(), which leaves
 * INT_MIN unchanged.  gfx950's own conversion saturates, so it is spelled out. */
SNH_HD int snh_ref_iabs(double x) {
    int t = (x > -2147483649.0 && x < 2147483648.0) ? (int)x : INT_MIN;
    if (t == INT_MIN) return INT_MIN;
    return t < 0 ? -t : t;
}

/* int + int of two abs() results, wrapping as the x86 add does */
SNH_HD int snh_iadd_wrap(int a, int b) { return (int)((unsigned)a + (unsigned)b); }

/* hypot(x, y) correctly rounded in practice: sqrt of the fused sum of
 * squares, then one residual correction computed exactly with fused
 * multiply-adds (C. F. Borges, "An improved algorithm for hypot(a,b)",
 * arXiv:1904.09481, fused variant).  Power-of-two scaling keeps the squares
 * in range. */
SNH_HD double snh_hypot(double x, double y) {
    double ax = fabs(x), ay = fabs(y);
    if (isinf(ax) || isinf(ay)) return INFINITY;
    if (isnan(ax) || isnan(ay)) return ax + ay;
    if (ax < ay) {
        double t = ax;
        ax = ay;
        ay = t;
    }
    if (ay == 0.0) return ax;
    /* (ay/ax)^2/2 below 2^-61: the sum of squares rounds to ax^2 */
    if (ay <= ax * 0x1p-30) return ax;
    double scale = 1.0;
    if (ax > 0x1p+500) {
        ax *= 0x1p-600;
        ay *= 0x1p-600;
        scale = 0x1p+600;
    } else if (ay < 0x1p-500) {
        ax *= 0x1p+600;
        ay *= 0x1p+600;
        scale = 0x1p-600;
    }
    double h = sqrt(fma(ax, ax, ay * ay));
    double h_sq = h * h;
    double ax_sq = ax * ax;
    double e = fma(-ay, ay, h_sq - ax_sq) + fma(h, h, -h_sq) - fma(ax, ax, -ax_sq);
    h = h - e / (2.0 * h);
    return h * scale;
}

#endif
