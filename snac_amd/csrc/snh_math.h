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

/* hypot(x, y) as the reference's libm computes it.  The reference calls the
 * system hypot in tql2 (plastic_BI/Constitutive.c:430,462); on x86-64 with
 * glibc 2.35 that is the unfused form of C. F. Borges' corrected square root
 * ("An improved algorithm for hypot(a,b)", arXiv:1904.09481, MyHypot3):
 * h = sqrt(ax^2 + ay^2) rounded, then one correction h -= (t1 + t2) / 2h
 * with the error terms of the branch h <= 2 ay or h > 2 ay, evaluated in
 * plain double arithmetic (no FMA).  Operands below 2^-459 or above 2^511
 * are scaled by 2^600; a term below 2^-54 of the other is added directly.
 * The result is faithful, not always correctly rounded (about 2 in 10^3
 * results are one ulp from the correctly rounded value), and this restatement
 * gives the system hypot's bits: checked on the host against the libm
 * function.  Requires FMA contraction off. */
SNH_HD double snh_hypot_kernel(double ax, double ay) {
    double h = sqrt(ax * ax + ay * ay);
    double t1, t2;
    if (h <= 2.0 * ay) {
        const double delta = h - ay;
        t1 = ax * (2.0 * delta - ax);
        t2 = (delta - 2.0 * (ax - ay)) * delta;
    } else {
        const double delta = h - ax;
        t1 = 2.0 * delta * (ax - 2.0 * ay);
        t2 = (4.0 * delta - ay) * ay + delta * delta;
    }
    h -= (t1 + t2) / (2.0 * h);
    return h;
}

SNH_HD double snh_hypot(double x, double y) {
    if (!isfinite(x) || !isfinite(y)) {
        if (isinf(x) || isinf(y)) return INFINITY;
        return x + y;
    }
    x = fabs(x);
    y = fabs(y);
    const double ax = x < y ? y : x, ay = x < y ? x : y;
    if (ax > 0x1p+511) {
        if (ay <= ax * 0x1p-54) return ax + ay;
        return snh_hypot_kernel(ax * 0x1p-600, ay * 0x1p-600) / 0x1p-600;
    }
    if (ay < 0x1p-459) {
        if (ax >= ay / 0x1p-54) return ax + ay;
        return snh_hypot_kernel(ax / 0x1p-600, ay / 0x1p-600) * 0x1p-600;
    }
    if (ay <= ax * 0x1p-54) return ax + ay;
    return snh_hypot_kernel(ax, ay);
}

/* RN(x / 3) without a division.  x = M 2^e with integer M, and ulp(x/3) <=
 * ulp(x), so x/3 measured in ulps of the result has fractional part 0, 1/3 or
 * 2/3: the exact quotient is at least 1/6 ulp from every rounding midpoint.
 * q = RN(x * RN(1/3)) is within 1.5 ulp; r = x - 3q is then a multiple of
 * ulp(q) of magnitude <= 3 ulp(q), so the FMA computes it exactly; and
 * q + r*RN(1/3) differs from x/3 by r*(RN(1/3) - 1/3) < 2^-53 ulp, which
 * cannot move it across a midpoint.  Hence one rounding gives RN(x/3), for all
 * normal x (zero keeps its magnitude; -0 may come out as +0). */
SNH_HD double snh_div3(double x) {
    const double y = 0x1.5555555555555p-2; /* RN(1/3) */
    double q = x * y;
    double r = fma(-3.0, q, x);
    return fma(r, y, q);
}

/* RN(x / 6) = RN(x / 3) / 2 exactly (halving commutes with rounding in the
 * normal range) */
SNH_HD double snh_div6(double x) { return snh_div3(x) * 0.5; }

/* RN(a / b) from y = RN(1 / b) (P. Markstein, "Computation of elementary
 * functions on the IBM RISC System/6000", IBM J. Res. Dev. 34 (1990); and
 * Muller et al., Handbook of Floating-Point Arithmetic, 2nd ed., sec. 4.7):
 * q0 = RN(a y) is within 1.5 ulp; one residual step makes it faithful; with
 * a faithful q and the correctly rounded reciprocal, r = a - b q is exact and
 * RN(q + r y) is the correctly rounded quotient.  Normal range only. */
SNH_HD double snh_div_by(double a, double b, double y) {
    double q = a * y;
    double r = fma(-b, q, a);
    q = fma(r, y, q);
    r = fma(-b, q, a);
    return fma(r, y, q);
}

#endif
