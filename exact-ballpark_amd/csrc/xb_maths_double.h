/* SPDX-License-Identifier: LGPL-2.1-or-later
 *
 * The GNU C Library 2.34's double sin, cos, sincos (IBM Accurate Mathematical Library, s_sin.c, s_sincos.c,
 * branred.c), asin and atan2 (the same library's e_asin.c, e_atan2.c), log and exp (Arm optimized-routines, e_log.c,
 * e_exp.c) as glibc's aarch64 build computes them: every
 * step a correctly rounded IEEE operation, with a fused multiply-add exactly at the steps GCC fused there and nowhere
 * else. The algorithms and constants are glibc's; the reference restates them at src/exact_ballpark/dlibm.py:429-691
 * (sin, cos, sincos, log, branred) and src/exact_ballpark/vec_exp.py:137-183 (exp).
 *
 * One source for the GPU and the host: each function is `static inline XB_HD`, where the including file defines XB_HD
 * (`__device__` in the HIP kernels, nothing in host builds) and XB_TABLE (`__constant__` on the GPU). Files that
 * include this must be compiled without floating-point contraction (-ffp-contract=off) and without fast-math, so that
 * only the explicit xb_fma calls below fuse.
 */
#ifndef XB_MATHS_DOUBLE_H
#define XB_MATHS_DOUBLE_H

#include <stdint.h>

#ifndef XB_HD
#define XB_HD
#endif

#include "xb_glibc_tables.h"

static inline XB_HD uint64_t xb_bits(double x) {
    union { double d; uint64_t u; } v;
    v.d = x;
    return v.u;
}

static inline XB_HD double xb_dbl(uint64_t u) {
    union { double d; uint64_t u; } v;
    v.u = u;
    return v.d;
}

#define XB_BITS_DOUBLE(u) xb_dbl(u)

static inline XB_HD double xb_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
/* aarch64's fmsub (c - a b) and fnmsub (a b - c), each rounded once */
static inline XB_HD double xb_fmsub(double a, double b, double c) { return __builtin_fma(-a, b, c); }
static inline XB_HD double xb_fnmsub(double a, double b, double c) { return __builtin_fma(a, b, -c); }
static inline XB_HD double xb_fabs(double x) { return xb_dbl(xb_bits(x) & 0x7fffffffffffffffULL); }
static inline XB_HD double xb_copysign(double m, double s) {
    return xb_dbl((xb_bits(m) & 0x7fffffffffffffffULL) | (xb_bits(s) & 0x8000000000000000ULL));
}
static inline XB_HD double xb_nan(void) { return xb_dbl(0x7ff8000000000000ULL); }
static inline XB_HD double xb_inf(void) { return xb_dbl(0x7ff0000000000000ULL); }

/* ------------------------------------------------------------------------------------------- sin and cos */

/* TAYLOR_SIN: a + ((POLY(a a) a - da/2) a a + da), the polynomial's Horner steps and the last two steps fused */
static inline XB_HD double xb_taylor_sin(double a, double da) {
    double xx = a * a;
    double poly = xb_fma(xx, XB_DL_S5, XB_DL_S4);
    poly = xb_fma(xx, poly, XB_DL_S3);
    poly = xb_fma(xx, poly, XB_DL_S2);
    poly = xb_fma(xx, poly, XB_DL_S1);
    double t = xb_fnmsub(poly, a, da * 0.5);
    return xb_fma(xx, t, da) + a;
}

/* The table cell of |x| (u = big + |x|, the cell its low word): the rest |x| - (u - big) and the cell's sine, its
   tail, cosine and tail. */
static inline XB_HD double xb_sc_cell(double x, double* sn, double* ssn, double* cs, double* ccs) {
    double ax = xb_fabs(x);
    double u = XB_DL_BIG + ax;
    uint32_t cell = (uint32_t)(xb_bits(u) & 0xffffffffULL) << 2;
    *sn = XB_DL_SINCOSTAB[cell];
    *ssn = XB_DL_SINCOSTAB[cell + 1];
    *cs = XB_DL_SINCOSTAB[cell + 2];
    *ccs = XB_DL_SINCOSTAB[cell + 3];
    return ax - (u - XB_DL_BIG);
}

static inline XB_HD double xb_do_sin(double x, double dx) {
    if (xb_fabs(x) < XB_DL_TAYLOR_LIMIT) return xb_taylor_sin(x, dx);
    if (x <= 0.0) dx = -dx;
    double sn, ssn, cs, ccs;
    double r = xb_sc_cell(x, &sn, &ssn, &cs, &ccs);
    double xx = r * r;
    double pc = xb_fma(xx, XB_DL_CS6, XB_DL_CS4);
    double ps = xb_fma(xx, XB_DL_SN5, XB_DL_SN3);
    double r3 = r * xx;
    double c1 = xb_fma(xx, pc, XB_DL_CS2);
    double s = r + xb_fma(r3, ps, dx);
    double c = xb_fma(r, dx, xx * c1);
    double cor = xb_fma(s, cs, xb_fmsub(c, sn, xb_fma(s, ccs, ssn)));
    return xb_copysign(sn + cor, x);
}

static inline XB_HD double xb_do_cos(double x, double dx) {
    if (x < 0.0) dx = -dx;
    double sn, ssn, cs, ccs;
    double r = xb_sc_cell(x, &sn, &ssn, &cs, &ccs) + dx;
    double xx = r * r;
    double pc = xb_fma(xx, XB_DL_CS6, XB_DL_CS4);
    double ps = xb_fma(xx, XB_DL_SN5, XB_DL_SN3);
    double r3 = r * xx;
    double c1 = xb_fma(xx, pc, XB_DL_CS2);
    double s = xb_fma(r3, ps, r);
    double c = xx * c1;
    double cor = xb_fmsub(s, sn, xb_fmsub(c, cs, xb_fmsub(s, ssn, ccs)));
    return cs + cor;
}

/* reduce_sincos: x = n pi/2 + a + da for |x| below 105414350, fused as compiled; returns n & 3 */
static inline XB_HD int xb_reduce_sincos(double x, double* a, double* da) {
    double t = xb_fma(x, XB_DL_HPINV, XB_DL_TOINT);
    double xn = t - XB_DL_TOINT;
    double y = xb_fmsub(xn, XB_DL_MP2, xb_fmsub(xn, XB_DL_MP1, x));
    double t2 = xb_fmsub(xn, XB_DL_PP3, y);
    double b = xb_fmsub(xn, XB_DL_PP4, t2);
    double db = xb_fmsub(xn, XB_DL_PP3, y - t2);
    db = db + xb_fmsub(xn, XB_DL_PP4, t2 - b);
    *a = b;
    *da = db;
    return (int)(xb_bits(t) & 3ULL);
}

/* One half of __branred: xh times 2/pi from six 24-bit pieces of toverp, the integer part split off. branred.c is
   compiled without contraction: nothing here is fused. */
static inline XB_HD void xb_branred_half(double xh, double* whole, double* b_out, double* bb_out) {
    int e = (int)((xb_bits(xh) >> 52) & 0x7ffULL);
    int k = 0;
    double first = XB_DL_BR_TOVERP0;
    if (e > XB_DL_BR_EXP_LIMIT) {
        k = (e - XB_DL_BR_EXP_BIAS) / 24;                 /* C's division, toward zero: 0 below the bias */
        first = XB_DL_BR_TOVERP[k];
    }
    double gor = xb_dbl((uint64_t)((uint32_t)XB_DL_BR_T576_HI - ((uint32_t)(3 * k) << 23)) << 32);
    double r[6];
    for (int i = 0; i < 6; ++i) {
        r[i] = (xh * (i == 0 ? first : XB_DL_BR_TOVERP[k + i])) * gor;
        gor = gor * XB_DL_BR_TM24;
    }
    double sum = 0.0;
    for (int i = 0; i < 3; ++i) {
        double s = (r[i] + XB_DL_BR_BIG) - XB_DL_BR_BIG;
        sum = sum + s;
        r[i] = r[i] - s;
    }
    double t = 0.0;
    for (int i = 5; i >= 0; --i) t = t + r[i];
    double bb = (((((r[0] - t) + r[1]) + r[2]) + r[3]) + r[4]) + r[5];
    double s = (t + XB_DL_BR_BIG) - XB_DL_BR_BIG;
    sum = s + sum;
    t = t - s;
    s = (sum + XB_DL_BR_BIG1) - XB_DL_BR_BIG1;
    sum = sum - s;
    double b = bb + t;
    bb = (t - b) + bb;
    *whole = sum;
    *b_out = b;
    *bb_out = bb;
}

/* __branred: x = n pi/2 + a + aa for huge |x| (scaled by 2^-600 and split in two halves Dekker's way) */
static inline XB_HD int xb_branred(double x, double* a, double* aa) {
    x = x * XB_DL_BR_TM600;
    double t0 = x * XB_DL_BR_SPLIT;
    double x1 = t0 - (t0 - x);
    double x2 = x - x1;
    double sum1, b1, bb1, sum2, b2, bb2;
    xb_branred_half(x1, &sum1, &b1, &bb1);
    xb_branred_half(x2, &sum2, &b2, &bb2);
    double sum = sum1 + sum2;
    double b = b1 + b2;
    double bb = xb_fabs(b1) > xb_fabs(b2) ? (b1 - b) + b2 : (b2 - b) + b1;
    if (b > 0.5) {
        b = b - 1.0;
        sum = sum + 1.0;
    } else if (b < -0.5) {
        b = b + 1.0;
        sum = sum - 1.0;
    }
    double s = ((bb + bb1) + bb2) + b;
    double t = ((b - s) + bb) + (bb1 + bb2);
    double sp = s * XB_DL_BR_SPLIT;
    double t1 = sp - (sp - s);
    double t2 = s - t1;
    b = s * XB_DL_BR_HP0;
    double bbx = (((t1 * XB_DL_BR_MP1 - b) + t1 * XB_DL_BR_MP2) + t2 * XB_DL_BR_MP1) +
                 ((t2 * XB_DL_BR_MP2 + s * XB_DL_BR_HP1) + t * XB_DL_BR_HP0);
    s = b + bbx;
    *a = s;
    *aa = (b - s) + bbx;
    return ((int)sum) & 3;
}

/* do_sincos: the cosine branch for odd n, the sine branch for even; negated when n & 2 */
static inline XB_HD double xb_sincos_quadrant(double a, double da, int n) {
    double v = (n & 1) ? xb_do_cos(a, da) : xb_do_sin(a, da);
    return (n & 2) ? -v : v;
}

static inline XB_HD uint32_t xb_sin_key(double x) { return (uint32_t)(xb_bits(x) >> 32) & 0x7fffffffu; }

static inline XB_HD double xb_sin(double x) {
    uint32_t k = xb_sin_key(x);
    double a, da;
    int n;
    if (k < XB_DL_K_TINY_SIN) return x;
    if (k < XB_DL_K_DO_SIN) return xb_do_sin(x, 0.0);
    if (k < XB_DL_K_MID) return xb_copysign(xb_do_cos(XB_DL_HP0 - xb_fabs(x), XB_DL_HP1), x);
    if (k < XB_DL_K_REDUCE) {
        n = xb_reduce_sincos(x, &a, &da);
        return xb_sincos_quadrant(a, da, n);
    }
    if (k < XB_DL_K_HUGE) {
        n = xb_branred(x, &a, &da);
        return xb_sincos_quadrant(a, da, n);
    }
    return xb_nan();
}

static inline XB_HD double xb_cos(double x) {
    uint32_t k = xb_sin_key(x);
    double a, da;
    int n;
    if (k < XB_DL_K_TINY_COS) return 1.0;
    if (k < XB_DL_K_DO_SIN) return xb_do_cos(x, 0.0);
    if (k < XB_DL_K_MID) {
        double y = XB_DL_HP0 - xb_fabs(x);
        a = y + XB_DL_HP1;
        return xb_do_sin(a, (y - a) + XB_DL_HP1);
    }
    if (k < XB_DL_K_REDUCE) {
        n = xb_reduce_sincos(x, &a, &da);
        return xb_sincos_quadrant(a, da, n + 1);
    }
    if (k < XB_DL_K_HUGE) {
        n = xb_branred(x, &a, &da);
        return xb_sincos_quadrant(a, da, n + 1);
    }
    return xb_nan();
}

/* s_sincos.c's __sincos: __sin's and __cos's branches at the same bounds, except that between 0.855469 and 2.426265
   the sine is the cosine branch after the two-sum of hp0 - |x| and hp1 (as __cos's cosine there), where __sin takes
   it from hp0 - |x| and hp1 directly; the two differ at x = +-0x1.ad1fb54442d18p+0. */
static inline XB_HD void xb_sincos(double x, double* s, double* c) {
    uint32_t k = xb_sin_key(x);
    double a, da;
    int n;
    if (k < XB_DL_K_TINY_COS) {
        *s = x;
        *c = 1.0;
        return;
    }
    if (k < XB_DL_K_DO_SIN) {
        *s = xb_do_sin(x, 0.0);
        *c = xb_do_cos(x, 0.0);
        return;
    }
    if (k < XB_DL_K_MID) {
        double y = XB_DL_HP0 - xb_fabs(x);
        a = y + XB_DL_HP1;
        da = (y - a) + XB_DL_HP1;
        *s = xb_copysign(xb_do_cos(a, da), x);
        *c = xb_do_sin(a, da);
        return;
    }
    if (k < XB_DL_K_REDUCE) {
        n = xb_reduce_sincos(x, &a, &da);
        *s = xb_sincos_quadrant(a, da, n);
        *c = xb_sincos_quadrant(a, da, n + 1);
        return;
    }
    if (k < XB_DL_K_HUGE) {
        n = xb_branred(x, &a, &da);
        *s = xb_sincos_quadrant(a, da, n);
        *c = xb_sincos_quadrant(a, da, n + 1);
        return;
    }
    *s = xb_nan();
    *c = xb_nan();
}

/* ------------------------------------------------------------------------------------------------------ log */

/* e_log.c's __log: near 1 (1 - 2^-4 <= x < 1 + 0x1.09p-4) a degree-11 polynomial with r split at 2^27; elsewhere
   x = 2^k z, r = z/c - 1 by one fused step from a 128-cell table and a degree-5 polynomial. The fused steps and their
   order are glibc's aarch64 build's. */
static inline XB_HD double xb_log(double x) {
    uint64_t ix = xb_bits(x);
    if (ix + XB_DL_LOG_NEG_LO <= XB_DL_LOG_RANGE) {
        if (ix == 0x3ff0000000000000ULL) return 0.0;
        double r = x - 1.0;
        double rhi = xb_fma(r, XB_DL_LOG_TWO27, r);
        double r2 = r * r;
        double p78 = xb_fma(XB_DL_LOG_B[8], r, XB_DL_LOG_B[7]);
        rhi = xb_fmsub(r, XB_DL_LOG_TWO27, rhi);
        double r3 = r * r2;
        double p789 = xb_fma(XB_DL_LOG_B[9], r2, p78);
        double p45 = xb_fma(XB_DL_LOG_B[5], r, XB_DL_LOG_B[4]);
        double p710 = xb_fma(XB_DL_LOG_B[10], r3, p789);
        double rhi2 = rhi * rhi;
        double p456 = xb_fma(XB_DL_LOG_B[6], r2, p45);
        double b0 = XB_DL_LOG_B[0];
        double hi = xb_fma(rhi2, b0, r);
        double p12 = xb_fma(XB_DL_LOG_B[2], r, XB_DL_LOG_B[1]);
        double p410 = xb_fma(p710, r3, p456);
        double rlo = r - rhi;
        double rr = r + rhi;
        double p123 = xb_fma(XB_DL_LOG_B[3], r2, p12);
        double b0rlo = b0 * rlo;
        double lo = r - hi;
        lo = xb_fma(rhi2, b0, lo);
        double p = xb_fma(p410, r3, p123);
        lo = xb_fma(b0rlo, rr, lo);
        double y = xb_fma(p, r3, lo);
        return hi + y;
    }
    uint32_t top = (uint32_t)(ix >> 48);
    if (top - 0x0010u >= 0x7fe0u) {
        if ((ix << 1) == 0) return -xb_inf();
        if (ix == 0x7ff0000000000000ULL) return x;
        if ((top & 0x8000u) || (top & 0x7ff0u) == 0x7ff0u) return xb_nan();
        ix = xb_bits(x * XB_DL_LOG_TWO52) + XB_DL_LOG_NEG_52;      /* a subnormal normalised by 2^52 */
    }
    uint64_t tmp = ix + XB_DL_LOG_NEG_OFF;
    int i = (int)((tmp >> 45) & 127ULL);
    int k = (int)((int64_t)tmp >> 52);
    uint64_t iz = ix - (tmp & 0xfff0000000000000ULL);
    double kd = (double)k;
    double invc = XB_DL_LOG_TAB[2 * i], logc = XB_DL_LOG_TAB[2 * i + 1];
    double w = xb_fma(XB_DL_LOG_LN2HI, kd, logc);
    double r = xb_fma(xb_dbl(iz), invc, -1.0);
    double hi = r + w;
    double p34 = xb_fma(XB_DL_LOG_A[4], r, XB_DL_LOG_A[3]);
    double p12 = xb_fma(XB_DL_LOG_A[2], r, XB_DL_LOG_A[1]);
    double r2 = r * r;
    double lo = w - hi;
    double r3 = r * r2;
    lo = lo + r;
    double p = xb_fma(p34, r2, p12);
    lo = xb_fma(XB_DL_LOG_LN2LO, kd, lo);
    lo = xb_fma(XB_DL_LOG_A[0], r2, lo);
    lo = xb_fma(r3, p, lo);
    return lo + hi;
}

/* ------------------------------------------------------------------------------------------------------ exp */

/* round-half-away-from-zero to an integer value (aarch64's frinta), for |z| < 2^52 */
static inline XB_HD double xb_round_away(double z) { return __builtin_round(z); }

/* e_exp.c's __exp: k = round(x 128/ln2), r = x - k ln2/128 in two fused steps, a degree-5 polynomial around the table
   of 2^(i/128); the special cases for |x| below 2^-54, at or above 512 and at or above 1024. Seven fused steps in the
   main path; the subnormal-result case computes scale x tmp unfused. */
static inline XB_HD double xb_exp(double x) {
    uint64_t ix = xb_bits(x);
    uint32_t abstop = (uint32_t)((ix >> 52) & 0x7ffULL);
    if (abstop < 0x3c9u) return x + 1.0;
    if (abstop > 0x408u) {
        if (ix == 0xfff0000000000000ULL) return 0.0;
        if (abstop == 0x7ffu) return x + 1.0;
        return (ix >> 63) ? 0.0 : xb_inf();
    }
    double z = x * XB_EX_INVLN2N;
    double kd = xb_round_away(z);
    int64_t ki = (int64_t)kd;
    double r = xb_fma(kd, XB_EX_NEGLN2HIN, x);
    r = xb_fma(kd, XB_EX_NEGLN2LON, r);
    int idx = (int)((ki & 0x7f) * 2);
    double tail = xb_dbl(XB_EXP_TAB[idx]);
    uint64_t sbits = XB_EXP_TAB[idx + 1] + ((uint64_t)ki << 45);
    double p23 = xb_fma(r, XB_EX_C3, XB_EX_C2);
    double p45 = xb_fma(r, XB_EX_C5, XB_EX_C4);
    double r2 = r * r;
    double tmp = r + tail;
    double r4 = r2 * r2;
    tmp = xb_fma(p23, r2, tmp);
    tmp = xb_fma(r4, p45, tmp);
    if (abstop != 0x408u) {
        double scale = xb_dbl(sbits);
        return xb_fma(tmp, scale, scale);
    }
    if ((ki & 0x80000000LL) == 0) {
        double sc_hi = xb_dbl(sbits - (1009ULL << 52));
        return xb_fma(tmp, sc_hi, sc_hi) * XB_EX_TWO_1009;
    }
    double sc_lo = xb_dbl(sbits + (1022ULL << 52));
    double st = tmp * sc_lo;
    double y = sc_lo + st;
    if (!(y < 1.0)) return y * XB_EX_TWO_M1022;
    double lo = (sc_lo - y) + st;
    double hi = y + 1.0;
    lo = ((1.0 - hi) + y) + lo;
    double y2 = (lo + hi) - 1.0;
    if (y2 == 0.0) y2 = 0.0;
    return y2 * XB_EX_TWO_M1022;
}

/* ------------------------------------------------------------------------------------------------ asin */
/* IBM's asin as glibc 2.34's aarch64 build computes it (sysdeps/ieee754/dbl-64/e_asin.c, its tables asncs, inroot
   and powtwo; no slow paths): the argument below 2^-26; a Taylor polynomial below 0.125; six ranges of the asncs table
   (a point and a polynomial in the distance to it) below 0.96875; below 1 pi/2 - 2 asin(sqrt(z)), z = (1 - |x|) / 2,
   the square root by a Newton step from inroot and powtwo; pi/2 at 1; NaN beyond and for NaN. The fused steps are
   glibc's. The reference restates it at src/exact_ballpark/dlibm_inv.py:1626-1705. */

static inline XB_HD double xb_asin_table(double xa, int n, int top) {
    const double* a = XB_IV_ASIN_ASNCS;
    const double xx = xa - a[n];
    double poly = xb_fma(a[n + top], xx, a[n + top - 1]);
    for (int j = top - 2; j > 1; --j) poly = xb_fma(poly, xx, a[n + j]);
    const double p = xb_fma(xx * xx, poly, a[n + top + 1]);
    const double t = xb_fma(a[n + 1], xx, p);
    return t + a[n + top + 2];
}

static inline XB_HD double xb_asin(double x) {
    const uint64_t u = xb_bits(x);
    const int32_t m = (int32_t)(uint32_t)(u >> 32);
    const uint32_t k = (uint32_t)(u >> 32) & 0x7fffffffu;
    if (k < XB_IV_ASIN_K_TINY) return x;
    if (k < XB_IV_ASIN_K_TAYLOR) {
        const double x2 = x * x;
        double p = xb_fma(x2, XB_IV_ASIN_F6, XB_IV_ASIN_F5);
        const double t = x * x2;
        p = xb_fma(p, x2, XB_IV_ASIN_F4);
        p = xb_fma(p, x2, XB_IV_ASIN_F3);
        p = xb_fma(p, x2, XB_IV_ASIN_F2);
        p = xb_fma(p, x2, XB_IV_ASIN_F1);
        return xb_fma(p, t, x);
    }
    const double xa = m > 0 ? x : -x;
    double res;
    if (k < XB_IV_ASIN_K_HALF) {
        int n;
        if (k < XB_IV_ASIN_K_QUARTER) n = XB_IV_ASIN_N_Q1_MUL * (int)((k >> 15) & 0x1f);
        else n = XB_IV_ASIN_N_Q1_MUL * (int)((k >> 14) & 0x3f) + XB_IV_ASIN_N_Q2_OFF;
        res = xb_asin_table(xa, n, 6);
    } else if (k < XB_IV_ASIN_K_075) {
        res = xb_asin_table(xa, XB_IV_ASIN_N_H_OFF + 3 * (int)((k >> 11) & 0x1fc), 7);
    } else if (k < XB_IV_ASIN_K_0921875) {
        res = xb_asin_table(xa, XB_IV_ASIN_N_A_OFF + XB_IV_ASIN_N_A_MUL * (int)((k >> 13) & 0x7f), 8);
    } else if (k < XB_IV_ASIN_K_0953125) {
        res = xb_asin_table(xa, XB_IV_ASIN_N_B_OFF + 14 * (int)((k >> 13) & 0x7f), 9);
    } else if (k < XB_IV_ASIN_K_096875) {
        res = xb_asin_table(xa, XB_IV_ASIN_N_C_OFF + 15 * (int)((k >> 13) & 0x7f), 10);
    } else if (k < XB_IV_ASIN_K_NEAR_ONE) {
        const double z = (m > 0 ? XB_IV_ASIN_ONE - x : x + XB_IV_ASIN_ONE) * XB_IV_ASIN_HALF;
        const uint64_t zb = xb_bits(z);
        double t = XB_IV_ASIN_INROOT[(zb >> 46) & 0x7f] * XB_IV_ASIN_POWTWO[XB_IV_ASIN_POWTWO_TOP - (int)(zb >> 53)];
        double p = xb_fma(z, XB_IV_ASIN_F6, XB_IV_ASIN_F5);
        p = xb_fma(p, z, XB_IV_ASIN_F4);
        const double r = xb_fmsub(t * t, z, XB_IV_ASIN_ONE);
        p = xb_fma(p, z, XB_IV_ASIN_F3);
        double q = xb_fma(r, XB_IV_ASIN_RT3, XB_IV_ASIN_RT2);
        p = xb_fma(p, z, XB_IV_ASIN_F2);
        q = xb_fma(q, r, XB_IV_ASIN_RT1);
        p = xb_fma(p, z, XB_IV_ASIN_F1);
        q = xb_fma(q, r, XB_IV_ASIN_RT0);
        p = p * z;
        t = q * t;
        const double c = z * t;
        const double w = xb_fmsub(t * XB_IV_ASIN_HALF, c, XB_IV_ASIN_THREE_HALVES);
        const double y = (c + XB_IV_ASIN_T24) - XB_IV_ASIN_T24;
        const double ty = xb_fma(w, c, y);
        const double e = xb_fmsub(y, y, z);
        const double res1 = xb_fmsub(y, XB_IV_ASIN_TWO, XB_IV_ASIN_HP0);
        const double cc = e / ty;
        double s = y + cc;
        double cor = xb_fmsub(cc, XB_IV_ASIN_TWO, XB_IV_ASIN_HP1);
        s = s + s;
        cor = xb_fmsub(s, p, cor);
        res = cor + res1;
    } else if (k == XB_IV_ASIN_K_ONE && (uint32_t)u == 0u) {
        return m > 0 ? XB_IV_ASIN_HP0 : XB_IV_ASIN_NEG_HP0;
    } else if (k > XB_IV_ASIN_K_INF || (k == XB_IV_ASIN_K_INF && (uint32_t)u != 0u)) {
        return x + x;
    } else {
        return xb_nan();
    }
    return m > 0 ? res : -res;
}

/* ----------------------------------------------------------------------------------------------- atan2 */
/* IBM's atan2 as glibc 2.34's aarch64 build computes it (sysdeps/ieee754/dbl-64/e_atan2.c, its table cij; no slow
   paths): C99 Annex F's special values; by the exponents' distance pi/2, a quotient, or u = the smaller of |x| and |y|
   over the larger with its error du (one fused step), atan(u) from an odd polynomial below 1/16 or the cij table,
   placed in its quadrant with pi/2 or pi in two parts. The reference restates it at src/exact_ballpark/dlibm_inv.py:
   1708-1843. */

static inline XB_HD double xb_atan2_signed(double z, double y) {     /* z's magnitude with y's sign */
    return xb_dbl((xb_bits(y) & 0x8000000000000000ULL) | (xb_bits(z) & 0x7fffffffffffffffULL));
}

static inline XB_HD double xb_atan2_poly(double v) {
    double p = xb_fma(v, XB_IV_ATAN2_D13, XB_IV_ATAN2_D11);
    p = xb_fma(p, v, XB_IV_ATAN2_D9);
    p = xb_fma(p, v, XB_IV_ATAN2_D7);
    p = xb_fma(p, v, XB_IV_ATAN2_D5);
    return xb_fma(p, v, XB_IV_ATAN2_D3);
}

static inline XB_HD const double* xb_atan2_row(double u) {         /* cij's row nearest u (7 doubles a row) */
    const int i = (int)(xb_fma(u, XB_IV_ATAN2_TWO8, XB_IV_ATAN2_TWO52) - XB_IV_ATAN2_TWO52) - XB_IV_ATAN2_I_OFF;
    return XB_IV_ATAN2_CIJ + 7 * i;
}

static inline XB_HD double xb_atan2(double y, double x) {
    const uint64_t bx = xb_bits(x), by = xb_bits(y);
    const int32_t ux = (int32_t)(uint32_t)(bx >> 32), uy = (int32_t)(uint32_t)(by >> 32);
    const uint32_t dx = (uint32_t)bx, dy = (uint32_t)by;
    if ((ux & 0x7ff00000) == 0x7ff00000 && ((ux & 0xfffff) | dx) != 0) return x + y;
    const int y_inf = (uy & 0x7ff00000) == 0x7ff00000;
    if (y_inf && ((uy & 0xfffff) | dy) != 0) return y + y;
    if (!y_inf && uy == 0) {
        if (dy == 0) return ux < 0 ? XB_IV_ATAN2_OPI : 0.0;
        if (x == 0.0) return XB_IV_ATAN2_HPI;
    } else {
        if (dy == 0 && uy == (int32_t)0x80000000) return ux < 0 ? XB_IV_ATAN2_MOPI : -0.0;
        if (x == 0.0) return uy < 0 ? XB_IV_ATAN2_MHPI : XB_IV_ATAN2_HPI;
    }
    if (ux == 0x7ff00000 && dx == 0) {
        if (uy == 0x7ff00000 && dy == 0) return XB_IV_ATAN2_QPI;
        if (uy == (int32_t)0xfff00000 && dy == 0) return XB_IV_ATAN2_MQPI;
        return uy < 0 ? -0.0 : 0.0;
    }
    if (ux == (int32_t)0xfff00000 && dx == 0) {
        if (uy == 0x7ff00000 && dy == 0) return XB_IV_ATAN2_TQPI;
        if (uy == (int32_t)0xfff00000 && dy == 0) return XB_IV_ATAN2_MTQPI;
        return uy < 0 ? XB_IV_ATAN2_MOPI : XB_IV_ATAN2_OPI;
    }
    if (uy == 0x7ff00000) return XB_IV_ATAN2_HPI;
    if (uy == (int32_t)0xfff00000 && dy == 0) return XB_IV_ATAN2_MHPI;
    double ax = x < 0.0 ? -x : x, ay = y < 0.0 ? -y : y;
    const int32_t de = (int32_t)((uint32_t)(uy & 0x7ff00000) - (uint32_t)(ux & 0x7ff00000));
    if (de >= XB_IV_ATAN2_EP) return y > 0.0 ? XB_IV_ATAN2_HPI : XB_IV_ATAN2_MHPI;
    if (de <= XB_IV_ATAN2_EM) {
        if (x > 0.0) return xb_atan2_signed(ay / ax, y);
        return y > 0.0 ? XB_IV_ATAN2_OPI : XB_IV_ATAN2_MOPI;
    }
    if (ax < XB_IV_ATAN2_TWOM500 || ay < XB_IV_ATAN2_TWOM500) ax = ax * XB_IV_ATAN2_TWO500, ay = ay * XB_IV_ATAN2_TWO500;
    if (ax > XB_IV_ATAN2_TWO500 || ay > XB_IV_ATAN2_TWO500) ax = ax * XB_IV_ATAN2_TWOM500, ay = ay * XB_IV_ATAN2_TWOM500;
    double u, du;
    if (ay < ax) {
        u = ay / ax;
        const double v = ax * u, vv = xb_fnmsub(ax, u, v);
        du = ((ay - v) - vv) / ax;
    } else {
        u = ax / ay;
        const double v = ay * u, vv = xb_fnmsub(ay, u, v);
        du = ((ax - v) - vv) / ay;
    }
    if (x > 0.0) {
        if (ay < ax) {                                   /* atan(ay / ax) */
            if (u < XB_IV_ATAN2_INV16) {
                const double v = u * u;
                const double zz = xb_fma(u * v, xb_atan2_poly(v), du);
                return xb_atan2_signed(u + zz, y);
            }
            const double* row = xb_atan2_row(u);
            const double t3 = u - row[0];
            const double v = du + t3;
            const double dv = xb_fabs(t3) > xb_fabs(du) ? (t3 - v) + du : (du - v) + t3;
            double p = xb_fma(row[6], v, row[5]);
            p = xb_fma(p, v, row[4]);
            p = xb_fma(p, v, row[3]);
            double zz = (v * v) * p;
            zz = xb_fma(dv, row[2], zz);
            zz = xb_fma(v, row[2], zz);
            return xb_atan2_signed(zz + row[1], y);
        }
        if (u < XB_IV_ATAN2_INV16) {                     /* pi/2 - atan(ax / ay) */
            const double v = u * u;
            const double zz = (u * v) * xb_atan2_poly(v);
            const double t2 = XB_IV_ATAN2_HPI - u;
            const double cor = xb_fabs(u) < XB_IV_ATAN2_HPI ? (XB_IV_ATAN2_HPI - t2) - u : XB_IV_ATAN2_HPI - (u + t2);
            return xb_atan2_signed((((cor + XB_IV_ATAN2_HPI1) - du) - zz) + t2, y);
        }
        const double* row = xb_atan2_row(u);
        const double v = (u - row[0]) + du;
        double p = xb_fma(row[6], v, row[5]);
        p = xb_fma(p, v, row[4]);
        p = xb_fma(p, v, row[3]);
        p = xb_fma(p, v, row[2]);
        return xb_atan2_signed((XB_IV_ATAN2_HPI - row[1]) + xb_fmsub(p, v, XB_IV_ATAN2_HPI1), y);
    }
    if (ax < ay) {                                       /* pi/2 + atan(ax / ay) */
        if (u < XB_IV_ATAN2_INV16) {
            const double v = u * u;
            const double zz = (v * u) * xb_atan2_poly(v);
            const double t2 = u + XB_IV_ATAN2_HPI;
            const double cor = xb_fabs(u) < XB_IV_ATAN2_HPI ? (XB_IV_ATAN2_HPI - t2) + u : (u - t2) + XB_IV_ATAN2_HPI;
            return xb_atan2_signed((((cor + XB_IV_ATAN2_HPI1) + du) + zz) + t2, y);
        }
        const double* row = xb_atan2_row(u);
        const double v = (u - row[0]) + du;
        double p = xb_fma(row[6], v, row[5]);
        p = xb_fma(p, v, row[4]);
        p = xb_fma(p, v, row[3]);
        p = xb_fma(p, v, row[2]);
        return xb_atan2_signed((row[1] + XB_IV_ATAN2_HPI) + xb_fma(p, v, XB_IV_ATAN2_HPI1), y);
    }
    if (u < XB_IV_ATAN2_INV16) {                         /* pi - atan(ay / ax) */
        const double v = u * u;
        const double zz = (v * u) * xb_atan2_poly(v);
        const double t2 = XB_IV_ATAN2_OPI - u;
        const double cor = xb_fabs(u) < XB_IV_ATAN2_OPI ? (XB_IV_ATAN2_OPI - t2) - u : XB_IV_ATAN2_OPI - (t2 + u);
        return xb_atan2_signed((((cor + XB_IV_ATAN2_OPI1) - du) - zz) + t2, y);
    }
    const double* row = xb_atan2_row(u);
    const double v = (u - row[0]) + du;
    double p = xb_fma(row[6], v, row[5]);
    p = xb_fma(p, v, row[4]);
    p = xb_fma(p, v, row[3]);
    p = xb_fma(p, v, row[2]);
    return xb_atan2_signed((XB_IV_ATAN2_OPI - row[1]) + xb_fmsub(p, v, XB_IV_ATAN2_OPI1), y);
}

#endif
