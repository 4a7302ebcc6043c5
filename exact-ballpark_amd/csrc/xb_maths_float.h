/* SPDX-License-Identifier: LGPL-2.1-or-later
 *
 * The GNU C Library 2.34's float expf, logf, sinf, cosf (Arm optimized-routines: computed in double, rounded to float
 * once), acosf, atanf, atan2f (fdlibm, computed in float) and the double pow (Arm), as glibc's aarch64 build computes
 * them: each step a correctly rounded IEEE operation, fused exactly where GCC fused it there. The algorithms and
 * constants are glibc's; the reference restates them at src/exact_ballpark/vec_libm.py:348-737.
 *
 * Host and device from one source (XB_HD, XB_TABLE as in xb_maths_double.h). Float arithmetic is float: every
 * constant below is a float, so no expression is evaluated in double unless it says so.
 */
#ifndef XB_MATHS_FLOAT_H
#define XB_MATHS_FLOAT_H

#include "xb_maths_double.h"

static inline XB_HD uint32_t xb_fbits(float x) {
    union { float f; uint32_t u; } v;
    v.f = x;
    return v.u;
}
static inline XB_HD float xb_flt(uint32_t u) {
    union { float f; uint32_t u; } v;
    v.u = u;
    return v.f;
}
static inline XB_HD float xb_fmaf(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
/* the fdlibm functions' fused steps; a test build defines XB_FDLIBM_FMAF as two roundings to compare the same algorithm
   with a C library compiled without fused multiply-adds */
#ifndef XB_FDLIBM_FMAF
#define XB_FDLIBM_FMAF(a, b, c) xb_fmaf(a, b, c)
#endif
static inline XB_HD float xb_fabsf(float x) { return xb_flt(xb_fbits(x) & 0x7fffffffu); }
static inline XB_HD float xb_nanf(void) { return xb_flt(0x7fc00000u); }
static inline XB_HD float xb_inff(void) { return xb_flt(0x7f800000u); }
/* a double's nearest float (round to nearest even; IEEE conversion) */
static inline XB_HD float xb_tof(double x) { return (float)x; }

#define XB_FK(name) ((float)(name))      /* a float constant held as its exact double literal */

/* ---------------------------------------------------------------------------------------------------- expf */

static inline XB_HD float xb_expf(float x) {
    uint32_t ix = xb_fbits(x);
    uint32_t abstop = (ix >> 20) & 0x7ffu;
    if (abstop > 0x42au && !((x <= XB_FK(XB_LF_EXPF_OFLOW)) && (x >= XB_FK(XB_LF_EXPF_MAY_UFLOW)))) {
        if (ix == 0xff800000u) return 0.0f;
        if (abstop >= 0x7f8u) return x + x;
        if (x > XB_FK(XB_LF_EXPF_OFLOW)) return xb_inff();
        if (x < XB_FK(XB_LF_EXPF_UFLOW)) return 0.0f;
        return XB_FK(XB_LF_MAY_UFLOWF);
    }
    double z = (double)x * XB_LF_EXPF_INVLN2;
    double kd = xb_round_away(z);
    int64_t ki = (int64_t)kd;
    double r = z - kd;
    double s = xb_dbl(XB_LF_EXP2F_TAB[ki & 31] + ((uint64_t)ki << 47));
    double zc = xb_fma(r, XB_LF_EXPF_C[0], XB_LF_EXPF_C[1]);
    double y = xb_fma(r, XB_LF_EXPF_C[2], 1.0);
    double r2 = r * r;
    y = xb_fma(zc, r2, y);
    return xb_tof(y * s);
}

/* ---------------------------------------------------------------------------------------------------- logf */

static inline XB_HD float xb_logf(float x) {
    uint32_t ix = xb_fbits(x);
    if (ix == 0x3f800000u) return 0.0f;
    if (ix - 0x00800000u >= 0x7f000000u) {
        if (!(ix > 0u && ix < 0x00800000u)) {
            if ((ix & 0x7fffffffu) == 0u) return xb_flt(0xff800000u);
            if (ix == 0x7f800000u) return xb_inff();
            return xb_nanf();
        }
        ix = xb_fbits(x * 8388608.0f) - (23u << 23);
    }
    uint32_t tmp = ix - 0x3f330000u;
    uint32_t i = (tmp >> 19) & 15u;
    int32_t k = (int32_t)tmp >> 23;
    float z = xb_flt(ix - (tmp & 0xff800000u));
    double invc = XB_LF_LOGF_TAB[2 * i], logc = XB_LF_LOGF_TAB[2 * i + 1];
    double r = xb_fma((double)z, invc, -1.0);
    double y0 = xb_fma((double)k, XB_LF_LOGF_LN2, logc);
    double y = xb_fma(r, XB_LF_LOGF_A[1], XB_LF_LOGF_A[2]);
    double r2 = r * r;
    double ry0 = r + y0;
    y = xb_fma(r2, XB_LF_LOGF_A[0], y);
    return xb_tof(xb_fma(y, r2, ry0));
}

/* ------------------------------------------------------------------------------------------- sinf, cosf */

static inline XB_HD double xb_sinf_poly(double xv, double x2) {
    double x3 = xv * x2;
    double s1 = xb_fma(x2, XB_LF_SINF_S[2], XB_LF_SINF_S[1]);
    double x7 = x2 * x3;
    double s = xb_fma(x3, XB_LF_SINF_S[0], xv);
    return xb_fma(s1, x7, s);
}

static inline XB_HD double xb_cosf_poly(double x2) {
    double c1 = xb_fma(x2, XB_LF_COSF_C[1], XB_LF_COSF_C[0]);
    double c2 = xb_fma(x2, XB_LF_COSF_C[4], XB_LF_COSF_C[3]);
    double x4 = x2 * x2;
    double x6 = x2 * x4;
    double c = xb_fma(x4, XB_LF_COSF_C[2], c1);
    return xb_fma(c2, x6, c);
}

/* reduce_large: x 4/pi in 62-bit fixed point from __inv_pio4, integers only */
static inline XB_HD void xb_sincosf_reduce_large(uint32_t xi, int64_t* np, double* xr) {
    uint32_t idx = (xi >> 26) & 15u;
    uint32_t shift = (xi >> 23) & 7u;
    uint32_t m = ((xi & 0x7fffffu) | 0x800000u) << shift;
    uint64_t res0 = (uint64_t)(uint32_t)(m * (uint32_t)XB_LF_INV_PIO4[idx]);
    uint64_t res1 = (uint64_t)m * XB_LF_INV_PIO4[idx + 4];
    uint64_t res2 = (uint64_t)m * XB_LF_INV_PIO4[idx + 8];
    res0 = (res2 >> 32) | (res0 << 32);
    res0 += res1;
    uint64_t n = (res0 + (1ULL << 61)) >> 62;
    res0 -= n << 62;
    *np = (int64_t)n;
    *xr = (double)(int64_t)res0 * XB_LF_PI63;
}

static inline XB_HD float xb_sincosf_one(float x, int cosine) {
    uint32_t ix = xb_fbits(x);
    uint32_t abstop = (ix >> 20) & 0x7ffu;
    if (abstop >= 0x7f8u) return xb_nanf();
    if (abstop < 0x3f4u) {
        if (abstop < 0x398u) return cosine ? 1.0f : x;
        double xd = (double)x;
        double xx = xd * xd;
        return xb_tof(cosine ? xb_cosf_poly(xx) : xb_sinf_poly(xd, xx));
    }
    int64_t n, q;
    double xr;
    if (abstop < 0x42fu) {
        double xd = (double)x;
        double kd = xb_round_away(xd * XB_LF_HPI_INV);
        xr = xb_fma(kd, -XB_LF_HPI, xd);
        n = (int64_t)kd;
        q = n;
    } else {
        xb_sincosf_reduce_large(ix, &n, &xr);
        q = n + (int64_t)(ix >> 31);
    }
    double s = XB_LF_SINCOSF_SIGN[q & 3];
    double x2 = xr * xr;
    int use_cos = cosine ? ((n & 1) == 0) : ((n & 1) == 1);
    double v = use_cos ? xb_cosf_poly(x2) * (((q & 2) != 0) ? -1.0 : 1.0) : xb_sinf_poly(xr * s, x2);
    return xb_tof(v);
}

static inline XB_HD float xb_sinf(float x) { return xb_sincosf_one(x, 0); }
static inline XB_HD float xb_cosf(float x) { return xb_sincosf_one(x, 1); }

/* -------------------------------------------------------------------------------------------------- acosf */

static inline XB_HD float xb_acosf_r(float z) {
    float p4 = XB_FDLIBM_FMAF(z, XB_FK(XB_LF_ACOSF_PS5), XB_FK(XB_LF_ACOSF_PS4));
    float q3 = XB_FDLIBM_FMAF(z, XB_FK(XB_LF_ACOSF_QS4), XB_FK(XB_LF_ACOSF_QS3));
    float p3 = XB_FDLIBM_FMAF(p4, z, XB_FK(XB_LF_ACOSF_PS3));
    float q2 = XB_FDLIBM_FMAF(q3, z, XB_FK(XB_LF_ACOSF_QS2));
    float p2 = XB_FDLIBM_FMAF(p3, z, XB_FK(XB_LF_ACOSF_PS2));
    float q1 = XB_FDLIBM_FMAF(q2, z, XB_FK(XB_LF_ACOSF_QS1));
    float p1 = XB_FDLIBM_FMAF(p2, z, XB_FK(XB_LF_ACOSF_PS1));
    float q = XB_FDLIBM_FMAF(q1, z, 1.0f);
    float p0 = XB_FDLIBM_FMAF(p1, z, XB_FK(XB_LF_ACOSF_PS0));
    float p = p0 * z;
    return p / q;
}

static inline XB_HD float xb_acosf(float x) {
    uint32_t hx = xb_fbits(x), ix = hx & 0x7fffffffu;
    if (ix > 0x3f800000u) return xb_nanf();
    if (ix == 0x3f800000u) return (hx >> 31) ? XB_FK(XB_LF_ACOSF_PI) + XB_FK(XB_LF_ACOSF_TWO_PIO2_LO) : 0.0f;
    if (ix <= 0x32800000u) return XB_FK(XB_LF_ACOSF_PIO2_HI) + XB_FK(XB_LF_ACOSF_PIO2_LO);
    if (ix < 0x3f000000u) {
        float z = x * x;
        float t = XB_FDLIBM_FMAF(-x, xb_acosf_r(z), XB_FK(XB_LF_ACOSF_PIO2_LO));
        return XB_FK(XB_LF_ACOSF_PIO2_HI) - (x - t);
    }
    if (hx >> 31) {
        float z = (x + 1.0f) * 0.5f;
        float r = xb_acosf_r(z);
        float s = __builtin_sqrtf(z);
        float w = XB_FDLIBM_FMAF(s, r, XB_FK(XB_LF_ACOSF_NEG_PIO2_LO));
        return XB_FDLIBM_FMAF(-(s + w), 2.0f, XB_FK(XB_LF_ACOSF_PI));
    }
    float z = (1.0f - x) * 0.5f;
    float s = __builtin_sqrtf(z);
    float df = xb_flt(xb_fbits(s) & 0xfffff000u);
    float c = XB_FDLIBM_FMAF(-df, df, z) / (s + df);
    float w = XB_FDLIBM_FMAF(s, xb_acosf_r(z), c);
    float dw = df + w;
    return dw + dw;
}

/* ------------------------------------------------------------------------------------------ atanf, atan2f */

static inline XB_HD float xb_atanf(float x) {
    uint32_t hx = xb_fbits(x), ix = hx & 0x7fffffffu;
    if (ix > 0x7f800000u) return x + x;
    if (ix > 0x4bffffffu)
        return (hx >> 31) ? -XB_FK(XB_LF_ATANF_HI3) - XB_FK(XB_LF_ATANF_LO3)
                          : XB_FK(XB_LF_ATANF_HI3) + XB_FK(XB_LF_ATANF_LO3);
    if (ix <= 0x30ffffffu) return x;
    int below = ix <= 0x3edfffffu;
    int k = 0;
    float xr = x;
    if (!below) {
        float ax = xb_fabsf(x);
        if (ix <= 0x3f97ffffu) {
            if (ix <= 0x3f2fffffu) {
                k = 0;
                xr = XB_FDLIBM_FMAF(ax, 2.0f, -1.0f) / (ax + 2.0f);
            } else {
                k = 1;
                xr = (ax - 1.0f) / (ax + 1.0f);
            }
        } else if (ix <= 0x401bffffu) {
            k = 2;
            xr = (ax - 1.5f) / XB_FDLIBM_FMAF(ax, 1.5f, 1.0f);
        } else {
            k = 3;
            xr = -1.0f / ax;
        }
    }
    float z = xr * xr;
    float w = z * z;
    float t8 = XB_FDLIBM_FMAF(w, XB_FK(XB_LF_ATANF_T10), XB_FK(XB_LF_ATANF_T8));
    float t7 = XB_FDLIBM_FMAF(w, XB_FK(XB_LF_ATANF_T9), XB_FK(XB_LF_ATANF_T7));
    float t6 = XB_FDLIBM_FMAF(t8, w, XB_FK(XB_LF_ATANF_T6));
    float t5 = XB_FDLIBM_FMAF(t7, w, XB_FK(XB_LF_ATANF_T5));
    float t4 = XB_FDLIBM_FMAF(t6, w, XB_FK(XB_LF_ATANF_T4));
    float t3 = XB_FDLIBM_FMAF(t5, w, XB_FK(XB_LF_ATANF_T3));
    float t2 = XB_FDLIBM_FMAF(t4, w, XB_FK(XB_LF_ATANF_T2));
    float t1 = XB_FDLIBM_FMAF(t3, w, XB_FK(XB_LF_ATANF_T1));
    float t0 = XB_FDLIBM_FMAF(t2, w, XB_FK(XB_LF_ATANF_T0));
    float s12 = XB_FDLIBM_FMAF(t0, z, t1 * w);
    float xs12 = xr * s12;
    if (below) return xr - xs12;
    float zz = XB_FK(XB_LF_ATANF_HI[k]) - ((xs12 - XB_FK(XB_LF_ATANF_LO[k])) - xr);
    return (hx >> 31) ? -zz : zz;
}

static inline XB_HD float xb_atan2f(float y, float x) {
    uint32_t hy = xb_fbits(y), hx = xb_fbits(x);
    uint32_t iy = hy & 0x7fffffffu, ix = hx & 0x7fffffffu;
    if (ix > 0x7f800000u || iy > 0x7f800000u) return x + y;
    if (hx == 0x3f800000u) return xb_atanf(y);
    uint32_t m = ((hx >> 30) & 2u) | (hy >> 31);
    const float pi = XB_FK(XB_LF_ATAN2F_PI), tiny = XB_FK(XB_LF_ATAN2F_TINY);
    float pi_t = pi + tiny;
    float mpi_t = XB_FK(XB_LF_ATAN2F_NEG_PI) - tiny;
    if (iy == 0u) return m <= 1u ? y : (m == 2u ? pi_t : mpi_t);
    float pio2_t = (hy >> 31) ? XB_FK(XB_LF_ATAN2F_NEG_PIO2) - tiny : XB_FK(XB_LF_ATAN2F_PIO2) + tiny;
    if (ix == 0u) return pio2_t;
    if (ix == 0x7f800000u) {
        if (iy == 0x7f800000u) {
            if (m == 0u) return XB_FK(XB_LF_ATAN2F_PIO4) + tiny;
            if (m == 1u) return XB_FK(XB_LF_ATAN2F_NEG_PIO4) - tiny;
            if (m == 2u) return XB_FDLIBM_FMAF(3.0f, XB_FK(XB_LF_ATAN2F_PIO4), tiny);
            return XB_FDLIBM_FMAF(-3.0f, XB_FK(XB_LF_ATAN2F_PIO4), XB_FK(XB_LF_ATAN2F_NEG_TINY));
        }
        if (m == 0u) return 0.0f;
        if (m == 1u) return -0.0f;
        return m == 2u ? pi_t : mpi_t;
    }
    if (iy == 0x7f800000u) return pio2_t;
    int32_t k = ((int32_t)iy - (int32_t)ix) >> 23;
    float z;
    if (k > 60) z = XB_FK(XB_LF_ATAN2F_PIO2) - XB_FK(XB_LF_ATAN2F_NEG_HALF_PI_LO);
    else if ((hx >> 31) && k < -60) z = 0.0f;
    else z = xb_atanf(xb_fabsf(y / x));
    float zl = z + XB_FK(XB_LF_ATAN2F_NEG_PI_LO);
    if (m == 0u) return z;
    if (m == 1u) return -z;
    if (m == 2u) return pi - zl;
    return zl - pi;
}

/* ---------------------------------------------------------------------------------------------- pow (double) */

static inline XB_HD int xb_pow_checkint(uint64_t iy) {
    int e = (int)((iy >> 52) & 0x7ff);
    if (e < 0x3ff) return 0;
    if (e > 0x3ff + 52) return 2;
    if (iy & ((1ULL << (0x3ff + 52 - e)) - 1)) return 0;
    if (iy & (1ULL << (0x3ff + 52 - e))) return 1;
    return 2;
}

static inline XB_HD double xb_pow_log(uint64_t ix, double* tail) {
    uint64_t tmp = ix - XB_LF_POW_OFF;
    int i = (int)((tmp >> 45) & 127);
    double kd = (double)((int64_t)tmp >> 52);
    double z = xb_dbl(ix - (tmp & 0xfff0000000000000ULL));
    double invc = XB_LF_POW_TAB[3 * i], logc = XB_LF_POW_TAB[3 * i + 1], logctail = XB_LF_POW_TAB[3 * i + 2];
    double t1 = xb_fma(kd, XB_LF_POW_LN2HI, logc);
    double r = xb_fma(z, invc, -1.0);
    double lo1 = xb_fma(kd, XB_LF_POW_LN2LO, logctail);
    double ar = r * XB_LF_POW_A[0];
    double a56 = xb_fma(r, XB_LF_POW_A[6], XB_LF_POW_A[5]);
    double t2 = r + t1;
    double a34 = xb_fma(r, XB_LF_POW_A[4], XB_LF_POW_A[3]);
    double a12 = xb_fma(r, XB_LF_POW_A[2], XB_LF_POW_A[1]);
    double ar2 = r * ar;
    double d = t1 - t2;
    double lo3 = xb_fma(ar, r, -ar2);
    double a3456 = xb_fma(ar2, a56, a34);
    double hi = t2 + ar2;
    double lo2 = d + r;
    double ar3 = r * ar2;
    double e = t2 - hi;
    double s = lo1 + lo2;
    double poly = xb_fma(ar2, a3456, a12);
    double lo4 = e + ar2;
    s = s + lo3;
    s = s + lo4;
    double lo = xb_fma(ar3, poly, s);
    double y = hi + lo;
    *tail = (hi - y) + lo;
    return y;
}

static inline XB_HD double xb_pow_exp(double ehi, double elo, int negative) {
    uint32_t abstop = (uint32_t)((xb_bits(ehi) >> 52) & 0x7ff);
    if (abstop < 0x3c9u) return negative ? -(ehi + 1.0) : ehi + 1.0;
    if (abstop > 0x408u) {
        double flow = ehi < 0.0 ? 0.0 : xb_inf();
        return negative ? -flow : flow;
    }
    double kd = xb_round_away(ehi * XB_EX_INVLN2N);
    int64_t ki = (int64_t)kd;
    double r = xb_fma(kd, XB_EX_NEGLN2HIN, ehi);
    r = xb_fma(kd, XB_EX_NEGLN2LON, r);
    int idx = (int)((ki & 0x7f) * 2);
    double tail = xb_dbl(XB_EXP_TAB[idx]);
    uint64_t base = XB_EXP_TAB[idx + 1] + ((uint64_t)ki << 45);
    uint64_t sgn = negative ? 0x8000000000000000ULL : 0ULL;
    r = elo + r;
    double p23 = xb_fma(r, XB_EX_C3, XB_EX_C2);
    double p45 = xb_fma(r, XB_EX_C5, XB_EX_C4);
    double r2 = r * r;
    double t = r + tail;
    double r4 = r2 * r2;
    double tmp = xb_fma(r2, p23, t);
    tmp = xb_fma(r4, p45, tmp);
    if (abstop != 0x408u) {
        double scale = xb_dbl(base ^ sgn);
        return xb_fma(tmp, scale, scale);
    }
    if ((ki & 0x80000000LL) == 0) {
        double sc_hi = xb_dbl((base - (1009ULL << 52)) ^ sgn);
        return xb_fma(tmp, sc_hi, sc_hi) * XB_EX_TWO_1009;
    }
    uint64_t sb_lo = (base + (1022ULL << 52)) ^ sgn;
    double sc_lo = xb_dbl(sb_lo);
    double st = tmp * sc_lo;
    double y = sc_lo + st;
    if (!(xb_fabs(y) < 1.0)) return y * XB_EX_TWO_M1022;
    double one = y < 0.0 ? -1.0 : 1.0;
    double lo = (sc_lo - y) + st;
    double hi = y + one;
    lo = ((one - hi) + y) + lo;
    double y2 = (lo + hi) - one;
    if (y2 == 0.0) y2 = xb_dbl(sb_lo & 0x8000000000000000ULL);
    return y2 * XB_EX_TWO_M1022;
}

static inline XB_HD double xb_pow(double x, double y) {
    const uint64_t SIGN = 0x8000000000000000ULL, ONE = 0x3ff0000000000000ULL, INF = 0x7ff0000000000000ULL;
    uint64_t ix = xb_bits(x);
    uint64_t iy = xb_bits(y);
    uint32_t topx = (uint32_t)(ix >> 52);
    uint32_t topy = (uint32_t)(iy >> 52);
    int negative = 0;
    if (topx - 0x001u >= 0x7ffu - 0x001u || (topy & 0x7ffu) - 0x3beu >= 0x43eu - 0x3beu) {
        if ((topy & 0x7ffu) == 0x7ffu || (iy & ~SIGN) == 0) {
            if ((iy & ~SIGN) == 0) return 1.0;
            if (ix == ONE) return 1.0;
            if ((ix & ~SIGN) > INF || (iy & ~SIGN) > INF) return x + y;
            if ((ix & ~SIGN) == ONE) return 1.0;
            if (((ix & ~SIGN) < ONE) == !(iy >> 63)) return 0.0;
            return y * y;
        }
        if ((topx & 0x7ffu) == 0x7ffu || (ix & ~SIGN) == 0) {
            double x2 = x * x;
            if ((ix >> 63) && xb_pow_checkint(iy) == 1) x2 = -x2;
            return (iy >> 63) ? 1.0 / x2 : x2;
        }
        if (ix >> 63) {
            int yint = xb_pow_checkint(iy);
            if (yint == 0) return xb_nan();
            if (yint == 1) negative = 1;
            ix &= ~SIGN;
            topx &= 0x7ffu;
        }
        if ((topy & 0x7ffu) - 0x3beu >= 0x43eu - 0x3beu) {
            if (ix == ONE) return 1.0;
            if ((topy & 0x7ffu) < 0x3beu) return ix > ONE ? 1.0 + y : 1.0 - y;
            return (ix > ONE) == (topy < 0x800u) ? xb_inf() : 0.0;
        }
        if (topx == 0u) {
            ix = xb_bits(xb_dbl(ix) * 0x1p52) & ~SIGN;
            ix -= 52ULL << 52;
        }
    }
    double lo;
    double hi = xb_pow_log(ix, &lo);
    double ehi = y * hi;
    double elo = xb_fma(y, lo, xb_fma(y, hi, -ehi));
    return xb_pow_exp(ehi, elo, negative);
}

/* ------------------------------------------------------------------------------------------------ powf */
/* powf from ARM's optimized-routines as glibc 2.34's aarch64 build computes it (sysdeps/ieee754/flt-32/e_powf.c, its
   tables __powf_log2_data and __exp2f_data): log2 x scaled by 32 from one of 16 points and a quintic in double, times
   y, then 2^(y log2 x) from a 32-entry table and a cubic, rounded to float once. The scaled exponent is rounded to the
   nearest integer with ties away from zero (the TOINT_INTRINSICS build: frinta and fcvtas), not by the 1.5 2^52 shift
   of other builds, which rounds ties to even. Nine multiply-adds are fused, as that build fuses them. The special
   cases, the parity of y for a negative x, the scaling of a subnormal x and the overflow and underflow thresholds are
   that routine's; the flags and errno are not modelled, and the rounding mode is to nearest, so the directed-rounding
   check before the overflow threshold is never taken. The reference restates it at src/exact_ballpark/dlibm_powf.py:
   193-295. */

static inline XB_HD int xb_powf_signaling(uint32_t u) {
    return ((u ^ (uint32_t)XB_PF_POWF_QNAN_BIT) & 0x7fffffffu) > (uint32_t)XB_PF_POWF_QNAN;
}

static inline XB_HD float xb_powf_xflow(uint32_t sign, uint32_t tbits) {     /* (sign ? -t : t) * t in float */
    const float t = xb_flt(tbits);
    return (sign ? -t : t) * t;
}

static inline XB_HD double xb_powf_log2(uint32_t ix) {        /* 32 log2 x for the bits of a positive normal x */
    const uint32_t tmp = ix + (uint32_t)XB_PF_POWF_NEG_OFF;
    const int i = (int)((tmp >> XB_PF_POWF_I_SHIFT) & ((1u << XB_PF_POWF_I_BITS) - 1u));
    const uint32_t top = tmp & (uint32_t)XB_PF_POWF_TOP_MASK;
    const uint32_t iz = ix - top;
    const int32_t k = (int32_t)top >> XB_PF_POWF_K_SHIFT;
    const double invc = XB_PF_POWF_LOG2_INVC[i], logc = XB_PF_POWF_LOG2_LOGC[i];
    const double z = (double)xb_flt(iz);
    const double r = xb_fma(invc, z, -1.0);
    const double y0 = (double)k + logc;
    const double p = xb_fma(r, XB_PF_POWF_A[2], XB_PF_POWF_A[3]);
    double q = xb_fma(r, XB_PF_POWF_A[4], y0);
    const double r2 = r * r;
    const double y = xb_fma(r, XB_PF_POWF_A[0], XB_PF_POWF_A[1]);
    const double r4 = r2 * r2;
    q = xb_fma(r2, p, q);
    return xb_fma(y, r4, q);
}

static inline XB_HD float xb_powf_exp2(double xd, uint32_t sign_bias) {   /* 2^(xd / 32) in float, its sign */
    const double kd = xb_round_away(xd);
    const int64_t ki = (int64_t)kd;
    const double r = xd - kd;
    double y = xb_fma(r, XB_PF_POWF_EXP2_C[2], 1.0);
    const double z = xb_fma(r, XB_PF_POWF_EXP2_C[0], XB_PF_POWF_EXP2_C[1]);
    const double r2 = r * r;
    uint64_t t = XB_PF_POWF_EXP2_TAB[ki & 31];
    t += ((uint64_t)ki + (uint64_t)sign_bias) << XB_PF_POWF_EXP2_SHIFT;
    y = xb_fma(z, r2, y);
    return (float)(y * xb_dbl(t));
}

static inline XB_HD float xb_powf_y_special(float x, float y, uint32_t ix, uint32_t iy) {  /* y zero, inf or NaN */
    const uint32_t two_iy = iy << 1;
    if (two_iy == 0u) return xb_powf_signaling(ix) ? x + y : 1.0f;
    if (ix == (uint32_t)XB_PF_POWF_ONE_BITS) return xb_powf_signaling(iy) ? x + y : 1.0f;
    const uint32_t two_ix = ix << 1;
    if (two_iy != (uint32_t)XB_PF_POWF_INF2 || two_ix > (uint32_t)XB_PF_POWF_INF2) return x + y;
    if (two_ix == (uint32_t)XB_PF_POWF_ONE2) return 1.0f;
    if ((two_ix < (uint32_t)XB_PF_POWF_ONE2) == !(iy >> 31)) return 0.0f;
    return y * y;
}

static inline XB_HD float xb_powf(float x, float y) {
    uint32_t ix = xb_fbits(x);
    const uint32_t iy = xb_fbits(y);
    const int zinf_y = (iy << 1) - 1u >= (uint32_t)XB_PF_POWF_ZINF;
    uint32_t sign_bias = 0;
    if (ix - (uint32_t)XB_PF_POWF_X_MIN > (uint32_t)XB_PF_POWF_X_SPAN) {   /* below 2^-126, negative, inf or NaN */
        if (zinf_y) return xb_powf_y_special(x, y, ix, iy);
        const uint32_t two_ix = ix << 1;
        if (two_ix - 1u >= (uint32_t)XB_PF_POWF_ZINF) {                  /* x zero, infinite or NaN */
            float x2 = x * x;
            uint32_t sb = 0;
            if (ix >> 31) {                                                /* y an odd integer: the sign stays */
                const uint32_t e = (iy >> 23) & 0xffu;
                if (e - (uint32_t)XB_PF_POWF_E_ONE <= (uint32_t)XB_PF_POWF_E_SPAN) {
                    const uint32_t bit = 1u << (XB_PF_POWF_INT_E - e);
                    if (!(iy & (bit - 1u)) && (iy & bit)) x2 = -x2, sb = 1;
                }
            }
            if (two_ix == 0u && (iy >> 31)) return (sb ? -1.0f : 1.0f) / 0.0f;
            return (iy >> 31) ? 1.0f / x2 : x2;
        }
        if (ix >> 31) {                                                    /* a finite x < 0: y must be an integer */
            const uint32_t e = (iy >> 23) & 0xffu;
            if (e <= (uint32_t)XB_PF_POWF_E_HALF) return (x - x) / (x - x);
            if (e <= (uint32_t)XB_PF_POWF_INT_E) {
                const uint32_t bit = 1u << (XB_PF_POWF_INT_E - e);
                if (iy & (bit - 1u)) return (x - x) / (x - x);
                if (iy & bit) sign_bias = XB_PF_POWF_SIGN_BIAS;
            }
            ix &= 0x7fffffffu;
        }
        if (ix <= (uint32_t)XB_PF_POWF_SUBNORMAL_MAX)                     /* subnormal: scaled by 2^23, then back */
            ix = (xb_fbits(x * xb_flt((uint32_t)XB_PF_POWF_TWO23_F)) & 0x7fffffffu) - (uint32_t)XB_PF_POWF_NORM_SUB;
    } else if (zinf_y) {
        return xb_powf_y_special(x, y, ix, iy);
    }
    const double ylogx = (double)y * xb_powf_log2(ix);
    if (((xb_bits(ylogx) >> 47) & 0xffffu) > (uint64_t)XB_PF_POWF_BIG) {  /* |y log2 x| >= 126 */
        if (ylogx > XB_PF_POWF_OFLOW) return xb_powf_xflow(sign_bias, (uint32_t)XB_PF_POWF_OFLOWF_X);
        if (ylogx <= xb_dbl(XB_PF_POWF_UFLOW)) return xb_powf_xflow(sign_bias, (uint32_t)XB_PF_POWF_UFLOWF_X);
        if (ylogx < xb_dbl(XB_PF_POWF_MAY_UFLOW)) return xb_powf_xflow(sign_bias, (uint32_t)XB_PF_POWF_MAY_UFLOWF_X);
    }
    return xb_powf_exp2(ylogx, sign_bias);
}

#endif
