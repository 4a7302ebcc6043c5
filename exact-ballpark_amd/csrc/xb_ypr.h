/* SPDX-License-Identifier: MIT
 *
 * The display orientation of a ball without dynamical orientation: Ball::CalculateYawPitchRoll (the reference restates
 * it at src/exact_ballpark/full.py:2043-2089, Modulo2pi at full.py:893-905). The yaw and pitch of the velocity (of the
 * way to the goto point at rest; mixed with it for a moving ball that is not a missile), reduced modulo 2 pi toward
 * the old values and smoothed by the mass; the roll driven by the yaw's change and damped. Snapped (no smoothing, the
 * roll reset) when a free ball is added and when its velocity is set; evolved at each tick's commit. One source for
 * the gfx950 kernels and their host build, included by xb_modes_tick.h; every operation is one IEEE double operation
 * in the reference's order, and atan2, asin and sin are glibc's (xb_maths_double.h).
 */
#ifndef XB_YPR_H
#define XB_YPR_H

#include <math.h>

#define XB_HALF_PI 1.57079632679489655              /* Ball.cpp:13-15 */
#define XB_TWO_PI 6.2831853071795862
#define XB_ROLL_SPEED_ACCELERATION (-3.0)           /* Ballpark.cpp:115-118 */
#define XB_ROLL_SPEED_DECAY (-0.9)
#define XB_ROLL_ACCELERATION 1.0
#define XB_ROLL_DECAY (-0.5)

/* the float64 rows of the display orientation block */
#define XB_Y_YAW 0
#define XB_Y_PITCH 1
#define XB_Y_ROLL 2
#define XB_Y_OLD_YAW 3
#define XB_Y_OLD_PITCH 4
#define XB_Y_OLD_ROLL 5
#define XB_Y_ROLL_SPEED 6
#define XB_Y_OLD_ROLL_SPEED 7
#define XB_Y_YAW_DELTA 8
#define XB_NY 9

struct XbYpr {
    double yaw, pitch, roll, old_yaw, old_pitch, old_roll, roll_speed, old_roll_speed, yaw_delta;
};

static inline XB_HD double xb_fmax_std(double a, double b) { return a < b ? b : a; }   /* std::max */
static inline XB_HD double xb_fmin_std(double a, double b) { return b < a ? b : a; }   /* std::min */

/* Modulo2pi: both angles reduced by fmod, then b moved by 2 pi toward a (the nearest of b, b + 2 pi, b - 2 pi). fmod's
   result is exact by definition (C's, and on the GPU the device library's, which computes it exactly). */
static inline XB_HD void xb_modulo2pi(double* a, double* b) {
    const double bb = fmod(*b, XB_TWO_PI), aa = fmod(*a, XB_TWO_PI);
    const double d0 = __builtin_fabs(bb - aa), dp = __builtin_fabs(bb + XB_TWO_PI - aa);
    const double dm = __builtin_fabs(bb - XB_TWO_PI - aa);
    *a = aa;
    if (d0 < dp) *b = (d0 < dm) ? bb : bb - XB_TWO_PI;
    else *b = (dp < dm) ? bb + XB_TWO_PI : bb - XB_TWO_PI;
}

/* CalculateYawPitchRoll for a ball as the tick's commit (or a setter) leaves it: `now` the world's clock in this tick,
   `hits` whether its collision list of this tick is not empty (an aligning warp then takes its velocity). */
static inline XB_HD void xb_yaw_pitch_roll(XbYpr& y, const XbShip& s, bool hits, int64_t now, bool snap,
                                          const XbPark& P) {
    y.old_pitch = y.pitch, y.old_yaw = y.yaw, y.old_roll = y.roll, y.old_roll_speed = y.roll_speed;
    double mm = s.mass / P.friction;
    mm = (mm != 0.0) ? 100.0 / mm : 100.0;
    const double vel2 = v3_lensq(s.vel);
    XbV3 d;
    if (vel2 == 0.0) {
        d = v3_sub(s.gp, s.pos);
    } else {
        d = v3(0.0, 0.0, 0.0);
        if (s.mode == XB_MODE_WARP) {
            if (s.effect_stamp < 0) {
                d = hits ? s.vel : v3_sub(s.gp, s.pos);
            } else if (mm < 1.0 && now != s.effect_stamp) {
                const int64_t inner_i = (now - s.effect_stamp) * 2;
                const double inner = 15.0 < (double)inner_i ? 15.0 : (double)inner_i;
                const double x = mm * inner;
                mm = 1.0 < x ? 1.0 : x;
            }
        } else {
            d = s.vel;
        }
        if (s.mode != XB_MODE_MISSILE && !snap) {
            const XbV3 off = v3_unit(v3_sub(s.gp, s.pos));
            const double sc = __builtin_sqrt(vel2) * 0.3;
            d = v3(d.x + off.x * sc, d.y + off.y * sc, d.z + off.z * sc);
        }
    }
    y.yaw = xb_atan2(d.x, d.z);
    const double r = v3_len(d);
    y.pitch = (r > 0.0) ? -xb_asin(xb_fmin_std(1.0, xb_fmax_std(-1.0, d.y / r))) : 0.0;
    xb_modulo2pi(&y.old_yaw, &y.yaw);
    xb_modulo2pi(&y.old_pitch, &y.pitch);
    if (snap) {
        y.roll = y.old_roll = y.roll_speed = y.old_roll_speed = y.yaw_delta = 0.0;
        return;
    }
    const double tmp = 1.0 / (1.0 + mm * 0.5);
    y.yaw = (y.old_yaw + mm * 0.5 * y.yaw) * tmp;
    y.pitch = (y.old_pitch + mm * 0.5 * y.pitch) * tmp;
    y.yaw_delta = xb_asin(xb_sin(y.yaw - y.old_yaw));
    y.roll_speed = y.roll_speed + y.yaw_delta * XB_ROLL_SPEED_ACCELERATION;
    y.roll_speed = y.roll_speed + y.old_roll_speed * XB_ROLL_SPEED_DECAY * P.dt;
    y.roll_speed = xb_fmax_std(-XB_HALF_PI, xb_fmin_std(XB_HALF_PI, y.roll_speed));
    y.roll = y.roll + y.roll_speed * XB_ROLL_ACCELERATION * P.dt;
    y.roll = y.roll + y.old_roll * XB_ROLL_DECAY * P.dt;
    y.roll = xb_fmax_std(-XB_HALF_PI, xb_fmin_std(XB_HALF_PI, y.roll));
}

static inline XB_HD void xb_ypr_load(const double* blk, int64_t stride, int64_t r, XbYpr& y) {
    y.yaw = blk[XB_Y_YAW * stride + r], y.pitch = blk[XB_Y_PITCH * stride + r], y.roll = blk[XB_Y_ROLL * stride + r];
    y.old_yaw = blk[XB_Y_OLD_YAW * stride + r], y.old_pitch = blk[XB_Y_OLD_PITCH * stride + r];
    y.old_roll = blk[XB_Y_OLD_ROLL * stride + r], y.roll_speed = blk[XB_Y_ROLL_SPEED * stride + r];
    y.old_roll_speed = blk[XB_Y_OLD_ROLL_SPEED * stride + r], y.yaw_delta = blk[XB_Y_YAW_DELTA * stride + r];
}

static inline XB_HD void xb_ypr_store(double* blk, int64_t stride, int64_t r, const XbYpr& y) {
    blk[XB_Y_YAW * stride + r] = y.yaw, blk[XB_Y_PITCH * stride + r] = y.pitch, blk[XB_Y_ROLL * stride + r] = y.roll;
    blk[XB_Y_OLD_YAW * stride + r] = y.old_yaw, blk[XB_Y_OLD_PITCH * stride + r] = y.old_pitch;
    blk[XB_Y_OLD_ROLL * stride + r] = y.old_roll, blk[XB_Y_ROLL_SPEED * stride + r] = y.roll_speed;
    blk[XB_Y_OLD_ROLL_SPEED * stride + r] = y.old_roll_speed, blk[XB_Y_YAW_DELTA * stride + r] = y.yaw_delta;
}

#endif
