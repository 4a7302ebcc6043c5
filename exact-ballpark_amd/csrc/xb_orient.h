// SPDX-License-Identifier: MIT
//
// Dynamical orientation (MOV-728) for one ship, in float, as destiny computes it: the heading torque
// (GotoTorqueBreaks), the roll torque (GotoRollBreaks), the rotation and angular velocity a tick on (ApplyTorque), the
// thrust laid along the heading, SnapOrientation, and the old rotation of a new ball; the reference's scalar statement
// is src/exact_ballpark/full.py:2091-2485 (with carbonengine's vectors, quaternions and matrices at full.py:179-325).
//
// Every float operation here is one IEEE binary32 operation (-ffp-contract=off: nothing fused; division and square
// root correctly rounded), which is what full.py's double operation rounded to float gives for + - * / sqrt of floats.
// Where full.py computes in double (a float times the double tick, comparisons with double literals, the length of a
// double thrust) the code says (double). The transcendentals are glibc's float functions and its double pow
// (xb_maths_float.h). std::max and std::min are written as their comparisons, so NaNs go where they go in C++.
//
// Include after xb_modes_tick.h's vector helpers and park (it is included from there).
#pragma once

#include "xb_maths_float.h"

struct XoV3 { float x, y, z; };
struct XoQ { float x, y, z, w; };
struct XoM { float m00, m01, m02, m10, m11, m12, m20, m21, m22; };   // the fourth row and column: 0 0 0 1

/* A ship's turning state: rotation, angular velocity and the turning values in; the old rotation and angular velocity,
   the torque and the roll torque out (floats held in the park's float64 block rows). */
struct XbTurn {
    XoQ rot, old_rot;
    XoV3 w, old_w, torque;
    float roll, max_ang_vel, rot_agility;
};

static inline XB_HD float xo_max(float a, float b) { return a < b ? b : a; }     // std::max
static inline XB_HD float xo_min(float a, float b) { return b < a ? b : a; }     // std::min
static inline XB_HD float xo_abs(float x) { return xb_fabsf(x); }
static inline XB_HD bool xo_finite(float x) { return (xb_fbits(x) & 0x7f800000u) != 0x7f800000u; }
static inline XB_HD int xo_to_int32(float x) {                                    // int(x) as aarch64 converts it
    if (x != x) return 0;
    if ((double)x >= 2147483648.0) return 2147483647;
    if ((double)x < -2147483648.0) return -2147483647 - 1;
    return (int)x;
}

static inline XB_HD XoV3 xo_v(float x, float y, float z) { XoV3 v = {x, y, z}; return v; }
static inline XB_HD XoV3 xo_of(XbV3 d) { return xo_v((float)d.x, (float)d.y, (float)d.z); }
static inline XB_HD XoV3 xo_add(XoV3 a, XoV3 b) { return xo_v(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline XB_HD XoV3 xo_sub(XoV3 a, XoV3 b) { return xo_v(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline XB_HD XoV3 xo_mul(XoV3 a, float f) { return xo_v(a.x * f, a.y * f, a.z * f); }
static inline XB_HD XoV3 xo_div(XoV3 a, float f) { return xo_mul(a, 1.0f / f); }    // Vector3 / float: times 1/f
static inline XB_HD float xo_dot(XoV3 a, XoV3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline XB_HD float xo_len(XoV3 a) { return __builtin_sqrtf(xo_dot(a, a)); }
static inline XB_HD XoV3 xo_cross(XoV3 a, XoV3 b) {
    return xo_v(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline XB_HD XoV3 xo_normalize(XoV3 v) {         // by the largest magnitude, then by the length
    const float mx = xo_max(xo_max(xo_abs(v.x), xo_abs(v.y)), xo_abs(v.z));
    const XoV3 mini = xo_div(v, mx != 0.0f ? mx : 1.0f);
    float length = xo_len(mini);
    if (length != 0.0f) length = 1.0f / length;
    return xo_mul(mini, length);
}

static inline XB_HD XoQ xo_q(float x, float y, float z, float w) { XoQ q = {x, y, z, w}; return q; }
static inline XB_HD XoQ xo_qmul(XoQ q, XoQ o) {          // Quaternion::operator*: q * o
    return xo_q(((o.w * q.x + o.x * q.w) + o.y * q.z) - o.z * q.y, ((o.w * q.y - o.x * q.z) + o.y * q.w) + o.z * q.x,
                ((o.w * q.z + o.x * q.y) - o.y * q.x) + o.z * q.w, ((o.w * q.w - o.x * q.x) - o.y * q.y) - o.z * q.z);
}
static inline XB_HD float xo_qlensq(XoQ q) { return ((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w; }
static inline XB_HD float xo_qlen(XoQ q) { return __builtin_sqrtf(xo_qlensq(q)); }
static inline XB_HD XoQ xo_qdiv(XoQ q, float f) {
    const float d = 1.0f / f;
    return xo_q(q.x * d, q.y * d, q.z * d, q.w * d);
}
static inline XB_HD XoQ xo_conj(XoQ q) { return xo_q(-q.x, -q.y, -q.z, q.w); }
static inline XB_HD XoQ xo_rotation(XoV3 axis, float angle) {    // RotationQuaternion(axis, angle)
    const XoV3 t = xo_normalize(axis);
    const float h = angle / 2.0f;
    const float s = xb_sinf(h);
    return xo_q(s * t.x, s * t.y, s * t.z, xb_cosf(h));
}
static inline XB_HD XoM xo_matrix(XoQ q) {                   // RotationMatrix(Quaternion)
    const float x = q.x, y = q.y, z = q.z, w = q.w;
    XoM m;
    m.m00 = 1.0f - 2.0f * (y * y + z * z), m.m01 = 2.0f * (x * y + z * w), m.m02 = 2.0f * (x * z - y * w);
    m.m10 = 2.0f * (x * y - z * w), m.m11 = 1.0f - 2.0f * (x * x + z * z), m.m12 = 2.0f * (y * z + x * w);
    m.m20 = 2.0f * (x * z + y * w), m.m21 = 2.0f * (y * z - x * w), m.m22 = 1.0f - 2.0f * (x * x + y * y);
    return m;
}
static inline XB_HD XoV3 xo_transform(XoV3 v, const XoM& m) {    // TransformCoord: every term kept (signs of zero)
    const float norm = ((0.0f * v.x + 0.0f * v.y) + 0.0f * v.z) + 1.0f;
    if (!(norm != 0.0f)) return xo_v(0.0f, 0.0f, 0.0f);
    return xo_v((((v.x * m.m00 + v.y * m.m10) + v.z * m.m20) + 0.0f) / norm,
                (((v.x * m.m01 + v.y * m.m11) + v.z * m.m21) + 0.0f) / norm,
                (((v.x * m.m02 + v.y * m.m12) + v.z * m.m22) + 0.0f) / norm);
}
static inline XB_HD XoV3 xo_heading(XoQ rot) { return xo_transform(xo_v(0.0f, 0.0f, 1.0f), xo_matrix(rot)); }

// RotationQuaternion(yaw, pitch, roll) of floats (DirectXMath's XMQuaternionRotationRollPitchYaw, scalar path;
// full.py:268-280): the half angles' glibc cosf and sinf, the four sums of products left to right.
static inline XB_HD XoQ xo_rotation_ypr(float yaw, float pitch, float roll) {
    const float hp = pitch * 0.5f, hy = yaw * 0.5f, hr = roll * 0.5f;
    const float cp = xb_cosf(hp), sp = xb_sinf(hp), cy = xb_cosf(hy), sy = xb_sinf(hy);
    const float cr = xb_cosf(hr), sr = xb_sinf(hr);
    return xo_q((cr * sp) * cy + (sr * cp) * sy, (cr * cp) * sy - (sr * sp) * cy, (sr * cp) * cy - (cr * sp) * sy,
                (cr * cp) * cy + (sr * sp) * sy);
}

/* ---- the searches (full.py:2133-2290) ---- */

struct XoShip {
    float max_ang_vel, t_scale, tf, fdt;
    double dt;
};

static inline XB_HD float xo_n2dt(int n, double dt) { return (float)((double)n * dt); }

/* CalculateN2Alpha2: the braking ticks n2, kappa2 = tf^n2, and the braking torque; the vector form */
static inline XB_HD void xo_n2_alpha2(const XoShip& k, XoV3 omega0, float len_omega0, int* n2, float* kappa2,
                                      XoV3* alpha2) {
    if ((double)len_omega0 > 1e-10) {
        const float factor = k.max_ang_vel / len_omega0;
        *n2 = xo_to_int32(__builtin_ceilf(xb_logf(factor / (1.0f + factor)) / (-k.fdt * k.t_scale)));
        *kappa2 = (float)xb_pow((double)k.tf, (double)*n2);
        *alpha2 = xo_mul(xo_v(-omega0.x, -omega0.y, -omega0.z), (k.t_scale * *kappa2) / (1.0f - *kappa2));
    } else {
        *n2 = 0, *kappa2 = 1.0f, *alpha2 = xo_v(0.0f, 0.0f, 0.0f);
    }
}

/* the float form (the roll) */
static inline XB_HD void xo_n2_alpha2_f(const XoShip& k, float omega0, int* n2, float* kappa2, float* alpha2) {
    const float len_omega0 = xo_abs(omega0);
    if ((double)len_omega0 > 1e-10) {
        const float factor = k.max_ang_vel / len_omega0;
        *n2 = xo_to_int32(__builtin_ceilf(xb_logf(factor / (1.0f + factor)) / (-k.fdt * k.t_scale)));
        *kappa2 = (float)xb_pow((double)k.tf, (double)*n2);
        *alpha2 = -omega0 * ((k.t_scale * *kappa2) / (1.0f - *kappa2));
    } else {
        *n2 = 0, *kappa2 = 1.0f, *alpha2 = 0.0f;
    }
}

/* GotoTorqueBreaks (full.py:2155-2222): the heading torque toward `direction`; `iters` counts the search's steps */
static inline XB_HD XoV3 xo_torque_breaks(const XoShip& k, XoV3 direction, XoQ rot, XoV3 w, int* iters) {
    const XoV3 zero = xo_v(0.0f, 0.0f, 0.0f);
    const float max_alpha = k.max_ang_vel * k.t_scale;
    if (max_alpha <= 0.0f) return zero;
    const float len_dir = xo_len(direction);
    if (len_dir == 0.0f) return zero;
    const XoV3 heading = xo_normalize(xo_heading(rot));
    XoV3 target = xo_div(xo_cross(heading, direction), len_dir);
    const float rot_len = xo_len(target);
    const float rot_angle = xb_acosf(xo_max(-1.0f, xo_min(1.0f, xo_dot(direction, heading) / len_dir)));
    if ((double)rot_len > 1e-12) target = xo_mul(target, rot_angle / rot_len);
    else if (rot_angle > 1.0f) target = xo_v(-rot_angle, 0.0f, 0.0f);
    const XoV3 omega0 = xo_sub(w, xo_mul(heading, xo_dot(heading, w)));
    const float len_omega0 = xo_len(omega0);
    if ((double)len_omega0 * k.dt < 1e-4 && (double)rot_angle < 1e-4) return zero;
    const float tau = 1.0f / k.t_scale;
    int n1 = 0, n2;
    XoV3 alpha1 = zero, alpha2;
    float kappa1 = 1.0f, kappa2;
    xo_n2_alpha2(k, omega0, len_omega0, &n2, &kappa2, &alpha2);
    for (int round = 0; round < 5; ++round) {
        if (n2 > 0) {
            const float n2dt = xo_n2dt(n2, k.dt);
            const XoV3 theta2 = xo_add(xo_mul(xo_mul(alpha2, tau), n2dt - tau * (1.0f - kappa2)),
                                       xo_mul(xo_mul(omega0, tau), 1.0f - kappa2));
            if ((double)xo_len(xo_sub(target, theta2)) < 1e-3) {
                n1 = 0;                                     // braking alone gets there
                break;
            }
        }
        kappa1 = 1.0f;
        n1 = 1;
        while (n1 < 100) {
            kappa1 = kappa1 * k.tf;
            const float n1dt = xo_n2dt(n1, k.dt);
            if (n2 == 0) {
                alpha1 = xo_div(xo_sub(target, xo_mul(omega0, tau * (1.0f - kappa1))),
                                tau * (n1dt - tau * (1.0f - kappa1)));
            } else {
                const float n2dt = xo_n2dt(n2, k.dt);
                const float num = tau - ((kappa1 * kappa2) * n2dt) / (1.0f - kappa2);
                const float den = tau * (n1dt - ((n2dt * kappa2) * (1.0f - kappa1)) / (1.0f - kappa2));
                alpha1 = xo_div(xo_sub(target, xo_mul(omega0, num)), den);
            }
            if (xo_len(alpha1) < max_alpha) break;
            n1 += 1;
        }
        *iters += n1;
        const int n2_old = n2;
        const XoV3 omega1 = xo_add(xo_mul(alpha1, tau * (1.0f - kappa1)), xo_mul(omega0, kappa1));
        xo_n2_alpha2(k, omega1, xo_len(omega1), &n2, &kappa2, &alpha2);
        if (n2_old == n2) break;
    }
    if (n1 > 0) return n1 > 1 ? xo_mul(xo_normalize(alpha1), max_alpha) : alpha1;
    if (n2 > 0) return alpha2;
    return zero;
}

/* GotoRollBreaks (full.py:2224-2290): the roll torque, into strong turns, otherwise upright; `iters` counts the search's
   steps */
static inline XB_HD float xo_roll_breaks(const XoShip& k, XoV3 heading_torque, XoQ rot, XoV3 w, int* iters) {
    const float max_alpha = k.t_scale * k.max_ang_vel;
    if (max_alpha <= 0.0f) return 0.0f;
    const XoM mat = xo_matrix(xo_conj(rot));
    const XoV3 ship_omega = xo_transform(w, mat);
    const float omega_roll = ship_omega.z;
    const float len_torque = xo_len(heading_torque);
    float turn_roll = 0.0f, weight = 0.0f;
    if ((double)len_torque > 0.05 * (double)max_alpha) {
        const XoV3 ship_torque = xo_transform(heading_torque, mat);
        const XoV3 flat = xo_v(ship_omega.x, ship_omega.y, 0.0f);
        if (xo_dot(flat, ship_torque) >= 0.0f || xo_len(flat) < 0.01f * k.max_ang_vel) {
            turn_roll = xb_atan2f(-ship_torque.y, -ship_torque.x);
            weight = len_torque / k.t_scale;
            if (weight > 1.0f) weight = 1.0f;
        }
    }
    const XoV3 up = xo_transform(xo_v(0.0f, 1.0f, 0.0f), mat);
    const float up_roll = up.z < 0.95f ? xb_atan2f(-up.x, up.y) : 0.0f;
    const float desired = weight < 0.05f ? up_roll : weight * turn_roll;
    if ((double)xo_abs(omega_roll) < 1e-4 && (double)xo_abs(desired) < 1e-4) return 0.0f;
    const float tau = 1.0f / k.t_scale;
    int n1 = 0, n2;
    float alpha1 = 0.0f, kappa1 = 1.0f, kappa2, alpha2;
    xo_n2_alpha2_f(k, omega_roll, &n2, &kappa2, &alpha2);
    for (int round = 0; round < 5; ++round) {
        if (n2 > 0) {
            const float n2dt = xo_n2dt(n2, k.dt);
            const float theta2 = (tau * alpha2) * (n2dt - tau * (1.0f - kappa2)) + (tau * omega_roll) * (1.0f - kappa2);
            if ((double)xo_abs(desired - theta2) < 1e-3) {
                n1 = 0;
                break;
            }
        }
        kappa1 = 1.0f;
        n1 = 1;
        while (n1 < 100) {
            kappa1 = kappa1 * k.tf;
            const float n1dt = xo_n2dt(n1, k.dt);
            if (n2 == 0) {
                alpha1 = (desired - omega_roll * (tau * (1.0f - kappa1))) / (tau * (n1dt - tau * (1.0f - kappa1)));
            } else {
                const float n2dt = xo_n2dt(n2, k.dt);
                const float num = tau - ((kappa1 * kappa2) * n2dt) / (1.0f - kappa2);
                const float den = tau * (n1dt - ((n2dt * kappa2) * (1.0f - kappa1)) / (1.0f - kappa2));
                alpha1 = (desired - omega_roll * num) / den;
            }
            if (xo_abs(alpha1) < max_alpha) break;
            n1 += 1;
        }
        *iters += n1;
        const int n2_old = n2;
        const float omega1 = (alpha1 * tau) * (1.0f - kappa1) + omega_roll * kappa1;
        xo_n2_alpha2_f(k, omega1, &n2, &kappa2, &alpha2);
        if (n2_old == n2) break;
    }
    if (n1 > 0) return n1 > 1 ? (alpha1 >= 0.0f ? 1.0f : -1.0f) * max_alpha : alpha1;
    if (n2 > 0) return alpha2;
    return 0.0f;
}

/* ApplyTorque (full.py:2292-2318) over one tick: the rotation and angular velocity a tick on, heading and roll apart */
static inline XB_HD void xo_apply_torque(XoQ rotation, XoV3 angvel, XoV3 torque, float roll, float tau, float fdt,
                                         XoQ* rot_out, XoV3* w_out) {
    const float tf = xb_expf(-fdt / tau);
    const XoV3 heading = xo_heading(rotation);
    const float omega_roll = xo_dot(angvel, heading);
    const XoV3 omega_heading = xo_sub(angvel, xo_mul(heading, omega_roll));
    const float span = fdt - tau * (1.0f - tf);
    const float rot_roll = tau * (roll * span + omega_roll * (1.0f - tf));
    const XoV3 rot_heading = xo_mul(xo_add(xo_mul(torque, span), xo_mul(omega_heading, 1.0f - tf)), tau);
    const float angle_heading = xo_len(rot_heading);
    const bool apply = (double)angle_heading > 1e-10;
    XoQ quat_heading = xo_q(0.0f, 0.0f, 0.0f, 1.0f);
    XoV3 heading_rotated = heading;
    if (apply) {
        quat_heading = xo_rotation(rot_heading, angle_heading);
        heading_rotated = xo_transform(heading, xo_matrix(quat_heading));
    }
    XoV3 w = xo_add(xo_mul(torque, tau * (1.0f - tf)), xo_mul(omega_heading, tf));
    w = xo_add(w, xo_mul(heading_rotated, tau * roll - (tau * roll - omega_roll) * tf));
    if (apply) rotation = xo_qmul(rotation, quat_heading);
    rotation = xo_qmul(rotation, xo_rotation(heading_rotated, rot_roll));
    const float norm = xo_qlen(rotation);
    if (norm > 0.0f) rotation = xo_qdiv(rotation, norm);
    *rot_out = rotation;
    *w_out = w;
}

/* EvolveBehaviorForBall's dynamical-orientation branch (full.py:2459-2485): above 1 % of the maximum acceleration the
   torque and roll are searched and the thrust laid along the heading a tick on, scaled by cos^2 of its angle (zero
   at 90 degrees or more); below it the thrust is kept, a ball in warp proper turning toward its exit point. The
   rotation and angular velocity a tick on (the same ApplyTorque the tick's commit makes) come back in rot1, w1; the
   two searches' steps in *iters (the tick's cost the scheduler learns, never a result). */
static inline XB_HD XbV3 xo_orient(const XbShip& s, XbTurn& t, XbV3 a, double friction, double dt, XoQ* rot1,
                                   XoV3* w1, int* iters) {
    const float accel = (float)v3_len(a);
    const float ff = (float)friction;
    const float fmass = (float)s.mass, ag = (float)s.agility, sf = (float)s.sf, mv = (float)s.max_vel;
    const float max_accel = ((ff * sf) * mv) / (fmass * ag);
    const bool thrust = (double)accel > 1e-2 * (double)max_accel;
    XoShip k;
    k.max_ang_vel = t.max_ang_vel;
    k.t_scale = ff / ((fmass * ag) * t.rot_agility);
    k.fdt = (float)dt;
    k.dt = dt;
    k.tf = xb_expf(-k.fdt * k.t_scale);
    const XoV3 a32 = xo_of(a);
    XoV3 torque = xo_v(0.0f, 0.0f, 0.0f);
    if (thrust) torque = xo_torque_breaks(k, a32, t.rot, t.w, iters);
    else if (s.mode == XB_MODE_WARP && !(s.effect_stamp < 0))
        torque = xo_torque_breaks(k, xo_of(v3_sub(s.gp, s.pos)), t.rot, t.w, iters);
    const float roll = xo_roll_breaks(k, torque, t.rot, t.w, iters);
    const float tau = ((fmass * ag) * t.rot_agility) / ff;
    xo_apply_torque(t.rot, t.w, torque, roll, tau, k.fdt, rot1, w1);
    t.torque = torque;
    t.roll = roll;
    if (!thrust) return a;
    const XoV3 heading = xo_heading(*rot1);
    const float cosangle = xo_min(1.0f, xo_dot(heading, a32) / accel);
    if (cosangle <= 0.0f) return v3(0.0, 0.0, 0.0);
    const XoV3 h = xo_mul(heading, (cosangle * cosangle) * accel);
    return v3((double)h.x, (double)h.y, (double)h.z);
}

/* The tick's commit of the rotation (full.py:2558-2563): the new values where finite, else the old kept (and the new
   ones put in the old members) */
static inline XB_HD void xo_commit(XbTurn& t, XoQ rot1, XoV3 w1) {
    const XoQ r0 = t.rot;
    const XoV3 w0 = t.w;
    const bool rot_ok = xo_finite(xo_qlensq(rot1)), w_ok = xo_finite(xo_dot(w1, w1));
    t.rot = rot_ok ? rot1 : r0;
    t.old_rot = rot_ok ? r0 : rot1;
    t.w = w_ok ? w1 : w0;
    t.old_w = w_ok ? w0 : w1;
}

/* SnapOrientation (full.py:2091-2110): turned to the velocity and rolled upright, the rotation stopped; nothing below
   a speed of 1e-6 m/s */
static inline XB_HD void xo_snap(XbV3 vel, XbTurn& t) {
    if (!(v3_lensq(vel) > 1e-12)) return;
    const XoV3 dir = xo_of(vel);
    const float len_dir = xo_len(dir);
    if ((double)len_dir < 1e-12) {
        t.rot = xo_q(0.0f, 0.0f, 0.0f, 1.0f);
    } else {
        const XoV3 vec = xo_cross(xo_v(0.0f, 0.0f, 1.0f), dir);
        const float angle = xb_acosf(xo_max(-1.0f, xo_min(1.0f, dir.z / len_dir)));
        XoQ rot = xo_rotation(vec, angle);
        const XoV3 up = xo_transform(xo_v(0.0f, 1.0f, 0.0f), xo_matrix(xo_conj(rot)));
        const float up_roll = up.z < 0.99f ? xb_atan2f(-up.x, up.y) : 0.0f;
        const XoV3 heading = xo_heading(rot);
        t.rot = xo_qmul(rot, xo_rotation(heading, up_roll));
    }
    t.old_rot = t.rot;
    t.w = t.old_w = xo_v(0.0f, 0.0f, 0.0f);
}

/* AddDynamicallyOrientedBall's old rotation of a new ball, rot x RotationQuaternion(angVel, -dt |angVel|)
   (full.py:1298-1300) */
static inline XB_HD XoQ xo_new_ball_old_rot(XoQ rot, XoV3 w, float fdt) {
    return xo_qmul(rot, xo_rotation(w, -fdt * xo_len(w)));
}
