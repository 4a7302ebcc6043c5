/* SPDX-License-Identifier: MIT
 *
 * One flight-modes tick of one ball row: destiny's Ballpark::Evolve (GOTO, STOP, FOLLOW, ORBIT under both laws, WARP
 * with its alignment, activation, warp path and drop-out, the release of a warper's followers, the closed-form
 * integrator), as the reference's batched park runs it per row (src/exact_ballpark/vec_modes.py:740-939 on tensors,
 * src/exact_ballpark/vec_modes_cuda.cu:540-717 on CUDA) and its scalar world defines it (src/exact_ballpark/modes.py:
 * 359-745, ref.py:151-211).
 *
 * A row's state is loaded into registers once (`xb_row_load`), ticked, and stored once (`xb_row_store`); a tick of a
 * group of whole worlds runs in three steps, each over every row of the group:
 *   A. `xb_record_a`: what the row shows the others in phase 1 -- its tick-start position and velocity, its id and
 *      radius, whether it starts warp proper at its own turn (`act`);
 *   B. `xb_record_b`: its vertical velocity after its own turn (STOP's damping, counting a release before its turn);
 *   C. `xb_tick_row`: the row's whole tick, reading other rows only through those records.
 * A row's follow target is a row of its own world, so a group of whole worlds needs nothing outside it: the GPU keeps
 * the records of a workgroup's worlds in LDS, the host in arrays.
 *
 * Exactness: every operation is one correctly rounded IEEE operation in the reference's order; files including this
 * are compiled with -ffp-contract=off and without fast-math, so nothing fuses except the maths header's explicit
 * xb_fma steps (glibc's own). Float fields are held in doubles and only read, compared and stored here.
 */
#ifndef XB_MODES_TICK_H
#define XB_MODES_TICK_H

#include <stdint.h>

#include "xb_layout.h"
#include "xb_maths_double.h"

#define XB_AU 0.1495978707e12                   /* Ballpark.h (modes.py:72) */
#define XB_ORBITAL_PRECESSION 0.001             /* modes.py:73 */
#define XB_WARP_AU_PER_SECOND 0.001             /* modes.py:74 */
#define XB_WARP_ACCELERATION (1.0 / 1000)       /* modes.py:75 */
#define XB_WARP_DECELERATION (1.0 / 3000)       /* modes.py:76 */
#define XB_MAX_ALIGN_TICKS 180                  /* modes.py:77 */
#define XB_STOP_DAMP 0.9345794392523364485981308411215   /* ref.py:37 */
#define XB_STOP_DECAY 0.07                      /* EvolveStop, ref.py:203 */
#define XB_TRUNC7 10000000.0                    /* modes.py:94-96 */
#define XB_RF_SCALE 40000.0                     /* modes.py:467 */
#define XB_GOTO_FAR (10.0 * XB_AU)              /* modes.py:477 */
#define XB_NORMAL_MIN 1e-12                     /* modes.py:503 */
#define XB_ANGLE_SCALE 0.1                      /* modes.py:519, 526 */
#define XB_ANGLE_ZERO 1e-10                     /* modes.py:531, 542 */
#define XB_FACTOR_FLOOR (-0.999)                /* modes.py:539 */
#define XB_LOG_FLOOR 0.001                      /* modes.py:541 */
#define XB_SLOW_TARGET 5e-2                     /* modes.py:560 */
#define XB_THRUST_OFF 0.1                       /* modes.py:598 */
#define XB_ALIGN_COS 0.01                       /* modes.py:348 */
#define XB_ALIGN_SPEED2 0.5625                  /* modes.py:349 */
#define XB_DROP_SPEED 100.0                     /* modes.py:650 */
#define XB_DIR_DIST 1.0e17                      /* modes.py:257 */
#define XB_TAYLOR_LIMIT 1e-10                   /* ref.py:168 */

/* ----------------------------------------------------------------------------------- double 3-vectors */

struct XbV3 {
    double x, y, z;
};

static inline XB_HD XbV3 v3(double x, double y, double z) { XbV3 v = {x, y, z}; return v; }
static inline XB_HD XbV3 v3_add(XbV3 a, XbV3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline XB_HD XbV3 v3_sub(XbV3 a, XbV3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline XB_HD XbV3 v3_scale(XbV3 a, double c) { return v3(a.x * c, a.y * c, a.z * c); }
static inline XB_HD double v3_lensq(XbV3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline XB_HD double v3_len(XbV3 a) { return __builtin_sqrt(v3_lensq(a)); }
static inline XB_HD XbV3 v3_unit(XbV3 a) {       /* Vector3d::Normalize: times 1/|a|; zero stays zero */
    double norm = v3_len(a);
    if (norm == 0.0) return a;
    return v3_scale(a, 1.0 / norm);
}
static inline XB_HD XbV3 v3_cross(XbV3 u, XbV3 v) {
    return v3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
static inline XB_HD bool xb_finite(double x) {
    return (xb_bits(x) & 0x7ff0000000000000ULL) != 0x7ff0000000000000ULL;
}
static inline XB_HD bool v3_finite(XbV3 a) { return xb_finite(a.x) && xb_finite(a.y) && xb_finite(a.z); }

/* (int64_t)x as aarch64 converts it: toward zero, saturating, NaN to 0 */
static inline XB_HD int64_t xb_to_int64(double x) {
    if (x != x) return 0;
    if (x >= 9223372036854775808.0) return INT64_MAX;
    if (x < -9223372036854775808.0) return INT64_MIN;
    return (int64_t)x;
}

static inline XB_HD double xb_trunc7(double x) { return (double)xb_to_int64(x * XB_TRUNC7) / XB_TRUNC7; }

/* ----------------------------------------------------------------------------------------- the park */

struct XbPark {
    double* f;
    int64_t* q;
    uint8_t* b;
    int64_t stride;
    double friction, dt;
    double k2, ook2, ook;          /* friction^2, 1/friction^2, 1/friction: the integrator's per-park constants */
    int new_orbit;
    int dynamical;                 /* dynamical orientation (MOV-728) */
    double* cf;                    /* the collision blocks (with the gradient method): float64 mLastC */
    int64_t* cq;                   /* int64 filing and the collision list's length */
    int64_t* hits;                 /* the collision lists, hits_cap rows */
    int hits_cap;
    int statics;                   /* the static collidables take part */
    const int32_t* st_range;       /* each row's world's first static collidable and how many (2 rows) */
    const double* st_f;            /* the static table (xb_statics.h) */
    const int64_t* st_q;
    const int64_t* st_keys;
    int64_t st_stride;
    int64_t tick_ms;               /* the tick interval in milliseconds (a missile's straight flight) */
    const double* form;            /* the formations' slot offsets (3 doubles a slot, form_stride slots a formation) */
    const int32_t* form_len;
    int n_form, form_stride;
    const int64_t* fld;            /* the force fields' keys [3][stride] (harmonic, corporation, alliance), or null */
    int64_t* fslots;               /* the master's formation reservations [stride] (a leader's bits), or null */
};

/* A row's tick-start records, as the others read them in phase 1 (indexed by the row's place in its group). */
struct XbRecords {
    uint8_t* act;
    double *p0x, *p0y, *p0z, *v0x, *v0y, *v0z, *vy_after;
    int64_t* id;
    double* radius;
    float* fq;                     /* with formations: each leader's rotation at the tick's start (4 floats a place) */
    int8_t* ffid;                  /* and its formation (-1 none) */
};

static inline XB_HD XbV3 xb_ld3(const XbPark& P, int f, int64_t r) {
    return v3(P.f[f * P.stride + r], P.f[(f + 1) * P.stride + r], P.f[(f + 2) * P.stride + r]);
}
static inline XB_HD void xb_st3(const XbPark& P, int f, int64_t r, XbV3 v) {
    P.f[f * P.stride + r] = v.x;
    P.f[(f + 1) * P.stride + r] = v.y;
    P.f[(f + 2) * P.stride + r] = v.z;
}
#define XB_F(P, fld, r) ((P).f[(int64_t)(fld) * (P).stride + (r)])
#define XB_Q(P, fld, r) ((P).q[(int64_t)(fld) * (P).stride + (r)])
#define XB_B(P, fld, r) ((P).b[(int64_t)(fld) * (P).stride + (r)])

/* One ball as its thread holds it during the tick. */
struct XbShip {
    XbV3 pos, vel, gp, last_g;
    double mass, tf, last_collision, radius, max_vel, agility, sf, follow_range;
    int64_t id, mode, effect_stamp, follow_id, owner_id, follow_ptr, clock;
    bool massive, interactive, sensor;
    int formation_id;              /* a leader's formation (-1 none) */
};

#include "xb_orient.h"

/* ------------------------------------------------------------------------------------------- modes */

/* GotoThrust (ref.py:151-160; full.py:2320-2331): toward the target at maxThrust, tapered by (L^2 / D^2)^2 within one
   tick's distance unless a missile's */
static inline XB_HD XbV3 xb_goto_thrust(const XbShip& s, XbV3 target, const XbPark& P, bool missile = false) {
    XbV3 a = v3_sub(target, s.pos);
    double length2 = v3_lensq(a);
    double dist = s.sf * s.max_vel * P.dt;
    double max_thrust = P.friction * s.sf * s.max_vel / (s.mass * s.agility);
    XbV3 unit = v3_unit(a);
    double dd = dist * dist;
    if (!missile && length2 < dd) {
        double coff = length2 / dd;
        return v3_scale(unit, max_thrust * coff * coff);
    }
    return v3_scale(unit, max_thrust);
}

#define XB_MISSILE_STRAIGHT_MS 800                /* Ballpark.cpp:974 (full.py:84) */

/* EvolveMissile (full.py:2337-2356): straight on toward mGoto while (clock - launch) x tick <= 800 ms; then at full
   thrust, untapered, toward the target's position predicted dist / maxVel ahead along its velocity (as this ball's turn
   sees it), offset by followRange + both radii back toward the missile (mGoto kept). */
static inline XB_HD XbV3 xb_missile(XbShip& s, bool target_ok, XbV3 op, XbV3 ov, double other_radius,
                                    const XbPark& P) {
    if ((s.clock - s.effect_stamp) * P.tick_ms <= XB_MISSILE_STRAIGHT_MS) return xb_goto_thrust(s, s.gp, P);
    if (!target_ok) return v3(0.0, 0.0, 0.0);
    const XbV3 delta = v3_sub(s.pos, op);
    const double dist = v3_len(delta);
    const double ct = dist / s.max_vel;
    XbV3 target = v3_add(op, v3_scale(ov, ct));
    const double r = s.follow_range + s.radius + other_radius;
    if (dist == 0.0) target = v3(op.x + 1.0 * r, op.y + 0.0 * r, op.z + 0.0 * r);
    else target = v3_add(target, v3_scale(v3_scale(delta, r), 1.0 / dist));
    s.gp = target;
    return xb_goto_thrust(s, target, P, true);
}

/* EvolveFollow (modes.py:428-442): the point at followRange + both radii from the target's tick-start position */
static inline XB_HD XbV3 xb_follow(XbShip& s, XbV3 op, double other_radius, const XbPark& P) {
    XbV3 delta = v3_sub(s.pos, op);
    double dist = v3_len(delta);
    double r = s.follow_range + s.radius + other_radius;
    if (dist == 0.0) {
        s.gp = v3(op.x + 1.0 * r, op.y + 0.0 * r, op.z + 0.0 * r);
    } else {
        double inv = 1.0 / dist;
        s.gp = v3(op.x + delta.x * r * inv, op.y + delta.y * r * inv, op.z + delta.z * r * inv);
    }
    return xb_goto_thrust(s, s.gp, P);
}

/* DefaultOrbitalNormal / the old orbit's w (modes.py:454-458, 480-486): truncated to 7 decimals, not normalised */
static inline XB_HD XbV3 xb_precessing(const XbShip& s) {
    double t = (double)s.clock;
    double phi1 = t * XB_ORBITAL_PRECESSION;
    double phi2 = (double)(s.id & 0xFFFFLL) + t * XB_ORBITAL_PRECESSION;
    double s1, c1, s2, c2;
    xb_sincos(phi1, &s1, &c1);
    xb_sincos(phi2, &s2, &c2);
    return v3(xb_trunc7(c1 * c2), xb_trunc7(s2), xb_trunc7(s1 * c2));
}

/* EvolveOldStyleOrbit (modes.py:444-478) */
static inline XB_HD XbV3 xb_old_orbit(XbShip& s, XbV3 op, double other_radius, const XbPark& P) {
    double cruise = s.sf * s.max_vel;
    double max_thrust = P.friction * cruise / (s.mass * s.agility);
    double r = s.follow_range + s.radius + other_radius;
    XbV3 to = v3_sub(op, s.pos);
    double dist = v3_len(to);
    to = v3_unit(to);
    XbV3 w = v3_unit(v3_cross(xb_precessing(s), to));
    double to_comp = dist * dist - r * r;
    if (to_comp >= 0.0) {
        double rad_comp = r * __builtin_sqrt(to_comp) / dist;
        double sc = to_comp / dist;
        to = v3_unit(v3(to.x * sc + w.x * rad_comp, to.y * sc + w.y * rad_comp, to.z * sc + w.z * rad_comp));
    }
    double dr = r - dist;
    double rf = xb_trunc7(xb_exp(-dr * dr / XB_RF_SCALE));
    double phi = (-to.x) * w.x + (-to.y) * w.y + (-to.z) * w.z;
    double tf = 1.0 + rf * rf * (phi * phi - 1.0);
    if (tf > 0.0) tf = rf * phi + __builtin_sqrt(tf);
    else tf = rf * phi;
    tf = tf * ((dist - r >= 0.0) ? 1.0 : -1.0);
    XbV3 a = v3((w.x * rf + to.x * tf) * max_thrust, (w.y * rf + to.y * tf) * max_thrust,
                (w.z * rf + to.z * tf) * max_thrust);
    s.gp = v3_add(s.pos, v3_scale(a, XB_GOTO_FAR));
    return a;
}

/* OrbitThrust (modes.py:586-601) */
static inline XB_HD XbV3 xb_orbit_thrust(const XbShip& s, XbV3 target, const XbPark& P) {
    double fom = P.friction / (s.mass * s.agility);
    double max_accel = fom * s.sf * s.max_vel;
    if (max_accel <= 0.0) return v3(0.0, 0.0, 0.0);
    XbV3 a = v3((target.x - s.pos.x) * fom - s.vel.x, (target.y - s.pos.y) * fom - s.vel.y,
                (target.z - s.pos.z) * fom - s.vel.z);
    double tacc = v3_len(a) / max_accel;
    if (tacc > 1.0) a = v3_scale(a, 1.0 / tacc);
    if (tacc < XB_THRUST_OFF) return v3(0.0, 0.0, 0.0);
    return a;
}

/* GotoThrustFollow with weight 1 (modes.py:557-584); the new orbit is its only caller */
static inline XB_HD XbV3 xb_goto_thrust_follow(const XbShip& s, XbV3 target, XbV3 tv, const XbPark& P) {
    double speed = v3_len(tv);
    if (speed < XB_SLOW_TARGET * s.max_vel * s.sf) {
        if (P.new_orbit && s.mode == XB_MODE_ORBIT) return xb_orbit_thrust(s, target, P);
        return xb_goto_thrust(s, s.gp, P);
    }
    double fom = P.friction / (s.mass * s.agility);
    double max_accel = fom * s.sf * s.max_vel;
    if (max_accel <= 0.0) return v3(0.0, 0.0, 0.0);
    double one_minus = 1.0 - s.tf;
    double inv = 1.0 / (P.dt * fom - one_minus);
    XbV3 a = v3(((target.x - s.pos.x) * fom - s.vel.x * one_minus) * fom * inv,
                ((target.y - s.pos.y) * fom - s.vel.y * one_minus) * fom * inv,
                ((target.z - s.pos.z) * fom - s.vel.z * one_minus) * fom * inv);
    inv = 1.0 / (1.0 - s.tf);
    XbV3 av = v3((tv.x - s.vel.x * s.tf) * fom * inv, (tv.y - s.vel.y * s.tf) * fom * inv,
                 (tv.z - s.vel.z * s.tf) * fom * inv);
    const double weight = 1.0;
    inv = 1.0 / (1.0 + weight);
    a = v3((a.x + av.x * weight) * inv, (a.y + av.y * weight) * inv, (a.z + av.z * weight) * inv);
    double accel = v3_lensq(a);
    if (accel > max_accel * max_accel) a = v3_scale(a, max_accel / __builtin_sqrt(accel));
    return a;
}

/* EvolveNewStyleOrbit with GetOrbitalNormal (modes.py:488-555); ov: the target's velocity as this ball sees it at its
   own turn; rot: with dynamical orientation the ball's tick-start rotation, whose heading gives the normal when the
   velocity's does not (full.py:2385-2414), else NULL */
static inline XB_HD XbV3 xb_new_orbit(XbShip& s, XbV3 op, XbV3 ov, double other_radius, const XbPark& P,
                                      const XoQ* rot) {
    double cruise = s.sf * s.max_vel;
    double tau = (s.mass * s.agility) / P.friction;
    double max_thrust = cruise / tau;
    double r = s.follow_range + s.radius + other_radius;
    XbV3 to = v3_sub(op, s.pos);
    to = v3(to.x + ov.x * P.dt, to.y + ov.y * P.dt, to.z + ov.z * P.dt);
    double dist = v3_len(to);
    if (dist > 0.0) to = v3_scale(to, 1.0 / dist);
    else to = v3(0.0, 0.0, 1.0);
    XbV3 o = v3_cross(to, s.vel);
    double length2 = v3_lensq(o);
    XbV3 normal;
    if (length2 > XB_NORMAL_MIN) {
        normal = v3_scale(o, 1.0 / __builtin_sqrt(length2));
    } else if (!rot) {
        normal = xb_precessing(s);
    } else {
        const XoV3 h = xo_heading(*rot);
        o = v3_cross(to, v3((double)h.x, (double)h.y, (double)h.z));
        length2 = v3_lensq(o);
        normal = (length2 <= XB_NORMAL_MIN) ? xb_precessing(s) : v3_scale(o, 1.0 / __builtin_sqrt(length2));
    }
    XbV3 radial = v3_unit(v3_cross(normal, to));
    double speed = v3_len(s.vel);
    double max_speed = (cruise < speed) ? speed : cruise;
    double angle = XB_ANGLE_SCALE * (r - dist) / max_speed;
    if (angle < -1.0) angle = -1.0;
    if (angle > 1.0) angle = 1.0;
    if (dist < XB_ANGLE_SCALE) dist = XB_ANGLE_SCALE;
    double sq = __builtin_sqrt(dist * max_thrust);
    double est_v = (sq < max_speed) ? sq : max_speed;
    double est_d = est_v * P.dt;
    XbV3 gp, gv;
    double st, ct;
    if (xb_fabs(angle) < XB_ANGLE_ZERO) {
        double t = est_d / dist;
        xb_sincos(t, &st, &ct);
        double cp = dist * (1.0 - ct), sp = dist * st, cv = ct * est_v, sv = st * est_v;
        gp = v3(to.x * cp + radial.x * sp, to.y * cp + radial.y * sp, to.z * cp + radial.z * sp);
        gv = v3(radial.x * cv + to.x * sv, radial.y * cv + to.y * sv, radial.z * cv + to.z * sv);
    } else {
        double factor = est_d * angle / dist / __builtin_sqrt(angle * angle + 1.0);
        double t;
        if (factor < XB_FACTOR_FLOOR) t = xb_log(XB_LOG_FLOOR) / angle;
        else if (xb_fabs(factor) > XB_ANGLE_ZERO) t = xb_log(1.0 + factor) / angle;
        else t = factor / angle;
        double dexp = dist * xb_exp(angle * t);
        xb_sincos(t, &st, &ct);
        double cp = dist - dexp * ct, sp = dexp * st;
        gp = v3(to.x * cp + radial.x * sp, to.y * cp + radial.y * sp, to.z * cp + radial.z * sp);
        double ca = -angle * ct + st, sa = angle * st + ct;
        gv = v3_scale(v3_unit(v3(to.x * ca + radial.x * sa, to.y * ca + radial.y * sa, to.z * ca + radial.z * sa)),
                      est_v);
    }
    s.gp = v3_add(s.pos, gp);
    return xb_goto_thrust_follow(s, s.gp, v3_add(gv, ov), P);
}

/* ------------------------------------------------------------------------------------------------ warp */

/* IsAlignedForWarp of an aligning ball (modes.py:342-355) */
static inline XB_HD bool xb_aligned(XbV3 gp, XbV3 pos, XbV3 vel, double max_vel, int64_t effect_stamp) {
    XbV3 d = v3_unit(v3_sub(gp, pos));
    XbV3 u = v3_unit(vel);
    if (xb_fabs(1.0 - (u.x * d.x + u.y * d.y + u.z * d.z)) < XB_ALIGN_COS &&
        v3_lensq(vel) > XB_ALIGN_SPEED2 * max_vel * max_vel)
        return true;
    return (effect_stamp < 0 ? -effect_stamp : effect_stamp) > XB_MAX_ALIGN_TICKS;
}

/* RealWarp, the warper's own part (modes.py:391-403): the exit point min range short of the goto point */
static inline XB_HD void xb_real_warp(XbShip& s) {
    XbV3 dst = s.gp;
    XbV3 delta = v3_unit(v3_sub(s.pos, dst));
    double mr = xb_dbl((uint64_t)s.follow_id);
    dst = v3_add(dst, v3_scale(delta, mr));
    delta = v3_sub(dst, s.pos);
    s.gp = dst;
    s.effect_stamp = s.clock;
    s.last_collision = v3_len(delta);
    s.massive = false;
    s.sensor = false;
}

/* SetupWarpConstants and WarpDistance (modes.py:605-666): the new position and velocity on the warp line toward the
   exit point; drops the ball out of warp (massive and sensing again, STOP) when slow enough or past deceleration */
static inline XB_HD void xb_warp_path(XbShip& s, XbV3* pp, XbV3* vp, double t) {
    double wf = (double)s.owner_id;
    double wd = s.last_collision;
    double sp = wf * XB_WARP_AU_PER_SECOND * XB_AU;
    double acc = wf * XB_WARP_ACCELERATION;
    double d0 = wf * XB_WARP_DECELERATION;
    double dec = (2.0 < d0) ? 2.0 : d0;
    double cap = (wd + 1.0) * acc * dec / (acc + dec);
    sp = (cap < sp) ? cap : sp;
    double t_a = xb_log(sp / acc) / acc;
    double t_c = (wd / sp) - (1.0 / acc) - (1.0 / dec) + (1.0 / sp);
    double t_d = xb_log(sp / dec) / dec;
    double d_a = sp / acc, d_c = sp * t_c, d_d = sp / dec - 1.0;
    XbV3 d = v3_unit(v3_sub(s.gp, *pp));
    double total = d_a + d_c + d_d;
    bool drop = false;
    if (t < t_a) {
        double e = xb_exp(acc * t);
        double speed = acc * e;
        double cur = v3_len(*vp);
        if (cur > speed) speed = cur;
        *vp = v3_scale(d, speed);
        *pp = v3_sub(s.gp, v3_scale(d, total - e));
    } else if ((t - t_a) < t_c) {
        *vp = v3_scale(d, sp);
        *pp = v3_sub(s.gp, v3_scale(d, total - (sp * (t - t_a) + d_a)));
    } else if ((t - t_c - t_a) < (t_d + 1.0)) {
        double e = xb_exp(-dec * (t - t_c - t_a));
        double speed = sp * e;
        *vp = v3_scale(d, speed);
        double sd = sp / dec;
        *pp = v3_sub(s.gp, v3_scale(d, total - (sd - sd * e + d_a + d_c)));
        double half = s.max_vel / 2.0;
        drop = speed < ((XB_DROP_SPEED < half) ? XB_DROP_SPEED : half);
    } else {
        *pp = s.gp;
        *vp = v3(0.0, 0.0, 0.0);
        drop = true;
    }
    if (drop) {
        s.massive = true;
        s.sensor = true;
        s.mode = XB_MODE_STOP;
    }
}

/* ------------------------------------------------------------------------- StopAllFollowers, Stop(Ball*) */

static inline XB_HD void xb_clear_follow(XbShip& s) {
    s.effect_stamp = 0;
    s.follow_ptr = -1;
    s.follow_id = 0;
    s.owner_id = 0;
    s.follow_range = 0.0;
}

/* FreeFormationSlot((size_t) mEffectStamp) of a FORMATION follower's Stop(Ball*) on the master (Ballpark.cpp:4629-4633,
   full.py:1850-1857; the reference's kernel vec_combat_cuda.cu:200-224): its slot's bit cleared in its leader's
   reservations, nothing without the leader's loaded formation or for a slot past its end (an unsigned comparison).
   Followers of one leader may stop in one launch: the clear is an atomic AND, which commutes. A slot of 16 or more
   (Destiny's bitset<16> throws) is refused before it can be reached: FormationFollow on the master takes slots 0-15. */
static inline XB_HD void xb_free_formation_slot(const XbPark& P, int64_t leader, int64_t slot) {
    if (!P.fslots || leader < 0) return;
    const int fid = (int)(int8_t)P.b[XB_B_FORMATION * P.stride + leader];
    if (fid < 0 || fid > P.n_form - 1) return;
    const uint64_t s = (uint64_t)slot;
    if (s > (uint64_t)(int64_t)(P.form_len[fid] - 1) || s >= 16) return;
    const uint64_t keep = ~((uint64_t)1 << s);
#if defined(__HIP_DEVICE_COMPILE__)
    atomicAnd(reinterpret_cast<unsigned long long*>(P.fslots + leader), (unsigned long long)keep);
#else
    __atomic_fetch_and(reinterpret_cast<uint64_t*>(P.fslots + leader), keep, __ATOMIC_RELAXED);
#endif
}

/* The release of one follower of a ball that starts warp proper (modes.py:410-426): an interactive orbiter flies on
   by GotoDirection along its velocity (refused, so left as it is, when the velocity or the point is not finite), any
   other member stops (Stop(Ball*) clears the follow fields of a FOLLOW, ORBIT or FORMATION ball). Returns the mode it
   leaves. */
static inline XB_HD int64_t xb_release(XbShip& s, bool apply, const XbPark* P = nullptr) {
    if (s.mode == XB_MODE_MISSILE || (s.interactive && s.mode == XB_MODE_ORBIT)) {
        if (apply && s.mode == XB_MODE_MISSILE) s.massive = false;    /* a missile flies on, non-massive */
        if (!v3_finite(s.vel)) return s.mode;
        XbV3 p = v3_add(s.pos, v3_scale(v3_unit(s.vel), XB_DIR_DIST));
        if (!v3_finite(p)) return s.mode;
        if (apply) {
            xb_clear_follow(s);
            s.gp = p;
            if (s.sf == 0.0) s.sf = 1.0;
            s.mode = XB_MODE_GOTO;
        }
        return XB_MODE_GOTO;
    }
    if (apply) {
        if (P && s.mode == XB_MODE_FORMATION) xb_free_formation_slot(*P, s.follow_ptr, s.effect_stamp);
        if (s.mode == XB_MODE_FOLLOW || s.mode == XB_MODE_ORBIT || s.mode == XB_MODE_FORMATION) xb_clear_follow(s);
        s.mode = XB_MODE_STOP;
    }
    return XB_MODE_STOP;
}

/* ---------------------------------------------------------------------------------------- integrator */

/* Ballpark::Integrate over one tick (ref.py:162-176 with t = dt, the cached decay factor) */
static inline XB_HD void xb_integrate(XbV3* pp, XbV3* vp, XbV3 a, double m, double tf, const XbPark& P) {
    const double k = P.friction, t = P.dt;
    if (t == 0.0) return;
    XbV3 p = *pp, v = *vp;
    XbV3 pn;
    if (k < XB_TAYLOR_LIMIT * m * t) {
        double inv_m = 1.0 / m;
        double kt = -k * t;
        pn = v3((m * (a.x * kt + k * (a.x * t + v.x * k * inv_m * t)) + p.x * P.k2) * P.ook2,
                (m * (a.y * kt + k * (a.y * t + v.y * k * inv_m * t)) + p.y * P.k2) * P.ook2,
                (m * (a.z * kt + k * (a.z * t + v.z * k * inv_m * t)) + p.z * P.k2) * P.ook2);
    } else {
        double tf1 = tf - 1.0;
        pn = v3((m * (m * a.x * tf1 + k * (a.x * t + v.x - tf * v.x)) + p.x * P.k2) * P.ook2,
                (m * (m * a.y * tf1 + k * (a.y * t + v.y - tf * v.y)) + p.y * P.k2) * P.ook2,
                (m * (m * a.z * tf1 + k * (a.z * t + v.z - tf * v.z)) + p.z * P.k2) * P.ook2);
    }
    *vp = v3((m * a.x - (m * a.x - v.x * k) * tf) * P.ook, (m * a.y - (m * a.y - v.y * k) * tf) * P.ook,
             (m * a.z - (m * a.z - v.z * k) * tf) * P.ook);
    *pp = pn;
}

/* ------------------------------------------------------------------------------------ the three steps */


/* A row's whole state as its lane holds it across the ticks of one launch: loaded once, ticked in registers, stored
   once. The flags are the row's own (free, moribund, interactive, massive, sensor). */
struct XbRow {
    XbShip s;
    XbTurn t;                      /* with dynamical orientation: the turning state */
    bool turned;                   /* the turning state was committed (stored back then) */
    XbV3 old_pos, old_vel;
    bool free_, moribund;
    bool grown;                    /* a mushroom's turn changed its radius or removed it (stored back then) */
    uint8_t path;                  /* the path class its last tick executed (XB_PATH_*) */
    uint8_t iters;                 /* its orientation searches' steps in its last tick, at most 255 */
    int64_t row;                   /* its physical row */
};

/* Path classes: what a row's tick executes beyond the integrator, from cheapest to dearest in the double functions it
   calls. The built-in scheduling key predicts the class from the tick-start state; the tick records the one it ran. */
#define XB_PATH_IDLE 0             /* not free, or moribund: nothing */
#define XB_PATH_COAST 1            /* STOP, inert modes, a troll: the integrator alone */
#define XB_PATH_GOTO 2             /* GOTO, WARP aligning: GotoThrust */
#define XB_PATH_FOLLOW 3           /* FOLLOW */
#define XB_PATH_ORBIT_OLD 4        /* the old orbit: two sincos, exp */
#define XB_PATH_ORBIT_NEW 5        /* the new orbit: sincos, exp, log */
#define XB_PATH_WARP 6             /* warp proper: the warp path, two logs and an exp */
#define XB_PATH_WARP_EDGE 7        /* activation (RealWarp) or dropping out: the warp path and more */

static inline XB_HD int xb_branch_key(int64_t mode, int64_t effect_stamp, bool live, int new_orbit) {
    if (!live) return XB_PATH_IDLE;
    switch (mode) {
        case XB_MODE_GOTO: return XB_PATH_GOTO;
        case XB_MODE_FOLLOW: return XB_PATH_FOLLOW;
        case XB_MODE_ORBIT: return new_orbit ? XB_PATH_ORBIT_NEW : XB_PATH_ORBIT_OLD;
        case XB_MODE_WARP: return effect_stamp < 0 ? XB_PATH_GOTO : XB_PATH_WARP;
        default: return XB_PATH_COAST;
    }
}

static inline XB_HD void xb_row_load(const XbPark& P, int64_t r, XbRow& w) {
    XbShip& s = w.s;
    s.pos = xb_ld3(P, XB_F_POS, r);
    s.vel = xb_ld3(P, XB_F_VEL, r);
    s.gp = xb_ld3(P, XB_F_GOTO, r);
    s.last_g = xb_ld3(P, XB_F_LAST_G, r);
    w.old_pos = xb_ld3(P, XB_F_OLD_POS, r);
    w.old_vel = xb_ld3(P, XB_F_OLD_VEL, r);
    s.mass = XB_F(P, XB_F_MASS, r);
    s.tf = XB_F(P, XB_F_TIME_FACTOR, r);
    s.last_collision = XB_F(P, XB_F_LAST_COLLISION, r);
    s.radius = XB_F(P, XB_F_RADIUS, r);
    s.max_vel = XB_F(P, XB_F_MAX_VEL, r);
    s.agility = XB_F(P, XB_F_AGILITY, r);
    s.sf = XB_F(P, XB_F_SPEED_FRACTION, r);
    s.follow_range = XB_F(P, XB_F_FOLLOW_RANGE, r);
    s.id = XB_Q(P, XB_I_IDS, r);
    s.mode = XB_Q(P, XB_I_MODE, r);
    s.effect_stamp = XB_Q(P, XB_I_EFFECT_STAMP, r);
    s.follow_id = XB_Q(P, XB_I_FOLLOW_ID, r);
    s.owner_id = XB_Q(P, XB_I_OWNER_ID, r);
    s.follow_ptr = XB_Q(P, XB_I_FOLLOW_PTR, r);
    s.clock = XB_Q(P, XB_I_CLOCK, r);
    w.free_ = XB_B(P, XB_B_FREE, r) != 0;
    w.moribund = XB_B(P, XB_B_MORIBUND, r) != 0;
    s.massive = XB_B(P, XB_B_MASSIVE, r) != 0;
    s.interactive = XB_B(P, XB_B_INTERACTIVE, r) != 0;
    s.sensor = XB_B(P, XB_B_SENSOR, r) != 0;
    s.formation_id = (int)(int8_t)XB_B(P, XB_B_FORMATION, r);
    w.path = XB_PATH_IDLE;
    w.iters = 0;
    w.row = r;
    w.turned = false;
    w.grown = false;
}

/* Every field the tick may change (mass, agility, maximum speed, decay, ids: never; the radius and the moribund flag
   only at a mushroom's turn) */
static inline XB_HD void xb_row_store(const XbPark& P, const XbRow& w) {
    const int64_t r = w.row;
    const XbShip& s = w.s;
    xb_st3(P, XB_F_POS, r, s.pos);
    xb_st3(P, XB_F_VEL, r, s.vel);
    xb_st3(P, XB_F_GOTO, r, s.gp);
    xb_st3(P, XB_F_LAST_G, r, s.last_g);
    xb_st3(P, XB_F_OLD_POS, r, w.old_pos);
    xb_st3(P, XB_F_OLD_VEL, r, w.old_vel);
    XB_F(P, XB_F_LAST_COLLISION, r) = s.last_collision;
    XB_F(P, XB_F_SPEED_FRACTION, r) = s.sf;
    XB_F(P, XB_F_FOLLOW_RANGE, r) = s.follow_range;
    XB_Q(P, XB_I_MODE, r) = s.mode;
    XB_Q(P, XB_I_EFFECT_STAMP, r) = s.effect_stamp;
    XB_Q(P, XB_I_FOLLOW_ID, r) = s.follow_id;
    XB_Q(P, XB_I_OWNER_ID, r) = s.owner_id;
    XB_Q(P, XB_I_FOLLOW_PTR, r) = s.follow_ptr;
    XB_Q(P, XB_I_CLOCK, r) = s.clock;
    XB_B(P, XB_B_FREE, r) = w.free_ ? 1 : 0;
    XB_B(P, XB_B_MASSIVE, r) = s.massive ? 1 : 0;
    XB_B(P, XB_B_INTERACTIVE, r) = s.interactive ? 1 : 0;
    XB_B(P, XB_B_SENSOR, r) = s.sensor ? 1 : 0;
    if (w.grown) {
        XB_F(P, XB_F_RADIUS, r) = s.radius;
        XB_B(P, XB_B_MORIBUND, r) = w.moribund ? 1 : 0;
    }
}

/* The turning state (dynamical orientation): floats held in the float64 block rows */
static inline XB_HD void xb_turn_load(const XbPark& P, int64_t r, XbTurn& t) {
    t.rot = xo_q((float)XB_F(P, XB_F_ROT, r), (float)XB_F(P, XB_F_ROT + 1, r), (float)XB_F(P, XB_F_ROT + 2, r),
                 (float)XB_F(P, XB_F_ROT + 3, r));
    t.w = xo_v((float)XB_F(P, XB_F_ANG_VEL, r), (float)XB_F(P, XB_F_ANG_VEL + 1, r),
               (float)XB_F(P, XB_F_ANG_VEL + 2, r));
    t.max_ang_vel = (float)XB_F(P, XB_F_MAX_ANG_VEL, r);
    t.rot_agility = (float)XB_F(P, XB_F_ROT_AGILITY, r);
}
static inline XB_HD void xb_turn_store(const XbPark& P, int64_t r, const XbTurn& t) {
    XB_F(P, XB_F_ROT, r) = t.rot.x, XB_F(P, XB_F_ROT + 1, r) = t.rot.y;
    XB_F(P, XB_F_ROT + 2, r) = t.rot.z, XB_F(P, XB_F_ROT + 3, r) = t.rot.w;
    XB_F(P, XB_F_OLD_ROT, r) = t.old_rot.x, XB_F(P, XB_F_OLD_ROT + 1, r) = t.old_rot.y;
    XB_F(P, XB_F_OLD_ROT + 2, r) = t.old_rot.z, XB_F(P, XB_F_OLD_ROT + 3, r) = t.old_rot.w;
    XB_F(P, XB_F_ANG_VEL, r) = t.w.x, XB_F(P, XB_F_ANG_VEL + 1, r) = t.w.y, XB_F(P, XB_F_ANG_VEL + 2, r) = t.w.z;
    XB_F(P, XB_F_OLD_ANG_VEL, r) = t.old_w.x, XB_F(P, XB_F_OLD_ANG_VEL + 1, r) = t.old_w.y;
    XB_F(P, XB_F_OLD_ANG_VEL + 2, r) = t.old_w.z;
    XB_F(P, XB_F_TORQUE, r) = t.torque.x, XB_F(P, XB_F_TORQUE + 1, r) = t.torque.y;
    XB_F(P, XB_F_TORQUE + 2, r) = t.torque.z;
    XB_F(P, XB_F_ROT_ROLL, r) = t.roll;
}

static inline XB_HD bool xb_row_live(const XbRow& w) { return w.free_ && !w.moribund; }

#include "xb_ypr.h"

/* Step A for the row at place j of its group */
static inline XB_HD void xb_record_a(const XbRow& w, const XbRecords& R, int j) {
    const XbShip& s = w.s;
    bool act = xb_row_live(w) && s.mode == XB_MODE_WARP && s.effect_stamp < 0 &&
               xb_aligned(s.gp, s.pos, s.vel, s.max_vel, s.effect_stamp);
    R.act[j] = act ? 1 : 0;
    R.p0x[j] = s.pos.x, R.p0y[j] = s.pos.y, R.p0z[j] = s.pos.z;
    R.v0x[j] = s.vel.x, R.v0y[j] = s.vel.y, R.v0z[j] = s.vel.z;
    R.id[j] = s.id;
    R.radius[j] = s.radius;
}

/* Step A with formations: the row's formation and, for a leader, its rotation at the tick's start: with dynamical
   orientation its own, else RotationQuaternion of its display yaw, pitch and roll as floats (`y`; Ball::
   GetRotatedVector, full.py:2112-2117). */
template <bool DO>
static inline XB_HD void xb_record_form(const XbRow& w, const XbRecords& R, int j, const XbYpr* y) {
    const int fid = w.s.formation_id;
    R.ffid[j] = (int8_t)fid;
    if (fid < 0) return;
    XoQ q = xo_q(0.0f, 0.0f, 0.0f, 1.0f);
    if (DO) q = w.t.rot;
    else if (y) q = xo_rotation_ypr((float)y->yaw, (float)y->pitch, (float)y->roll);
    R.fq[4 * j] = q.x, R.fq[4 * j + 1] = q.y, R.fq[4 * j + 2] = q.z, R.fq[4 * j + 3] = q.w;
}

/* EvolveFormation (full.py:2359-2382) for a FORMATION row whose leader is at place tp: the leader's slot offset
   (slot: the row's effect stamp as a char), scaled to both radii plus its length, as a float vector turned by the
   leader's rotation, added to the leader's position; GotoThrust toward it (mGoto kept). No thrust without the leader's
   formation or past its slots. */
static inline XB_HD XbV3 xb_formation(XbShip& s, const XbRecords& R, int tp, const XbPark& P) {
    const int fid = R.ffid[tp];
    if (fid < 0 || fid > P.n_form - 1) return v3(0.0, 0.0, 0.0);
    const int slot = (int)(int8_t)(s.effect_stamp & 0xFF);
    if (slot > P.form_len[fid] - 1 || slot < 0) return v3(0.0, 0.0, 0.0);
    const double* o = P.form + 3 * ((int64_t)fid * P.form_stride + slot);
    const XbV3 offset = v3(o[0], o[1], o[2]);
    double len = v3_len(offset);
    if (len == 0.0) len = 1.0;
    const double sc = s.radius + R.radius[tp] + len;
    const double inv = 1.0 / len;
    const XoV3 t = xo_v((float)((offset.x * sc) * inv), (float)((offset.y * sc) * inv), (float)((offset.z * sc) * inv));
    const XoQ q = xo_q(R.fq[4 * tp], R.fq[4 * tp + 1], R.fq[4 * tp + 2], R.fq[4 * tp + 3]);
    const XoV3 v = xo_transform(t, xo_matrix(q));
    const XbV3 op = v3(R.p0x[tp] + (double)v.x, R.p0y[tp] + (double)v.y, R.p0z[tp] + (double)v.z);
    s.gp = op;
    return xb_goto_thrust(s, op, P);
}

/* The place of the row's follow target in its group (a world's rows keep their order and spacing in the group), or -1
   when it has none or the target is not in the follower set as StopAllFollowers tests it (its id is the follow id) */
static inline XB_HD int xb_member_place(const XbRow& w, const XbRecords& R, int j, int lo, int hi) {
    int64_t t = w.s.follow_ptr;
    if (t < 0) return -1;
    int64_t tj = (int64_t)j + (t - w.row);
    if (tj < lo || tj >= hi) return -1;
    return R.id[tj] == w.s.follow_id ? (int)tj : -1;
}

/* Step B: the vertical velocity the row shows after its own turn (STOP's damping, without dynamical orientation) */
static inline XB_HD void xb_record_b(const XbRow& w, const XbRecords& R, int j, int lo, int hi, const XbPark& P) {
    double vy = w.s.vel.y;
    if (xb_row_live(w)) {
        int64_t m = w.s.mode;
        int tj = xb_member_place(w, R, j, lo, hi);
        if (tj >= 0 && R.id[tj] < w.s.id && R.act[tj]) {
            XbShip f;
            f.mode = m;
            f.interactive = w.s.interactive;
            f.pos = w.s.pos;
            f.vel = w.s.vel;
            m = xb_release(f, false);
        }
        if (m == XB_MODE_STOP && !P.dynamical) vy = (vy - XB_STOP_DECAY * vy) * XB_STOP_DAMP;
    }
    R.vy_after[j] = vy;
}

/* Step C: the row's whole tick (phase 1, phases 3 and 4, its clock); the group's places are [lo, hi). Other rows are
   read only through the records. */

/* Phase 1 of the row at place j (Ballpark::Evolve's behaviours, modes.py:707-716): a release when the target starts
   warp proper, before or after this ball's own turn by their ids; the behaviour (mLastG; with dynamical orientation
   the rotation a tick on, into rot1 and w1); mLastC cleared. */
template <bool DO>
static inline XB_HD void xb_phase1(XbRow& w, const XbRecords& R, int j, int lo, int hi, const XbPark& P, XoQ& rot1,
                                   XoV3& w1) {
    XbShip& s = w.s;
    const bool live = xb_row_live(w);
    const bool act_self = R.act[j] != 0;
    bool rel_before = false, rel_after = false;
    int tj = xb_member_place(w, R, j, lo, hi);
    if (tj >= 0 && R.act[tj]) {
        if (R.id[tj] < s.id) rel_before = true;
        else rel_after = true;
    }
    if (rel_before) xb_release(s, true, &P);
    w.path = live ? XB_PATH_COAST : XB_PATH_IDLE;
    if (live) {
        XbV3 a = v3(0.0, 0.0, 0.0);
        int64_t t = s.follow_ptr;
        int64_t tp = (t >= 0) ? (int64_t)j + (t - w.row) : -1;
        bool target_ok = t >= 0 && tp >= lo && tp < hi;
        if (s.mode == XB_MODE_WARP) {
            if (s.effect_stamp >= 0) {
                w.path = XB_PATH_WARP;
            } else if (act_self) {
                xb_real_warp(s);
                w.path = XB_PATH_WARP_EDGE;
            } else {
                s.effect_stamp -= 1;
                a = xb_goto_thrust(s, s.gp, P);
                w.path = XB_PATH_GOTO;
            }
        } else if (s.mode == XB_MODE_GOTO) {
            a = xb_goto_thrust(s, s.gp, P);
            w.path = XB_PATH_GOTO;
        } else if (s.mode == XB_MODE_FORMATION) {
            if (target_ok && R.ffid) a = xb_formation(s, R, (int)tp, P);
            w.path = XB_PATH_FOLLOW;
        } else if (s.mode == XB_MODE_MISSILE) {
            XbV3 op = v3(0.0, 0.0, 0.0), ov = op;
            double orad = 0.0;
            if (target_ok) {
                op = v3(R.p0x[tp], R.p0y[tp], R.p0z[tp]);
                ov = v3(R.v0x[tp], (R.id[tp] < s.id) ? R.vy_after[tp] : R.v0y[tp], R.v0z[tp]);
                orad = R.radius[tp];
            }
            a = xb_missile(s, target_ok, op, ov, orad, P);
            w.path = XB_PATH_FOLLOW;
        } else if ((s.mode == XB_MODE_FOLLOW || s.mode == XB_MODE_ORBIT) && !target_ok) {
        } else if (s.mode == XB_MODE_FOLLOW) {
            a = xb_follow(s, v3(R.p0x[tp], R.p0y[tp], R.p0z[tp]), R.radius[tp], P);
            w.path = XB_PATH_FOLLOW;
        } else if (s.mode == XB_MODE_ORBIT) {
            w.path = P.new_orbit ? XB_PATH_ORBIT_NEW : XB_PATH_ORBIT_OLD;
            XbV3 op = v3(R.p0x[tp], R.p0y[tp], R.p0z[tp]);
            if (P.new_orbit) {
                XbV3 ov = v3(R.v0x[tp], (R.id[tp] < s.id) ? R.vy_after[tp] : R.v0y[tp], R.v0z[tp]);
                a = xb_new_orbit(s, op, ov, R.radius[tp], P, DO ? &w.t.rot : (const XoQ*)0);
            } else {
                a = xb_old_orbit(s, op, R.radius[tp], P);
            }
        } else if (s.mode == XB_MODE_STOP) {
            if (!DO) s.vel.y = (s.vel.y - XB_STOP_DECAY * s.vel.y) * XB_STOP_DAMP;
        }
        if (DO) {
            int it = 0;
            a = xo_orient(s, w.t, a, P.friction, P.dt, &rot1, &w1, &it);
            w.iters = (uint8_t)(it < 255 ? it : 255);
        }
        s.last_g = a;
    }
    if (rel_after) xb_release(s, true, &P);
}

/* Phase 3 of a row in warp proper (modes.py:724-732): its warp path; dropping out (TRV-714), the drop-out point
   becomes its state and it takes one more step from there with a (mLastG + mLastC). */
static inline XB_HD void xb_phase3_warp(XbRow& w, XbV3 a, const XbPark& P, XbV3* pp, XbV3* vp) {
    XbShip& s = w.s;
    XbV3 p = s.pos, v = s.vel;
    xb_warp_path(s, &p, &v, (double)(s.clock - s.effect_stamp) * P.dt);
    if (s.mode == XB_MODE_STOP) {
        s.pos = p;
        s.vel = v;
        xb_integrate(&p, &v, a, s.mass * s.agility, s.tf, P);
        w.path = XB_PATH_WARP_EDGE;
    }
    *pp = p;
    *vp = v;
}

/* Phase 4 (modes.py:735-743): the start-of-tick state becomes the old state, the end-of-tick state the state. */
template <bool DO>
static inline XB_HD void xb_phase4(XbRow& w, XbV3 p, XbV3 v, XoQ rot1, XoV3 w1) {
    XbShip& s = w.s;
    w.old_pos = s.pos;
    w.old_vel = s.vel;
    s.pos = p;
    s.vel = v;
    if (DO) {
        xo_commit(w.t, rot1, w1);
        w.turned = true;
    }
}

/* A mushroom's turn in the last phase (full.py:2545-2555), in place of the swap: the end-of-tick state stays in the old
   members (with dynamical orientation the rotation a tick on in the old rotation, as the turn before the swap leaves
   it, full.py:2526-2532); the fraction of its lifetime passed, (float)((now - start) tick) x 0.001f / (float)lifetime
   in float; growing, mRadius = range x powf(fraction, 0.25f) (glibc's powf) and true: to be re-filed; its lifetime
   passed, RemoveBall inside Evolve (full.py:1633-1657): stopped, moribund, non-massive, stamped with the clock (its
   followers would be released: there are none, the park refuses to make a mushroom a target). */
template <bool DO>
static inline XB_HD bool xb_mushroom_turn(XbRow& w, XbV3 p, XbV3 v, XoQ rot1, XoV3 w1, const XbPark& P) {
    XbShip& s = w.s;
    w.old_pos = p;
    w.old_vel = v;
    if (DO) {
        w.t.old_rot = rot1;
        w.t.old_w = w1;
        w.turned = true;
    }
    w.grown = true;
    const float elapsed = (float)(double)((s.clock - s.effect_stamp) * (int64_t)P.tick_ms);
    const float fraction = (elapsed * 0.001f) / (float)s.gp.x;
    if ((double)fraction < 1.0) {
        s.radius = (double)(float)(s.follow_range * (double)xb_powf(fraction, 0.25f));
        return true;
    }
    s.mode = XB_MODE_STOP;
    w.moribund = true;
    s.massive = false;
    s.effect_stamp = s.clock;
    return false;
}

/* The end of a row's tick: PetrifyTroll of a troll due (full.py:1800-1807: SetBallFree(false), so fixed, at rest,
   mLastG cleared; SetBallInteractive(false); RIGID), and the world's clock. Returns whether the row was petrified:
   the caller then does SetBallFree's rest, the display orientation snapped to the rest velocity and the filing. */
static inline XB_HD bool xb_tick_tail(XbRow& w) {
    XbShip& s = w.s;
    bool petrified = false;
    if (xb_row_live(w) && s.mode == XB_MODE_TROLL && !(s.effect_stamp > s.clock)) {
        s.vel = v3(0.0, 0.0, 0.0);
        w.old_vel = s.vel;
        s.last_g = s.vel;
        w.free_ = false;
        s.interactive = false;
        s.mode = XB_MODE_RIGID;
        petrified = true;
    }
    s.clock = s.clock + 1;
    return petrified;
}

/* One tick of the row at place j of its group [lo, hi), without collisions (mLastC stays zero). With YPR (and without
   dynamical orientation) the display orientation `y` evolves at the commit. */
template <bool DO, bool YPR = false>
static inline XB_HD void xb_tick_row(XbRow& w, const XbRecords& R, int j, int lo, int hi, const XbPark& P,
                                     XbYpr* y = nullptr) {
    XbShip& s = w.s;
    XoQ rot1 = xo_q(0.0f, 0.0f, 0.0f, 1.0f);     /* with dynamical orientation: the rotation and angular velocity */
    XoV3 w1 = xo_v(0.0f, 0.0f, 0.0f);           /* a tick on */
    xb_phase1<DO>(w, R, j, lo, hi, P, rot1, w1);
    if (xb_row_live(w)) {
        if (s.massive) s.last_collision = -1.0;
        XbV3 p = s.pos, v = s.vel;
        XbV3 a = v3(s.last_g.x + 0.0, s.last_g.y + 0.0, s.last_g.z + 0.0);     /* mLastG + mLastC (zero) */
        if (s.mode == XB_MODE_WARP && s.effect_stamp >= 0) xb_phase3_warp(w, a, P, &p, &v);
        else xb_integrate(&p, &v, a, s.mass * s.agility, s.tf, P);
        if (s.mode == XB_MODE_MUSHROOM) {
            xb_mushroom_turn<DO>(w, p, v, rot1, w1, P);
        } else {
            xb_phase4<DO>(w, p, v, rot1, w1);
            if (YPR && !DO) xb_yaw_pitch_roll(*y, s, false, s.clock, false, P);
        }
    }
    if (xb_tick_tail(w) && YPR && !DO) xb_yaw_pitch_roll(*y, s, false, s.clock, true, P);
}

#include "xb_collide.h"
#include "xb_statics.h"

/* ------------------------------------------------------------------------------ the tick with collisions */

/* A row's collision state as its lane holds it: mLastC, the filing, the collision list's length, space junk. */
struct XbCollRow {
    XbV3 last_c;
    XbFiling fil;
    int64_t nh, cloaked;
    bool junk;
};

static inline XB_HD void xb_coll_load(const XbPark& P, int64_t r, XbCollRow& c) {
    const int64_t st = P.stride;
    c.last_c = v3(P.cf[(XB_C_LAST_C + 0) * st + r], P.cf[(XB_C_LAST_C + 1) * st + r], P.cf[(XB_C_LAST_C + 2) * st + r]);
    c.fil.level = (int)P.cq[XB_C_LEVEL * st + r];
    for (int q = 0; q < 3; ++q) {
        c.fil.key[q] = P.cq[(XB_C_KEY + q) * st + r];
        c.fil.mod[q] = (int)P.cq[(XB_C_SIDE + q) * st + r];
    }
    c.nh = P.cq[XB_C_HITS_N * st + r];
    c.cloaked = XB_Q(P, XB_I_CLOAKED, r);
    c.junk = XB_B(P, XB_B_JUNK, r) != 0;
}

static inline XB_HD void xb_coll_store(const XbPark& P, int64_t r, const XbCollRow& c) {
    const int64_t st = P.stride;
    P.cf[(XB_C_LAST_C + 0) * st + r] = c.last_c.x;
    P.cf[(XB_C_LAST_C + 1) * st + r] = c.last_c.y;
    P.cf[(XB_C_LAST_C + 2) * st + r] = c.last_c.z;
    P.cq[XB_C_LEVEL * st + r] = c.fil.level;
    for (int q = 0; q < 3; ++q) {
        P.cq[(XB_C_KEY + q) * st + r] = c.fil.key[q];
        P.cq[(XB_C_SIDE + q) * st + r] = c.fil.mod[q];
    }
    P.cq[XB_C_HITS_N * st + r] = c.nh;
}

/* The row's tick-start view for the others' gradients (after the behaviours): position, velocity, mLastG, the tick's
   straight prediction, masses, decay, filing, flags. */
static inline XB_HD void xb_coll_record(const XbRow& w, const XbCollRow& c, const XbCollRecs& C, int j,
                                        const XbPark& P) {
    const XbShip& s = w.s;
    const double m = w.free_ ? s.mass * s.agility : XB_FIXED_MASS;
    XbV3 e = s.pos, ev = s.vel;
    xb_integrate(&e, &ev, s.last_g, m, s.tf, P);
    C.px[j] = s.pos.x, C.py[j] = s.pos.y, C.pz[j] = s.pos.z;
    C.vx[j] = s.vel.x, C.vy[j] = s.vel.y, C.vz[j] = s.vel.z;
    C.gx[j] = s.last_g.x, C.gy[j] = s.last_g.y, C.gz[j] = s.last_g.z;
    C.ex[j] = e.x, C.ey[j] = e.y, C.ez[j] = e.z;
    C.mass[j] = s.mass, C.m[j] = m, C.tf[j] = s.tf;
    C.owner[j] = s.owner_id;
    C.fid[j] = s.follow_id;
    C.kind[j] = (uint8_t)((s.mode == XB_MODE_FIELD ? XB_CK_FIELD : 0) | (s.mode == XB_MODE_MINIBALL ? XB_CK_MINIBALL : 0));
    if (C.fh) {
        C.fh[j] = P.fld[w.row], C.fc[j] = P.fld[P.stride + w.row], C.fa[j] = P.fld[2 * P.stride + w.row];
    }
    xb_rec_put_filing(C, j, c.fil);
    uint8_t f = 0;
    if (s.massive) f |= XB_CF_MASSIVE;
    if (w.free_) f |= XB_CF_FREE;
    if (w.moribund || c.cloaked != 0) f |= XB_CF_SKIP;
    if (s.mode == XB_MODE_MISSILE) f |= XB_CF_MISSILE;
    if (c.junk) f |= XB_CF_JUNK;
    if (s.mode == XB_MODE_MUSHROOM) f |= XB_CF_MUSHROOM;
    C.flags[j] = f;
}

/* Phase 3 of a row in warp proper, with collisions: as xb_phase3_warp; a row that drops out is massive from then on,
   for the balls after it in id order (its record updated: the drop-out point, its prediction, its filing when it
   came to rest and was filed again, space.py's InsertBallInBoxes). */
static inline XB_HD void xb_coll_phase3_warp(XbRow& w, XbCollRow& c, const XbCollRecs& C, int j, const XbPark& P,
                                             XbV3* pp, XbV3* vp) {
    XbShip& s = w.s;
    const XbV3 a = v3(s.last_g.x + c.last_c.x, s.last_g.y + c.last_c.y, s.last_g.z + c.last_c.z);
    xb_phase3_warp(w, a, P, pp, vp);
    if (s.mode != XB_MODE_STOP) return;
    if (v3_lensq(s.vel) < 0.1) xb_file(s, w.free_, P, c.fil);
    const double m = s.mass * s.agility;
    XbV3 e = s.pos, ev = s.vel;
    xb_integrate(&e, &ev, s.last_g, m, s.tf, P);
    C.px[j] = s.pos.x, C.py[j] = s.pos.y, C.pz[j] = s.pos.z;
    C.vx[j] = s.vel.x, C.vy[j] = s.vel.y, C.vz[j] = s.vel.z;
    C.ex[j] = e.x, C.ey[j] = e.y, C.ez[j] = e.z;
    xb_rec_put_filing(C, j, c.fil);
    C.flags[j] = (uint8_t)((C.flags[j] & ~XB_CF_MASSIVE) | XB_CF_DROPPED);
}

/* Phase 3's gradient and end state for a row not in warp proper, then phase 4 and the refiling of a row that moved
   at least sqrt(0.1) m (modes.py:717-743). The row's world is at places [wlo, whi) of the group. */
template <bool DO, bool YPR = false, bool ST = false>
static inline XB_HD void xb_coll_phase34(XbRow& w, XbCollRow& c, bool warped, XbV3 p, XbV3 v, const XbCollRecs& C,
                                         const XbRecords& R, int j, int wlo, int whi, const XbPark& P, XoQ rot1,
                                         XoV3 w1, XbYpr* y = nullptr) {
    XbShip& s = w.s;
    if (!xb_row_live(w)) return;
    if (!warped) {
        if (s.massive) {
            s.last_collision = -1.0;
            XbCollMe me;
            me.pos = s.pos, me.vel = s.vel, me.g = s.last_g;
            me.m = s.mass * s.agility, me.mass = s.mass, me.tf = s.tf, me.radius = s.radius;
            me.id = s.id, me.mode = s.mode, me.follow_id = s.follow_id, me.owner_id = s.owner_id;
            me.last_c = c.last_c, me.last_collision = s.last_collision;
            me.nh = c.nh, me.hits = P.hits + w.row, me.stride = P.stride, me.cap = P.hits_cap;
            if (s.mode == XB_MODE_MISSILE && s.owner_id >= 0) {
                const int64_t tp = s.follow_ptr >= 0 ? (int64_t)j + (s.follow_ptr - w.row) : -1;
                xb_missile_gradient(me, c.junk, C, R.id, R.radius, j, (tp >= wlo && tp < whi) ? (int)tp : -1, P);
            } else {
                xb_gradient(me, c.fil, c.junk, C, R.id, R.radius, j, wlo, whi, P);
            }
            if (ST && P.statics && w.free_ && s.mode != XB_MODE_MISSILE && s.mode != XB_MODE_MUSHROOM) {
                const int cnt = P.st_range[P.stride + w.row];
                if (cnt > 0) {
                    XbStatics S;
                    S.f = P.st_f, S.q = P.st_q, S.keys = P.st_keys, S.stride = P.st_stride;
                    xb_statics_gradient(me, c.fil, S, P.st_range[w.row], cnt, P);
                }
            }
            c.last_c = me.last_c;
            s.last_collision = me.last_collision;
            c.nh = me.nh;
        }
        p = s.pos, v = s.vel;
        const XbV3 a = v3(s.last_g.x + c.last_c.x, s.last_g.y + c.last_c.y, s.last_g.z + c.last_c.z);
        xb_integrate(&p, &v, a, s.mass * s.agility, s.tf, P);
    }
    if (s.mode == XB_MODE_MUSHROOM) {
        if (xb_mushroom_turn<DO>(w, p, v, rot1, w1, P)) xb_file(s, w.free_, P, c.fil);
        return;
    }
    xb_phase4<DO>(w, p, v, rot1, w1);
    if (YPR && !DO) xb_yaw_pitch_roll(*y, s, c.nh > 0, s.clock, false, P);
    const XbV3 d = v3_sub(s.pos, w.old_pos);
    if (!(v3_lensq(d) < 0.1)) xb_file(s, w.free_, P, c.fil);
}

#endif
