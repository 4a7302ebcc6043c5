// SPDX-License-Identifier: MIT
//
// Commands and setters between ticks, one record per listed row: the drop-in park's masked API (stop, goto_point,
// goto_direction, follow, orbit, warp_to, entity_warp_in, the SetBall* setters, the orientation setters,
// missile_follow, launch_missile) as one interpreter, run by the host build and by the gfx950 kernels alike.
//
// What each command does to a row is Destiny's (Stop(Ball*), GotoPoint, FollowBall, WarpTo, EntityWarpIn,
// StopAllFollowers, SetBall*, LaunchMissile, MissileFollow: src/exact_ballpark/modes.py:206-426, full.py:1555-1736,
// vec_orient.py:403-478, as the reference's batched park applies them to masked rows, vec_modes.py:286-483 and
// vec_combat.py:339-398). How a batch runs is this file's:
//
//   gather  every listed row copies what it needs of OTHER rows (a target's position, world, bubble, flags; a
//           launcher's position, velocity, rotation) into its own record, and reports a refusal (a mushroom as a
//           target, a launcher at rest with no heading) in a status word -- nothing is written to the park;
//   apply   every listed row runs its command on its own row alone, reading other rows only through its record, so
//           the lanes of a batch never race and the result is that of reading every row before writing any (the
//           batched park's semantics); a command that stops a ball's followers marks the ball;
//   release every row of the park whose follow target is marked (and still names it by id) is released as
//           StopAllFollowers releases it: a missile or an interactive orbiter flies on along its velocity, any other
//           follower stops;
//   unmark  the marks of the listed rows are cleared (the mark row is all zero again).
//
// Include after xb_modes_tick.h (the park, the accessors, the vector helpers), xb_collide.h (the filing) and xb_ypr.h
// (the display orientation).
#pragma once

#include "../../include/xb.h"

#define XB_CMD_NONE 0
#define XB_CMD_GOTO 1                 /* x[0..2] the point */
#define XB_CMD_GOTO_DIRECTION 2       /* x[0..2] the direction */
#define XB_CMD_STOP 3
#define XB_CMD_FOLLOW 4               /* i[0] the target row, x[0] the range */
#define XB_CMD_ORBIT 5                /* i[0] the target row, x[0] the range */
#define XB_CMD_WARP 6                 /* x[0..2] the point, x[3] the minimum range, i[0] the warp factor */
#define XB_CMD_WARP_IN 7              /* x[0..2] the point, i[0] the warp factor; marks the ball */
#define XB_CMD_STOP_RELEASE 8         /* Stop(Ball*) and StopAllFollowers (RemoveBall's first steps); marks the ball */
#define XB_CMD_RELEASE 9              /* StopAllFollowers alone; marks the ball */
#define XB_CMD_SET_MASS 10            /* x[0] */
#define XB_CMD_SET_AGILITY 11         /* x[0] */
#define XB_CMD_SET_MAX_VEL 12         /* x[0]; re-filed */
#define XB_CMD_SET_SPEED_FRACTION 13  /* x[0] */
#define XB_CMD_SET_RADIUS 14          /* x[0]; re-filed */
#define XB_CMD_SET_VELOCITY 15        /* x[0..2]; the orientation snapped */
#define XB_CMD_SET_MAX_ANG_SPEED 16   /* x[0] */
#define XB_CMD_SET_ROT_AGILITY 17     /* x[0] */
#define XB_CMD_SET_ANG_VEL 18         /* x[0..2] */
#define XB_CMD_SET_ROTATION 19        /* x[0..3] the quaternion x, y, z, w */
#define XB_CMD_SET_ROTATION_YPR 20    /* x[0..2] yaw, pitch, roll */
#define XB_CMD_SET_FREE 21            /* flag */
#define XB_CMD_SET_RIGID 22
#define XB_CMD_SET_BOID 23            /* flag */
#define XB_CMD_SET_INTERACTIVE 24     /* flag */
#define XB_CMD_SET_TROLL 25           /* i[0] the delay in ticks */
#define XB_CMD_MISSILE_FOLLOW 26      /* i[0] the target row (-1 none), i[1] the target's id, i[2] the owner's id */
#define XB_CMD_LAUNCH_MISSILE 27      /* i[0] the launcher row (-1: no launch), i[1] the target row (-1 none), i[2]
                                         the owner's id; flag: aimed | massive << 8 (chars) */
#define XB_CMD_OPS 28

/* refusals the gather step reports (status bits): the batch must not be applied */
#define XB_REFUSE_MUSHROOM_TARGET 1   /* a mushroom as a follow, orbit or missile target (not batched) */
#define XB_REFUSE_NO_HEADING 2        /* a launcher at rest, and neither dynamical nor display orientation */

#define XB_DIR_DISTANCE 1.0e17        /* GotoDirection's point: that far along the direction */
#define XB_AU 0.1495978707e12         /* EntityWarpIn's velocity: one AU a second toward the point */
#define XB_MIN_WARP_DISTANCE_SQ 10000000000.0   /* WarpTo under 100 km is a GotoPoint */
#define XB_TROLL_MIN_DELAY 1
#define XB_MISSILE_LAUNCH_SPEED 150.0
#define XB_MISSILE_AHEAD 1.0e16       /* MissileFollow's straight flight: its point that far along the velocity */
#define XB_EMPTY_ID (-4611686018427387904LL)   /* an empty row's id (-(2^62)): no ball */

/* A listed row's command (include/xb.h's xb_command): the caller fills op .. x (the other words zero); the gather step
   fills the rest. */
typedef xb_command XbCommand;
#define XB_T_ROW 1                    /* t_state: the target is a row of the park */
#define XB_T_MORIBUND 2
#define XB_T_MUSHROOM 4
#define XB_L_ROW 8                    /* t_state: the launcher is a row of the park */
#define XB_L_HEADING 16               /* t_state: l_rot holds the launcher's rotation */
#define XB_ACCEPTED 32                /* t_state, set by the apply step: SetMaxSpeed or SetBallRadius took the value */

static inline XB_HD bool xb_follows(int64_t mode) {
    return mode == XB_MODE_FOLLOW || mode == XB_MODE_ORBIT || mode == XB_MODE_MISSILE || mode == XB_MODE_FORMATION;
}

/* ------------------------------------------------------------------------------------------------- gather */

static inline XB_HD int xb_cmd_gather(const XbPark& P, int64_t n, int64_t r, XbCommand& c, const double* ypr) {
    int refuse = 0;
    const int64_t op = c.op;
    int64_t t = -1;
    if (op == XB_CMD_FOLLOW || op == XB_CMD_ORBIT || op == XB_CMD_MISSILE_FOLLOW) t = c.i[0];
    if (op == XB_CMD_LAUNCH_MISSILE) t = c.i[1];
    c.t_state = 0;
    if (t >= 0 && t < n) {
        c.t_state |= XB_T_ROW;
        const XbV3 p = xb_ld3(P, XB_F_POS, t);
        c.t_pos[0] = p.x, c.t_pos[1] = p.y, c.t_pos[2] = p.z;
        c.t_radius = XB_F(P, XB_F_RADIUS, t);
        c.t_id = XB_Q(P, XB_I_IDS, t);
        c.t_world = XB_Q(P, XB_I_WORLD, t);
        c.t_bubble = XB_Q(P, XB_I_BUBBLE, t);
        c.t_cloaked = XB_Q(P, XB_I_CLOAKED, t);
        if (XB_B(P, XB_B_MORIBUND, t)) c.t_state |= XB_T_MORIBUND;
        if (XB_Q(P, XB_I_MODE, t) == XB_MODE_MUSHROOM) c.t_state |= XB_T_MUSHROOM;
    }
    if (op == XB_CMD_LAUNCH_MISSILE) {
        const int64_t l = c.i[0];
        if (l >= 0 && l < n) {
            c.t_state |= XB_L_ROW;
            const XbV3 p = xb_ld3(P, XB_F_POS, l), v = xb_ld3(P, XB_F_VEL, l);
            c.l_pos[0] = p.x, c.l_pos[1] = p.y, c.l_pos[2] = p.z;
            c.l_vel[0] = v.x, c.l_vel[1] = v.y, c.l_vel[2] = v.z;
            c.l_max_vel = XB_F(P, XB_F_MAX_VEL, l);
            if (P.dynamical) {
                for (int k = 0; k < 4; ++k) c.l_rot[k] = XB_F(P, XB_F_ROT + k, l);
                c.t_state |= XB_L_HEADING;
            } else if (ypr) {     /* RotationQuaternion of the display yaw, pitch and roll */
                const XoQ q = xo_rotation_ypr((float)ypr[0 * P.stride + l], (float)ypr[1 * P.stride + l],
                                              (float)ypr[2 * P.stride + l]);
                c.l_rot[0] = q.x, c.l_rot[1] = q.y, c.l_rot[2] = q.z, c.l_rot[3] = q.w;
                c.t_state |= XB_L_HEADING;
            }
        }
    }
    /* the refusals, where the apply step would accept the command otherwise */
    const bool mush = (c.t_state & XB_T_MUSHROOM) != 0;
    const bool t_ok = (c.t_state & XB_T_ROW) && !(c.t_state & XB_T_MORIBUND);
    if ((op == XB_CMD_FOLLOW || op == XB_CMD_ORBIT) && mush && t_ok && c.i[0] != r &&
        xb_finite(xb_f32(c.x[0])) && c.t_world == XB_Q(P, XB_I_WORLD, r) && c.t_bubble == XB_Q(P, XB_I_BUBBLE, r) &&
        c.t_cloaked == 0)
        refuse |= XB_REFUSE_MUSHROOM_TARGET;
    if (op == XB_CMD_MISSILE_FOLLOW && mush && t_ok) {
        const int64_t me = XB_Q(P, XB_I_IDS, r), dst = c.i[1], own = c.i[2];
        if (dst >= 0 && me != dst && me != own && dst != own) refuse |= XB_REFUSE_MUSHROOM_TARGET;
    }
    if (op == XB_CMD_LAUNCH_MISSILE && (c.t_state & XB_L_ROW)) {
        if (mush) refuse |= XB_REFUSE_MUSHROOM_TARGET;
        const int aimed = (int)(int8_t)(uint8_t)(c.flag & 0xFF);
        const bool aim = (c.t_state & XB_T_ROW) && aimed == 1 && c.t_cloaked == 0;
        const XbV3 dir = v3_unit(v3(c.l_vel[0], c.l_vel[1], c.l_vel[2]));
        if (!aim && v3_lensq(dir) == 0.0 && !(c.t_state & XB_L_HEADING)) refuse |= XB_REFUSE_NO_HEADING;
    }
    return refuse;
}

/* -------------------------------------------------------------------------------------------- own-row steps */

/* Stop(Ball*): a follower forgets its target (stamp, pointer, ids, range); the mode becomes STOP. */
static inline XB_HD void xb_row_stop(const XbPark& P, int64_t r) {
    if (XB_Q(P, XB_I_MODE, r) == XB_MODE_FORMATION)               /* the master's FreeFormationSlot first */
        xb_free_formation_slot(P, XB_Q(P, XB_I_FOLLOW_PTR, r), XB_Q(P, XB_I_EFFECT_STAMP, r));
    if (xb_follows(XB_Q(P, XB_I_MODE, r))) {
        XB_Q(P, XB_I_EFFECT_STAMP, r) = 0;
        XB_Q(P, XB_I_FOLLOW_PTR, r) = -1;
        XB_Q(P, XB_I_FOLLOW_ID, r) = 0;
        XB_Q(P, XB_I_OWNER_ID, r) = 0;
        XB_F(P, XB_F_FOLLOW_RANGE, r) = 0.0;
    }
    XB_Q(P, XB_I_MODE, r) = XB_MODE_STOP;
}

static inline XB_HD void xb_row_goto(const XbPark& P, int64_t r, XbV3 p) {
    if (!v3_finite(p)) return;
    xb_row_stop(P, r);
    xb_st3(P, XB_F_GOTO, r, p);
    if (XB_F(P, XB_F_SPEED_FRACTION, r) == 0.0) XB_F(P, XB_F_SPEED_FRACTION, r) = 1.0;
    XB_Q(P, XB_I_MODE, r) = XB_MODE_GOTO;
}

static inline XB_HD void xb_row_goto_direction(const XbPark& P, int64_t r, XbV3 d) {
    if (!v3_finite(d)) return;
    xb_row_goto(P, r, v3_add(xb_ld3(P, XB_F_POS, r), v3_scale(v3_unit(d), XB_DIR_DISTANCE)));
}

/* SetBallTimeFactor: 0 when mass x agility is not positive, else glibc's exp(-k dt / (mass x agility)). */
static inline XB_HD void xb_row_time_factor(const XbPark& P, int64_t r) {
    const double m = XB_F(P, XB_F_MASS, r) * XB_F(P, XB_F_AGILITY, r);
    XB_F(P, XB_F_TIME_FACTOR, r) = (m <= 0.0) ? 0.0 : xb_exp((-P.friction * P.dt) / m);
}

/* InsertBallInBoxes' filing from the row's state (with the gradient method only). */
static inline XB_HD void xb_row_refile(const XbPark& P, int64_t r) {
    if (!P.cq) return;
    XbRow w;
    xb_row_load(P, r, w);
    XbCollRow c;
    xb_coll_load(P, r, c);
    xb_file(w.s, w.free_, P, c.fil);
    xb_coll_store(P, r, c);
}

/* SnapOrientation (dynamical orientation) or CalculateYawPitchRoll with snap (the display orientation) to the row's
   velocity, as SetBallVelocity does. */
static inline XB_HD void xb_row_snap(const XbPark& P, int64_t r, double* ypr) {
    if (P.dynamical) {
        XbTurn t;
        xb_turn_load(P, r, t);
        t.old_rot = xo_q((float)XB_F(P, XB_F_OLD_ROT, r), (float)XB_F(P, XB_F_OLD_ROT + 1, r),
                         (float)XB_F(P, XB_F_OLD_ROT + 2, r), (float)XB_F(P, XB_F_OLD_ROT + 3, r));
        t.old_w = xo_v((float)XB_F(P, XB_F_OLD_ANG_VEL, r), (float)XB_F(P, XB_F_OLD_ANG_VEL + 1, r),
                       (float)XB_F(P, XB_F_OLD_ANG_VEL + 2, r));
        t.torque = xo_v((float)XB_F(P, XB_F_TORQUE, r), (float)XB_F(P, XB_F_TORQUE + 1, r),
                        (float)XB_F(P, XB_F_TORQUE + 2, r));
        t.roll = (float)XB_F(P, XB_F_ROT_ROLL, r);
        xo_snap(xb_ld3(P, XB_F_VEL, r), t);
        xb_turn_store(P, r, t);
    } else if (ypr) {
        XbRow w;
        xb_row_load(P, r, w);
        XbYpr y;
        xb_ypr_load(ypr, P.stride, r, y);
        xb_yaw_pitch_roll(y, w.s, false, w.s.clock, true, P);
        xb_ypr_store(ypr, P.stride, r, y);
    }
}

static inline XB_HD void xb_row_set_velocity(const XbPark& P, int64_t r, XbV3 v, double* ypr) {
    xb_st3(P, XB_F_VEL, r, v);
    xb_st3(P, XB_F_OLD_VEL, r, v);
    xb_row_snap(P, r, ypr);
}

/* SetBallRotation: the float quaternion normalised when longer than 1e-20 (else ignored), the rotation stopped. */
static inline XB_HD void xb_row_set_rotation(const XbPark& P, int64_t r, XoQ q) {
    const float len = xo_qlen(q);
    if (!((double)len > 1e-20)) return;
    const float inv = 1.0f / len;
    const XoQ u = xo_q(q.x * inv, q.y * inv, q.z * inv, q.w * inv);
    const float c[4] = {u.x, u.y, u.z, u.w};
    for (int k = 0; k < 4; ++k) XB_F(P, XB_F_ROT + k, r) = c[k], XB_F(P, XB_F_OLD_ROT + k, r) = c[k];
    for (int k = 0; k < 3; ++k) XB_F(P, XB_F_ANG_VEL + k, r) = 0.0, XB_F(P, XB_F_OLD_ANG_VEL + k, r) = 0.0;
}

/* SetBallFree: a ball fixed is stopped, at rest (its orientation snapped), re-filed and its thrust cleared. */
static inline XB_HD void xb_row_set_free(const XbPark& P, int64_t r, bool on, double* ypr) {
    if (XB_Q(P, XB_I_IDS, r) == XB_EMPTY_ID || (XB_B(P, XB_B_FREE, r) != 0) == on) return;
    XB_B(P, XB_B_FREE, r) = on ? 1 : 0;
    if (on) return;
    xb_row_stop(P, r);
    xb_row_set_velocity(P, r, v3(0.0, 0.0, 0.0), ypr);
    xb_row_refile(P, r);
    xb_st3(P, XB_F_LAST_G, r, v3(0.0, 0.0, 0.0));
}

/* WarpTo: a finite point; under 100 km a GotoPoint; nothing if already warping to that very point; else the warp
   waits for alignment (stamp -1), its minimum range's bits in follow_id and its factor in owner_id. */
static inline XB_HD void xb_row_warp(const XbPark& P, int64_t r, XbV3 p, double min_range, int64_t factor) {
    if (!v3_finite(p)) return;
    const double mr = min_range < 0.0 ? 0.0 : min_range;
    const int64_t wf = factor <= 0 ? 1 : factor;
    if (v3_lensq(v3_sub(xb_ld3(P, XB_F_POS, r), p)) < XB_MIN_WARP_DISTANCE_SQ) {
        xb_row_goto(P, r, p);
        return;
    }
    if (XB_Q(P, XB_I_MODE, r) == XB_MODE_WARP) {
        const XbV3 g = xb_ld3(P, XB_F_GOTO, r);
        if (g.x == p.x && g.y == p.y && g.z == p.z) return;
    }
    xb_row_stop(P, r);
    xb_st3(P, XB_F_GOTO, r, p);
    XB_Q(P, XB_I_EFFECT_STAMP, r) = -1;
    XB_F(P, XB_F_FOLLOW_RANGE, r) = 0.0;
    XB_Q(P, XB_I_FOLLOW_ID, r) = (int64_t)xb_bits(mr);
    XB_Q(P, XB_I_OWNER_ID, r) = wf;
    XB_Q(P, XB_I_MODE, r) = XB_MODE_WARP;
}

/* FollowBall / Orbit of a gathered target: a row of the park, not this one, in the same world and bubble, not
   moribund, not cloaked; the range a finite float. */
static inline XB_HD void xb_row_follow(const XbPark& P, int64_t r, const XbCommand& c, int64_t mode) {
    const double rng = xb_f32(c.x[0]);
    if (!xb_finite(rng) || !(c.t_state & XB_T_ROW) || c.i[0] == r) return;
    if (c.t_world != XB_Q(P, XB_I_WORLD, r) || (c.t_state & XB_T_MORIBUND) || c.t_bubble != XB_Q(P, XB_I_BUBBLE, r) ||
        c.t_cloaked != 0)
        return;
    xb_row_stop(P, r);
    XB_Q(P, XB_I_FOLLOW_ID, r) = c.t_id;
    XB_Q(P, XB_I_FOLLOW_PTR, r) = c.i[0];
    XB_F(P, XB_F_FOLLOW_RANGE, r) = rng;
    XB_Q(P, XB_I_MODE, r) = mode;
}

/* MissileFollow (target t_id at row i[0], owner i[2]; a negative owner is a defender): refused for a target that is
   not a row, the missile itself, its owner, or moribund; otherwise stopped, and, unless the target is in another
   bubble or cloaked, sent at it: the follow range minus both radii (a float), straight on for 800 ms toward a point far
   along its velocity at full speed fraction. */
static inline XB_HD void xb_row_missile_follow(const XbPark& P, int64_t r, const XbCommand& c, int64_t dst) {
    const int64_t own = c.i[2], me = XB_Q(P, XB_I_IDS, r);
    if (dst < 0 || me == dst || me == own || dst == own || !(c.t_state & XB_T_ROW) || (c.t_state & XB_T_MORIBUND))
        return;
    xb_row_stop(P, r);
    if (XB_Q(P, XB_I_BUBBLE, r) != c.t_bubble || c.t_cloaked != 0) return;
    XB_Q(P, XB_I_FOLLOW_ID, r) = dst;
    XB_Q(P, XB_I_FOLLOW_PTR, r) = c.op == XB_CMD_LAUNCH_MISSILE ? c.i[1] : c.i[0];
    XB_F(P, XB_F_FOLLOW_RANGE, r) = -xb_f32(XB_F(P, XB_F_RADIUS, r) + c.t_radius);
    XB_Q(P, XB_I_OWNER_ID, r) = own;
    XB_Q(P, XB_I_EFFECT_STAMP, r) = XB_Q(P, XB_I_CLOCK, r);
    xb_st3(P, XB_F_GOTO, r, v3_scale(v3_unit(xb_ld3(P, XB_F_VEL, r)), XB_MISSILE_AHEAD));
    XB_F(P, XB_F_SPEED_FRACTION, r) = 1.0;
    XB_Q(P, XB_I_MODE, r) = XB_MODE_MISSILE;
}

/* LaunchMissile from the gathered launcher: at its position, aimed at the target (unit speed) or along its velocity
   (its heading backwards when at rest) at max(150, its maximum speed); massive if asked; the position set (and
   re-filed) and the velocity set (snapped); then MissileFollow, or GotoDirection without a target. */
static inline XB_HD void xb_row_launch(const XbPark& P, int64_t r, const XbCommand& c, double* ypr) {
    if (!(c.t_state & XB_L_ROW)) return;
    const int aimed = (int)(int8_t)(uint8_t)(c.flag & 0xFF), massive = (int)(int8_t)(uint8_t)((c.flag >> 8) & 0xFF);
    const bool has_t = (c.t_state & XB_T_ROW) != 0;
    const XbV3 ps = v3(c.l_pos[0], c.l_pos[1], c.l_pos[2]), vs = v3(c.l_vel[0], c.l_vel[1], c.l_vel[2]);
    const bool aim = has_t && aimed == 1 && c.t_cloaked == 0;
    const double max_velocity = aim ? 0.0 : c.l_max_vel;
    XbV3 dir = v3_unit(vs);
    if (!aim && v3_lensq(dir) == 0.0 && (c.t_state & XB_L_HEADING)) {
        const XoQ q = xo_q((float)c.l_rot[0], (float)c.l_rot[1], (float)c.l_rot[2], (float)c.l_rot[3]);
        const XoV3 back = xo_transform(xo_v(0.0f, 0.0f, -1.0f), xo_matrix(q));
        dir = v3(back.x, back.y, back.z);
    }
    const double speed = XB_MISSILE_LAUNCH_SPEED < max_velocity ? max_velocity : XB_MISSILE_LAUNCH_SPEED;
    const XbV3 v0 = aim ? v3_unit(v3_sub(v3(c.t_pos[0], c.t_pos[1], c.t_pos[2]), ps)) : v3_add(vs, v3_scale(dir, speed));
    if (massive != 0) XB_B(P, XB_B_MASSIVE, r) = 1;
    xb_st3(P, XB_F_POS, r, ps);                       /* SetBallPosition */
    xb_st3(P, XB_F_OLD_POS, r, ps);
    xb_row_refile(P, r);
    xb_row_set_velocity(P, r, v0, ypr);
    if (!has_t) {
        if (max_velocity > 0.0) xb_row_goto_direction(P, r, v0);
    } else if (c.t_cloaked == 0) {
        xb_row_missile_follow(P, r, c, c.t_id);
    }
}

/* EntityWarpIn: WarpTo the point with no minimum range; a ball that is not then going to a point is marked for
   StopAllFollowers and put into warp proper from its exit point, moving one AU a second, neither massive nor sensing,
   its stamp five ticks back (not before 0) and its warp length the distance to its exit. */
static inline XB_HD bool xb_row_warp_in(const XbPark& P, int64_t r, XbV3 point, int64_t factor) {
    xb_row_warp(P, r, point, 0.0, factor);
    if (XB_Q(P, XB_I_MODE, r) == XB_MODE_GOTO) return false;
    const XbV3 pos = xb_ld3(P, XB_F_POS, r), g0 = xb_ld3(P, XB_F_GOTO, r);
    const double mr = xb_dbl((uint64_t)XB_Q(P, XB_I_FOLLOW_ID, r));
    const XbV3 g = v3_add(g0, v3_scale(v3_unit(v3_sub(pos, g0)), mr));
    xb_st3(P, XB_F_GOTO, r, g);
    XB_B(P, XB_B_MASSIVE, r) = 0;
    XB_B(P, XB_B_SENSOR, r) = 0;
    xb_st3(P, XB_F_VEL, r, v3_scale(v3_unit(point), XB_AU));
    if (P.dynamical) xb_row_snap(P, r, nullptr);
    const int64_t back = XB_Q(P, XB_I_CLOCK, r) - 5;
    XB_Q(P, XB_I_EFFECT_STAMP, r) = back < 0 ? 0 : back;
    XB_F(P, XB_F_LAST_COLLISION, r) = v3_len(v3_sub(pos, g));
    return true;
}

/* --------------------------------------------------------------------------------------------------- apply */

/* Row r's command (its record gathered); returns true when the row's followers are to be released. */
static inline XB_HD bool xb_cmd_run(const XbPark& P, int64_t r, XbCommand& c, double* ypr) {
    const XbV3 p = v3(c.x[0], c.x[1], c.x[2]);
    const double a = c.x[0];
    switch (c.op) {
        case XB_CMD_GOTO: xb_row_goto(P, r, p); break;
        case XB_CMD_GOTO_DIRECTION: xb_row_goto_direction(P, r, p); break;
        case XB_CMD_STOP: xb_row_stop(P, r); break;
        case XB_CMD_FOLLOW: xb_row_follow(P, r, c, XB_MODE_FOLLOW); break;
        case XB_CMD_ORBIT: xb_row_follow(P, r, c, XB_MODE_ORBIT); break;
        case XB_CMD_WARP: xb_row_warp(P, r, p, c.x[3], c.i[0]); break;
        case XB_CMD_WARP_IN: return xb_row_warp_in(P, r, p, c.i[0]);
        case XB_CMD_STOP_RELEASE: xb_row_stop(P, r); return true;
        case XB_CMD_RELEASE: return true;
        case XB_CMD_SET_MASS:
            if (!(a <= 0.0)) XB_F(P, XB_F_MASS, r) = a, xb_row_time_factor(P, r);
            break;
        case XB_CMD_SET_AGILITY: {
            const double f = xb_f32(a);
            if (!(f <= 0.0)) XB_F(P, XB_F_AGILITY, r) = f, xb_row_time_factor(P, r);
            break;
        }
        case XB_CMD_SET_MAX_VEL: {
            const double f = xb_f32(a);
            if (!(f < 0.0)) XB_F(P, XB_F_MAX_VEL, r) = f, xb_row_refile(P, r), c.t_state |= XB_ACCEPTED;
            break;
        }
        case XB_CMD_SET_SPEED_FRACTION: {
            const double f = xb_f32(a);
            if (xb_finite(f)) XB_F(P, XB_F_SPEED_FRACTION, r) = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
            break;
        }
        case XB_CMD_SET_RADIUS: {
            const double f = xb_f32(a);
            if (!(f < 0.0)) XB_F(P, XB_F_RADIUS, r) = f, xb_row_refile(P, r), c.t_state |= XB_ACCEPTED;
            break;
        }
        case XB_CMD_SET_VELOCITY: xb_row_set_velocity(P, r, p, ypr); break;
        case XB_CMD_SET_MAX_ANG_SPEED: {
            const double f = xb_f32(a);
            if (!(f < 0.0)) XB_F(P, XB_F_MAX_ANG_VEL, r) = f;
            break;
        }
        case XB_CMD_SET_ROT_AGILITY: {
            const double f = xb_f32(a);
            if (!(f <= 0.0)) XB_F(P, XB_F_ROT_AGILITY, r) = f;
            break;
        }
        case XB_CMD_SET_ANG_VEL:
            for (int k = 0; k < 3; ++k) {
                const double f = xb_f32(c.x[k]);
                XB_F(P, XB_F_ANG_VEL + k, r) = f, XB_F(P, XB_F_OLD_ANG_VEL + k, r) = f;
            }
            break;
        case XB_CMD_SET_ROTATION:
            xb_row_set_rotation(P, r, xo_q((float)c.x[0], (float)c.x[1], (float)c.x[2], (float)c.x[3]));
            break;
        case XB_CMD_SET_ROTATION_YPR:
            xb_row_set_rotation(P, r, xo_rotation_ypr((float)c.x[0], (float)c.x[1], (float)c.x[2]));
            break;
        case XB_CMD_SET_FREE: xb_row_set_free(P, r, c.flag != 0, ypr); break;
        case XB_CMD_SET_RIGID:
            xb_row_stop(P, r);
            XB_Q(P, XB_I_MODE, r) = XB_MODE_RIGID;
            break;
        case XB_CMD_SET_BOID:
            if (c.flag) XB_Q(P, XB_I_MODE, r) = XB_MODE_BOID;
            else xb_row_stop(P, r);
            break;
        case XB_CMD_SET_INTERACTIVE: XB_B(P, XB_B_INTERACTIVE, r) = c.flag ? 1 : 0; break;
        case XB_CMD_SET_TROLL: {
            const int64_t d = c.i[0] < XB_TROLL_MIN_DELAY ? XB_TROLL_MIN_DELAY : c.i[0];
            xb_row_stop(P, r);
            xb_row_set_free(P, r, true, ypr);
            XB_B(P, XB_B_INTERACTIVE, r) = 1;
            XB_Q(P, XB_I_EFFECT_STAMP, r) = XB_Q(P, XB_I_CLOCK, r) + d;
            XB_Q(P, XB_I_MODE, r) = XB_MODE_TROLL;
            break;
        }
        case XB_CMD_MISSILE_FOLLOW: xb_row_missile_follow(P, r, c, c.i[1]); break;
        case XB_CMD_LAUNCH_MISSILE: xb_row_launch(P, r, c, ypr); break;
        default: break;
    }
    return false;
}

/* StopAllFollowers' part on one follower (row r): a missile (no longer massive) and an interactive orbiter fly on along
   their velocity; any other follower stops. */
static inline XB_HD void xb_row_released(const XbPark& P, int64_t r) {
    const int64_t mode = XB_Q(P, XB_I_MODE, r);
    const bool missile = mode == XB_MODE_MISSILE;
    if (missile) XB_B(P, XB_B_MASSIVE, r) = 0;
    if (missile || (mode == XB_MODE_ORBIT && XB_B(P, XB_B_INTERACTIVE, r)))
        xb_row_goto_direction(P, r, xb_ld3(P, XB_F_VEL, r));
    else
        xb_row_stop(P, r);
}

/* StopAllFollowers over the park: row r is released when its follow target is marked and still names it by id. */
static inline XB_HD void xb_cmd_release(const XbPark& P, int64_t n, int64_t r, const uint8_t* mark) {
    const int64_t t = XB_Q(P, XB_I_FOLLOW_PTR, r);
    if (t < 0 || t >= n || !mark[t] || XB_Q(P, XB_I_FOLLOW_ID, r) != XB_Q(P, XB_I_IDS, t)) return;
    xb_row_released(P, r);
}

/* The flight commands of the serving loop's batch (op, point, target row, range, minimum range, warp factor at row
   r of each array) as records, run in place: a command reads of another row only what no flight command writes
   (its world, moribund flag, bubble, cloaking and id), so gathering and applying in one lane is the batch's result. */
static inline XB_HD void xb_cmd_apply(const XbPark& P, int64_t n, int64_t r, const uint8_t* op, const double* point,
                                      const int64_t* target, const double* range, const double* min_range,
                                      const int64_t* warp_factor) {
    XbCommand c = {};
    c.op = op[r];
    if (c.op == XB_CMD_NONE) return;
    c.x[0] = point[r * 3], c.x[1] = point[r * 3 + 1], c.x[2] = point[r * 3 + 2];
    if (c.op == XB_CMD_FOLLOW || c.op == XB_CMD_ORBIT) c.i[0] = target[r], c.x[0] = range[r];
    if (c.op == XB_CMD_WARP) c.x[3] = min_range[r], c.i[0] = warp_factor[r];
    if (c.op != XB_CMD_GOTO && c.op != XB_CMD_STOP && c.op != XB_CMD_FOLLOW && c.op != XB_CMD_ORBIT &&
        c.op != XB_CMD_WARP)
        return;
    xb_cmd_gather(P, n, r, c, nullptr);
    xb_cmd_run(P, r, c, nullptr);
}
