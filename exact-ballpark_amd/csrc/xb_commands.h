// SPDX-License-Identifier: MIT
//
// The flight-mode commands on one row, as the drop-in park's batched commands compute them between ticks: Stop,
// GotoPoint, Follow, Orbit and WarpTo (the park's stop, goto_point, follow, orbit and warp_to; the reference's
// src/exact_ballpark/vec_modes.py:315-386 and the server's per-row cmd_stop, cmd_goto_point, cmd_follow and
// cmd_warp_to, src/exact_ballpark/vec_server_cuda.cu:117-186). A command writes only its own row and reads of its
// target only fields no command writes (world, moribund, bubble, cloaked, id), so rows with one command each may run
// in any order, or all at once, with the result of applying them op by op.
//
// Include after xb_modes_tick.h (the park, the state accessors and the vector helpers).
#pragma once

#define XB_CMD_GOTO 1
#define XB_CMD_STOP 3
#define XB_CMD_FOLLOW 4
#define XB_CMD_ORBIT 5
#define XB_CMD_WARP 6

#define XB_MIN_WARP_DISTANCE_SQ 10000000000.0      /* WarpTo's short warp: under 100 km it is a GotoPoint */

static inline XB_HD double xb_round_f32(double x) { return (double)(float)x; }

/* Stop(Ball*) on the row: a follower forgets its target; the mode becomes STOP. */
static inline XB_HD void xb_cmd_stop_row(const XbPark& P, int64_t r) {
    const int64_t m = XB_Q(P, XB_I_MODE, r);
    if (m == XB_MODE_FOLLOW || m == XB_MODE_ORBIT) {
        XB_Q(P, XB_I_EFFECT_STAMP, r) = 0;
        XB_Q(P, XB_I_FOLLOW_PTR, r) = -1;
        XB_Q(P, XB_I_FOLLOW_ID, r) = 0;
        XB_Q(P, XB_I_OWNER_ID, r) = 0;
        XB_F(P, XB_F_FOLLOW_RANGE, r) = 0.0;
    }
    XB_Q(P, XB_I_MODE, r) = XB_MODE_STOP;
}

static inline XB_HD void xb_cmd_stop(const XbPark& P, int64_t r) {
    if (XB_Q(P, XB_I_MODE, r) != XB_MODE_STOP) xb_cmd_stop_row(P, r);
}

static inline XB_HD void xb_cmd_goto_point(const XbPark& P, int64_t r, XbV3 p) {
    if (!v3_finite(p)) return;
    xb_cmd_stop_row(P, r);
    xb_st3(P, XB_F_GOTO, r, p);
    if (XB_F(P, XB_F_SPEED_FRACTION, r) == 0.0) XB_F(P, XB_F_SPEED_FRACTION, r) = 1.0;
    XB_Q(P, XB_I_MODE, r) = XB_MODE_GOTO;
}

/* FollowBall / Orbit: the target a row of the park (n rows), another ball of the same world and bubble, not
   moribund, not cloaked; the range rounded to float and finite. */
static inline XB_HD void xb_cmd_follow(const XbPark& P, int64_t n, int64_t r, int64_t t, double range, int64_t mode) {
    const double rng = xb_round_f32(range);
    if (!xb_finite(rng) || t < 0 || t >= n || t == r) return;
    if (XB_Q(P, XB_I_WORLD, t) != XB_Q(P, XB_I_WORLD, r) || XB_B(P, XB_B_MORIBUND, t) ||
        XB_Q(P, XB_I_BUBBLE, t) != XB_Q(P, XB_I_BUBBLE, r) || XB_Q(P, XB_I_CLOAKED, t) != 0)
        return;
    xb_cmd_stop_row(P, r);
    XB_Q(P, XB_I_FOLLOW_ID, r) = XB_Q(P, XB_I_IDS, t);
    XB_Q(P, XB_I_FOLLOW_PTR, r) = t;
    XB_F(P, XB_F_FOLLOW_RANGE, r) = rng;
    XB_Q(P, XB_I_MODE, r) = mode;
}

/* WarpTo: a finite point; under 100 km a GotoPoint; nothing if already warping to that very point; else the warp set
   up to start at the next tick (effect stamp -1), the minimum range's bits kept in follow_id and the warp factor in
   owner_id (min range below 0 taken as 0, factor below 1 as 1). */
static inline XB_HD void xb_cmd_warp_to(const XbPark& P, int64_t r, XbV3 p, double min_range, int64_t warp_factor) {
    if (!v3_finite(p)) return;
    const double mr = min_range < 0.0 ? 0.0 : min_range;
    const int64_t wf = warp_factor <= 0 ? 1 : warp_factor;
    const XbV3 d = v3_sub(xb_ld3(P, XB_F_POS, r), p);
    if (v3_lensq(d) < XB_MIN_WARP_DISTANCE_SQ) {
        xb_cmd_goto_point(P, r, p);
        return;
    }
    if (XB_Q(P, XB_I_MODE, r) == XB_MODE_WARP) {
        const XbV3 g = xb_ld3(P, XB_F_GOTO, r);
        if (g.x == p.x && g.y == p.y && g.z == p.z) return;
    }
    xb_cmd_stop_row(P, r);
    xb_st3(P, XB_F_GOTO, r, p);
    XB_Q(P, XB_I_EFFECT_STAMP, r) = -1;
    XB_F(P, XB_F_FOLLOW_RANGE, r) = 0.0;
    XB_Q(P, XB_I_FOLLOW_ID, r) = (int64_t)xb_bits(mr);
    XB_Q(P, XB_I_OWNER_ID, r) = wf;
    XB_Q(P, XB_I_MODE, r) = XB_MODE_WARP;
}

/* Row r's command of a batch: op 0 none, else one of XB_CMD_*; arguments at row r of each array. */
static inline XB_HD void xb_cmd_apply(const XbPark& P, int64_t n, int64_t r, const uint8_t* op, const double* point,
                                      const int64_t* target, const double* range, const double* min_range,
                                      const int64_t* warp_factor) {
    const int o = op[r];
    if (o == 0) return;
    const XbV3 p = (o == XB_CMD_GOTO || o == XB_CMD_WARP) ? v3(point[r * 3], point[r * 3 + 1], point[r * 3 + 2])
                                                          : v3(0.0, 0.0, 0.0);
    switch (o) {
        case XB_CMD_STOP: xb_cmd_stop(P, r); break;
        case XB_CMD_GOTO: xb_cmd_goto_point(P, r, p); break;
        case XB_CMD_FOLLOW: xb_cmd_follow(P, n, r, target[r], range[r], XB_MODE_FOLLOW); break;
        case XB_CMD_ORBIT: xb_cmd_follow(P, n, r, target[r], range[r], XB_MODE_ORBIT); break;
        case XB_CMD_WARP: xb_cmd_warp_to(P, r, p, min_range[r], warp_factor[r]); break;
        default: break;
    }
}
