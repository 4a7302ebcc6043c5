// SPDX-License-Identifier: MIT
//
// The server layer's pre-tick, one lane a world, run by the host build and the gfx950 kernels alike: Destiny's
// net/server for a world of one bubble (Actions' queue with its duplicate test, the Ticker's tick, BubbleUpdater's
// additions, the history's routes, ParkUpdateBatcher's send flags), as the reference's batched layer does it
// (src/exact_ballpark/vec_server.py:660-1066; its kernel vec_server_cuda.cu:445-595; the scalar layer server.py:397-1143
// with Destiny's _actions.py, _baseticker.py, _ticker.py, _bubbleupdater.py and _parkupdatebatcher.py).
//
// A world's calls run in queue order on its own lane, each applied to the park at once by the park's own commands and
// setters (xb_commands.h), so a call reads the rows the calls before it wrote, as Destiny's Ticker applies them one by
// one. Worlds share nothing: a lane reads and writes only its world's rows and its world's entries of every table.
//
// What a tick writes for the sends: each call's queued flag, rows and route; the bubble update's lists; each
// UncloakBall's recipients; the send flags; and, when asked (snapshot), stream records: the fields
// Ballpark::WriteBallToStream writes (Thunkers.cpp:3180-3377) of a world's balls at its bubble update and of an
// uncloaked ball as its UncloakBall left it. The host turns records into Destiny's byte streams.
//
// Include after xb_commands.h (the interpreter) and include/xb.h.
#pragma once

#include <stddef.h>

#define XB_SV_NONE 0
#define XB_SV_BUBBLE 1
#define XB_SV_CHARACTER 2

/* the calls' kinds (the reference's KINDS order, src/exact_ballpark/vec_server.py:120-125) */
#define XB_K_GOTO_POINT 1
#define XB_K_GOTO_DIRECTION 2
#define XB_K_STOP 3
#define XB_K_FOLLOW_BALL 4
#define XB_K_ORBIT 5
#define XB_K_WARP_TO 6
#define XB_K_ENTITY_WARP_IN 7
#define XB_K_SET_SPEED_FRACTION 8
#define XB_K_SET_MAX_SPEED 9
#define XB_K_SET_BALL_MASS 10
#define XB_K_SET_BALL_AGILITY 11
#define XB_K_SET_BALL_RADIUS 12
#define XB_K_SET_BALL_VELOCITY 13
#define XB_K_SET_BALL_MASSIVE 14
#define XB_K_SET_BALL_INTERACTIVE 15
#define XB_K_SET_BALL_POSITION 16
#define XB_K_CLOAK_BALL 17
#define XB_K_UNCLOAK_BALL 18
#define XB_K_SET_MAX_ANGULAR_SPEED 19
#define XB_K_SET_BALL_ANGULAR_VELOCITY 20
#define XB_K_SET_BALL_ROTATION 21
#define XB_K_SET_BALL_ANGULAR_AGILITY 22
#define XB_K_SET_MAX_ANGULAR_VELOCITY 23     /* queued and sent, never applied: the park has no such method */
#define XB_K_LAUNCH_MISSILE 24
#define XB_K_SET_BALL_TROLL 25
#define XB_K_SET_BALL_FREE 26

#define XB_DST_NORMAL_CLOAK 1                /* IDstConstants.h's DSTNORMALCLOAK */
#define XB_SV_REFUSED 1                      /* wflags: a call the batched park refuses (a mushroom as a target) */

/* A stream record's words (8 bytes each; doubles as their bits). */
#define XB_SR_ID 0
#define XB_SR_MODE 1
#define XB_SR_RADIUS 2
#define XB_SR_POS 3                          /* 3 words */
#define XB_SR_FLAGS 6                        /* bit 0 free, 1 massive, 2 interactive, 3 space junk, 4 moribund, 5 set */
#define XB_SR_MASS 7
#define XB_SR_CLOAKED 8
#define XB_SR_MAX_VEL 9
#define XB_SR_VEL 10                         /* 3 words */
#define XB_SR_AGILITY 13
#define XB_SR_SPEED_FRACTION 14
#define XB_SR_ROT 15                         /* 4 words */
#define XB_SR_MAX_ANG_VEL 19
#define XB_SR_ANG_VEL 20                     /* 3 words */
#define XB_SR_ROT_AGILITY 23
#define XB_SR_FORMATION 24
#define XB_SR_FOLLOW_ID 25
#define XB_SR_FOLLOW_RANGE 26
#define XB_SR_OWNER_ID 27
#define XB_SR_EFFECT_STAMP 28
#define XB_SR_GOTO 29                        /* 3 words */
#define XB_SR_LAST_COLLISION 32
#define XB_SR_ROW 33

struct XsvCtx {
    XbPark P;
    int64_t n;                 /* the park's rows */
    double* ypr;               /* the park's display orientation block, or null */
    const xb_server_state* S;
    const xb_server_queue* Q;
    int64_t w, W, N;
};

static inline XB_HD int64_t xsv_row(const XsvCtx& c, int64_t k) { return c.S->rows[k * c.W + c.w]; }

/* The world's row of the ball with Destiny id `id`, or -1 (ids are distinct within a world). */
static inline XB_HD int64_t xsv_lookup(const XsvCtx& c, int64_t id) {
    for (int64_t k = 0; k < c.N; ++k) {
        const int64_t x = xsv_row(c, k);
        if (x >= 0 && XB_Q(c.P, XB_I_IDS, x) == id) return x;
    }
    return -1;
}

/* StopAllFollowers of row r, over the world's rows. */
static inline XB_HD void xsv_release(const XsvCtx& c, int64_t r) {
    const int64_t id = XB_Q(c.P, XB_I_IDS, r);
    for (int64_t k = 0; k < c.N; ++k) {
        const int64_t x = xsv_row(c, k);
        if (x >= 0 && XB_Q(c.P, XB_I_FOLLOW_PTR, x) == r && XB_Q(c.P, XB_I_FOLLOW_ID, x) == id) xb_row_released(c.P, x);
    }
}

/* CloakBall (full.py:1744-1758): refused unless the mode, a signed char, is positive; on the master a normal cloak adds
   the ball's proximity sensor (AddProximitySensor's guard: float(radius + float(range)) not below 0); the followers
   released; cloaked with the mode's byte; non-massive. */
static inline XB_HD void xsv_cloak(const XsvCtx& c, int64_t r, int64_t mode_in, double range) {
    const int mode = (int)(int8_t)(uint8_t)(mode_in & 0xFF);
    if (r < 0 || mode <= 0) return;
    if (c.S->master && mode == XB_DST_NORMAL_CLOAK && !(xb_f32(XB_F(c.P, XB_F_RADIUS, r) + xb_f32(range)) < 0.0))
        XB_B(c.P, XB_B_SENSOR, r) = 1;
    xsv_release(c, r);
    XB_Q(c.P, XB_I_CLOAKED, r) = mode & 0xFF;
    XB_B(c.P, XB_B_MASSIVE, r) = 0;
}

/* UncloakBall (full.py:1760-1780): on the master a normal cloak's sensor removed; uncloaked; massive again unless in
   warp proper. */
static inline XB_HD void xsv_uncloak(const XsvCtx& c, int64_t r) {
    if (r < 0) return;
    if (c.S->master && XB_Q(c.P, XB_I_CLOAKED, r) == XB_DST_NORMAL_CLOAK) XB_B(c.P, XB_B_SENSOR, r) = 0;
    XB_Q(c.P, XB_I_CLOAKED, r) = 0;
    const bool warping = XB_Q(c.P, XB_I_MODE, r) == XB_MODE_WARP && XB_Q(c.P, XB_I_EFFECT_STAMP, r) >= 0;
    if (!warping) XB_B(c.P, XB_B_MASSIVE, r) = 1;
}

/* One queued call but CloakBall on row r (tr its target's row): the park method Destiny's Ticker calls through
   `_parent_<name>` (server.py:804-820), with the thunker's argument conversions (server.py:1538-1720). */
static inline XB_HD void xsv_apply(const XsvCtx& c, int kind, int64_t r, int64_t tr, const double* f, int64_t nn) {
    XbCommand cmd = {};
    for (int k = 0; k < 4; ++k) cmd.x[k] = f[k];
    switch (kind) {
        case XB_K_GOTO_POINT: cmd.op = XB_CMD_GOTO; break;
        case XB_K_GOTO_DIRECTION: cmd.op = XB_CMD_GOTO_DIRECTION; break;
        case XB_K_STOP: cmd.op = XB_CMD_STOP; break;
        case XB_K_FOLLOW_BALL: cmd.op = XB_CMD_FOLLOW, cmd.i[0] = tr; break;
        case XB_K_ORBIT: cmd.op = XB_CMD_ORBIT, cmd.i[0] = tr; break;
        case XB_K_WARP_TO: cmd.op = XB_CMD_WARP, cmd.i[0] = nn; break;
        case XB_K_ENTITY_WARP_IN: cmd.op = XB_CMD_WARP_IN, cmd.i[0] = nn; break;
        case XB_K_SET_SPEED_FRACTION: cmd.op = XB_CMD_SET_SPEED_FRACTION; break;
        case XB_K_SET_MAX_SPEED: cmd.op = XB_CMD_SET_MAX_VEL; break;
        case XB_K_SET_BALL_MASS: cmd.op = XB_CMD_SET_MASS; break;
        case XB_K_SET_BALL_AGILITY: cmd.op = XB_CMD_SET_AGILITY; break;
        case XB_K_SET_BALL_RADIUS: cmd.op = XB_CMD_SET_RADIUS; break;
        case XB_K_SET_BALL_VELOCITY: cmd.op = XB_CMD_SET_VELOCITY; break;
        case XB_K_SET_BALL_INTERACTIVE: cmd.op = XB_CMD_SET_INTERACTIVE, cmd.flag = nn != 0; break;
        case XB_K_SET_MAX_ANGULAR_SPEED: cmd.op = XB_CMD_SET_MAX_ANG_SPEED; break;
        case XB_K_SET_BALL_ANGULAR_VELOCITY: cmd.op = XB_CMD_SET_ANG_VEL; break;
        case XB_K_SET_BALL_ROTATION: cmd.op = XB_CMD_SET_ROTATION; break;
        case XB_K_SET_BALL_ANGULAR_AGILITY: cmd.op = XB_CMD_SET_ROT_AGILITY; break;
        case XB_K_SET_BALL_TROLL: cmd.op = XB_CMD_SET_TROLL, cmd.i[0] = nn; break;
        case XB_K_SET_BALL_FREE: cmd.op = XB_CMD_SET_FREE, cmd.flag = nn != 0; break;
        case XB_K_SET_BALL_MASSIVE:                     /* SetBallMassive(ball, flag != 0) */
            XB_B(c.P, XB_B_MASSIVE, r) = nn != 0;
            return;
        case XB_K_SET_BALL_POSITION:                    /* SetBallPosition: the new and old position, re-filed */
            xb_st3(c.P, XB_F_POS, r, v3(f[0], f[1], f[2]));
            xb_st3(c.P, XB_F_OLD_POS, r, v3(f[0], f[1], f[2]));
            xb_row_refile(c.P, r);
            return;
        case XB_K_UNCLOAK_BALL: xsv_uncloak(c, r); return;
        default: return;                                /* SetMaxAngularVelocity: no park method */
    }
    if (xb_cmd_gather(c.P, c.n, r, cmd, c.ypr)) {
        c.S->wflags[c.w] |= XB_SV_REFUSED;
        return;
    }
    if (xb_cmd_run(c.P, r, cmd, c.ypr)) xsv_release(c, r);
}

/* The fields WriteBallToStream writes of row x, into record `rec`. */
static inline XB_HD void xsv_record(const XsvCtx& c, int64_t x, int64_t* rec) {
    const XbPark& P = c.P;
    auto d = [&](int q, double v) { rec[q] = (int64_t)xb_bits(v); };
    rec[XB_SR_ID] = XB_Q(P, XB_I_IDS, x);
    rec[XB_SR_MODE] = XB_Q(P, XB_I_MODE, x);
    d(XB_SR_RADIUS, XB_F(P, XB_F_RADIUS, x));
    for (int k = 0; k < 3; ++k) {
        d(XB_SR_POS + k, XB_F(P, XB_F_POS + k, x));
        d(XB_SR_VEL + k, XB_F(P, XB_F_VEL + k, x));
        d(XB_SR_ANG_VEL + k, XB_F(P, XB_F_ANG_VEL + k, x));
        d(XB_SR_GOTO + k, XB_F(P, XB_F_GOTO + k, x));
    }
    for (int k = 0; k < 4; ++k) d(XB_SR_ROT + k, XB_F(P, XB_F_ROT + k, x));
    rec[XB_SR_FLAGS] = (XB_B(P, XB_B_FREE, x) ? 1 : 0) | (XB_B(P, XB_B_MASSIVE, x) ? 2 : 0) |
                       (XB_B(P, XB_B_INTERACTIVE, x) ? 4 : 0) | (XB_B(P, XB_B_JUNK, x) ? 8 : 0) |
                       (XB_B(P, XB_B_MORIBUND, x) ? 16 : 0) | 32;
    d(XB_SR_MASS, XB_F(P, XB_F_MASS, x));
    rec[XB_SR_CLOAKED] = XB_Q(P, XB_I_CLOAKED, x);
    d(XB_SR_MAX_VEL, XB_F(P, XB_F_MAX_VEL, x));
    d(XB_SR_AGILITY, XB_F(P, XB_F_AGILITY, x));
    d(XB_SR_SPEED_FRACTION, XB_F(P, XB_F_SPEED_FRACTION, x));
    d(XB_SR_MAX_ANG_VEL, XB_F(P, XB_F_MAX_ANG_VEL, x));
    d(XB_SR_ROT_AGILITY, XB_F(P, XB_F_ROT_AGILITY, x));
    rec[XB_SR_FORMATION] = (int64_t)(int8_t)XB_B(P, XB_B_FORMATION, x);
    rec[XB_SR_FOLLOW_ID] = XB_Q(P, XB_I_FOLLOW_ID, x);
    d(XB_SR_FOLLOW_RANGE, XB_F(P, XB_F_FOLLOW_RANGE, x));
    rec[XB_SR_OWNER_ID] = XB_Q(P, XB_I_OWNER_ID, x);
    rec[XB_SR_EFFECT_STAMP] = XB_Q(P, XB_I_EFFECT_STAMP, x);
    d(XB_SR_LAST_COLLISION, XB_F(P, XB_F_LAST_COLLISION, x));
    rec[XB_SR_ROW] = x;
}

/* The queue in order: each call tested against the world's last queued call (Actions._add_to_system_history's
   duplicate test, server.py:413-425, as the pre-tick meets the calls), its rows looked up, and, with `apply`, every
   queued call but CloakBall applied to the park (Ticker._apply_park_parent_methods_except_cloak_ball). */
static inline XB_HD void xsv_queue(const XsvCtx& c, bool apply) {
    const xb_server_state& S = *c.S;
    const xb_server_queue& Q = *c.Q;
    const int64_t w = c.w, W = c.W;
    bool has_last = false;
    int64_t lk = 0, lb = 0, lt = 0, ln = 0, ll = 0;
    double lf[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t j = 0; j < Q.J; ++j) {
        const int64_t i = j * W + w;
        const int kind = Q.kind[j];
        const int64_t b = Q.ball[i], t = Q.target[i], nn = Q.n[i];
        const double* f = Q.f + 4 * i;
        const int64_t r = xsv_lookup(c, b);
        const int64_t tr = (kind == XB_K_FOLLOW_BALL || kind == XB_K_ORBIT) ? xsv_lookup(c, t) : -1;
        const uint8_t mv = Q.mask[i];
        bool pushed = (mv & 1) != 0;
        int64_t loc = Q.loc_in[i];
        if (kind == XB_K_CLOAK_BALL) {          /* its bubble as queued: 0, or -1 before its old bubble is its bubble */
            loc = (r >= 0 && S.settled[r]) ? 0 : -1;
            pushed = pushed && r >= 0;          /* Actions.cloak_ball queues nothing for a ball not in the park */
        }
        const bool tested = (mv & 2) != 0;      /* an undo's outcome: queued as bit 0 says */
        const bool dup = !tested && has_last && lk == kind && lb == b && lt == t && lf[0] == f[0] && lf[1] == f[1] &&
                         lf[2] == f[2] && lf[3] == f[3] && ln == nn && ll == loc;
        const bool ok = pushed && !dup;
        Q.ok[i] = ok;
        Q.row[i] = r;
        Q.trow[i] = tr;
        Q.loc[i] = loc;
        if (!ok) continue;
        has_last = true;
        lk = kind, lb = b, lt = t, ln = nn, ll = loc;
        for (int k = 0; k < 4; ++k) lf[k] = f[k];
        if (apply && kind != XB_K_CLOAK_BALL && r >= 0) xsv_apply(c, kind, r, tr, f, nn);
    }
}

/* The stamps, the queue and the bubble update (Ticker.tick up to _update_bubbles, server.py:656-776). */
static inline XB_HD void xsv_pass1(const XsvCtx& c) {
    const xb_server_state& S = *c.S;
    const xb_server_queue& Q = *c.Q;
    const XbPark& P = c.P;
    const int64_t w = c.w, W = c.W;
    const int64_t first = xsv_row(c, 0);
    const int64_t clock = first >= 0 ? XB_Q(P, XB_I_CLOCK, first) : 0;
    S.stamp[w] = S.stamp_for_system[w];
    S.clock[w] = clock;
    S.stamp_for_system[w] = clock + 1;
    xsv_queue(c, true);
    /* BubbleUpdater: AdditionsAndDeletions once per park time (Thunkers.cpp:2570-2895 for one bubble): a user ball
       whose old bubble is not yet its bubble gets every uncloaked ball, itself included, when there is one */
    bool any_listed = false;
    for (int64_t k = 0; k < c.N; ++k) {
        const int64_t x = xsv_row(c, k);
        if (x >= 0 && XB_Q(P, XB_I_CLOAKED, x) == 0) any_listed = true;
    }
    const bool again = clock == S.bubble_time[w];
    bool any_fresh = false;
    for (int64_t k = 0; k < c.N; ++k) {
        const int64_t x = xsv_row(c, k);
        if (x < 0) continue;
        if (!again) {
            const bool user = XB_B(P, XB_B_INTERACTIVE, x) != 0;
            S.fresh[x] = user && !S.settled[x] && any_listed;
            S.listed[x] = XB_Q(P, XB_I_CLOAKED, x) == 0;
            S.settled[x] = S.settled[x] || user;
        }
        any_fresh = any_fresh || S.fresh[x];
    }
    S.bubble_time[w] = clock;
    if (S.snapshot && Q.rec_add && any_fresh) {
        for (int64_t k = 0; k < c.N; ++k) {
            const int64_t x = xsv_row(c, k);
            int64_t* rec = Q.rec_add + (k * W + w) * XB_SREC_WORDS;
            if (x >= 0)
                xsv_record(c, x, rec);
            else
                rec[XB_SR_FLAGS] = 0;
        }
    }
}

static inline XB_HD bool xsv_live(const XsvCtx& c) {
    for (int64_t k = 0; k < c.N; ++k) {
        const int64_t x = xsv_row(c, k);
        if (x >= 0 && XB_B(c.P, XB_B_INTERACTIVE, x)) return true;
    }
    return false;
}

/* An event about row r goes to the bubble; while the ball is cloaked (at this point of the history) to its character
   instead, or nowhere without one (Ticker._apply_history, server.py:781-802). */
static inline XB_HD int8_t xsv_route_of(const XsvCtx& c, int64_t r) {
    if (XB_Q(c.P, XB_I_CLOAKED, r) == 0) return XB_SV_BUBBLE;
    return c.S->pilot[r] >= 0 ? XB_SV_CHARACTER : XB_SV_NONE;
}

/* The history in queue order (Ticker._apply_history and its handlers, server.py:781-886). */
static inline XB_HD void xsv_pass2(const XsvCtx& c) {
    const xb_server_state& S = *c.S;
    const xb_server_queue& Q = *c.Q;
    const XbPark& P = c.P;
    const int64_t w = c.w, W = c.W;
    const bool live = xsv_live(c);
    for (int64_t k = 0; k < c.N; ++k) {        /* a character's history starts with its bubble additions */
        const int64_t x = xsv_row(c, k);
        if (x >= 0) S.history[x] = S.fresh[x];
    }
    bool bubble = false;
    for (int64_t j = 0; j < Q.J; ++j) {
        const int64_t i = j * W + w;
        const int kind = Q.kind[j];
        const bool ok = Q.ok[i] != 0;
        const int64_t r = Q.row[i];
        const bool valid = ok && r >= 0;
        int8_t route = XB_SV_NONE;
        if (kind == XB_K_CLOAK_BALL) {           /* to the bubble it was queued in, then applied */
            if (ok) {
                route = Q.loc[i] == 0 ? XB_SV_BUBBLE : XB_SV_NONE;
                xsv_cloak(c, r, Q.n[i], Q.f[4 * i]);
            }
        } else if (kind == XB_K_UNCLOAK_BALL) {  /* to its own character; AddBallsToPark to the others' characters */
            if (valid && live && XB_B(P, XB_B_INTERACTIVE, r) && S.pilot[r] >= 0) route = XB_SV_CHARACTER;
            for (int64_t k = 0; k < c.N; ++k) {
                const int64_t x = xsv_row(c, k);
                if (x < 0) continue;
                const bool to = valid && live && XB_B(P, XB_B_INTERACTIVE, x) && x != r && S.pilot[x] >= 0;
                Q.notice[j * c.n + x] = to;
                if (to) S.history[x] = 1;
            }
            if (S.snapshot && Q.rec_unc && valid) xsv_record(c, r, Q.rec_unc + i * XB_SREC_WORDS);
        } else if (kind == XB_K_SET_BALL_POSITION) {   /* a local ball's move is sent nowhere */
            if (valid && Q.loc[i] == 0 && live) route = xsv_route_of(c, r);
        } else if (valid && live && Q.ball[i] > 0) {
            route = xsv_route_of(c, r);
        }
        Q.route[i] = route;
        if (route == XB_SV_BUBBLE) bubble = true;
        if (route == XB_SV_CHARACTER) S.history[r] = 1;
    }
    S.bubble_history[w] = bubble;
}

/* ParkUpdateBatcher.send_batch's split (server.py:947-1001): a pilot gets a singlecast when its character has history
   (the bubble additions, a call routed to it, an uncloak's notice: `history`, from the pre-tick) and the bubble's
   narrowcast when the bubble has history and its ship is interactive (now: after the park's tick). */
static inline XB_HD void xsv_flags(const XsvCtx& c) {
    const xb_server_state& S = *c.S;
    const bool bubble = S.bubble_history[c.w] != 0;
    for (int64_t k = 0; k < c.N; ++k) {
        const int64_t x = xsv_row(c, k);
        if (x < 0) continue;
        const bool pilot = S.pilot[x] >= 0;
        S.has_character[x] = S.history[x] && pilot;
        S.gets_bubble[x] = bubble && XB_B(c.P, XB_B_INTERACTIVE, x) && pilot;
    }
}

/* A world without a tick (OnTick's mask): nothing of it changes and it sends nothing. */
static inline XB_HD void xsv_quiet(const XsvCtx& c) {
    for (int64_t k = 0; k < c.N; ++k) {
        const int64_t x = xsv_row(c, k);
        if (x < 0) continue;
        c.S->has_character[x] = 0;
        c.S->gets_bubble[x] = 0;
    }
}

/* World w's part of xb_server_pre_tick. */
static inline XB_HD void xsv_world(const XbPark& P, int64_t n, double* ypr, const xb_server_state& S,
                                   const xb_server_queue& Q, int phase, int64_t w) {
    const XsvCtx c{P, n, ypr, &S, &Q, w, S.W, S.N};
    if (phase == 2) {                          /* the test alone: the queued flags and rows, nothing applied */
        xsv_queue(c, false);
        return;
    }
    if (Q.on && !Q.on[w]) {
        xsv_quiet(c);
        return;
    }
    if (phase == 0) {
        xsv_pass1(c);
        xsv_pass2(c);
    }
    xsv_flags(c);
}

static inline bool xsv_args_ok(const xb_modes_state* park, const xb_server_state* s, const xb_server_queue* q,
                               int phase) {
    if (!park || !park->f64 || !park->i64 || !park->u8 || !s || !q || phase < 0 || phase > 2) return false;
    if (s->W < 0 || s->N < 0 || (s->W > 0 && (!s->rows || !s->pilot || !s->stamp_for_system || !s->stamp ||
                                               !s->clock || !s->bubble_time || !s->settled || !s->fresh ||
                                               !s->listed || !s->has_character || !s->gets_bubble || !s->wflags ||
                                               !s->history || !s->bubble_history)))
        return false;
    if (q->J < 0 || q->J > 4096) return false;
    if (q->J > 0 && (!q->kind || !q->mask || !q->ball || !q->target || !q->f || !q->n || !q->loc_in || !q->ok ||
                     !q->row || !q->trow || !q->loc || !q->route || !q->notice))
        return false;
    return true;
}

static inline void xsv_layout(int64_t* out) {
    out[0] = (int64_t)sizeof(xb_server_state);
    out[1] = (int64_t)offsetof(xb_server_state, wflags);
    out[2] = (int64_t)sizeof(xb_server_queue);
    out[3] = (int64_t)offsetof(xb_server_queue, rec_unc);
    out[4] = XB_SREC_WORDS;
}
