// SPDX-License-Identifier: MIT
//
// The master's space layer over a batched park: the partition's top boxes and their bubbles, the bubble dictionaries,
// the interactive and keep-alive sets, the dead bubbles, the proximity sensors, notification ranges and target tracking
// in libstdc++'s unordered_map order, glibc rand() for shuffled sensors, and the visibility and range queries, for W
// independent worlds of up to N slots each (the reference's VecSpace, src/exact_ballpark/vec_space.py:777-1180, which
// holds it to FullBallpark(is_master=True): space.py:569-1610). One sequential algorithm a world, run by one lane a
// world on gfx950 (or one loop iteration a world on the host build).
//
// The layout is the MI355X part: every per-world table is world-interleaved, entry k of world w at [k * W + w], so the
// 64 lanes of a wavefront (64 consecutive worlds) read entry k of their tables as one run of consecutive words. The
// boxes below the tops are not stored: a box exists while a ball is filed in it or below it, so a sensor's walk of
// GetNearbyBalls follows from the two filings (xb_order_key, as the gradient's contacts do). The tops that exist are a
// table with each one's bubble; a top exists while a filing reaches it.
//
// Exactness: the filing is xb_file's (the park's own, pinned); distances and the sensors' float sums are IEEE double
// and float operations in the reference's order; every order (bubble dictionaries, sensor lists, the proximity map,
// the ball dictionary) is integer bookkeeping.
//
// Include after xb_modes_tick.h (the park and its accessors; xb_collide.h's filing and order key) and xb_rand.h.
#pragma once

#include "../../include/xb.h"

#define XS_OP_NONE 0
#define XS_OP_ADD 1                     /* AddBall's filing of slot a (InsertBallInBoxes) */
#define XS_OP_REFILE 2                  /* SetBallRadius, SetMaxSpeed, SetBallPosition: slot a re-filed */
#define XS_OP_INIT_BUBBLES 3            /* a < 0: InitializeBubbles; else UpdateBallBubble of slot a */
#define XS_OP_COPY_BUBBLES 4
#define XS_OP_KEEP_ALIVE 5              /* slot a into bubbleKeepAliveBalls */
#define XS_OP_SET_INTERACTIVE 6         /* SetBallInteractive(slot a, b != 0) */
#define XS_OP_ADD_SENSOR 7              /* AddProximitySensor(slot a, range f[0], period f[1], interactives only b,
                                           shuffled f[2] != 0) */
#define XS_OP_REMOVE_SENSOR 8
#define XS_OP_NOTIFY_RANGE 9            /* SetNotificationRange(slot a, f[0]) */
#define XS_OP_TRACK 10                  /* SetTargetTracking(slot a, the target's id b (0: off), range f[0]) */
#define XS_OP_REMOVE_BALL 11            /* RemoveBall's partition side for slot a (after the park's own removal) */
#define XS_OP_DEAD 12                   /* phase 1's dead bubbles: their rows flagged moribund for the park's tick */
#define XS_OP_AFTER_EVOLVE 13           /* after the park's tick: the drop-outs' and commits' re-filings, proximities */
#define XS_OP_CLEAR_EVENTS 14
#define XS_OP_CLOAK_SENSOR 15           /* CloakBall's sensor on the master: mode b (a char), range f[0] */
#define XS_OP_UNCLOAK 16                /* UncloakBall's part on the master, before the park's: the normal cloak's sensor
                                           removed; unless in warp proper, slot a out of its bubble's sensor lists */
#define XS_OPS 17

#define XS_Q_VISIBILITY 1               /* CheckVisibility(slot a, slot b): out[0] the occluder's id, 0 or -2 */
#define XS_Q_RANGE_BALL 2               /* GetBallIdsInRange(slot a, range f[3], cloaked b): out[0] n (-1: no such ball),
                                           then the ids; outf the squared distances */
#define XS_Q_RANGE_POINT 3              /* GetBallIdsInRange(point f[0..2], range f[3], cloaked b) */
#define XS_Q_BUBBLE_AT 4                /* GetBubbleAtCoordinates(f[0..2]): out[0] the bubble, XS_NO_TOP for none */
#define XS_NO_TOP INT64_MIN
#define XS_Q_REMOTE_FOLLOWERS 5         /* GetRemoteFollowers: out[0] n, ids ascending */

#define XS_F_EVENTS 1                   /* the world's flags: a table past its room, or outside the layer's domain */
#define XS_F_TOPS 2
#define XS_F_ENTRIES 4
#define XS_F_DICTS 8
#define XS_F_SETS 16
#define XS_F_MAP 32
#define XS_F_HUGE 64
#define XS_F_BAD_OP 128

#define XS_EV_PROXIMITY 1
#define XS_EV_RANGE 2
#define XS_EV_TRACKING 3
#define XS_EV_FINAL_CLEANUP 4

/* a slot's flag bits */
#define XS_IN 1                         /* in the park's ball dictionary */
#define XS_DEAD 2                       /* in a dead bubble at this tick's start (moribund for the park's tick) */
#define XS_WARPING 4                    /* in warp proper at this tick's start */
#define XS_PROX 8                       /* Ball::mHasProximity */
#define XS_ONLY 16                      /* the sensor senses interactive balls only */
#define XS_TRK_IN 32
#define XS_NTF_IN 64
#define XS_KEEP 128                     /* in bubbleKeepAliveBalls */

#define XS_TOP_WIDTH (XB_BIGBOX)
#define XS_TRACKING_MIN 0.00001         /* Ball.cpp:598 */
#define XS_RAND_MAX 2147483647.0        /* floatrand's RAND_MAX */
#define XS_CLOAK_SENSOR_PERIOD 2.0      /* CloakBall's sensor period (Ballpark.cpp:5243) */

typedef xb_space_state XsState;

/* ------------------------------------------------------------------------------------------------ access */

struct XsCtx {
    const XsState* s;
    XbPark P;
    int64_t w, W, N;
};

#define XS_AT(c, k) ((int64_t)(k) * (c).W + (c).w)
#define XS_SLOT(c, fld, j) ((c).s->fld[XS_AT(c, j)])
#define XS_WORLD(c, fld) ((c).s->fld[(c).w])

static inline XB_HD int64_t xs_row(const XsCtx& c, int j) { return (int64_t)XS_SLOT(c, rows, j); }
static inline XB_HD bool xs_in(const XsCtx& c, int j) { return (XS_SLOT(c, sflags, j) & XS_IN) != 0; }
static inline XB_HD void xs_flag(const XsCtx& c, int64_t bit) { XS_WORLD(c, wflags) |= bit; }
static inline XB_HD int64_t xs_id(const XsCtx& c, int j) { return XS_SLOT(c, ids, j); }

static inline XB_HD XbFiling xs_filing(const XsCtx& c, int j) {
    XbFiling f;
    f.level = (int)XS_SLOT(c, level, j);
    for (int q = 0; q < 3; ++q) {
        f.key[q] = c.s->key[((int64_t)q * c.N + j) * c.W + c.w];
        f.mod[q] = (int)c.s->mod[((int64_t)q * c.N + j) * c.W + c.w];
    }
    return f;
}
static inline XB_HD void xs_put_filing(const XsCtx& c, int j, const XbFiling& f) {
    XS_SLOT(c, level, j) = (int8_t)f.level;
    for (int q = 0; q < 3; ++q) {
        c.s->key[((int64_t)q * c.N + j) * c.W + c.w] = f.key[q];
        c.s->mod[((int64_t)q * c.N + j) * c.W + c.w] = (int8_t)f.mod[q];
    }
}

/* the park row's ship as the filing reads it, at a given position and velocity */
static inline XB_HD XbShip xs_ship(const XsCtx& c, int64_t r, XbV3 pos, XbV3 vel) {
    XbShip s = {};
    s.pos = pos;
    s.vel = vel;
    s.radius = XB_F(c.P, XB_F_RADIUS, r);
    s.mass = XB_F(c.P, XB_F_MASS, r);
    s.agility = XB_F(c.P, XB_F_AGILITY, r);
    s.sf = XB_F(c.P, XB_F_SPEED_FRACTION, r);
    s.max_vel = XB_F(c.P, XB_F_MAX_VEL, r);
    s.tf = XB_F(c.P, XB_F_TIME_FACTOR, r);
    return s;
}

/* ------------------------------------------------------------------------------------- the unordered maps */

/* std::unordered_map<int64, ...>'s iteration order under libstdc++'s prime rehash policy (space.py:193-320): the keys
   (slots here, hashed by their balls' ids) in order, the bucket count, the policy's next resize. */
#define XS_N_PRIMES 61
XB_TABLE static const int32_t XS_PRIMES[XS_N_PRIMES] = {
    2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 103, 109, 113, 127,
    137, 139, 149, 157, 167, 179, 193, 199, 211, 227, 241, 257, 277, 293, 313, 337, 359, 383, 409, 439, 467, 503, 541,
    577, 619, 661, 709, 761, 823, 887, 953, 1031};
XB_TABLE static const int32_t XS_FAST_BKT[14] = {2, 2, 2, 3, 5, 5, 7, 7, 11, 11, 11, 11, 13, 13};

struct XsMap {                          /* a map's place in the state: its key row, length, buckets, next resize */
    int8_t* keys;                       /* key i at keys[i * stride] */
    int64_t stride;
    int32_t* len;
    int32_t* bkt;
    int32_t* next;
};

static inline XB_HD int32_t xs_next_bkt(int32_t* next, int64_t want) {
    if (want < 14) {
        if (want == 0) return 1;
        *next = XS_FAST_BKT[want];
        return XS_FAST_BKT[want];
    }
    for (int j = 6; j < XS_N_PRIMES; ++j)
        if (XS_PRIMES[j] >= want) {
            *next = XS_PRIMES[j];
            return XS_PRIMES[j];
        }
    return 0;
}

static inline XB_HD uint64_t xs_bucket_of(const XsCtx& c, int slot, int32_t bkt) {
    return (uint64_t)xs_id(c, slot) % (uint64_t)bkt;
}

/* the linking rule: before the first key of the order in the same bucket, else at the front */
static inline XB_HD void xs_map_link(const XsCtx& c, const XsMap& m, int slot) {
    const int32_t n = *m.len, bkt = *m.bkt;
    const uint64_t b = xs_bucket_of(c, slot, bkt);
    int at = 0;
    for (int i = 0; i < n; ++i)
        if (xs_bucket_of(c, m.keys[i * m.stride], bkt) == b) {
            at = i;
            break;
        }
    for (int i = n; i > at; --i) m.keys[i * m.stride] = m.keys[(i - 1) * m.stride];
    m.keys[at * m.stride] = (int8_t)slot;
    *m.len = n + 1;
}

static inline XB_HD int xs_map_find(const XsMap& m, int slot) {
    for (int i = 0; i < *m.len; ++i)
        if (m.keys[i * m.stride] == slot) return i;
    return -1;
}

/* insert(): nothing if present; else room asked for one more key, a growth re-linking every key in its old order */
static inline XB_HD bool xs_map_insert(const XsCtx& c, const XsMap& m, int slot, int cap) {
    if (xs_map_find(m, slot) >= 0) return true;
    const int32_t n = *m.len;
    if (n >= cap) return false;
    const int64_t total = (int64_t)n + 1;
    if (total > *m.next) {
        const int64_t floor_keys = *m.next == 0 ? 11 : 0;
        const double want_f = (double)(total > floor_keys ? total : floor_keys);
        if (want_f >= (double)*m.bkt) {
            const int64_t a = (int64_t)want_f + 1, b2 = (int64_t)*m.bkt * 2;
            const int32_t nb = xs_next_bkt(m.next, a > b2 ? a : b2);
            if (nb == 0) return false;
            *m.bkt = nb;
            int8_t old[64];
            for (int i = 0; i < n; ++i) old[i] = m.keys[i * m.stride];
            *m.len = 0;
            for (int i = 0; i < n; ++i) xs_map_link(c, m, old[i]);
        } else {
            *m.next = *m.bkt;
        }
    }
    xs_map_link(c, m, slot);
    return true;
}

static inline XB_HD void xs_map_erase(const XsMap& m, int slot) {
    const int i = xs_map_find(m, slot);
    if (i < 0) return;
    const int32_t n = *m.len;
    for (int k = i; k + 1 < n; ++k) m.keys[k * m.stride] = m.keys[(k + 1) * m.stride];
    *m.len = n - 1;
}

static inline XB_HD XsMap xs_prox_map(const XsCtx& c) {
    XsMap m;
    m.keys = c.s->p_keys + c.w;
    m.stride = c.W;
    m.len = c.s->p_len + c.w;
    m.bkt = c.s->p_bkt + c.w;
    m.next = c.s->p_next + c.w;
    return m;
}
static inline XB_HD XsMap xs_sensor_map(const XsCtx& c, int j) {
    XsMap m;
    m.keys = c.s->s_keys + ((int64_t)j * c.W + c.w);             /* key i of slot j at ((i N + j) W + w) */
    m.stride = c.N * c.W;
    m.len = c.s->s_len + XS_AT(c, j);
    m.bkt = c.s->s_bkt + XS_AT(c, j);
    m.next = c.s->s_next + XS_AT(c, j);
    return m;
}

/* ---------------------------------------------------------------------------------------------- the tops */

static inline XB_HD void xs_top_of(const XbFiling& f, int64_t box_x, int64_t box_y, int64_t box_z, int32_t t[3]) {
    const int sh = 2 * f.level;
    t[0] = (int32_t)(box_x >> sh), t[1] = (int32_t)(box_y >> sh), t[2] = (int32_t)(box_z >> sh);
}

static inline XB_HD int xs_top_find(const XsCtx& c, const int32_t t[3]) {
    const int32_t n = XS_WORLD(c, tn);
    const int64_t T = c.s->T;
    for (int i = 0; i < n; ++i) {
        const int64_t a = (int64_t)i * c.W + c.w;
        if (c.s->top_key[a] == t[0] && c.s->top_key[T * c.W + a] == t[1] && c.s->top_key[2 * T * c.W + a] == t[2])
            return i;
    }
    return -1;
}

/* a filing's tops (up to 8: the boxes' per-axis options, each shifted to level 0), distinct */
static inline XB_HD int xs_filing_tops(const XbFiling& f, int32_t out[8][3]) {
    int64_t o[3][2];
    int cn[3];
    for (int q = 0; q < 3; ++q) cn[q] = xb_axis_options(f, q, o[q]);
    int n = 0;
    for (int a = 0; a < cn[0]; ++a)
        for (int b = 0; b < cn[1]; ++b)
            for (int d = 0; d < cn[2]; ++d) {
                int32_t t[3];
                xs_top_of(f, o[0][a], o[1][b], o[2][d], t);
                bool dup = false;
                for (int k = 0; k < n; ++k) dup = dup || (out[k][0] == t[0] && out[k][1] == t[1] && out[k][2] == t[2]);
                if (!dup) out[n][0] = t[0], out[n][1] = t[1], out[n][2] = t[2], ++n;
            }
    return n;
}

static inline XB_HD bool xs_filing_reaches(const XbFiling& f, const int32_t t[3]) {
    if (f.level < 0) return false;
    int32_t tops[8][3];
    const int n = xs_filing_tops(f, tops);
    for (int k = 0; k < n; ++k)
        if (tops[k][0] == t[0] && tops[k][1] == t[1] && tops[k][2] == t[2]) return true;
    return false;
}

static inline XB_HD void xs_top_drop(const XsCtx& c, int i) {
    const int32_t n = XS_WORLD(c, tn);
    const int64_t T = c.s->T;
    for (int k = i; k + 1 < n; ++k) {
        const int64_t a = (int64_t)k * c.W + c.w, b = (int64_t)(k + 1) * c.W + c.w;
        for (int q = 0; q < 3; ++q) c.s->top_key[q * T * c.W + a] = c.s->top_key[q * T * c.W + b];
        c.s->top_bub[a] = c.s->top_bub[b];
        c.s->top_ref[a] = c.s->top_ref[b];
    }
    XS_WORLD(c, tn) = n - 1;
}

/* a top made (no bubble, no filing yet) unless it exists; its index, or -1 past the room */
static inline XB_HD int xs_top_make(const XsCtx& c, const int32_t t[3]) {
    const int i = xs_top_find(c, t);
    if (i >= 0) return i;
    const int32_t n = XS_WORLD(c, tn);
    if (n >= c.s->T) {
        xs_flag(c, XS_F_TOPS);
        return -1;
    }
    const int64_t T = c.s->T, a = (int64_t)n * c.W + c.w;
    for (int q = 0; q < 3; ++q) c.s->top_key[q * T * c.W + a] = t[q];
    c.s->top_bub[a] = -1;
    c.s->top_ref[a] = 0;
    XS_WORLD(c, tn) = n + 1;
    return n;
}

/* a filing's boxes counted into their tops (each top made as it is first reached) */
static inline XB_HD void xs_filing_acquire(const XsCtx& c, const XbFiling& f) {
    if (f.level < 0) return;
    int32_t tops[8][3];
    const int n = xs_filing_tops(f, tops);
    for (int k = 0; k < n; ++k) {
        const int i = xs_top_make(c, tops[k]);
        if (i >= 0) c.s->top_ref[(int64_t)i * c.W + c.w] += 1;
    }
}

/* a filing's boxes counted out of their tops; a top no filing reaches is released (its bubble lost), but `pin` (a box
   just made there keeps it) */
static inline XB_HD void xs_filing_release(const XsCtx& c, const XbFiling& f, const int32_t* pin) {
    if (f.level < 0) return;
    int32_t tops[8][3];
    const int n = xs_filing_tops(f, tops);
    for (int k = 0; k < n; ++k) {
        const int i = xs_top_find(c, tops[k]);
        if (i < 0) continue;
        int32_t& ref = c.s->top_ref[(int64_t)i * c.W + c.w];
        ref -= 1;
        const bool pinned = pin && tops[k][0] == pin[0] && tops[k][1] == pin[1] && tops[k][2] == pin[2];
        if (ref <= 0 && !pinned) xs_top_drop(c, i);
    }
}

static inline XB_HD int64_t xs_top_bubble(const XsCtx& c, const int32_t t[3]) {
    const int i = xs_top_find(c, t);
    return i < 0 ? -1 : c.s->top_bub[(int64_t)i * c.W + c.w];
}
static inline XB_HD void xs_top_set(const XsCtx& c, const int32_t t[3], int64_t bub) {
    const int i = xs_top_find(c, t);
    if (i >= 0) c.s->top_bub[(int64_t)i * c.W + c.w] = bub;
}

/* ------------------------------------------------------------------------ the dictionaries and the sets */

static inline XB_HD int xs_entry_find(const XsCtx& c, int64_t bub, int slot) {
    for (int i = 0; i < XS_WORLD(c, en); ++i) {
        const int64_t a = (int64_t)i * c.W + c.w;
        if (c.s->e_bub[a] == bub && c.s->e_slot[a] == slot) return i;
    }
    return -1;
}
static inline XB_HD void xs_entry_drop(const XsCtx& c, int i) {
    const int32_t n = XS_WORLD(c, en);
    for (int k = i; k + 1 < n; ++k) {
        const int64_t a = (int64_t)k * c.W + c.w, b = (int64_t)(k + 1) * c.W + c.w;
        c.s->e_bub[a] = c.s->e_bub[b], c.s->e_slot[a] = c.s->e_slot[b], c.s->e_act[a] = c.s->e_act[b];
    }
    XS_WORLD(c, en) = n - 1;
}
static inline XB_HD bool xs_dict_has(const XsCtx& c, int64_t bub) {
    for (int i = 0; i < XS_WORLD(c, dn); ++i)
        if (c.s->d_bub[(int64_t)i * c.W + c.w] == bub) return true;
    return false;
}
static inline XB_HD void xs_dict_ensure(const XsCtx& c, int64_t bub) {
    if (xs_dict_has(c, bub)) return;
    const int32_t n = XS_WORLD(c, dn);
    if (n >= c.s->D) {
        xs_flag(c, XS_F_DICTS);
        return;
    }
    c.s->d_bub[(int64_t)n * c.W + c.w] = bub;
    XS_WORLD(c, dn) = n + 1;
}

/* the interactive (set 0) and keep-alive (set 1) sets as (bubble, slot) pairs */
static inline XB_HD int64_t* xs_pb(const XsCtx& c, int set) { return set ? c.s->k_bub : c.s->i_bub; }
static inline XB_HD int8_t* xs_ps(const XsCtx& c, int set) { return set ? c.s->k_slot : c.s->i_slot; }
static inline XB_HD int32_t* xs_pn(const XsCtx& c, int set) { return (set ? c.s->kn : c.s->in_) + c.w; }

static inline XB_HD void xs_pair_add(const XsCtx& c, int set, int64_t bub, int slot) {
    int64_t* pb = xs_pb(c, set);
    int8_t* ps = xs_ps(c, set);
    int32_t* pn = xs_pn(c, set);
    for (int i = 0; i < *pn; ++i)
        if (pb[(int64_t)i * c.W + c.w] == bub && ps[(int64_t)i * c.W + c.w] == slot) return;
    if (*pn >= c.s->I) {
        xs_flag(c, XS_F_SETS);
        return;
    }
    pb[(int64_t)*pn * c.W + c.w] = bub, ps[(int64_t)*pn * c.W + c.w] = (int8_t)slot;
    *pn += 1;
}
static inline XB_HD void xs_pair_del(const XsCtx& c, int set, int64_t bub, int slot) {
    int64_t* pb = xs_pb(c, set);
    int8_t* ps = xs_ps(c, set);
    int32_t* pn = xs_pn(c, set);
    for (int i = 0; i < *pn; ++i)
        if (pb[(int64_t)i * c.W + c.w] == bub && ps[(int64_t)i * c.W + c.w] == slot) {
            for (int k = i; k + 1 < *pn; ++k) {
                pb[(int64_t)k * c.W + c.w] = pb[(int64_t)(k + 1) * c.W + c.w];
                ps[(int64_t)k * c.W + c.w] = ps[(int64_t)(k + 1) * c.W + c.w];
            }
            *pn -= 1;
            return;
        }
}
static inline XB_HD int xs_pair_count(const XsCtx& c, int set, int64_t bub) {
    const int64_t* pb = xs_pb(c, set);
    int n = 0;
    for (int i = 0; i < *xs_pn(c, set); ++i) n += pb[(int64_t)i * c.W + c.w] == bub;
    return n;
}

/* the slot's bubble written through to the park row (the park's commands compare bubbles) */
static inline XB_HD void xs_set_bubble(const XsCtx& c, int j, int64_t bub) {
    XS_SLOT(c, bubble, j) = bub;
    const int64_t r = xs_row(c, j);
    if (r >= 0) XB_Q(c.P, XB_I_BUBBLE, r) = bub;
}

/* Ballpark::AddToBubble (space.py:922-954): +1 a ball added, 0 kept, -1 removed (a just-added one dropped instead);
   then the interactive and keep-alive sets */
static inline XB_HD void xs_add_to_bubble(const XsCtx& c, int j) {
    const int64_t bub = XS_SLOT(c, bubble, j), old = XS_SLOT(c, old_bubble, j);
    if (bub != -1) {
        xs_dict_ensure(c, bub);
        const int e = xs_entry_find(c, bub, j);
        if (e < 0) {
            const int32_t n = XS_WORLD(c, en);
            if (n >= c.s->E) {
                xs_flag(c, XS_F_ENTRIES);
            } else {
                const int64_t a = (int64_t)n * c.W + c.w;
                c.s->e_bub[a] = bub, c.s->e_slot[a] = (int8_t)j, c.s->e_act[a] = 1;
                XS_WORLD(c, en) = n + 1;
            }
        } else if (c.s->e_act[(int64_t)e * c.W + c.w] == -1) {
            c.s->e_act[(int64_t)e * c.W + c.w] = 0;
        }
    }
    if (old != -1 && xs_dict_has(c, old)) {
        const int e = xs_entry_find(c, old, j);
        if (e >= 0) {
            int8_t& act = c.s->e_act[(int64_t)e * c.W + c.w];
            if (act == 0) act = -1;
            else if (act == 1) xs_entry_drop(c, e);
        }
    }
    const int64_t r = xs_row(c, j);
    if (r >= 0 && XB_B(c.P, XB_B_INTERACTIVE, r)) {
        if (bub != -1) xs_pair_add(c, 0, bub, j);
        if (old != -1) xs_pair_del(c, 0, old, j);
    }
    if (XS_SLOT(c, sflags, j) & XS_KEEP) {
        if (bub != -1) xs_pair_add(c, 1, bub, j);
        if (old != -1) xs_pair_del(c, 1, old, j);
    }
}

/* InDeadBubble (space.py:974-984) */
static inline XB_HD bool xs_dead(const XsCtx& c, int j) {
    const int64_t r = xs_row(c, j);
    if (XB_Q(c.P, XB_I_MODE, r) == XB_MODE_WARP) return false;
    const int64_t bub = XS_SLOT(c, bubble, j);
    return xs_pair_count(c, 0, bub) == 0 && xs_pair_count(c, 1, bub) == 0;
}

/* --------------------------------------------------------------------------------- InsertInBoxes, master */

static inline XB_HD int64_t xs_nearby_bubbles(const XsCtx& c, const int32_t t[3]) {   /* NearbyBubbles, i-j-k */
    for (int i = -1; i <= 1; ++i)
        for (int j = -1; j <= 1; ++j)
            for (int k = -1; k <= 1; ++k) {
                if (i == 0 && j == 0 && k == 0) continue;
                const int32_t u[3] = {t[0] + i, t[1] + j, t[2] + k};
                const int64_t b = xs_top_bubble(c, u);
                if (b != -1) return b;
            }
    return -1;
}

/* Ballpark::InsertInBoxes then Ball::InsertInBoxes (space.py:729-817) for slot j at (pos, vel) */
static inline XB_HD void xs_insert_at(const XsCtx& c, int j, XbV3 pos, XbV3 vel) {
    const int64_t r = xs_row(c, j);
    XbFiling f;
    const XbShip sh = xs_ship(c, r, pos, vel);
    const bool free_ = XB_B(c.P, XB_B_FREE, r) != 0;
    xb_file(sh, free_, c.P, f);
    if (f.level == 0 && sh.radius * 2 > xb_level_width(0)) xs_flag(c, XS_F_HUGE);
    int32_t top[3];
    xs_top_of(f, f.key[0], f.key[1], f.key[2], top);
    xs_top_make(c, top);                                               /* GetBox makes the box and its top */
    int64_t nb = xs_top_bubble(c, top);
    if (nb == -1) {
        XbFiling gone = xs_filing(c, j);                               /* DeleteFromBoxes */
        xs_filing_release(c, gone, top);
        gone.level = -1;
        xs_put_filing(c, j, gone);
        nb = xs_nearby_bubbles(c, top);
        if (nb == -1) nb = XS_WORLD(c, bubble_id)++;
        xs_top_set(c, top, nb);
    }
    if (XS_SLOT(c, bubble, j) != nb) {
        XS_SLOT(c, old_bubble, j) = XS_SLOT(c, bubble, j);
        xs_set_bubble(c, j, nb);
        xs_add_to_bubble(c, j);
    }
    const XbFiling now = xs_filing(c, j);
    if (now.level == f.level && now.key[0] == f.key[0] && now.key[1] == f.key[1] && now.key[2] == f.key[2] &&
        now.mod[0] == f.mod[0] && now.mod[1] == f.mod[1] && now.mod[2] == f.mod[2])
        return;                                                        /* the same box and sides */
    xs_filing_release(c, now, top);                                    /* out of the old boxes first */
    xs_put_filing(c, j, f);
    xs_filing_acquire(c, f);
    xs_top_set(c, top, nb);
}

static inline XB_HD void xs_insert(const XsCtx& c, int j) {
    const int64_t r = xs_row(c, j);
    xs_insert_at(c, j, xb_ld3(c.P, XB_F_POS, r), xb_ld3(c.P, XB_F_VEL, r));
}

/* ------------------------------------------------------------------------------------------ the bubbles */

/* the slots in the ball dictionary's order (the order they were added) */
static inline XB_HD int xs_dict_order(const XsCtx& c, int8_t* out) {
    int n = 0;
    for (int j = 0; j < c.N; ++j) {
        if (!xs_in(c, j)) continue;
        int k = n++;
        while (k > 0 && XS_SLOT(c, seq, out[k - 1]) > XS_SLOT(c, seq, j)) {
            out[k] = out[k - 1];
            --k;
        }
        out[k] = (int8_t)j;
    }
    return n;
}

/* Box::Traverse from top `start`: the seed written into every top reachable from it through neighbouring tops without
   a bubble (a flood to its fixed point; the recursion's visiting order changes nothing written) */
static inline XB_HD void xs_traverse(const XsCtx& c, int start, int64_t seed) {
    int64_t* bub = c.s->top_bub;
    const int64_t T = c.s->T;
    const int32_t n = XS_WORLD(c, tn);
    if (bub[(int64_t)start * c.W + c.w] != -1) return;
    bub[(int64_t)start * c.W + c.w] = seed;
    for (bool grew = true; grew;) {
        grew = false;
        for (int i = 0; i < n; ++i) {
            const int64_t a = (int64_t)i * c.W + c.w;
            if (bub[a] != -1) continue;
            for (int k = 0; k < n; ++k) {
                const int64_t b = (int64_t)k * c.W + c.w;
                if (bub[b] != seed || k == i) continue;
                bool near = true;
                for (int q = 0; q < 3; ++q) {
                    const int32_t d = c.s->top_key[q * T * c.W + a] - c.s->top_key[q * T * c.W + b];
                    near = near && d >= -1 && d <= 1;
                }
                if (near) {
                    bub[a] = seed;
                    grew = true;
                    break;
                }
            }
        }
    }
}

static inline XB_HD void xs_initialize_bubbles(const XsCtx& c) {       /* space.py:993-1015 */
    for (int i = 0; i < XS_WORLD(c, tn); ++i) c.s->top_bub[(int64_t)i * c.W + c.w] = -1;
    XS_WORLD(c, bubble_id) = 0;
    XS_WORLD(c, dn) = 0, XS_WORLD(c, en) = 0, XS_WORLD(c, in_) = 0, XS_WORLD(c, kn) = 0;
    for (int j = 0; j < c.N; ++j) XS_SLOT(c, sflags, j) &= (uint8_t)~XS_KEEP;
    int8_t order[64];
    const int n = xs_dict_order(c, order);
    for (int q = 0; q < n; ++q) {
        const int j = order[q];
        const XbFiling f = xs_filing(c, j);
        XS_SLOT(c, old_bubble, j) = XS_SLOT(c, bubble, j);
        if (f.level < 0) {
            xs_set_bubble(c, j, -1);
            continue;
        }
        int32_t t[3];
        xs_top_of(f, f.key[0], f.key[1], f.key[2], t);
        const int x = xs_top_find(c, t);
        const int64_t tb = c.s->top_bub[(int64_t)x * c.W + c.w];
        if (tb != -1) {
            xs_set_bubble(c, j, tb);
            xs_add_to_bubble(c, j);
            continue;
        }
        xs_traverse(c, x, XS_WORLD(c, bubble_id));
        xs_set_bubble(c, j, XS_WORLD(c, bubble_id));
        xs_add_to_bubble(c, j);
        XS_WORLD(c, bubble_id)++;
    }
}

static inline XB_HD void xs_update_ball_bubble(const XsCtx& c, int slot) {   /* space.py:1017-1033 */
    const double max_width = xb_level_width(0) * xb_level_width(0);
    int64_t hi_id = -1;
    const XbV3 me = xb_ld3(c.P, XB_F_POS, xs_row(c, slot));
    int8_t order[64];
    const int n = xs_dict_order(c, order);
    for (int q = 0; q < n; ++q) {
        const int j = order[q];
        const int64_t b = XS_SLOT(c, bubble, j);
        if (b > hi_id) hi_id = b;
        if (j == slot || b == -1) continue;
        if (v3_lensq(v3_sub(me, xb_ld3(c.P, XB_F_POS, xs_row(c, j)))) < max_width) {
            xs_set_bubble(c, slot, b);
            break;
        }
    }
    if (XS_SLOT(c, bubble, slot) == -1) xs_set_bubble(c, slot, hi_id + 1);
}

static inline XB_HD void xs_copy_bubbles(const XsCtx& c) {              /* space.py:1035-1051 */
    for (int d = 0; d < XS_WORLD(c, dn);) {
        const int64_t bub = c.s->d_bub[(int64_t)d * c.W + c.w];
        int left = 0;
        for (int i = 0; i < XS_WORLD(c, en);) {
            const int64_t a = (int64_t)i * c.W + c.w;
            if (c.s->e_bub[a] != bub) {
                ++i;
                continue;
            }
            if (c.s->e_act[a] == 1) c.s->e_act[a] = 0;
            if (c.s->e_act[a] == -1) {
                xs_entry_drop(c, i);
                continue;
            }
            ++left;
            ++i;
        }
        if (left == 0) {
            for (int k = d; k + 1 < XS_WORLD(c, dn); ++k)
                c.s->d_bub[(int64_t)k * c.W + c.w] = c.s->d_bub[(int64_t)(k + 1) * c.W + c.w];
            XS_WORLD(c, dn)--;
        } else {
            ++d;
        }
    }
}

/* --------------------------------------------------------------------------------------------- the sensors */

static inline XB_HD void xs_event(const XsCtx& c, int64_t kind, int64_t id, int64_t a, int64_t b) {
    const int32_t n = XS_WORLD(c, evn);
    if (n >= c.s->M) {
        xs_flag(c, XS_F_EVENTS);
        return;
    }
    const int64_t M = c.s->M, at = (int64_t)n * c.W + c.w;
    c.s->ev[at] = kind, c.s->ev[M * c.W + at] = id, c.s->ev[2 * M * c.W + at] = a, c.s->ev[3 * M * c.W + at] = b;
    XS_WORLD(c, evn) = n + 1;
}

static inline XB_HD void xs_proximity_status(const XsCtx& c, int j) {   /* Ball::UpdateProximityStatus */
    const int64_t r = xs_row(c, j);
    uint8_t& fl = XS_SLOT(c, sflags, j);
    if (XS_SLOT(c, trk_range, j) != 0.0 || XS_SLOT(c, ntf_range, j) != 0.0 || XB_B(c.P, XB_B_SENSOR, r)) {
        fl |= XS_PROX;
        if (!xs_map_insert(c, xs_prox_map(c), j, (int)c.N)) xs_flag(c, XS_F_MAP);
    } else {
        fl &= (uint8_t)~XS_PROX;
        xs_map_erase(xs_prox_map(c), j);
    }
}

static inline XB_HD void xs_add_sensor(const XsCtx& c, int j, double range_in, double period, bool only,
                                       bool shuffle) {                  /* space.py:1354-1371 */
    const int64_t r = xs_row(c, j);
    const float range = (float)range_in;
    if ((float)XB_F(c.P, XB_F_RADIUS, r) + range < 0.0f) return;
    *xs_sensor_map(c, j).len = 0;
    XS_SLOT(c, s_range, j) = range;
    XS_SLOT(c, s_period, j) = period;
    XB_B(c.P, XB_B_SENSOR, r) = 1;
    uint8_t& fl = XS_SLOT(c, sflags, j);
    fl = only ? (uint8_t)(fl | XS_ONLY) : (uint8_t)(fl & ~XS_ONLY);
    if (shuffle) {
        int32_t st[XB_RAND_WORDS];
        for (int k = 0; k < XB_RAND_WORDS; ++k) st[k] = c.s->rnd[(int64_t)k * c.W + c.w];
        const int32_t draw = xb_rand(st);
        for (int k = 0; k < XB_RAND_WORDS; ++k) c.s->rnd[(int64_t)k * c.W + c.w] = st[k];
        XS_SLOT(c, s_elapsed, j) = 0.0 + (double)draw / XS_RAND_MAX * (period - 0.0);
    }
    xs_proximity_status(c, j);
}

/* GetNearbyBalls for slot j's sensor (exclude_fixed unless cloaked; exclude_inert when it senses interactives only),
   the candidates in the walk's order: by the pair's walk key from the two filings, then by id */
static inline XB_HD int xs_sensor_candidates(const XsCtx& c, int j, int8_t* out) {
    const int64_t r = xs_row(c, j);
    const bool only = (XS_SLOT(c, sflags, j) & XS_ONLY) != 0;
    const bool exclude_fixed = XB_Q(c.P, XB_I_CLOAKED, r) == 0;
    const int64_t bub = XS_SLOT(c, bubble, j);
    if (only && xs_pair_count(c, 0, bub) == 0) return 0;
    const XbFiling me = xs_filing(c, j);
    int64_t keys[64];
    int n = 0;
    for (int k = 0; k < c.N; ++k) {
        if (k == j || !xs_in(c, k)) continue;
        const int64_t o = xs_row(c, k);
        if (XB_B(c.P, XB_B_MORIBUND, o) && !(XS_SLOT(c, sflags, k) & XS_DEAD)) continue;
        if (only) {
            if (!XB_B(c.P, XB_B_INTERACTIVE, o)) continue;
        } else {
            if (XS_SLOT(c, bubble, k) != bub) continue;
            if (exclude_fixed && !XB_B(c.P, XB_B_FREE, o)) continue;
        }
        const XbFiling of = xs_filing(c, k);
        if (of.level < 0) continue;
        const int64_t key = xb_order_key(me, of);
        if (key == XB_NOT_RELATED) continue;
        const int64_t id = xs_id(c, k);
        int at = n++;
        while (at > 0 && (keys[at - 1] > key || (keys[at - 1] == key && xs_id(c, out[at - 1]) > id))) {
            keys[at] = keys[at - 1];
            out[at] = out[at - 1];
            --at;
        }
        keys[at] = key;
        out[at] = (int8_t)k;
    }
    return n;
}

/* a ball's moribund flag as the master's HandleProximities reads it: the park's (the dead bubbles' rows are not) */
static inline XB_HD bool xs_moribund(const XsCtx& c, int j) {
    return XB_B(c.P, XB_B_MORIBUND, xs_row(c, j)) != 0 && !(XS_SLOT(c, sflags, j) & XS_DEAD);
}

static inline XB_HD void xs_set_radius(const XsCtx& c, int j, double rad, XbV3 vel) {   /* SetBallRadius(Ball*) */
    if (rad < 0.0) return;
    const int64_t r = xs_row(c, j);
    XB_F(c.P, XB_F_RADIUS, r) = rad;
    xs_insert_at(c, j, xb_ld3(c.P, XB_F_POS, r), vel);
}

static inline XB_HD void xs_check_sensor(const XsCtx& c, int j) {      /* space.py:1476-1522 */
    const int64_t r = xs_row(c, j);
    XS_SLOT(c, s_elapsed, j) += c.P.dt;
    if (XS_SLOT(c, s_elapsed, j) < XS_SLOT(c, s_period, j)) return;
    XS_SLOT(c, s_elapsed, j) = 0.0;
    const double radius = XB_F(c.P, XB_F_RADIUS, r);
    const XbV3 vel = xb_ld3(c.P, XB_F_VEL, r);
    xs_set_radius(c, j, (double)((float)radius + XS_SLOT(c, s_range, j)), v3(0.0, 0.0, 0.0));
    int8_t uni[64];
    const int nu = xs_sensor_candidates(c, j, uni);
    xs_set_radius(c, j, radius, vel);
    const XsMap m = xs_sensor_map(c, j);
    const int64_t me = xs_id(c, j);
    if (*m.len > 0) {
        int8_t keys[64];
        const int nk = *m.len;
        for (int i = 0; i < nk; ++i) keys[i] = m.keys[i * m.stride];
        for (int i = 0; i < nk; ++i) {
            bool near = keys[i] == j;
            for (int u = 0; u < nu && !near; ++u) near = uni[u] == keys[i];
            if (near) continue;
            xs_map_erase(m, keys[i]);
            xs_event(c, XS_EV_PROXIMITY, me, xs_id(c, keys[i]), 0);
            if (!XB_B(c.P, XB_B_SENSOR, r)) break;
        }
    }
    for (int u = 0; u < nu; ++u) {
        if (xs_moribund(c, j)) break;
        const int k = uni[u];
        if (k == j || xs_moribund(c, k)) continue;
        const int64_t o = xs_row(c, k);
        const bool in = xs_map_find(m, k) >= 0;
        const double d2 = v3_lensq(v3_sub(xb_ld3(c.P, XB_F_POS, r), xb_ld3(c.P, XB_F_POS, o)));
        const float ck = (XS_SLOT(c, s_range, j) + (float)XB_F(c.P, XB_F_RADIUS, o)) + (float)XB_F(c.P, XB_F_RADIUS, r);
        const double ckd = (double)ck;
        if (d2 > ckd * ckd) {
            if (in) {
                xs_map_erase(m, k);
                xs_event(c, XS_EV_PROXIMITY, me, xs_id(c, k), 0);
            }
        } else if (!in) {
            if (!xs_map_insert(c, m, k, (int)c.N)) xs_flag(c, XS_F_MAP);
            xs_event(c, XS_EV_PROXIMITY, me, xs_id(c, k), 1);
        }
        if (!XB_B(c.P, XB_B_SENSOR, r)) break;
    }
}

static inline XB_HD int xs_slot_of_id(const XsCtx& c, int64_t id) {
    for (int j = 0; j < c.N; ++j)
        if (xs_in(c, j) && xs_id(c, j) == id) return j;
    return -1;
}

static inline XB_HD void xs_check_notification(const XsCtx& c, int j) { /* space.py:1434-1458 */
    const int64_t r = xs_row(c, j);
    const double nr = XS_SLOT(c, ntf_range, j);
    uint8_t& fl = XS_SLOT(c, sflags, j);
    bool send = false;
    if (nr != 0.0) {
        bool inside = false;
        const int64_t mode = XB_Q(c.P, XB_I_MODE, r);
        const XbV3 pos = xb_ld3(c.P, XB_F_POS, r);
        const double rad = XB_F(c.P, XB_F_RADIUS, r);
        if (mode == XB_MODE_GOTO) {
            const double lim = nr + rad;
            inside = v3_lensq(v3_sub(pos, xb_ld3(c.P, XB_F_GOTO, r))) < lim * lim;
        } else if (mode == XB_MODE_FOLLOW || mode == XB_MODE_MISSILE || mode == XB_MODE_ORBIT) {
            const int64_t o = XB_Q(c.P, XB_I_FOLLOW_PTR, r);
            const double len2 = v3_lensq(v3_sub(pos, xb_ld3(c.P, XB_F_POS, o)));
            const double hi = XB_F(c.P, XB_F_RADIUS, o) + XB_F(c.P, XB_F_FOLLOW_RANGE, r) + nr + rad;
            double lo = hi - 2.0 * nr;
            if (lo < 0.0) lo = 0.0;
            inside = len2 < hi * hi && len2 > lo * lo;
        }
        if (inside) {
            if (!(fl & XS_NTF_IN)) send = true, fl |= XS_NTF_IN;
        } else if (fl & XS_NTF_IN) {
            send = true, fl &= (uint8_t)~XS_NTF_IN;
        }
    }
    if (send) xs_event(c, XS_EV_RANGE, xs_id(c, j), (fl & XS_NTF_IN) ? 1 : 0, 0);
}

static inline XB_HD void xs_check_tracking(const XsCtx& c, int j) {     /* space.py:1460-1474 */
    const int64_t r = xs_row(c, j);
    const int t = XS_SLOT(c, trk, j);
    uint8_t& fl = XS_SLOT(c, sflags, j);
    bool send = false;
    if (XS_SLOT(c, trk_range, j) >= XS_TRACKING_MIN && t >= 0) {
        const XbV3 tp = v3(XS_SLOT(c, trk_pos, t), c.s->trk_pos[((int64_t)c.N + t) * c.W + c.w],
                           c.s->trk_pos[((int64_t)2 * c.N + t) * c.W + c.w]);
        const XbV3 tpos = xs_in(c, t) ? xb_ld3(c.P, XB_F_POS, xs_row(c, t)) : tp;
        const double d2 = v3_lensq(v3_sub(xb_ld3(c.P, XB_F_POS, r), tpos));
        const double lim = XS_SLOT(c, trk_range, j) + XB_F(c.P, XB_F_RADIUS, r);
        if (d2 < lim * lim) {
            if (!(fl & XS_TRK_IN)) send = true, fl |= XS_TRK_IN;
        } else if (fl & XS_TRK_IN) {
            send = true, fl &= (uint8_t)~XS_TRK_IN;
        }
    }
    if (send) xs_event(c, XS_EV_TRACKING, xs_id(c, j), (fl & XS_TRK_IN) ? 1 : 0, 0);
}

static inline XB_HD void xs_check_for_proximity(const XsCtx& c, int j) {
    if (xs_moribund(c, j)) return;
    xs_check_notification(c, j);
    xs_check_tracking(c, j);
    if (!XB_B(c.P, XB_B_SENSOR, xs_row(c, j))) return;
    xs_check_sensor(c, j);
}

static inline XB_HD void xs_handle_proximities(const XsCtx& c) {       /* space.py:1406-1422 */
    const XsMap m = xs_prox_map(c);
    int8_t keys[64];
    const int nk = *m.len;
    for (int i = 0; i < nk; ++i) keys[i] = m.keys[i * m.stride];
    for (int i = 0; i < nk; ++i) {
        const int j = keys[i];
        if (xs_moribund(c, j)) continue;
        if (xs_dead(c, j)) {
            const XsMap sm = xs_sensor_map(c, j);
            if (*sm.len > 0) {
                XS_SLOT(c, s_elapsed, j) = XS_SLOT(c, s_period, j);
                xs_check_for_proximity(c, j);
            }
            *sm.len = 0;
            continue;
        }
        if (!(XS_SLOT(c, sflags, j) & XS_PROX)) continue;
        xs_check_for_proximity(c, j);
    }
}

/* ------------------------------------------------------------------------------------------ the operations */

static inline XB_HD void xs_remove_ball(const XsCtx& c, int j) {        /* space.py:1286-1317, the partition side */
    const int64_t r = xs_row(c, j);
    xs_event(c, XS_EV_FINAL_CLEANUP, xs_id(c, j), 0, 0);
    const XbV3 p = xb_ld3(c.P, XB_F_POS, r);                           /* kept for the tracking that outlives it */
    XS_SLOT(c, trk_pos, j) = p.x;
    c.s->trk_pos[((int64_t)c.N + j) * c.W + c.w] = p.y;
    c.s->trk_pos[((int64_t)2 * c.N + j) * c.W + c.w] = p.z;
    XbFiling f = xs_filing(c, j);
    xs_filing_release(c, f, nullptr);
    f.level = -1;
    xs_put_filing(c, j, f);
    XS_SLOT(c, old_bubble, j) = XS_SLOT(c, bubble, j);
    xs_set_bubble(c, j, -1);
    xs_add_to_bubble(c, j);
    if (XS_SLOT(c, sflags, j) & XS_PROX) xs_map_erase(xs_prox_map(c), j);
    XS_SLOT(c, sflags, j) &= (uint8_t)~XS_IN;
}

/* UncloakBall on the master (full.py:1760-1780), the layer's part: a normal cloak's proximity sensor removed
   (RemoveProximitySensor); unless the ball is in warp proper, it leaves the sensor list of every live proximity ball of
   its bubble. The park's part (the cloak byte, massive again) follows. */
static inline XB_HD void xs_uncloak(const XsCtx& c, int j) {
    const int64_t r = xs_row(c, j);
    if (XB_Q(c.P, XB_I_CLOAKED, r) == 1) {
        XB_B(c.P, XB_B_SENSOR, r) = 0;
        xs_proximity_status(c, j);
    }
    if (XB_Q(c.P, XB_I_MODE, r) == XB_MODE_WARP && XB_Q(c.P, XB_I_EFFECT_STAMP, r) >= 0) return;
    const int64_t bub = XS_SLOT(c, bubble, j);
    const XsMap prox = xs_prox_map(c);
    for (int i = 0; i < *prox.len; ++i) {
        const int k = prox.keys[(int64_t)i * prox.stride];
        if (XS_SLOT(c, bubble, k) != bub || xs_moribund(c, k) || !(XS_SLOT(c, sflags, k) & XS_PROX)) continue;
        xs_map_erase(xs_sensor_map(c, k), j);
    }
}

static inline XB_HD void xs_set_interactive(const XsCtx& c, int j, bool on) {
    const int64_t r = xs_row(c, j);
    const bool was = XB_B(c.P, XB_B_INTERACTIVE, r) != 0;
    const int64_t bub = XS_SLOT(c, bubble, j);
    if (on && !was) {
        XB_B(c.P, XB_B_INTERACTIVE, r) = 1;
        if (bub != -1) xs_pair_add(c, 0, bub, j);
    } else if (!on && was) {
        if (bub != -1) xs_pair_del(c, 0, bub, j);
        XB_B(c.P, XB_B_INTERACTIVE, r) = 0;
    }
}

/* phase 1 (before the park's tick): the dead bubbles' live rows flagged moribund for the tick, and the rows in warp
   proper noted (a drop-out is re-filed afterwards) */
static inline XB_HD void xs_mark_dead(const XsCtx& c) {
    for (int j = 0; j < c.N; ++j) {
        uint8_t& fl = XS_SLOT(c, sflags, j);
        fl &= (uint8_t)~(XS_DEAD | XS_WARPING);
        if (!(fl & XS_IN)) continue;
        const int64_t r = xs_row(c, j);
        if (!XB_B(c.P, XB_B_FREE, r) || XB_B(c.P, XB_B_MORIBUND, r)) continue;
        if (xs_dead(c, j)) {
            fl |= XS_DEAD;
            XB_B(c.P, XB_B_MORIBUND, r) = 1;
        } else if (XB_Q(c.P, XB_I_MODE, r) == XB_MODE_WARP && XB_Q(c.P, XB_I_EFFECT_STAMP, r) >= 0) {
            fl |= XS_WARPING;
        }
    }
}

static inline XB_HD void xs_after_evolve(const XsCtx& c) {
    int8_t order[64];
    int n = 0;
    for (int k = 0; k < c.N; ++k) {                                    /* the slots by ascending id */
        const int j = c.s->by_id[XS_AT(c, k)];
        if (j >= 0) order[n++] = (int8_t)j;
    }
    for (int q = 0; q < n; ++q) {                                      /* the dead rows' flag undone */
        const int j = order[q];
        if (XS_SLOT(c, sflags, j) & XS_DEAD) XB_B(c.P, XB_B_MORIBUND, xs_row(c, j)) = 0;
    }
    for (int q = 0; q < n; ++q) {                                      /* phase 3: the drop-outs at rest, re-filed */
        const int j = order[q];
        if (!xs_in(c, j) || !(XS_SLOT(c, sflags, j) & XS_WARPING)) continue;
        const int64_t r = xs_row(c, j);
        if (XB_Q(c.P, XB_I_MODE, r) == XB_MODE_WARP) continue;
        const XbV3 ov = xb_ld3(c.P, XB_F_OLD_VEL, r);
        if (v3_lensq(ov) < 0.1) xs_insert_at(c, j, xb_ld3(c.P, XB_F_OLD_POS, r), ov);
    }
    for (int q = 0; q < n; ++q) {                                      /* phase 4: the commits that moved */
        const int j = order[q];
        if (!xs_in(c, j) || (XS_SLOT(c, sflags, j) & XS_DEAD)) continue;
        const int64_t r = xs_row(c, j);
        if (!XB_B(c.P, XB_B_FREE, r) || XB_B(c.P, XB_B_MORIBUND, r)) continue;
        if (v3_lensq(v3_sub(xb_ld3(c.P, XB_F_POS, r), xb_ld3(c.P, XB_F_OLD_POS, r))) < 0.1) continue;
        xs_insert(c, j);
    }
    xs_handle_proximities(c);
    for (int j = 0; j < c.N; ++j) XS_SLOT(c, sflags, j) &= (uint8_t)~(XS_DEAD | XS_WARPING);
}

static inline XB_HD void xs_apply(const XsCtx& c, int64_t code, int64_t a, int64_t b, const double* f) {
    if (code == XS_OP_NONE) return;
    const bool slot = a >= 0 && a < c.N && xs_row(c, (int)a) >= 0;
    const int j = (int)a;
    switch (code) {
        case XS_OP_ADD:
            if (!slot) break;
            XS_SLOT(c, sflags, j) |= XS_IN;
            XS_SLOT(c, seq, j) = XS_WORLD(c, seq_next)++;
            XS_SLOT(c, ids, j) = XB_Q(c.P, XB_I_IDS, xs_row(c, j));
            xs_insert(c, j);
            break;
        case XS_OP_REFILE:
            if (slot && xs_in(c, j)) xs_insert(c, j);
            break;
        case XS_OP_INIT_BUBBLES:
            if (a < 0) xs_initialize_bubbles(c);
            else if (slot && xs_in(c, j)) xs_update_ball_bubble(c, j);
            break;
        case XS_OP_COPY_BUBBLES: xs_copy_bubbles(c); break;
        case XS_OP_KEEP_ALIVE:
            if (slot) XS_SLOT(c, sflags, j) |= XS_KEEP;
            break;
        case XS_OP_SET_INTERACTIVE:
            if (slot && xs_in(c, j)) xs_set_interactive(c, j, b != 0);
            break;
        case XS_OP_ADD_SENSOR:
            if (slot && xs_in(c, j)) xs_add_sensor(c, j, f[0], f[1], b != 0, f[2] != 0.0);
            break;
        case XS_OP_REMOVE_SENSOR:
            if (slot && xs_in(c, j)) XB_B(c.P, XB_B_SENSOR, xs_row(c, j)) = 0, xs_proximity_status(c, j);
            break;
        case XS_OP_NOTIFY_RANGE:
            if (slot && xs_in(c, j) && !(f[0] < 0.0)) {
                XS_SLOT(c, ntf_range, j) = f[0];
                xs_proximity_status(c, j);
                XS_SLOT(c, sflags, j) &= (uint8_t)~XS_NTF_IN;
            }
            break;
        case XS_OP_TRACK:
            if (!(slot && xs_in(c, j))) break;
            if (f[0] <= 0.0 || b == 0) {
                XS_SLOT(c, trk, j) = -1, XS_SLOT(c, trk_range, j) = 0.0;
                break;
            }
            {
                const int t = xs_slot_of_id(c, b);
                if (t < 0) break;
                XS_SLOT(c, trk, j) = t, XS_SLOT(c, trk_range, j) = f[0];
                xs_proximity_status(c, j);
                XS_SLOT(c, sflags, j) &= (uint8_t)~XS_TRK_IN;
            }
            break;
        case XS_OP_REMOVE_BALL:
            if (slot && xs_in(c, j)) xs_remove_ball(c, j);
            break;
        case XS_OP_DEAD: xs_mark_dead(c); break;
        case XS_OP_AFTER_EVOLVE: xs_after_evolve(c); break;
        case XS_OP_CLEAR_EVENTS: XS_WORLD(c, evn) = 0; break;
        case XS_OP_CLOAK_SENSOR:
            if (slot && xs_in(c, j) && (int)(int8_t)(uint8_t)(b & 0xFF) == 1)
                xs_add_sensor(c, j, f[0], XS_CLOAK_SENSOR_PERIOD, false, false);
            break;
        case XS_OP_UNCLOAK:
            if (slot && xs_in(c, j)) xs_uncloak(c, j);
            break;
        default: xs_flag(c, XS_F_BAD_OP); break;
    }
}

/* ---------------------------------------------------------------------------------------------- the queries */

static inline XB_HD int64_t xs_top_at(const XsCtx& c, XbV3 p, bool* found) {   /* GetBox(1e7, pos, create=false) */
    const double w = xb_level_width(0);
    const int64_t n = xb_level_grid(0);
    const double fac = 1.0 / w, half = 0.5 * w * (double)n;
    const double pos[3] = {p.x, p.y, p.z};
    int32_t t[3];
    for (int q = 0; q < 3; ++q) t[q] = (int32_t)xb_clamp_key(xb_to_int64((pos[q] + half) * fac), n);
    const int i = xs_top_find(c, t);
    *found = i >= 0;
    return i >= 0 ? c.s->top_bub[(int64_t)i * c.W + c.w] : -1;
}

/* out: [W][1 + N] int64 (word k of world w at out[k * W + w]), outf [W][N] */
static inline XB_HD void xs_query(const XsCtx& c, int64_t kind, int64_t a, int64_t b, const double* f, int64_t* out,
                                  double* outf) {
    const int64_t W = c.W;
    auto put = [&](int k, int64_t v) { out[(int64_t)k * W + c.w] = v; };
    switch (kind) {
        case XS_Q_VISIBILITY: {                                        /* space.py:1524-1548 */
            const bool ok = a >= 0 && a < c.N && b >= 0 && b < c.N && xs_in(c, (int)a) && xs_in(c, (int)b);
            if (!ok || XS_SLOT(c, bubble, (int)a) != XS_SLOT(c, bubble, (int)b)) {
                put(0, -2);
                break;
            }
            const XbV3 sp = xb_ld3(c.P, XB_F_POS, xs_row(c, (int)a));
            const XbV3 ray0 = v3_sub(xb_ld3(c.P, XB_F_POS, xs_row(c, (int)b)), sp);
            const double distance = v3_len(ray0);
            const XbV3 ray = v3_unit(ray0);
            int8_t order[64];
            const int n = xs_dict_order(c, order);
            put(0, 0);
            for (int q = 0; q < n; ++q) {
                const int k = order[q];
                const int64_t o = xs_row(c, k);
                if (XS_SLOT(c, bubble, k) != XS_SLOT(c, bubble, (int)a) || k == a || k == b ||
                    !XB_B(c.P, XB_B_MASSIVE, o) || XB_Q(c.P, XB_I_CLOAKED, o))
                    continue;
                const XbV3 tc = v3_sub(xb_ld3(c.P, XB_F_POS, o), sp);
                const double v = ray.x * tc.x + ray.y * tc.y + ray.z * tc.z;
                const double r = XB_F(c.P, XB_F_RADIUS, o);
                const double dsq = r * r - ((tc.x * tc.x + tc.y * tc.y + tc.z * tc.z) - v * v);
                if (dsq < 0.0) continue;
                const double d = __builtin_sqrt(dsq);
                if (v - d > 0.0 && v - d < distance) {
                    put(0, xs_id(c, k));
                    break;
                }
            }
            break;
        }
        case XS_Q_RANGE_BALL:
        case XS_Q_RANGE_POINT: {                                       /* space.py:1568-1610 */
            XbV3 centre;
            int64_t bub;
            int ref = -1;
            if (kind == XS_Q_RANGE_BALL) {
                if (!(a >= 0 && a < c.N && xs_in(c, (int)a))) {
                    put(0, -1);
                    break;
                }
                ref = (int)a;
                centre = xb_ld3(c.P, XB_F_POS, xs_row(c, ref));
                bub = XS_SLOT(c, bubble, ref);
            } else {
                centre = v3(f[0], f[1], f[2]);
                bool found;
                bub = xs_top_at(c, centre, &found);
                if (!found) bub = -1;
            }
            int n = 0;
            if (bub != -1 && xs_dict_has(c, bub)) {
                for (int i = 0; i < XS_WORLD(c, en); ++i) {
                    const int64_t at = (int64_t)i * c.W + c.w;
                    if (c.s->e_bub[at] != bub) continue;
                    const int k = c.s->e_slot[at];
                    if (!xs_in(c, k) || k == ref) continue;
                    const int64_t o = xs_row(c, k);
                    if (xs_moribund(c, k)) continue;
                    if (XB_Q(c.P, XB_I_CLOAKED, o) && !b) continue;
                    const double d2 = v3_lensq(v3_sub(xb_ld3(c.P, XB_F_POS, o), centre));
                    const double lim = f[3] + XB_F(c.P, XB_F_RADIUS, o);
                    if (d2 <= lim * lim && n < c.N) {
                        put(1 + n, xs_id(c, k));
                        if (outf) outf[(int64_t)n * W + c.w] = d2;
                        ++n;
                    }
                }
            }
            put(0, n);
            break;
        }
        case XS_Q_BUBBLE_AT: {
            bool found;
            const int64_t bub = xs_top_at(c, v3(f[0], f[1], f[2]), &found);
            put(0, found ? bub : XS_NO_TOP);
            break;
        }
        case XS_Q_REMOTE_FOLLOWERS: {                                  /* space.py:1062-1072 */
            int n = 0;
            for (int k = 0; k < c.N; ++k) {
                const int j = c.s->by_id[XS_AT(c, k)];
                if (j < 0 || !xs_in(c, j)) continue;
                const int64_t r = xs_row(c, j);
                if (!XB_B(c.P, XB_B_FREE, r)) continue;
                const int64_t m = XB_Q(c.P, XB_I_MODE, r);
                if (!(m == XB_MODE_FOLLOW || m == XB_MODE_ORBIT || m == XB_MODE_MISSILE || m == XB_MODE_FORMATION))
                    continue;
                const int t = xs_slot_of_id(c, XB_Q(c.P, XB_I_FOLLOW_ID, r));
                if (t >= 0 && XS_SLOT(c, bubble, t) == XS_SLOT(c, bubble, j)) continue;
                put(1 + n++, xs_id(c, j));
            }
            put(0, n);
            break;
        }
        default: put(0, 0); break;
    }
}
