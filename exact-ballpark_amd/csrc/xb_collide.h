/* SPDX-License-Identifier: MIT
 *
 * Destiny's gradient collision method and the partition filing it reads, one source for the gfx950 kernels and their
 * host build (XB_HD as in xb_modes_tick.h, which includes this after its ball and integrator types).
 *
 * The rules (the reference restates them in src/exact_ballpark/space.py and vec_collide.py):
 * - Filing (Ball::InsertInBoxes, space.py:710-817): a ball's inflated radius (a free ball's radius plus a tick at full
 *   thrust) picks the finest of 8 partition levels whose box width is at least three times it; the ball's box holds
 *   its centre, and on each axis where the inflated sphere reaches past the box's face the ball is also in the
 *   neighbour on that side: its boxes are the product of {key} or {key, key + side} per axis (clamped into the grid).
 * - Candidates (GetNearbyBalls with Gradient's filter, space.py:819-885 and 1650-1661): the other massive, uncloaked,
 *   live balls of the world that are not missiles, whose boxes are one of the ball's boxes, an ancestor or a
 *   descendant of one; less a mushroom the ball owns and space junk unless the ball is junk too.
 * - Their order: GetNearbyBalls walks the ball's boxes by key; for each, its non-empty ancestors bottom-up, its
 *   descendants in post-order (children by key), then the box itself; the balls of a box by id, each ball at its first
 *   box. That order is a sort key a pair computes from the two filings alone (xb_order_key).
 * - Potential (space.py:1663-1743) with each candidate in that order: the time of impact of the two straight
 *   predictions within the tick; an impact reflects off a fixed ball or exchanges momentum with a free one (the
 *   acceleration that reaches the post-impact path's end); an overlap pushes the two apart (recursing twice more for
 *   two free balls, with the light-ball rule and the early return), the acceleration that moves the ball out within
 *   the tick; 0.85 of it kept if not earlier than the impact held (the stronger at an equal time), and the other ball
 *   appended to the collision list.
 * Every operation is one IEEE double operation in the reference's order (no contraction), exp is glibc's.
 */
#ifndef XB_COLLIDE_H
#define XB_COLLIDE_H

#define XB_NUM_LEVELS 8                         /* Partition.cpp:20-35 (space.py:69-73) */
#define XB_BIGBOX (65536.0 * 480.0 * 0.25)      /* the top level's box width, metres */
#define XB_LIGHT_BALL_RATIO 25.0                /* Ballpark.cpp:2948 */
#define XB_RESPONSE_SCALE 0.85                  /* Ballpark.cpp:2985 */
#define XB_FIXED_MASS 1.0e34                    /* a fixed ball's mass in the response */
#define XB_CODE_BASE 65                         /* post-order digits: 64 children, and 64 for "past its level" */
#define XB_CODE_SPAN 4902227890625LL            /* 65^7: one digit per level below the ball's */
#define XB_NOT_RELATED (-1LL)

/* the collision block rows: float64 mLastC (3); int64 level, key (3), side (3), the collision list's length */
#define XB_C_LAST_C 0
#define XB_NCF 3
#define XB_C_LEVEL 0
#define XB_C_KEY 1
#define XB_C_SIDE 4
#define XB_C_HITS_N 7
#define XB_NCI 8

static inline XB_HD double xb_level_width(int i) { return XB_BIGBOX / (double)(1LL << (2 * i)); }
static inline XB_HD int64_t xb_level_grid(int i) { return 1LL << (2 * i + 23); }
static inline XB_HD int64_t xb_clamp_key(int64_t k, int64_t n) { return k < 0 ? 0 : k >= n ? n - 1 : k; }

struct XbFiling {
    int level;
    int mod[3];
    int64_t key[3];
};

/* Ball::GetInflatedRadius (space.py:710-719): a free ball's radius plus dt times the speed a tick at full thrust
   reaches from its velocity; a fixed ball's radius. */
static inline XB_HD double xb_inflated_radius(const XbShip& s, bool free_, const XbPark& P) {
    if (!free_) return s.radius;
    const double m = s.mass * s.agility;
    const double a = (P.friction * s.sf * s.max_vel) / m;
    const double v = __builtin_sqrt(v3_lensq(s.vel));
    const double k = P.friction;
    const double mps = (m * a - (m * a - v * k) * s.tf) / k;
    return s.radius + P.dt * mps;
}

/* InsertInBoxes's outcome (space.py:623-637, 756-771) */
static inline XB_HD void xb_file(const XbShip& s, bool free_, const XbPark& P, XbFiling& f) {
    const double infl = xb_inflated_radius(s, free_, P);
    int level = 0;
    for (int i = 0; i < XB_NUM_LEVELS; ++i) {
        if (3.0 * infl > xb_level_width(i)) break;
        level = i;
    }
    const double w = xb_level_width(level);
    const int64_t n = xb_level_grid(level);
    const double fac = 1.0 / w, half = 0.5 * w * (double)n, center = w * (double)n * 0.5;
    const double pos[3] = {s.pos.x, s.pos.y, s.pos.z};
    for (int q = 0; q < 3; ++q) {
        const int64_t k = xb_clamp_key(xb_to_int64((pos[q] + half) * fac), n);
        const double lo = (double)k * w - center, hi = (double)(k + 1) * w - center;
        f.key[q] = k;
        f.mod[q] = (lo + infl - pos[q] > 0.0) ? -1 : (hi - infl - pos[q] < 0.0) ? 1 : 0;
    }
    f.level = level;
}

/* A filing's boxes on one axis: its options in ascending order (1 or 2, clamped into the grid, distinct). */
static inline XB_HD int xb_axis_options(const XbFiling& f, int q, int64_t o[2]) {
    const int64_t n = xb_level_grid(f.level);
    const int64_t a = xb_clamp_key(f.key[q] + (f.mod[q] < 0 ? f.mod[q] : 0), n);
    const int64_t b = xb_clamp_key(f.key[q] + (f.mod[q] > 0 ? f.mod[q] : 0), n);
    o[0] = a;
    o[1] = b;
    return a == b ? 1 : 2;
}

/* The place of the other ball's first box in the ball's GetNearbyBalls walk, as a sort key (then the other's id), or
   XB_NOT_RELATED when none of its boxes is one of the ball's boxes, an ancestor or a descendant of one. The ball's
   boxes are indexed in key order, (ix cnt_y + iy) cnt_z + iz over the per-axis options. Its key:
   (segment x 3 + part) x 65^7 + sub, the segment the index of the ball's first box the other relates to; the part 0
   for an ancestor (sub: levels up, the walk climbs bottom-up), 2 for the same box, 1 for a descendant (sub: its
   post-order code below the ball's box: one base-65 digit a level from the ball's level + 1 down, a child's 16 dx +
   4 dy + dz, and 64 past the descendant's own level so that it sorts after its children). */
static inline XB_HD int64_t xb_order_key(const XbFiling& b, const XbFiling& n) {
    int64_t ob[3][2], on[3][2];
    int cb[3], cn[3];
    for (int q = 0; q < 3; ++q) {
        cb[q] = xb_axis_options(b, q, ob[q]);
        cn[q] = xb_axis_options(n, q, on[q]);
    }
    if (n.level <= b.level) {                          /* an ancestor, or the same level */
        const int sh = 2 * (b.level - n.level);
        int idx[3];
        for (int q = 0; q < 3; ++q) {
            idx[q] = -1;
            for (int i = cb[q] - 1; i >= 0; --i) {
                const int64_t a = ob[q][i] >> sh;
                if (a == on[q][0] || (cn[q] == 2 && a == on[q][1])) idx[q] = i;
            }
            if (idx[q] < 0) return XB_NOT_RELATED;
        }
        const int64_t seg = ((int64_t)idx[0] * cb[1] + idx[1]) * cb[2] + idx[2];
        if (n.level < b.level) return (seg * 3) * XB_CODE_SPAN + (b.level - n.level);
        return (seg * 3 + 2) * XB_CODE_SPAN;
    }
    const int sh = 2 * (n.level - b.level);            /* below: each of the other's boxes under one of the ball's */
    int64_t best = XB_NOT_RELATED;
    for (int i0 = 0; i0 < cn[0]; ++i0)
        for (int i1 = 0; i1 < cn[1]; ++i1)
            for (int i2 = 0; i2 < cn[2]; ++i2) {
                const int64_t x[3] = {on[0][i0], on[1][i1], on[2][i2]};
                int idx[3];
                bool under = true;
                for (int q = 0; q < 3; ++q) {
                    const int64_t a = x[q] >> sh;
                    idx[q] = (a == ob[q][0]) ? 0 : (cb[q] == 2 && a == ob[q][1]) ? 1 : -1;
                    under = under && idx[q] >= 0;
                }
                if (!under) continue;
                int64_t code = 0;
                for (int l = b.level + 1; l < XB_NUM_LEVELS; ++l) {
                    int64_t digit = XB_CODE_BASE - 1;
                    if (l <= n.level) {
                        const int s2 = 2 * (n.level - l);
                        digit = ((x[0] >> s2) & 3) * 16 + ((x[1] >> s2) & 3) * 4 + ((x[2] >> s2) & 3);
                    }
                    code = code * XB_CODE_BASE + digit;
                }
                const int64_t seg = ((int64_t)idx[0] * cb[1] + idx[1]) * cb[2] + idx[2];
                const int64_t key = (seg * 3 + 1) * XB_CODE_SPAN + code;
                if (best == XB_NOT_RELATED || key < best) best = key;
            }
    return best;
}

/* ------------------------------------------------------------------------------------- the pair's physics */

/* Ballpark::Integrate over a time t (space.py:1612-1628): nothing for t = 0; the cached decay factor for t = dt,
   else glibc's exp(-k/m t). */
static inline XB_HD void xb_integrate_t(XbV3* pp, XbV3* vp, XbV3 a, double m, double tf, double t, const XbPark& P) {
    const double k = P.friction;
    if (t == 0.0) return;
    if (t != P.dt) tf = xb_exp((-k / m) * t);
    const XbV3 p = *pp, v = *vp;
    XbV3 pn;
    if (k < XB_TAYLOR_LIMIT * m * t) {
        const double inv_m = 1.0 / m;
        const double kt = -k * t;
        pn = v3((m * (a.x * kt + k * (a.x * t + v.x * k * inv_m * t)) + p.x * P.k2) * P.ook2,
                (m * (a.y * kt + k * (a.y * t + v.y * k * inv_m * t)) + p.y * P.k2) * P.ook2,
                (m * (a.z * kt + k * (a.z * t + v.z * k * inv_m * t)) + p.z * P.k2) * P.ook2);
    } else {
        const double tf1 = tf - 1.0;
        pn = v3((m * (m * a.x * tf1 + k * (a.x * t + v.x - tf * v.x)) + p.x * P.k2) * P.ook2,
                (m * (m * a.y * tf1 + k * (a.y * t + v.y - tf * v.y)) + p.y * P.k2) * P.ook2,
                (m * (m * a.z * tf1 + k * (a.z * t + v.z - tf * v.z)) + p.z * P.k2) * P.ook2);
    }
    *vp = v3((m * a.x - (m * a.x - v.x * k) * tf) * P.ook, (m * a.y - (m * a.y - v.y * k) * tf) * P.ook,
             (m * a.z - (m * a.z - v.z * k) * tf) * P.ook);
    *pp = pn;
}

static inline XB_HD double v3_dot(XbV3 a, XbV3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* CollideTwoSpheres (space.py:1630-1648): 0 if already touching, else the earlier root in [0, 1], else -1. */
static inline XB_HD double xb_collide_two_spheres(XbV3 p0, XbV3 p1, XbV3 q0, XbV3 q1, double cr) {
    const XbV3 p0q0 = v3_sub(p0, q0);
    const double p0q0_2 = v3_dot(p0q0, p0q0);
    if (p0q0_2 <= cr * cr) return 0.0;
    const XbV3 dpq = v3((p1.x - p0.x) - (q1.x - q0.x), (p1.y - p0.y) - (q1.y - q0.y), (p1.z - p0.z) - (q1.z - q0.z));
    const double a = v3_dot(dpq, dpq), b = 2.0 * v3_dot(p0q0, dpq), c = p0q0_2 - cr * cr;
    if (a == 0.0 && b == 0.0) return -1.0;
    double det = b * b - 4.0 * a * c;
    if (det < 0.0) return -1.0;
    det = __builtin_sqrt(det);
    const double s2 = ((-b - det) * 0.5) / a;
    if (s2 >= 0.0 && s2 <= 1.0) return s2;
    return -1.0;
}

/* The responding ball during its gradient, and one candidate as the records hold it. */
struct XbCollMe {
    XbV3 pos, vel, g;              /* start of tick (after the behaviours), mLastG */
    double m, mass, tf, radius;    /* mass x agility, the base mass, the decay factor */
    int64_t id, mode, follow_id, owner_id;
    XbV3 last_c;
    double last_collision;
    int64_t nh;                    /* the collision list's length */
    int64_t* hits;                 /* its list: entry i at hits[i * stride] */
    int64_t stride;
    int cap;
};
struct XbCollOther {
    XbV3 pos, vel, g;
    double m, mass, tf, radius;
    int64_t id;
    bool free_;
};

static inline XB_HD void xb_coll_append(XbCollMe& me, int64_t id) {
    if (me.nh < me.cap) me.hits[me.nh * me.stride] = id;
    me.nh += 1;
}

/* Potential at recursion depth D, with the two balls at pm and po (space.py:1663-1743). */
template <int D>
static inline XB_HD void xb_potential(XbCollMe& me, const XbCollOther& o, XbV3 pm, XbV3 po, const XbPark& P) {
    const double cr = me.radius + o.radius;
    const double k = P.friction, dt = P.dt;
    XbV3 p1 = pm, pv = me.vel;
    xb_integrate(&p1, &pv, me.g, me.m, me.tf, P);
    XbV3 q1 = po, qv = o.vel;
    xb_integrate(&q1, &qv, o.g, o.m, o.tf, P);
    const double s = xb_collide_two_spheres(pm, p1, po, q1, cr);
    if (s == -1.0) return;
    XbV3 a1;
    if (s > 0.0) {                                      /* the impact within the tick */
        XbV3 pa = pm, va = me.vel;
        xb_integrate_t(&pa, &va, me.g, me.m, me.tf, s * dt, P);
        XbV3 qa = po, wa = o.vel;
        xb_integrate_t(&qa, &wa, o.g, o.m, o.tf, s * dt, P);
        const XbV3 nrm = v3_unit(v3_sub(qa, pa));
        const double v1 = v3_dot(va, nrm), v2 = v3_dot(wa, nrm);
        if (!o.free_) {                                 /* reflected off a fixed ball */
            const double c = 2.0 * v1;
            va = v3(va.x - nrm.x * c, va.y - nrm.y * c, va.z - nrm.z * c);
        } else {                                        /* exchanged with a free one (base masses) */
            const double mm1 = me.mass, mm2 = o.mass;
            const double v1p = (mm1 * v1 - mm2 * v1 + 2.0 * mm2 * v2) / (mm1 + mm2);
            const double dv = v1p - v1;
            va = v3(va.x + nrm.x * dv, va.y + nrm.y * dv, va.z + nrm.z * dv);
        }
        xb_integrate_t(&pa, &va, me.g, me.m, me.tf, (1.0 - s) * dt, P);
        const double m1 = me.m, tf = me.tf;
        const double inv_m = 1.0 / m1, inv_t = 1.0 / (tf - 1.0);
        const XbV3 g = me.g, v0 = me.vel;
        a1 = v3(-((((g.x * -m1) + g.x * (tf * m1)) - (v0.x * tf) * k) + va.x * k) * inv_m * inv_t,
                -((((g.y * -m1) + g.y * (tf * m1)) - (v0.y * tf) * k) + va.y * k) * inv_m * inv_t,
                -((((g.z * -m1) + g.z * (tf * m1)) - (v0.z * tf) * k) + va.z * k) * inv_m * inv_t);
    } else {                                            /* already overlapping */
        const XbV3 d = v3_sub(po, pm);
        const double d2 = v3_dot(d, d);
        const XbV3 nrm = (d2 == 0.0) ? v3(me.id > o.id ? 1.0 : -1.0, 0.0, 0.0) : v3_unit(d);
        const double dist = cr - __builtin_sqrt(d2) + 1;
        const double d1 = (o.m / (me.m + o.m)) * dist;
        const double dd2 = (me.m / (me.m + o.m)) * dist;
        if (D < 2 && o.free_) {                         /* two free balls: pushed apart and tried again */
            if constexpr (D < 2) {
                xb_potential<D + 1>(me, o, v3(pm.x - nrm.x * d1, pm.y - nrm.y * d1, pm.z - nrm.z * d1),
                                    v3(po.x + nrm.x * dd2, po.y + nrm.y * dd2, po.z + nrm.z * dd2), P);
            }
            if (o.m / me.m > XB_LIGHT_BALL_RATIO) {
                const XbV3 u = v3_unit(me.last_c);
                me.last_c = v3(u.x * dist, u.y * dist, u.z * dist);
            }
            const XbV3 dv = v3_sub(me.vel, o.vel);
            if (v3_lensq(dv) > 0.0001) return;
        }
        const double nc = v3_dot(me.g, nrm);
        const double tmp = 1.0 / (me.m + dt * k);
        const double c = -d1 / (dt * tmp * me.m);
        const double inv_dt = 1.0 / dt;
        a1 = v3((nrm.x * c - me.vel.x) * inv_dt - nrm.x * nc, (nrm.y * c - me.vel.y) * inv_dt - nrm.y * nc,
                (nrm.z * c - me.vel.z) * inv_dt - nrm.z * nc);
    }
    if ((me.mode != XB_MODE_MISSILE || o.id != me.follow_id) && s >= me.last_collision) {
        const XbV3 lc = v3(a1.x * XB_RESPONSE_SCALE, a1.y * XB_RESPONSE_SCALE, a1.z * XB_RESPONSE_SCALE);
        if (s == me.last_collision) {
            if (v3_lensq(lc) > v3_lensq(me.last_c)) me.last_c = lc;
        } else {
            me.last_c = lc;
            me.last_collision = s;
        }
    }
    xb_coll_append(me, o.id);
}

/* ---------------------------------------------------------------------------------- the group's records */

/* What the gradient reads of each ball of a group (a tile of whole worlds), written after the behaviours: indexed by
   the ball's place in the group. */
struct XbCollRecs {
    double *px, *py, *pz, *vx, *vy, *vz, *gx, *gy, *gz;   /* position, velocity, mLastG */
    double *ex, *ey, *ez;                                 /* the tick's straight prediction (depth 0) */
    double *mass, *m, *tf;                                /* base mass; mass x agility (1e34 if fixed); decay */
    int64_t *kx, *ky, *kz, *owner, *fid;                  /* the filing's key; the owner id; the follow id */
    int8_t *level, *mx, *my, *mz;                         /* the filing's level and sides */
    uint8_t* flags;                                       /* XB_CF_* */
    uint8_t* kind;                                        /* XB_CK_*: a force field, a miniball */
    int64_t *fh, *fc, *fa;                                /* with force fields: harmonic, corporation, alliance */
};
#define XB_CK_FIELD 1
#define XB_CK_MINIBALL 2
#define XB_CF_MASSIVE 1            /* a candidate of every ball */
#define XB_CF_DROPPED 2            /* dropped out of warp in this tick's phase 3: a candidate of the balls after it */
#define XB_CF_FREE 4
#define XB_CF_SKIP 8               /* never a candidate: moribund, cloaked, a missile */
#define XB_CF_JUNK 16
#define XB_CF_MUSHROOM 32
#define XB_CF_MISSILE 64           /* a candidate of defender missiles only */

static inline XB_HD XbFiling xb_rec_filing(const XbCollRecs& C, int j) {
    XbFiling f;
    f.level = C.level[j];
    f.key[0] = C.kx[j], f.key[1] = C.ky[j], f.key[2] = C.kz[j];
    f.mod[0] = C.mx[j], f.mod[1] = C.my[j], f.mod[2] = C.mz[j];
    return f;
}

static inline XB_HD void xb_rec_put_filing(const XbCollRecs& C, int j, const XbFiling& f) {
    C.level[j] = (int8_t)f.level;
    C.kx[j] = f.key[0], C.ky[j] = f.key[1], C.kz[j] = f.key[2];
    C.mx[j] = (int8_t)f.mod[0], C.my[j] = (int8_t)f.mod[1], C.mz[j] = (int8_t)f.mod[2];
}

static inline XB_HD XbCollOther xb_rec_other(const XbCollRecs& C, const int64_t* ids, const double* radius, int k) {
    XbCollOther o;
    o.pos = v3(C.px[k], C.py[k], C.pz[k]);
    o.vel = v3(C.vx[k], C.vy[k], C.vz[k]);
    o.g = v3(C.gx[k], C.gy[k], C.gz[k]);
    o.m = C.m[k], o.mass = C.mass[k], o.tf = C.tf[k], o.radius = radius[k];
    o.id = ids[k];
    o.free_ = (C.flags[k] & XB_CF_FREE) != 0;
    return o;
}

/* Gradient's skips for the pair (me, the ball at place k) (full.py:2633-2638): my own missile or mushroom, a missile
   at me, junk unless I am junk; for a missile, its owner (the signed owner id: a defender's launcher is not skipped)
   and its owner's miniballs. */
static inline XB_HD bool xb_coll_skipped(const XbCollRecs& C, const int64_t* ids, int j, int k, bool me_junk,
                                         bool me_missile, int64_t me_owner) {
    const uint8_t f = C.flags[k];
    if ((f & (XB_CF_MUSHROOM | XB_CF_MISSILE)) && C.owner[k] == ids[j]) return true;
    if ((f & XB_CF_MISSILE) && C.fid[k] == ids[j]) return true;
    if ((f & XB_CF_JUNK) && !me_junk) return true;
    return me_missile && (ids[k] == me_owner || ((C.kind[k] & XB_CK_MINIBALL) && C.owner[k] == me_owner));
}

/* GetNearbyBalls' force field filter (full.py:1162-1174): a ball in mode FIELD is no candidate of `me` (at place j)
   when my harmonic is -2 (the pass), or its harmonic, corporation or alliance (not -1) is mine. */
static inline XB_HD bool xb_field_passed(const XbCollRecs& C, int j, int k) {
    if (!(C.kind[k] & XB_CK_FIELD) || !C.fh) return false;
    if (C.fh[j] == -2) return true;
    if (C.fh[k] != -1 && C.fh[k] == C.fh[j]) return true;
    if (C.fc[k] != -1 && C.fc[k] == C.fc[j]) return true;
    return C.fa[k] != -1 && C.fa[k] == C.fa[j];
}

/* BallSortComparer (full.py:888-890): within a box two local (negative) ids by their owners first, otherwise by id */
static inline XB_HD bool xb_ball_before(int64_t a, int64_t a_owner, int64_t b, int64_t b_owner) {
    if (a < 0 && b < 0) return a_owner < b_owner || (a_owner == b_owner && a < b);
    return a < b;
}

/* Whether the ball at place k is one of the gradient's candidates of `me` (at place j) among GetNearbyBalls' (not a
   normal missile's): live, massive (or dropped out of warp before me in id order), uncloaked, not a missile unless
   `me` is a defender missile, and not skipped. */
static inline XB_HD bool xb_coll_candidate(const XbCollRecs& C, const int64_t* ids, int j, int k, bool me_junk,
                                           bool me_defender, int64_t me_owner) {
    if (k == j) return false;
    const uint8_t f = C.flags[k];
    if (f & XB_CF_SKIP) return false;
    if ((f & XB_CF_MISSILE) && !me_defender) return false;
    if (!(f & XB_CF_MASSIVE) && !((f & XB_CF_DROPPED) && ids[k] < ids[j])) return false;
    if (xb_field_passed(C, j, k)) return false;
    return !xb_coll_skipped(C, ids, j, k, me_junk, me_defender, me_owner);
}

/* Gradient (space.py:1650-1661) for the ball at place j of a group whose world occupies places [wlo, whi) (at most 64
   places): the candidates whose depth-0 prediction touches are found first (a pair with no contact at depth 0 changes
   nothing), then Potential runs on them in the walk's order. */
static inline XB_HD void xb_gradient(XbCollMe& me, const XbFiling& mf, bool me_junk, const XbCollRecs& C,
                                     const int64_t* ids, const double* radius, int j, int wlo, int whi,
                                     const XbPark& P) {
    const XbV3 me_end = v3(C.ex[j], C.ey[j], C.ez[j]);
    const bool defender = me.mode == XB_MODE_MISSILE;     /* a normal missile's gradient is xb_missile_gradient */
    uint64_t active = 0;
    for (int k = wlo; k < whi; ++k) {
        if (!xb_coll_candidate(C, ids, j, k, me_junk, defender, me.owner_id)) continue;
        if (xb_order_key(mf, xb_rec_filing(C, k)) == XB_NOT_RELATED) continue;
        const double cr = me.radius + radius[k];
        const double s = xb_collide_two_spheres(me.pos, me_end, v3(C.px[k], C.py[k], C.pz[k]),
                                                v3(C.ex[k], C.ey[k], C.ez[k]), cr);
        if (s != -1.0) active |= 1ULL << (k - wlo);
    }
    while (active) {                        /* the contacts in (walk key, BallSortComparer) order */
        int pick = -1;
        int64_t pk = 0;
        for (uint64_t a = active; a; a &= a - 1) {
            const int k = wlo + __builtin_ctzll(a);
            const int64_t key = xb_order_key(mf, xb_rec_filing(C, k));
            bool take = pick < 0 || key < pk;
            if (!take && key == pk) take = xb_ball_before(ids[k], C.owner[k], ids[pick], C.owner[pick]);
            if (take) {
                pick = k;
                pk = key;
            }
        }
        active &= ~(1ULL << (pick - wlo));
        const XbCollOther o = xb_rec_other(C, ids, radius, pick);
        xb_potential<0>(me, o, me.pos, o.pos, P);
    }
}

/* A normal missile's Gradient (Partition::GetCollisionCandidates, full.py:1202-1209): its one candidate is its target
   at place tp, whatever its flags and boxes (its bubble is the missile's: MissileFollow refuses another's, and no
   bubble changes in this park), less Gradient's skips. */
static inline XB_HD void xb_missile_gradient(XbCollMe& me, bool me_junk, const XbCollRecs& C, const int64_t* ids,
                                             const double* radius, int j, int tp, const XbPark& P) {
    if (tp < 0 || tp == j || xb_coll_skipped(C, ids, j, tp, me_junk, true, me.owner_id)) return;
    const XbCollOther o = xb_rec_other(C, ids, radius, tp);
    xb_potential<0>(me, o, me.pos, o.pos, P);
}

#endif
