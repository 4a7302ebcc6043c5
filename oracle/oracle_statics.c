/* SPDX-License-Identifier: MIT
 *
 * TEST INFRASTRUCTURE: part of the oracle (included by oracle_modes.c), never linked into the product.
 *
 * FullBallpark's static collidables restated in scalar C, one object at a time in the reference's structure: capsules
 * and oriented boxes (src/exact_ballpark/full.py:345-705: Collision.cpp's segment and triangle tests, BoxShape,
 * RectangleShape, Triangle, AABB), their filing in the partition (Capsule::InsertInBoxes, OrientedBox::InsertInBoxes,
 * full.py:998-1100) and their response in the gradient (Capsule::CollideWithBall, HandleCollisionNonIteratively,
 * ReactToCollision; OrientedBox's; full.py:2646-2800). Vectors are double[3]; every operation is the reference's, in
 * its order.
 */

#define EJECT_MARGIN 1.0                       /* Capsule.cpp:427, OrientedBox.cpp:106 (full.py:94) */
#define DBL_MAX_ 1.7976931348623157e308

enum { S_FRONT, S_BACK, S_LEFT, S_RIGHT, S_TOP, S_BOTTOM };
static const int CONNECTED_SIDES[8][3] = {                /* BoxShape.cpp:189 (full.py:473-476) */
    {S_BACK, S_LEFT, S_BOTTOM}, {S_BACK, S_RIGHT, S_BOTTOM}, {S_BACK, S_LEFT, S_TOP}, {S_FRONT, S_LEFT, S_BOTTOM},
    {S_BACK, S_RIGHT, S_TOP}, {S_FRONT, S_LEFT, S_TOP}, {S_FRONT, S_RIGHT, S_BOTTOM}, {S_FRONT, S_RIGHT, S_TOP}};
static const int SIDE_MAPPING[3][2] = {{S_LEFT, S_RIGHT}, {S_BOTTOM, S_TOP}, {S_BACK, S_FRONT}};   /* BoxShape.cpp:137 */
/* each side's corners a, b, c, d as corner indices A..H = 0..7 (BoxShape::BoxShape, full.py:512-519) */
static const int SIDE_CORNERS[6][4] = {{7, 5, 3, 6}, {2, 4, 1, 0}, {5, 2, 0, 3}, {4, 7, 6, 1}, {2, 5, 7, 4},
                                        {1, 6, 3, 0}};

static void sub3(const double a[3], const double b[3], double o[3]) { o[0] = a[0] - b[0], o[1] = a[1] - b[1], o[2] = a[2] - b[2]; }
static void add3(const double a[3], const double b[3], double o[3]) { o[0] = a[0] + b[0], o[1] = a[1] + b[1], o[2] = a[2] + b[2]; }
static void scale3(const double a[3], double c, double o[3]) { o[0] = a[0] * c, o[1] = a[1] * c, o[2] = a[2] * c; }
static void copy3(const double a[3], double o[3]) { o[0] = a[0], o[1] = a[1], o[2] = a[2]; }
static double sqrt_(double x) { return sqrt(x); }                       /* full.py:167-169: NaN below zero */
static double fmax_s(double a, double b) { return a < b ? b : a; }      /* std::max */

/* ------------------------------------------------------------------- Collision.cpp (full.py:347-397) */

static int quadratic(double a, double b, double c, double* v1, double* v2) {   /* full.py:347-354 */
    double det = b * b - 4.0 * a * c;
    if (det < 0.0) return 0;
    det = sqrt_(det);
    *v1 = ieee_div((-b + det) * 0.5, a);
    *v2 = ieee_div((-b - det) * 0.5, a);
    return 1;
}

static void point_on_ball(const double pos[3], double radius, const double n[3], double o[3]) {   /* :362-363 */
    double s[3];
    scale3(n, radius, s);
    sub3(pos, s, o);
}

/* IntersectSegmentPlane (full.py:366-375): found; *t written when `t_set` says so (even when not found). */
static int segment_plane(const double p0[3], const double p1[3], double plane_d, const double n[3], double point[3],
                         double* t, int* t_set) {
    double p0p1[3];
    sub3(p1, p0, p0p1);
    double ndab = dot3(n, p0p1);
    *t_set = 0;
    if (ndab == 0.0) return 0;
    *t = ieee_div(plane_d - dot3(n, p0), ndab);
    *t_set = 1;
    if (0.0 <= *t && *t <= 1.0) {
        double s[3];
        scale3(p0p1, *t, s);
        add3(p0, s, point);
        return 1;
    }
    return 0;
}

/* IntersectSegmentSphere (full.py:378-397) */
static int segment_sphere(const double p0[3], const double p1[3], const double sp[3], double sr, double point[3],
                          double* t, int* t_set) {
    double dir[3], nd[3], m[3];
    sub3(p1, p0, dir);
    normalized(dir, nd);
    sub3(p0, sp, m);
    double b = dot3(m, nd);
    double c = dot3(m, m) - sr * sr;
    *t_set = 0;
    if (c > 0.0 && b > 0.0) return 0;
    double disc = b * b - c;
    if (disc < 0.0) return 0;
    double tt = -b - sqrt_(disc);
    if (tt < 0.0) tt = 0.0;
    else tt = ieee_div(tt, length(dir));
    *t = tt;
    *t_set = 1;
    if (tt > 1.0) return 0;
    double s[3];
    scale3(dir, tt, s);
    add3(p0, s, point);
    return 1;
}

/* -------------------------------------------------------------------------- Triangle (full.py:400-447) */

static void tri_normal(const double a[3], const double b[3], const double c[3], double n[3]) {
    double ab[3], ac[3], x[3];
    sub3(b, a, ab);
    sub3(c, a, ac);
    cross(ab, ac, x);
    normalized(x, n);
}

static void tri_closest_point(const double a[3], const double b[3], const double c[3], const double p[3],
                              double o[3]) {
    double ab[3], ac[3], ap[3], bp[3], cp[3], s[3], t[3];
    sub3(b, a, ab);
    sub3(c, a, ac);
    sub3(p, a, ap);
    double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { copy3(a, o); return; }
    sub3(p, b, bp);
    double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) { copy3(b, o); return; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double v = ieee_div(d1, d1 - d3);
        scale3(ab, v, s);
        add3(a, s, o);
        return;
    }
    sub3(p, c, cp);
    double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 > 0.0 && d5 <= d6) { copy3(c, o); return; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = ieee_div(d2, d2 - d6);
        scale3(ac, w, s);
        add3(a, s, o);
        return;
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && d5 - d6 >= 0.0) {
        double w = ieee_div(d4 - d3, (d4 - d3) + (d5 - d6));
        double cb[3];
        sub3(c, b, cb);
        scale3(cb, w, s);
        add3(b, s, o);
        return;
    }
    double denom = ieee_div(1.0, va + vb + vc);
    double v = vb * denom, w = vc * denom;
    scale3(ab, v, s);
    add3(a, s, t);
    scale3(ac, w, s);
    add3(t, s, o);
}

static int tri_contains(const double a[3], const double b[3], const double c[3], const double p[3]) {
    double v0[3], v1[3], v2[3];
    sub3(c, a, v0);
    sub3(b, a, v1);
    sub3(p, a, v2);
    double dot00 = dot3(v0, v0), dot01 = dot3(v0, v1), dot02 = dot3(v0, v2);
    double dot11 = dot3(v1, v1), dot12 = dot3(v1, v2);
    double inv = ieee_div(1.0, dot00 * dot11 - dot01 * dot01);
    double u = (dot11 * dot02 - dot01 * dot12) * inv;
    double v = (dot00 * dot12 - dot01 * dot02) * inv;
    return u >= 0 && v >= 0 && u + v < 1;
}

typedef struct { double point[3], time, normal[3]; } OHit;            /* the `out` of the shape tests */

/* IntersectSphereTriangle (full.py:450-470) */
static int sphere_triangle(const double sp[3], double radius, const double vel[3], const double a[3],
                           const double b[3], const double c[3], OHit* out) {
    double n[3], p0[3], p1[3], point[3], t;
    int t_set;
    tri_normal(a, b, c, n);
    point_on_ball(sp, radius, n, p0);
    add3(p0, vel, p1);
    int found = segment_plane(p0, p1, dot3(a, n), n, point, &t, &t_set);
    if (t_set) out->time = t;
    if (!found) return 0;
    copy3(point, out->point);
    if (tri_contains(a, b, c, point)) {
        copy3(n, out->normal);
        return 1;
    }
    double q[3], back[3], impact[3], d[3];
    tri_closest_point(a, b, c, point, q);
    copy3(q, out->point);
    out->normal[0] = -vel[0], out->normal[1] = -vel[1], out->normal[2] = -vel[2];
    sub3(q, vel, back);
    found = segment_sphere(q, back, sp, radius, impact, &t, &t_set);
    if (t_set) out->time = t;
    if (!found) return 0;
    sub3(sp, impact, d);
    normalized(d, out->normal);
    return 1;
}

/* ------------------------------------------------------------- RectangleShape, BoxShape (full.py:479-620) */

static int rect_collide(const OStatic* s, int side, const double sp[3], double radius, const double vel[3],
                        OHit* out) {                                    /* full.py:494-500 */
    const double *a = s->corner[SIDE_CORNERS[side][0]], *b = s->corner[SIDE_CORNERS[side][1]];
    const double *c = s->corner[SIDE_CORNERS[side][2]], *d = s->corner[SIDE_CORNERS[side][3]];
    double pa[3], pc[3], m[3];
    sub3(a, sp, pa);
    sub3(c, sp, pc);
    cross(pc, pa, m);
    if (dot3(vel, m) <= 0.0) return sphere_triangle(sp, radius, vel, a, b, c, out);
    return sphere_triangle(sp, radius, vel, a, c, d, out);
}

static void box_center(const OStatic* s, double o[3]) {                /* full.py:523-524 */
    double d[3], h[3];
    sub3(s->corner[7], s->corner[0], d);
    scale3(d, ieee_div(1.0, 2.0), h);
    add3(s->corner[0], h, o);
}

static int box_closest_corner(const OStatic* s, const double p[3]) {   /* full.py:529-535 */
    double d[3];
    sub3(s->corner[0], p, d);
    int best = 0;
    double min_d = lensq(d);
    for (int i = 1; i < 8; ++i) {
        sub3(s->corner[i], p, d);
        double x = lensq(d);
        if (x < min_d) min_d = x, best = i;
    }
    return best;
}

static void box_closest_point(const OStatic* s, const double p[3], double q[3]) {   /* full.py:537-549 */
    double d[3], u[3], t[3];
    box_center(s, q);
    sub3(p, q, d);
    for (int i = 0; i < 3; ++i) {
        double half = length(s->local[i]) / 2.0;
        normalized(s->local[i], u);
        double dist = dot3(d, u);
        if (dist > half) dist = half;
        else if (dist < -half) dist = -half;
        scale3(u, dist, t);
        add3(q, t, q);
    }
}

/* RayIntersect (full.py:551-579): found; the entry side and distance, the exit-axis side and distance */
static int box_ray_intersect(const OStatic* s, const double origin[3], const double ray[3], int* side1, double* d1,
                             int* side2, double* d2) {
    double tmin[3], tmax[3], o[3], c[3];
    int revert[3] = {0, 0, 0};
    box_center(s, c);
    sub3(origin, c, o);
    for (int i = 0; i < 3; ++i) {
        const double* arm = s->local[i];
        double arm_len_sq = lensq(arm), arm_origin = dot3(arm, o), arm_ray = dot3(arm, ray);
        if (arm_ray < 0) {
            revert[i] = 1;
            tmin[i] = ieee_div(0.5 * arm_len_sq - arm_origin, arm_ray);
            tmax[i] = ieee_div(-0.5 * arm_len_sq - arm_origin, arm_ray);
        } else if (arm_ray > 0) {
            tmin[i] = ieee_div(-0.5 * arm_len_sq - arm_origin, arm_ray);
            tmax[i] = ieee_div(0.5 * arm_len_sq - arm_origin, arm_ray);
        } else {
            tmin[i] = -DBL_MAX_, tmax[i] = DBL_MAX_;
        }
    }
    int i_min = 0, i_max = 0;
    for (int i = 1; i < 3; ++i)
        if (tmin[i_min] < tmin[i]) i_min = i;
    for (int i = 1; i < 3; ++i)
        if (tmax[i] < tmax[i_max]) i_max = i;
    *d1 = tmin[i_min], *d2 = tmax[i_max];
    if (*d1 < *d2) {
        *side1 = SIDE_MAPPING[i_min][revert[i_min]];
        *side2 = SIDE_MAPPING[i_min][1 - revert[i_min]];
        return 1;
    }
    return 0;
}

/* CollideWithSphere (full.py:581-620) */
static int box_collide(const OStatic* s, const double sp[3], double radius, const double vel[3], OHit* out) {
    int closest = box_closest_corner(s, sp);
    int could_be_inside = 1, closest_half = -1;
    double shortest = -1.0;
    for (int i = 0; i < 3; ++i) {
        int side = CONNECTED_SIDES[closest][i];
        double point[3], d[3];
        point_on_ball(sp, radius, s->normal[side], point);
        sub3(s->corner[SIDE_CORNERS[side][0]], point, d);
        double hs = dot3(s->normal[side], d);
        if (hs >= 0.0) {
            if (hs < shortest || shortest == -1.0) shortest = hs, closest_half = side;
        } else {
            could_be_inside = 0;
            break;
        }
    }
    if (could_be_inside) {
        double cp[3], normal[3];
        box_closest_point(s, sp, cp);
        sub3(sp, cp, normal);
        copy3(normal, out->normal);
        double n_len = length(normal);
        if (n_len < radius) {
            if (n_len < 1e-10) {
                double speed = length(vel);
                if (speed < 1e-10) {
                    scale3(s->normal[closest_half], fmax_s(shortest, 1e-5), normal);
                } else {
                    int side1, side2;
                    double d1, d2;
                    if (box_ray_intersect(s, sp, vel, &side1, &d1, &side2, &d2)) {
                        if (fabs(d1) < fabs(d2)) scale3(s->normal[side1], fmax_s(fabs(d1) * speed, 1e-5), normal);
                        else scale3(s->normal[side2], fmax_s(fabs(d2) * speed, 1e-5), normal);
                    } else {
                        scale3(s->normal[closest_half], shortest, normal);
                    }
                }
            } else {
                double extrication = fabs(radius - n_len), u[3];
                normalized(normal, u);
                scale3(u, extrication, normal);
            }
            copy3(normal, out->normal);
            out->time = 0.0;
            return 1;
        }
    }
    for (int i = 0; i < 3; ++i)
        if (rect_collide(s, CONNECTED_SIDES[closest][i], sp, radius, vel, out)) return 1;
    return 0;
}

static int aabb_can_exclude(const OStatic* s, const double p0[3], const double p1[3], double r) {   /* :634-638 */
    const double *lo = s->lo, *hi = s->hi;
    return ((p0[0] - r > hi[0] && p1[0] - r > hi[0]) || (p0[0] + r < lo[0] && p1[0] + r < lo[0]) ||
            (p0[1] - r > hi[1] && p1[1] - r > hi[1]) || (p0[1] + r < lo[1] && p1[1] + r < lo[1]) ||
            (p0[2] - r > hi[2] && p1[2] - r > hi[2]) || (p0[2] + r < lo[2] && p1[2] + r < lo[2]));
}

/* ----------------------------------------------------------------------- the filing (full.py:966-1100) */

/* StaticCollidable::InsertInBox: the box once in its set */
static void static_in_box(OPark* p, OStatic* s, OBox x) {
    for (int j = 0; j < s->nbox; ++j)
        if (box_eq(&s->boxes[j], &x)) return;
    if (s->nbox >= MAX_SBOXES) {
        p->overflow = 1;
        return;
    }
    s->boxes[s->nbox++] = x;
}

static void box_at(int level, int64_t ix, int64_t iy, int64_t iz, OBox* x) {   /* GetBox's clamped key */
    int64_t n = level_grid(level);
    x->level = level;
    x->k[0] = clamp_key(ix, n), x->k[1] = clamp_key(iy, n), x->k[2] = clamp_key(iz, n);
}

static void box_for(double radius, const double pos[3], OBox* x) {     /* GetBox(radius, pos), space.py:623-643 */
    int level = level_for(radius);
    double w = level_width(level);
    int64_t n = level_grid(level);
    double fac = 1.0 / w, half = 0.5 * w * (double)n;
    box_at(level, to_int64_((pos[0] + half) * fac), to_int64_((pos[1] + half) * fac), to_int64_((pos[2] + half) * fac),
           x);
}

static void box_bounds(OBox x, double lo[3], double hi[3]) {           /* Box::Box (space.py:606-609) */
    double w = level_width(x.level);
    int64_t n = level_grid(x.level);
    double center = w * (double)n * 0.5;
    for (int q = 0; q < 3; ++q) lo[q] = (double)x.k[q] * w - center, hi[q] = (double)(x.k[q] + 1) * w - center;
}

static void insert_with_sides(OPark* p, OStatic* s, OBox x, const int mod[3]) {
    OBox y;
    int64_t ix = x.k[0], iy = x.k[1], iz = x.k[2];
    int level = x.level;
    static_in_box(p, s, x);
    if (mod[0]) box_at(level, ix + mod[0], iy, iz, &y), static_in_box(p, s, y);
    if (mod[1]) box_at(level, ix, iy + mod[1], iz, &y), static_in_box(p, s, y);
    if (mod[2]) box_at(level, ix, iy, iz + mod[2], &y), static_in_box(p, s, y);
    if (mod[0] && mod[1]) box_at(level, ix + mod[0], iy + mod[1], iz, &y), static_in_box(p, s, y);
    if (mod[0] && mod[2]) box_at(level, ix + mod[0], iy, iz + mod[2], &y), static_in_box(p, s, y);
    if (mod[1] && mod[2]) box_at(level, ix, iy + mod[1], iz + mod[2], &y), static_in_box(p, s, y);
    if (mod[0] && mod[1] && mod[2]) box_at(level, ix + mod[0], iy + mod[1], iz + mod[2], &y), static_in_box(p, s, y);
}

/* Capsule::InsertInBoxes at its first insert (full.py:998-1052): none without volume; else the boxes at its radius's
   level along the segment, sampled every 2r from A while the float length lasts, each with the adjacent boxes its
   radius reaches. A walk that never ends (Destiny's result undefined) sets the park's overflow flag. */
static void capsule_file(OPark* p, OStatic* c) {
    if (c->radius <= 0.0) return;
    double interp[3], dir[3], u[3];
    copy3(c->a, interp);
    sub3(c->b, c->a, dir);
    double len = F(length(dir));
    normalized(dir, u);
    for (int q = 0; q < 3; ++q) dir[q] = u[q] * c->radius * 2.0;
    double step = F(c->radius * 2.0);
    while (len >= 0.0) {
        OBox x;
        double lo[3], hi[3];
        int mod[3] = {0, 0, 0};
        box_for(c->radius, interp, &x);
        box_bounds(x, lo, hi);
        for (int q = 0; q < 3; ++q) {
            if (lo[q] + c->radius - interp[q] > 0.0) mod[q] = -1;
            else if (hi[q] - c->radius - interp[q] < 0.0) mod[q] = 1;
        }
        insert_with_sides(p, c, x, mod);
        add3(interp, dir, interp);
        double shorter = F(len - step);
        if (shorter == len) {
            p->overflow = 1;
            return;
        }
        len = shorter;
    }
}

/* OrientedBox::InsertInBoxes at its first insert (full.py:1054-1100) */
static void obox_file(OPark* p, OStatic* s) {
    double lw = fmax_s(s->hi[0] - s->lo[0], fmax_s(s->hi[1] - s->lo[1], s->hi[2] - s->lo[2]));   /* :640-642 */
    double r = lw / 2.0;
    OBox x;
    double lo[3], hi[3];
    int mod[3] = {0, 0, 0};
    box_for(r, s->pos, &x);
    box_bounds(x, lo, hi);
    for (int q = 0; q < 3; ++q) {
        if (lo[q] - s->lo[q] > 0.0) mod[q] = -1;
        else if (hi[q] - s->hi[q] < 0.0) mod[q] = 1;
    }
    if (x.level == 0 && F(s->bounding_radius_f * 2.0) > level_width(0)) {
        static_in_box(p, s, x);
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    if (dx == 0 && dy == 0 && dz == 0) continue;
                    OBox y;
                    box_at(0, x.k[0] + dx, x.k[1] + dy, x.k[2] + dz, &y);
                    static_in_box(p, s, y);
                }
        return;
    }
    insert_with_sides(p, s, x, mod);
}

/* ------------------------------------------------------------------- the response (full.py:2646-2800) */

static void static_response(OPark* p, OBall* b, int64_t s_id, const double last_c[3], double toi) {   /* :2735-2746 */
    if (toi == b->last_collision) {
        if (lensq(last_c) > lensq(b->last_c)) copy3(last_c, b->last_c);
    } else {
        copy3(last_c, b->last_c);
        b->last_collision = toi;
    }
    add_hit(p, b, s_id);
}

/* ReactToCollision's s > 0 branch (full.py:2748-2758) */
static void static_reflect(const OPark* p, const OBall* b, const double pos0[3], const double vel0[3], double m1,
                           const double normal[3], double toi, double acc[3]) {
    double k = p->friction, tf = b->tf, dt = p->dt, pos[3], vel[3], pn[3], vn[3];
    const double* g = b->last_g;
    integrate_t(p, pos0, vel0, g, m1, tf, toi * dt, pos, vel);
    double v1 = dot3(vel, normal), s[3];
    scale3(normal, 2.0 * v1, s);
    sub3(vel, s, vel);
    integrate_t(p, pos, vel, g, m1, tf, (1.0 - toi) * dt, pn, vn);
    double inv_m = ieee_div(1.0, m1), inv_t = ieee_div(1.0, tf - 1.0);
    for (int i = 0; i < 3; ++i)
        acc[i] = -((((g[i] * -m1) + g[i] * (tf * m1)) - (b->vel[i] * tf) * k) + vn[i] * k) * inv_m * inv_t;
}

/* ReactToCollision's s = 0 branch (full.py:2760-2767) */
static void static_eject(const OPark* p, const OBall* b, double m1, const double normal[3], double normal_comp,
                         double dist, double acc[3]) {
    double dt = p->dt;
    double tmp = ieee_div(1.0, m1 + dt * p->friction);
    double c = ieee_div(dist, dt * tmp * m1);
    double inv_dt = ieee_div(1.0, dt);
    for (int i = 0; i < 3; ++i) acc[i] = (normal[i] * c - b->vel[i]) * inv_dt - normal[i] * normal_comp;
}

static void capsule_react(OPark* p, const OStatic* c, OBall* b, const double pos[3], const double vel[3], double m1,
                          const double normal0[3], double toi) {        /* full.py:2769-2786 */
    if (toi == -1.0) return;
    double center_distance = length(normal0), normal[3], acc[3], lc[3];
    normalized(normal0, normal);
    double normal_comp = dot3(b->last_g, normal);
    if (toi > 0.0) {
        if (normal_comp > 0.0) return;
        static_reflect(p, b, pos, vel, m1, normal, toi, acc);
    } else {
        double dist = c->radius + b->radius - center_distance + EJECT_MARGIN;
        static_eject(p, b, m1, normal, normal_comp, dist, acc);
    }
    scale3(acc, RESPONSE_SCALE, lc);
    static_response(p, b, c->id, lc, toi);
}

static void obox_react(OPark* p, const OStatic* s, OBall* b, const double pos[3], const double vel[3], double m1,
                       const double normal0[3], double toi) {           /* full.py:2788-2800 */
    if (toi == -1.0) return;
    double depth = length(normal0), normal[3], acc[3], lc[3];
    normalized(normal0, normal);
    if (toi > 0.0) static_reflect(p, b, pos, vel, m1, normal, toi, acc);
    else static_eject(p, b, m1, normal, dot3(b->last_g, normal), depth + EJECT_MARGIN, acc);
    scale3(acc, RESPONSE_SCALE, lc);
    static_response(p, b, s->id, lc, toi);
}

/* Capsule::HandleCollisionNonIteratively (full.py:2669-2733) */
static void capsule_non_iterative(OPark* p, const OStatic* c, OBall* b, const double p0[3], const double p1[3],
                                  double m1, const double vp1[3]) {
    double rad = b->radius + c->radius;
    double d[3], ao[3], q[3], r[3], t3[3], normal[3] = {0.0, 0.0, 0.0};
    copy3(vp1, d);
    sub3(p0, c->a, ao);
    double abab = dot3(c->ab, c->ab);
    double m = ieee_div(dot3(c->ab, d), abab);
    double n = ieee_div(dot3(c->ab, ao), abab);
    scale3(c->ab, m, t3);
    sub3(d, t3, q);
    scale3(c->ab, n, t3);
    sub3(ao, t3, r);
    double a = dot3(q, q);
    double bq = 2.0 * dot3(q, r);
    double cq = dot3(r, r) - rad * rad;
    double rad_sq = rad * rad;
    double s = -1.0, toi = -1.0;
    if (a == 0.0) {
        double sa = collide_two_spheres(p0, p1, c->a, c->a, rad);
        if (sa != -1.0 && (toi == -1.0 || sa < toi)) toi = sa, sub3(p0, c->a, normal);
        double sb = collide_two_spheres(p0, p1, c->b, c->b, rad);
        if (sb != -1.0 && (toi == -1.0 || sb < toi)) toi = sb, sub3(p0, c->b, normal);
        capsule_react(p, c, b, p1, vp1, m1, normal, toi);
        return;
    }
    double s1, s2;
    if (!quadratic(a, bq, cq, &s1, &s2)) return;
    if (0.0 <= s1 && s1 <= 1.0) s = s1;
    if (0.0 <= s2 && s2 <= 1.0 && s2 < s1) {
        s = s2;
    } else if (s == -1.0) {
        double pa[3];
        sub3(p0, c->a, pa);
        double t = ieee_div(dot3(pa, c->ab), lensq(c->ab));
        if (0.0 <= t && t <= 1.0) {
            double proj[3], dist[3];
            scale3(c->ab, t, t3);
            add3(c->a, t3, proj);
            sub3(p0, proj, dist);
            if (lensq(dist) < rad_sq) {
                capsule_react(p, c, b, p0, b->vel, m1, dist, 0.0);
                return;
            }
        }
    }
    double tk[2] = {(s1 * m) + n, (s2 * m) + n};
    const double* far[2] = {p0, p1};
    for (int i = 0; i < 2; ++i) {
        if (tk[i] < 0.0) {
            double sa = collide_two_spheres(p0, p1, c->a, c->a, rad);
            if (sa != -1.0 && (toi == -1.0 || sa < toi)) toi = sa, sub3(p0, c->a, normal);
        } else if (tk[i] > 1.0) {
            double sb = collide_two_spheres(p0, p1, c->b, c->b, rad);
            if (sb != -1.0 && (toi == -1.0 || sb < toi)) toi = sb, sub3(far[i], c->b, normal);
        } else {
            if (s != -1.0 && (toi == -1.0 || s < toi)) {
                toi = s;
                double proj[3];
                scale3(c->ab, ieee_div(dot3(ao, c->ab), lensq(c->ab)), proj);
                sub3(ao, proj, normal);
                if (lensq(normal) < rad_sq) toi = 0.0;
            }
        }
    }
    capsule_react(p, c, b, p0, b->vel, m1, normal, toi);
}

/* Capsule::CollideWithBall and OrientedBox::CollideWithBall (full.py:2646-2667): the free ball's straight path over
   the tick, predicted with mLastG */
static void collide_with_ball(OPark* p, const OStatic* s, OBall* b) {
    if (!b->is_free) return;
    double m1 = b->mass * b->agility, p0[3], p1[3], vp1[3];
    copy3(b->pos, p0);
    integrate(p, b->pos, b->vel, b->last_g, m1, b->tf, p1, vp1);
    if (s->kind == 0) {
        capsule_non_iterative(p, s, b, p0, p1, m1, vp1);
        return;
    }
    if (aabb_can_exclude(s, p0, p1, b->radius)) return;
    OHit out;
    memset(&out, 0, sizeof(out));
    double v[3];
    sub3(p1, p0, v);
    if (box_collide(s, p0, b->radius, v, &out)) obox_react(p, s, b, p0, b->vel, m1, out.normal, out.time);
}
