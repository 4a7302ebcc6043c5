/* SPDX-License-Identifier: MIT
 *
 * Destiny's static collidables in the gradient method: capsules and oriented boxes, one source for the gfx950 kernels
 * and their host build (included by xb_modes_tick.h after xb_collide.h).
 *
 * The rules (the reference restates them in src/exact_ballpark/full.py:345-705, 998-1100 and 2628-2800):
 * - A static collidable is fixed: AddCapsule takes the segment A-B and a float radius, AddOrientedBox a corner and
 *   three edge vectors (BoxShape: corners A..H, six rectangular sides, their normals; its AABB).
 * - Filing (InsertInBoxes, once): a capsule without volume nowhere; else at its radius's level, the boxes along the
 *   segment sampled every 2r from A while the float length lasts, each with the adjacent boxes its radius reaches. A
 *   box at the level for half its AABB's longest width, with the adjacent boxes its AABB reaches (all 26 around it at
 *   the top level when its bounding diameter exceeds a top box).
 * - Candidates: after a ball's Potential with every ball candidate, Gradient runs each static collidable of the ball's
 *   GetNearbyBalls walk, at its first box, by (parent id, id) within a box: a static collidable's place in the walk is
 *   the least xb_order_key over its boxes (each box taken as a filing of one box).
 * - CollideWithBall (free balls only): the ball's straight path over the tick, predicted with mLastG. A capsule's
 *   HandleCollisionNonIteratively (the swept-sphere quadratic; the caps by CollideTwoSpheres; the end state passed on
 *   when the end velocity is parallel to the axis; inside the cylinder an ejection); a box's AABB exclusion, then
 *   BoxShape::CollideWithSphere (inside: the closest point, a ray out, or the nearest face; else the three faces at the
 *   nearest corner, each as two triangles). ReactToCollision: a reflection at the time of impact (a capsule ignores one
 *   whose mLastG points out of it), or at time 0 an ejection past the surface plus 1 m; 0.85 of the acceleration that
 *   does it, the stronger at an equal time, else the latest (no earlier-time test), and the id appended to the list.
 * A static collidable's response depends on the ball's state at the tick's start and its mLastG only, so the responses
 * are found first (each in any order) and applied in the walk's order.
 * Every operation is one IEEE double operation in the reference's order (no contraction).
 */
#ifndef XB_STATICS_H
#define XB_STATICS_H

#define XB_EJECT_MARGIN 1.0                     /* Capsule.cpp:427, OrientedBox.cpp:106 */
#define XB_DBL_MAX 1.7976931348623157e308
#define XB_MAX_WORLD_STATICS 64                 /* one 64-bit response mask a ball */
#define XB_ST_CAPSULE 0
#define XB_ST_BOX 1

/* the static table's float64 rows (a union by kind). Capsule: A, B - A, B, the radius, |B - A|^2. Box: the AABB's
   low and high corners, the corners A..H, the sides' normals (triangle (a, b, c)) and their triangles' (a, c, d)
   normals, the centre, the edge vectors, each normalised, half their lengths and their squared lengths. */
#define XB_SF_A 0
#define XB_SF_AB 3
#define XB_SF_B 6
#define XB_SF_RAD 9
#define XB_SF_ABAB 10
#define XB_SF_LO 0
#define XB_SF_HI 3
#define XB_SF_CORNER 6
#define XB_SF_NORMAL 30
#define XB_SF_TNORMAL 48
#define XB_SF_CENTER 66
#define XB_SF_LOCAL 69
#define XB_SF_UNIT 78
#define XB_SF_HALF 87
#define XB_SF_LENSQ 90
#define XB_NSF 93
/* the int64 rows: kind, id, parent id, the filing's level, how many boxes */
#define XB_SI_KIND 0
#define XB_SI_ID 1
#define XB_SI_PARENT 2
#define XB_SI_LEVEL 3
#define XB_SI_NBOX 4
#define XB_NSI 5

enum { XB_S_FRONT, XB_S_BACK, XB_S_LEFT, XB_S_RIGHT, XB_S_TOP, XB_S_BOTTOM };

/* BoxShape's tables (BoxShape.cpp:137, 189): the three sides at each corner, each axis's sides (back, front), each
   side's corners a, b, c, d as corner indices A..H = 0..7. */
static inline XB_HD int xb_connected_side(int corner, int i) {
    const int t[8][3] = {{XB_S_BACK, XB_S_LEFT, XB_S_BOTTOM}, {XB_S_BACK, XB_S_RIGHT, XB_S_BOTTOM},
                         {XB_S_BACK, XB_S_LEFT, XB_S_TOP},    {XB_S_FRONT, XB_S_LEFT, XB_S_BOTTOM},
                         {XB_S_BACK, XB_S_RIGHT, XB_S_TOP},   {XB_S_FRONT, XB_S_LEFT, XB_S_TOP},
                         {XB_S_FRONT, XB_S_RIGHT, XB_S_BOTTOM}, {XB_S_FRONT, XB_S_RIGHT, XB_S_TOP}};
    return t[corner][i];
}
static inline XB_HD int xb_side_of_axis(int axis, int far_) {
    const int t[3][2] = {{XB_S_LEFT, XB_S_RIGHT}, {XB_S_BOTTOM, XB_S_TOP}, {XB_S_BACK, XB_S_FRONT}};
    return t[axis][far_];
}
static inline XB_HD int xb_side_corner(int side, int i) {
    const int t[6][4] = {{7, 5, 3, 6}, {2, 4, 1, 0}, {5, 2, 0, 3}, {4, 7, 6, 1}, {2, 5, 7, 4}, {1, 6, 3, 0}};
    return t[side][i];
}

/* the static table as a tick reads it */
struct XbStatics {
    const double* f;               /* XB_NSF rows */
    const int64_t* q;              /* XB_NSI rows */
    const int64_t* keys;           /* box b's key on axis a at row 3 b + a */
    int64_t stride;
};
#define XB_SFV(S, row, k) v3((S).f[(row) * (S).stride + (k)], (S).f[((row) + 1) * (S).stride + (k)], \
                             (S).f[((row) + 2) * (S).stride + (k)])
#define XB_SFD(S, row, k) ((S).f[(row) * (S).stride + (k)])
#define XB_SQ(S, row, k) ((S).q[(row) * (S).stride + (k)])

static inline XB_HD double xb_f32(double x) { return (double)(float)x; }
static inline XB_HD double xb_dmax(double a, double b) { return a < b ? b : a; }   /* std::max */

/* ------------------------------------------------------------------ Collision.cpp's tests (full.py:347-397) */

static inline XB_HD XbV3 xb_point_on_ball(XbV3 pos, double radius, XbV3 n) { return v3_sub(pos, v3_scale(n, radius)); }

/* IntersectSegmentPlane: found; *t written when *t_set (also when not found) */
static inline XB_HD bool xb_segment_plane(XbV3 p0, XbV3 p1, double plane_d, XbV3 n, XbV3* point, double* t,
                                          bool* t_set) {
    const XbV3 p0p1 = v3_sub(p1, p0);
    const double ndab = v3_dot(n, p0p1);
    *t_set = false;
    if (ndab == 0.0) return false;
    *t = (plane_d - v3_dot(n, p0)) / ndab;
    *t_set = true;
    if (0.0 <= *t && *t <= 1.0) {
        *point = v3_add(p0, v3_scale(p0p1, *t));
        return true;
    }
    return false;
}

/* IntersectSegmentSphere */
static inline XB_HD bool xb_segment_sphere(XbV3 p0, XbV3 p1, XbV3 sp, double sr, XbV3* point, double* t,
                                           bool* t_set) {
    const XbV3 dir = v3_sub(p1, p0);
    const XbV3 nd = v3_unit(dir);
    const XbV3 m = v3_sub(p0, sp);
    const double b = v3_dot(m, nd);
    const double c = v3_dot(m, m) - sr * sr;
    *t_set = false;
    if (c > 0.0 && b > 0.0) return false;
    const double disc = b * b - c;
    if (disc < 0.0) return false;
    double tt = -b - __builtin_sqrt(disc);
    if (tt < 0.0) tt = 0.0;
    else tt = tt / v3_len(dir);
    *t = tt;
    *t_set = true;
    if (tt > 1.0) return false;
    *point = v3_add(p0, v3_scale(dir, tt));
    return true;
}

/* Triangle::ClosestPoint (the Voronoi regions of a, b, c, the edges, the face) */
static inline XB_HD XbV3 xb_tri_closest(XbV3 a, XbV3 b, XbV3 c, XbV3 p) {
    const XbV3 ab = v3_sub(b, a), ac = v3_sub(c, a), ap = v3_sub(p, a);
    const double d1 = v3_dot(ab, ap), d2 = v3_dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;
    const XbV3 bp = v3_sub(p, b);
    const double d3 = v3_dot(ab, bp), d4 = v3_dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return b;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return v3_add(a, v3_scale(ab, d1 / (d1 - d3)));
    const XbV3 cp = v3_sub(p, c);
    const double d5 = v3_dot(ab, cp), d6 = v3_dot(ac, cp);
    if (d6 > 0.0 && d5 <= d6) return c;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return v3_add(a, v3_scale(ac, d2 / (d2 - d6)));
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && d5 - d6 >= 0.0)
        return v3_add(b, v3_scale(v3_sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    return v3_add(v3_add(a, v3_scale(ab, v)), v3_scale(ac, w));
}

/* Triangle::Contains (barycentric, the edge b-c open) */
static inline XB_HD bool xb_tri_contains(XbV3 a, XbV3 b, XbV3 c, XbV3 p) {
    const XbV3 v0 = v3_sub(c, a), v1 = v3_sub(b, a), v2 = v3_sub(p, a);
    const double dot00 = v3_dot(v0, v0), dot01 = v3_dot(v0, v1), dot02 = v3_dot(v0, v2);
    const double dot11 = v3_dot(v1, v1), dot12 = v3_dot(v1, v2);
    const double inv = 1.0 / (dot00 * dot11 - dot01 * dot01);
    const double u = (dot11 * dot02 - dot01 * dot12) * inv;
    const double v = (dot00 * dot12 - dot01 * dot02) * inv;
    return u >= 0 && v >= 0 && u + v < 1;
}

struct XbShapeHit {                /* the shape tests' outputs as the reference writes them */
    double time;
    XbV3 normal;
};

/* IntersectSphereTriangle, the triangle's normal n given */
static inline XB_HD bool xb_sphere_triangle(XbV3 sp, double radius, XbV3 vel, XbV3 a, XbV3 b, XbV3 c, XbV3 n,
                                            XbShapeHit* out) {
    const XbV3 p0 = xb_point_on_ball(sp, radius, n);
    const XbV3 p1 = v3_add(p0, vel);
    XbV3 point;
    double t;
    bool t_set;
    bool found = xb_segment_plane(p0, p1, v3_dot(a, n), n, &point, &t, &t_set);
    if (t_set) out->time = t;
    if (!found) return false;
    if (xb_tri_contains(a, b, c, point)) {
        out->normal = n;
        return true;
    }
    const XbV3 q = xb_tri_closest(a, b, c, point);
    out->normal = v3(-vel.x, -vel.y, -vel.z);
    XbV3 impact;
    found = xb_segment_sphere(q, v3_sub(q, vel), sp, radius, &impact, &t, &t_set);
    if (t_set) out->time = t;
    if (!found) return false;
    out->normal = v3_unit(v3_sub(sp, impact));
    return true;
}

/* ------------------------------------------------------------------------------------ BoxShape (full.py:479-620) */

static inline XB_HD XbV3 xb_box_corner(const XbStatics& S, int64_t k, int i) { return XB_SFV(S, XB_SF_CORNER + 3 * i, k); }
static inline XB_HD XbV3 xb_box_normal(const XbStatics& S, int64_t k, int side) {
    return XB_SFV(S, XB_SF_NORMAL + 3 * side, k);
}

/* RectangleShape::CollideWithSphere: the triangle (a, b, c) or (a, c, d) on the velocity's side of the diagonal */
static inline XB_HD bool xb_rect_collide(const XbStatics& S, int64_t k, int side, XbV3 sp, double radius, XbV3 vel,
                                         XbShapeHit* out) {
    const XbV3 a = xb_box_corner(S, k, xb_side_corner(side, 0)), c = xb_box_corner(S, k, xb_side_corner(side, 2));
    const XbV3 m = v3_cross(v3_sub(c, sp), v3_sub(a, sp));
    if (v3_dot(vel, m) <= 0.0)
        return xb_sphere_triangle(sp, radius, vel, a, xb_box_corner(S, k, xb_side_corner(side, 1)), c,
                                  xb_box_normal(S, k, side), out);
    return xb_sphere_triangle(sp, radius, vel, a, c, xb_box_corner(S, k, xb_side_corner(side, 3)),
                              XB_SFV(S, XB_SF_TNORMAL + 3 * side, k), out);
}

/* BoxShape::RayIntersect (the slabs): found; the entry side and distance, the exit distance and the side opposite the
   entry on the entry's axis */
static inline XB_HD bool xb_box_ray(const XbStatics& S, int64_t k, XbV3 origin, XbV3 ray, int* side1, double* d1,
                                    int* side2, double* d2) {
    const XbV3 o = v3_sub(origin, XB_SFV(S, XB_SF_CENTER, k));
    double tmin[3], tmax[3];
    int revert[3];
    for (int i = 0; i < 3; ++i) {
        const XbV3 arm = XB_SFV(S, XB_SF_LOCAL + 3 * i, k);
        const double arm_len_sq = XB_SFD(S, XB_SF_LENSQ + i, k), arm_origin = v3_dot(arm, o);
        const double arm_ray = v3_dot(arm, ray);
        revert[i] = 0;
        if (arm_ray < 0) {
            revert[i] = 1;
            tmin[i] = (0.5 * arm_len_sq - arm_origin) / arm_ray;
            tmax[i] = (-0.5 * arm_len_sq - arm_origin) / arm_ray;
        } else if (arm_ray > 0) {
            tmin[i] = (-0.5 * arm_len_sq - arm_origin) / arm_ray;
            tmax[i] = (0.5 * arm_len_sq - arm_origin) / arm_ray;
        } else {
            tmin[i] = -XB_DBL_MAX, tmax[i] = XB_DBL_MAX;
        }
    }
    int i_min = 0, i_max = 0;
    for (int i = 1; i < 3; ++i)
        if (tmin[i_min] < tmin[i]) i_min = i;
    for (int i = 1; i < 3; ++i)
        if (tmax[i] < tmax[i_max]) i_max = i;
    *d1 = tmin[i_min], *d2 = tmax[i_max];
    if (*d1 < *d2) {
        *side1 = xb_side_of_axis(i_min, revert[i_min]);
        *side2 = xb_side_of_axis(i_min, 1 - revert[i_min]);
        return true;
    }
    return false;
}

/* BoxShape::CollideWithSphere */
static inline XB_HD bool xb_box_collide(const XbStatics& S, int64_t k, XbV3 sp, double radius, XbV3 vel,
                                        XbShapeHit* out) {
    int closest = 0;
    double min_d = v3_lensq(v3_sub(xb_box_corner(S, k, 0), sp));
    for (int i = 1; i < 8; ++i) {                            /* ClosestCorner: the first nearest */
        const double d = v3_lensq(v3_sub(xb_box_corner(S, k, i), sp));
        if (d < min_d) min_d = d, closest = i;
    }
    bool could_be_inside = true;
    double shortest = -1.0;
    int closest_half = 0;
    for (int i = 0; i < 3; ++i) {
        const int side = xb_connected_side(closest, i);
        const XbV3 n = xb_box_normal(S, k, side);
        const XbV3 point = xb_point_on_ball(sp, radius, n);
        const double hs = v3_dot(n, v3_sub(xb_box_corner(S, k, xb_side_corner(side, 0)), point));
        if (hs >= 0.0) {
            if (hs < shortest || shortest == -1.0) shortest = hs, closest_half = side;
        } else {
            could_be_inside = false;
            break;
        }
    }
    if (could_be_inside) {
        XbV3 q = XB_SFV(S, XB_SF_CENTER, k);                 /* ClosestPoint */
        const XbV3 d = v3_sub(sp, q);
        for (int i = 0; i < 3; ++i) {
            const double half = XB_SFD(S, XB_SF_HALF + i, k);
            const XbV3 u = XB_SFV(S, XB_SF_UNIT + 3 * i, k);
            double dist = v3_dot(d, u);
            if (dist > half) dist = half;
            else if (dist < -half) dist = -half;
            q = v3_add(q, v3_scale(u, dist));
        }
        XbV3 normal = v3_sub(sp, q);
        out->normal = normal;
        const double n_len = v3_len(normal);
        if (n_len < radius) {
            if (n_len < 1e-10) {
                const double speed = v3_len(vel);
                if (speed < 1e-10) {
                    normal = v3_scale(xb_box_normal(S, k, closest_half), xb_dmax(shortest, 1e-5));
                } else {
                    int side1, side2;
                    double d1, d2;
                    if (xb_box_ray(S, k, sp, vel, &side1, &d1, &side2, &d2)) {
                        if (__builtin_fabs(d1) < __builtin_fabs(d2))
                            normal = v3_scale(xb_box_normal(S, k, side1), xb_dmax(__builtin_fabs(d1) * speed, 1e-5));
                        else
                            normal = v3_scale(xb_box_normal(S, k, side2), xb_dmax(__builtin_fabs(d2) * speed, 1e-5));
                    } else {
                        normal = v3_scale(xb_box_normal(S, k, closest_half), shortest);
                    }
                }
            } else {
                normal = v3_scale(v3_unit(normal), __builtin_fabs(radius - n_len));
            }
            out->normal = normal;
            out->time = 0.0;
            return true;
        }
    }
    for (int i = 0; i < 3; ++i)
        if (xb_rect_collide(S, k, xb_connected_side(closest, i), sp, radius, vel, out)) return true;
    return false;
}

/* ------------------------------------------------------------------------ the response (full.py:2646-2800) */

struct XbStaticResp {
    double toi;
    XbV3 lc;                       /* 0.85 of the acceleration */
};

/* ReactToCollision: the reflection at toi > 0 (from `pos`, `vel`), else the ejection by `dist` along the normal */
static inline XB_HD XbV3 xb_static_reflect(const XbCollMe& me, XbV3 pos, XbV3 vel, XbV3 normal, double toi,
                                           const XbPark& P) {
    const double k = P.friction, tf = me.tf, dt = P.dt, m1 = me.m;
    const XbV3 g = me.g;
    xb_integrate_t(&pos, &vel, g, m1, tf, toi * dt, P);
    const double v1 = v3_dot(vel, normal);
    vel = v3_sub(vel, v3_scale(normal, 2.0 * v1));
    xb_integrate_t(&pos, &vel, g, m1, tf, (1.0 - toi) * dt, P);
    const double inv_m = 1.0 / m1, inv_t = 1.0 / (tf - 1.0);
    const XbV3 v0 = me.vel;
    return v3(-((((g.x * -m1) + g.x * (tf * m1)) - (v0.x * tf) * k) + vel.x * k) * inv_m * inv_t,
              -((((g.y * -m1) + g.y * (tf * m1)) - (v0.y * tf) * k) + vel.y * k) * inv_m * inv_t,
              -((((g.z * -m1) + g.z * (tf * m1)) - (v0.z * tf) * k) + vel.z * k) * inv_m * inv_t);
}
static inline XB_HD XbV3 xb_static_eject(const XbCollMe& me, XbV3 normal, double normal_comp, double dist,
                                         const XbPark& P) {
    const double dt = P.dt, m1 = me.m;
    const double tmp = 1.0 / (m1 + dt * P.friction);
    const double c = dist / (dt * tmp * m1);
    const double inv_dt = 1.0 / dt;
    return v3((normal.x * c - me.vel.x) * inv_dt - normal.x * normal_comp,
              (normal.y * c - me.vel.y) * inv_dt - normal.y * normal_comp,
              (normal.z * c - me.vel.z) * inv_dt - normal.z * normal_comp);
}

/* Capsule::ReactToCollision: none at toi -1 or when mLastG points out at an impact */
static inline XB_HD bool xb_capsule_react(const XbCollMe& me, double cap_radius, XbV3 pos, XbV3 vel, XbV3 normal0,
                                          double toi, const XbPark& P, XbStaticResp* r) {
    if (toi == -1.0) return false;
    const double center_distance = v3_len(normal0);
    const XbV3 normal = v3_unit(normal0);
    const double normal_comp = v3_dot(me.g, normal);
    XbV3 acc;
    if (toi > 0.0) {
        if (normal_comp > 0.0) return false;
        acc = xb_static_reflect(me, pos, vel, normal, toi, P);
    } else {
        const double dist = cap_radius + me.radius - center_distance + XB_EJECT_MARGIN;
        acc = xb_static_eject(me, normal, normal_comp, dist, P);
    }
    r->toi = toi;
    r->lc = v3(acc.x * XB_RESPONSE_SCALE, acc.y * XB_RESPONSE_SCALE, acc.z * XB_RESPONSE_SCALE);
    return true;
}

/* Capsule::HandleCollisionNonIteratively: the response, if any */
static inline XB_HD bool xb_capsule_response(const XbStatics& S, int64_t k, const XbCollMe& me, XbV3 p0, XbV3 p1,
                                             XbV3 vp1, const XbPark& P, XbStaticResp* r) {
    const double cap_radius = XB_SFD(S, XB_SF_RAD, k);
    const double rad = me.radius + cap_radius;
    const XbV3 A = XB_SFV(S, XB_SF_A, k), ab = XB_SFV(S, XB_SF_AB, k);
    const XbV3 d = vp1;
    const XbV3 ao = v3_sub(p0, A);
    const double abab = XB_SFD(S, XB_SF_ABAB, k);
    const double m = v3_dot(ab, d) / abab;
    const double n = v3_dot(ab, ao) / abab;
    const XbV3 q = v3_sub(d, v3_scale(ab, m));
    const XbV3 rr = v3_sub(ao, v3_scale(ab, n));
    const double a = v3_dot(q, q);
    const double bq = 2.0 * v3_dot(q, rr);
    const double cq = v3_dot(rr, rr) - rad * rad;
    const double rad_sq = rad * rad;
    double s = -1.0, toi = -1.0;
    XbV3 normal = v3(0.0, 0.0, 0.0);
    const XbV3 B = XB_SFV(S, XB_SF_B, k);
    if (a == 0.0) {                                          /* parallel to the axis: the caps alone */
        const double sa = xb_collide_two_spheres(p0, p1, A, A, rad);
        if (sa != -1.0 && (toi == -1.0 || sa < toi)) toi = sa, normal = v3_sub(p0, A);
        const double sb = xb_collide_two_spheres(p0, p1, B, B, rad);
        if (sb != -1.0 && (toi == -1.0 || sb < toi)) toi = sb, normal = v3_sub(p0, B);
        return xb_capsule_react(me, cap_radius, p1, vp1, normal, toi, P, r);
    }
    double det = bq * bq - 4.0 * a * cq;                     /* Quadratic */
    if (det < 0.0) return false;
    det = __builtin_sqrt(det);
    const double s1 = ((-bq + det) * 0.5) / a, s2 = ((-bq - det) * 0.5) / a;
    if (0.0 <= s1 && s1 <= 1.0) s = s1;
    if (0.0 <= s2 && s2 <= 1.0 && s2 < s1) {
        s = s2;
    } else if (s == -1.0) {
        const double t = v3_dot(v3_sub(p0, A), ab) / v3_lensq(ab);
        if (0.0 <= t && t <= 1.0) {
            const XbV3 dist = v3_sub(p0, v3_add(A, v3_scale(ab, t)));
            if (v3_lensq(dist) < rad_sq) return xb_capsule_react(me, cap_radius, p0, me.vel, dist, 0.0, P, r);
        }
    }
    for (int i = 0; i < 2; ++i) {
        const double tk = ((i == 0 ? s1 : s2) * m) + n;
        if (tk < 0.0) {
            const double sa = xb_collide_two_spheres(p0, p1, A, A, rad);
            if (sa != -1.0 && (toi == -1.0 || sa < toi)) toi = sa, normal = v3_sub(p0, A);
        } else if (tk > 1.0) {
            const double sb = xb_collide_two_spheres(p0, p1, B, B, rad);
            if (sb != -1.0 && (toi == -1.0 || sb < toi)) toi = sb, normal = v3_sub(i == 0 ? p0 : p1, B);
        } else if (s != -1.0 && (toi == -1.0 || s < toi)) {
            toi = s;
            normal = v3_sub(ao, v3_scale(ab, v3_dot(ao, ab) / v3_lensq(ab)));
            if (v3_lensq(normal) < rad_sq) toi = 0.0;
        }
    }
    return xb_capsule_react(me, cap_radius, p0, me.vel, normal, toi, P, r);
}

/* OrientedBox::CollideWithBall after the AABB test, and its ReactToCollision (the normal's length the depth) */
static inline XB_HD bool xb_box_response(const XbStatics& S, int64_t k, const XbCollMe& me, XbV3 p0, XbV3 p1,
                                         const XbPark& P, XbStaticResp* r) {
    const XbV3 lo = XB_SFV(S, XB_SF_LO, k), hi = XB_SFV(S, XB_SF_HI, k);
    const double rad = me.radius;
    if ((p0.x - rad > hi.x && p1.x - rad > hi.x) || (p0.x + rad < lo.x && p1.x + rad < lo.x) ||
        (p0.y - rad > hi.y && p1.y - rad > hi.y) || (p0.y + rad < lo.y && p1.y + rad < lo.y) ||
        (p0.z - rad > hi.z && p1.z - rad > hi.z) || (p0.z + rad < lo.z && p1.z + rad < lo.z))
        return false;                                        /* AABB::CanExclude */
    XbShapeHit h;
    h.time = 0.0;
    h.normal = v3(0.0, 0.0, 0.0);
    if (!xb_box_collide(S, k, p0, rad, v3_sub(p1, p0), &h)) return false;
    const double toi = h.time;
    if (toi == -1.0) return false;
    const double depth = v3_len(h.normal);
    const XbV3 normal = v3_unit(h.normal);
    XbV3 acc;
    if (toi > 0.0) acc = xb_static_reflect(me, p0, me.vel, normal, toi, P);
    else acc = xb_static_eject(me, normal, v3_dot(me.g, normal), depth + XB_EJECT_MARGIN, P);
    r->toi = toi;
    r->lc = v3(acc.x * XB_RESPONSE_SCALE, acc.y * XB_RESPONSE_SCALE, acc.z * XB_RESPONSE_SCALE);
    return true;
}

static inline XB_HD bool xb_static_response(const XbStatics& S, int64_t k, const XbCollMe& me, XbV3 p0, XbV3 p1,
                                            XbV3 vp1, const XbPark& P, XbStaticResp* r) {
    if (XB_SQ(S, XB_SI_KIND, k) == XB_ST_CAPSULE) return xb_capsule_response(S, k, me, p0, p1, vp1, P, r);
    return xb_box_response(S, k, me, p0, p1, P, r);
}

/* The static collidable's place in the ball's walk: the least xb_order_key over its boxes, or XB_NOT_RELATED. */
static inline XB_HD int64_t xb_static_key(const XbStatics& S, int64_t k, const XbFiling& mf) {
    const int nbox = (int)XB_SQ(S, XB_SI_NBOX, k);
    XbFiling f;
    f.level = (int)XB_SQ(S, XB_SI_LEVEL, k);
    f.mod[0] = f.mod[1] = f.mod[2] = 0;
    int64_t best = XB_NOT_RELATED;
    for (int b = 0; b < nbox; ++b) {
        f.key[0] = S.keys[(3 * b) * S.stride + k];
        f.key[1] = S.keys[(3 * b + 1) * S.stride + k];
        f.key[2] = S.keys[(3 * b + 2) * S.stride + k];
        const int64_t key = xb_order_key(mf, f);
        if (key != XB_NOT_RELATED && (best == XB_NOT_RELATED || key < best)) best = key;
    }
    return best;
}

/* Whether static collidable k comes before static collidable l in the ball's walk: by walk key, then parent id, then
   id. */
static inline XB_HD bool xb_static_before(const XbStatics& S, int64_t k, int64_t l, const XbFiling& mf) {
    const int64_t kk = xb_static_key(S, k, mf), kl = xb_static_key(S, l, mf);
    if (kk != kl) return kk < kl;
    const int64_t pk = XB_SQ(S, XB_SI_PARENT, k), pl = XB_SQ(S, XB_SI_PARENT, l);
    if (pk != pl) return pk < pl;
    return XB_SQ(S, XB_SI_ID, k) < XB_SQ(S, XB_SI_ID, l);
}

/* Gradient's static collidables for a free ball (`me` after its Potential with the balls), its world's statics
   [first, first + count): the responses found, then applied in (walk key, parent id, id) order. */
static inline XB_HD void xb_statics_gradient(XbCollMe& me, const XbFiling& mf, const XbStatics& S, int64_t first,
                                             int count, const XbPark& P) {
    XbV3 p1 = me.pos, vp1 = me.vel;
    xb_integrate(&p1, &vp1, me.g, me.m, me.tf, P);
    const XbV3 p0 = me.pos;
    uint64_t resp = 0;
    for (int i = 0; i < count; ++i) {
        XbStaticResp r;
        if (xb_static_response(S, first + i, me, p0, p1, vp1, P, &r) && xb_static_key(S, first + i, mf) != XB_NOT_RELATED)
            resp |= 1ULL << i;
    }
    while (resp) {
        /* Only the index is carried through the search; the tie-breaks are read again from the table. (Carrying the
           key, parent and id beside the index in one conditional update was compiled for gfx950 so that the index
           kept an earlier candidate while the id moved on.) */
        int pick = __builtin_ctzll(resp);
        for (uint64_t a = resp & (resp - 1); a; a &= a - 1) {
            const int i = __builtin_ctzll(a);
            if (xb_static_before(S, first + i, first + pick, mf)) pick = i;
        }
        const int64_t pid = XB_SQ(S, XB_SI_ID, first + pick);
        resp &= ~(1ULL << pick);
        XbStaticResp r;
        xb_static_response(S, first + pick, me, p0, p1, vp1, P, &r);
        if (r.toi == me.last_collision) {                    /* the stronger at an equal time, else the latest */
            if (v3_lensq(r.lc) > v3_lensq(me.last_c)) me.last_c = r.lc;
        } else {
            me.last_c = r.lc;
            me.last_collision = r.toi;
        }
        xb_coll_append(me, pid);
    }
}

/* ------------------------------------------------------------------ set-up at load (full.py:650-705, 998-1100) */

/* the box at a level for a position (GetBox(radius, pos)'s key, clamped) and its bounds */
static inline XB_HD void xb_static_box(int level, const double pos[3], int64_t key[3]) {
    const double w = xb_level_width(level);
    const int64_t n = xb_level_grid(level);
    const double fac = 1.0 / w, half = 0.5 * w * (double)n;
    for (int q = 0; q < 3; ++q) key[q] = xb_clamp_key(xb_to_int64((pos[q] + half) * fac), n);
}
static inline XB_HD void xb_box_bounds(int level, const int64_t key[3], double lo[3], double hi[3]) {
    const double w = xb_level_width(level);
    const double center = w * (double)xb_level_grid(level) * 0.5;
    for (int q = 0; q < 3; ++q) lo[q] = (double)key[q] * w - center, hi[q] = (double)(key[q] + 1) * w - center;
}
static inline XB_HD int xb_level_for_radius(double radius) {
    int level = 0;
    for (int i = 0; i < XB_NUM_LEVELS; ++i) {
        if (3.0 * radius > xb_level_width(i)) break;
        level = i;
    }
    return level;
}

/* A static collidable's boxes as a set (the order of insertion kept; duplicates dropped), at most `cap` of them
   stored; returns how many there are, or -1 when the capsule's walk never ends (Capsule.cpp:499, 578). */
struct XbBoxSet {
    int64_t* keys;                 /* 3 a box */
    int n, cap;
};
static inline XB_HD void xb_boxset_add(XbBoxSet& s, int level, int64_t x, int64_t y, int64_t z) {
    const int64_t g = xb_level_grid(level);
    x = xb_clamp_key(x, g), y = xb_clamp_key(y, g), z = xb_clamp_key(z, g);
    const int stored = s.n < s.cap ? s.n : s.cap;
    for (int i = 0; i < stored; ++i)
        if (s.keys[3 * i] == x && s.keys[3 * i + 1] == y && s.keys[3 * i + 2] == z) return;
    if (s.n < s.cap) s.keys[3 * s.n] = x, s.keys[3 * s.n + 1] = y, s.keys[3 * s.n + 2] = z;
    s.n += 1;
}
static inline XB_HD void xb_boxset_with_sides(XbBoxSet& s, int level, const int64_t k[3], const int mod[3]) {
    xb_boxset_add(s, level, k[0], k[1], k[2]);
    if (mod[0]) xb_boxset_add(s, level, k[0] + mod[0], k[1], k[2]);
    if (mod[1]) xb_boxset_add(s, level, k[0], k[1] + mod[1], k[2]);
    if (mod[2]) xb_boxset_add(s, level, k[0], k[1], k[2] + mod[2]);
    if (mod[0] && mod[1]) xb_boxset_add(s, level, k[0] + mod[0], k[1] + mod[1], k[2]);
    if (mod[0] && mod[2]) xb_boxset_add(s, level, k[0] + mod[0], k[1], k[2] + mod[2]);
    if (mod[1] && mod[2]) xb_boxset_add(s, level, k[0], k[1] + mod[1], k[2] + mod[2]);
    if (mod[0] && mod[1] && mod[2]) xb_boxset_add(s, level, k[0] + mod[0], k[1] + mod[1], k[2] + mod[2]);
}

/* Capsule::InsertInBoxes */
static inline XB_HD int xb_capsule_boxes(const double a[3], const double b[3], double radius, int* level,
                                         XbBoxSet& s) {
    *level = xb_level_for_radius(radius);
    s.n = 0;
    if (radius <= 0.0) return 0;
    XbV3 interp = v3(a[0], a[1], a[2]);
    const XbV3 ab = v3(b[0] - a[0], b[1] - a[1], b[2] - a[2]);
    double len = xb_f32(v3_len(ab));
    const XbV3 u = v3_unit(ab);
    const XbV3 dir = v3(u.x * radius * 2.0, u.y * radius * 2.0, u.z * radius * 2.0);
    const double step = xb_f32(radius * 2.0);
    while (len >= 0.0) {
        const double p[3] = {interp.x, interp.y, interp.z};
        int64_t k[3];
        double lo[3], hi[3];
        xb_static_box(*level, p, k);
        xb_box_bounds(*level, k, lo, hi);
        int mod[3] = {0, 0, 0};
        for (int q = 0; q < 3; ++q) {
            if (lo[q] + radius - p[q] > 0.0) mod[q] = -1;
            else if (hi[q] - radius - p[q] < 0.0) mod[q] = 1;
        }
        xb_boxset_with_sides(s, *level, k, mod);
        interp = v3_add(interp, dir);
        const double shorter = xb_f32(len - step);
        if (shorter == len) return -1;
        len = shorter;
    }
    return s.n;
}

/* OrientedBox::InsertInBoxes, from its AABB, centre and float bounding radius */
static inline XB_HD int xb_obox_boxes(const double lo_[3], const double hi_[3], const double pos[3], double brad_f,
                                      int* level, XbBoxSet& s) {
    const double lw = xb_dmax(hi_[0] - lo_[0], xb_dmax(hi_[1] - lo_[1], hi_[2] - lo_[2]));
    *level = xb_level_for_radius(lw / 2.0);
    s.n = 0;
    int64_t k[3];
    double lo[3], hi[3];
    xb_static_box(*level, pos, k);
    xb_box_bounds(*level, k, lo, hi);
    int mod[3] = {0, 0, 0};
    for (int q = 0; q < 3; ++q) {
        if (lo[q] - lo_[q] > 0.0) mod[q] = -1;
        else if (hi[q] - hi_[q] < 0.0) mod[q] = 1;
    }
    if (*level == 0 && xb_f32(brad_f * 2.0) > xb_level_width(0)) {
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) xb_boxset_add(s, 0, k[0] + dx, k[1] + dy, k[2] + dz);
        return s.n;
    }
    xb_boxset_with_sides(s, *level, k, mod);
    return s.n;
}

#endif
