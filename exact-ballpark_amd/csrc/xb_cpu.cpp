// SPDX-License-Identifier: MIT
//
// The CPU batched path: the same tick source as the gfx950 kernels (xb_modes_tick.h) built for the host, the worlds'
// groups run in parallel with OpenMP. It serves as the host cross-check of the kernels and as the CPU baseline the
// benchmark times beside the GPU. Built with g++ -ffp-contract=off, no fast-math.
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>

#include <vector>

#define XB_HD
#include "xb_modes_tick.h"
#include "xb_maths_float.h"
#include "xb_commands.h"
#include "xb_rand.h"
#include "../../include/xb.h"
#include "xb_park.h"
#include "xb_space.h"
#include "xb_server.h"

namespace {

struct Scratch {
    std::vector<double> d;
    std::vector<int64_t> id;
    std::vector<uint8_t> act;
    std::vector<float> fq;
    std::vector<int8_t> ffid;
    XbRecords R;
    /* the formation records, when the park has formations */
    void size_form(int m, bool on) {
        if (!on) {
            R.fq = nullptr, R.ffid = nullptr;
            return;
        }
        if ((int)ffid.size() < m) fq.assign((size_t)4 * m, 0.0f), ffid.assign(m, -1);
        R.fq = fq.data(), R.ffid = ffid.data();
    }
    void size(int m) {
        if ((int)act.size() >= m) return;
        d.assign((size_t)8 * m, 0.0);
        id.assign(m, 0);
        act.assign(m, 0);
        R.p0x = d.data();
        R.p0y = R.p0x + m;
        R.p0z = R.p0y + m;
        R.v0x = R.p0z + m;
        R.v0y = R.v0x + m;
        R.v0z = R.v0y + m;
        R.vy_after = R.v0z + m;
        R.radius = R.vy_after + m;
        R.id = id.data();
        R.act = act.data();
    }
    /* the collision records (XbCollRecs) of a group */
    std::vector<double> cd;
    std::vector<int64_t> cq;
    std::vector<int8_t> cb;
    std::vector<uint8_t> cflags, ckind;
    std::vector<int64_t> cfld;
    XbCollRecs C;
    void size_coll(int m) {
        if ((int)cflags.size() >= m) return;
        cd.assign((size_t)15 * m, 0.0);
        cq.assign((size_t)5 * m, 0);
        cb.assign((size_t)4 * m, 0);
        cflags.assign(m, 0);
        double* d0 = cd.data();
        C.px = d0, C.py = d0 + m, C.pz = d0 + 2 * m, C.vx = d0 + 3 * m, C.vy = d0 + 4 * m, C.vz = d0 + 5 * m;
        C.gx = d0 + 6 * m, C.gy = d0 + 7 * m, C.gz = d0 + 8 * m, C.ex = d0 + 9 * m, C.ey = d0 + 10 * m;
        C.ez = d0 + 11 * m, C.mass = d0 + 12 * m, C.m = d0 + 13 * m, C.tf = d0 + 14 * m;
        C.kx = cq.data(), C.ky = C.kx + m, C.kz = C.ky + m, C.owner = C.kz + m, C.fid = C.owner + m;
        C.level = cb.data(), C.mx = C.level + m, C.my = C.mx + m, C.mz = C.my + m;
        C.flags = cflags.data();
        ckind.assign(m, 0);
        C.kind = ckind.data();
        cfld.assign((size_t)3 * m, -1);
        C.fh = C.fc = C.fa = nullptr;
    }
    void size_fields(bool on) {                 // after size_coll: the force fields' keys, when the park has them
        C.fh = on ? cfld.data() : nullptr;
        C.fc = on ? cfld.data() + cflags.size() : nullptr;
        C.fa = on ? cfld.data() + 2 * cflags.size() : nullptr;
    }
};

/* A group's ticks with the gradient collision method: xb_collide_tiled_body's phases, each over the whole group
   before the next (each world's rows consecutive, at most 64). */
template <bool DO>
void run_group_coll(const XbPark& P, const xb_modes_state& s, Scratch& S, std::vector<XbRow>& rows, int64_t lo,
                    int64_t cnt) {
    S.size((int)cnt);
    S.size_coll((int)cnt);
    S.size_fields(P.fld != nullptr);
    S.size_form((int)cnt, P.n_form > 0);
    rows.resize((size_t)cnt);
    std::vector<XbCollRow> cr((size_t)cnt);
    std::vector<int> wlo((size_t)cnt), whi((size_t)cnt);
    std::vector<XoQ> rot1((size_t)cnt);
    std::vector<XoV3> w1((size_t)cnt);
    std::vector<XbV3> pe((size_t)cnt), ve((size_t)cnt);
    std::vector<char> warped((size_t)cnt);
    const bool ypr = !DO && s.ypr;
    std::vector<XbYpr> ys(ypr ? (size_t)cnt : 0);
    for (int64_t k = 0; k < cnt; ++k) {
        xb_row_load(P, s.rowmap ? (int64_t)s.rowmap[lo + k] : lo + k, rows[k]);
        if (DO) xb_turn_load(P, rows[k].row, rows[k].t);
        xb_coll_load(P, rows[k].row, cr[k]);
        if (ypr) xb_ypr_load(s.ypr, s.stride, rows[k].row, ys[k]);
    }
    std::vector<char> on((size_t)cnt);               /* OnTick's mask: a row outside it keeps every field */
    for (int64_t k = 0; k < cnt; ++k) on[k] = !s.active || s.active[rows[k].row];
    for (int64_t k = 0; k < cnt; ++k) {
        const int64_t wk = XB_Q(P, XB_I_WORLD, rows[k].row);
        int a = (int)k, b = (int)k + 1;
        while (a > 0 && XB_Q(P, XB_I_WORLD, rows[a - 1].row) == wk) --a;
        while (b < cnt && XB_Q(P, XB_I_WORLD, rows[b].row) == wk) ++b;
        wlo[k] = a, whi[k] = b;
    }
    const int ticks = s.ticks > 0 ? s.ticks : 1;
    for (int t = 0; t < ticks; ++t) {
        for (int64_t k = 0; k < cnt; ++k) {
            if (!on[k]) continue;
            xb_record_a(rows[k], S.R, (int)k);
            if (S.R.ffid) xb_record_form<DO>(rows[k], S.R, (int)k, ypr ? &ys[k] : nullptr);
        }
        for (int64_t k = 0; k < cnt; ++k)
            if (on[k]) xb_record_b(rows[k], S.R, (int)k, 0, (int)cnt, P);
        for (int64_t k = 0; k < cnt; ++k) {
            if (!on[k]) continue;
            rot1[k] = xo_q(0.0f, 0.0f, 0.0f, 1.0f);
            w1[k] = xo_v(0.0f, 0.0f, 0.0f);
            xb_phase1<DO>(rows[k], S.R, (int)k, 0, (int)cnt, P, rot1[k], w1[k]);
            if (xb_row_live(rows[k])) {
                cr[k].last_c = v3(0.0, 0.0, 0.0);
                cr[k].nh = 0;
            }
            xb_coll_record(rows[k], cr[k], S.C, (int)k, P);
        }
        for (int64_t k = 0; k < cnt; ++k) {
            if (!on[k]) continue;
            const XbShip& sh = rows[k].s;
            warped[k] = xb_row_live(rows[k]) && sh.mode == XB_MODE_WARP && sh.effect_stamp >= 0;
            pe[k] = v3(0.0, 0.0, 0.0), ve[k] = pe[k];
            if (warped[k]) xb_coll_phase3_warp(rows[k], cr[k], S.C, (int)k, P, &pe[k], &ve[k]);
        }
        for (int64_t k = 0; k < cnt; ++k) {
            if (!on[k]) continue;
            if (ypr)
                xb_coll_phase34<DO, true, true>(rows[k], cr[k], warped[k] != 0, pe[k], ve[k], S.C, S.R, (int)k,
                                                wlo[k], whi[k], P, rot1[k], w1[k], &ys[k]);
            else
                xb_coll_phase34<DO, false, true>(rows[k], cr[k], warped[k] != 0, pe[k], ve[k], S.C, S.R, (int)k,
                                                 wlo[k], whi[k], P, rot1[k], w1[k]);
            if (xb_tick_tail(rows[k])) {            /* a petrified troll: SetBallFree(false)'s snap and filing */
                if (ypr) xb_yaw_pitch_roll(ys[k], rows[k].s, false, rows[k].s.clock, true, P);
                xb_file(rows[k].s, rows[k].free_, P, cr[k].fil);
            }
        }
    }
    for (int64_t k = 0; k < cnt; ++k) {
        if (!on[k]) continue;
        xb_row_store(P, rows[k]);
        if (DO && rows[k].turned) xb_turn_store(P, rows[k].row, rows[k].t);
        xb_coll_store(P, rows[k].row, cr[k]);
        if (ypr) xb_ypr_store(s.ypr, s.stride, rows[k].row, ys[k]);
        if (s.path) s.path[rows[k].row] = rows[k].path;
        if (s.iters) s.iters[rows[k].row] = rows[k].iters;
    }
}

void run_group(const XbPark& P, const xb_modes_state& s, Scratch& S, std::vector<XbRow>& rows, int64_t lo,
               int64_t cnt) {
    S.size((int)cnt);
    S.size_form((int)cnt, P.n_form > 0);
    rows.resize((size_t)cnt);
    const bool dyn = s.dynamical != 0;
    const bool ypr = !dyn && s.ypr;
    std::vector<XbYpr> ys(ypr ? (size_t)cnt : 0);
    for (int64_t k = 0; k < cnt; ++k) {
        xb_row_load(P, s.rowmap ? (int64_t)s.rowmap[lo + k] : lo + k, rows[k]);
        if (dyn) xb_turn_load(P, rows[k].row, rows[k].t);
        if (ypr) xb_ypr_load(s.ypr, s.stride, rows[k].row, ys[k]);
    }
    std::vector<char> on((size_t)cnt);               /* OnTick's mask: a row outside it keeps every field */
    for (int64_t k = 0; k < cnt; ++k) on[k] = !s.active || s.active[rows[k].row];
    const int ticks = s.ticks > 0 ? s.ticks : 1;
    for (int t = 0; t < ticks; ++t) {
        for (int64_t k = 0; k < cnt; ++k) {
            if (!on[k]) continue;
            xb_record_a(rows[k], S.R, (int)k);
            if (S.R.ffid) {
                if (dyn) xb_record_form<true>(rows[k], S.R, (int)k, nullptr);
                else xb_record_form<false>(rows[k], S.R, (int)k, ypr ? &ys[k] : nullptr);
            }
        }
        for (int64_t k = 0; k < cnt; ++k)
            if (on[k]) xb_record_b(rows[k], S.R, (int)k, 0, (int)cnt, P);
        for (int64_t k = 0; k < cnt; ++k) {
            if (!on[k]) continue;
            if (dyn) xb_tick_row<true>(rows[k], S.R, (int)k, 0, (int)cnt, P);
            else if (ypr) xb_tick_row<false, true>(rows[k], S.R, (int)k, 0, (int)cnt, P, &ys[k]);
            else xb_tick_row<false>(rows[k], S.R, (int)k, 0, (int)cnt, P);
        }
    }
    for (int64_t k = 0; k < cnt; ++k) {
        if (!on[k]) continue;
        xb_row_store(P, rows[k]);
        if (dyn && rows[k].turned) xb_turn_store(P, rows[k].row, rows[k].t);
        if (ypr) xb_ypr_store(s.ypr, s.stride, rows[k].row, ys[k]);
        if (s.path) s.path[rows[k].row] = rows[k].path;
        if (s.iters) s.iters[rows[k].row] = rows[k].iters;
    }
}

}  // namespace

extern "C" {

int xb_cpu_abi(void) { return XB_ABI_VERSION; }

// OpenMP's default team size, read once before any call sets a size of its own (a call's `threads` applies to that
// call only).
static int default_threads() {
    static const int n = omp_get_max_threads();
    return n;
}

// One tick of the park in host memory; tiles as xb_modes_tick's (NULL: the whole park is one group). threads <= 0:
// OpenMP's default.
int xb_cpu_modes_tick(const xb_modes_state* state, int threads) {
    const int team = threads > 0 ? threads : default_threads();
    if (!state || state->n < 0 || state->stride < state->n) return -1;
    if (state->n == 0) return 0;
    if (state->dynamical && state->ypr) return -1;
    if (state->form && state->n_form > 0 &&
        (!state->form_len || state->form_stride <= 0 || !state->tiles || (!state->dynamical && !state->ypr)))
        return -1;
    const XbPark P = xb_park(*state);
    if (state->collisions) {
        if (!state->tiles || state->n_tiles <= 0 || !state->coll_f64 || !state->coll_i64 || state->hits_cap < 0 ||
            (state->hits_cap > 0 && !state->hits))
            return -1;
        const int nt = state->n_tiles;
        const bool dyn = state->dynamical != 0;
#pragma omp parallel num_threads(team)
        {
            Scratch S;
            std::vector<XbRow> rows;
#pragma omp for schedule(dynamic, 16)
            for (int g = 0; g < nt; ++g) {
                const int64_t lo = state->tiles[g], cnt = (int64_t)state->tiles[g + 1] - state->tiles[g];
                if (dyn) run_group_coll<true>(P, *state, S, rows, lo, cnt);
                else run_group_coll<false>(P, *state, S, rows, lo, cnt);
            }
        }
        return 0;
    }
    if (!state->tiles) {
        Scratch S;
        std::vector<XbRow> rows;
        run_group(P, *state, S, rows, 0, state->n);
        return 0;
    }
    const int nt = state->n_tiles;
#pragma omp parallel num_threads(team)
    {
        Scratch S;
        std::vector<XbRow> rows;
#pragma omp for schedule(dynamic, 16)
        for (int g = 0; g < nt; ++g)
            run_group(P, *state, S, rows, state->tiles[g], (int64_t)state->tiles[g + 1] - state->tiles[g]);
    }
    return 0;
}

int xb_cpu_collide_file(const xb_modes_state* state, const int32_t* rows, int64_t n_rows) {
    if (!state || !state->f64 || !state->i64 || !state->u8 || !state->coll_f64 || !state->coll_i64 || state->n < 0 ||
        state->stride < state->n || n_rows < 0 || (!rows && n_rows > state->n))
        return -1;
    const XbPark P = xb_park(*state);
    for (int64_t j = 0; j < n_rows; ++j) {
        const int64_t r = rows ? (int64_t)rows[j] : j;
        if (r < 0 || r >= state->n) continue;
        XbRow w;
        xb_row_load(P, r, w);
        XbCollRow c;
        xb_coll_load(P, r, c);
        xb_file(w.s, w.free_, P, c.fil);
        xb_coll_store(P, r, c);
    }
    return 0;
}

int xb_cpu_statics_setup(int64_t n, const int64_t* kind, const int64_t* id, const int64_t* parent,
                         const double* geom, double* st_f64, int64_t* st_i64, int64_t* st_keys, int64_t st_stride,
                         int32_t max_boxes) {
    if (n < 0 || st_stride < n || max_boxes < 0 || (n > 0 && (!kind || !id || !parent || !geom || !st_f64 ||
                                                               !st_i64 || (max_boxes > 0 && !st_keys))))
        return -1;
    int most = 0;
    std::vector<int64_t> keys(3 * (size_t)(max_boxes > 0 ? max_boxes : 1));
    for (int64_t k = 0; k < n; ++k) {
        const double* g = geom + 12 * k;
        double* f = st_f64;
        auto put = [&](int row, XbV3 v) {
            f[row * st_stride + k] = v.x, f[(row + 1) * st_stride + k] = v.y, f[(row + 2) * st_stride + k] = v.z;
        };
        for (int row = 0; row < XB_NSF; ++row) f[row * st_stride + k] = 0.0;
        XbBoxSet set;
        set.keys = keys.data(), set.cap = max_boxes, set.n = 0;
        int level = 0, cnt;
        if (kind[k] == XB_ST_CAPSULE) {                  /* Capsule::Initialize: the radius a float */
            const XbV3 a = v3(g[0], g[1], g[2]), b = v3(g[3], g[4], g[5]);
            const double radius = xb_f32(g[6]);
            const XbV3 ab = v3_sub(b, a);
            put(XB_SF_A, a), put(XB_SF_AB, ab), put(XB_SF_B, b);
            f[XB_SF_RAD * st_stride + k] = radius;
            f[XB_SF_ABAB * st_stride + k] = v3_dot(ab, ab);
            const double pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z};
            cnt = xb_capsule_boxes(pa, pb, radius, &level, set);
            if (cnt < 0) return -2;
        } else if (kind[k] == XB_ST_BOX) {               /* BoxShape::BoxShape, OrientedBox::Initialize */
            const XbV3 A = v3(g[0], g[1], g[2]), lx = v3(g[3], g[4], g[5]), ly = v3(g[6], g[7], g[8]);
            const XbV3 lz = v3(g[9], g[10], g[11]);
            XbV3 c[8];
            c[0] = A, c[1] = v3_add(A, lx), c[2] = v3_add(A, ly), c[3] = v3_add(A, lz);
            c[4] = v3_add(c[1], ly), c[5] = v3_add(c[2], lz), c[6] = v3_add(c[1], lz), c[7] = v3_add(c[4], lz);
            for (int i = 0; i < 8; ++i) put(XB_SF_CORNER + 3 * i, c[i]);
            for (int side = 0; side < 6; ++side) {
                const XbV3 sa = c[xb_side_corner(side, 0)], sb = c[xb_side_corner(side, 1)];
                const XbV3 sc = c[xb_side_corner(side, 2)], sd = c[xb_side_corner(side, 3)];
                put(XB_SF_NORMAL + 3 * side, v3_unit(v3_cross(v3_sub(sb, sa), v3_sub(sc, sa))));
                put(XB_SF_TNORMAL + 3 * side, v3_unit(v3_cross(v3_sub(sc, sa), v3_sub(sd, sa))));
            }
            const XbV3 centre = v3_add(A, v3_scale(v3_sub(c[7], A), 1.0 / 2.0));
            put(XB_SF_CENTER, centre);
            const XbV3 loc[3] = {lx, ly, lz};
            for (int i = 0; i < 3; ++i) {
                put(XB_SF_LOCAL + 3 * i, loc[i]);
                put(XB_SF_UNIT + 3 * i, v3_unit(loc[i]));
                f[(XB_SF_HALF + i) * st_stride + k] = v3_len(loc[i]) / 2.0;
                f[(XB_SF_LENSQ + i) * st_stride + k] = v3_lensq(loc[i]);
            }
            double lo[3] = {A.x, A.y, A.z}, hi[3] = {A.x, A.y, A.z};
            for (int i = 1; i < 8; ++i) {                /* AABB::Update */
                const double v[3] = {c[i].x, c[i].y, c[i].z};
                for (int q = 0; q < 3; ++q) {
                    if (v[q] > hi[q]) hi[q] = v[q];
                    else if (v[q] < lo[q]) lo[q] = v[q];
                }
            }
            put(XB_SF_LO, v3(lo[0], lo[1], lo[2])), put(XB_SF_HI, v3(hi[0], hi[1], hi[2]));
            const double brad_f = xb_f32(v3_len(v3_sub(c[7], A)) / 2.0);
            const double pos[3] = {centre.x, centre.y, centre.z};
            cnt = xb_obox_boxes(lo, hi, pos, brad_f, &level, set);
        } else {
            return -1;
        }
        st_i64[XB_SI_KIND * st_stride + k] = kind[k];
        st_i64[XB_SI_ID * st_stride + k] = id[k];
        st_i64[XB_SI_PARENT * st_stride + k] = parent[k];
        st_i64[XB_SI_LEVEL * st_stride + k] = level;
        st_i64[XB_SI_NBOX * st_stride + k] = cnt < max_boxes ? cnt : max_boxes;
        if (cnt > most) most = cnt;
        const int stored = cnt < max_boxes ? cnt : max_boxes;
        for (int b = 0; b < stored; ++b)
            for (int q = 0; q < 3; ++q) st_keys[(3 * b + q) * st_stride + k] = keys[3 * b + q];
    }
    return most;
}

int xb_cpu_modes_ypr_snap(const xb_modes_state* state, const int32_t* rows, int64_t n_rows) {
    if (!state || !state->f64 || !state->i64 || !state->u8 || !state->ypr || state->n < 0 ||
        state->stride < state->n || n_rows < 0 || (!rows && n_rows > state->n))
        return -1;
    const XbPark P = xb_park(*state);
    for (int64_t j = 0; j < n_rows; ++j) {
        const int64_t r = rows ? (int64_t)rows[j] : j;
        if (r < 0 || r >= state->n) continue;
        XbRow w;
        xb_row_load(P, r, w);
        XbYpr y;
        xb_ypr_load(state->ypr, state->stride, r, y);
        xb_yaw_pitch_roll(y, w.s, false, w.s.clock, true, P);
        xb_ypr_store(state->ypr, state->stride, r, y);
    }
    return 0;
}

static bool command_args_ok(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, const xb_command* cmds) {
    return state && state->f64 && state->i64 && state->u8 && state->n >= 0 && state->stride >= state->n &&
           n_rows >= 0 && cmds && (rows || n_rows <= state->n);
}

int xb_cpu_modes_command_check(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, xb_command* cmds,
                               int32_t* status) {
    if (!command_args_ok(state, rows, n_rows, cmds) || !status) return -1;
    const XbPark P = xb_park(*state);
    for (int64_t j = 0; j < n_rows; ++j) {
        const int64_t r = rows ? (int64_t)rows[j] : j;
        if (r >= 0 && r < state->n) *status |= xb_cmd_gather(P, state->n, r, cmds[j], state->ypr);
    }
    return 0;
}

int xb_cpu_modes_command_apply(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, xb_command* cmds,
                               uint8_t* mark) {
    if (!command_args_ok(state, rows, n_rows, cmds)) return -1;
    const XbPark P = xb_park(*state);
    const int64_t n = state->n;
    for (int64_t j = 0; j < n_rows; ++j) {                  /* gather: every read of another row first */
        const int64_t r = rows ? (int64_t)rows[j] : j;
        if (r >= 0 && r < n) xb_cmd_gather(P, n, r, cmds[j], state->ypr);
    }
    for (int64_t j = 0; j < n_rows; ++j) {                  /* apply: each row's own */
        const int64_t r = rows ? (int64_t)rows[j] : j;
        if (r >= 0 && r < n && xb_cmd_run(P, r, cmds[j], state->ypr) && mark) mark[r] = 1;
    }
    if (!mark) return 0;
    for (int64_t r = 0; r < n; ++r) xb_cmd_release(P, n, r, mark);
    for (int64_t j = 0; j < n_rows; ++j) {
        const int64_t r = rows ? (int64_t)rows[j] : j;
        if (r >= 0 && r < n) mark[r] = 0;
    }
    return 0;
}

static bool space_args_ok(const xb_modes_state* park, const xb_space_state* space) {
    return park && park->f64 && park->i64 && park->u8 && space && space->W >= 0 && space->N >= 1 && space->N <= 64 &&
           space->rows && space->by_id && space->ids;
}

int xb_cpu_space_tape(const xb_modes_state* park, const xb_space_state* space, int64_t K, const int64_t* code,
                      const int64_t* a, const int64_t* b, const double* f) {
    if (!space_args_ok(park, space) || K < 0 || (K > 0 && (!code || !a || !b || !f))) return -1;
    const XbPark P = xb_park(*park);
    const int64_t W = space->W;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t w = 0; w < W; ++w) {
        XsCtx c{space, P, w, W, space->N};
        for (int64_t k = 0; k < K; ++k) {
            const double fk[4] = {f[(0 * K + k) * W + w], f[(1 * K + k) * W + w], f[(2 * K + k) * W + w],
                                  f[(3 * K + k) * W + w]};
            xs_apply(c, code[k * W + w], a[k * W + w], b[k * W + w], fk);
        }
    }
    return 0;
}

int xb_cpu_server_layout(int64_t* out) {
    xsv_layout(out);
    return 0;
}

int xb_cpu_server_pre_tick(const xb_modes_state* park, const xb_server_state* srv, const xb_server_queue* q,
                           int phase) {
    if (!xsv_args_ok(park, srv, q, phase)) return -1;
    const XbPark P = xb_park(*park);
    const int64_t W = srv->W;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t w = 0; w < W; ++w) xsv_world(P, park->n, park->ypr, *srv, *q, srv->space, phase, w);
    return 0;
}

int xb_cpu_space_query(const xb_modes_state* park, const xb_space_state* space, int64_t kind, const int64_t* a,
                       const int64_t* b, const double* f, int64_t* out, double* outf) {
    if (!space_args_ok(park, space) || !a || !b || !f || !out) return -1;
    const XbPark P = xb_park(*park);
    const int64_t W = space->W;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t w = 0; w < W; ++w) {
        XsCtx c{space, P, w, W, space->N};
        const double fw[5] = {f[0 * W + w], f[1 * W + w], f[2 * W + w], f[3 * W + w], f[4 * W + w]};
        xs_query(c, kind, a[w], b[w], fw, out, outf);
    }
    return 0;
}

int xb_cpu_modes_commands(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, const uint8_t* op,
                          const double* point, const int64_t* target, const double* range, const double* min_range,
                          const int64_t* warp_factor) {
    if (!state || !state->f64 || !state->i64 || !state->u8 || state->n < 0 || state->stride < state->n || n_rows < 0 ||
        !op || !point || !target || !range || !min_range || !warp_factor || (!rows && n_rows > state->n))
        return -1;
    const XbPark P = xb_park(*state);
    for (int64_t j = 0; j < n_rows; ++j) {
        const int64_t r = rows ? (int64_t)rows[j] : j;
        if (r >= 0 && r < state->n) xb_cmd_apply(P, state->n, r, op, point, target, range, min_range, warp_factor);
    }
    return 0;
}

int xb_cpu_modes_orient_setup(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, int what) {
    if (!state || !state->f64 || state->n < 0 || state->stride < state->n || n_rows < 0 || what < 0 || what > 1 ||
        (!rows && n_rows > state->n))
        return -1;
    const XbPark P = xb_park(*state);
    for (int64_t j = 0; j < n_rows; ++j) {
        const int64_t r = rows ? (int64_t)rows[j] : j;
        if (r < 0 || r >= state->n) continue;
        XbTurn t;
        xb_turn_load(P, r, t);
        t.old_rot = xo_q((float)XB_F(P, XB_F_OLD_ROT, r), (float)XB_F(P, XB_F_OLD_ROT + 1, r),
                         (float)XB_F(P, XB_F_OLD_ROT + 2, r), (float)XB_F(P, XB_F_OLD_ROT + 3, r));
        t.old_w = xo_v((float)XB_F(P, XB_F_OLD_ANG_VEL, r), (float)XB_F(P, XB_F_OLD_ANG_VEL + 1, r),
                       (float)XB_F(P, XB_F_OLD_ANG_VEL + 2, r));
        t.torque = xo_v((float)XB_F(P, XB_F_TORQUE, r), (float)XB_F(P, XB_F_TORQUE + 1, r),
                        (float)XB_F(P, XB_F_TORQUE + 2, r));
        t.roll = (float)XB_F(P, XB_F_ROT_ROLL, r);
        if (what == 1) t.old_rot = xo_new_ball_old_rot(t.rot, t.w, state->fdt);
        xo_snap(xb_ld3(P, XB_F_VEL, r), t);
        xb_turn_store(P, r, t);
    }
    return 0;
}

int xb_cpu_rand_seed(const uint32_t* seeds, int64_t n_worlds, int32_t* states) {
    if (n_worlds < 0 || (n_worlds > 0 && (!seeds || !states))) return -1;
    for (int64_t w = 0; w < n_worlds; ++w) xb_srand(seeds[w], states + w * XB_RAND_WORDS);
    return 0;
}

int xb_cpu_rand_draw(int32_t* states, int64_t n_worlds, int32_t draws, int32_t* out) {
    if (n_worlds < 0 || draws < 0 || (n_worlds > 0 && (!states || (draws > 0 && !out)))) return -1;
    for (int64_t w = 0; w < n_worlds; ++w)
        for (int32_t k = 0; k < draws; ++k) out[w * (int64_t)draws + k] = xb_rand(states + w * XB_RAND_WORDS);
    return 0;
}

int xb_cpu_pack_tiles(const int32_t* sizes, int64_t n, int32_t block, int32_t* tiles, int64_t* n_tiles) {
    if (!sizes || !tiles || !n_tiles || n < 0 || block <= 0) return -1;
    int64_t k = 0, acc = 0, row = 0;
    tiles[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t sz = sizes[i];
        if (sz < 0 || sz > block) return -1;
        if (acc + sz > block) {
            tiles[++k] = (int32_t)row;
            acc = 0;
        }
        acc += sz;
        row += sz;
    }
    if (row > INT32_MAX) return -1;
    if (n > 0) tiles[++k] = (int32_t)row;
    *n_tiles = k;
    return 0;
}

int xb_cpu_maths_eval(int func, int64_t n, const double* x, double* out) {
    for (int64_t i = 0; i < n; ++i) {
        double sn, cs;
        switch (func) {
            case 0: out[i] = xb_sin(x[i]); break;
            case 1: out[i] = xb_cos(x[i]); break;
            case 2: out[i] = xb_log(x[i]); break;
            case 3: out[i] = xb_exp(x[i]); break;
            case 4: xb_sincos(x[i], &sn, &cs); out[i] = sn; break;
            case 5: xb_sincos(x[i], &sn, &cs); out[i] = cs; break;
            case 6: out[i] = __builtin_sqrt(x[i]); break;
            case 7: out[i] = x[i] / x[n + i]; break;
            case 8: out[i] = (double)xb_expf((float)x[i]); break;
            case 9: out[i] = (double)xb_logf((float)x[i]); break;
            case 10: out[i] = (double)xb_sinf((float)x[i]); break;
            case 11: out[i] = (double)xb_cosf((float)x[i]); break;
            case 12: out[i] = (double)xb_acosf((float)x[i]); break;
            case 13: out[i] = (double)xb_atanf((float)x[i]); break;
            case 14: out[i] = (double)xb_atan2f((float)x[i], (float)x[n + i]); break;
            case 15: out[i] = xb_pow(x[i], x[n + i]); break;
            case 16: out[i] = xb_asin(x[i]); break;
            case 17: out[i] = xb_atan2(x[i], x[n + i]); break;
            case 18: out[i] = (double)xb_powf((float)x[i], (float)x[n + i]); break;
            default: return -1;
        }
    }
    return 0;
}

}  // extern "C"
