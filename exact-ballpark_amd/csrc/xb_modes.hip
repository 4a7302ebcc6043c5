// SPDX-License-Identifier: MIT
//
// The flight-modes tick on gfx950: one lane per ball row, one workgroup per tile of whole worlds.
//
// Tiled launch (the park's worlds laid out contiguously): a workgroup reads its rows' tick-start records into LDS
// (step A), derives each row's vertical velocity after its own turn (step B), then runs every row's tick (step C)
// with targets read from LDS. One pass over HBM: each state field is read once and written once, structure of
// arrays, so a wavefront's loads and stores of one field are whole cache lines. The general launch (any row order)
// runs the same three steps as three launches through caller-owned scratch.
//
// Exactness: compiled with -ffp-contract=off, no fast-math, denormals kept; the only fused operations are the maths
// header's explicit ones. The tick logic is xb_modes_tick.h, the same source the host build (xb_cpu.cpp) compiles.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <mutex>

#define XB_HD __device__
#define XB_TABLE __constant__
#include "xb_modes_tick.h"
#include "xb_maths_float.h"
#include "xb_commands.h"
#include "../../include/xb.h"
#include "xb_park.h"

namespace {

constexpr int XB_LDS_BYTES_PER_ROW = 9 * 8 + 4 + 1;     // the records, the lane schedule (int32), act
constexpr int XB_FORM_LDS_BYTES_PER_ROW = 4 * 4 + 1;   // with formations: a leader's rotation (4 floats), its formation

// The formation records after a kernel's other LDS (`used` bytes a lane), when the park has formations.
__device__ __forceinline__ void xb_form_lds(XbRecords& R, const XbPark& P, int nb, int used) {
    R.fq = nullptr;
    R.ffid = nullptr;
    if (P.n_form <= 0) return;
    extern __shared__ double xb_lds[];
    char* base = reinterpret_cast<char*>(xb_lds) + (size_t)used * nb;
    R.fq = reinterpret_cast<float*>(base);
    R.ffid = reinterpret_cast<int8_t*>(R.fq + 4 * nb);
}

// Waves a SIMD each tiled kernel is built for. A row's whole state lives in registers across a launch's ticks, and
// with dynamical orientation or collisions the compiler would take every register a lane has: one wave a SIMD, so a
// CU's one workgroup idles at each barrier while its slowest wave finishes. Held to more waves, it keeps the hottest
// state in registers and moves the rest to scratch. Measured on 65,536 worlds (ms a tick, 1 / 2 / 3 / 4 waves): the
// flight park 0.033 / 0.033 / 0.029 / 0.033, with orientation 0.66 / 0.43 / 0.41 / 0.36, with collisions 0.30 / 0.20 /
// 0.29 / 0.41, with both 1.21 / 0.76 / 0.88 / 1.02 (8 ticks a launch). The ragged flight park's single tick, lanes by
// branch key: 72 us at two waves, 67 at three.
#ifndef XB_WAVES_FLIGHT
#define XB_WAVES_FLIGHT 3
#endif
#ifndef XB_WAVES_TURNING
#define XB_WAVES_TURNING 4
#endif
#ifndef XB_WAVES_COLLIDE
#define XB_WAVES_COLLIDE 2
#endif
constexpr int XB_MAX_WAVES = 4;                          // 256 lanes
constexpr int XB_KEY_BUCKETS = 16;

// A row's orientation searches' steps in its last tick as one of 16 buckets, two a power of two (0; 1; 2, 3; 4-5, 6-7;
// ...; from 128 all in the last): lanes of similar search lengths share a wavefront, whose search loop runs as long as
// its longest lane's.
__device__ __forceinline__ int xb_steps_bucket(int v) {
    if (v <= 0) return 0;
    const int m = 31 - __clz(v);
    const int half = m >= 1 ? (v >> (m - 1)) & 1 : 0;
    const int b = 1 + 2 * m + half;
    return b < XB_KEY_BUCKETS - 1 ? b : XB_KEY_BUCKETS - 1;
}

// The place (row of the tile) lane j runs: j itself in row order; otherwise the lanes sorted by their rows' keys, a
// stable counting sort by ballot (a lane's rank among equal keys: its wave's lower lanes, then the lower waves), so
// the assignment is deterministic. Places past the tile's row count take the last bucket.
__device__ __forceinline__ int xb_schedule_place(const xb_modes_state& s, const XbPark& P, int j, int cnt, int lo,
                                                 int* perm) {
    __shared__ int cnt_w[XB_MAX_WAVES][XB_KEY_BUCKETS];
    __shared__ int off_w[XB_MAX_WAVES][XB_KEY_BUCKETS];
    int key = XB_KEY_BUCKETS - 1;
    if (j < cnt) {
        const int64_t r = s.rowmap ? (int64_t)s.rowmap[lo + j] : (int64_t)(lo + j);
        if (s.schedule == 2) {
            key = s.keys[r];
        } else if (s.schedule == 3) {
            key = xb_steps_bucket(s.iters[r]);
        } else {
            const bool live = XB_B(P, XB_B_FREE, r) && !XB_B(P, XB_B_MORIBUND, r);
            key = xb_branch_key(XB_Q(P, XB_I_MODE, r), XB_Q(P, XB_I_EFFECT_STAMP, r), live, P.new_orbit);
        }
        key = key < XB_KEY_BUCKETS - 1 ? key : XB_KEY_BUCKETS - 2;
    }
    const int wave = j >> 6, lane = j & 63;
    const unsigned long long below = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    int rank = 0;
    for (int b = 0; b < XB_KEY_BUCKETS; ++b) {
        const unsigned long long m = __ballot(key == b);
        if (lane == 0) cnt_w[wave][b] = __popcll(m);
        if (key == b) rank = __popcll(m & below);
    }
    __syncthreads();
    if (j == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        int total = 0;
        for (int b = 0; b < XB_KEY_BUCKETS; ++b)
            for (int w = 0; w < nw; ++w) {
                off_w[w][b] = total;
                total += cnt_w[w][b];
            }
    }
    __syncthreads();
    perm[off_w[wave][key] + rank] = j;
    __syncthreads();
    return perm[j];
}

template <bool DO, bool YPR>
__device__ __forceinline__ void xb_modes_tiled_body(const xb_modes_state& s) {
    extern __shared__ double xb_lds[];
    const int nb = blockDim.x;
    const int g = blockIdx.x;
    const int lo = s.tiles[g];
    const int cnt = s.tiles[g + 1] - lo;
    XbRecords R;
    R.p0x = xb_lds;
    R.p0y = R.p0x + nb;
    R.p0z = R.p0y + nb;
    R.v0x = R.p0z + nb;
    R.v0y = R.v0x + nb;
    R.v0z = R.v0y + nb;
    R.vy_after = R.v0z + nb;
    R.radius = R.vy_after + nb;
    R.id = reinterpret_cast<int64_t*>(R.radius + nb);
    int* perm = reinterpret_cast<int*>(R.id + nb);
    R.act = reinterpret_cast<uint8_t*>(perm + nb);
    const XbPark P = xb_park(s);
    xb_form_lds(R, P, nb, XB_LDS_BYTES_PER_ROW);
    const int j = s.schedule ? xb_schedule_place(s, P, (int)threadIdx.x, cnt, lo, perm) : (int)threadIdx.x;
    const int64_t r = j < cnt ? (s.rowmap ? (int64_t)s.rowmap[lo + j] : (int64_t)(lo + j)) : 0;
    const bool on = j < cnt && (!s.active || s.active[r]);     /* a row outside OnTick's mask: nothing */
    const int ticks = s.ticks > 0 ? s.ticks : 1;
    XbRow w;
    XbYpr y;
    if (on) {
        xb_row_load(P, r, w);
        if (DO) xb_turn_load(P, r, w.t);
        if (YPR) xb_ypr_load(s.ypr, s.stride, r, y);
    }
    for (int t = 0; t < ticks; ++t) {
        if (on) {
            xb_record_a(w, R, j);
            if (R.ffid) xb_record_form<DO>(w, R, j, YPR ? &y : nullptr);
        }
        __syncthreads();
        if (on) xb_record_b(w, R, j, 0, cnt, P);
        __syncthreads();
        if (on) xb_tick_row<DO, YPR>(w, R, j, 0, cnt, P, &y);
        __syncthreads();
    }
    if (on) {
        xb_row_store(P, w);
        if (DO && w.turned) xb_turn_store(P, r, w.t);
        if (YPR) xb_ypr_store(s.ypr, s.stride, r, y);
        if (s.path) s.path[r] = w.path;
        if (s.iters) s.iters[r] = w.iters;
    }
}

// The collision records of a tile in LDS, after the tick's own records (XB_LDS_BYTES_PER_ROW a lane).
constexpr int XB_COLL_LDS_BYTES_PER_ROW = 15 * 8 + 5 * 8 + 4 + 1 + 1;
constexpr int XB_FIELD_LDS_BYTES_PER_ROW = 3 * 8;      // with force fields: harmonic, corporation, alliance

__device__ __forceinline__ XbCollRecs xb_coll_lds(double* base, int nb) {
    XbCollRecs C;
    double* d = base;
    C.px = d, C.py = d + nb, C.pz = d + 2 * nb, C.vx = d + 3 * nb, C.vy = d + 4 * nb, C.vz = d + 5 * nb;
    C.gx = d + 6 * nb, C.gy = d + 7 * nb, C.gz = d + 8 * nb, C.ex = d + 9 * nb, C.ey = d + 10 * nb;
    C.ez = d + 11 * nb, C.mass = d + 12 * nb, C.m = d + 13 * nb, C.tf = d + 14 * nb;
    int64_t* q = reinterpret_cast<int64_t*>(d + 15 * nb);
    C.kx = q, C.ky = q + nb, C.kz = q + 2 * nb, C.owner = q + 3 * nb, C.fid = q + 4 * nb;
    int8_t* b = reinterpret_cast<int8_t*>(q + 5 * nb);
    C.level = b, C.mx = b + nb, C.my = b + 2 * nb, C.mz = b + 3 * nb;
    C.flags = reinterpret_cast<uint8_t*>(b + 4 * nb);
    C.kind = C.flags + nb;
    C.fh = C.fc = C.fa = nullptr;
    return C;
}

// The force fields' keys after a kernel's other LDS (`used` bytes a lane, a multiple of 8 bytes for 64 lanes), when
// the park has them.
__device__ __forceinline__ void xb_field_lds(XbCollRecs& C, const XbPark& P, int nb, int used) {
    if (!P.fld) return;
    extern __shared__ double xb_lds[];
    int64_t* q = reinterpret_cast<int64_t*>(reinterpret_cast<char*>(xb_lds) + (size_t)used * nb);
    C.fh = q, C.fc = q + nb, C.fa = q + 2 * nb;
}

// The tiled tick with the gradient collision method: the phases of Ballpark::Evolve with a barrier between them, as
// the reference's order needs (modes.py:702-745): (A, B) the tick-start records; (1) the behaviours, then each row's
// view for the gradients; (3a) the rows in warp proper, whose drop-out makes them massive for the balls after them;
// (3b) each massive row's gradient over its world's rows, its end state; (4) the commit and the refiling.
template <bool DO, bool YPR, bool ST>
__device__ __forceinline__ void xb_collide_tiled_body(const xb_modes_state& s) {
    extern __shared__ double xb_lds[];
    const int nb = blockDim.x;
    const int g = blockIdx.x;
    const int lo = s.tiles[g];
    const int cnt = s.tiles[g + 1] - lo;
    XbRecords R;
    R.p0x = xb_lds;
    R.p0y = R.p0x + nb;
    R.p0z = R.p0y + nb;
    R.v0x = R.p0z + nb;
    R.v0y = R.v0x + nb;
    R.v0z = R.v0y + nb;
    R.vy_after = R.v0z + nb;
    R.radius = R.vy_after + nb;
    R.id = reinterpret_cast<int64_t*>(R.radius + nb);
    int* perm = reinterpret_cast<int*>(R.id + nb);
    R.act = reinterpret_cast<uint8_t*>(perm + nb);
    XbCollRecs C = xb_coll_lds(xb_lds + (XB_LDS_BYTES_PER_ROW * nb) / 8, nb);
    const XbPark P = xb_park(s);
    xb_form_lds(R, P, nb, XB_LDS_BYTES_PER_ROW + XB_COLL_LDS_BYTES_PER_ROW);
    xb_field_lds(C, P, nb, XB_LDS_BYTES_PER_ROW + XB_COLL_LDS_BYTES_PER_ROW +
                               (P.n_form > 0 ? XB_FORM_LDS_BYTES_PER_ROW : 0));
    const int j = s.schedule ? xb_schedule_place(s, P, (int)threadIdx.x, cnt, lo, perm) : (int)threadIdx.x;
    const int64_t r = j < cnt ? (s.rowmap ? (int64_t)s.rowmap[lo + j] : (int64_t)(lo + j)) : 0;
    const bool on = j < cnt && (!s.active || s.active[r]);     /* a row outside OnTick's mask: nothing */
    const int ticks = s.ticks > 0 ? s.ticks : 1;
    XbRow w;
    XbCollRow c;
    XbYpr y;
    if (on) {
        xb_row_load(P, r, w);
        if (DO) xb_turn_load(P, r, w.t);
        xb_coll_load(P, r, c);
        if (YPR) xb_ypr_load(s.ypr, s.stride, r, y);
    }
    C.kx[j] = on ? XB_Q(P, XB_I_WORLD, r) : -1;     // the world's places: a run of equal world ids
    __syncthreads();
    int wlo = j, whi = j + 1;
    if (on) {
        while (wlo > 0 && C.kx[wlo - 1] == C.kx[j]) --wlo;
        while (whi < cnt && C.kx[whi] == C.kx[j]) ++whi;
    }
    __syncthreads();
    for (int t = 0; t < ticks; ++t) {
        if (on) {
            xb_record_a(w, R, j);
            if (R.ffid) xb_record_form<DO>(w, R, j, YPR ? &y : nullptr);
        }
        __syncthreads();
        if (on) xb_record_b(w, R, j, 0, cnt, P);
        __syncthreads();
        XoQ rot1 = xo_q(0.0f, 0.0f, 0.0f, 1.0f);
        XoV3 w1 = xo_v(0.0f, 0.0f, 0.0f);
        if (on) {
            xb_phase1<DO>(w, R, j, 0, cnt, P, rot1, w1);
            if (xb_row_live(w)) {
                c.last_c = v3(0.0, 0.0, 0.0);
                c.nh = 0;
            }
            xb_coll_record(w, c, C, j, P);
        }
        __syncthreads();
        const bool warped = on && xb_row_live(w) && w.s.mode == XB_MODE_WARP && w.s.effect_stamp >= 0;
        XbV3 p = v3(0.0, 0.0, 0.0), v = p;
        if (warped) xb_coll_phase3_warp(w, c, C, j, P, &p, &v);
        __syncthreads();
        if (on) {
            xb_coll_phase34<DO, YPR, ST>(w, c, warped, p, v, C, R, j, wlo, whi, P, rot1, w1, &y);
            if (xb_tick_tail(w)) {                  /* a petrified troll: SetBallFree(false)'s snap and filing */
                if (YPR && !DO) xb_yaw_pitch_roll(y, w.s, false, w.s.clock, true, P);
                xb_file(w.s, w.free_, P, c.fil);
            }
        }
        __syncthreads();
    }
    if (on) {
        xb_row_store(P, w);
        if (DO && w.turned) xb_turn_store(P, r, w.t);
        xb_coll_store(P, r, c);
        if (YPR) xb_ypr_store(s.ypr, s.stride, r, y);
        if (s.path) s.path[r] = w.path;
        if (s.iters) s.iters[r] = w.iters;
    }
}

// One lane per listed row: Ball::InsertInBoxes's outcome into the filing rows.
__global__ void xb_collide_file_kernel(xb_modes_state s, const int32_t* __restrict__ rows, int64_t n_rows) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_rows) return;
    const int64_t r = rows ? (int64_t)rows[j] : j;
    if (r < 0 || r >= s.n) return;
    const XbPark P = xb_park(s);
    XbRow w;
    xb_row_load(P, r, w);
    XbCollRow c;
    xb_coll_load(P, r, c);
    xb_file(w.s, w.free_, P, c.fil);
    xb_coll_store(P, r, c);
}

// One lane per listed row: the display orientation snapped to the row's state (its clock as `now`; no hits).
__global__ void xb_ypr_snap_kernel(xb_modes_state s, const int32_t* __restrict__ rows, int64_t n_rows) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_rows) return;
    const int64_t r = rows ? (int64_t)rows[j] : j;
    if (r < 0 || r >= s.n) return;
    const XbPark P = xb_park(s);
    XbRow w;
    xb_row_load(P, r, w);
    XbYpr y;
    xb_ypr_load(s.ypr, s.stride, r, y);
    xb_yaw_pitch_roll(y, w.s, false, w.s.clock, true, P);
    xb_ypr_store(s.ypr, s.stride, r, y);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(XB_WAVES_FLIGHT)))
xb_modes_tiled_flight(xb_modes_state s) {
    xb_modes_tiled_body<false, false>(s);
}
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(XB_WAVES_TURNING)))
xb_modes_tiled_turning(xb_modes_state s) {
    xb_modes_tiled_body<true, false>(s);
}
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(XB_WAVES_COLLIDE)))
xb_collide_tiled_flight(xb_modes_state s) {
    xb_collide_tiled_body<false, false, false>(s);
}
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(XB_WAVES_COLLIDE)))
xb_collide_tiled_turning(xb_modes_state s) {
    xb_collide_tiled_body<true, false, false>(s);
}
// With static collidables: the statics' responses compiled in (kept out of the other kernels: their working set
// costs registers and scratch every tick).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(XB_WAVES_COLLIDE)))
xb_collide_tiled_flight_st(xb_modes_state s) {
    xb_collide_tiled_body<false, false, true>(s);
}
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(XB_WAVES_COLLIDE)))
xb_collide_tiled_turning_st(xb_modes_state s) {
    xb_collide_tiled_body<true, false, true>(s);
}
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(XB_WAVES_COLLIDE)))
xb_collide_tiled_flight_ypr_st(xb_modes_state s) {
    xb_collide_tiled_body<false, true, true>(s);
}
// The flight tick keeping the display orientation (without dynamical orientation only).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(XB_WAVES_FLIGHT)))
xb_modes_tiled_flight_ypr(xb_modes_state s) {
    xb_modes_tiled_body<false, true>(s);
}
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(XB_WAVES_COLLIDE)))
xb_collide_tiled_flight_ypr(xb_modes_state s) {
    xb_collide_tiled_body<false, true, false>(s);
}

__device__ __forceinline__ XbRecords xb_global_records(const xb_modes_state& s) {
    XbRecords R;
    R.p0x = s.scratch_f64 + XB_S_POS0 * s.stride;
    R.p0y = R.p0x + s.stride;
    R.p0z = R.p0y + s.stride;
    R.v0x = s.scratch_f64 + XB_S_VEL0 * s.stride;
    R.v0y = R.v0x + s.stride;
    R.v0z = R.v0y + s.stride;
    R.vy_after = s.scratch_f64 + XB_S_VY_AFTER * s.stride;
    R.radius = s.f64 + XB_F_RADIUS * s.stride;
    R.id = s.i64 + XB_I_IDS * s.stride;
    R.act = s.scratch_u8;
    R.fq = nullptr;                   // the general launch has no formations (refused at the entry point)
    R.ffid = nullptr;
    return R;
}

__global__ void __launch_bounds__(256) xb_modes_step_a(xb_modes_state s) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= s.n || (s.active && !s.active[r])) return;
    XbRecords R = xb_global_records(s);
    const XbPark P = xb_park(s);
    XbRow w;
    xb_row_load(P, r, w);
    const XbShip& sh = w.s;
    bool act = xb_row_live(w) && sh.mode == XB_MODE_WARP && sh.effect_stamp < 0 &&
               xb_aligned(sh.gp, sh.pos, sh.vel, sh.max_vel, sh.effect_stamp);
    R.act[r] = act ? 1 : 0;           // the id and radius records are the park's own rows, read only
    R.p0x[r] = sh.pos.x, R.p0y[r] = sh.pos.y, R.p0z[r] = sh.pos.z;
    R.v0x[r] = sh.vel.x, R.v0y[r] = sh.vel.y, R.v0z[r] = sh.vel.z;
}

__global__ void __launch_bounds__(256) xb_modes_step_b(xb_modes_state s) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= s.n || (s.active && !s.active[r])) return;
    XbRow w;
    xb_row_load(xb_park(s), r, w);
    xb_record_b(w, xb_global_records(s), (int)r, 0, (int)s.n, xb_park(s));
}

template <bool DO, bool YPR>
__global__ void __launch_bounds__(256) xb_modes_step_c(xb_modes_state s) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= s.n || (s.active && !s.active[r])) return;
    const XbPark P = xb_park(s);
    XbRow w;
    xb_row_load(P, r, w);
    if (DO) xb_turn_load(P, r, w.t);
    XbYpr y;
    if (YPR) xb_ypr_load(s.ypr, s.stride, r, y);
    xb_tick_row<DO, YPR>(w, xb_global_records(s), (int)r, 0, (int)s.n, P, &y);
    xb_row_store(P, w);
    if (DO && w.turned) xb_turn_store(P, r, w.t);
    if (YPR) xb_ypr_store(s.ypr, s.stride, r, y);
    if (s.path) s.path[r] = w.path;
    if (s.iters) s.iters[r] = w.iters;
}

// Dynamical orientation's set-up of the listed rows (rows NULL: rows 0 .. n_rows - 1): what 0 SnapOrientation to the
// row's velocity; 1 a new ball's old rotation (rot x RotationQuaternion(angVel, -dt |angVel|)), then the snap.
__global__ void xb_orient_setup_kernel(xb_modes_state s, const int32_t* __restrict__ rows, int64_t n_rows, int what) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_rows) return;
    const int64_t r = rows ? (int64_t)rows[j] : j;
    if (r < 0 || r >= s.n) return;
    const XbPark P = xb_park(s);
    XbTurn t;
    xb_turn_load(P, r, t);
    t.old_rot = xo_q((float)XB_F(P, XB_F_OLD_ROT, r), (float)XB_F(P, XB_F_OLD_ROT + 1, r),
                     (float)XB_F(P, XB_F_OLD_ROT + 2, r), (float)XB_F(P, XB_F_OLD_ROT + 3, r));
    t.old_w = xo_v((float)XB_F(P, XB_F_OLD_ANG_VEL, r), (float)XB_F(P, XB_F_OLD_ANG_VEL + 1, r),
                   (float)XB_F(P, XB_F_OLD_ANG_VEL + 2, r));
    t.torque = xo_v((float)XB_F(P, XB_F_TORQUE, r), (float)XB_F(P, XB_F_TORQUE + 1, r),
                    (float)XB_F(P, XB_F_TORQUE + 2, r));
    t.roll = (float)XB_F(P, XB_F_ROT_ROLL, r);
    if (what == 1) t.old_rot = xo_new_ball_old_rot(t.rot, t.w, s.fdt);
    xo_snap(xb_ld3(P, XB_F_VEL, r), t);
    xb_turn_store(P, r, t);
}

// One lane per listed row (rows NULL: rows 0 .. n_rows - 1): the row's command of the batch, if any.
__global__ void xb_commands_kernel(xb_modes_state s, const int32_t* __restrict__ rows, int64_t n_rows,
                                   const uint8_t* __restrict__ op, const double* __restrict__ point,
                                   const int64_t* __restrict__ target, const double* __restrict__ range,
                                   const double* __restrict__ min_range, const int64_t* __restrict__ warp_factor) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_rows) return;
    const int64_t r = rows ? (int64_t)rows[j] : j;
    if (r < 0 || r >= s.n) return;
    xb_cmd_apply(xb_park(s), s.n, r, op, point, target, range, min_range, warp_factor);
}

// The command batch's steps, one lane per listed row (rows NULL: rows 0 .. n_rows - 1), or per park row (release).
__global__ void xb_command_gather_kernel(xb_modes_state s, const int32_t* __restrict__ rows, int64_t n_rows,
                                         xb_command* __restrict__ cmds, int32_t* __restrict__ status) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_rows) return;
    const int64_t r = rows ? (int64_t)rows[j] : j;
    if (r < 0 || r >= s.n) return;
    const int refuse = xb_cmd_gather(xb_park(s), s.n, r, cmds[j], s.ypr);
    if (refuse && status) atomicOr(status, refuse);
}

__global__ void xb_command_run_kernel(xb_modes_state s, const int32_t* __restrict__ rows, int64_t n_rows,
                                      xb_command* __restrict__ cmds, uint8_t* __restrict__ mark) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_rows) return;
    const int64_t r = rows ? (int64_t)rows[j] : j;
    if (r < 0 || r >= s.n) return;
    if (xb_cmd_run(xb_park(s), r, cmds[j], s.ypr) && mark) mark[r] = 1;
}

__global__ void xb_command_release_kernel(xb_modes_state s, const uint8_t* __restrict__ mark) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= s.n) return;
    xb_cmd_release(xb_park(s), s.n, r, mark);
}

__global__ void xb_command_unmark_kernel(int64_t n, const int32_t* __restrict__ rows, int64_t n_rows,
                                         uint8_t* __restrict__ mark) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_rows) return;
    const int64_t r = rows ? (int64_t)rows[j] : j;
    if (r >= 0 && r < n) mark[r] = 0;
}

__global__ void xb_maths_kernel(int func, int64_t n, const double* __restrict__ x, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
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
        default: out[i] = xb_nan(); break;
    }
}

int xb_launch_rc() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

bool xb_block_ok(int block) { return block > 0 && block <= 256 && block % 64 == 0; }

// The allocation-free tick's scratch: the general launch's records (XB_NS float64 rows, then one uint8 row) for up to
// XB_NOALLOC_MAX_ROWS rows, in one array of the library's own on each device that loads it (about 15 MB, held until
// the process ends). Host code never passes the array's name to a kernel: the name is a host-side stand-in, so the
// device address of each device's copy is taken once per device (hipGetSymbolAddress, with that device current) and
// kept in scratch_ptr.
constexpr size_t XB_NOALLOC_F64_BYTES = (size_t)XB_NS * sizeof(double) * XB_NOALLOC_MAX_ROWS;
static __device__ __align__(256) unsigned char scratch[XB_NOALLOC_F64_BYTES + XB_NOALLOC_MAX_ROWS];

constexpr int XB_MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_ptr[XB_MAX_DEVICES] = {};

// The current device's copy of `scratch`: its address looked up on the first call on that device, then reused.
int xb_scratch_of_current_device(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= XB_MAX_DEVICES) return -1;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return -1;
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" {

int xb_abi(void) { return XB_ABI_VERSION; }

const char* xb_error_string(int rc) {
    if (rc == 0) return "success";
    if (rc == XB_ERR_TOO_MANY_ROWS)
        return "refused before launch: the park's stride is above XB_NOALLOC_MAX_ROWS, the most rows the "
               "allocation-free tick's scratch holds";
    if (rc < 0) return "refused before launch (block size, sizes or arguments)";
    return hipGetErrorString((hipError_t)rc);
}

int xb_layout(int64_t* out) {
    out[0] = XB_NF;
    out[1] = XB_NI;
    out[2] = XB_NB;
    out[3] = XB_NS;
    out[4] = (int64_t)sizeof(xb_modes_state);
    return 0;
}

int xb_modes_tick(const xb_modes_state* state, void* stream, int block) {
    if (!state || state->n < 0 || state->stride < state->n || !xb_block_ok(block)) return -1;
    if (state->n == 0) return 0;
    if (!state->f64 || !state->i64 || !state->u8) return -1;
    const bool dyn = state->dynamical != 0;
    const bool ypr = state->ypr != nullptr;
    hipStream_t st = (hipStream_t)stream;
    if (state->schedule < 0 || state->schedule > 3 || (state->schedule == 2 && !state->keys) ||
        (state->schedule == 3 && !state->iters))
        return -1;
    if (dyn && ypr) return -1;      // dynamical orientation keeps the rotation, not the display orientation
    const bool forms = state->form && state->n_form > 0;
    if (forms && (!state->form_len || state->form_stride <= 0 || !state->tiles || (!dyn && !ypr))) return -1;
    const size_t form_lds = forms ? (size_t)block * XB_FORM_LDS_BYTES_PER_ROW : 0;
    if (state->collisions) {
        if (!state->tiles || state->n_tiles <= 0 || !state->coll_f64 || !state->coll_i64 ||
            (state->hits_cap > 0 && !state->hits) || state->hits_cap < 0)
            return -1;
        const size_t lds = (size_t)block * (XB_LDS_BYTES_PER_ROW + XB_COLL_LDS_BYTES_PER_ROW) + form_lds +
                           (state->fld ? (size_t)block * XB_FIELD_LDS_BYTES_PER_ROW : 0);
        const bool stat = state->statics && state->st_f64 && state->st_range;
        if (stat && (!state->st_i64 || !state->st_keys || state->st_stride <= 0)) return -1;
        if (dyn) {
            if (stat) xb_collide_tiled_turning_st<<<state->n_tiles, block, lds, st>>>(*state);
            else xb_collide_tiled_turning<<<state->n_tiles, block, lds, st>>>(*state);
        } else if (ypr) {
            if (stat) xb_collide_tiled_flight_ypr_st<<<state->n_tiles, block, lds, st>>>(*state);
            else xb_collide_tiled_flight_ypr<<<state->n_tiles, block, lds, st>>>(*state);
        } else {
            if (stat) xb_collide_tiled_flight_st<<<state->n_tiles, block, lds, st>>>(*state);
            else xb_collide_tiled_flight<<<state->n_tiles, block, lds, st>>>(*state);
        }
        return xb_launch_rc();
    }
    if (state->tiles) {
        if (state->n_tiles <= 0) return -1;
        const size_t lds = (size_t)block * XB_LDS_BYTES_PER_ROW + form_lds;
        if (dyn) xb_modes_tiled_turning<<<state->n_tiles, block, lds, st>>>(*state);
        else if (ypr) xb_modes_tiled_flight_ypr<<<state->n_tiles, block, lds, st>>>(*state);
        else xb_modes_tiled_flight<<<state->n_tiles, block, lds, st>>>(*state);
        return xb_launch_rc();
    }
    if (!state->scratch_f64 || !state->scratch_u8 || state->ticks > 1 || state->schedule) return -1;
    const unsigned grid = (unsigned)((state->n + block - 1) / block);
    xb_modes_step_a<<<grid, block, 0, st>>>(*state);
    int rc = xb_launch_rc();
    if (rc) return rc;
    xb_modes_step_b<<<grid, block, 0, st>>>(*state);
    rc = xb_launch_rc();
    if (rc) return rc;
    if (dyn) xb_modes_step_c<true, false><<<grid, block, 0, st>>>(*state);
    else if (ypr) xb_modes_step_c<false, true><<<grid, block, 0, st>>>(*state);
    else xb_modes_step_c<false, false><<<grid, block, 0, st>>>(*state);
    return xb_launch_rc();
}

int xb_modes_tick_noalloc(const xb_modes_state* state, void* stream, int block) {
    if (!state || state->tiles || state->collisions || state->ticks > 1 || state->schedule) return -1;
    if (state->stride > XB_NOALLOC_MAX_ROWS) return XB_ERR_TOO_MANY_ROWS;
    void* base = nullptr;
    const int rc = xb_scratch_of_current_device(&base);
    if (rc) return rc;
    xb_modes_state s = *state;
    s.scratch_f64 = (double*)base;                                   // XB_NS rows of s.stride doubles
    s.scratch_u8 = (uint8_t*)base + XB_NOALLOC_F64_BYTES;            // one row of s.stride bytes
    return xb_modes_tick(&s, stream, block);
}

int xb_modes_orient_setup(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, int what, void* stream) {
    if (!state || !state->f64 || state->n < 0 || state->stride < state->n || n_rows < 0 || what < 0 || what > 1 ||
        (!rows && n_rows > state->n))
        return -1;
    if (n_rows == 0) return 0;
    xb_orient_setup_kernel<<<(unsigned)((n_rows + 255) / 256), 256, 0, (hipStream_t)stream>>>(*state, rows, n_rows,
                                                                                               what);
    return xb_launch_rc();
}

int xb_collide_file(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, void* stream) {
    if (!state || !state->f64 || !state->i64 || !state->u8 || !state->coll_f64 || !state->coll_i64 || state->n < 0 ||
        state->stride < state->n || n_rows < 0 || (!rows && n_rows > state->n))
        return -1;
    if (n_rows == 0) return 0;
    xb_collide_file_kernel<<<(unsigned)((n_rows + 255) / 256), 256, 0, (hipStream_t)stream>>>(*state, rows, n_rows);
    return xb_launch_rc();
}

int xb_modes_ypr_snap(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, void* stream) {
    if (!state || !state->f64 || !state->i64 || !state->u8 || !state->ypr || state->n < 0 ||
        state->stride < state->n || n_rows < 0 || (!rows && n_rows > state->n))
        return -1;
    if (n_rows == 0) return 0;
    xb_ypr_snap_kernel<<<(unsigned)((n_rows + 255) / 256), 256, 0, (hipStream_t)stream>>>(*state, rows, n_rows);
    return xb_launch_rc();
}

int xb_modes_commands(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, const uint8_t* op,
                      const double* point, const int64_t* target, const double* range, const double* min_range,
                      const int64_t* warp_factor, void* stream) {
    if (!state || !state->f64 || !state->i64 || !state->u8 || state->n < 0 || state->stride < state->n || n_rows < 0 ||
        !op || !point || !target || !range || !min_range || !warp_factor || (!rows && n_rows > state->n))
        return -1;
    if (n_rows == 0) return 0;
    xb_commands_kernel<<<(unsigned)((n_rows + 255) / 256), 256, 0, (hipStream_t)stream>>>(
        *state, rows, n_rows, op, point, target, range, min_range, warp_factor);
    return xb_launch_rc();
}

static bool xb_command_args_ok(const xb_modes_state* state, const int32_t* rows, int64_t n_rows,
                               const xb_command* cmds) {
    return state && state->f64 && state->i64 && state->u8 && state->n >= 0 && state->stride >= state->n &&
           n_rows >= 0 && cmds && (rows || n_rows <= state->n);
}

int xb_modes_command_check(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, xb_command* cmds,
                           int32_t* status, void* stream) {
    if (!xb_command_args_ok(state, rows, n_rows, cmds) || !status) return -1;
    if (n_rows == 0) return 0;
    xb_command_gather_kernel<<<(unsigned)((n_rows + 255) / 256), 256, 0, (hipStream_t)stream>>>(*state, rows, n_rows,
                                                                                                 cmds, status);
    return xb_launch_rc();
}

int xb_modes_command_apply(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, xb_command* cmds,
                           uint8_t* mark, void* stream) {
    if (!xb_command_args_ok(state, rows, n_rows, cmds)) return -1;
    if (n_rows == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const unsigned grid = (unsigned)((n_rows + 255) / 256);
    xb_command_gather_kernel<<<grid, 256, 0, st>>>(*state, rows, n_rows, cmds, nullptr);
    int rc = xb_launch_rc();
    if (rc) return rc;
    xb_command_run_kernel<<<grid, 256, 0, st>>>(*state, rows, n_rows, cmds, mark);
    rc = xb_launch_rc();
    if (rc || !mark) return rc;
    xb_command_release_kernel<<<(unsigned)((state->n + 255) / 256), 256, 0, st>>>(*state, mark);
    rc = xb_launch_rc();
    if (rc) return rc;
    xb_command_unmark_kernel<<<grid, 256, 0, st>>>(state->n, rows, n_rows, mark);
    return xb_launch_rc();
}

int xb_maths_eval(int func, int64_t n, const double* x, double* out, void* stream, int block) {
    if (n < 0 || func < 0 || func > 18 || !xb_block_ok(block)) return -1;
    if (n == 0) return 0;
    xb_maths_kernel<<<(unsigned)((n + block - 1) / block), block, 0, (hipStream_t)stream>>>(func, n, x, out);
    return xb_launch_rc();
}

int xb_modes_info(int which, int64_t* out) {
    hipFuncAttributes fa;
    hipError_t e;
    if (which == 0) e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&xb_modes_tiled_flight));
    else if (which == 1) e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&xb_modes_step_c<false, false>));
    else if (which == 3) e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&xb_modes_tiled_turning));
    else if (which == 4) e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&xb_collide_tiled_flight));
    else if (which == 5) e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&xb_collide_tiled_turning));
    else if (which == 6) e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&xb_modes_tiled_flight_ypr));
    else if (which == 7) e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&xb_collide_tiled_flight_ypr));
    else if (which == 8) e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&xb_collide_tiled_flight_st));
    else if (which == 9) e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&xb_collide_tiled_turning_st));
    else if (which == 2) e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&xb_modes_step_a));
    else return -1;
    if (e != hipSuccess) return (int)e;
    out[0] = fa.numRegs;
    out[1] = (int64_t)fa.localSizeBytes;
    out[2] = (int64_t)fa.sharedSizeBytes;
    out[3] = fa.maxThreadsPerBlock;
    return 0;
}

}  // extern "C"
