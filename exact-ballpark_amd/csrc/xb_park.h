// SPDX-License-Identifier: MIT
//
// The tick's view of a park (XbPark) from the C ABI's xb_modes_state: the blocks, the integrator's per-park constants
// (friction^2, 1/friction^2, 1/friction), the optional tables. One builder for the gfx950 kernels and the host build.
// Include after xb_modes_tick.h and include/xb.h.
#pragma once

static inline XB_HD XbPark xb_park(const xb_modes_state& s) {
    XbPark P;
    P.f = s.f64;
    P.q = s.i64;
    P.b = s.u8;
    P.stride = s.stride;
    P.friction = s.friction;
    P.dt = s.dt;
    P.k2 = s.friction * s.friction;
    P.ook2 = 1.0 / P.k2;
    P.ook = 1.0 / s.friction;
    P.new_orbit = s.new_orbit;
    P.dynamical = s.dynamical;
    P.cf = s.coll_f64;
    P.cq = s.coll_i64;
    P.hits = s.hits;
    P.hits_cap = s.hits_cap;
    P.statics = s.statics && s.st_f64 && s.st_range;
    P.st_range = s.st_range;
    P.st_f = s.st_f64;
    P.st_q = s.st_i64;
    P.st_keys = s.st_keys;
    P.st_stride = s.st_stride;
    P.tick_ms = (int64_t)(s.dt * 1000.0 + 0.5);
    P.form = s.form;
    P.form_len = s.form_len;
    P.n_form = s.form ? s.n_form : 0;
    P.form_stride = s.form_stride;
    P.fld = s.fld;
    P.fslots = P.n_form > 0 ? s.form_slots : nullptr;
    return P;
}
