// The orientation searches with their screen against the same searches without it (exact-ballpark_amd/csrc/
// xb_orient.h, XO_SCREEN): built twice by tests/test_orient_screen.py, once as-is (-DSIDE=screened) and once with
// -DXO_SCREEN=0 -DSIDE=plain, and linked together with -DMAIN. Seeded states over the ranges the tick meets and
// beyond (angular speeds to 3 rad/s and none, decay factors near 0 and near 1, tiny and huge directions, zero
// components); every torque and roll torque compared bit for bit. Prints one JSON object.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#define XB_HD
#include "../exact-ballpark_amd/csrc/xb_modes_tick.h"

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)

struct In { float mav, tscale, fdt, dir[3], rot[4], w[3], ht[3]; };

extern "C" void CAT(torque, SIDE)(const In* in, float* out) {
    XoShip k;
    k.max_ang_vel = in->mav, k.t_scale = in->tscale, k.fdt = in->fdt, k.dt = (double)in->fdt;
    k.tf = xb_expf(-k.fdt * k.t_scale);
    const XoQ q = xo_q(in->rot[0], in->rot[1], in->rot[2], in->rot[3]);
    const XoV3 w = xo_v(in->w[0], in->w[1], in->w[2]);
    const XoV3 t = xo_torque_breaks(k, xo_v(in->dir[0], in->dir[1], in->dir[2]), q, w);
    out[0] = t.x, out[1] = t.y, out[2] = t.z;
    out[3] = xo_roll_breaks(k, t, q, w);
    out[4] = xo_roll_breaks(k, xo_v(in->ht[0], in->ht[1], in->ht[2]), q, w);
}

#ifdef MAIN
extern "C" void torque_plain(const In* in, float* out);
static uint64_t s = 0x9E3779B97F4A7C15ULL;
static double u01() {
    s ^= s << 13, s ^= s >> 7, s ^= s << 17;
    return (double)(s >> 11) * 0x1p-53;
}
static float uf(double a, double b) { return (float)(a + (b - a) * u01()); }
int main(int argc, char** argv) {
    const long n = argc > 1 ? atol(argv[1]) : 1000000;
    long bad = 0, nonzero = 0;
    for (long i = 0; i < n; ++i) {
        In in;
        in.mav = u01() < 0.05 ? 0.0f : uf(0.05, 3.0);
        in.tscale = (float)pow(10.0, -4.5 + 6.0 * u01());
        const float dts[4] = {0.25f, 0.5f, 1.0f, 2.0f};
        in.fdt = dts[s % 4];
        const double mag = pow(10.0, -8.0 + 16.0 * u01());
        for (int c = 0; c < 3; ++c) in.dir[c] = u01() < 0.1 ? 0.0f : (float)(mag * (2.0 * u01() - 1.0));
        float qn = 0.0f;
        for (int c = 0; c < 4; ++c) in.rot[c] = uf(-1.0, 1.0), qn += in.rot[c] * in.rot[c];
        qn = sqrtf(qn);
        for (int c = 0; c < 4; ++c) in.rot[c] /= qn;
        const double wm = u01() < 0.2 ? 0.0 : 3.0 * u01();
        for (int c = 0; c < 3; ++c) in.w[c] = (float)(wm * (2.0 * u01() - 1.0));
        for (int c = 0; c < 3; ++c) in.ht[c] = uf(-3.0, 3.0) * in.mav * in.tscale;
        float a[5], b[5];
        torque_screened(&in, a);
        torque_plain(&in, b);
        if (memcmp(a, b, sizeof a) != 0) {
            bool nan_same = true;
            for (int c = 0; c < 5; ++c) nan_same &= (a[c] == b[c]) || (a[c] != a[c] && b[c] != b[c]);
            if (!nan_same) {
                if (bad < 5) fprintf(stderr, "differ at %ld: %a %a %a %a %a vs %a %a %a %a %a\n", i, a[0], a[1], a[2], a[3],
                                     a[4], b[0], b[1], b[2], b[3], b[4]);
                ++bad;
            }
        }
        nonzero += (a[0] != 0.0f || a[1] != 0.0f || a[2] != 0.0f) && a[3] != 0.0f;
    }
    printf("{\"cases\": %ld, \"differ\": %ld, \"torque_and_roll_nonzero\": %ld}\n", n, bad, nonzero);
    return 0;
}
#endif
