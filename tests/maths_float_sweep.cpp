// The float maths (exact-ballpark_amd/csrc/xb_maths_float.h) swept against this machine's C library: every
// `stride`-th float bit pattern for expf, logf, sinf, cosf, acosf, atanf and powf(x, 0.25f) (the mushroom growth's
// call); 2e6 seeded pairs for atan2f, pow, powf (any two bit patterns, and a negative x with an integral y), the
// double asin (any bit pattern, and [-1, 1]) and atan2 (any two bit patterns, and two normal magnitudes). Built
// by tests/test_maths_cpu.py twice: as shipped (aarch64's fused steps) and with the fdlibm functions' fused steps
// rounded twice (-DXB_FDLIBM_FMAF), the way this x86-64 C library's fdlibm code is compiled. Prints one JSON object.
// An expf mismatch is counted apart when x 32/ln2 is exactly halfway between integers: aarch64 rounds that tie away
// from zero (frinta), x86-64's build to even.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#define XB_HD
#include "../exact-ballpark_amd/csrc/xb_maths_float.h"

static float fl(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bi(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static bool same(float a, float b) { return bi(a) == bi(b) || (a != a && b != b); }

int main(int argc, char** argv) {
    const long long stride = argc > 1 ? atoll(argv[1]) : 1;
    long long bad[7] = {0, 0, 0, 0, 0, 0, 0}, expf_ties = 0;
#pragma omp parallel for reduction(+ : bad[:7], expf_ties) schedule(static, 4096)
    for (long long i = 0; i < (1LL << 32); i += stride) {
        const float x = fl((uint32_t)i);
        if (!same(xb_expf(x), expf(x))) {
            const double z = (double)x * XB_LF_EXPF_INVLN2;
            if (z - std::trunc(z) == 0.5 || z - std::trunc(z) == -0.5) expf_ties++;
            else bad[0]++;
        }
        bad[1] += !same(xb_logf(x), logf(x));
        bad[2] += !same(xb_sinf(x), sinf(x));
        bad[3] += !same(xb_cosf(x), cosf(x));
        bad[4] += !same(xb_acosf(x), acosf(x));
        bad[5] += !same(xb_atanf(x), atanf(x));
        bad[6] += !same(xb_powf(x, 0.25f), powf(x, 0.25f));
    }
    long long bad_atan2 = 0, bad_pow = 0, bad_powf = 0, bad_asin = 0, bad_atan2d = 0;
    uint64_t s = 88172645463325252ULL;
    for (int i = 0; i < 2000000; ++i) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        const float y = fl((uint32_t)s), x = fl((uint32_t)(s >> 32));
        bad_atan2 += !same(xb_atan2f(y, x), atan2f(y, x));
        const double px = std::fabs((double)fl((uint32_t)s)), py = (double)(int)((s >> 40) % 200) - 100;
        const double pa = xb_pow(px, py), pb = pow(px, py);
        bad_pow += memcmp(&pa, &pb, 8) != 0 && !(pa != pa && pb != pb);
        const float qy = (float)((int)((s >> 24) % 61) - 30), qx = -fl((uint32_t)(s >> 8) & 0x3fffffffu);
        bad_powf += !same(xb_powf(x, y), powf(x, y)) + !same(xb_powf(qx, qy), powf(qx, qy));
        double dx, dy, du, dv;
        const uint64_t t = s * 0x9E3779B97F4A7C15ULL;
        memcpy(&dx, &s, 8);
        memcpy(&dy, &t, 8);
        du = (double)(s >> 11) * 0x1p-52 - 1.0;
        dv = (double)(t >> 11) * 0x1p-53 * (double)((int)(s % 41) - 20);
        const double a0 = xb_asin(dx), a1 = asin(dx), a2 = xb_asin(du), a3 = asin(du);
        bad_asin += (memcmp(&a0, &a1, 8) != 0 && !(a0 != a0 && a1 != a1)) + (memcmp(&a2, &a3, 8) != 0);
        const double b0 = xb_atan2(dy, dx), b1 = atan2(dy, dx), b2 = xb_atan2(dv, du), b3 = atan2(dv, du);
        bad_atan2d += (memcmp(&b0, &b1, 8) != 0 && !(b0 != b0 && b1 != b1)) + (memcmp(&b2, &b3, 8) != 0);
    }
    printf("{\"expf\": %lld, \"expf_ties\": %lld, \"logf\": %lld, \"sinf\": %lld, \"cosf\": %lld, \"acosf\": %lld, "
           "\"atanf\": %lld, \"atan2f\": %lld, \"pow\": %lld, \"powf_quarter\": %lld, \"powf\": %lld, "
           "\"asin\": %lld, \"atan2\": %lld}\n",
           bad[0], expf_ties, bad[1], bad[2], bad[3], bad[4], bad[5], bad_atan2, bad_pow, bad[6], bad_powf, bad_asin,
           bad_atan2d);
    return 0;
}
