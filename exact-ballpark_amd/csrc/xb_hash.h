// SPDX-License-Identifier: MIT
//
// The serving loop's counter-based draws: SplitMix64's output function (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014; the finaliser constants of Vigna's public-domain splitmix64.c) applied
// to (seed, field, ship, tick) in three chained rounds, and a uniform double in [0, 1) from the top 53 bits. Host and
// device share this header; exact_ballpark_amd/serve.py restates it with numpy's wrapping uint64 arithmetic.
#pragma once

#include <cstdint>

#ifndef XB_HASH_HD
#if defined(__HIPCC__)
#define XB_HASH_HD __host__ __device__
#else
#define XB_HASH_HD
#endif
#endif

// the draws' fields
enum {
    XB_EP_SHIPS = 0, XB_EP_LENGTH = 1, XB_EP_MASS = 2, XB_EP_RADIUS = 3, XB_EP_MAX_VEL = 4, XB_EP_AGILITY = 5,
    XB_EP_POS = 6,      // 6, 7, 8: x, y, z
    XB_EP_INTERACTIVE = 9, XB_EP_GAP = 10, XB_EP_OP = 11,
    XB_EP_POINT = 12,   // 12, 13, 14
    XB_EP_TARGET = 15, XB_EP_RANGE = 16
};

#define XB_EP_POINT_HALF 5.0e9          // a command's point: uniform in the cube of half-side 5e9 m (within 1e7 km)

// the policy's five commands, by draw: stop, goto, follow, orbit, warp (the park's command codes)
#if defined(__HIPCC__)
__constant__
#endif
static const uint8_t xb_policy_ops[5] = {3, 1, 4, 5, 6};

static inline XB_HASH_HD uint64_t xb_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline XB_HASH_HD uint64_t xb_key(uint64_t seed, uint64_t field, uint64_t ship, uint64_t tick) {
    const uint64_t g = 0x9E3779B97F4A7C15ull;
    uint64_t h = xb_mix64(seed + g * (field + 1));
    h = xb_mix64(h + g * (ship + 1));
    return xb_mix64(h + g * (tick + 1));
}

static inline XB_HASH_HD double xb_u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }
