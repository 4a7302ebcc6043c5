/* SPDX-License-Identifier: LGPL-2.1-or-later
 *
 * The GNU C Library's srand() and rand() (srandom_r and random_r with the default TYPE_3 table: 31 words, separation
 * 3), as destiny's floatrand draws it for a shuffled proximity sensor (Ball.cpp:19), one generator per world. The
 * algorithm is glibc's (stdlib/random_r.c): the table seeded by the Park-Miller minimal standard generator 16807 x mod
 * 2^31 - 1 (Schrage's decomposition, 127773 = m / a, 2836 = m mod a), 310 draws discarded, then each draw the sum of
 * the words at the front and rear pointers modulo 2^32, stored at the front, shifted right by one. Integer operations
 * only, host and device. The reference restates it at src/exact_ballpark/glibc_rand.py:14-45.
 *
 * State of one generator: XB_RAND_WORDS int32 = the 31 table words, then the front and rear indices.
 */
#pragma once

#include <stdint.h>

#ifndef XB_RAND_HD
#if defined(__HIPCC__)
#define XB_RAND_HD __host__ __device__
#else
#define XB_RAND_HD
#endif
#endif

#define XB_RAND_DEG 31
#define XB_RAND_SEP 3
#define XB_RAND_WORDS (XB_RAND_DEG + 2)

static inline XB_RAND_HD int32_t xb_rand(int32_t* st) {
    int32_t f = st[XB_RAND_DEG], r = st[XB_RAND_DEG + 1];
    const uint32_t val = (uint32_t)st[f] + (uint32_t)st[r];
    st[f] = (int32_t)val;
    if (++f >= XB_RAND_DEG) {
        f = 0;
        ++r;
    } else if (++r >= XB_RAND_DEG) {
        r = 0;
    }
    st[XB_RAND_DEG] = f;
    st[XB_RAND_DEG + 1] = r;
    return (int32_t)(val >> 1);
}

static inline XB_RAND_HD void xb_srand(uint32_t seed, int32_t* st) {
    int32_t word = (int32_t)(seed == 0u ? 1u : seed);
    st[0] = word;
    for (int i = 1; i < XB_RAND_DEG; ++i) {
        const int32_t hi = word / 127773, lo = word % 127773;     /* C division: toward zero */
        word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        st[i] = word;
    }
    st[XB_RAND_DEG] = XB_RAND_SEP;
    st[XB_RAND_DEG + 1] = 0;
    for (int i = 0; i < 10 * XB_RAND_DEG; ++i) (void)xb_rand(st);
}
