// SPDX-License-Identifier: LGPL-2.1-or-later
//
// glibc's srand() and rand() per world on the device (xb_rand.h): one lane per world.
#include <hip/hip_runtime.h>

#include "../../include/xb.h"
#include "xb_rand.h"

namespace {

__global__ void xb_rand_seed_kernel(const uint32_t* __restrict__ seeds, int64_t n, int32_t* __restrict__ states) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    xb_srand(seeds[w], states + w * XB_RAND_WORDS);
}

__global__ void xb_rand_draw_kernel(int32_t* __restrict__ states, int64_t n, int32_t draws, int32_t* __restrict__ out) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    int32_t st[XB_RAND_WORDS];
    for (int i = 0; i < XB_RAND_WORDS; ++i) st[i] = states[w * XB_RAND_WORDS + i];
    for (int32_t k = 0; k < draws; ++k) out[w * (int64_t)draws + k] = xb_rand(st);
    for (int i = 0; i < XB_RAND_WORDS; ++i) states[w * XB_RAND_WORDS + i] = st[i];
}

int rc_of() {
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

}  // namespace

extern "C" {

int xb_rand_seed(const uint32_t* seeds, int64_t n_worlds, int32_t* states, void* stream) {
    if (n_worlds < 0 || (n_worlds > 0 && (!seeds || !states))) return -1;
    if (n_worlds == 0) return 0;
    xb_rand_seed_kernel<<<(unsigned)((n_worlds + 255) / 256), 256, 0, (hipStream_t)stream>>>(seeds, n_worlds, states);
    return rc_of();
}

int xb_rand_draw(int32_t* states, int64_t n_worlds, int32_t draws, int32_t* out, void* stream) {
    if (n_worlds < 0 || draws < 0 || (n_worlds > 0 && (!states || (draws > 0 && !out)))) return -1;
    if (n_worlds == 0 || draws == 0) return 0;
    xb_rand_draw_kernel<<<(unsigned)((n_worlds + 255) / 256), 256, 0, (hipStream_t)stream>>>(states, n_worlds, draws,
                                                                                            out);
    return rc_of();
}

}  // extern "C"
