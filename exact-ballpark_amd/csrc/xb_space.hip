// SPDX-License-Identifier: MIT
//
// The master's space layer on gfx950 (xb_space.h): one lane a world, the world-interleaved tables read by a
// wavefront's 64 worlds as consecutive words; the tape's operations and the queries as C ABI entry points.
#include <hip/hip_runtime.h>

#define XB_HD __host__ __device__
#define XB_RAND_HD __host__ __device__
#define XB_TABLE __constant__
#include "xb_modes_tick.h"
#include "xb_rand.h"
#include "../../include/xb.h"
#include "xb_park.h"
#include "xb_space.h"

namespace {

__global__ void __launch_bounds__(64) xb_space_tape_kernel(xb_modes_state park, xb_space_state space, int64_t K,
                                                           const int64_t* __restrict__ code,
                                                           const int64_t* __restrict__ a,
                                                           const int64_t* __restrict__ b,
                                                           const double* __restrict__ f) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t W = space.W;
    if (w >= W) return;
    const XsCtx c{&space, xb_park(park), w, W, space.N};
    for (int64_t k = 0; k < K; ++k) {
        const double fk[4] = {f[(0 * K + k) * W + w], f[(1 * K + k) * W + w], f[(2 * K + k) * W + w],
                              f[(3 * K + k) * W + w]};
        xs_apply(c, code[k * W + w], a[k * W + w], b[k * W + w], fk);
    }
}

__global__ void __launch_bounds__(64) xb_space_query_kernel(xb_modes_state park, xb_space_state space, int64_t kind,
                                                            const int64_t* __restrict__ a,
                                                            const int64_t* __restrict__ b,
                                                            const double* __restrict__ f, int64_t* __restrict__ out,
                                                            double* __restrict__ outf) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t W = space.W;
    if (w >= W) return;
    const XsCtx c{&space, xb_park(park), w, W, space.N};
    const double fw[5] = {f[0 * W + w], f[1 * W + w], f[2 * W + w], f[3 * W + w], f[4 * W + w]};
    xs_query(c, kind, a[w], b[w], fw, out, outf);
}

bool space_args_ok(const xb_modes_state* park, const xb_space_state* space) {
    return park && park->f64 && park->i64 && park->u8 && space && space->W >= 0 && space->N >= 1 && space->N <= 64 &&
           space->rows && space->by_id && space->ids;
}

}  // namespace

extern "C" {

int xb_space_layout(int64_t* out) {
    out[0] = (int64_t)sizeof(xb_space_state);
    out[1] = XS_OPS;
    out[2] = 64;
    return 0;
}

int xb_space_tape(const xb_modes_state* park, const xb_space_state* space, int64_t K, const int64_t* code,
                  const int64_t* a, const int64_t* b, const double* f, void* stream) {
    if (!space_args_ok(park, space) || K < 0 || (K > 0 && (!code || !a || !b || !f))) return -1;
    if (K == 0 || space->W == 0) return 0;
    const unsigned grid = (unsigned)((space->W + 63) / 64);
    xb_space_tape_kernel<<<grid, 64, 0, (hipStream_t)stream>>>(*park, *space, K, code, a, b, f);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

int xb_space_query(const xb_modes_state* park, const xb_space_state* space, int64_t kind, const int64_t* a,
                   const int64_t* b, const double* f, int64_t* out, double* outf, void* stream) {
    if (!space_args_ok(park, space) || !a || !b || !f || !out) return -1;
    if (space->W == 0) return 0;
    const unsigned grid = (unsigned)((space->W + 63) / 64);
    xb_space_query_kernel<<<grid, 64, 0, (hipStream_t)stream>>>(*park, *space, kind, a, b, f, out, outf);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

}  // extern "C"
