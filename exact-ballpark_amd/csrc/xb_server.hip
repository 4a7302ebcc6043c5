// SPDX-License-Identifier: MIT
//
// The server layer's pre-tick on gfx950 (xb_server.h): one lane a world, a wavefront's 64 lanes 64 consecutive worlds,
// so each table entry of a call (its mask, ball, arguments, queued flag, route) is one contiguous read or write across
// the wavefront; the park's rows a world's calls touch are that world's own, and so are its entries of the
// partition's tables (the same world-interleaved layout) when the server runs over the partition.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#define XB_HD __device__
#define XB_TABLE __constant__
#include "xb_modes_tick.h"
#include "xb_maths_float.h"
#include "xb_commands.h"
#include "xb_rand.h"
#include "../../include/xb.h"
#include "xb_park.h"
#include "xb_space.h"
#include "xb_server.h"

namespace {

/* worlds of one bubble: the partition's code folds away (its context null), so this kernel keeps its own registers */
__global__ void __launch_bounds__(64) xb_server_kernel(xb_modes_state park, xb_server_state srv, xb_server_queue q,
                                                       int phase) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= srv.W) return;
    xsv_world(xb_park(park), park.n, park.ypr, srv, q, nullptr, phase, w);
}

/* over the partition: its tables by value beside the server's */
__global__ void __launch_bounds__(64) xb_server_space_kernel(xb_modes_state park, xb_server_state srv,
                                                             xb_server_queue q, xb_space_state sp, int phase) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= srv.W) return;
    xsv_world(xb_park(park), park.n, park.ypr, srv, q, &sp, phase, w);
}

}  // namespace

extern "C" {

int xb_server_layout(int64_t* out) {
    xsv_layout(out);
    return 0;
}

int xb_server_pre_tick(const xb_modes_state* park, const xb_server_state* srv, const xb_server_queue* q, int phase,
                       void* stream) {
    if (!xsv_args_ok(park, srv, q, phase)) return -1;
    if (srv->W == 0) return 0;
    const unsigned grid = (unsigned)((srv->W + 63) / 64);
    if (srv->space)                               /* the partition's tables by value (srv->space is a host address) */
        xb_server_space_kernel<<<grid, 64, 0, (hipStream_t)stream>>>(*park, *srv, *q, *srv->space, phase);
    else
        xb_server_kernel<<<grid, 64, 0, (hipStream_t)stream>>>(*park, *srv, *q, phase);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : (int)e;
}

}  // extern "C"
