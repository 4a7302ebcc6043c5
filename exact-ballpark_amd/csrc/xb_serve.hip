// SPDX-License-Identifier: MIT
//
// The rollout-serving loop's seeded random policy on the device: for every live row of the resident park, whether its
// ship issues a mode command at this tick of its episode, and which (stop, goto a point, follow or orbit another ship
// of its episode, warp to a point), with the command's arguments. Every draw is a counter-based hash of (episode seed,
// field, ship, episode tick), so an episode's commands depend on nothing but the episode: not on the step it was
// admitted at, its rows, or the other episodes in the park. exact_ballpark_amd/serve.py restates the same draws on the
// host (the episode generator and the replay of an episode's commands).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/xb.h"
#include "xb_hash.h"

namespace {

__global__ void xb_policy_kernel(xb_episode_rows e, const int32_t* __restrict__ live, int64_t n_live, int64_t step,
                                 uint8_t* __restrict__ op, double* __restrict__ point, int64_t* __restrict__ target,
                                 double* __restrict__ range) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_live) return;
    const int64_t r = live[j];
    const uint64_t seed = e.seed[r];
    const int64_t ship = e.ship[r];
    const int64_t t = step - e.admit[r] + 1;                       // the episode's tick about to run (first: 1)
    uint8_t o = 0;
    if (t == e.next_cmd[r]) {
        const uint64_t ut = (uint64_t)t, us = (uint64_t)ship;
        e.next_cmd[r] = t + 1 + (int64_t)(xb_key(seed, XB_EP_GAP, us, ut) % 10u);
        o = xb_policy_ops[xb_key(seed, XB_EP_OP, us, ut) % 5u];
        for (int a = 0; a < 3; ++a)
            point[r * 3 + a] = (2.0 * xb_u01(xb_key(seed, XB_EP_POINT + a, us, ut)) - 1.0) * XB_EP_POINT_HALF;
        const int64_t size = e.size[r];
        int64_t other = (int64_t)(xb_key(seed, XB_EP_TARGET, us, ut) % (uint64_t)(size > 1 ? size - 1 : 1));
        if (other >= ship) other += 1;
        target[r] = e.base[r] + (other < size ? other : size - 1);
        range[r] = 500.0 + xb_u01(xb_key(seed, XB_EP_RANGE, us, ut)) * 19500.0;
    }
    op[r] = o;
}

}  // namespace

extern "C" {

int xb_serve_policy(const xb_episode_rows* e, const int32_t* live, int64_t n_live, int64_t step, uint8_t* op,
                    double* point, int64_t* target, double* range, void* stream) {
    if (!e || !e->seed || !e->ship || !e->size || !e->base || !e->admit || !e->next_cmd || n_live < 0 || !op ||
        !point || !target || !range || (n_live > 0 && !live))
        return -1;
    if (n_live == 0) return 0;
    const int64_t grid = (n_live + 255) / 256;
    xb_policy_kernel<<<(unsigned)grid, 256, 0, (hipStream_t)stream>>>(*e, live, n_live, step, op, point, target,
                                                                        range);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : (int)err;
}

}  // extern "C"
