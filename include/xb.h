/* SPDX-License-Identifier: MIT
 *
 * The C ABI of libxb_hip.so (the batched destiny tick on gfx950) and libxb_cpu.so (the same tick source built for the
 * host). Plain pointers and sizes: the caller owns every buffer, nothing here allocates or frees, no call synchronises
 * the host, and every launch goes on the caller's HIP stream, so a caller may capture the launches in a HIP graph.
 *
 * Return codes: 0 success; -1 refused before any launch (a bad block size, size or argument; see xb_error_string);
 * a positive value is the hipError_t of the launch.
 *
 * Each entry point names the reference interface it replaces (the ctypes-bound functions of the reference's CUDA
 * sources, src/exact_ballpark/vec_*_cuda.cu) and where it is declared there.
 */
#ifndef XB_H
#define XB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XB_ABI_VERSION 9

/* A flight-modes park in device (or, for libxb_cpu.so, host) memory: the structure-of-arrays blocks laid out as
   exact-ballpark_amd/csrc/xb_layout.h defines them (field f of row r at block[f * stride + r]). Replaces the
   reference's `struct Modes` of 38 tensor pointers (src/exact_ballpark/vec_modes_cuda.cu:83-96).

   Two launch shapes:
   - tiled (`tiles` not NULL): one launch; workgroup g runs the rows tiles[g] .. tiles[g+1] - 1 (at most `block` rows,
     no world split across tiles: every follow target of a row lies in its tile), its worlds' tick-start records held
     in LDS. `rowmap`, when not NULL, maps the tiled index space to physical rows (each world's rows consecutive and in
     order in both), so worlds can be regrouped without moving their state.
   - general (`tiles` NULL): any row order; three launches through the caller's scratch (XB_NS float64 rows and one
     uint8 row of `stride` each), one tick.
   `ticks` > 1 (tiled only) runs that many ticks in the one launch: each lane keeps its row's state in registers from
   the first tick to the last, so every field is read from and written to HBM once per launch, not once per tick.
   `schedule` assigns a tile's rows to its lanes so that each wavefront runs rows of one branch class (a row's result
   never depends on its lane: it reads other rows only through the tile's records). */
typedef struct xb_modes_state {
    double* f64;
    int64_t* i64;
    uint8_t* u8;
    double* scratch_f64;
    uint8_t* scratch_u8;
    const int32_t* tiles;
    const int32_t* rowmap;
    int64_t n;
    int64_t stride;
    int32_t n_tiles;
    int32_t new_orbit;
    double friction;
    double dt;
    float ffriction;
    float fdt;
    int32_t dynamical;             /* dynamical orientation (MOV-728): the turning rows of f64 (rot, ang_vel, max_ang_vel,
                                      rot_agility in; old_rot, old_ang_vel, torque, rot_roll out), floats held in
                                      doubles; STOP without its damping; the new orbit's normal from the heading */
    int32_t ticks;                 /* ticks per launch (tiled shape only; 0 or 1: one) */
    int32_t schedule;              /* lanes of a tile: 0 in row order; 1 sorted by the built-in branch key;
                                      2 sorted by keys[row]; 3 sorted by the orientation searches' steps each row took
                                      in its last recorded tick (iters[row], bucketed) (tiled shape only) */
    int32_t n_keys;                /* the keys' bucket count, at most 16 */
    const uint8_t* keys;           /* schedule 2: each row's bucket, below n_keys */
    uint8_t* path;                 /* if not NULL: each row's executed path class in its last tick (0..7) */
    uint8_t* iters;                /* if not NULL: each row's orientation searches' steps in its last tick (GotoTorque-
                                      Breaks' and GotoRollBreaks' inner iterations, at most 255; 0 without dynamical
                                      orientation): a cost counter for the scheduler, never a result */
    const uint8_t* active;         /* if not NULL: OnTick's mask, each row's byte (a world's rows alike): a row whose
                                      byte is 0 does not tick and keeps every field bit for bit */
    /* The gradient collision method (SpaceBallpark's partition and Gradient, a client's view), tiled shape only, on
       when `collisions` is set: each world at most 64 rows, consecutive in the tile. coll_f64 holds XB_NCF rows
       (mLastC), coll_i64 XB_NCI rows (the filing: level, key x3, side x3; the tick's collision list's length), hits
       hits_cap rows (the tick's collision list in order: entry i of row r at hits[i * stride + r]; a list past
       hits_cap keeps counting, its entries past it dropped). Space junk is u8 row XB_B_JUNK, cloaking i64 row
       XB_I_CLOAKED. The filing must be current before the first tick (xb_collide_file). */
    double* coll_f64;
    int64_t* coll_i64;
    int64_t* hits;
    int32_t hits_cap;
    int32_t collisions;
    /* If not NULL (and `dynamical` not set): the display orientation Ball::CalculateYawPitchRoll keeps, XB_NY float64
       rows (yaw, pitch, roll, old_yaw, old_pitch, old_roll, roll_speed, old_roll_speed, yaw_delta), evolved at every
       live row's commit; snapped by xb_modes_ypr_snap. */
    double* ypr;
    /* FullBallpark's static collidables (capsules and oriented boxes) in the gradient method, on when `statics` is set
       (with `collisions`): a table of every world's, XB_NSF float64 rows, XB_NSI int64 rows (kind, id, parent id, the
       filing's level, its box count) and 3 st_boxes int64 rows of box keys, st_stride columns, as
       xb_cpu_statics_setup writes it; st_range two int32 rows of `stride` columns: each row's world's first static
       collidable in the table and how many (at most 64). The collision lists' room counts them too. */
    const double* st_f64;
    const int64_t* st_i64;
    const int64_t* st_keys;
    const int32_t* st_range;
    int64_t st_stride;
    int32_t st_boxes;
    int32_t statics;
    /* The park's formations (LoadFormations), when n_form > 0 (tiled shape only; without dynamical orientation the
       leaders' rotations are their display orientation's: `ypr` needed): form holds each formation's slot offsets,
       formation f's slot k at form[3 (f form_stride + k) ..], form_len[f] its slot count. A leader's formation is u8
       row XB_B_FORMATION (a char). */
    const double* form;
    const int32_t* form_len;
    int32_t n_form;
    int32_t form_stride;
    /* The force fields' keys (SetBallHarmonic), or NULL for none: int64 [3][stride], each row's harmonic, corporation
       and alliance (-1 for none). A candidate in mode FIELD is left out of the gradient when the responder's harmonic
       is -2 or one of its keys (not -1) is the responder's. */
    const int64_t* fld;
    /* The master's formation reservations (std::bitset<16> a leader, ReserveFormationSlot's bits), int64 [stride], or
       NULL on a client's view: on the master, Stop(Ball*) of a FORMATION follower (a command's, or StopAllFollowers
       inside the tick) clears its slot's bit in its leader's word (FreeFormationSlot; an atomic AND, so followers of
       one leader stopping in one launch commute). Read with form_len and the leader's formation. */
    int64_t* form_slots;
} xb_modes_state;

/* The ABI version (XB_ABI_VERSION). Replaces xbpm_abi (src/exact_ballpark/vec_modes_cuda.cu:750). */
int xb_abi(void);

/* A message for a return code. Replaces xbpo_error_string (src/exact_ballpark/vec_orient_cuda.cu:497-503). */
const char* xb_error_string(int rc);

/* The state layout as the library was built: out[0..4] = XB_NF, XB_NI, XB_NB, XB_NS, sizeof(xb_modes_state).
   Plays the part of xbps_layout (src/exact_ballpark/vec_space.cu:1853-1868) for the modes state. */
int xb_layout(int64_t* out);

/* One tick of every row of the park (Ballpark::Evolve per world, in place), on `stream` (a hipStream_t; NULL is the
   null stream). block: a multiple of 64, at most 256. Replaces xbpm_tick (src/exact_ballpark/vec_modes_cuda.cu:754). */
int xb_modes_tick(const xb_modes_state* state, void* stream, int block);

/* ---- the allocation-free tick ---- */

#define XB_NOALLOC_MAX_ROWS (1 << 18)   /* the most rows (the state's stride) xb_modes_tick_noalloc takes */
#define XB_ERR_TOO_MANY_ROWS (-2)       /* its refusal of a larger park (before any launch) */

/* No counterpart in the reference: for callers that cannot call hipMalloc (code under HIP graph capture, tight
   serving loops, harnesses that forbid allocation). xb_modes_tick's general shape (state->tiles NULL: any row order,
   one tick, three launches) with its records kept in a static device array of the library's own (about 15 MB on each
   device the library runs on, held until the process ends) in place of the caller's scratch_f64 and scratch_u8, which
   are ignored. Allocates nothing: no hipMalloc, hipFree or allocating library call. Refuses a tiled or collision
   state, ticks > 1 or a lane schedule (-1), and a stride above XB_NOALLOC_MAX_ROWS (XB_ERR_TOO_MANY_ROWS). Each
   launch writes every scratch word the tick reads before reading it, so nothing needs clearing between calls; but
   every call on one device shares that device's one array, so calls that overlap on one device (other streams or
   threads) race. */
int xb_modes_tick_noalloc(const xb_modes_state* state, void* stream, int block);

/* The tick's double functions alone, out[i] = f(x[i]) (func 0 sin, 1 cos, 2 log, 3 exp, 4 sincos's sine, 5 sincos's
   cosine, 6 IEEE sqrt, 7 IEEE division x[i] / x[n + i] with out of n elements; on floats held in doubles 8 expf, 9
   logf, 10 sinf, 11 cosf, 12 acosf, 13 atanf, 14 atan2f(x[i], x[n + i]); 15 the double pow(x[i], x[n + i]); 16 the
   double asin, 17 the double atan2(x[i], x[n + i]); 18 powf(x[i], x[n + i]) on floats held in doubles). Replaces
   xbpm_dlibm (src/exact_ballpark/vec_modes_cuda.cu:779), xbpo_libm (vec_orient_cuda.cu:520), xbpy_dlibm
   (vec_ypr_cuda.cu:102) and xbpc_powf (vec_combat_cuda.py:321). */
int xb_maths_eval(int func, int64_t n, const double* x, double* out, void* stream, int block);

/* Dynamical orientation's set-up of the listed rows (rows NULL: rows 0 .. n_rows - 1), in place: what 0
   SnapOrientation (the rotation turned to the row's velocity and rolled upright, the rotation stopped; nothing below
   1e-6 m/s), as SetBallVelocity and EntityWarpIn do it; what 1 a new ball's old rotation, rot x RotationQuaternion(
   angVel, -dt |angVel|), then the snap, as the Python AddBall does it. Replaces the batched _snap_orientation and the
   load's turning set-up (src/exact_ballpark/vec_orient.py:556-575, src/exact_ballpark/vec_modes.py:256-264). */
int xb_modes_orient_setup(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, int what, void* stream);

/* The partition filing of the listed rows (rows NULL: rows 0 .. n_rows - 1), in place: Ball::InsertInBoxes's outcome
   (the level of the inflated radius, the key of the box holding the centre, the side on each axis where the inflated
   sphere reaches the next box) into coll_i64, as AddBall, SetBallRadius, SetMaxSpeed, SetBallPosition and fixing a
   ball file it. Replaces vec_collide_cuda.filing (src/exact_ballpark/vec_collide_cuda.cu:1214-1243, `filing_of`). */
int xb_collide_file(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, void* stream);

/* The display orientation of the listed rows (rows NULL: rows 0 .. n_rows - 1) snapped to their state, as AddBall does
   for a free ball and SetBallVelocity for any (Ball::CalculateYawPitchRoll with snap: no smoothing, the roll reset).
   Needs state->ypr. Replaces the reference's `_yaw_pitch_roll(mask, snap=True)` (src/exact_ballpark/vec_orient.py:
   479-535) and vec_ypr_cuda's device maths for it. */
int xb_modes_ypr_snap(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, void* stream);

/* One flight-mode command per listed row, applied in place between ticks: for each row r of rows[0 .. n_rows - 1]
   (rows NULL: rows 0 .. n_rows - 1), op[r] 0 none, 3 Stop, 1 GotoPoint(point[r][3]), 4 Follow and 5 Orbit (target row
   target[r], range[r]), 6 WarpTo(point[r], min_range[r], warp_factor[r]); a row listed twice is a race. The result is
   the park's op-by-op commands' (a command writes only its own row). Replaces the flight-mode commands of the
   server's pre-tick kernel (cmd_stop, cmd_goto_point, cmd_follow, cmd_warp_to;
   src/exact_ballpark/vec_server_cuda.cu:117-186, applied by `apply` at :317). */
int xb_modes_commands(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, const uint8_t* op,
                      const double* point, const int64_t* target, const double* range, const double* min_range,
                      const int64_t* warp_factor, void* stream);

/* One listed row's command or setter between ticks (exact-ballpark_amd/csrc/xb_commands.h numbers the ops and gives
   each one's arguments): the caller fills op, flag, i and x and leaves the rest zero; the gather step of
   xb_modes_command_check and xb_modes_command_apply fills the rest from other rows (a target's, a launcher's). Every
   word is 8 bytes. */
typedef struct xb_command {
    int64_t op, flag;
    int64_t i[3];
    double x[4];
    double l_pos[3], l_vel[3], l_max_vel, l_rot[4];     /* gathered: the launcher */
    double t_pos[3], t_radius;                          /* gathered: the target */
    int64_t t_id, t_world, t_bubble, t_cloaked, t_state;
} xb_command;

/* The gather step alone for rows[j] with cmds[j] (j < n_rows): fills each record from the other rows it names and ORs
   into *status (device memory, cleared by the caller) the refusals the batch would meet (1 a mushroom as a follow,
   orbit or missile target; 2 a missile launcher at rest without dynamical or display orientation); writes nothing to
   the park. The reference refuses these in its masked commands before any change (vec_combat.py:339-398). */
int xb_modes_command_check(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, xb_command* cmds,
                           int32_t* status, void* stream);

/* The batch: gather, then each listed row's command on its own row (other rows read only through its record, so the
   result is that of reading every row before writing any); then, when `mark` is not NULL (n bytes, all zero on entry
   and on return), StopAllFollowers for the balls an EntityWarpIn or a removal's stop marked, over every row of the
   park. One row listed twice is a race. Replaces the reference's masked commands and setters on device tensors
   (src/exact_ballpark/vec_modes.py:286-483, vec_orient.py:403-478, vec_combat.py:339-398) and the server kernel's
   per-row commands (vec_server_cuda.cu:117-186). */
int xb_modes_command_apply(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, xb_command* cmds,
                           uint8_t* mark, void* stream);

/* Kernel attributes as loaded on the current device: which 0 tiled, 1 general main, 2 general prepass, 3 tiled with
   dynamical orientation, 4 tiled with collisions, 5 with both, 6 tiled keeping the display orientation, 7 that with
   collisions, 8 tiled with collisions and static collidables, 9 that with dynamical orientation;
   out[0..3] = vector registers a lane (hipFuncAttributes numRegs), private (scratch) bytes a lane, static LDS bytes,
   maximum threads a workgroup. Replaces xbpm_info (src/exact_ballpark/vec_modes_cuda.cu:766). */
int xb_modes_info(int which, int64_t* out);

/* ---- the master's space layer (the partition's tops and bubbles, the sensors, the queries) ---- */

/* W worlds of N slots (N <= 64), each world's tables world-interleaved: entry k of world w's table at [k * W + w]
   (k counting the table's rows; a per-slot table [q][N][W] has entry (q, j) at [(q * N + j) * W + w]). Replaces the
   reference's `struct Space` of [W, ...] tensors (src/exact_ballpark/vec_space.cu:171-237). Room per world: T tops,
   E bubble dictionary entries, D dictionaries, I pairs of each set, M events; a world past a room is flagged in
   wflags (and left outside the layer's domain). The caller zeroes every table but the following, set at creation:
   rows and by_id (as given), level -1, bubble and old_bubble -1, trk -1, s_period 2, s_bkt and p_bkt 1, rnd each
   world's glibc srand state (xb_rand_seed's). */
typedef struct xb_space_state {
    int64_t W, N, T, E, D, I, M;
    const int32_t* rows;      /* [N][W] slot j's park row (-1: none), the slots in the order the world adds them */
    const int32_t* by_id;     /* [N][W] the slots by ascending id, -1 after the last */
    int64_t* ids;             /* [N][W] each slot's ball id, from its ADD */
    int8_t* level;            /* [N][W] the filing's level, -1 for none */
    int64_t* key;             /* [3][N][W] */
    int8_t* mod;              /* [3][N][W] */
    int64_t* bubble;          /* [N][W] */
    int64_t* old_bubble;      /* [N][W] */
    int64_t* seq;             /* [N][W] the ball dictionary's place */
    uint8_t* sflags;          /* [N][W] in the park, dead, in warp proper, proximity, interactives only, in range,
                                 in tracking range, kept alive (xb_space.h's XS_* bits) */
    float* s_range;           /* [N][W] the proximity sensor's range (a float) */
    double* s_period;         /* [N][W] */
    double* s_elapsed;        /* [N][W] */
    int32_t* trk;             /* [N][W] the tracked slot, -1: none */
    double* trk_range;        /* [N][W] */
    double* ntf_range;        /* [N][W] */
    double* trk_pos;          /* [3][N][W] a removed slot's last position (the tracking pointer outlives it) */
    int8_t* s_keys;           /* [N][N][W] slot j's sensor list, key i at [(i N + j) W + w], unordered_map order */
    int32_t* s_len;           /* [N][W] */
    int32_t* s_bkt;           /* [N][W] its bucket count */
    int32_t* s_next;          /* [N][W] its rehash policy's next resize */
    int32_t* tn;              /* [W] the tops that exist */
    int32_t* top_key;         /* [3][T][W] */
    int64_t* top_bub;         /* [T][W] */
    int32_t* top_ref;         /* [T][W] the filings' boxes under it (a top with none is released) */
    int32_t* en;              /* [W] the bubble dictionaries' entries, in insertion order */
    int64_t* e_bub;           /* [E][W] */
    int8_t* e_slot;           /* [E][W] */
    int8_t* e_act;            /* [E][W] +1 added, 0 kept, -1 removed */
    int32_t* dn;              /* [W] the dictionaries, in insertion order */
    int64_t* d_bub;           /* [D][W] */
    int32_t* in_;             /* [W] bubbleInteractives as (bubble, slot) pairs */
    int64_t* i_bub;           /* [I][W] */
    int8_t* i_slot;           /* [I][W] */
    int32_t* kn;              /* [W] bubbleKeepAlives as pairs */
    int64_t* k_bub;           /* [I][W] */
    int8_t* k_slot;           /* [I][W] */
    int8_t* p_keys;           /* [N][W] the proximity balls, unordered_map order */
    int32_t* p_len;           /* [W] */
    int32_t* p_bkt;           /* [W] */
    int32_t* p_next;          /* [W] */
    int64_t* bubble_id;       /* [W] mBubbleId */
    int64_t* seq_next;        /* [W] */
    int32_t* evn;             /* [W] events since the last CLEAR_EVENTS */
    int64_t* ev;              /* [4][M][W] kind (1 DoProximity, 2 DoRange, 3 DoTargetTracking, 4 DoFinalCleanup),
                                 ball id, a, b */
    int32_t* rnd;             /* [33][W] glibc's TYPE_3 table, then its front and rear indices */
    int64_t* wflags;          /* [W] rooms passed, refusals (XS_F_*) */
} xb_space_state;

/* The layout as built: out[0] = sizeof(xb_space_state), out[1] = the op count, out[2] = the largest N. */
int xb_space_layout(int64_t* out);

/* Per world, K operations in order (op k of world w: code, a (a slot) and b (an int) at [k * W + w], its floats f at
   [(q * K + k) * W + w], q < 4), one lane a world, over the park `park` (the tick state the slots' rows index: the
   layer reads and writes its rows' bubble, interactive, sensor and, around the park's tick, moribund fields). The ops
   are xb_space.h's XS_OP_*. Replaces xbps_tape_cuda (src/exact_ballpark/vec_space.cu:1871). */
int xb_space_tape(const xb_modes_state* park, const xb_space_state* space, int64_t K, const int64_t* code,
                  const int64_t* a, const int64_t* b, const double* f, void* stream);

/* One query a world (kind XS_Q_*, a and b at [w], f [5][W]): out [1 + N][W] (word k of world w at [k * W + w]), outf
   [N][W] (the range queries' squared distances; may be NULL). Replaces xbps_query_cuda (vec_space.cu:1895). */
int xb_space_query(const xb_modes_state* park, const xb_space_state* space, int64_t kind, const int64_t* a,
                   const int64_t* b, const double* f, int64_t* out, double* outf, void* stream);

/* ---- the server layer's pre-tick (Destiny's Actions queue, Ticker, bubble update and send flags) ---- */

/* The server layer's own state for W worlds of one bubble each, over a park (the world tables world-interleaved: entry
   k of world w at [k * W + w]; the per-ball tables by park row). Replaces the reference's `struct Srv`
   (src/exact_ballpark/vec_server_cuda.cu:83-91). */
typedef struct xb_server_state {
    int64_t W, N;                 /* worlds; the most balls a world holds */
    const int32_t* rows;          /* [N][W] the park rows of each world's balls in the order they were added, -1 after */
    const int64_t* pilot;         /* [n] the pilot flying each row's ship, -1 for none */
    int64_t* stamp_for_system;    /* [W] the stamp the next queued actions get (Actions._stamp_for_system) */
    int64_t* stamp;               /* [W] written: the stamp of the actions this pre-tick applied */
    int64_t* clock;               /* [W] written: the park's time at this pre-tick */
    int64_t* bubble_time;         /* [W] BubbleUpdater's last park time (-1 before the first) */
    uint8_t* settled;             /* [n] the ball's old bubble is its bubble (AdditionsAndDeletions has seen it) */
    uint8_t* fresh;               /* [n] a user ball getting every uncloaked ball of its bubble this tick */
    uint8_t* listed;              /* [n] the ball is in those lists (uncloaked at the bubble update) */
    uint8_t* has_character;       /* [n] written: the row's pilot gets a singlecast this tick */
    uint8_t* gets_bubble;         /* [n] written: the row's pilot gets the bubble's narrowcast this tick */
    uint8_t* history;             /* [n] written: the row's character has history this tick (additions, a call routed
                                     to it, an uncloak's notice), before the pilot gate */
    uint8_t* bubble_history;      /* [W] written: the world's bubble has history this tick */
    int64_t* wflags;              /* [W] OR-ed: 1 a command the batched park refuses (a mushroom as a target) */
    int32_t master;               /* the server's view: a normal cloak adds the ball's proximity sensor */
    int32_t snapshot;             /* write the stream records (rec_add, rec_unc) the decode of the sends needs */
    /* The master's partition (the partition server; NULL: one bubble a world): the space layer's state over the same
       park and the same worlds (its `rows` the server's). With it the pre-tick also queues the follower clean-up's
       Stops, runs each call's part in the partition (re-filings, the interactive sets, the cloak's and uncloak's
       sensors), AdditionsAndDeletions across old and new bubbles in place of the one-bubble additions, and routes
       every event to its ball's bubble; the flags give each pilot the narrowcast of its ship's bubble. */
    const xb_space_state* space;
    int8_t* ad_kind;              /* [n] the bubble update's lists per row: 0 none, 1 its bubble's, 2 its own in a
                                     bubble of no pre-built list, 3 its own across its old and new bubbles */
    int32_t* ad_addn;             /* [n] */
    int32_t* ad_deln;             /* [n] */
    int8_t* ad_add;               /* [N][n] the additions as world slots, in the dictionaries' order (NULL: counts) */
    int8_t* ad_del;               /* [N][n] the deletions */
    int64_t* narrow;              /* [n] written: the bubble whose narrowcast the row's pilot gets, -1 none */
} xb_server_state;

/* One tick's queue: J calls, each for every world ([J][W], call j of world w at [j * W + w]; f [J][W][4]). kind[j] is
   the call's action (the reference's KINDS numbering, src/exact_ballpark/vec_server.py:120-125); mask bit 0: pushed
   for the world, bit 1: already tested by an undo (bit 0 is then its queued flag). Replaces `struct Queue`
   (vec_server_cuda.cu:70-80). */
typedef struct xb_server_queue {
    int64_t J;
    int64_t J_user;               /* with a partition, the pushed calls: slots J_user .. J - 1 (kind Stop) take the
                                     follower clean-up's Stops, which the kernel writes (ball, mask); else J */
    const int8_t* kind;           /* [J] */
    uint8_t* mask;                /* [J][W] */
    int64_t* ball;                /* [J][W] Destiny ids */
    const int64_t* target;        /* [J][W] FollowBall's and Orbit's target ids */
    const double* f;              /* [J][W][4] the float arguments */
    const int64_t* n;             /* [J][W] the int argument */
    const int64_t* loc_in;        /* [J][W] the local argument (SetBallPosition's is_local_ball) */
    const uint8_t* on;            /* [W] the worlds this pre-tick runs for (OnTick's mask); NULL: every world */
    uint8_t* ok;                  /* [J][W] written: queued (pushed and not a duplicate of the world's last) */
    int64_t* row;                 /* [J][W] written: the ball's park row, -1 for none */
    int64_t* trow;                /* [J][W] written: the target's park row, -1 for none */
    int64_t* loc;                 /* [J][W] written: the local argument as queued (CloakBall's bubble, 0 or -1) */
    int8_t* route;                /* [J][W] written: 0 sent nowhere, 1 to the bubble, 2 to the ball's character */
    uint8_t* notice;              /* [J][n] written for UncloakBall calls: the rows whose pilot gets AddBallsToPark */
    int64_t* rec_add;             /* [N][W][XB_SREC_WORDS] with snapshot: a world's balls at its bubble update */
    int64_t* rec_unc;             /* [J][W][XB_SREC_WORDS] with snapshot: an uncloaked ball as its UncloakBall left it */
    int64_t* rbubble;             /* [J][W] written with a partition: the bubble a call routed to the bubble goes to */
} xb_server_queue;

#define XB_SREC_WORDS 34          /* a stream record (exact-ballpark_amd/csrc/xb_server.h numbers its words) */

/* The layout as built, for the loader's check: out[0] = sizeof(xb_server_state), out[1] its `narrow` offset, out[2] =
   sizeof(xb_server_queue), out[3] its `rbubble` offset, out[4] = XB_SREC_WORDS. Plays the part of the reference's ABI
   check at load (src/exact_ballpark/vec_server_cuda.py's xbpv_abi). */
int xb_server_layout(int64_t* out);

/* The pre-tick of every world whose `on` byte is set, one lane a world: phase 0 the stamps, the queue in order (the
   duplicate test against the world's last queued call, each ball's and target's row, every call but CloakBall applied
   to the park by the park's own commands and setters), the bubble update, the history (CloakBall applied there; each
   call's route; UncloakBall's notices) and the send flags; phase 1 the send flags alone (after the park's tick, which
   can change a ball's interactivity); phase 2 the queue's test alone (each call's queued flag, rows and local
   argument, as phase 0 would find them; nothing else written: the pending queries' view of the queue). Worlds whose
   `on` byte is clear are not touched and their flags are cleared (phases 0 and 1). With the partition (srv->space),
   phase 0 also writes the follower clean-up's Stops into slots J_user .. J - 1, applies each call's part in the
   partition, runs AdditionsAndDeletions across old and new bubbles (ad_*), writes each BUBBLE route's bubble (rbubble)
   and routes by bubble; the flags give each pilot its ship's bubble's narrowcast (narrow). The partition's tables are
   read and written as xb_space_tape's are, on the same stream.
   Replaces xbpv_pre_tick (src/exact_ballpark/vec_server_cuda.cu:615) and xbpq_pre_tick
   (vec_server_space_cuda.cu:399). */
int xb_server_pre_tick(const xb_modes_state* park, const xb_server_state* srv, const xb_server_queue* q, int phase,
                       void* stream);

/* ---- the learned scheduler (no counterpart in the reference: it orders work, never changes a result) ---- */

#define XB_MLP_IN 12
#define XB_MLP_HIDDEN 64
#define XB_MLP_OUT 8
#define XB_MLP_PARAMS (XB_MLP_IN * XB_MLP_HIDDEN + XB_MLP_HIDDEN + XB_MLP_HIDDEN * XB_MLP_HIDDEN + XB_MLP_HIDDEN + \
                       XB_MLP_HIDDEN * XB_MLP_OUT + XB_MLP_OUT)
#define XB_MLP_CHUNKS 64 /* the backward pass sums each gradient over this many runs of samples, then the runs in order */

/* Each world's 12 features (float32 [n_worlds][12]): its mode histogram (stop and inert, goto, follow, orbit, warp; as
   fractions of its ships), ships / 32, aligning warpers, idle rows, log10(1 + the nearest ship's distance to the
   centroid) / 10, log10(1 + the mean distance) / 10, the divergence cost of its last tick (the summed weights of the
   distinct path classes its rows executed) / 38 and its rows' summed path weights per ship / 9.
   World w is the rows world_start[w] .. world_start[w] + world_count[w] - 1; `path` may be NULL. */
int xb_world_features(const xb_modes_state* state, const int32_t* world_start, const int32_t* world_count,
                      int32_t n_worlds, const uint8_t* path, float* features, void* stream);

/* The network 12 -> 64 -> 64 -> 8 (ReLU, ReLU, logits), float32. params: W1[12][64], b1[64], W2[64][64], b2[64],
   W3[64][8], b3[8] (XB_MLP_PARAMS). h1, h2 ([n][64]) keep the hidden activations for the backward pass (NULL:
   not kept). */
int xb_mlp_forward(const float* params, const float* x, int32_t n, float* h1, float* h2, float* logits, void* stream);

/* The mean softmax cross-entropy over n samples against labels (0..7): per-sample loss[n] (each over n), the
   per-sample deltas dl[n][8], d2[n][64], d1[n][64] (scratch), and grads[XB_MLP_PARAMS]. Each gradient is summed over
   XB_MLP_CHUNKS consecutive runs of samples in sample order into work[XB_MLP_CHUNKS][XB_MLP_PARAMS] (scratch), then
   over the runs in order: the result depends on n alone, never on timing. */
int xb_mlp_backward(const float* params, const float* x, const float* h1, const float* h2, const float* logits,
                    const int32_t* labels, int32_t n, float* dl, float* d2, float* d1, float* loss, float* work,
                    float* grads, void* stream);

/* One Adam step (beta 0.9, 0.999, eps 1e-8) of learning rate lr at step >= 1, moments m and v in place. */
int xb_mlp_adam(float* params, const float* grads, float* m, float* v, int32_t step, float lr, void* stream);

/* ---- the rollout-serving loop (no counterpart in the reference: episodes admitted into a resident park) ---- */

/* Per row of the resident park (capacity rows each): its episode's seed, the row's ship index in the episode, the
   episode's ship count and first row, the serving step the episode was admitted at, and the episode tick of the row's
   next command (updated by xb_serve_policy). */
typedef struct xb_episode_rows {
    const uint64_t* seed;
    const int32_t* ship;
    const int32_t* size;
    const int64_t* base;
    const int64_t* admit;
    int64_t* next_cmd;
} xb_episode_rows;

/* The seeded random policy for serving step `step`, for the rows live[0 .. n_live - 1]: a row whose episode tick
   (step - admit + 1) is its next command tick gets op[row] = the command (3 stop, 1 goto, 4 follow, 5 orbit, 6 warp)
   with point[row][3], target[row] (a row of another ship of its episode) and range[row], and its next command tick
   moves on by 1 + (a draw mod 10); every other listed row gets op[row] = 0. Each draw is the counter-based hash of
   exact-ballpark_amd/csrc/xb_hash.h on (episode seed, field, ship, episode tick). */
int xb_serve_policy(const xb_episode_rows* rows, const int32_t* live, int64_t n_live, int64_t step, uint8_t* op,
                    double* point, int64_t* target, double* range, void* stream);

/* ---- glibc's rand() per world (destiny's floatrand for shuffled sensors) ---- */

/* srand(seeds[w]) for each world: states[w][33] int32 = glibc's TYPE_3 table (31 words), then its front and rear
   indices. Replaces the reference's seeding from its own generator (src/exact_ballpark/vec_space.py:148-158). */
int xb_rand_seed(const uint32_t* seeds, int64_t n_worlds, int32_t* states, void* stream);

/* `draws` rand() values from each world's generator, out[w][draws], the states advanced in place. Replaces the space
   kernel's glibc_rand (src/exact_ballpark/glibc_rand.cuh:9-26). */
int xb_rand_draw(int32_t* states, int64_t n_worlds, int32_t draws, int32_t* out, void* stream);

/* ---- libxb_cpu.so: the host build of the tick and the host runtime's helpers ---- */

/* The ABI version of the host library (XB_ABI_VERSION). */
int xb_cpu_abi(void);

/* xb_modes_tick on a park in host memory (OpenMP over the tiles; threads <= 0: OpenMP's default). The CPU batched
   path, in place of the reference's torch tick on the CPU (src/exact_ballpark/vec_modes.py:943-960). */
int xb_cpu_modes_tick(const xb_modes_state* state, int threads);

/* xb_modes_orient_setup on a park in host memory (no stream). */
int xb_cpu_modes_orient_setup(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, int what);

/* xb_collide_file on a park in host memory (no stream). */
int xb_cpu_collide_file(const xb_modes_state* state, const int32_t* rows, int64_t n_rows);

/* The static collidables' table from their arguments, once at load: n of them, kind[i] 0 a capsule (AddCapsule:
   geom[12 i ..] = A (3), B (3), the radius, rounded to a float here) or 1 an oriented box (AddOrientedBox: the corner,
   the edge vectors lx, ly, lz), id[i], parent[i]; written into st_f64 (XB_NSF rows), st_i64 (XB_NSI rows) and the
   keys of the first max_boxes boxes of each into st_keys (3 max_boxes rows), st_stride columns (>= n). Returns the
   most boxes any of them is filed in (call again with more rows if above max_boxes), -1 for bad arguments, -2 for a
   capsule whose filing never ends (twice its radius below half a float ulp of its length: Destiny's result undefined).
   Replaces the reference's `static_record` (src/exact_ballpark/vec_collide.py:261-330) and AddCapsule /
   AddOrientedBox's InsertInBoxes (full.py:998-1100). */
int xb_cpu_statics_setup(int64_t n, const int64_t* kind, const int64_t* id, const int64_t* parent,
                         const double* geom, double* st_f64, int64_t* st_i64, int64_t* st_keys, int64_t st_stride,
                         int32_t max_boxes);

/* xb_modes_ypr_snap on a park in host memory (no stream). */
int xb_cpu_modes_ypr_snap(const xb_modes_state* state, const int32_t* rows, int64_t n_rows);

/* xb_modes_command_check and xb_modes_command_apply on a park in host memory (no stream; status in host memory). */
int xb_cpu_modes_command_check(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, xb_command* cmds,
                               int32_t* status);
int xb_cpu_modes_command_apply(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, xb_command* cmds,
                               uint8_t* mark);

/* xb_modes_commands on a park in host memory (no stream). */
int xb_cpu_modes_commands(const xb_modes_state* state, const int32_t* rows, int64_t n_rows, const uint8_t* op,
                          const double* point, const int64_t* target, const double* range, const double* min_range,
                          const int64_t* warp_factor);

/* xb_space_tape and xb_space_query on a park and layer in host memory (no stream). */
int xb_cpu_space_tape(const xb_modes_state* park, const xb_space_state* space, int64_t K, const int64_t* code,
                      const int64_t* a, const int64_t* b, const double* f);
int xb_cpu_space_query(const xb_modes_state* park, const xb_space_state* space, int64_t kind, const int64_t* a,
                       const int64_t* b, const double* f, int64_t* out, double* outf);

/* xb_server_layout and xb_server_pre_tick on a park and server layer in host memory (no stream). */
int xb_cpu_server_layout(int64_t* out);
int xb_cpu_server_pre_tick(const xb_modes_state* park, const xb_server_state* srv, const xb_server_queue* q,
                           int phase);

/* xb_rand_seed and xb_rand_draw on the host (no stream). */
int xb_cpu_rand_seed(const uint32_t* seeds, int64_t n_worlds, int32_t* states);
int xb_cpu_rand_draw(int32_t* states, int64_t n_worlds, int32_t draws, int32_t* out);

/* xb_maths_eval's functions on the host (same codes, no stream). */
int xb_cpu_maths_eval(int func, int64_t n, const double* x, double* out);

/* Greedy packing of worlds in the given order into tiles of at most `block` rows (a world never split):
   tiles[0] = 0, tiles[k + 1] = the first row after tile k; *n_tiles = the tile count. tiles must hold n + 1 entries.
   -1 if a world is wider than the block. */
int xb_cpu_pack_tiles(const int32_t* sizes, int64_t n, int32_t block, int32_t* tiles, int64_t* n_tiles);

#ifdef __cplusplus
}
#endif

#endif
