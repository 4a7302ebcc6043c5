/* SPDX-License-Identifier: MIT
 *
 * The park's state in device memory: structure of arrays, one block per element type, one block row per field.
 * Field f of ball row r is at block[f * stride + r]; stride is the row count rounded up to 64 so every field row starts
 * 512-byte aligned and a wavefront's 64 lanes read one field as whole cache lines. The fields are the reference's
 * per-ball tensors (src/exact_ballpark/vec_modes_cuda.py:281-290, `STATE_TENSORS`), a vector field taking one block
 * row per component; floats (radius .. follow_range, the turning values) are held in doubles as the reference holds
 * them, so the drop-in tensors keep its dtypes.
 */
#ifndef XB_LAYOUT_H
#define XB_LAYOUT_H

/* float64 block */
#define XB_F_POS 0          /* 3 rows: x, y, z */
#define XB_F_VEL 3
#define XB_F_OLD_POS 6
#define XB_F_OLD_VEL 9
#define XB_F_GOTO 12
#define XB_F_LAST_G 15
#define XB_F_MASS 18
#define XB_F_TIME_FACTOR 19
#define XB_F_LAST_COLLISION 20
#define XB_F_RADIUS 21
#define XB_F_MAX_VEL 22
#define XB_F_AGILITY 23
#define XB_F_SPEED_FRACTION 24
#define XB_F_FOLLOW_RANGE 25
#define XB_F_MAX_ANG_VEL 26
#define XB_F_ROT_AGILITY 27
#define XB_F_ROT_ROLL 28
#define XB_F_ROT 29         /* 4 rows: x, y, z, w */
#define XB_F_OLD_ROT 33
#define XB_F_ANG_VEL 37
#define XB_F_OLD_ANG_VEL 40
#define XB_F_TORQUE 43
#define XB_NF 46

/* int64 block */
#define XB_I_IDS 0
#define XB_I_MODE 1
#define XB_I_EFFECT_STAMP 2
#define XB_I_FOLLOW_ID 3
#define XB_I_OWNER_ID 4
#define XB_I_FOLLOW_PTR 5   /* the target's row in this park, -1 for none */
#define XB_I_CLOCK 6
#define XB_I_WORLD 7
#define XB_I_CLOAKED 8
#define XB_I_BUBBLE 9
#define XB_NI 10

/* uint8 block (0 or 1) */
#define XB_B_FREE 0
#define XB_B_MORIBUND 1
#define XB_B_MASSIVE 2
#define XB_B_INTERACTIVE 3
#define XB_B_SENSOR 4
#define XB_B_JUNK 5         /* space junk (collides only with space junk) */
#define XB_B_FORMATION 6    /* a leader's formation: a char (-1 none), two's complement */
#define XB_NB 7

/* the general path's scratch (caller-owned): float64 rows pos0 (3), vel0 (3), vy_after; then one uint8 row act */
#define XB_S_POS0 0
#define XB_S_VEL0 3
#define XB_S_VY_AFTER 6
#define XB_NS 7

/* DSTBALLMODE (IDstConstants.h, as the reference's modes.py:65-67 lists it) */
#define XB_MODE_GOTO 0
#define XB_MODE_FOLLOW 1
#define XB_MODE_STOP 2
#define XB_MODE_WARP 3
#define XB_MODE_ORBIT 4
#define XB_MODE_MISSILE 5
#define XB_MODE_MUSHROOM 6
#define XB_MODE_BOID 7
#define XB_MODE_TROLL 8
#define XB_MODE_MINIBALL 9
#define XB_MODE_FIELD 10
#define XB_MODE_RIGID 11
#define XB_MODE_FORMATION 12

#endif
