#!/usr/bin/env python3
"""Writes tests/golden/orient_golden.npz: the reference's own scalar ballpark with dynamical orientation
(src/exact_ballpark/full.py, `FullBallpark(dynamical_orientation=True)`, a client's view) run on the seeded turning
worlds of tests/scenarios.py (`do_cases`), every ball's state after every tick, as bits. Rotations given as yaw, pitch and roll go through the
reference's RotationQuaternion(yaw, pitch, roll) (full.py:268-280, DirectXMath's scalar path) into SetBallRotation.

Run only where the reference is mounted (/root/reference); the fixture it writes is committed and nothing else reads
the reference. The reference takes the C library's float functions through ctypes (full.py:121-133) and its double
exp, log, sincos and pow from Python's math (the C library); here each goes to the reference's own transcription of
glibc 2.34 aarch64 (vec_libm.expf, logf, acosf, sinf, cosf, atan2f, pow; vec_exp.exp; dlibm.log, dlibm.sincos),
the functions its kernels are held to, so the vectors do not depend on this machine's C library. Python 3.10 has no
math.fma, which the transcriptions take at import: an exact rational a b + c rounded once stands in for it. The balls
are not massive, so the collision gradient has nothing to do; the worlds' orientation, modes and setters are the
subject.
"""

from __future__ import annotations

import json
import math
import sys
import types
from fractions import Fraction
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, "/root/reference/src")

import scenarios as S  # noqa: E402


def _exact_fma(a: float, b: float, c: float) -> float:
    if not (math.isfinite(a) and math.isfinite(b) and math.isfinite(c)):
        return a * b + c
    r = Fraction(a) * Fraction(b) + Fraction(c)
    if r == 0:
        return (a * b) + c if (a * b == 0.0 and c == 0.0) else 0.0
    return float(r)


math.fma = _exact_fma                                       # before the transcriptions are imported
import torch  # noqa: E402
import exact_ballpark.dlibm as dlibm  # noqa: E402
import exact_ballpark.full as full  # noqa: E402
import exact_ballpark.modes as modes  # noqa: E402
import exact_ballpark.ref as ref  # noqa: E402
import exact_ballpark.space as space  # noqa: E402
import exact_ballpark.vec_exp as vec_exp  # noqa: E402
import exact_ballpark.vec_libm as vec_libm  # noqa: E402

F64 = torch.float64


def _t(*xs):
    return [torch.tensor([float(x)], dtype=F64) for x in xs]


def _f1(fn):
    return lambda x: float(fn(vec_libm.f32(*_t(x))).item())


def _glibc_exp(x: float) -> float:
    return float(vec_exp.exp(*_t(x)).item())


def _glibc_pow(x: float, y: float) -> float:
    return float(vec_libm.pow(*_t(x, y)).item())


def _refused(*_a):
    raise RuntimeError("a transcendental the turning worlds do not reach")


_maths = types.SimpleNamespace(**{k: getattr(math, k) for k in dir(math) if not k.startswith("_")})
_maths.exp = _glibc_exp
_maths.log = dlibm.log
_maths.pow = _glibc_pow
_maths.sin = _maths.asin = _maths.atan2 = _refused
for mod in (ref, modes, space, full):
    mod.math = _maths
modes._sincos = dlibm.sincos
full._expf, full._logf, full._acosf = _f1(vec_libm.expf), _f1(vec_libm.logf), _f1(vec_libm.acosf)
full._sinf, full._cosf = _f1(vec_libm.sinf), _f1(vec_libm.cosf)
full._atan2f = lambda y, x: float(vec_libm.atan2f(vec_libm.f32(*_t(y)), vec_libm.f32(*_t(x))).item())
full._asinf = full._powf = _refused

DOUBLES = ("px", "py", "pz", "vx", "vy", "vz", "gx", "gy", "gz", "ax", "ay", "az", "opx", "opy", "opz", "ovx", "ovy",
           "ovz", "mass", "radius", "max_vel", "agility", "sf", "follow_range", "time_factor", "last_collision",
           "rx", "ry", "rz", "rw", "orx", "ory", "orz", "orw", "wx", "wy", "wz", "owx", "owy", "owz", "tx", "ty", "tz",
           "rot_roll", "max_ang_vel", "rot_agility")
INTS = ("id", "mode", "effect_stamp", "follow_id", "owner_id", "follow_ptr", "massive", "sensor")


def _apply(park, c) -> None:
    _, op, a, b, x, y, z, rr, w = c
    if op == S.GOTO_OP:
        park.goto_point(a, x, y, z)
    elif op == S.DIR_OP:
        park.goto_direction(a, x, y, z)
    elif op == S.STOP_OP:
        park.stop(a)
    elif op == S.FOLLOW_OP:
        park.follow_ball(a, b, rr)
    elif op == S.ORBIT_OP:
        park.orbit(a, b, rr)
    elif op == S.WARP_OP:
        park.warp_to(a, x, y, z, rr, w)
    elif op == S.WARPIN_OP:
        park.entity_warp_in(a, x, y, z, w)
    elif op == S.VEL_OP:
        park.set_velocity(a, x, y, z)
    elif op == S.ROT_OP:
        park.set_ball_rotation(a, x, y, z, rr)
    elif op == S.ANGVEL_OP:
        park.set_ball_angular_velocity(a, x, y, z)
    elif op == S.MAXANG_OP:
        park.set_max_angular_speed(a, rr)
    elif op == S.ROTAG_OP:
        park.set_ball_angular_agility(a, rr)
    elif op == S.YPR_OP:                              # RotationQuaternion(yaw, pitch, roll) of the floats
        park.set_ball_rotation(a, *full.q_rotation_ypr(full._f(x), full._f(y), full._f(z)))
    else:
        raise ValueError(op)


def _state(b) -> tuple:
    d = [*b.pos, *b.vel, *b.goto, *b.last_g, *b.old_pos, *b.old_vel, b.mass, b.radius, b.max_vel, b.agility,
         b.speed_fraction, b.follow_range, b.time_factor, b.last_collision, *b.rot, *b.old_rot, *b.ang_vel,
         *b.old_ang_vel, *b.torque, b.rot_roll, b.max_ang_vel, b.rot_agility]
    q = [b.id, b.mode, b.effect_stamp, b.follow_id, b.owner_id, -1 if b.follow_ptr is None else b.follow_ptr,
         int(b.is_massive), int(b.sensor_active)]
    return [float(x) for x in d], [int(x) for x in q]


def run_world(w: dict) -> list:
    """[(step, ball index, doubles, ints, followers)] at step 0 (the balls as added) and after every tick."""
    park = full.FullBallpark(friction=w["friction"], tick_interval_ms=w["tick"], use_new_orbit=w["new_orbit"],
                             dynamical_orientation=True)
    for b in w["balls"]:
        park.add_ball(b["id"], b["mass"], b["radius"], b["max_vel"], is_free=b["free"], is_massive=b["massive"],
                      is_interactive=b["interactive"], x=b["pos"][0], y=b["pos"][1], z=b["pos"][2], vx=b["vel"][0],
                      vy=b["vel"][1], vz=b["vel"][2], agility=b["agility"], speed_fraction=b["sf"],
                      max_ang_vel=b["max_ang_vel"], rot_agility=b["rot_agility"])
    order = sorted(b["id"] for b in w["balls"])
    out = []
    for k, bid in enumerate(order):
        d, q = _state(park.balls[bid])
        out.append((0, k, d, q, sorted(park.balls[bid].followers)))
    for step in range(1, w["steps"] + 1):
        for c in w["cmds"]:
            if c[0] == step:
                _apply(park, c)
        park.evolve()
        for k, bid in enumerate(order):
            d, q = _state(park.balls[bid])
            out.append((step, k, d, q, sorted(park.balls[bid].followers)))
    return out


def main(seed: int = 1, scale: int = 1, name: str = "orient_golden.npz") -> None:
    worlds = S.do_cases(seed, scale)
    kept, skipped = [], []
    rec_world, rec_step, rec_ball, rec_d, rec_q, fol_off, fol_ids = [], [], [], [], [], [0], []
    for wi, w in enumerate(worlds):
        try:
            rows = run_world(w)
        except (ZeroDivisionError, ValueError, OverflowError) as e:
            skipped.append([wi, type(e).__name__, str(e)[:80]])
            continue
        kept.append(wi)
        for step, k, d, q, fol in rows:
            rec_world.append(wi)
            rec_step.append(step)
            rec_ball.append(k)
            rec_d.append(d)
            rec_q.append(q)
            fol_ids.extend(fol)
            fol_off.append(len(fol_ids))
    d = np.array(rec_d, dtype=np.float64).view(np.uint64)
    np.savez_compressed(HERE / name, seed=np.int64(seed), worlds=np.array(json.dumps(worlds)),
                        kept=np.array(kept, dtype=np.int64), skipped=np.array(json.dumps(skipped)),
                        world=np.array(rec_world, dtype=np.int32), step=np.array(rec_step, dtype=np.int32),
                        ball=np.array(rec_ball, dtype=np.int32), d=d, q=np.array(rec_q, dtype=np.int64),
                        fol_off=np.array(fol_off, dtype=np.int64), fol_ids=np.array(fol_ids, dtype=np.int64),
                        doubles=np.array(DOUBLES), ints=np.array(INTS))
    print(f"{name}: {len(kept)} worlds kept, {len(skipped)} skipped {skipped[:5]}, {len(rec_d)} records",
          file=sys.stderr)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 1, int(sys.argv[2]) if len(sys.argv) > 2 else 1,
         sys.argv[3] if len(sys.argv) > 3 else "orient_golden.npz")
