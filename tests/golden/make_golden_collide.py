#!/usr/bin/env python3
"""Writes tests/golden/collide_golden.npz: the reference's own scalar ballpark with Destiny's partition and the
gradient collision method (src/exact_ballpark/space.py, `SpaceBallpark`, a client's view) run on the seeded collision
worlds of tests/scenarios.py (`collide_cases`), every ball's state after every tick, as bits: the flight fields, the
collision acceleration (mLastC) and time (mLastCollision), the collision list in order, and the filing (the level and
key of the ball's box and its side on each axis). With `--do` it writes collide_do_golden.npz: the same worlds with
dynamical orientation (`do_collide_cases`) on the reference's FullBallpark(dynamical_orientation=True), the turning
state recorded too; rotations given as yaw, pitch and roll go through its RotationQuaternion into SetBallRotation.

Run only where the reference is mounted (/root/reference); the fixture it writes is committed and nothing else reads
the reference. The reference's double exp, log, sincos and pow go to its own transcriptions of glibc 2.34 aarch64
(vec_exp.exp, dlibm.log, dlibm.sincos, vec_libm.pow), the functions its kernels are held to, so the vectors do not
depend on this machine's C library; so do FullBallpark's float functions (vec_libm's expf, logf, acosf, sinf, cosf,
atan2f). Python 3.10 has no math.fma, which the transcriptions take at import: an exact
rational a b + c rounded once stands in for it.
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


def _glibc_exp(x: float) -> float:
    return float(vec_exp.exp(*_t(x)).item())


def _glibc_pow(x: float, y: float) -> float:
    return float(vec_libm.pow(*_t(x, y)).item())


def _refused(*_a):
    raise RuntimeError("a transcendental the collision worlds do not reach")


_maths = types.SimpleNamespace(**{k: getattr(math, k) for k in dir(math) if not k.startswith("_")})
_maths.exp = _glibc_exp
_maths.log = dlibm.log
_maths.pow = _glibc_pow
_maths.sin = _maths.asin = _maths.atan2 = _refused
for mod in (ref, modes, space, full):
    mod.math = _maths
modes._sincos = dlibm.sincos


def _f1(fn):
    return lambda x: float(fn(vec_libm.f32(*_t(x))).item())


full._expf, full._logf, full._acosf = _f1(vec_libm.expf), _f1(vec_libm.logf), _f1(vec_libm.acosf)
full._sinf, full._cosf = _f1(vec_libm.sinf), _f1(vec_libm.cosf)
full._atan2f = lambda y, x: float(vec_libm.atan2f(vec_libm.f32(*_t(y)), vec_libm.f32(*_t(x))).item())
full._asinf = full._powf = _refused
TURNING = ("rx", "ry", "rz", "rw", "orx", "ory", "orz", "orw", "wx", "wy", "wz", "owx", "owy", "owz", "tx", "ty", "tz",
           "rot_roll", "max_ang_vel", "rot_agility")

DOUBLES = ("px", "py", "pz", "vx", "vy", "vz", "gx", "gy", "gz", "ax", "ay", "az", "opx", "opy", "opz", "ovx", "ovy",
           "ovz", "mass", "radius", "max_vel", "agility", "sf", "follow_range", "time_factor", "last_collision",
           "cx", "cy", "cz")
INTS = ("id", "mode", "effect_stamp", "follow_id", "owner_id", "follow_ptr", "massive", "sensor", "level", "kx", "ky",
        "kz", "mx", "my", "mz")


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
    elif op == S.YPR_OP:                              # RotationQuaternion(yaw, pitch, roll) of the floats
        park.set_ball_rotation(a, *full.q_rotation_ypr(full._f(x), full._f(y), full._f(z)))
    else:
        raise ValueError(op)


def _state(b, turning: bool = False) -> tuple:
    d = [*b.pos, *b.vel, *b.goto, *b.last_g, *b.old_pos, *b.old_vel, b.mass, b.radius, b.max_vel, b.agility,
         b.speed_fraction, b.follow_range, b.time_factor, b.last_collision, *b.last_c]
    if turning:
        d += [*b.rot, *b.old_rot, *b.ang_vel, *b.old_ang_vel, *b.torque, b.rot_roll, b.max_ang_vel, b.rot_agility]
    box = b.my_box
    q = [b.id, b.mode, b.effect_stamp, b.follow_id, b.owner_id, -1 if b.follow_ptr is None else b.follow_ptr,
         int(b.is_massive), int(b.sensor_active), box.level, *box.key, *b.mod]
    return [float(x) for x in d], [int(x) for x in q]


def run_world(w: dict) -> list:
    """[(step, ball index, doubles, ints, collision list)] at step 0 (the balls as added) and after every tick."""
    do = bool(w.get("dynamical"))
    if do:
        park = full.FullBallpark(friction=w["friction"], tick_interval_ms=w["tick"], use_new_orbit=w["new_orbit"],
                                 dynamical_orientation=True)
    else:
        park = space.SpaceBallpark(friction=w["friction"], tick_interval_ms=w["tick"], use_new_orbit=w["new_orbit"])
    park.current_time = w["start"]
    for b in w["balls"]:
        kw = {"max_ang_vel": b["max_ang_vel"], "rot_agility": b["rot_agility"]} if do else {}
        park.add_ball(b["id"], b["mass"], b["radius"], b["max_vel"], is_free=b["free"], is_massive=b["massive"],
                      is_interactive=b["interactive"], is_space_junk=b["junk"], x=b["pos"][0], y=b["pos"][1],
                      z=b["pos"][2], vx=b["vel"][0], vy=b["vel"][1], vz=b["vel"][2], agility=b["agility"],
                      speed_fraction=b["sf"], **kw)
        if b["cloaked"]:
            park.set_cloaked(b["id"], b["cloaked"])
    order = sorted(b["id"] for b in w["balls"])
    out = []
    for k, bid in enumerate(order):
        d, q = _state(park.balls[bid], do)
        out.append((0, k, d, q, list(park.balls[bid].collisions)))
    for step in range(1, w["steps"] + 1):
        for c in w["cmds"]:
            if c[0] == step:
                _apply(park, c)
        park.evolve()
        for k, bid in enumerate(order):
            d, q = _state(park.balls[bid], do)
            out.append((step, k, d, q, list(park.balls[bid].collisions)))
    return out


def main(seed: int = 1, scale: int = 1, name: str = "collide_golden.npz", do: bool = False) -> None:
    worlds = S.do_collide_cases(seed, scale) if do else S.collide_cases(seed, scale)
    kept, skipped = [], []
    rec_world, rec_step, rec_ball, rec_d, rec_q, hit_off, hit_ids = [], [], [], [], [], [0], []
    for wi, w in enumerate(worlds):
        try:
            rows = run_world(w)
        except (ZeroDivisionError, ValueError, OverflowError, space.DestinyUndefined, NotImplementedError) as e:
            skipped.append([wi, type(e).__name__, str(e)[:80]])
            continue
        kept.append(wi)
        for step, k, d, q, hits in rows:
            rec_world.append(wi)
            rec_step.append(step)
            rec_ball.append(k)
            rec_d.append(d)
            rec_q.append(q)
            hit_ids.extend(hits)
            hit_off.append(len(hit_ids))
    d = np.array(rec_d, dtype=np.float64).view(np.uint64)
    np.savez_compressed(HERE / name, seed=np.int64(seed), worlds=np.array(json.dumps(worlds)),
                        kept=np.array(kept, dtype=np.int64), skipped=np.array(json.dumps(skipped)),
                        world=np.array(rec_world, dtype=np.int32), step=np.array(rec_step, dtype=np.int32),
                        ball=np.array(rec_ball, dtype=np.int32), d=d, q=np.array(rec_q, dtype=np.int64),
                        hit_off=np.array(hit_off, dtype=np.int64), hit_ids=np.array(hit_ids, dtype=np.int64),
                        doubles=np.array(DOUBLES + (TURNING if do else ())), ints=np.array(INTS))
    n_hits = sum(1 for i in range(1, len(hit_off)) if hit_off[i] > hit_off[i - 1])
    print(f"{name}: {len(kept)} worlds kept, {len(skipped)} skipped {skipped[:5]}, {len(rec_d)} records, "
          f"{n_hits} with collisions", file=sys.stderr)


if __name__ == "__main__":
    args = [x for x in sys.argv[1:] if x != "--do"]
    do = "--do" in sys.argv[1:]
    main(int(args[0]) if len(args) > 0 else 1, int(args[1]) if len(args) > 1 else 1,
         args[2] if len(args) > 2 else ("collide_do_golden.npz" if do else "collide_golden.npz"), do)
