"""The oracle's old position and velocity (mOldPos, mOldVel), which the reference's ModeBallpark does not keep and so
the golden vectors do not hold, pinned by the closed forms FullBallpark gives them (src/exact_ballpark/full.py):

- AddBall: old position = position - velocity x dt, old velocity = velocity (full.py:1247; vec_modes.py:248);
- each tick's commit: a ball that ticked holds, as the old ones, its position and velocity after its mode's behaviour
  (full.py:2531, 2556-2557): the start-of-tick position; the start-of-tick velocity for GOTO, FOLLOW and ORBIT (their
  behaviours only thrust); for STOP the velocity with EvolveStop's damping of y, (vy - 0.07 vy) x 0.93457943925...
  (ref.py:203, modes.py:374); a ball that did not tick (not free, or moribund) keeps both.
Warping balls and balls whose mode changed in the tick are left out of the velocity check (a release, a drop-out from
the drop-out point, TRV-714, modes.py:724-732). Checked bit for bit on every golden scenario world, at every tick.
"""

import numpy as np

import oracle as O
import scenarios as S

P, V, OP, OV = slice(0, 3), slice(3, 6), slice(12, 15), slice(15, 18)
MODE, FREE, MORIBUND = 1, 9, 10
GOTO, FOLLOW, STOP, WARP, ORBIT = 0, 1, 2, 3, 4
STOP_DAMP = 0.9345794392523364485981308411215


def _bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)


def test_addball_and_commit_closed_forms():
    checked = 0
    for w in S.cases(1):
        o = O.World(w["friction"], w["tick"], w["new_orbit"], w["start"])
        for b in w["balls"]:
            o.add(b)
        d, q, _ = o.state()
        dt = w["tick"] * 0.001
        assert (_bits(d[:, OP]) == _bits(d[:, P] - d[:, V] * dt)).all()
        assert (_bits(d[:, OV]) == _bits(d[:, V])).all()
        for step in range(1, w["steps"] + 1):
            for c in w["cmds"]:
                if c[0] == step:
                    o.command(c)
            d0, q0, _ = o.state()
            o.evolve()
            d1, q1, _ = o.state()
            ticked = (q0[:, FREE] != 0) & (q0[:, MORIBUND] == 0)
            plain = ticked & (q0[:, MODE] != WARP)
            assert (_bits(d1[plain][:, OP]) == _bits(d0[plain][:, P])).all(), (w["kind"], step)
            kept = ticked & (q0[:, MODE] == q1[:, MODE])
            thrust = kept & np.isin(q0[:, MODE], (GOTO, FOLLOW, ORBIT))
            assert (_bits(d1[thrust][:, OV]) == _bits(d0[thrust][:, V])).all(), (w["kind"], step)
            stop = kept & (q0[:, MODE] == STOP)
            v0 = d0[stop][:, V]
            damped = np.stack([v0[:, 0], (v0[:, 1] - 0.07 * v0[:, 1]) * STOP_DAMP, v0[:, 2]], axis=1)
            assert (_bits(d1[stop][:, OV]) == _bits(damped)).all(), (w["kind"], step)
            checked += int(thrust.sum()) + int(stop.sum())
            still = ~ticked
            assert (_bits(d1[still][:, OP]) == _bits(d0[still][:, OP])).all()
            assert (_bits(d1[still][:, OV]) == _bits(d0[still][:, OV])).all()
    assert checked > 1000
