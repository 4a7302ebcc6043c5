"""Pins of the double asin and atan2 (exact-ballpark_amd/csrc/xb_maths_double.h) and of powf (xb_maths_float.h), the
maths of the roll's yaw delta and display yaw-pitch-roll (Ball.cpp's asin(sin(yaw - oldYaw)) and atan2 of a heading)
and of the mushroom growth (Ballpark.cpp:569, range * powf(timeFraction, 0.25f)), held against this machine's C
library bit for bit (glibc 2.35's x86-64 FMA builds of the same IBM and Arm routines the reference restates at
src/exact_ballpark/dlibm_inv.py and dlibm_powf.py): on arguments over every branch, each bound from both sides, the
special values and seeded samples. The branch of each argument is computed here from the routines' documented
thresholds (the exponent words glibc compares), not from the formulas, so that a sample set that missed a branch
fails. The seeded sweeps of tests/maths_float_sweep.cpp cover more arguments; the device is held to the host build in
tests/test_modes_gpu.py."""

import ctypes
import math
import random
import struct

import numpy as np
import pytest
import torch

from exact_ballpark_amd import _native as N

_C = ctypes.CDLL("libm.so.6")
_C.asin.restype = _C.atan2.restype = ctypes.c_double
_C.asin.argtypes = [ctypes.c_double]
_C.atan2.argtypes = [ctypes.c_double, ctypes.c_double]
_C.powf.restype = ctypes.c_float
_C.powf.argtypes = [ctypes.c_float, ctypes.c_float]

# asin's bounds on the high word of |x|: 2^-26, 0.125, 0.25, 0.5, 0.75, 0.921875, 0.953125, 0.96875, 1, inf
ASIN_BOUNDS = (0x3E500000, 0x3FC00000, 0x3FD00000, 0x3FE00000, 0x3FE80000, 0x3FED8000, 0x3FEE8000, 0x3FEF0000,
               0x3FF00000, 0x7FF00000)
ASIN_NAMES = ("tiny", "taylor", "q1", "q2", "h", "a", "b", "c", "near_one")


def d64(u: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", u & ((1 << 64) - 1)))[0]


def bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def same(a, b) -> np.ndarray:
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return (a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))


def asin_arguments(seed: int, n: int) -> list:
    r = random.Random(seed)
    out = [0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 1.0 + 2 ** -52, -1.0 - 2 ** -52, 1 - 2 ** -53, -1 + 2 ** -53, 2.0, -3.5,
           5e-324, -5e-324, 2.2250738585072014e-308, 1e300, math.inf, -math.inf, math.nan]
    for k in ASIN_BOUNDS:
        for low in (0, 1, 0xFFFFFFFF):
            for hi in (k, k - 1):
                out += [d64((hi << 32) | low), -d64((hi << 32) | low)]
    for j in range(n):
        kind = j % 6
        if kind == 0:
            x = d64(r.getrandbits(64))
        elif kind == 1:
            x = r.uniform(-1.0, 1.0)
        elif kind == 2:
            x = r.choice([-1, 1]) * (1.0 - 10 ** r.uniform(-17, -1.5))
        elif kind == 3:
            x = r.choice([-1, 1]) * 10 ** r.uniform(-12, 0)
        elif kind == 4:
            x = r.uniform(0.9, 1.0) * r.choice([-1, 1])
        else:
            x = math.sin(r.uniform(-4.0, 4.0))
        out.append(x)
    return out


def asin_branch(x: float) -> str:
    u = bits(x)
    k, lo = (u >> 32) & 0x7FFFFFFF, u & 0xFFFFFFFF
    for name, bound in zip(ASIN_NAMES, ASIN_BOUNDS):
        if k < bound:
            return name
    if k == 0x3FF00000 and lo == 0:
        return "one"
    if k > 0x7FF00000 or (k == 0x7FF00000 and lo):
        return "nan"
    return "beyond"


def atan2_arguments(seed: int, n: int) -> list:
    r = random.Random(seed)
    special = [0.0, -0.0, math.inf, -math.inf, math.nan, 1.0, -1.0, 5e-324, -5e-324, 1.5e-310, 1e300, -1e300, 0.3, -7.0]
    out = [(a, b) for a in special for b in special]
    for e in (56, 57, 58, 500, 501, 520, 600, 1000):      # the exponents' distance and the rescaling by 2^+-500
        for s in (1.0, -1.0):
            for t in (1.0, -1.0):
                out += [(s * 2.0 ** e * 1.25, t * 0.75), (s * 0.75, t * 2.0 ** e * 1.25), (s * 2.0 ** -e, t * 1.5),
                        (s * 1.5, t * 2.0 ** -e), (s * 2.0 ** (e - 1050), t * 2.0 ** -1060)]
    for j in range(n):
        kind = j % 7
        if kind == 0:
            y, x = d64(r.getrandbits(64)), d64(r.getrandbits(64))
        elif kind == 1:
            y, x = r.uniform(-1e3, 1e3), r.uniform(-1e3, 1e3)
        elif kind == 2:
            y, x = r.gauss(0, 1) * 10 ** r.uniform(-300, 300), r.gauss(0, 1) * 10 ** r.uniform(-300, 300)
        elif kind == 3:                                    # |y| near |x|
            x = r.uniform(-1, 1)
            y = x * r.uniform(-1.2, 1.2)
        elif kind == 4:                                    # the smaller over the larger below 1/16
            x = r.gauss(0, 1)
            y = x * r.uniform(-0.07, 0.07)
            if r.random() < 0.5:
                x, y = y, x
        elif kind == 5:                                    # near the table's points k/256
            x = r.choice([-1, 1]) * r.uniform(0.5, 2.0)
            y = x * (r.randint(16, 256) / 256 + r.uniform(-2 ** -10, 2 ** -10)) * r.choice([-1, 1])
            if r.random() < 0.5:
                x, y = y, x
        else:                                              # subnormal and tiny
            y, x = d64(r.randrange(1, 1 << 60)) * r.choice([-1, 1]), d64(r.randrange(1, 1 << 60)) * r.choice([-1, 1])
        out.append((y, x))
    return out


def atan2_branch(y: float, x: float) -> str:
    """The path of IBM's atan2 a pair takes: special values; the exponents' distance beyond 2^57 (+-pi/2) or below
    2^-57 (a quotient or +-pi); else the quadrant case (i) x > 0, |y| < |x|; (ii) x > 0, |y| >= |x|; (iii) x < 0,
    |x| < |y|; (iv) x < 0, |x| >= |y|, each below 1/16 (polynomial) or not (table), and whether 2^+-500 rescaled."""
    if x != x or y != y:
        return "nan"
    if y == 0.0 or x == 0.0 or math.isinf(x) or math.isinf(y):
        return "special"
    de = ((bits(y) >> 32) & 0x7FF00000) - ((bits(x) >> 32) & 0x7FF00000)
    if de >= 0x3900000:
        return "de_ep"
    if de <= -0x3900000:
        return "de_em_pos" if x > 0 else "de_em_neg"
    ax, ay, tag = abs(x), abs(y), ""
    if ax < 2.0 ** -500 or ay < 2.0 ** -500:
        ax, ay, tag = ax * 2.0 ** 500, ay * 2.0 ** 500, "+up"
    if ax > 2.0 ** 500 or ay > 2.0 ** 500:
        ax, ay, tag = ax * 2.0 ** -500, ay * 2.0 ** -500, tag + "+down"
    u = ay / ax if ay < ax else ax / ay
    case = ("i" if ay < ax else "ii") if x > 0 else ("iii" if ax < ay else "iv")
    return case + ("_small" if u < 1 / 16 else "_table") + tag


def powf_arguments(seed: int, n: int) -> tuple:
    r = random.Random(seed)
    f = lambda u: struct.unpack("<f", struct.pack("<I", u & 0xFFFFFFFF))[0]
    special = [0.0, -0.0, 1.0, -1.0, math.inf, -math.inf, math.nan, f(0x7FA00000), 0.5, 2.0, -2.0, 3.0, -3.0, 0.25,
               -0.5, f(0x00000001), f(0x807FFFFF), f(0x00400000), f(0x7F7FFFFF), f(0xFF7FFFFF), f(0x00800000),
               f(0x33000000), f(0x3F7FFFFF), f(0x3F800001), 1e10, -1e10, 127.0, 128.0, -149.0, -150.0, 2.5, -2.5]
    xs = [a for a in special for _ in special]
    ys = [b for _ in special for b in special]
    for j in range(n):
        kind = j % 6
        if kind == 0:
            x, y = f(r.getrandbits(32)), f(r.getrandbits(32))
        elif kind == 1:                                    # the mushroom growth's call
            x, y = r.uniform(0.0, 1.0), 0.25
        elif kind == 2:
            x, y = r.uniform(0.0, 100.0), r.uniform(-30.0, 30.0)
        elif kind == 3:                                    # a negative x: y integral, odd or even, or not
            x, y = -r.uniform(0.0, 10.0), float(r.randint(-40, 40)) + r.choice([0.0, 0.0, 0.5])
        elif kind == 4:                                    # near the overflow and underflow thresholds
            x, y = r.uniform(1.5, 4.0), r.choice([-1, 1]) * r.uniform(60.0, 330.0)
        else:                                              # a subnormal x
            x, y = f(r.randrange(1, 0x00800000)) * r.choice([-1, 1]), r.uniform(-2.0, 2.0)
        xs.append(x)
        ys.append(y)
    return np.array(xs, np.float32), np.array(ys, np.float32)


def test_asin_matches_c_library_on_every_branch():
    xs = asin_arguments(1, 60000)
    got = N.maths("asin", torch.tensor(xs, dtype=torch.float64)).numpy()
    want = np.array([_C.asin(v) for v in xs])
    ok = same(got, want)
    assert ok.all(), [xs[i].hex() for i in np.nonzero(~ok)[0][:5]]
    seen = {}
    for x in xs:
        b = asin_branch(x)
        seen[b] = seen.get(b, 0) + 1
    assert set(seen) == set(ASIN_NAMES) | {"one", "nan", "beyond"}, seen
    assert min(seen.values()) >= 4, seen


def test_atan2_matches_c_library_on_every_branch():
    pairs = atan2_arguments(2, 60000)
    y = torch.tensor([p[0] for p in pairs], dtype=torch.float64)
    x = torch.tensor([p[1] for p in pairs], dtype=torch.float64)
    got = N.maths("atan2", y, x).numpy()
    want = np.array([_C.atan2(a, b) for a, b in pairs])
    ok = same(got, want)
    assert ok.all(), [(pairs[i][0].hex(), pairs[i][1].hex()) for i in np.nonzero(~ok)[0][:5]]
    seen = {}
    for a, b in pairs:
        k = atan2_branch(a, b)
        seen[k] = seen.get(k, 0) + 1
    base = {k.split("+")[0] for k in seen}
    assert base >= {"nan", "special", "de_ep", "de_em_pos", "de_em_neg", "i_small", "i_table", "ii_small", "ii_table",
                    "iii_small", "iii_table", "iv_small", "iv_table"}, sorted(seen)
    assert any("+up" in k for k in seen) and any("+down" in k for k in seen), sorted(seen)
    assert min(seen[k] for k in seen if "+" not in k) >= 20, seen


def test_powf_matches_c_library_on_every_branch():
    xs, ys = powf_arguments(3, 60000)
    got = N.maths("powf", torch.from_numpy(xs.astype(np.float64)), torch.from_numpy(ys.astype(np.float64)))
    got = got.numpy().astype(np.float32)
    want = np.array([_C.powf(float(a), float(b)) for a, b in zip(xs, ys)], np.float32)
    ok = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    assert ok.all(), [(float(xs[i]), float(ys[i])) for i in np.nonzero(~ok)[0][:5]]
    # every kind of result was reached: NaN, both infinities, both zeros, subnormals, negatives, ordinary values
    assert np.isnan(got).sum() > 100 and (got == np.inf).sum() > 100 and (got == -np.inf).sum() > 10
    assert ((got == 0) & ~np.signbit(got)).sum() > 100 and ((got == 0) & np.signbit(got)).sum() > 10
    assert ((np.abs(got) < np.finfo(np.float32).tiny) & (got != 0)).sum() > 10 and (got < 0).sum() > 1000


def test_powf_known_answers():
    """Exact powers: a power of two to an integer, a perfect square's root, and the quarter powers the mushroom takes
    of 0, 1 and 16."""
    cases = [(2.0, 10.0, 1024.0), (2.0, -3.0, 0.125), (-2.0, 3.0, -8.0), (-2.0, 2.0, 4.0), (9.0, 0.5, 3.0),
             (16.0, 0.25, 2.0), (0.0, 0.25, 0.0), (1.0, 0.25, 1.0), (0.0625, 0.25, 0.5), (4.0, -0.5, 0.5)]
    x = torch.tensor([c[0] for c in cases], dtype=torch.float64)
    y = torch.tensor([c[1] for c in cases], dtype=torch.float64)
    assert N.maths("powf", x, y).tolist() == [c[2] for c in cases]


def test_asin_and_atan2_identities():
    """asin(+-1) = +-pi/2 and atan2(1, 1) = pi/4 as glibc rounds them; asin odd; atan2(y, x) = -atan2(-y, x)."""
    hp = float.fromhex("0x1.921fb54442d18p+0")
    a = N.maths("asin", torch.tensor([1.0, -1.0, 0.5], dtype=torch.float64)).tolist()
    assert a[0] == hp and a[1] == -hp and a[2] == float.fromhex("0x1.0c152382d7366p-1")
    t = N.maths("atan2", torch.tensor([1.0, -1.0], dtype=torch.float64), torch.tensor([1.0, -1.0],
                                                                                       dtype=torch.float64)).tolist()
    assert t[0] == float.fromhex("0x1.921fb54442d18p-1") and t[1] == -float.fromhex("0x1.2d97c7f3321d2p+1")
    r = np.random.default_rng(4)
    x = torch.from_numpy(r.uniform(-1, 1, 20000))
    assert (N.maths("asin", -x).numpy() == -N.maths("asin", x).numpy()).all()
    y, w = torch.from_numpy(r.normal(size=20000)), torch.from_numpy(r.normal(size=20000))
    assert (N.maths("atan2", -y, w).numpy() == -N.maths("atan2", y, w).numpy()).all()
