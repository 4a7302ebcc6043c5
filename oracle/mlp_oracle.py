"""THE ORACLE for the learned scheduler's network -- test infrastructure only (tests/). A float64 numpy restatement of
the 12 -> 64 -> 64 -> 8 multilayer perceptron the GPU kernels compute (exact-ballpark_amd/csrc/xb_sched.hip): forward
pass (ReLU, ReLU, logits), mean softmax cross-entropy, its gradient with respect to every parameter, and one Adam step.

What pins it: tests/test_mlp.py holds its gradient to central finite differences of its own loss (float64, a
mathematical check independent of the formulas below), and the GPU's float32 results to it within stated tolerances.
The network has no counterpart in the reference.
"""

from __future__ import annotations

import numpy as np

NI, NH, NO = 12, 64, 8
SIZES = [("W1", (NI, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NO)), ("b3", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p: np.ndarray) -> dict:
    out, k = {}, 0
    for name, shape in SIZES:
        m = int(np.prod(shape))
        out[name] = p[k:k + m].reshape(shape)
        k += m
    return out


def forward(p: np.ndarray, x: np.ndarray) -> tuple:
    q = unpack(p.astype(np.float64))
    x = x.astype(np.float64)
    h1 = np.maximum(x @ q["W1"] + q["b1"], 0.0)
    h2 = np.maximum(h1 @ q["W2"] + q["b2"], 0.0)
    return h1, h2, h2 @ q["W3"] + q["b3"]


def loss(p: np.ndarray, x: np.ndarray, y: np.ndarray) -> float:
    _, _, z = forward(p, x)
    z = z - z.max(axis=1, keepdims=True)
    lse = np.log(np.exp(z).sum(axis=1))
    return float((lse - z[np.arange(len(y)), y]).mean())


def grad(p: np.ndarray, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    q = unpack(p.astype(np.float64))
    x = x.astype(np.float64)
    h1, h2, z = forward(p, x)
    e = np.exp(z - z.max(axis=1, keepdims=True))
    sm = e / e.sum(axis=1, keepdims=True)
    n = len(y)
    dz = sm.copy()
    dz[np.arange(n), y] -= 1.0
    dz /= n
    d2 = (dz @ q["W3"].T) * (h2 > 0)
    d1 = (d2 @ q["W2"].T) * (h1 > 0)
    g = {"W1": x.T @ d1, "b1": d1.sum(0), "W2": h1.T @ d2, "b2": d2.sum(0), "W3": h2.T @ dz, "b3": dz.sum(0)}
    return np.concatenate([g[name].reshape(-1) for name, _ in SIZES])


def adam(p: np.ndarray, g: np.ndarray, m: np.ndarray, v: np.ndarray, step: int, lr: float) -> tuple:
    b1, b2, eps = 0.9, 0.999, 1e-8
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


PATH_COST = (0.0, 1.0, 2.0, 3.0, 8.0, 9.0, 7.0, 8.0)
GOTO, FOLLOW, WARP, ORBIT = 0, 1, 3, 4


def features(pos, mode, effect_stamp, free, moribund, path, start, count) -> np.ndarray:
    """Each world's 12 features, as xb_world_features defines them (float64 here, float32 there): mode histogram
    (stop and inert, goto, follow, orbit, warp) over the world's ships, ships / 32, aligning warpers and idle rows as
    fractions, log10(1 + nearest and mean distance to the world's centroid) / 10, the last tick's divergence cost (the
    summed weights of the distinct path classes) / 38 and its summed path weights per ship / 9."""
    W = len(start)
    out = np.zeros((W, NI), dtype=np.float64)
    for w in range(W):
        lo, c = int(start[w]), int(count[w])
        if c == 0:
            continue
        r = slice(lo, lo + c)
        m = mode[r]
        h = np.select([m == GOTO, m == FOLLOW, m == ORBIT, m == WARP], [1, 2, 3, 4], 0)
        out[w, :5] = np.bincount(h, minlength=5) / c
        out[w, 5] = c / 32.0
        out[w, 6] = np.sum((m == WARP) & (effect_stamp[r] < 0)) / c
        out[w, 7] = np.sum(~(free[r].astype(bool) & ~moribund[r].astype(bool))) / c
        d = np.sqrt(((pos[r] - pos[r].mean(axis=0)) ** 2).sum(axis=1))
        out[w, 8] = np.log10(1.0 + d.min()) / 10.0
        out[w, 9] = np.log10(1.0 + d.mean()) / 10.0
        pc = path[r].astype(np.int64) & 7 if path is not None else np.zeros(0, dtype=np.int64)
        out[w, 10] = float(np.sum(np.asarray(PATH_COST)[np.unique(pc)])) / 38.0
        out[w, 11] = float(np.sum(np.asarray(PATH_COST)[pc])) / c / 9.0
    return out
