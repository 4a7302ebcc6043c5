"""The learned scheduler: a 12 -> 64 -> 64 -> 8 FP32 network, trained on the GPU by the library's own kernels
(xb_world_features, xb_mlp_forward, xb_mlp_backward, xb_mlp_adam in libxb_hip.so), predicts each world's cost bucket for
the coming ticks from features of its state and the path classes its last tick recorded; the park then regroups worlds
of one bucket into the same workgroups (`rowmap` and tiles; no state moves). It orders work and never changes a result:
a world's bits depend on no other world and on no workgroup, which the tests hold bit for bit against the plain order.

Labels: a world's divergence cost in a tick, the summed weights (`PATH_COST`: the double functions and divisions each
class calls) of the distinct path classes its rows executed -- what a wavefront holding the world pays, since a
wavefront runs the union of its lanes' branches -- bucketed into 8 by quantiles of the training ticks' costs. Worlds of
one bucket executed the same costly branches, so grouping them makes wavefronts that diverge little.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as N

NI, NH, NO = 12, 64, 8
NPARAM = NI * NH + NH + NH * NH + NH + NH * NO + NO
CHUNKS = 64                                  # runs of samples the backward pass sums separately (XB_MLP_CHUNKS)
PATH_COST = (0, 1, 2, 3, 8, 9, 7, 8)        # idle, coast, goto, follow, old orbit, new orbit, warp, warp edge
_bound = False


def _lib():
    global _bound
    L = N.hip()
    if not _bound:
        vp, i32, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_float
        L.xb_world_features.argtypes = [ctypes.POINTER(N.ModesState), vp, vp, i32, vp, vp, vp]
        L.xb_mlp_forward.argtypes = [vp, vp, i32, vp, vp, vp, vp]
        L.xb_mlp_backward.argtypes = [vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp]
        L.xb_mlp_adam.argtypes = [vp, vp, vp, vp, i32, f32, vp]
        _bound = True
    return L


def _ptr(t):
    return None if t is None else t.data_ptr()


def _stream(dev):
    return torch.cuda.current_stream(dev).cuda_stream


def world_ranges(park) -> tuple:
    """(start, count) int32 device tensors of a world-contiguous park's worlds, in row order."""
    w = park.world
    change = torch.ones(park.n, dtype=torch.bool, device=park.dev)
    change[1:] = w[1:] != w[:-1]
    start = torch.nonzero(change).reshape(-1)
    count = torch.diff(torch.cat((start, torch.tensor([park.n], device=park.dev))))
    return start.to(torch.int32), count.to(torch.int32)


class Network:
    """The network's parameters and Adam moments on the device, and its kernels."""

    def __init__(self, dev, seed: int = 0):
        rng = np.random.default_rng(seed)
        parts = []
        for fan_in, shape in ((NI, (NI, NH)), (None, (NH,)), (NH, (NH, NH)), (None, (NH,)), (NH, (NH, NO)),
                              (None, (NO,))):
            parts.append(rng.normal(0.0, np.sqrt(2.0 / fan_in), shape).reshape(-1) if fan_in else np.zeros(shape))
        self.dev = torch.device(dev)
        self.params = torch.tensor(np.concatenate(parts), dtype=torch.float32, device=self.dev)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0

    def forward(self, x: torch.Tensor, keep: bool = False) -> tuple:
        n = x.shape[0]
        logits = torch.empty(n, NO, dtype=torch.float32, device=self.dev)
        h1 = torch.empty(n, NH, dtype=torch.float32, device=self.dev) if keep else None
        h2 = torch.empty(n, NH, dtype=torch.float32, device=self.dev) if keep else None
        N.check(_lib().xb_mlp_forward(_ptr(self.params), _ptr(x), n, _ptr(h1), _ptr(h2), _ptr(logits),
                                      _stream(self.dev)), "xb_mlp_forward")
        return logits, h1, h2

    def gradient(self, x: torch.Tensor, labels: torch.Tensor, check: bool = True) -> tuple:
        """(mean loss, gradient [NPARAM]) at the current parameters; labels int32 in 0..7 (checked unless check=False,
        for a caller that checked them once)."""
        n = x.shape[0]
        labels = labels.to(device=self.dev, dtype=torch.int32).contiguous()
        if n == 0 or labels.numel() != n:
            raise ValueError("xb_mlp_backward: needs one label for each of at least one sample")
        if check and (int(labels.min()) < 0 or int(labels.max()) >= NO):
            raise ValueError(f"xb_mlp_backward: labels must lie in 0..{NO - 1}")
        logits, h1, h2 = self.forward(x, keep=True)
        dl = torch.empty(n, NO, dtype=torch.float32, device=self.dev)
        d2 = torch.empty(n, NH, dtype=torch.float32, device=self.dev)
        d1 = torch.empty(n, NH, dtype=torch.float32, device=self.dev)
        loss = torch.empty(n, dtype=torch.float32, device=self.dev)
        work = torch.empty(CHUNKS, NPARAM, dtype=torch.float32, device=self.dev)
        grads = torch.empty(NPARAM, dtype=torch.float32, device=self.dev)
        N.check(_lib().xb_mlp_backward(_ptr(self.params), _ptr(x), _ptr(h1), _ptr(h2), _ptr(logits), _ptr(labels), n,
                                       _ptr(dl), _ptr(d2), _ptr(d1), _ptr(loss), _ptr(work), _ptr(grads),
                                       _stream(self.dev)), "xb_mlp_backward")
        return loss.sum(), grads

    def adam(self, grads: torch.Tensor, lr: float) -> None:
        self.step += 1
        N.check(_lib().xb_mlp_adam(_ptr(self.params), _ptr(grads), _ptr(self.m), _ptr(self.v), self.step, float(lr),
                                   _stream(self.dev)), "xb_mlp_adam")

    def train_step(self, x: torch.Tensor, labels: torch.Tensor, lr: float = 3e-3, check: bool = True) -> torch.Tensor:
        """One Adam step on a minibatch; returns its mean loss as a device scalar (no host synchronisation)."""
        loss, grads = self.gradient(x, labels, check)
        self.adam(grads, lr)
        return loss

    def predict(self, x: torch.Tensor) -> torch.Tensor:
        return self.forward(x)[0].argmax(dim=1).to(torch.uint8)


def features(park, start: torch.Tensor, count: torch.Tensor) -> torch.Tensor:
    """Each world's 12 features (xb_world_features), from the park's state and its recorded paths."""
    W = start.numel()
    out = torch.empty(W, NI, dtype=torch.float32, device=park.dev)
    s = park.state()
    N.check(_lib().xb_world_features(ctypes.byref(s), _ptr(start), _ptr(count), W, _ptr(park.path), _ptr(out),
                                     _stream(park.dev)), "xb_world_features")
    return out


def world_cost(park, start: torch.Tensor, count: torch.Tensor) -> torch.Tensor:
    """Each world's divergence cost in the last recorded tick: the summed weights of the distinct path classes its rows
    executed (integers)."""
    W = start.numel()
    world_of = torch.repeat_interleave(torch.arange(W, device=park.dev), count.to(torch.int64))
    path = park.path.to(torch.int64)
    cost = torch.zeros(W, dtype=torch.int64, device=park.dev)
    for c, wt in enumerate(PATH_COST):
        if wt:
            seen = torch.zeros(W, dtype=torch.int64, device=park.dev)
            seen.scatter_reduce_(0, world_of, (path == c).to(torch.int64), reduce="amax")
            cost += seen * wt
    return cost


class LearnedScheduler:
    """Trains the network on a park's own ticks and regroups the park's worlds by predicted cost bucket. The park must
    be tiled (every world contiguous and no wider than park.block); call regroup() after choosing the launch shape
    (the autotuner restores the row order) and before capturing a ModesGraph (a graph keeps the order it was made
    with)."""

    def __init__(self, park, seed: int = 0):
        if park.tiles is None:
            raise ValueError("the learned scheduler regroups tiles: the park needs contiguous worlds within its block")
        self.park = park
        self.start, self.count = world_ranges(park)
        self.net = Network(park.dev, seed)
        self.edges = None

    def collect(self, ticks: int) -> tuple:
        """(features, costs) over `ticks` ticks of the park (which advances): each world's features before a tick, its
        cost in that tick after it."""
        park = self.park
        if park.path is None:
            park.record_paths(True)
        xs, cs = [], []
        for _ in range(ticks):
            xs.append(features(park, self.start, self.count))
            park.evolve()
            cs.append(world_cost(park, self.start, self.count))
        return torch.cat(xs), torch.cat(cs)

    def bucket(self, cost: torch.Tensor) -> torch.Tensor:
        return torch.bucketize(cost, self.edges).clamp_(0, NO - 1)

    def fit(self, x: torch.Tensor, cost: torch.Tensor, epochs: int = 20, batch: int = 8192, lr: float = 3e-3,
            seed: int = 0) -> dict:
        """Buckets the costs by their quantiles and trains on (features, bucket) in shuffled minibatches."""
        q = torch.linspace(0, 1, NO + 1, dtype=torch.float64, device=cost.device)[1:-1]
        self.edges = torch.unique(torch.quantile(cost.to(torch.float64), q).round().to(torch.int64))
        y = self.bucket(cost).to(torch.int32)
        batch = min(batch, x.shape[0])
        g = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        for _ in range(epochs):
            perm = torch.randperm(x.shape[0], generator=g).to(x.device)
            for k in range(0, x.shape[0] - batch + 1, batch):
                idx = perm[k:k + batch]
                losses.append(self.net.train_step(x[idx].contiguous(), y[idx].contiguous(), lr, check=False))
        losses = torch.stack(losses).cpu()
        pred = self.net.predict(x)
        return {"samples": int(x.shape[0]), "steps": len(losses), "loss_first": float(losses[0]),
                "loss_last": float(losses[-1]), "accuracy": float((pred.to(torch.int64) == y).to(torch.float64).mean()),
                "edges": self.edges.tolist()}

    def train(self, ticks: int = 8, **kw) -> dict:
        x, c = self.collect(ticks)
        return self.fit(x, c, **kw)

    def predict(self) -> torch.Tensor:
        """Each world's predicted cost bucket for its next tick, from its current state and last recorded paths."""
        return self.net.predict(features(self.park, self.start, self.count))

    def regroup(self, buckets: torch.Tensor | None = None) -> dict:
        """Orders the park's worlds by bucket (the network's prediction by default; stable: ties keep row order),
        packs them into tiles of at most park.block rows and installs the order (park.rowmap, park.tiles)."""
        if buckets is None:
            buckets = self.predict()
        order = torch.sort(buckets.to(torch.int64), stable=True).indices
        return install_order(self.park, self.start, self.count, order)


def install_order(park, start: torch.Tensor, count: torch.Tensor, order: torch.Tensor | None) -> dict:
    """Installs a world order on a tiled park: rowmap = the worlds' rows in that order (each world's rows consecutive
    and in row order), tiles packed greedily (whole worlds, at most park.block rows). order=None restores the row
    order."""
    if order is None:
        park._set_tiles()
        return {"tiles": None if park.tiles is None else int(park.tiles.numel() - 1)}
    cnt = count.to(torch.int64)[order]
    if int(cnt.max()) > park.block:
        raise ValueError(f"a world is wider than the block ({park.block} rows)")
    st = start.to(torch.int64)[order]
    ends = torch.cumsum(cnt, 0)
    if int(ends[-1]) != park.n:
        raise ValueError("the order must hold every world once")
    rowmap = torch.repeat_interleave(st - (ends - cnt), cnt) + torch.arange(park.n, device=park.dev)
    if not torch.equal(torch.sort(rowmap).values, torch.arange(park.n, device=park.dev)):
        raise ValueError("the order must hold every world once")
    sizes = cnt.to(torch.int32).cpu().numpy()
    tiles = np.zeros(sizes.size + 1, dtype=np.int32)
    nt = ctypes.c_int64()
    N.check(N.cpu().xb_cpu_pack_tiles(sizes.ctypes.data, sizes.size, park.block, tiles.ctypes.data, ctypes.byref(nt)),
            "xb_cpu_pack_tiles")
    park.tiles = torch.from_numpy(tiles[:nt.value + 1].copy()).to(park.dev)
    park.rowmap = rowmap.to(torch.int32)
    park._state = None
    return {"tiles": int(nt.value)}


def modeled_cost(park, block: int = 64) -> int:
    """A model of the tiled launch's time for the last recorded tick: a wavefront of `block` lanes executes the union of
    its lanes' path classes (the sum of their weights), a workgroup waits for its slowest wavefront, and the launch
    costs the sum over workgroups. Lanes in a workgroup in the park's lane schedule (0: row order; otherwise by path
    class, as the branch-key sort places them)."""
    path = park.path.cpu().numpy().astype(np.int64)
    tiles = park.tiles.cpu().numpy()
    rows = park.rowmap.cpu().numpy() if park.rowmap is not None else np.arange(park.n)
    w = np.asarray(PATH_COST, dtype=np.int64)
    total = 0
    for g in range(len(tiles) - 1):
        p = path[rows[tiles[g]:tiles[g + 1]]]
        if park.schedule != 0:
            p = np.sort(p, kind="stable")
        total += max(int(w[np.unique(p[k:k + block])].sum()) for k in range(0, len(p), block))
    return total
