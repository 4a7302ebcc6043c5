#!/usr/bin/env python3
"""Benchmark: bit-exact world-ticks/s of the batched destiny tick on one MI355X.

Workload: the flight-modes park, 65,536 worlds x 4 ships (exact_ballpark_amd.workloads.flight_park, seed 1): modes
uniform over stop, goto, follow, orbit and warp given at tick 1, then ticked. A step is one tick of every world, one
replay of the park's HIP graph (one kernel launch). The state is resident in HBM before the timed region.

Printed as one JSON line (rank 0): throughput, per-tick latency percentiles (HIP events around single ticks), the
kernel's HBM roofline fraction (algorithmic bytes: 361 B a ship-tick, each persistent per-ship field the tick reads or
writes counted once at destiny's sizes), and the CPU baseline: the oracle (a scalar C port of the reference's flight
world, one core) timed on a bounded sample of the same worlds, whose final states also check the GPU's bit for bit.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT / "exact-ballpark_amd", ROOT / "oracle", ROOT / "tests"):
    sys.path.insert(0, str(_p))

METRIC = "bit-exact world-ticks/s on 1 MI355X (ragged composed park); p99 tick latency"
BYTES_PER_SHIP_TICK = 361          # 164 read + 197 written, flight-modes park without dynamical orientation
HBM_PEAK_GBS = 8000.0              # MI355X HBM3E
TRAFFIC_FILE = ROOT / "profiles" / "flight_modes_traffic.json"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--worlds", type=int, default=65536)
    ap.add_argument("--ships", type=int, default=4)
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="about how long the oracle baseline runs")
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    world_size = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world_size != 1:
        print(f"bench.py runs on one GPU only (asked for {max(a.gpus, world_size)}): the worlds are independent and "
              "this build measures one GPU's share; see DESIGN.md for how the park would shard", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from exact_ballpark_amd import _native as N
    from exact_ballpark_amd import workloads as WL
    from exact_ballpark_amd.vec_modes import ModesGraph, VecModesPark

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    wl = WL.flight_park(1, a.worlds, a.ships)
    park = VecModesPark(wl["n"], device=dev, block=a.block)
    WL.load_flight_park(park, wl)
    graph = ModesGraph(park)
    print(f"# park {a.worlds} worlds x {a.ships} ships, tiled={park.tiles is not None}; warmup {a.warmup}",
          file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        graph.tick()
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream(dev)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(a.steps):
        graph.tick()
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_time(end)
    ms_per_step = gpu_ms / a.steps

    # per-tick latency: each tick between its own pair of events
    lat_ticks = max(50, min(a.steps, 200))
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(lat_ticks)]
    for e0, e1 in evs:
        e0.record(stream)
        graph.tick()
        e1.record(stream)
    torch.cuda.synchronize()
    lat = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    total_ticks = a.warmup + a.steps + lat_ticks

    n = wl["n"]
    value = a.worlds * a.steps / (gpu_ms / 1000.0)
    kernel_s = float(lat.mean()) / 1000.0
    achieved = BYTES_PER_SHIP_TICK * n / kernel_s / 1e9
    traffic = None
    if TRAFFIC_FILE.exists():
        tr = json.loads(TRAFFIC_FILE.read_text())
        if tr.get("worlds") == a.worlds and tr.get("ships") == a.ships:
            traffic = tr.get("bytes_per_launch")
    out = {
        "metric": METRIC, "value": value, "unit": "world-ticks/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"flight-modes park, {a.worlds} worlds x {a.ships} ships, modes uniform over "
                               "stop/goto/follow/orbit/warp, no dynamical orientation, HIP graph replay per tick",
                   "worlds": a.worlds, "ships_per_world": a.ships, "block": a.block, "seed": 1},
        "ship_ticks_per_s": value * a.ships,
        "wall_s_timed": wall,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max()), "ticks": lat_ticks},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch_algorithmic": BYTES_PER_SHIP_TICK * n,
                     "kernel_ms_mean": kernel_s * 1000.0},
        "kernel": N.kernel_info(0),
    }
    if not a.no_cpu:
        out["cpu_baseline"], out["exact"] = cpu_baseline(park, wl, total_ticks, a.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(park, wl: dict, ticks: int, seconds: float) -> tuple:
    """The oracle on one core over a sample of the benchmark's own worlds for the same number of ticks; its final
    states are compared with the GPU park's bit for bit."""
    import numpy as np

    import oracle as O
    import park_driver as D
    from exact_ballpark_amd import workloads as WL

    probe = [WL.world_scenario(wl, wi, ticks) for wi in range(min(64, wl["worlds"]))]
    dt, _ = O.time_worlds(probe, ticks)
    per_world = dt / len(probe)
    m = int(max(64, min(wl["worlds"], 16384, seconds / max(per_world, 1e-9))))
    idx = np.linspace(0, wl["worlds"] - 1, m).astype(int).tolist()
    sample = [WL.world_scenario(wl, wi, ticks) for wi in idx]
    print(f"# cpu baseline: oracle on {m} worlds x {ticks} ticks", file=sys.stderr, flush=True)
    dt, objs = O.time_worlds(sample, ticks)
    f = park._f[:, :park.n].cpu().numpy()
    q = park._q[:, :park.n].cpu().numpy()
    b = park._b[:, :park.n].cpu().numpy()
    s = wl["ships"]
    rows = np.array([wi * s + k for wi in idx for k in range(s)])      # ids 1..s in row order, as the oracle's
    fcols = [0, 1, 2, 3, 4, 5, 12, 13, 14, 15, 16, 17, 6, 7, 8, 9, 10, 11, 18, 21, 22, 23, 24, 25, 19, 20]
    gd = f[fcols][:, rows].T
    fp = q[5, rows]
    gq = np.stack([q[0, rows], q[1, rows], q[2, rows], q[3, rows], q[4, rows],
                   np.where(fp >= 0, q[0, np.maximum(fp, 0)], -1), b[2, rows], b[4, rows], b[3, rows], b[0, rows],
                   b[1, rows]], axis=1)
    states = [o.state() for o in objs]
    od = np.concatenate([st[0][:, D.D_COLS] for st in states])
    oq = np.concatenate([st[1][:, D.Q_COLS] for st in states])
    same = D._same_bits(gd, od).all(axis=1) & (gq == oq).all(axis=1)
    bad = int((~same.reshape(m, s).all(axis=1)).sum())
    cpu = {"value": m * ticks / dt, "unit": "world-ticks/s", "cores": 1, "kind": "port",
           "sample": f"{m} of the benchmark's worlds (evenly spaced) x {ticks} ticks, the oracle's C loop, 1 thread"}
    exact = {"worlds_checked": m, "ticks": ticks, "worlds_differing": bad, "bit_exact": bad == 0}
    return cpu, exact


if __name__ == "__main__":
    sys.exit(main())
