"""The partition park (the master's space layer over the flight modes: ships in Gaussian clusters across top boxes,
bubbles, dead bubbles, proximity sensors of 1-250 km, shuffled sensors, follows refused across bubbles, warps to other
clusters) on the host build against the oracle's master world by world: every ball's flight state and the layer's
filing, and every master structure, event and query, after the load and after each of 24 ticks."""

import space_driver as SD
from exact_ballpark_amd import workloads as WL
from exact_ballpark_amd.vec_modes import VecModesPark
from exact_ballpark_amd.vec_space import SpaceGraph, VecSpace
from test_oracle_space_golden import run_oracle_space


def build(wl, device, events=256):
    """(park, layer) for a partition park on device, loaded."""
    park = VecModesPark(wl["n"], device=device, block=256, display_ypr=True, master=True)
    return park, WL.load_partition_park(park, wl, lambda *a, **k: VecSpace(*a, events=events, **k))


def run_product(wl, device, ticks, every=True, graph=False):
    park, space = build(wl, device)
    ws = [WL.partition_scenario(wl, wi, ticks) for wi in range(wl["worlds"])]
    rows_of = [list(WL.world_rows(wl, wi)) for wi in range(wl["worlds"])]
    out = {}
    g = None
    for s in range(0, ticks + 1):
        if s and graph:
            g = g or SpaceGraph(space)
            g.tick()
        elif s:
            space.tick()
        if every or s == ticks:
            masters = SD.master_states(park, space, ws, None)
            for wi in range(wl["worlds"]):
                d, q = SD._flight(park, space, wi, rows_of[wi])
                out[(wi, s)] = (d, q, masters[wi])
    space.check()
    return out, ws


def compare_oracle(got, ws, worlds, ticks, every=True):
    bad = []
    for wi in worlds:
        for step, d, q, m in run_oracle_space(ws[wi]):
            if not every and step != ticks:
                continue
            gd, gq, gm = got[(wi, step)]
            ok = d.shape == gd.shape and SD.D._same_bits(d, gd).all() and (q == gq).all()
            keys = [k for k in m if gm.get(k) != m[k]]
            if not ok or keys:
                bad.append((wi, step, "balls" if not ok else "", keys[:4]))
                break
    return bad


def test_partition_park_matches_oracle():
    wl = WL.partition_park(3, 48, 4, 40)
    got, ws = run_product(wl, "cpu", 24)
    bad = compare_oracle(got, ws, range(48), 24)
    assert not bad, f"{len(bad)} worlds differ, first {bad[:6]}"
    ev = sum(len(got[(wi, s)][2]["events"]) for wi in range(48) for s in range(25))
    assert ev > 50
