"""Per-tile lane order of the merge (GPU).

A full run orders the gates of each range per tile by their pin events in
that tile, so that the lanes of a wave carry similar work.  Only the lane
assignment changes: every partial goes to the gate's schedule position, so
the results must be bit-identical to the schedule order (GTP_TILE_ORDER=0,
read when a run is prepared) and within the engine parity bar of the oracle
(1e-6 of the term magnitudes, counts exact).
"""
import numpy as np
import pytest

import oracle
from gtpower_amd import synthetic as S
from gtpower_amd.power import HipPower

pytestmark = pytest.mark.gpu


def same(a, b, what):
    assert np.array_equal(a.gate_toggle_x2, b.gate_toggle_x2), f"{what}: toggle counts differ"
    assert np.array_equal(a.gate_glitch, b.gate_glitch), f"{what}: glitch counts differ"
    assert np.array_equal(a.gate_power, b.gate_power), f"{what}: gate power differs"
    assert np.array_equal(a.cycle_power, b.cycle_power), f"{what}: cycle power differs"


@pytest.mark.parametrize("tile_cycles,gates_per_wg,block_lanes", [
    (64, 128, 0), (256, 512, 0), (512, 1024, 0), (128, 192, 0), (128, 512, 64), (128, 100, 0)])
def test_tile_order_equals_schedule_order(monkeypatch, tile_cycles, gates_per_wg, block_lanes):
    # 192 gates: three waves, one pass; 64 lanes: eight passes of one wave;
    # 100 gates: lanes not whole waves (ranks stay in sorted order)
    print(f"[order] 6000 gates, tile {tile_cycles}, {gates_per_wg} gates per range, lanes {block_lanes}", flush=True)
    d = S.make_design(6000, dff_fraction=0.15, clock_ns=2.0, seed=81)
    tr = d.trace(1500, toggle_p=0.2, seed=181)
    res = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("GTP_TILE_ORDER", flag)
        with HipPower(tile_cycles=tile_cycles, gates_per_wg=gates_per_wg, block_lanes=block_lanes) as eng:
            eng.set_design(d.library, d)
            eng.read_power_activities(tr)
            res[flag] = eng.power()
            again = eng.power()
        same(again, res[flag], f"second evaluation, GTP_TILE_ORDER={flag}")
    monkeypatch.delenv("GTP_TILE_ORDER")
    same(res["1"], res["0"], "tile order vs schedule order")
    ref = oracle.run(d.library, d, tr, threads=0)
    r = res["1"]
    assert np.array_equal(r.gate_toggle_x2, ref["gate_toggle_x2"])
    assert np.array_equal(r.gate_glitch, ref["gate_glitch"])
    assert np.all(np.abs(r.cycle_power - ref["cycle_power"]) <= 1e-6 * ref["cycle_abs"] + 1e-30)
    assert np.all(np.abs(r.gate_power - ref["gate_power"]) <= 1e-6 * ref["gate_abs"] + 1e-30)


def test_tile_order_follows_a_new_trace():
    # the order is rebuilt with the tile index when another trace is loaded
    print("[order] two traces on one context", flush=True)
    d = S.make_design(3000, dff_fraction=0.15, clock_ns=2.0, seed=83)
    t1 = d.trace(800, toggle_p=0.1, seed=183)
    t2 = d.trace(1100, toggle_p=0.3, seed=184)
    with HipPower(tile_cycles=128, gates_per_wg=256) as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(t1)
        r1 = eng.power()
        eng.read_power_activities(t2)
        r2 = eng.power()
    for r, t in ((r1, t1), (r2, t2)):
        with HipPower(tile_cycles=128, gates_per_wg=256) as eng:
            eng.set_design(d.library, d)
            eng.read_power_activities(t)
            same(r, eng.power(), "fresh context")


@pytest.mark.parametrize("round_tiles", [0, 3])
def test_tile_order_in_batched_runs(monkeypatch, round_tiles):
    # per work tile (trace, tile), over rounds of a few work tiles
    print(f"[order] batched run of 4 traces, round_tiles {round_tiles}", flush=True)
    d = S.make_design(3000, dff_fraction=0.15, clock_ns=2.0, seed=85)
    traces = [d.trace(c, toggle_p=p, seed=185 + k) for k, (c, p) in enumerate(
        [(300, 0.1), (900, 0.25), (130, 0.3), (600, 0.15)])]
    res = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("GTP_TILE_ORDER", flag)
        with HipPower(tile_cycles=128, gates_per_wg=256, round_tiles=round_tiles) as eng:
            eng.set_design(d.library, d)
            eng.read_power_activities(traces[0])
            res[flag] = eng.power_batch(traces)
    monkeypatch.delenv("GTP_TILE_ORDER")
    for k, (a, b) in enumerate(zip(res["1"], res["0"])):
        same(a, b, f"batched trace {k}")
    with HipPower(tile_cycles=128, gates_per_wg=256) as eng:
        eng.set_design(d.library, d)
        for k, tr in enumerate(traces):
            eng.read_power_activities(tr)
            same(res["1"][k], eng.power(), f"trace {k} alone")
