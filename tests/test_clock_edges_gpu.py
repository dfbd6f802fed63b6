"""Clock edges taken once per work item (GPU).

Flip-flops whose clock pin is on their work item's clock net leave the clock
events out of the per-gate merge; the workgroup stages the tile's edges and
adds their energies, toggles and leakage steps for all those gates at once.
GTP_CLOCK_EDGES=0 (read at every launch) makes every gate step through its
clock events.  Both must give the same integers, so every result is compared
bit-exactly between the two, and against the oracle as in the engine parity
tests (1e-6 of the term magnitudes, counts exact).

Cases: a regular clock; a clock that starts unknown (first event after time
0), goes X mid-trace and has three edges in one cycle (that tile falls back
to per-gate stepping); a gated clock with long gaps and edges on cycle
boundaries; D and Q events at the same times as clock edges; flip-flops on
three clock nets sharing work items; tiles of 64 to 1024 cycles (above 512
the launch steps through the edges); batched and tile-list runs.
"""
import numpy as np
import pytest

import oracle
from gtpower_amd import synthetic as S
from gtpower_amd.power import HipPower

pytestmark = pytest.mark.gpu

T = 1000  # ticks per cycle (1 ps ticks, 1 ns clock)
N_CYC = 700


def same(a, b, what):
    assert np.array_equal(a.gate_toggle_x2, b.gate_toggle_x2), f"{what}: toggle counts differ"
    assert np.array_equal(a.gate_glitch, b.gate_glitch), f"{what}: glitch counts differ"
    assert np.array_equal(a.gate_power, b.gate_power), f"{what}: gate power differs"
    assert np.array_equal(a.cycle_power, b.cycle_power), f"{what}: cycle power differs"


def near_oracle(res, d, tr):
    ref = oracle.run(d.library, d, tr, threads=0)
    assert np.array_equal(res.gate_toggle_x2, ref["gate_toggle_x2"])
    assert np.array_equal(res.gate_glitch, ref["gate_glitch"])
    err = np.abs(res.cycle_power - ref["cycle_power"])
    assert np.all(err <= 1e-6 * ref["cycle_abs"] + 1e-30)
    gerr = np.abs(res.gate_power - ref["gate_power"])
    assert np.all(gerr <= 1e-6 * ref["gate_abs"] + 1e-30)


def _clock_waves(rng):
    # net 0: 1 at 0, falls at k*T + T/2, rises at (k+1)*T
    c0 = [(0, 1)]
    for k in range(N_CYC):
        c0 += [(k * T + T // 2, 0), ((k + 1) * T, 1)]
    # net 1: unknown until 37, phase-shifted, X for two cycles, three edges in cycle 300
    c1 = [(37, 0)]
    for k in range(N_CYC):
        if k in (120, 121):
            continue
        if k == 119:
            c1 += [(k * T + 250, 1), (k * T + 750, 2)]
            continue
        if k == 300:
            c1 += [(k * T + 250, 1), (k * T + 500, 0), (k * T + 600, 1), (k * T + 750, 0)]
            continue
        c1 += [(k * T + 250, 1), (k * T + 750, 0)]
    # net 2: gated clock, edges on cycle boundaries, long gaps
    c2 = [(0, 0)]
    on = rng.random(N_CYC) < 0.3
    for k in range(1, N_CYC):
        if on[k]:
            c2 += [(k * T, 1), (k * T + 400, 0)]
    return [c0, c1, c2]


def _data_wave(rng, clock, tie_p=0.3, p=0.25, q_delay=None):
    """Events on a D or Q net: random times, some exactly on clock edge times
    (q_delay: the edge time plus a delay, or 0 for the same time)."""
    edges = [t for t, _ in clock[1:]]
    times = set()
    for k in range(N_CYC):
        if rng.random() < p:
            times.add(int(k * T + rng.integers(1, T)))
    for t in edges:
        if rng.random() < tie_p:
            times.add(int(t + (q_delay or 0)))
    times = sorted(t for t in times if t < N_CYC * T)
    wave = [(0, int(rng.integers(0, 2)))]
    v = wave[0][1]
    for t in times:
        if t == 0:
            continue
        v = 1 - v
        wave.append((t, v))
    return wave


def _design(library, seed, n_ff=150, d_to_q_arc=False):
    rng = np.random.default_rng(seed)
    lib = library
    ff, nand = lib.index("dfxtp_1"), lib.index("nand2_1")
    clocks = _clock_waves(rng)
    waves = list(clocks)
    cells, nets, arcs = [], [], []
    for g in range(n_ff):
        ck = int(rng.choice(3, p=[0.6, 0.25, 0.15]))
        dn = len(waves)
        waves.append(_data_wave(rng, clocks[ck], tie_p=0.2))
        qn = len(waves)
        waves.append(_data_wave(rng, clocks[ck], tie_p=0.5, p=0.05, q_delay=int(rng.choice([0, 0, 35]))))
        cells.append(ff)
        nets += [ck, dn, qn]
        arcs += [50e-12, 52e-12, 48e-12, 51e-12]  # CLK -> Q
        arcs += [40e-12, 41e-12, 39e-12, 42e-12] if d_to_q_arc else [np.nan] * 4
    # a few combinational gates on the flip-flop outputs
    for g in range(40):
        a_, b_ = (int(x) for x in rng.integers(3, len(waves), 2))
        yn = len(waves)
        waves.append(_data_wave(rng, clocks[0], tie_p=0.1, p=0.2, q_delay=60))
        cells.append(nand)
        nets += [a_, b_, yn]
        arcs += [30e-12] * 8
    P = len(nets)
    slew = rng.uniform(20e-12, 90e-12, 2 * P)
    cap = np.where(np.arange(P) % 3 == 2, rng.uniform(1e-15, 8e-15, P), 0.0)
    d = S.ArrayDesign(lib, cells, nets, slew, cap, np.array(arcs), len(waves))
    tr = S.trace_from_lists(waves, N_CYC * T, 1e-12, 1e-9)
    return d, tr


def _runs(d, traces, tile_cycles, gates_per_wg, block_lanes, tiles=None):
    with HipPower(tile_cycles=tile_cycles, gates_per_wg=gates_per_wg, block_lanes=block_lanes) as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(traces[0])
        res = eng.power()
        bat = eng.power_batch(traces) if len(traces) > 1 else []
        sub = eng.power_tiles(tile_cycles, np.array(tiles, np.int32)) if tiles is not None else None
    return res, bat, sub


@pytest.mark.parametrize("tile_cycles,gates_per_wg,block_lanes", [
    (64, 64, 64), (256, 128, 128), (512, 256, 256), (1024, 64, 64)])
def test_clock_edges_equal_per_gate_stepping(monkeypatch, library, tile_cycles, gates_per_wg, block_lanes):
    print(f"[clock] 3 clock nets, tile {tile_cycles}, {gates_per_wg} gates per range, {block_lanes} lanes",
          flush=True)
    d, tr = _design(library, 5)
    _, tr2 = _design(library, 6)
    tiles = {64: [0, 2, 4, 10], 256: [0, 2], 512: [1]}.get(tile_cycles)  # (a tile list needs tiles within the trace)
    on, on_b, on_t = _runs(d, [tr, tr2], tile_cycles, gates_per_wg, block_lanes, tiles)
    monkeypatch.setenv("GTP_CLOCK_EDGES", "0")
    off, off_b, off_t = _runs(d, [tr, tr2], tile_cycles, gates_per_wg, block_lanes, tiles)
    monkeypatch.delenv("GTP_CLOCK_EDGES")
    assert on.gate_toggle_x2.sum() > 0
    same(on, off, "resident run")
    for k, (a, b) in enumerate(zip(on_b, off_b)):
        same(a, b, f"batched trace {k}")
    if tiles is not None:
        assert np.array_equal(on_t[1], off_t[1]), "tile-list run differs"
    near_oracle(on, d, tr)


def test_clock_edges_with_d_to_q_arcs(monkeypatch, library):
    # the related-input search sees the latest consumed clock edge and D events
    print("[clock] D -> Q arcs, tile 128", flush=True)
    d, tr = _design(library, 9, d_to_q_arc=True)
    on, _, _ = _runs(d, [tr], 128, 64, 64)
    monkeypatch.setenv("GTP_CLOCK_EDGES", "0")
    off, _, _ = _runs(d, [tr], 128, 64, 64)
    monkeypatch.delenv("GTP_CLOCK_EDGES")
    same(on, off, "resident run")
    near_oracle(on, d, tr)


def test_clock_edges_synthetic_netlist(monkeypatch):
    print("[clock] synthetic netlist, 4000 gates, 20% flip-flops", flush=True)
    d = S.make_design(4000, dff_fraction=0.2, clock_ns=2.0, seed=71)
    tr = d.trace(600, toggle_p=0.2, seed=171)
    res = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("GTP_CLOCK_EDGES", flag)
        with HipPower(tile_cycles=128) as eng:
            eng.set_design(d.library, d)
            eng.read_power_activities(tr)
            res[flag] = eng.power()
            sl = eng.power_sliced(tr, max_events=tr.n_event // 2 + 4 * d.n_net)
        same(sl, res[flag], f"sliced run, GTP_CLOCK_EDGES={flag}")
    monkeypatch.delenv("GTP_CLOCK_EDGES")
    same(res["1"], res["0"], "resident run")
    near_oracle(res["1"], d, tr)
