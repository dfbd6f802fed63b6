"""Parity of the HIP engine with the FP64 oracle (GPU).

Tolerances (written here, checked per gate and per cycle):
  * toggle and glitch counts: bit-exact;
  * every power value: |gpu - oracle| <= 1e-6 * scale + 1e-30, where `scale`
    is the oracle's sum of |term| magnitudes for that gate / cycle / component
    (for internal energy the interpolation of |table values| with |weights|,
    the error-propagation scale of an FP32 table lookup).  For one-signed
    quantities (switching, leakage) this is 1e-6 relative.
"""
import numpy as np
import pytest

import oracle
from gtpower_amd import synthetic as S
from gtpower_amd.power import HipPower

pytestmark = pytest.mark.gpu

REL = 1e-6


def compare(res, ref, what=""):
    G = res.gate_power.shape[1]
    assert np.array_equal(res.gate_toggle_x2, ref["gate_toggle_x2"]), f"{what}: toggle counts differ"
    assert np.array_equal(res.gate_glitch, ref["gate_glitch"]), f"{what}: glitch counts differ"
    gerr = np.abs(res.gate_power - ref["gate_power"])
    gtol = REL * ref["gate_abs"] + 1e-30
    bad = np.argwhere(gerr > gtol)
    assert len(bad) == 0, f"{what}: gate power mismatch at {bad[:5].tolist()} " \
        f"gpu={res.gate_power[tuple(bad[0])]!r} ref={ref['gate_power'][tuple(bad[0])]!r}"
    cerr = np.abs(res.cycle_power - ref["cycle_power"])
    ctol = REL * ref["cycle_abs"] + 1e-30
    bad = np.argwhere(cerr > ctol)
    assert len(bad) == 0, f"{what}: cycle power mismatch at {bad[:5].tolist()} " \
        f"gpu={res.cycle_power[tuple(bad[0])]!r} ref={ref['cycle_power'][tuple(bad[0])]!r}"
    tot = ref["gate_power"].sum(axis=1)
    assert np.allclose(res.total, tot, rtol=1e-6, atol=1e-30)


def run_both(design, trace, **opts):
    with HipPower(**opts) as eng:
        eng.set_design(design.library, design)
        eng.read_power_activities(trace)
        res = eng.power()
    ref = oracle.run(design.library, design, trace, threads=0)
    return res, ref


@pytest.mark.parametrize("n_gate,n_cycle,seed", [(200, 400, 1), (2000, 2000, 2)])
def test_random_design_parity(n_gate, n_cycle, seed):
    print(f"[parity] random design n_gate={n_gate} n_cycle={n_cycle} seed={seed}", flush=True)
    d = S.make_design(n_gate, dff_fraction=0.1, clock_ns=10.0, seed=seed)
    tr = d.trace(n_cycle, toggle_p=0.2, seed=seed + 100)
    res, ref = run_both(d, tr)
    compare(res, ref, f"random {n_gate}x{n_cycle}")
    assert res.gate_glitch.sum() > 0  # the workload exercises glitch filtering


@pytest.mark.parametrize("tile", [1, 7, 64, 512])
def test_tile_lengths_agree(tile):
    print(f"[parity] tile_cycles={tile}", flush=True)
    d = S.make_design(300, dff_fraction=0.15, clock_ns=2.0, seed=11)
    tr = d.trace(300, toggle_p=0.15, seed=12)
    res, ref = run_both(d, tr, tile_cycles=tile, gates_per_wg=64)
    compare(res, ref, f"tile {tile}")


def test_deterministic():
    print("[parity] determinism", flush=True)
    d = S.make_design(500, dff_fraction=0.1, clock_ns=10.0, seed=5)
    tr = d.trace(500, toggle_p=0.2, seed=6)
    with HipPower() as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(tr)
        a = eng.power()
        b = eng.power()
    assert np.array_equal(a.gate_power, b.gate_power)
    assert np.array_equal(a.cycle_power, b.cycle_power)


def test_random_activity_mode():
    print("[parity] random activity (skew) mode", flush=True)
    d = S.make_design(400, dff_fraction=0.15, clock_ns=2.0, seed=21)
    tr = d.trace(400, mode=1, lognorm_mu=-3.0, lognorm_sigma=1.5, hot_fraction=0.05, hot_share=0.6,
                 burst_fraction=0.1, burst_mult=8.0, max_rate=16.0, seed=22)
    res, ref = run_both(d, tr)
    compare(res, ref, "skew")


def test_trace_end_on_cycle_boundary_and_partial_cycle():
    print("[parity] t_max edge cases", flush=True)
    d = S.make_design(100, dff_fraction=0.1, clock_ns=1.0, seed=31)
    tr = d.trace(50, toggle_p=0.3, seed=32)
    # a trace whose end is not a whole cycle, with events at t_max
    T = tr.ticks_per_cycle
    t_max = 49 * T + T // 3
    off = tr.net_event_offset
    keep = tr.event_time <= t_max
    counts = np.array([keep[off[n]:off[n + 1]].sum() for n in range(len(off) - 1)])
    noff = np.zeros(len(off), np.int64)
    noff[1:] = np.cumsum(counts)
    tt = tr.event_time[keep].copy()
    vv = tr.event_value[keep].copy()
    tr2 = S.ArrayTrace(noff, tt, vv, t_max, tr.time_scale, tr.clk_period, T)
    res, ref = run_both(d, tr2, tile_cycles=16)
    compare(res, ref, "partial last cycle")
    assert res.n_cycle == 50
