"""Kernels that stride over per-event work (GPU).

A dispatch counts its work-items in 32 bits, so a grid of one 256-thread
block per 256-event chunk stops covering the event store above about 4.3e9
events.  The glitch annotation and the toggle correlation therefore take
chunks in a grid stride under a capped grid.  These tests force a tiny cap
(GTP_STRIDE_BLOCKS, read at every launch) so that small inputs run the stride
loops, and run one trace above 2^32 events, where the resident run is checked
against the sliced run (whose loads each stay below 2^32 events).

Tolerances: bit-exact between capped and uncapped runs and between the
resident and the sliced run; the capped run against the oracle as in the
engine parity tests (1e-6 of the term magnitudes, counts exact).
"""
import numpy as np
import pytest

import oracle
from gtpower_amd import synthetic as S
from gtpower_amd.power import HipPower
from gtpower_amd.surrogate import cycle_total

pytestmark = pytest.mark.gpu

CAP = "3"


def same(a, b, what):
    assert np.array_equal(a.gate_toggle_x2, b.gate_toggle_x2), f"{what}: toggle counts differ"
    assert np.array_equal(a.gate_glitch, b.gate_glitch), f"{what}: glitch counts differ"
    assert np.array_equal(a.gate_power, b.gate_power), f"{what}: gate power differs"
    assert np.array_equal(a.cycle_power, b.cycle_power), f"{what}: cycle power differs"


def _run(d, traces, batch):
    with HipPower() as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(traces[0])
        res = eng.power()
        out = eng.power_batch(traces) if batch else None
    return res, out


def test_capped_grid_annotation(monkeypatch):
    print("[stride] glitch annotation under a 3-block grid", flush=True)
    d = S.make_design(2000, dff_fraction=0.1, clock_ns=10.0, seed=31)
    t1 = d.trace(1500, toggle_p=0.2, seed=131)
    t2 = d.trace(900, toggle_p=0.3, seed=132)
    full, full_b = _run(d, [t1, t2], batch=True)
    monkeypatch.setenv("GTP_STRIDE_BLOCKS", CAP)
    capped, capped_b = _run(d, [t1, t2], batch=True)
    monkeypatch.delenv("GTP_STRIDE_BLOCKS")
    assert capped.gate_glitch.sum() > 0
    same(capped, full, "resident run")
    for k, (a, b) in enumerate(zip(capped_b, full_b)):
        same(a, b, f"batched trace {k}")
    ref = oracle.run(d.library, d, t1, threads=0)
    assert np.array_equal(capped.gate_glitch, ref["gate_glitch"])
    err = np.abs(capped.cycle_power - ref["cycle_power"])
    assert np.all(err <= 1e-6 * ref["cycle_abs"] + 1e-30)


def test_capped_grid_correlation(monkeypatch):
    print("[stride] toggle correlation under a 3-block grid", flush=True)
    d = S.make_design(1500, dff_fraction=0.1, clock_ns=10.0, seed=33)
    tr = d.trace(1200, toggle_p=0.2, seed=133)
    with HipPower() as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(tr)
        y = cycle_total(eng.power(want=("cycle",)).cycle_power)
        mask = np.zeros(y.size, np.uint8)
        mask[1::2] = 1
        full = eng.feature_correlation(y, mask)
        monkeypatch.setenv("GTP_STRIDE_BLOCKS", CAP)
        capped = eng.feature_correlation(y, mask)
        monkeypatch.delenv("GTP_STRIDE_BLOCKS")
    assert np.count_nonzero(full) > 0
    assert np.array_equal(capped, full)


def _burst_waves(seed):
    """Nets whose cycle groups are longer than a 256-event chunk and cross
    chunk boundaries (T = 1000 ticks, 400 cycles)."""
    rng = np.random.default_rng(seed)

    def alt(times, v0=1):
        return [(int(t), (v0 + k) % 2) for k, t in enumerate(times)]

    w0 = [(0, 0)] + alt(2000 + np.arange(700)) + alt(5100 + np.arange(257), 0)[:257] + \
        alt([9100, 9200, 9300]) + alt(10000 + np.arange(256) * 3)
    w0 = [w0[0]] + [(t, k % 2) for k, (t, _) in enumerate(w0[1:], start=1)]
    w1 = alt(np.arange(1, 251) * 1000 + 500) + alt(260000 + np.arange(300)) + alt(np.arange(270, 280) * 1000 + 7)
    t2 = np.sort(rng.choice(np.arange(1, 400000), size=1500, replace=False))
    w2 = alt(t2)
    w3 = [(0, 1), (123456, 0), (223456, 1)]
    w4 = alt(np.arange(1, 800) * 500)
    return [w0, w1, w2, w3, w4]


@pytest.mark.parametrize("cap", ["1", "2", "3"])
def test_long_cycle_groups_under_capped_grids(monkeypatch, library, cap):
    # groups of 700, 257, 256 and 300 events, some across chunk boundaries;
    # small grids start waves in the middle of nets and of groups
    print(f"[stride] long cycle groups, GTP_STRIDE_BLOCKS={cap}", flush=True)
    lib = library
    cm, ca = lib.index("mux2_1"), lib.index("a21oi_1")
    nets = [0, 1, 2, 3, 1, 2, 0, 4]
    slew = np.linspace(30e-12, 90e-12, 16)
    cap_f = np.array([0, 0, 0, 6e-15, 0, 0, 0, 3e-15])
    arcs = np.full(24, 40e-12)
    d = S.ArrayDesign(lib, [cm, ca], nets, slew, cap_f, arcs, 5)
    tr = S.trace_from_lists(_burst_waves(7), 400000, 1e-12, 1e-9)
    with HipPower() as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(tr)
        full = eng.power()
    monkeypatch.setenv("GTP_STRIDE_BLOCKS", cap)
    with HipPower() as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(tr)
        capped = eng.power()
        batched = eng.power_batch([tr, tr])
    monkeypatch.delenv("GTP_STRIDE_BLOCKS")
    assert capped.gate_glitch.sum() > 500
    same(capped, full, "resident run")
    for k, b in enumerate(batched):
        same(b, full, f"batched copy {k}")
    ref = oracle.run(d.library, d, tr, threads=0)
    assert np.array_equal(full.gate_glitch, ref["gate_glitch"])
    assert np.array_equal(full.gate_toggle_x2, ref["gate_toggle_x2"])
    err = np.abs(full.cycle_power - ref["cycle_power"])
    assert np.all(err <= 1e-6 * ref["cycle_abs"] + 1e-30)
    gerr = np.abs(full.gate_power - ref["gate_power"])
    assert np.all(gerr <= 1e-6 * ref["gate_abs"] + 1e-30)


def test_above_2e32_events_resident_equals_sliced():
    # 200k gates x 100k cycles: 4.56e9 events, one event store of 18 GB
    print("[stride] 200000 gates x 100000 cycles, resident vs sliced", flush=True)
    d, tr = S.workload("exact_200k")
    try:
        assert tr.n_event > 2 ** 32
        with HipPower() as eng:
            eng.set_design(d.library, d)
            eng.read_power_activities(tr)
            res = eng.power(want=("cycle", "counts"))
            budget = tr.n_event // 4 + 4 * d.n_net
            sl = eng.power_sliced(tr, max_events=budget, want=("cycle", "counts"))
        print(f"[stride] {tr.n_event} events, {sl.schedule['n_interval']} intervals", flush=True)
        assert sl.schedule["n_interval"] > 1
        assert np.array_equal(res.gate_toggle_x2, sl.gate_toggle_x2)
        assert np.array_equal(res.gate_glitch, sl.gate_glitch)
        assert np.array_equal(res.cycle_power, sl.cycle_power)
    finally:
        tr.close()
