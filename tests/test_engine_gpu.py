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
from gtpower_amd.abi import GtpError
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
    # design totals: the same bar over all terms (internal energies may cancel)
    tot = ref["gate_power"].sum(axis=1)
    assert np.all(np.abs(res.total - tot) <= REL * ref["gate_abs"].sum(axis=1) + 1e-30), \
        f"{what}: totals gpu={res.total!r} ref={tot!r}"



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


# ---------------------------------------------------------------------------
# hand-built edge cases
# ---------------------------------------------------------------------------
def _mux_design(lib, nets, n_net):
    """Two gates: a mux2_1 and an a21oi_1 on the given pin nets."""
    cm, ca = lib.index("mux2_1"), lib.index("a21oi_1")
    P = 8
    slew = np.linspace(30e-12, 90e-12, 2 * P)
    cap = np.array([0, 0, 0, 6e-15, 0, 0, 0, 3e-15])
    arcs = np.full(2 * 12, np.nan)
    rng = np.random.default_rng(0)
    for g, (cell, nin) in enumerate(((cm, 3), (ca, 3))):
        for i in range(nin):
            for fr in range(2):
                for to in range(2):
                    if lib.cells[cell]["arcs"].get((0, i, fr, to)):
                        arcs[g * 12 + (i * 2 + fr) * 2 + to] = rng.uniform(20e-12, 80e-12)
    return S.ArrayDesign(lib, [cm, ca], nets, slew, cap, arcs, n_net)


@pytest.mark.parametrize("case", ["x_and_z", "late_first_event", "shared_net", "unconnected", "coincident"])
def test_hand_cases(library, case):
    print(f"[parity] hand case {case}", flush=True)
    lib = library
    T = 1000
    if case == "x_and_z":
        waves = [[(0, 2), (150, 0), (900, 1), (1300, 3), (2100, 2), (2200, 0)],
                 [(0, 1), (400, 0), (420, 1), (2500, 2)],
                 [(0, 0), (700, 1), (710, 0), (730, 1), (3300, 0)],
                 [(0, 2), (190, 0), (950, 1), (1350, 2), (2250, 0)],
                 [(0, 0), (2700, 1)]]
        nets = [0, 1, 2, 3, 1, 2, 4, 4]
    elif case == "late_first_event":
        waves = [[(300, 1), (1500, 0)], [(0, 1)], [(1200, 1), (1250, 0), (1260, 1)], [(350, 0), (1600, 1)],
                 [(2000, 1)]]
        nets = [0, 1, 2, 3, 0, 2, 1, 4]
    elif case == "shared_net":
        waves = [[(0, 0), (100, 1), (180, 0), (1100, 1)], [(0, 1), (300, 0)], [(0, 0), (130, 1), (200, 0)],
                 [(0, 1), (150, 0), (1150, 1)], []]
        nets = [0, 0, 1, 2, 0, 0, 0, 3]
    elif case == "unconnected":
        waves = [[(0, 0), (100, 1), (1100, 0)], [(0, 1), (130, 0), (1160, 1)], [], [], []]
        nets = [0, -1, -1, 1, 0, -1, 4, 4]
    else:  # coincident events on several pins, in and out at one time
        waves = [[(0, 0), (500, 1), (1500, 0)], [(0, 1), (500, 0), (1500, 1)], [(0, 0), (500, 1), (2500, 0)],
                 [(0, 1), (500, 0), (540, 1), (1500, 0)], [(0, 0), (540, 1)]]
        nets = [0, 1, 2, 3, 2, 1, 0, 4]
    d = _mux_design(lib, nets, len(waves))
    tr = S.trace_from_lists(waves, 4000, 1e-12, T * 1e-12)
    res, ref = run_both(d, tr, tile_cycles=2)
    compare(res, ref, case)


def test_when_conditioned_library_uses_the_generic_kernel(library):
    """A library whose nand2 input power depends on the state (a `when`
    condition) is evaluated by the generic kernel with state-indexed energies."""
    print("[parity] when-conditioned library", flush=True)
    from gtpower_amd.library import CellLibrary
    lib = CellLibrary()
    c = lib.index("nand2_1")
    ipa = lib.ipower_array
    # nand2 pin A: power only when B is 1 (states 2, 3, 6, 7)
    k = [i for i in range(len(ipa)) if ipa[i]["cell"] == c and ipa[i]["pin"] == 0][0]
    ipa[k]["when_truth"] = sum(1 << s for s in (2, 3, 6, 7))
    # nand2 output, related A: only when Y was 1 before (states 4..7)
    k2 = [i for i in range(len(ipa)) if ipa[i]["cell"] == c and ipa[i]["pin"] == 2 and ipa[i]["related_pin"] == 0][0]
    ipa[k2]["when_truth"] = sum(1 << s for s in range(4, 8))
    lib.ipowers = [tuple(r) for r in ipa[["cell", "pin", "related_pin", "rise_table", "fall_table", "when_truth"]]]
    d = S.make_design(400, dff_fraction=0.1, clock_ns=2.0, seed=41, library=lib)
    tr = d.trace(300, toggle_p=0.3, seed=42)
    res, ref = run_both(d, tr)
    compare(res, ref, "when")


def _batch_traces(d, lengths, seed0):
    return [d.trace(n, toggle_p=p, seed=seed0 + i) for i, (n, p) in enumerate(lengths)]


@pytest.mark.parametrize("opts", [dict(), dict(round_tiles=3), dict(max_event_num=400000, round_tiles=2)])
def test_batch_equals_individual_runs(opts):
    """gtp_run_batch over traces of uneven length (some shorter than a tile)
    gives each trace exactly what gtp_run gives it: counts and every power
    value bit-identical (exact integer sums), across device loads and rounds."""
    print(f"[batch] 1500 gates, 6 traces, opts={opts}", flush=True)
    d = S.make_design(1500, dff_fraction=0.15, clock_ns=2.0, seed=41)
    lengths = [(700, 0.15), (37, 0.3), (1200, 0.05), (129, 0.2), (1, 0.2), (400, 0.25)]
    trs = _batch_traces(d, lengths, 500)
    singles = []
    with HipPower(tile_cycles=64) as eng:
        eng.set_design(d.library, d)
        for tr in trs:
            eng.read_power_activities(tr)
            singles.append(eng.power())
    with HipPower(tile_cycles=64, **opts) as eng:
        eng.set_design(d.library, d)
        batch = eng.power_batch(trs)
    for k, (a, b) in enumerate(zip(singles, batch)):
        assert b.n_cycle == a.n_cycle
        assert np.array_equal(a.gate_toggle_x2, b.gate_toggle_x2), k
        assert np.array_equal(a.gate_glitch, b.gate_glitch), k
        assert np.array_equal(a.cycle_power, b.cycle_power), k
        assert np.array_equal(a.gate_power, b.gate_power), k
        assert np.array_equal(a.total, b.total), k


@pytest.mark.parametrize("parts,buffers", [(1, 2), (3, 2), (11, 2), (11, 1)])
def test_sliced_equals_full_run(parts, buffers, monkeypatch):
    """gtp_run_sliced with an event budget that forces about `parts`
    intervals gives exactly what one gtp_run over the whole trace gives:
    counts and every power value bit-identical (the slices carry the events
    that the state, glitch flags and related-pin choice at their edges need;
    per-gate sums are exact integers over intervals).  With two device
    buffers the next interval is packed and uploaded while one runs; with one
    they take turns."""
    if buffers == 1:
        monkeypatch.setenv("GTP_SLICE_BUFFERS", "1")
    print(f"[sliced] 2000 gates x 1500 cycles, ~{parts} intervals, {buffers} buffer(s)", flush=True)
    d = S.make_design(2000, dff_fraction=0.15, clock_ns=2.0, seed=45)
    tr = d.trace(1500, toggle_p=0.2, seed=46)
    with HipPower(tile_cycles=64) as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(tr)
        a = eng.power()
        b = eng.power_sliced(tr, max_events=(tr.n_event + parts - 1) // parts + 4 * d.n_net if parts > 1 else 0)
    if parts > 1:
        assert b.schedule["n_interval"] >= min(parts, 2)
    assert b.n_cycle == a.n_cycle
    assert np.array_equal(a.gate_toggle_x2, b.gate_toggle_x2)
    assert np.array_equal(a.gate_glitch, b.gate_glitch)
    assert np.array_equal(a.cycle_power, b.cycle_power)
    assert np.array_equal(a.gate_power, b.gate_power)
    assert np.array_equal(a.total, b.total)


def test_sliced_generic_kernel_equals_full_run():
    """The same with nand2 evaluated by the generic kernel (a `when`
    condition on its input power), which reads the glitch widths the slice's
    annotation writes."""
    from gtpower_amd.library import CellLibrary
    lib = CellLibrary()
    c = lib.index("nand2_1")
    ipa = lib.ipower_array
    k = [i for i in range(len(ipa)) if ipa[i]["cell"] == c and ipa[i]["pin"] == 0][0]
    ipa[k]["when_truth"] = sum(1 << s for s in (2, 3, 6, 7))
    lib.ipowers = [tuple(r) for r in ipa[["cell", "pin", "related_pin", "rise_table", "fall_table", "when_truth"]]]
    print("[sliced] generic kernel, ~7 intervals", flush=True)
    d = S.make_design(1200, dff_fraction=0.1, clock_ns=2.0, seed=49, library=lib)
    tr = d.trace(900, toggle_p=0.3, seed=50)
    with HipPower(tile_cycles=32) as eng:
        eng.set_design(lib, d)
        eng.read_power_activities(tr)
        a = eng.power()
        b = eng.power_sliced(tr, max_events=tr.n_event // 7 + 4 * d.n_net)
    assert b.schedule["n_interval"] >= 2
    for key in ("gate_toggle_x2", "gate_glitch", "cycle_power", "gate_power", "total"):
        assert np.array_equal(getattr(a, key), getattr(b, key)), key


def test_sliced_rejects_a_budget_below_one_tile():
    d = S.make_design(300, dff_fraction=0.1, clock_ns=2.0, seed=47)
    tr = d.trace(300, toggle_p=0.2, seed=48)
    with HipPower(tile_cycles=64) as eng:
        eng.set_design(d.library, d)
        with pytest.raises(GtpError):
            eng.power_sliced(tr, max_events=10)


def test_batch_matches_oracle():
    print("[batch] oracle parity of a 3-trace batch", flush=True)
    d = S.make_design(800, dff_fraction=0.1, clock_ns=10.0, seed=43)
    trs = _batch_traces(d, [(300, 0.2), (90, 0.3), (500, 0.1)], 700)
    with HipPower() as eng:
        eng.set_design(d.library, d)
        batch = eng.power_batch(trs)
    for k, tr in enumerate(trs):
        compare(batch[k], oracle.run(d.library, d, tr, threads=0), f"batch trace {k}")


def test_batch_rejects_mixed_clocks():
    d = S.make_design(200, dff_fraction=0.1, clock_ns=10.0, seed=44)
    t1 = d.trace(50, toggle_p=0.2, seed=1)
    d2 = S.make_design(200, dff_fraction=0.1, clock_ns=5.0, seed=44)
    t2 = d2.trace(50, toggle_p=0.2, seed=2)
    from gtpower_amd import abi
    with HipPower() as eng:
        eng.set_design(d.library, d)
        with pytest.raises(abi.GtpError):
            eng.power_batch([t1, t2])


def test_autotune_keeps_results_and_feeds_batches():
    """The autotuner only changes the schedule: results are bit-identical to
    the default schedule's, its choice is one of its candidates, and batched
    runs of the same context use it."""
    print("[autotune] 3000 gates x 1500 cycles", flush=True)
    d = S.make_design(3000, dff_fraction=0.15, clock_ns=2.0, seed=45)
    tr = d.trace(1500, toggle_p=0.15, seed=46)
    with HipPower() as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(tr)
        ref = eng.power()
    with HipPower(autotune=1) as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(tr)
        res = eng.power()
        sched = res.schedule
        assert sched["autotune_ms"] > 0
        assert sched["tile_cycles"] in (32, 64, 128, 256, 512)
        assert sched["gates_per_wg"] in (256, 512, 1024, 2048, 4096, 8192)
        batch = eng.power_batch([tr])
        assert batch[0].schedule["tile_cycles"] == sched["tile_cycles"]
    for r in (res, batch[0]):
        assert np.array_equal(r.cycle_power, ref.cycle_power)
        assert np.array_equal(r.gate_power, ref.gate_power)
        assert np.array_equal(r.gate_toggle_x2, ref.gate_toggle_x2)


def test_silent_trace_and_a_single_partial_cycle():
    """A trace with no events at all (leakage of the default states only, an
    empty event store, no annotation chunks) and one shorter than a cycle,
    resident, sliced and batched, against the oracle."""
    print("[parity] silent trace and single partial cycle", flush=True)
    d = S.make_design(300, dff_fraction=0.1, clock_ns=1.0, seed=41)
    tr = d.trace(20, toggle_p=0.3, seed=42)
    T = tr.ticks_per_cycle
    N = len(tr.net_event_offset) - 1
    silent = S.ArrayTrace(np.zeros(N + 1, np.int64), np.zeros(0, np.int64), np.zeros(0, np.uint8),
                          20 * T, tr.time_scale, tr.clk_period, T)
    short = S.slice_trace(tr, 1)
    short = S.ArrayTrace(short.net_event_offset, short.event_time, short.event_value, T // 2 + 7,
                         tr.time_scale, tr.clk_period, T)
    # keep only the events before the shortened end (inclusive, as at t_max)
    off = short.net_event_offset
    keep = short.event_time <= short.t_max
    counts = np.array([keep[off[n]:off[n + 1]].sum() for n in range(N)])
    noff = np.zeros(N + 1, np.int64)
    noff[1:] = np.cumsum(counts)
    short = S.ArrayTrace(noff, short.event_time[keep], short.event_value[keep], short.t_max,
                         tr.time_scale, tr.clk_period, T)
    for name, t in (("silent", silent), ("single partial cycle", short)):
        res, ref = run_both(d, t, tile_cycles=4)
        compare(res, ref, name)
        with HipPower(tile_cycles=4) as eng:
            eng.set_design(d.library, d)
            eng.read_power_activities(t)
            full = eng.power()
            sl = eng.power_sliced(t, max_events=max(t.n_event, 1) + 4 * d.n_net)
            both = eng.power_batch([t, tr])
        for other, what in ((sl, "sliced"), (both[0], "batched")):
            assert np.array_equal(other.gate_glitch, full.gate_glitch), f"{name} {what}"
            assert np.array_equal(other.gate_toggle_x2, full.gate_toggle_x2), f"{name} {what}"
            assert np.array_equal(other.cycle_power, full.cycle_power), f"{name} {what}"
    assert res.n_cycle == 1


@pytest.mark.parametrize("staged", [False, True])
def test_invalid_traces_are_rejected_and_the_context_stays_usable(monkeypatch, staged):
    """Each malformed trace is refused with an error status (never a crash),
    by the resident, sliced and streamed-batch entry points; a valid trace
    works after.
    staged: the resident load goes through the page-locked staging ring,
    which checks each event as it packs it (large traces take that path)."""
    if staged:
        monkeypatch.setenv("GTP_STAGED_MIN_EVENTS", "1")
    print(f"[errors] malformed traces (staged={staged})", flush=True)
    d = S.make_design(60, dff_fraction=0.1, clock_ns=1.0, seed=51)
    tr = d.trace(10, toggle_p=0.3, seed=52)
    T = tr.ticks_per_cycle
    off, tm, val = tr.net_event_offset, tr.event_time, tr.event_value
    n_busy = int(np.argmax(np.diff(off) >= 2))  # a net with at least two events

    def variant(off=off, tm=tm, val=val, t_max=tr.t_max, clk=tr.clk_period):
        return S.ArrayTrace(off.copy(), tm.copy(), val.copy(), t_max, tr.time_scale, clk, T)

    cases = {}
    o = off.copy(); o[0] = 1
    cases["offset 0 not zero"] = variant(off=o)
    o = off.copy(); o[n_busy + 1] = o[n_busy] - 1 if o[n_busy] > 0 else -1
    cases["decreasing offsets"] = variant(off=o)
    t = tm.copy(); a = off[n_busy]; t[a], t[a + 1] = t[a + 1], t[a]
    cases["unsorted times"] = variant(tm=t)
    t = tm.copy(); t[off[n_busy]] = -5
    cases["negative time"] = variant(tm=t)
    t = tm.copy(); t[off[n_busy + 1] - 1] = tr.t_max + 1
    cases["time past t_max"] = variant(tm=t)
    v = val.copy(); v[off[n_busy]] = 7
    cases["value above 3"] = variant(val=v)
    cases["t_max zero"] = variant(t_max=0)
    with HipPower() as eng:
        eng.set_design(d.library, d)
        for name, bad in cases.items():
            with pytest.raises(GtpError):
                eng.read_power_activities(bad)
            with pytest.raises(GtpError):
                eng.power_sliced(bad, max_events=10 ** 6)
            with pytest.raises(GtpError):
                eng.power_batch_stream([tr, bad])
        eng.read_power_activities(tr)
        res = eng.power()
        streamed, _ = eng.power_batch_stream([tr, tr])
        for r in streamed:
            assert np.array_equal(r.cycle_power, res.cycle_power) and np.array_equal(r.gate_power, res.gate_power)
    ref = oracle.run(d.library, d, tr, threads=0)
    compare(res, ref, "after rejected traces")


def test_long_clock_period_uses_the_generic_kernel():
    """100 000 ticks per cycle (100 ns at 1 ps): every cell goes to the generic
    kernel (the specialised one compares arrivals on 32 bits, below 2^15 ticks
    per cycle) and the cycle multiply-shift uses its large-period constants;
    resident, sliced and batched runs against the oracle."""
    print("[parity] long clock period", flush=True)
    d = S.make_design(300, dff_fraction=0.1, clock_ns=100.0, seed=61)
    tr = d.trace(300, toggle_p=0.3, seed=62)
    assert tr.ticks_per_cycle >= 32768
    res, ref = run_both(d, tr, tile_cycles=32)
    compare(res, ref, "long clock period")
    assert res.gate_glitch.sum() > 0
    with HipPower(tile_cycles=32) as eng:
        eng.set_design(d.library, d)
        eng.read_power_activities(tr)
        full = eng.power()
        sl = eng.power_sliced(tr, max_events=tr.n_event // 3 + 4 * d.n_net)
        both = eng.power_batch([tr, tr])
    assert sl.schedule["n_interval"] > 1
    for other, what in ((sl, "sliced"), (both[1], "batched")):
        assert np.array_equal(other.gate_glitch, full.gate_glitch), what
        assert np.array_equal(other.cycle_power, full.cycle_power), what
