"""Traces of 2^29 ticks or more (GPU): evaluated in intervals whose event
words hold times relative to each interval's origin (gtp_run_sliced; the
Python HipPower routes such traces there by itself), against the oracle at the
engine parity bar (counts exact, values within 1e-6 of the term magnitudes).
The reference keeps int64 event times (util/Types.hh:16), so any length is
valid input."""
import numpy as np
import pytest

import oracle
from gtpower_amd import synthetic as S
from gtpower_amd.power import LONG_TRACE_TICKS, HipPower

from test_engine_gpu import _mux_design, compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("parts", [1, 5])
def test_long_trace_random_design(parts):
    """100 000-tick cycles x 6000 cycles: t_max 6e8 ticks."""
    print(f"[long] 250 gates x 6000 cycles of 100000 ticks, event budget 1/{parts}", flush=True)
    d = S.make_design(250, dff_fraction=0.1, clock_ns=100.0, seed=71)
    tr = d.trace(6000, toggle_p=0.2, seed=171)
    assert tr.t_max > LONG_TRACE_TICKS
    with HipPower() as eng:
        eng.set_design(d.library, d)
        if parts == 1:
            eng.read_power_activities(tr)
            res = eng.power()
        else:
            res = eng.power_sliced(tr, max_events=tr.n_event // parts + 4 * d.n_net)
        assert res.schedule["n_interval"] >= 2
    ref = oracle.run(d.library, d, tr, threads=0)
    compare(res, ref, f"long trace, 1/{parts}")
    assert res.gate_glitch.sum() > 0


def test_long_trace_quiet_spans_and_glitch_groups(library):
    """Pins quiet for thousands of cycles on both sides of interval
    boundaries, glitch groups just before and after them, and events at the
    trace end: the events next to an interval that its load moves closer must
    leave every result unchanged."""
    print("[long] hand-built: quiet spans, glitch groups next to interval edges", flush=True)
    T = 1000
    n_cycle = 600_000   # t_max 6e8 ticks
    rng = np.random.default_rng(5)
    waves = []
    for net in range(5):
        w, v = [(0, net & 1)], net & 1
        t = 0
        while True:
            # bursts: a glitch group of 2-4 events in one cycle, then a quiet
            # gap of 1 to 40 000 cycles
            c = t // T + int(rng.integers(1, 40_000))
            if c >= n_cycle:
                break
            k = int(rng.integers(1, 5))
            offs = np.sort(rng.choice(np.arange(10, T - 10), size=k, replace=False))
            for o in offs:
                v ^= 1
                w.append((c * T + int(o), v))
            t = c * T + T
        waves.append(w)
    waves[3].append((n_cycle * T, 1 - waves[3][-1][1]))   # an event at t_max
    nets = [0, 1, 2, 3, 1, 2, 4, 4]
    d = _mux_design(library, nets, len(waves))
    tr = S.trace_from_lists(waves, n_cycle * T, 1e-12, T * 1e-12)
    assert tr.t_max > LONG_TRACE_TICKS
    ref = oracle.run(library, d, tr, threads=1)
    with HipPower(tile_cycles=512) as eng:
        eng.set_design(library, d)
        for budget in (0, tr.n_event // 3):
            res = eng.power_sliced(tr, max_events=budget + (4 * d.n_net if budget else 0))
            assert res.schedule["n_interval"] >= 2
            compare(res, ref, f"hand long trace, budget {budget}")
    assert ref["gate_glitch"].sum() > 0


def test_long_trace_refused_by_the_resident_store():
    """gtp_set_trace still refuses it (its words hold absolute times)."""
    import ctypes as C
    from gtpower_amd.abi import GtpError
    d = S.make_design(50, dff_fraction=0.1, clock_ns=100.0, seed=73)
    tr = d.trace(5400, toggle_p=0.1, seed=173)
    with HipPower() as eng:
        eng.set_design(d.library, d)
        tv = tr.as_abi()
        st = eng._lib.gtp_set_trace(eng._h, C.byref(tv))
        assert st == -5 and b"gtp_run_sliced" in eng._lib.gtp_last_error(eng._h)
