"""Streamed batches (GPU): traces handed to the engine one at a time
(gtp_batch_begin / gtp_batch_add / gtp_batch_end), each packed onto the
device as it arrives, so that the caller holds one trace at a time.  Every
result must equal the one-trace run of the same trace bit for bit (exact
integer sums), across loads, with traces that need cycle intervals of their
own (larger than a load, or of 2^29 ticks or more), and match the oracle."""
import numpy as np
import pytest

import oracle
from gtpower_amd import abi
from gtpower_amd import synthetic as S
from gtpower_amd.power import LONG_TRACE_TICKS, HipPower

from test_engine_gpu import compare

pytestmark = pytest.mark.gpu


def _same(a, b, k):
    assert b.n_cycle == a.n_cycle, k
    assert np.array_equal(a.gate_toggle_x2, b.gate_toggle_x2), k
    assert np.array_equal(a.gate_glitch, b.gate_glitch), k
    assert np.array_equal(a.cycle_power, b.cycle_power), k
    assert np.array_equal(a.gate_power, b.gate_power), k
    assert np.array_equal(a.total, b.total), k


@pytest.mark.parametrize("budget", [0, 400_000, 60_000])
def test_stream_equals_individual_runs(budget):
    """Six traces of uneven length (one shorter than a tile, one of a single
    cycle); with a budget of 400k words they form several loads, with 60k the
    longest ones each run in intervals of their own.  The generator frees
    each trace as soon as the engine has taken it."""
    print(f"[stream] 1500 gates, 6 traces, event budget {budget}", flush=True)
    d = S.make_design(1500, dff_fraction=0.15, clock_ns=2.0, seed=41)
    lengths = [(700, 0.15), (37, 0.3), (1200, 0.05), (129, 0.2), (1, 0.2), (400, 0.25)]
    singles = []
    with HipPower(tile_cycles=64) as eng:
        eng.set_design(d.library, d)
        for i, (n, p) in enumerate(lengths):
            tr = d.trace(n, toggle_p=p, seed=500 + i)
            eng.read_power_activities(tr)
            singles.append(eng.power())
            tr.close()
        eng.trace = None
    held = []

    def gen():
        for i, (n, p) in enumerate(lengths):
            tr = d.trace(n, toggle_p=p, seed=500 + i)
            held.append(tr)
            assert sum(t._h is not None for t in held) == 1, "more than one trace held"
            yield tr

    with HipPower(tile_cycles=64, max_event_num=budget) as eng:
        eng.set_design(d.library, d)
        res, ms = eng.power_batch_stream(gen(), release=lambda t: t.close())
        sched = res[0].schedule
    assert ms > 0
    if budget:
        assert sched["n_interval"] >= 2, sched
    for k, (a, b) in enumerate(zip(singles, res)):
        _same(a, b, k)


def test_stream_matches_oracle_with_a_long_trace():
    """A batch of two short traces and one of 100 000-tick cycles x 6000
    cycles (t_max 6e8 ticks, evaluated in intervals by itself), against the
    oracle at the engine parity bar."""
    print("[stream] 250 gates: 2 short traces + one of 6e8 ticks, against the oracle", flush=True)
    d = S.make_design(250, dff_fraction=0.1, clock_ns=100.0, seed=71)
    trs = [d.trace(300, toggle_p=0.2, seed=11), d.trace(6000, toggle_p=0.2, seed=171),
           d.trace(45, toggle_p=0.3, seed=12)]
    assert trs[1].t_max > LONG_TRACE_TICKS
    with HipPower() as eng:
        eng.set_design(d.library, d)
        res, _ = eng.power_batch_stream(trs)
    for k, tr in enumerate(trs):
        compare(res[k], oracle.run(d.library, d, tr, threads=0), f"streamed trace {k}")


def test_stream_rejects_mixed_clocks_and_frees():
    """A trace of another clock period is refused; later adds return the same
    error and the batch still ends (and frees) cleanly."""
    d = S.make_design(200, dff_fraction=0.1, clock_ns=10.0, seed=44)
    t1 = d.trace(50, toggle_p=0.2, seed=1)
    d2 = S.make_design(200, dff_fraction=0.1, clock_ns=5.0, seed=44)
    t2 = d2.trace(50, toggle_p=0.2, seed=2)
    with HipPower() as eng:
        eng.set_design(d.library, d)
        with pytest.raises(abi.GtpError):
            eng.power_batch_stream([t1, t2, t1])
        # the context stays usable
        res, _ = eng.power_batch_stream([t1])
        eng.read_power_activities(t1)
        _same(eng.power(), res[0], 0)


def test_stream_refuses_a_design_change():
    """Loading another design while a streamed batch is open makes the next
    add fail (the batch's per-gate sums are sized for the first design)."""
    d = S.make_design(200, dff_fraction=0.1, clock_ns=10.0, seed=44)
    d2 = S.make_design(300, dff_fraction=0.1, clock_ns=10.0, seed=45)
    t1 = d.trace(50, toggle_p=0.2, seed=1)
    with HipPower() as eng:
        eng.set_design(d.library, d)

        def gen():
            yield t1
            eng.set_design(d2.library, d2)
            yield d2.trace(50, toggle_p=0.2, seed=2)

        with pytest.raises(abi.GtpError, match="design changed"):
            eng.power_batch_stream(gen())
