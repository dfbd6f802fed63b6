"""Host-side mirror of the reference's power interface over the C ABI.

The reference selects a power backend in Sta::makePower (search/Sta.cc:502-509)
and calls ``power(corner, total, sequential, combinational, clock, macro, pad)``
then ``printCellRes()`` (power/Power.hh:80-109).  ``HipPower`` offers the same
two operations on explicit inputs:

    engine = HipPower(options)
    engine.set_design(library, design)        # initGateData + LUT construction
    engine.read_power_activities(trace)       # read_power_activities -vcd (event store)
    res = engine.power()                      # report_power: totals, classes, waveforms
    engine.print_cell_res(path, names)        # cuda_cell_power.txt
    engine.print_per_cycle_results(path, top) # cuda_per_cycle_waveform.txt

There is no CPU fallback: every call goes to libgtpower_hip.so, and a missing
library or device raises.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi

# Event times at or beyond this many ticks do not fit the resident event store;
# such traces are evaluated in intervals from host memory (gtp_run_sliced).
LONG_TRACE_TICKS = (1 << 29) - 1


@dataclass
class PowerResult:
    """Totals and waveforms of one run (reference: include/sta/PowerClass.hh:95-161)."""
    gate_power: np.ndarray          # [5, G]: internal, glitch internal, switching, glitch switching, leakage
    cycle_power: np.ndarray         # [5, C]: leakage, internal, glitch internal, switching, glitch switching
    total: np.ndarray               # [5]
    class_total: np.ndarray         # [5 classes, 5]
    gate_toggle_x2: np.ndarray      # [G]
    gate_glitch: np.ndarray         # [G]
    n_cycle: int
    kernel_ms: float
    total_ms: float
    schedule: dict = field(default_factory=dict)

    @property
    def total_power(self):
        return float(self.total.sum())


def options(**kw):
    o = abi.GtpOptions()
    lib = abi.load_engine()
    lib.gtp_options_init(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class VcdClock:
    """The clock grid of a trace loaded from a VCD (what power() needs of it)."""

    def __init__(self, info, clk_period):
        self.t_max = int(info["t_max"])
        self.time_scale = float(info["time_scale"])
        self.clk_period = float(clk_period)
        self.n_event = int(info["n_event"])
        ratio = self.clk_period / self.time_scale
        T = int(np.ceil(ratio))
        if abs(ratio - round(ratio)) < 1e-9 * max(1.0, ratio):
            T = int(round(ratio))
        self.ticks_per_cycle = T
        self.n_cycle = -(-self.t_max // T)


class HipPower:
    """One engine context bound to one HIP device."""

    def __init__(self, opts=None, **kw):
        self._lib = abi.load_engine()
        self._opts = opts if opts is not None else options(**kw)
        h = C.c_void_p()
        st = self._lib.gtp_create(C.byref(self._opts), C.byref(h))
        if st != 0:
            raise abi.GtpError(st, "gtp_create", self._lib.gtp_create_error().decode())
        self._h = h
        self._keep = []
        self.n_gate = 0

    def _check(self, st, where):
        if st != 0:
            raise abi.GtpError(st, where, self._lib.gtp_last_error(self._h).decode())

    def set_design(self, library, design):
        self._keep = [library, design]
        lv = library.as_abi()
        dv = design.as_abi()
        self._check(self._lib.gtp_set_design(self._h, C.byref(lv), C.byref(dv)), "gtp_set_design")
        self.n_gate = dv.n_gate
        self.library = library
        self.design = design

    def read_power_activities(self, trace):
        """Load the trace's event store on the device; a trace of 2^29 - 1
        ticks or more stays in host memory and power() evaluates it in
        intervals (gtp_run_sliced)."""
        self._long = trace.t_max >= LONG_TRACE_TICKS
        if not self._long:
            tv = trace.as_abi()
            self._check(self._lib.gtp_set_trace(self._h, C.byref(tv)), "gtp_set_trace")
        self.trace = trace

    def read_vcd(self, vcd, clk_period):
        """Load a VCD parsed on the device (gtpower_amd.vcd.VcdText) as the
        trace: the design's nets are the VCD's (id, bit) streams."""
        self._check(self._lib.gtp_set_trace_vcd(self._h, vcd.handle, float(clk_period)), "gtp_set_trace_vcd")
        self.trace = VcdClock(vcd.info, clk_period)

    def power(self, want=("gate", "cycle", "counts")):
        if getattr(self, "_long", False):
            return self.power_sliced(self.trace, 0, want)
        G = self.n_gate
        T = self.trace.ticks_per_cycle
        Cn = -(-self.trace.t_max // T)
        gp = np.zeros(5 * G) if "gate" in want else None
        cp = np.zeros(5 * Cn) if "cycle" in want else None
        tot = np.zeros(5)
        cls = np.zeros(25)
        tg = np.zeros(G, np.uint64) if "counts" in want else None
        gl = np.zeros(G, np.uint64) if "counts" in want else None
        r = abi.GtpResult(abi.ptr(gp, C.c_double), abi.ptr(cp, C.c_double), abi.ptr(tot, C.c_double),
                          abi.ptr(cls, C.c_double), abi.ptr(tg, C.c_uint64), abi.ptr(gl, C.c_uint64),
                          0, 0.0, 0.0)
        self._check(self._lib.gtp_run(self._h, C.byref(r)), "gtp_run")
        si = abi.GtpScheduleInfo()
        self._check(self._lib.gtp_get_schedule(self._h, C.byref(si)), "gtp_get_schedule")
        sched = {k: getattr(si, k) for k, _ in abi.GtpScheduleInfo._fields_}
        self.last = PowerResult(gp.reshape(5, G) if gp is not None else None,
                                cp.reshape(5, Cn) if cp is not None else None, tot, cls.reshape(5, 5),
                                tg, gl, r.n_cycle, r.kernel_ms, r.total_ms, sched)
        return self.last

    def autotune(self):
        """Pick tile length and gate-range size by timing the loaded trace
        (gtp_autotune); returns the chosen schedule."""
        si = abi.GtpScheduleInfo()
        self._check(self._lib.gtp_autotune(self._h, C.byref(si)), "gtp_autotune")
        return {k: getattr(si, k) for k, _ in abi.GtpScheduleInfo._fields_}

    def power_batch(self, traces, want=("gate", "cycle", "counts")):
        """Evaluate several traces of the loaded design in one batched run
        (gtp_run_batch); returns one PowerResult per trace."""
        G = self.n_gate
        views, outs, rs = [], [], []
        for tr in traces:
            v = tr.as_abi()
            Cn = -(-tr.t_max // tr.ticks_per_cycle)
            o = dict(gp=np.zeros(5 * G) if "gate" in want else None,
                     cp=np.zeros(5 * Cn) if "cycle" in want else None,
                     tot=np.zeros(5), cls=np.zeros(25),
                     tg=np.zeros(G, np.uint64) if "counts" in want else None,
                     gl=np.zeros(G, np.uint64) if "counts" in want else None, C=Cn)
            views.append(v)
            outs.append(o)
            rs.append(abi.GtpResult(abi.ptr(o["gp"], C.c_double), abi.ptr(o["cp"], C.c_double),
                                    abi.ptr(o["tot"], C.c_double), abi.ptr(o["cls"], C.c_double),
                                    abi.ptr(o["tg"], C.c_uint64), abi.ptr(o["gl"], C.c_uint64), 0, 0.0, 0.0))
        tv = (abi.GtpTrace * len(views))(*views)
        rv = (abi.GtpResult * len(rs))(*rs)
        self._check(self._lib.gtp_run_batch(self._h, len(views), tv, rv), "gtp_run_batch")
        si = abi.GtpScheduleInfo()
        self._check(self._lib.gtp_get_schedule(self._h, C.byref(si)), "gtp_get_schedule")
        sched = {k: getattr(si, k) for k, _ in abi.GtpScheduleInfo._fields_}
        res = []
        for o, r in zip(outs, rv):
            res.append(PowerResult(o["gp"].reshape(5, G) if o["gp"] is not None else None,
                                   o["cp"].reshape(5, o["C"]) if o["cp"] is not None else None,
                                   o["tot"], o["cls"].reshape(5, 5), o["tg"], o["gl"], r.n_cycle, r.kernel_ms,
                                   r.total_ms, sched))
        return res

    def power_batch_stream(self, traces, want=("gate", "cycle", "counts"), release=None):
        """Evaluate traces handed over one at a time (gtp_batch_begin / add /
        end): `traces` may be a generator, each trace is packed onto the
        device when it is added and `release(trace)` (if given) is called right
        after, so that only one trace need be held at a time.  Returns (one
        PowerResult per trace, the batch's device time in ms)."""
        G = self.n_gate
        h = C.c_void_p()
        outs, rs = [], []
        it = iter(traces)
        first = next(it, None)
        if first is None:
            return [], 0.0
        self._check(self._lib.gtp_batch_begin(self._h, first.time_scale, first.clk_period, C.byref(h)),
                    "gtp_batch_begin")
        ended = False
        try:
            tr = first
            while tr is not None:
                v = tr.as_abi()
                Cn = -(-tr.t_max // tr.ticks_per_cycle)
                o = dict(gp=np.zeros(5 * G) if "gate" in want else None,
                         cp=np.zeros(5 * Cn) if "cycle" in want else None,
                         tot=np.zeros(5), cls=np.zeros(25),
                         tg=np.zeros(G, np.uint64) if "counts" in want else None,
                         gl=np.zeros(G, np.uint64) if "counts" in want else None, C=Cn)
                r = abi.GtpResult(abi.ptr(o["gp"], C.c_double), abi.ptr(o["cp"], C.c_double),
                                  abi.ptr(o["tot"], C.c_double), abi.ptr(o["cls"], C.c_double),
                                  abi.ptr(o["tg"], C.c_uint64), abi.ptr(o["gl"], C.c_uint64), 0, 0.0, 0.0)
                outs.append(o)
                rs.append(r)
                self._check(self._lib.gtp_batch_add(h, C.byref(v), C.byref(r)), "gtp_batch_add")
                if release is not None:
                    release(tr)
                tr = next(it, None)
            ms = C.c_double(0.0)
            ended = True
            self._check(self._lib.gtp_batch_end(h, C.byref(ms)), "gtp_batch_end")
        finally:
            if not ended:
                self._lib.gtp_batch_end(h, None)
        si = abi.GtpScheduleInfo()
        self._check(self._lib.gtp_get_schedule(self._h, C.byref(si)), "gtp_get_schedule")
        sched = {k: getattr(si, k) for k, _ in abi.GtpScheduleInfo._fields_}
        res = [PowerResult(o["gp"].reshape(5, G) if o["gp"] is not None else None,
                           o["cp"].reshape(5, o["C"]) if o["cp"] is not None else None,
                           o["tot"], o["cls"].reshape(5, 5), o["tg"], o["gl"], r.n_cycle, r.kernel_ms,
                           r.total_ms, sched) for o, r in zip(outs, rs)]
        return res, ms.value

    def power_sliced(self, trace, max_events=0, want=("gate", "cycle", "counts")):
        """Evaluate a host trace in cycle intervals whose device event store
        stays within max_events (gtp_run_sliced): for traces larger than
        device memory; bit-identical to power() over the whole trace."""
        G = self.n_gate
        v = trace.as_abi()
        Cn = -(-trace.t_max // trace.ticks_per_cycle)
        gp = np.zeros(5 * G) if "gate" in want else None
        cp = np.zeros(5 * Cn) if "cycle" in want else None
        tot, cls = np.zeros(5), np.zeros(25)
        tg = np.zeros(G, np.uint64) if "counts" in want else None
        gl = np.zeros(G, np.uint64) if "counts" in want else None
        r = abi.GtpResult(abi.ptr(gp, C.c_double), abi.ptr(cp, C.c_double), abi.ptr(tot, C.c_double),
                          abi.ptr(cls, C.c_double), abi.ptr(tg, C.c_uint64), abi.ptr(gl, C.c_uint64), 0, 0.0, 0.0)
        self._check(self._lib.gtp_run_sliced(self._h, C.byref(v), int(max_events), C.byref(r)), "gtp_run_sliced")
        si = abi.GtpScheduleInfo()
        self._check(self._lib.gtp_get_schedule(self._h, C.byref(si)), "gtp_get_schedule")
        sched = {k: getattr(si, k) for k, _ in abi.GtpScheduleInfo._fields_}
        return PowerResult(gp.reshape(5, G) if gp is not None else None, cp.reshape(5, Cn) if cp is not None else None,
                           tot, cls.reshape(5, 5), tg, gl, r.n_cycle, r.kernel_ms, r.total_ms, sched)

    @property
    def n_cycle(self):
        T = self.trace.ticks_per_cycle
        return -(-self.trace.t_max // T)

    def power_tiles(self, tile_cycles, tiles):
        """Exact per-cycle power of the cycles of `tiles` (tile j = cycles
        [j*tile_cycles, (j+1)*tile_cycles)): returns (cycles [n], waves [5, n], ms),
        cycles past the end of the trace dropped."""
        tiles = np.ascontiguousarray(tiles, dtype=np.int32)
        n_out = len(tiles) * tile_cycles
        cp = np.zeros(5 * n_out)
        ms = C.c_double()
        self._check(self._lib.gtp_run_tiles(self._h, int(tile_cycles), len(tiles), abi.ptr(tiles, C.c_int32),
                                            abi.ptr(cp, C.c_double), C.byref(ms)), "gtp_run_tiles")
        cyc = (tiles.astype(np.int64)[:, None] * tile_cycles + np.arange(tile_cycles)[None, :]).reshape(-1)
        keep = cyc < self.n_cycle
        return cyc[keep], cp.reshape(5, n_out)[:, keep], ms.value

    def feature_correlation(self, y, mask=None):
        """Pearson correlation of each feature's per-cycle bit with y, indexed by
        feature id = kind * n_net + net (kind 0: value change, 1: rise)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        assert y.shape == (self.n_cycle,)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        corr = np.zeros(2 * self.design.n_net, np.float32)
        self._check(self._lib.gtp_feature_correlation(self._h, abi.ptr(y, C.c_double), abi.ptr(m, C.c_uint8),
                                                      abi.ptr(corr, C.c_float)), "gtp_feature_correlation")
        return corr

    def toggle_features(self, nets, cycle_begin=0, n_cycle=None):
        """Per-cycle bits of the features `nets` (ids kind * n_net + net):
        uint32 [n_cycle][len(nets)/32]."""
        nets = np.ascontiguousarray(nets, dtype=np.int32)
        n = self.n_cycle - cycle_begin if n_cycle is None else n_cycle
        bits = np.zeros((n, len(nets) // 32), np.uint32)
        self._check(self._lib.gtp_toggle_features(self._h, len(nets), abi.ptr(nets, C.c_int32), int(cycle_begin),
                                                  int(n), abi.ptr(bits, C.c_uint32)), "gtp_toggle_features")
        return bits

    def net_event_counts(self):
        """Events per net of the loaded trace (int64 [n_net])."""
        out = np.zeros(self.design.n_net, np.int64)
        self._check(self._lib.gtp_net_event_counts(self._h, abi.ptr(out, C.c_int64)), "gtp_net_event_counts")
        return out

    def net_weight_sum(self, nets, weights, cycle_begin=0, n_cycle=None):
        """Per cycle, sum over k of weights[k] x value changes of nets[k]
        (device, exact to 2^-20 of the largest weight); returns (sums, ms)."""
        nets = np.ascontiguousarray(nets, dtype=np.int32)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        assert nets.shape == w.shape
        n = self.n_cycle - cycle_begin if n_cycle is None else n_cycle
        out = np.zeros(n, np.float64)
        ms = C.c_double()
        self._check(self._lib.gtp_net_weight_sum(self._h, len(nets), abi.ptr(nets, C.c_int32), abi.ptr(w, C.c_double),
                                                 int(cycle_begin), int(n), abi.ptr(out, C.c_double), C.byref(ms)),
                    "gtp_net_weight_sum")
        return out, ms.value

    def print_cell_res(self, path, gate_names=None):
        """Write cuda_cell_power.txt (power/cuda/CudaPower.cu:3134-3165)."""
        res = self.last
        gp = np.ascontiguousarray(res.gate_power.reshape(-1))
        names = None
        if gate_names is not None:
            names = b"".join(n.encode() + b"\0" for n in gate_names)
        cells = b"".join(n.encode() + b"\0" for n in self.library.full_names) + b"\0"
        gc = np.ascontiguousarray(self.design.gate_cell, dtype=np.int32)
        st = self._lib.gtp_write_cell_power(str(path).encode(), self.n_gate, abi.ptr(gp, C.c_double), names,
                                            abi.ptr(gc, C.c_int32), cells)
        if st != 0:
            raise abi.GtpError(st, "gtp_write_cell_power")

    def print_per_cycle_results(self, path, top_name="top"):
        """Write the per-cycle waveform file (include/sta/PowerClass.hh:123-149)."""
        cp = np.ascontiguousarray(self.last.cycle_power.reshape(-1))
        st = self._lib.gtp_write_per_cycle(str(path).encode(), top_name.encode(), self.last.n_cycle,
                                           abi.ptr(cp, C.c_double))
        if st != 0:
            raise abi.GtpError(st, "gtp_write_per_cycle")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.gtp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def reference_cycles_per_thread(n_event, n_cycle, e_target=8, parallelism_floor=512):
    """The reference's per-gate cycles-per-thread (power/cuda/Gate.cuh:354-376)."""
    lib = abi.load_engine()
    ne = np.ascontiguousarray(n_event, dtype=np.int64)
    out = np.zeros(len(ne), dtype=np.int64)
    st = lib.gtp_reference_cycles_per_thread(len(ne), abi.ptr(ne, C.c_int64), n_cycle, e_target,
                                             parallelism_floor, abi.ptr(out, C.c_int64))
    if st != 0:
        raise abi.GtpError(st, "gtp_reference_cycles_per_thread")
    return out
