"""FP64 CPU restatement of the reference's time-based power path.

TEST INFRASTRUCTURE ONLY: the oracle for the HIP engine (tests/,
__graft_entry__.smoke()) and the CPU baseline of bench.py.  Product code never
imports this package.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

from gtpower_amd import abi

SO = Path(os.environ.get("GTPOWER_ORACLE_SO", Path(__file__).resolve().parent / "liboracle.so"))
SO_F32 = Path(__file__).resolve().parent / "liboracle_f32.so"


class OracleResult(C.Structure):
    _fields_ = [("gate_power", abi.p_f64), ("gate_abs", abi.p_f64), ("cycle_power", abi.p_f64),
                ("cycle_abs", abi.p_f64), ("gate_toggle_x2", abi.p_u64), ("gate_glitch", abi.p_u64),
                ("n_cycle", C.c_int64), ("wall_s", C.c_double)]


_LIB = None
_LIB_F32 = None


def lib_f32():
    """The power path built with FP32 values (energy terms, per-gate sums and
    per-cycle waveforms in float, as the reference's CPU path): the
    FP32-versus-FP64 error-budget diagnostic, never the oracle."""
    global _LIB_F32
    if _LIB_F32 is None:
        if not SO_F32.exists():
            raise RuntimeError(f"FP32 oracle build missing: {SO_F32} (python -m gtpower_amd.build --oracle)")
        L = C.CDLL(str(SO_F32))
        L.oracle_run.argtypes = [C.POINTER(abi.GtpLibrary), C.POINTER(abi.GtpDesign),
                                 C.POINTER(abi.GtpTrace), C.c_int32, C.c_int64, C.c_int64,
                                 C.POINTER(OracleResult)]
        L.oracle_run.restype = C.c_int32
        _LIB_F32 = L
    return _LIB_F32


def lib():
    global _LIB
    if _LIB is None:
        if not SO.exists():
            raise RuntimeError(f"oracle not built: {SO} (python -m gtpower_amd.build --oracle)")
        L = C.CDLL(str(SO))
        L.oracle_run.argtypes = [C.POINTER(abi.GtpLibrary), C.POINTER(abi.GtpDesign),
                                 C.POINTER(abi.GtpTrace), C.c_int32, C.c_int64, C.c_int64,
                                 C.POINTER(OracleResult)]
        L.oracle_run.restype = C.c_int32
        L.oracle_glitch_flags.argtypes = [abi.p_i64, C.c_int64, C.c_int64, abi.p_u8]
        L.oracle_glitch_flags.restype = None
        L.oracle_glitch_ratio.argtypes = [abi.p_i64, C.c_int64, abi.p_u8, C.c_int64, C.c_double,
                                          C.c_double, C.c_double]
        L.oracle_glitch_ratio.restype = C.c_double
        L.oracle_table_value.argtypes = [C.POINTER(abi.GtpLibrary), C.c_int32, C.c_double, C.c_double]
        L.oracle_table_value.restype = C.c_double
        L.oracle_cycles_per_thread.argtypes = [C.c_int64] * 4
        L.oracle_cycles_per_thread.restype = C.c_int64
        L.oracle_toggle_x2.argtypes = [C.c_int32, C.c_int32]
        L.oracle_toggle_x2.restype = C.c_int32
        L.oracle_rise_x2.argtypes = [C.c_int32, C.c_int32]
        L.oracle_rise_x2.restype = C.c_int32
        L.oracle_vcd_parse.argtypes = [C.c_char_p, C.c_int64, C.c_char_p, C.c_int64]
        L.oracle_vcd_parse.restype = C.c_void_p
        L.oracle_vcd_info.argtypes = [C.c_void_p, abi.p_i64, C.POINTER(C.c_double)]
        L.oracle_vcd_info.restype = None
        for f, r in (("name", C.c_char_p), ("id", C.c_char_p), ("width", C.c_int32), ("stream", C.c_int64)):
            fn = getattr(L, "oracle_vcd_var_" + f)
            fn.argtypes = [C.c_void_p, C.c_int64]
            fn.restype = r
        L.oracle_vcd_events.argtypes = [C.c_void_p, abi.p_i64, abi.p_i64, abi.p_u8]
        L.oracle_vcd_events.restype = None
        L.oracle_vcd_free.argtypes = [C.c_void_p]
        L.oracle_vcd_free.restype = None
        _LIB = L
    return _LIB


def run(library, design, trace, threads=0, gate_begin=0, gate_end=-1, fp32=False):
    """Evaluate the design on the trace; returns a dict of numpy arrays.
    fp32=True runs the FP32 diagnostic build instead of the FP64 oracle."""
    G = design.n_gate if hasattr(design, "n_gate") else design.as_abi().n_gate
    T = trace.ticks_per_cycle
    Cn = -(-trace.t_max // T)
    out = dict(gate_power=np.zeros(5 * G), gate_abs=np.zeros(5 * G),
               cycle_power=np.zeros(5 * Cn), cycle_abs=np.zeros(5 * Cn),
               gate_toggle_x2=np.zeros(G, np.uint64), gate_glitch=np.zeros(G, np.uint64))
    r = OracleResult(abi.ptr(out["gate_power"], C.c_double), abi.ptr(out["gate_abs"], C.c_double),
                     abi.ptr(out["cycle_power"], C.c_double), abi.ptr(out["cycle_abs"], C.c_double),
                     abi.ptr(out["gate_toggle_x2"], C.c_uint64), abi.ptr(out["gate_glitch"], C.c_uint64),
                     0, 0.0)
    lv, dv, tv = library.as_abi(), design.as_abi(), trace.as_abi()
    L = lib_f32() if fp32 else lib()
    st = L.oracle_run(C.byref(lv), C.byref(dv), C.byref(tv), threads, gate_begin, gate_end, C.byref(r))
    if st != 0:
        raise abi.GtpError(st, "oracle_run")
    assert r.n_cycle == Cn, (r.n_cycle, Cn)
    for k in ("gate_power", "gate_abs"):
        out[k] = out[k].reshape(5, G)
    for k in ("cycle_power", "cycle_abs"):
        out[k] = out[k].reshape(5, Cn)
    out["n_cycle"] = Cn
    out["wall_s"] = r.wall_s
    return out


def glitch_flags(times, ticks_per_cycle):
    t = np.ascontiguousarray(times, dtype=np.int64)
    f = np.zeros(len(t), np.uint8)
    lib().oracle_glitch_flags(abi.ptr(t, C.c_int64), len(t), ticks_per_cycle, abi.ptr(f, C.c_uint8))
    return f


def glitch_ratio(times, flags, i, time_scale, slew_rise, slew_fall):
    t = np.ascontiguousarray(times, dtype=np.int64)
    f = np.ascontiguousarray(flags, dtype=np.uint8)
    return lib().oracle_glitch_ratio(abi.ptr(t, C.c_int64), len(t), abi.ptr(f, C.c_uint8), i,
                                     time_scale, slew_rise, slew_fall)


def table_value(library, table, slew, cap):
    lv = library.as_abi()
    return lib().oracle_table_value(C.byref(lv), table, slew, cap)


def cycles_per_thread(n_event, n_cycle, e_target=8, parallelism_floor=512):
    return lib().oracle_cycles_per_thread(n_event, n_cycle, e_target, parallelism_floor)


class VcdError(Exception):
    pass


def vcd_parse(text):
    """Read a VCD text (bytes) as the reference does (oracle/vcd.cpp).

    Returns dict(vars=[(name, id, width, first stream)], offsets [S+1],
    times [E], values [E], t_max, time_scale, n_id); raises VcdError."""
    L = lib()
    buf = bytes(text)
    err = C.create_string_buffer(512)
    h = L.oracle_vcd_parse(buf, len(buf), err, 512)
    if not h:
        raise VcdError(err.value.decode())
    try:
        info = np.zeros(5, np.int64)
        ts = C.c_double()
        L.oracle_vcd_info(h, abi.ptr(info, C.c_int64), C.byref(ts))
        n_var, n_id, S, E, t_max = (int(x) for x in info)
        vars_ = [(L.oracle_vcd_var_name(h, i).decode(), L.oracle_vcd_var_id(h, i).decode(),
                  int(L.oracle_vcd_var_width(h, i)), int(L.oracle_vcd_var_stream(h, i))) for i in range(n_var)]
        off = np.zeros(S + 1, np.int64)
        times = np.zeros(max(E, 1), np.int64)
        vals = np.zeros(max(E, 1), np.uint8)
        L.oracle_vcd_events(h, abi.ptr(off, C.c_int64), abi.ptr(times, C.c_int64), abi.ptr(vals, C.c_uint8))
    finally:
        L.oracle_vcd_free(h)
    return dict(vars=vars_, offsets=off, times=times[:E], values=vals[:E], t_max=t_max,
                time_scale=ts.value, n_id=n_id)
