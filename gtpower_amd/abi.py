"""ctypes mirror of include/gtpower.h, and the loader of the built libraries.

The structures below are field-for-field copies of the C declarations; a test
checks their sizes against the compiled library.  `load_engine()` loads the
HIP engine (`libgtpower_hip.so`) and fails loudly when it was not built: the
engine has no CPU fallback.
"""
import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
ENGINE_SO = PKG_DIR / "libgtpower_hip.so"
SYNTH_SO = PKG_DIR / "libgtpower_synth.so"

GTP_OK = 0
STATUS_NAMES = {0: "GTP_OK", -1: "GTP_ERR_ARG", -2: "GTP_ERR_DEVICE", -3: "GTP_ERR_OOM",
                -4: "GTP_ERR_STATE", -5: "GTP_ERR_UNSUPPORTED", -6: "GTP_ERR_LIMIT"}

N_COMPONENT = 5
N_WAVE = 5
N_CLASS = 5
COMPONENTS = ("internal", "glitch_internal", "switching", "glitch_switching", "leakage")
WAVES = ("leakage", "internal", "glitch_internal", "switching", "glitch_switching")
CLASS_COMB, CLASS_SEQ, CLASS_CLOCK, CLASS_MACRO, CLASS_PAD = range(5)
AXIS_SLEW, AXIS_CAP = 0, 1
MAX_CELL_PINS = 16
MAX_WHEN_PINS = 6        # internal-power `when` tables: 2^6 states in a u64

p_i32 = C.POINTER(C.c_int32)
p_u32 = C.POINTER(C.c_uint32)
p_i64 = C.POINTER(C.c_int64)
p_u64 = C.POINTER(C.c_uint64)
p_f64 = C.POINTER(C.c_double)
p_i8 = C.POINTER(C.c_int8)
p_u8 = C.POINTER(C.c_uint8)


class GtpOptions(C.Structure):
    _fields_ = [("device", C.c_int32), ("tile_cycles", C.c_int32), ("block_lanes", C.c_int32),
                ("gates_per_wg", C.c_int32), ("autotune", C.c_int32), ("host_threads", C.c_int32),
                ("max_event_num", C.c_int64), ("round_tiles", C.c_int32), ("pad_", C.c_int32)]


class GtpTable(C.Structure):
    _fields_ = [("n1", C.c_int32), ("n2", C.c_int32), ("var1", C.c_int32), ("var2", C.c_int32),
                ("axis1", C.c_int64), ("axis2", C.c_int64), ("values", C.c_int64)]


class GtpInternalPower(C.Structure):
    _fields_ = [("cell", C.c_int32), ("pin", C.c_int32), ("related_pin", C.c_int32),
                ("rise_table", C.c_int32), ("fall_table", C.c_int32), ("pad_", C.c_int32),
                ("when_truth", C.c_uint64)]


class GtpLibrary(C.Structure):
    _fields_ = [("n_cell", C.c_int32), ("cell_n_pin", p_i32), ("cell_n_in", p_i32),
                ("cell_class", p_i32), ("cell_clock_mask", p_u32), ("cell_leak_offset", p_i64),
                ("leak_values", p_f64), ("cell_leak_default", p_f64), ("n_table", C.c_int32),
                ("tables", C.POINTER(GtpTable)), ("table_data", p_f64), ("n_ipower", C.c_int32),
                ("ipowers", C.POINTER(GtpInternalPower))]


class GtpDesign(C.Structure):
    _fields_ = [("n_gate", C.c_int64), ("n_pin", C.c_int64), ("n_net", C.c_int64),
                ("gate_cell", p_i32), ("gate_pin_offset", p_i64), ("gate_arc_offset", p_i64),
                ("arc_delay", p_f64), ("pin_net", p_i32), ("pin_slew", p_f64),
                ("pin_load_cap", p_f64), ("pin_voltage", p_f64), ("pin_default_state", p_i8)]


class GtpTrace(C.Structure):
    _fields_ = [("n_net", C.c_int64), ("net_event_offset", p_i64), ("event_time", p_i64),
                ("event_value", p_u8), ("t_max", C.c_int64), ("time_scale", C.c_double),
                ("clk_period", C.c_double)]


class GtpResult(C.Structure):
    _fields_ = [("gate_power", p_f64), ("cycle_power", p_f64), ("total", p_f64),
                ("class_total", p_f64), ("gate_toggle_x2", p_u64), ("gate_glitch", p_u64),
                ("n_cycle", C.c_int64), ("kernel_ms", C.c_double), ("total_ms", C.c_double)]


class GtpScheduleInfo(C.Structure):
    _fields_ = [("tile_cycles", C.c_int32), ("block_lanes", C.c_int32), ("gates_per_wg", C.c_int32),
                ("n_launch", C.c_int32), ("n_interval", C.c_int64), ("n_unit", C.c_int64),
                ("n_workgroup", C.c_int64), ("n_tile", C.c_int64), ("n_event", C.c_int64),
                ("n_pin_event", C.c_int64), ("autotune_ms", C.c_double)]


class GtpMlpConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("n_in", C.c_int32), ("n_hidden", C.c_int32),
                ("pad_", C.c_int32), ("seed", C.c_uint64)]


class GtpTrainConfig(C.Structure):
    _fields_ = [("batch", C.c_int32), ("epochs", C.c_int32), ("lr", C.c_float),
                ("weight_decay", C.c_float), ("seed", C.c_uint64), ("val_rows", C.c_int32),
                ("input_dropout", C.c_int32)]


class GtpVcdInfo(C.Structure):
    _fields_ = [("n_var", C.c_int64), ("n_id", C.c_int64), ("n_stream", C.c_int64), ("n_event", C.c_int64),
                ("t_max", C.c_int64), ("time_scale", C.c_double), ("n_token", C.c_int64),
                ("n_change", C.c_int64), ("body_bytes", C.c_int64), ("device_ms", C.c_double),
                ("total_ms", C.c_double), ("n_part", C.c_int64)]


class GtpVcdVarInfo(C.Structure):
    _fields_ = [("name", C.c_char_p), ("id", C.c_char_p), ("width", C.c_int32), ("id_index", C.c_int32),
                ("first_stream", C.c_int64)]


class GtpError(RuntimeError):
    def __init__(self, status, where, text=""):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {text}")


def ptr(arr, ctype):
    """Pointer to a contiguous numpy array (kept alive by the caller)."""
    if arr is None:
        return None
    assert arr.flags["C_CONTIGUOUS"], "array must be contiguous"
    return arr.ctypes.data_as(C.POINTER(ctype))


_ENGINE = None


def load_engine():
    """Load libgtpower_hip.so.  Raises when it is missing: there is no fallback."""
    global _ENGINE
    if _ENGINE is not None:
        return _ENGINE
    path = Path(os.environ.get("GTPOWER_ENGINE_SO", ENGINE_SO))
    if not path.exists():
        raise RuntimeError(f"HIP engine not built: {path} is missing "
                           "(run __graft_entry__.build() or gtpower_amd/build.py)")
    lib = C.CDLL(str(path))
    lib.gtp_options_init.argtypes = [C.POINTER(GtpOptions)]
    lib.gtp_options_init.restype = None
    lib.gtp_create.argtypes = [C.POINTER(GtpOptions), C.POINTER(C.c_void_p)]
    lib.gtp_destroy.argtypes = [C.c_void_p]
    lib.gtp_destroy.restype = None
    lib.gtp_last_error.argtypes = [C.c_void_p]
    lib.gtp_last_error.restype = C.c_char_p
    lib.gtp_create_error.argtypes = []
    lib.gtp_create_error.restype = C.c_char_p
    lib.gtp_set_design.argtypes = [C.c_void_p, C.POINTER(GtpLibrary), C.POINTER(GtpDesign)]
    lib.gtp_set_trace.argtypes = [C.c_void_p, C.POINTER(GtpTrace)]
    lib.gtp_run.argtypes = [C.c_void_p, C.POINTER(GtpResult)]
    lib.gtp_run_batch.argtypes = [C.c_void_p, C.c_int32, C.POINTER(GtpTrace), C.POINTER(GtpResult)]
    lib.gtp_run_sliced.argtypes = [C.c_void_p, C.POINTER(GtpTrace), C.c_int64, C.POINTER(GtpResult)]
    lib.gtp_batch_begin.argtypes = [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_void_p)]
    lib.gtp_batch_add.argtypes = [C.c_void_p, C.POINTER(GtpTrace), C.POINTER(GtpResult)]
    lib.gtp_batch_end.argtypes = [C.c_void_p, p_f64]
    lib.gtp_get_schedule.argtypes = [C.c_void_p, C.POINTER(GtpScheduleInfo)]
    lib.gtp_reference_cycles_per_thread.argtypes = [C.c_int64, p_i64, C.c_int64, C.c_int64,
                                                    C.c_int64, p_i64]
    lib.gtp_write_cell_power.argtypes = [C.c_char_p, C.c_int64, p_f64, C.c_char_p, p_i32, C.c_char_p]
    lib.gtp_write_per_cycle.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, p_f64]
    lib.gtp_autotune.argtypes = [C.c_void_p, C.POINTER(GtpScheduleInfo)]
    lib.gtp_run_tiles.argtypes = [C.c_void_p, C.c_int32, C.c_int32, p_i32, p_f64, p_f64]
    p_f32 = C.POINTER(C.c_float)
    lib.gtp_mlp_create.argtypes = [C.POINTER(GtpMlpConfig), C.POINTER(C.c_void_p)]
    lib.gtp_mlp_destroy.argtypes = [C.c_void_p]
    lib.gtp_mlp_destroy.restype = None
    lib.gtp_mlp_last_error.argtypes = [C.c_void_p]
    lib.gtp_mlp_last_error.restype = C.c_char_p
    lib.gtp_mlp_param_count.argtypes = [C.c_void_p]
    lib.gtp_mlp_param_count.restype = C.c_int64
    lib.gtp_mlp_get_params.argtypes = [C.c_void_p, p_f32, p_f32]
    lib.gtp_mlp_set_params.argtypes = [C.c_void_p, p_f32, p_f32]
    lib.gtp_keep_mask.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    lib.gtp_keep_mask.restype = C.c_uint32
    lib.gtp_mlp_gradients.argtypes = [C.c_void_p, p_u32, p_f32, C.c_int32, C.c_uint32, p_f32, p_f32]
    lib.gtp_mlp_train.argtypes = [C.c_void_p, p_u32, p_f32, C.c_int64, C.POINTER(GtpTrainConfig), p_f32]
    lib.gtp_mlp_predict.argtypes = [C.c_void_p, p_u32, C.c_int64, p_f32]
    lib.gtp_mlp_predict_trace.argtypes = [C.c_void_p, C.c_void_p, p_i32, C.c_int64, C.c_int64, p_f32, p_f64]
    lib.gtp_feature_correlation.argtypes = [C.c_void_p, p_f64, p_u8, p_f32]
    lib.gtp_toggle_features.argtypes = [C.c_void_p, C.c_int32, p_i32, C.c_int64, C.c_int64, p_u32]
    lib.gtp_net_event_counts.argtypes = [C.c_void_p, p_i64]
    lib.gtp_net_weight_sum.argtypes = [C.c_void_p, C.c_int32, p_i32, p_f64, C.c_int64, C.c_int64, p_f64,
                                       C.POINTER(C.c_double)]
    lib.gtp_mlp_gemm_test.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, p_f32]
    lib.gtp_vcd_create.argtypes = [C.c_int32, C.POINTER(C.c_void_p)]
    lib.gtp_vcd_destroy.argtypes = [C.c_void_p]
    lib.gtp_vcd_destroy.restype = None
    lib.gtp_vcd_last_error.argtypes = [C.c_void_p]
    lib.gtp_vcd_last_error.restype = C.c_char_p
    lib.gtp_vcd_parse.argtypes = [C.c_void_p, C.c_char_p, C.c_int64, C.POINTER(GtpVcdInfo)]
    lib.gtp_vcd_var.argtypes = [C.c_void_p, C.c_int64, C.POINTER(GtpVcdVarInfo)]
    lib.gtp_vcd_events.argtypes = [C.c_void_p, p_i64, p_i64, p_u8]
    lib.gtp_set_trace_vcd.argtypes = [C.c_void_p, C.c_void_p, C.c_double]
    for name in ("gtp_create", "gtp_set_design", "gtp_set_trace", "gtp_run", "gtp_run_batch", "gtp_run_sliced",
                 "gtp_get_schedule", "gtp_reference_cycles_per_thread", "gtp_write_cell_power",
                 "gtp_write_per_cycle", "gtp_run_tiles", "gtp_autotune", "gtp_mlp_create", "gtp_mlp_get_params",
                 "gtp_mlp_set_params", "gtp_mlp_gradients", "gtp_mlp_train", "gtp_mlp_predict",
                 "gtp_mlp_predict_trace", "gtp_feature_correlation", "gtp_toggle_features", "gtp_net_weight_sum",
                 "gtp_net_event_counts",
                 "gtp_mlp_gemm_test", "gtp_vcd_create", "gtp_vcd_parse", "gtp_vcd_var", "gtp_vcd_events",
                 "gtp_set_trace_vcd"):
        getattr(lib, name).restype = C.c_int32
    _ENGINE = lib
    return lib


_SYNTH = None


def load_synth():
    """Load the synthetic design/trace generator library (CPU only)."""
    global _SYNTH
    if _SYNTH is not None:
        return _SYNTH
    path = Path(os.environ.get("GTPOWER_SYNTH_SO", SYNTH_SO))
    if not path.exists():
        raise RuntimeError(f"generator not built: {path} is missing (run gtpower_amd/build.py)")
    _SYNTH = C.CDLL(str(path))
    return _SYNTH
