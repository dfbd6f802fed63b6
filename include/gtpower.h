/*
 * gtpower.h -- C ABI of the MI355X time-based (per-cycle) power engine.
 *
 * This is the drop-in boundary for GTPower's time-based power hot path.  In the
 * reference the path is reached through the gtpower::Power virtual interface
 *   virtual void Power::power(const Corner*, PowerResult& total, ...)   power/Power.hh:80-87
 *   virtual void Power::printCellRes() const                            power/Power.hh:109
 * which CudaPower overrides (power/cuda/CudaPower.hh:24-33) and Sta::makePower
 * selects (search/Sta.cc:502-509).  Its inputs are implicit OpenSTA state
 * (netlist, slews, load caps, arc delays, Liberty tables, VCD waveforms).  Here
 * every input is an explicit, caller-owned flat array, and every output goes to
 * caller-owned buffers.
 *
 * Conventions
 *   - Plain C types only; no exceptions cross the ABI; functions return a
 *     gtp_status (0 = ok, < 0 = error) and never call exit().  The text of the
 *     last error of a context is returned by gtp_last_error().
 *   - The caller owns every input and output array; the library copies what it
 *     keeps and owns its device memory.
 *   - One context per host thread.  A context is bound to one HIP device.
 *   - Units are SI: seconds, farads, volts, joules, watts.  Event times are
 *     integer ticks of `time_scale` seconds (the VCD $timescale).
 *   - Logic values are 0, 1, 2 (X) and 3 (Z), as the reference's VCD value map
 *     (power/Power.hh:114).
 */
#ifndef GTPOWER_H
#define GTPOWER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GTP_ABI_VERSION 2

typedef int32_t gtp_status;
enum {
  GTP_OK = 0,
  GTP_ERR_ARG = -1,         /* malformed input (sizes, offsets, values) */
  GTP_ERR_DEVICE = -2,      /* a HIP runtime call failed */
  GTP_ERR_OOM = -3,         /* device or host allocation failed, or the event budget is exceeded */
  GTP_ERR_STATE = -4,       /* call order violated (e.g. run before set_trace) */
  GTP_ERR_UNSUPPORTED = -5, /* valid input outside what this library implements */
  GTP_ERR_LIMIT = -6        /* a size limit of this library is exceeded */
};

/* Power components, in the column order of the reference's cell report
 * (power/cuda/CudaPower.cu:3134-3165): regular internal (glitch excluded),
 * glitch internal, regular switching (glitch excluded), glitch switching,
 * leakage. */
enum { GTP_INTERNAL = 0, GTP_GLITCH_INTERNAL = 1, GTP_SWITCHING = 2,
       GTP_GLITCH_SWITCHING = 3, GTP_LEAKAGE = 4, GTP_N_COMPONENT = 5 };

/* Per-cycle waveforms, in the section order of the reference's per-cycle file
 * (include/sta/PowerClass.hh:123-149): leakage, internal (glitch included),
 * glitch internal, switching (glitch included), glitch switching. */
enum { GTP_WAVE_LEAKAGE = 0, GTP_WAVE_INTERNAL = 1, GTP_WAVE_GLITCH_INTERNAL = 2,
       GTP_WAVE_SWITCHING = 3, GTP_WAVE_GLITCH_SWITCHING = 4, GTP_N_WAVE = 5 };

/* Instance classes for the grouped totals of report_power
 * (power/cuda/CudaPower.cu:3037-3132 finalizeResult). */
enum { GTP_CLASS_COMBINATIONAL = 0, GTP_CLASS_SEQUENTIAL = 1, GTP_CLASS_CLOCK = 2,
       GTP_CLASS_MACRO = 3, GTP_CLASS_PAD = 4, GTP_N_CLASS = 5 };

/* Options.  Mirrors the power flags of the reference CLI
 * (app/StaMain.cc:71-249, util/GlobalConfig.hh:16-46) that apply to this path. */
typedef struct {
  int32_t device;            /* HIP device ordinal; -1 = the current device */
  int32_t tile_cycles;       /* cycles per work tile; 0 = chosen by the autotuner */
  int32_t block_lanes;       /* lanes (gates) per workgroup pass; 0 = autotuned */
  int32_t gates_per_wg;      /* gates a workgroup walks per tile; 0 = autotuned */
  int32_t autotune;          /* 1 = time candidate schedules on the loaded trace */
  int32_t host_threads;      /* host worker threads (-multi_thread_number); 0 = all cores */
  int64_t max_event_num;     /* device event budget (-max_event_num); 0 = no limit */
  int32_t round_tiles;       /* batched runs: work tiles per round; 0 = sized by device memory */
  int32_t pad_;
} gtp_options;

/* A Liberty NLDM table, 1-D (n2 == 0) or 2-D, row-major values[i1 * n2 + i2].
 * Offsets index gtp_library.table_data.  var1/var2 name what each axis is
 * (reference: liberty/InternalPower.cc:178-229 axisValue). */
enum { GTP_AXIS_SLEW = 0, GTP_AXIS_CAP = 1 };
typedef struct {
  int32_t n1, n2;
  int32_t var1, var2;
  int64_t axis1, axis2, values;
} gtp_table;

/* One Liberty internal_power group of a cell pin.  Groups of a (cell, pin) are
 * listed in Liberty order: the first one whose `when` holds is used
 * (reference: power/TimeBasedPower.cc:373-378, :489-502).
 * when_truth bit s is set when the condition holds for gate state index s,
 * the state index being sum over cell pins p of (value_p == 1) << p
 * (reference: power/PowerUtils.cc:39-48, :155-187).  Cells of more than 6
 * pins take unconditional groups only (when_truth = all ones); a `when` there
 * is GTP_ERR_UNSUPPORTED. */
typedef struct {
  int32_t cell;
  int32_t pin;           /* cell pin index: inputs first, then outputs */
  int32_t related_pin;   /* cell pin index of the related input; -1 = none */
  int32_t rise_table;    /* index into tables[]; -1 = none */
  int32_t fall_table;
  int32_t pad_;
  uint64_t when_truth;
} gtp_internal_power;

/* The cell library, flattened.  Pins of a cell are ordered inputs first, then
 * outputs, as power/TimeBasedPower.cc:223-256 orders them.  Cells may have at
 * most GTP_MAX_CELL_PINS pins, the reference's limit for state-indexed
 * leakage and internal power (util/GlobalConfig.hh:42-43); its cells of more
 * pins are counted as fallbacks (power/TimeBasedPower.cc:100-102). */
#define GTP_MAX_CELL_PINS 16
typedef struct {
  int32_t n_cell;
  const int32_t* cell_n_pin;        /* [n_cell] */
  const int32_t* cell_n_in;         /* [n_cell] */
  const int32_t* cell_class;        /* [n_cell] GTP_CLASS_* */
  const uint32_t* cell_clock_mask;  /* [n_cell] bit p: pin p is a clock pin (no glitch flags) */
  /* State-indexed leakage (reference getLeakagePower, power/TimeBasedPower.cc:258-301):
   * cell c uses leak_values[cell_leak_offset[c] .. +2^n_pin) when the cell has
   * leakage_power groups, else its default below.  */
  const int64_t* cell_leak_offset;  /* [n_cell + 1] */
  const double* leak_values;
  const double* cell_leak_default;  /* [n_cell] cell_leakage_power, or 0 when absent */
  int32_t n_table;
  const gtp_table* tables;
  const double* table_data;
  int32_t n_ipower;
  const gtp_internal_power* ipowers;
} gtp_library;

/* The design: gates (leaf instances), their pins and the pins' nets.
 * Pins of gate g are pins[gate_pin_offset[g] .. gate_pin_offset[g+1]) in the
 * cell's pin order.  Arc delays of gate g are
 *   arc_delay[gate_arc_offset[g] + ((o * n_in + i) * 2 + from_rf) * 2 + to_rf]
 * for output o (0-based among outputs), input i, rf 0 = rise, 1 = fall, in
 * seconds; NaN = no such timing arc (reference: Gate::cellArcDelay,
 * power/cuda/Gate.cuh:344-348, and graph_->gateEdgeArc,
 * power/TimeBasedPower.cc:432-446). */
typedef struct {
  int64_t n_gate;
  int64_t n_pin;
  int64_t n_net;
  const int32_t* gate_cell;         /* [n_gate] */
  const int64_t* gate_pin_offset;   /* [n_gate + 1] */
  const int64_t* gate_arc_offset;   /* [n_gate + 1] */
  const double* arc_delay;
  const int32_t* pin_net;           /* [n_pin] net index, -1 = no activity */
  const double* pin_slew;           /* [2 * n_pin] rise, fall slew per pin (s); +inf allowed */
  const double* pin_load_cap;       /* [n_pin] total load capacitance seen by an output pin (F) */
  const double* pin_voltage;        /* [n_pin] supply voltage of an output pin (V) */
  const int8_t* pin_default_state;  /* [n_pin] state of a pin without activity (0,1,2) */
} gtp_design;

/* One activity trace as per-net event lists (compressed sparse rows).
 * Events of net n are [net_event_offset[n], net_event_offset[n+1]) in strictly
 * increasing time.  As the reference's VCD model guarantees
 * (power/Vcd.cc:203-216), consecutive events of a net carry different values
 * and a net whose first change is after 0 starts with an X at time 0. */
typedef struct {
  int64_t n_net;
  const int64_t* net_event_offset;  /* [n_net + 1] */
  const int64_t* event_time;        /* [n_event] ticks */
  const uint8_t* event_value;       /* [n_event] 0,1,2,3 */
  int64_t t_max;                    /* end of the trace (ticks); the VCD's last time */
  double time_scale;                /* seconds per tick */
  double clk_period;                /* seconds */
} gtp_trace;

/* Results.  Every pointer may be NULL when that output is not wanted. */
typedef struct {
  double* gate_power;        /* [GTP_N_COMPONENT * n_gate], component-major */
  double* cycle_power;       /* [GTP_N_WAVE * n_cycle], waveform-major */
  double* total;             /* [GTP_N_COMPONENT] design totals */
  double* class_total;       /* [GTP_N_CLASS * GTP_N_COMPONENT] */
  uint64_t* gate_toggle_x2;  /* [n_gate] sum over pin events of 2 * n_toggle */
  uint64_t* gate_glitch;     /* [n_gate] pin events with a glitch scaling ratio > 0 */
  int64_t n_cycle;           /* out: ceil(t_max / ticks_per_cycle) */
  double kernel_ms;          /* out: device time of the power evaluation */
  double total_ms;           /* out: device time of the whole run call */
} gtp_result;

typedef struct gtp_ctx gtp_ctx;

/* Fill `opt` with defaults. */
void gtp_options_init(gtp_options* opt);

/* Create a context on opt->device.  Replaces CudaPower's constructor plus
 * utils::cuda::setDevice (util/cuda/DeviceSetting.cu:6-15). */
gtp_status gtp_create(const gtp_options* opt, gtp_ctx** out);
void gtp_destroy(gtp_ctx* ctx);
const char* gtp_last_error(const gtp_ctx* ctx);
/* Error text of the last failed gtp_create on this thread. */
const char* gtp_create_error(void);

/* Load the cell library and the design.  Replaces CudaPower::initGateData
 * (power/cuda/CudaPower.cu:1601-1816) and the LUT construction of
 * getInputInternalPower / getOutputInternalPower (:2347-2624). */
gtp_status gtp_set_design(gtp_ctx* ctx, const gtp_library* lib, const gtp_design* design);

/* Load one activity trace and build the device event store.  Replaces
 * getGateWaveformRangeMultiThread + copyWaveformToDeviceSide
 * (power/cuda/CudaPower.cu:1874-2130) and setPeriodAndInitPerCyclePower
 * (:2656-2701). */
gtp_status gtp_set_trace(gtp_ctx* ctx, const gtp_trace* trace);

/* Evaluate per-gate and per-cycle power of the loaded trace.  Replaces
 * CudaPower::power (power/cuda/CudaPower.cu:1227-1337): schedule,
 * kernel1AllPowerCalculationTimeRangePartitionedByCycleUnit (:317-584),
 * kernel2MergeAllPowerResult (:586-606), recordResult and finalizeResult
 * (:2996-3132). */
gtp_status gtp_run(gtp_ctx* ctx, gtp_result* res);

/* Exact per-cycle power of a subset of the cycles: the tiles tiles[0..n_tiles)
 * of tile_cycles cycles each (tile j covers cycles [j*tile_cycles,
 * (j+1)*tile_cycles)).  cycle_power is [GTP_N_WAVE][n_tiles * tile_cycles],
 * column l*tile_cycles + i holding cycle tiles[l]*tile_cycles + i (0 past the
 * last cycle); the values equal those gtp_run returns for the same cycles.
 * This is the judge of the learned surrogate.  No counterpart in the
 * reference (new method). */
gtp_status gtp_run_tiles(gtp_ctx* ctx, int32_t tile_cycles, int32_t n_tiles, const int32_t* tiles,
                         double* cycle_power, double* kernel_ms);

/* Evaluate several traces of one design (same time scale and clock) in
 * persistent launches over all their tiles.  traces[k] and results[k] pair
 * up; each result equals gtp_run's for that trace, and kernel_ms / total_ms
 * hold the device time of the whole batch.  Traces are loaded in groups under
 * the event budget (max_event_num, else device memory), each group processed
 * in rounds of round_tiles work tiles.  The resident trace of gtp_set_trace is
 * not touched.  No counterpart in the reference (new method). */
gtp_status gtp_run_batch(gtp_ctx* ctx, int32_t n_trace, const gtp_trace* traces,
                         gtp_result* results);

/* Evaluate a trace held in host memory in cycle intervals whose device event
 * store stays within max_events (<= 0: the device memory budget), for traces
 * larger than device memory.  Each interval's per-net slice carries the
 * events before it that the state, the glitch flags and the related-pin
 * choice at its start depend on, and the two after its end; per-gate sums
 * are exact integers over intervals, so the result is bit-identical to one
 * gtp_run over the whole trace.  The resident trace of gtp_set_trace is not
 * touched.  Replaces the reference's multi-round interval loop
 * (power/cuda/CudaPower.cu:1261-1314, getTimeIntervals / renewGateWaveform
 * :1421-1599); the interval count is reported in gtp_schedule_info.n_interval.
 * kernel_ms sums the intervals' device time; total_ms is the call's wall time
 * (host slicing and uploads included). */
gtp_status gtp_run_sliced(gtp_ctx* ctx, const gtp_trace* trace, int64_t max_events, gtp_result* result);

/* Streamed batch: traces handed over one at a time, so that the caller holds
 * only the trace it is adding (a batch larger than host memory can be read
 * trace by trace from its files).  gtp_batch_add packs the trace into event
 * words on the device (through page-locked staging) before it returns, after
 * which the caller may free the trace; its result is filled when the load it
 * joined is evaluated -- when the next trace does not fit the event budget
 * (max_event_num, else device memory), or in gtp_batch_end -- so result
 * pointers must stay valid until gtp_batch_end returns.  A trace of 2^29 - 1
 * ticks or more, or larger than one load, is evaluated by itself in cycle
 * intervals (as gtp_run_sliced).  Every result equals gtp_run's for its trace;
 * a result's kernel_ms is the device time of the load it was evaluated in.
 * After an error, later calls return it; gtp_batch_end frees the batch in
 * every case and returns the first error, or fills *kernel_ms (may be NULL)
 * with the device time of the whole batch.  The design must stay the one
 * loaded at gtp_batch_begin (GTP_ERR_STATE otherwise); the resident trace of
 * gtp_set_trace is not touched.  Replaces the reference's input rounds
 * (power/cuda/CudaPower.cu:1261-1314, the reader's bounded buffers
 * power/VcdReader.cc:292-300) for many traces; the batch itself is a new
 * method. */
typedef struct gtp_batch gtp_batch;
gtp_status gtp_batch_begin(gtp_ctx* ctx, double time_scale, double clk_period, gtp_batch** out);
gtp_status gtp_batch_add(gtp_batch* batch, const gtp_trace* trace, gtp_result* result);
gtp_status gtp_batch_end(gtp_batch* batch, double* kernel_ms);

/* Schedule of the last run. */
typedef struct {
  int32_t tile_cycles;   /* cycles per work tile */
  int32_t block_lanes;   /* lanes (gates) per workgroup pass */
  int32_t gates_per_wg;  /* gates of one work item (a gate range of one tile) */
  int32_t n_launch;      /* power-kernel launches of one evaluation (one per cell kernel) */
  int64_t n_interval;    /* gtp_run_sliced: cycle intervals loaded one after another; else 1 */
  int64_t n_unit;        /* (gate, tile) pairs */
  int64_t n_workgroup;   /* work items (gate range, tile) */
  int64_t n_tile;        /* work tiles */
  int64_t n_event;       /* net events */
  int64_t n_pin_event;   /* pin events (net events x the pins reading them) */
  double autotune_ms;    /* time the autotuner spent */
} gtp_schedule_info;
gtp_status gtp_get_schedule(const gtp_ctx* ctx, gtp_schedule_info* info);

/* Autotuner: times evaluations of the loaded trace over candidate tile lengths
 * (32..512 cycles) and gate-range sizes (512..8192 gates per workgroup item),
 * and keeps the fastest pair for later gtp_run and gtp_run_batch calls until
 * the next gtp_set_trace (options that fix tile_cycles / gates_per_wg win).
 * gtp_run calls it by itself when options.autotune is set.  Replaces the
 * reference's fixed cycles-per-thread heuristic (power/cuda/Gate.cuh:354-376)
 * as the schedule choice. */
gtp_status gtp_autotune(gtp_ctx* ctx, gtp_schedule_info* best);

/* The reference's event-density heuristic for cycles per thread
 * (Gate::setNCyclePerThreadByEventCount, power/cuda/Gate.cuh:354-376), for
 * n gates with n_event[g] events each over n_cycle cycles. */
gtp_status gtp_reference_cycles_per_thread(int64_t n, const int64_t* n_event, int64_t n_cycle,
                                           int64_t e_target, int64_t parallelism_floor,
                                           int64_t* out_ncpt);

/* ---------------------------------------------------------------------------
 * Learned per-cycle power surrogate.  No counterpart in the reference (a new
 * method on top of its path): an MLP over per-cycle toggle bits of selected
 * nets, n_in -> n_hidden -> n_hidden -> 1 with GELU (x * Phi(x)), trained and
 * run on the device with BF16 matrix-core kernels accumulating in FP32.
 * Training targets come from gtp_run; gtp_run_tiles judges it.
 *
 * Bit layout: row c of a feature matrix is n_in / 32 words; bit k of the row
 * is bit (k % 32) of word k / 32.
 * Parameter layout (FP32): W1[n_hidden][n_in], b1[n_hidden],
 * W2[n_hidden][n_hidden], b2[n_hidden], w3[n_hidden], b3[1].
 * ------------------------------------------------------------------------- */
typedef struct {
  int32_t device;    /* HIP device ordinal; -1 = the current device */
  int32_t n_in;      /* feature bits, a multiple of 128 */
  int32_t n_hidden;  /* hidden width, a multiple of 128 */
  int32_t pad_;
  uint64_t seed;     /* initialisation seed */
} gtp_mlp_config;

typedef struct {
  int32_t batch;        /* rows per step, a multiple of 32 */
  int32_t epochs;
  float lr;             /* Adam peak learning rate (cosine decay to 0) */
  float weight_decay;   /* decoupled weight decay (AdamW); 0 = none */
  uint64_t seed;        /* shuffling seed */
  int32_t val_rows;     /* the last val_rows rows are held out; the best epoch's parameters are kept */
  int32_t input_dropout; /* 1: each input bit of a training row is dropped with probability 1/2 and
                            kept ones count 2 (inverted dropout); 0: none */
} gtp_train_config;

typedef struct gtp_mlp gtp_mlp;

gtp_status gtp_mlp_create(const gtp_mlp_config* cfg, gtp_mlp** out);
void gtp_mlp_destroy(gtp_mlp* m);
const char* gtp_mlp_last_error(const gtp_mlp* m);
int64_t gtp_mlp_param_count(const gtp_mlp* m);
/* y_mean_std (may be NULL): the target standardisation, mean and std. */
gtp_status gtp_mlp_get_params(gtp_mlp* m, float* params, float* y_mean_std);
gtp_status gtp_mlp_set_params(gtp_mlp* m, const float* params, const float* y_mean_std);
/* Gradients of the mean squared error over n rows (n a multiple of 32), with
 * targets y already standardised; the model is not changed.  drop_seed != 0
 * applies the training-time input dropout: bit b of word w of row r is kept
 * when bit b of gtp_keep_mask(drop_seed, r, w) is set, and counts 2. */
gtp_status gtp_mlp_gradients(gtp_mlp* m, const uint32_t* bits, const float* y, int32_t n,
                             uint32_t drop_seed, float* grads, float* loss);
/* The dropout keep bits of (seed, row, word) (a murmur3-style finaliser of
 * seed ^ row * 0x9E3779B1 ^ word * 0x85EBCA77). */
uint32_t gtp_keep_mask(uint32_t seed, uint32_t row, uint32_t word);
/* Train on n rows with raw targets y (standardised inside with the training
 * rows' mean and std).  epoch_loss (may be NULL) is [2 * epochs]: the mean
 * training loss and the validation loss (NaN without validation rows) of each
 * epoch, on standardised targets. */
gtp_status gtp_mlp_train(gtp_mlp* m, const uint32_t* bits, const float* y, int64_t n,
                         const gtp_train_config* tc, float* epoch_loss);
/* Predictions in target units for n rows of host feature bits. */
gtp_status gtp_mlp_predict(gtp_mlp* m, const uint32_t* bits, int64_t n, float* out);
/* gtp_mlp_predict without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * rows, hidden activations and outputs of each chunk of rows are staged in a
 * fixed 16 MiB device array that the library holds on every GPU it runs on
 * until the process ends; rows are processed in chunks of as many as fit it
 * (10880 for n_in 4096, n_hidden 256).  GTP_ERR_LIMIT when a 128-row chunk of
 * the model does not fit.  Results equal gtp_mlp_predict's.  Calls on one GPU
 * share that array: calls that overlap on one GPU (from several threads or
 * models) race and must be serialised by the caller. */
gtp_status gtp_mlp_predict_noalloc(gtp_mlp* m, const uint32_t* bits, int64_t n, float* out);
/* Features are per-cycle toggle bits of (net, kind) pairs, identified as
 * kind * n_net + net: kind GTP_FEAT_CHANGE sets the bit of a cycle in which
 * the net changes value, kind GTP_FEAT_RISE of one in which it changes to 1. */
enum { GTP_FEAT_CHANGE = 0, GTP_FEAT_RISE = 1, GTP_N_FEAT_KIND = 2 };
/* Predictions for cycles [cycle_begin, +n_cycle) of the trace loaded in ctx,
 * features built on the device from its event store (feats[k] = feature id
 * of input k); ms (may be NULL) receives the device time of both. */
gtp_status gtp_mlp_predict_trace(gtp_mlp* m, gtp_ctx* ctx, const int32_t* feats, int64_t cycle_begin,
                                 int64_t n_cycle, float* out, double* ms);
/* Pearson correlation of every feature's per-cycle bit with y[n_cycle], over
 * the cycles with mask[c] != 0 (NULL: all); corr is [GTP_N_FEAT_KIND * n_net],
 * indexed by feature id. */
gtp_status gtp_feature_correlation(gtp_ctx* ctx, const double* y, const uint8_t* mask, float* corr);
/* Bits of the features feats[0..n_sel) (n_sel a multiple of 32) for cycles
 * [cycle_begin, +n_cycle) into bits[n_cycle][n_sel / 32]. */
gtp_status gtp_toggle_features(gtp_ctx* ctx, int32_t n_sel, const int32_t* feats, int64_t cycle_begin,
                               int64_t n_cycle, uint32_t* bits);
/* Events per net of the trace loaded in ctx (counts[n_net]; the first event
 * of a net is its initial value). */
gtp_status gtp_net_event_counts(gtp_ctx* ctx, int64_t* counts);
/* Weighted toggle sums of the trace loaded in ctx: out[c] = sum over k of
 * weights[k] x (value changes of net nets[k] in cycle cycle_begin + c), for
 * n_cycle cycles (the surrogate's linear base term; no counterpart in the
 * reference).  Weights are held to 2^-20 of the largest |weight| and summed
 * exactly.  Runs on the context's auxiliary stream, so it may overlap a
 * power run of the same context issued from another host thread; ms (may be
 * NULL) receives the device time. */
gtp_status gtp_net_weight_sum(gtp_ctx* ctx, int32_t n_sel, const int32_t* nets, const double* weights,
                              int64_t cycle_begin, int64_t n_cycle, double* out, double* ms);
/* The surrogate's matrix-core GEMM on host arrays, C[M][N] = A[M][K] B[K][N]
 * in FP32: a_mode 0 A row-major bf16, 1 A stored [K][M] bf16, 2 A bits
 * [M][K/32]; b_mode 0 B row-major bf16, 1 B stored [N][K] bf16, 2 B bits
 * [K][N/32].  N a multiple of 128, K of 32.  For layout tests. */
gtp_status gtp_mlp_gemm_test(int32_t a_mode, int32_t b_mode, int32_t M, int32_t N, int32_t K,
                             const void* A, int64_t a_bytes, const void* B, int64_t b_bytes, float* C);

/* ---- Device VCD ingest --------------------------------------------------
 * VCD text -> the per-(id, bit) event store, parsed on the device.  Replaces
 * the reference's single-threaded reader (VcdReader::read / parseVarValues,
 * power/VcdReader.cc:80-131, :238-285, with Vcd::varAppendValue /
 * varAppendBusValue, power/Vcd.cc:203-232) and the host event-store build
 * that cuts per-bit ranges out of it (power/cuda/CudaPower.cu:1874-2086, value
 * table :1129-1145).  Same token rules, same dropping of a change equal to
 * its id's last value, same initial X (scalar) or 0 (bus) at time 0.  One
 * event stream per (unique id, bit), ids numbered in declaration order, bits
 * 0..w-1 (w = the widest declaration of the id, at most 64); values 0, 1,
 * X = 2, Z = 3, U = 2.  Errors where the reference reports a file error, and
 * when an id's kept times do not increase strictly or reach 2^29 - 1 ticks. */
typedef struct gtp_vcd gtp_vcd;
typedef struct {
  int64_t n_var;        /* declared variables ($var) */
  int64_t n_id;         /* unique ids */
  int64_t n_stream;     /* event streams (sum of id widths): the trace's n_net */
  int64_t n_event;      /* events over all streams */
  int64_t t_max;        /* the last time of the text (the trace's t_max) */
  double time_scale;    /* seconds per tick ($timescale) */
  int64_t n_token;      /* tokens of the value-change section */
  int64_t n_change;     /* value changes before dropping repeats */
  int64_t body_bytes;   /* bytes of the value-change section */
  double device_ms;     /* device time of the parse (text resident on the device) */
  double total_ms;      /* wall time including the text upload */
  int64_t n_part;       /* parts the text was parsed in (cut at whitespace) */
} gtp_vcd_info;
typedef struct {
  const char* name;     /* scope path joined by '/', as the reference names it */
  const char* id;
  int32_t width;        /* declared width */
  int32_t id_index;     /* unique id */
  int64_t first_stream; /* stream of bit 0 of its id */
} gtp_vcd_var_info;
gtp_status gtp_vcd_create(int32_t device, gtp_vcd** out);
void gtp_vcd_destroy(gtp_vcd* v);
const char* gtp_vcd_last_error(const gtp_vcd* v);
/* Parse text[0, len); the result stays on the device.  A text longer than
 * the part size (1 GiB; GTP_VCD_PART_BYTES in the environment sets another)
 * is parsed in parts cut at whitespace (not after a token that starts with
 * 'b' or 'B', which may take the next token as its id): each part goes
 * through the device pipeline with the time and every id's last kept change
 * carried over, and the parts' stores are appended per stream at the end
 * (device memory: the parts' stores and the final store at once). */
gtp_status gtp_vcd_parse(gtp_vcd* v, const char* text, int64_t len, gtp_vcd_info* info);
/* Declared variable i; the strings live until the next parse. */
gtp_status gtp_vcd_var(const gtp_vcd* v, int64_t i, gtp_vcd_var_info* out);
/* Copy the streams out: offsets [n_stream + 1], times and values [n_event]. */
gtp_status gtp_vcd_events(gtp_vcd* v, int64_t* offsets, int64_t* times, uint8_t* values);
/* Load a parsed VCD as ctx's trace (device to device; the design's n_net
 * must equal n_stream): gtp_set_trace for text input. */
gtp_status gtp_set_trace_vcd(gtp_ctx* ctx, const gtp_vcd* v, double clk_period);

/* Writers for the reference's two result files.
 * gtp_write_cell_power: `cuda_cell_power.txt` (power/cuda/CudaPower.cu:3134-3165),
 *   one row per gate: name, internal, glitch internal, switching, glitch
 *   switching, leakage, total, cell name; tab separated, 6 significant digits.
 * gtp_write_per_cycle: the per-cycle waveform file
 *   (include/sta/PowerClass.hh:123-149) for one top instance.
 * Names are NUL-separated strings, one per gate (resp. cell type). */
gtp_status gtp_write_cell_power(const char* path, int64_t n_gate, const double* gate_power,
                                const char* gate_names, const int32_t* gate_cell,
                                const char* cell_names);
gtp_status gtp_write_per_cycle(const char* path, const char* top_name, int64_t n_cycle,
                               const double* cycle_power);

#ifdef __cplusplus
}
#endif
#endif /* GTPOWER_H */
