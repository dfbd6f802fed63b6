/*
 * oracle.h -- FP64 CPU restatement of GTPower's time-based power algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle for the HIP engine and
 * the CPU baseline of the benchmark; no product code path calls it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * It restates, in double precision, the reference's multi-threaded CPU path
 * (the one selected by -disable_cuda_power_analysis):
 *   Power::power                        power/Power.cc:226-386
 *   Power::findTimeBasedAllPower        power/TimeBasedPower.cc:703-901
 * on the same explicit inputs as the engine (include/gtpower.h).
 */
#ifndef GTPOWER_ORACLE_H
#define GTPOWER_ORACLE_H

#include <stdint.h>
#include "../include/gtpower.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double* gate_power;        /* [5 * n_gate]  component-major, GTP_INTERNAL.. order */
  double* gate_abs;          /* [5 * n_gate]  sum of |term| magnitudes (tolerance scale) */
  double* cycle_power;       /* [5 * n_cycle] waveform-major, GTP_WAVE_LEAKAGE.. order */
  double* cycle_abs;         /* [5 * n_cycle] */
  uint64_t* gate_toggle_x2;  /* [n_gate] */
  uint64_t* gate_glitch;     /* [n_gate] */
  int64_t n_cycle;           /* out */
  double wall_s;             /* out: wall time of the power loop */
} oracle_result;

/* Evaluate gates [gate_begin, gate_end) of the design on the trace with
 * `threads` worker threads (0 = all cores). */
int32_t oracle_run(const gtp_library* lib, const gtp_design* d, const gtp_trace* tr,
                   int32_t threads, int64_t gate_begin, int64_t gate_end, oracle_result* out);

/* Clock-cycle-based glitch flags of one waveform
 * (power/TimeBasedPower.cc:589-637 getEventClockCycleBasedGlitchFlags). */
void oracle_glitch_flags(const int64_t* t, int64_t n, int64_t ticks_per_cycle, uint8_t* flags);
/* Glitch scaling ratio of event i, or -1 when it is not a glitch
 * (power/TimeBasedPower.cc:303-339, power/PowerUtils.cc:285-294). */
double oracle_glitch_ratio(const int64_t* t, int64_t n, const uint8_t* flags, int64_t i,
                           double time_scale, double slew_rise, double slew_fall);
/* Liberty table interpolation in double (liberty/TableModel.cc:858-1062). */
double oracle_table_value(const gtp_library* lib, int32_t table, double slew, double cap);
/* Reference cycles-per-thread heuristic (power/cuda/Gate.cuh:354-376). */
int64_t oracle_cycles_per_thread(int64_t n_event, int64_t n_cycle, int64_t e_target,
                                 int64_t parallelism_floor);
/* Toggle and rise counts (power/PowerUtils.cc:213-256), times 2 (exact integers). */
int32_t oracle_toggle_x2(int32_t prev, int32_t cur);
int32_t oracle_rise_x2(int32_t prev, int32_t cur);

#ifdef __cplusplus
}
#endif
#endif
