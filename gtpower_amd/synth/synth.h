/*
 * synth.h -- seeded synthetic designs and activity traces for the power engine.
 *
 * Test and benchmark input only: a random gate-level netlist over a cell
 * library (gtp_library plus per-arc NLDM delay/slew table indices), its load
 * capacitances, slews and arc delays from the NLDM tables, and event traces
 * from a transport-delay logic simulation, so that unequal path delays create
 * glitches.  Every array uses the layout of gtp_design / gtp_trace.
 */
#ifndef GTPOWER_SYNTH_H
#define GTPOWER_SYNTH_H

#include <stdint.h>
#include "../../include/gtpower.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int64_t n_gate;            /* instances */
  double dff_fraction;       /* fraction of gates that are flip-flops */
  int32_t depth;             /* logic levels */
  int32_t fanout_min, fanout_max;
  double wire_cap_min, wire_cap_max;   /* F */
  int64_t n_pi;              /* primary inputs; 0 = n_gate / 50 (at least 8) */
  double pi_slew;            /* s */
  double time_scale;         /* s per tick */
  int64_t ticks_per_cycle;   /* clock period in ticks */
  uint64_t seed;
  int32_t n_mix;             /* combinational cell mix */
  const int32_t* mix_cells;
  const double* mix_weights;
  int32_t dff_cell;          /* cell index of the flip-flop, -1 = none */
  int32_t threads;           /* 0 = all cores */
} gtps_design_params;

typedef struct {
  int64_t n_cycle;
  double toggle_p;           /* primary-input toggle probability per cycle */
  uint64_t seed;
  /* Random-activity mode (mode = 1): every net toggles independently, with a
   * per-net rate (events per cycle) drawn lognormal(mu, sigma); `hot_fraction`
   * of the nets are scaled up so that they carry `hot_share` of the events;
   * `burst_fraction` of the cycles have `burst_mult` times the rate. */
  int32_t mode;              /* 0 = logic simulation, 1 = random activity */
  double lognorm_mu, lognorm_sigma;
  double hot_fraction, hot_share;
  double burst_fraction, burst_mult;
  double max_rate;           /* cap on events per cycle of one net */
  int32_t threads;
} gtps_trace_params;

typedef struct gtps_design gtps_design;
typedef struct gtps_trace gtps_trace;

/* arc_tables[((cell * GTP_MAX_CELL_PINS + in) * 2 + from_rf) * 2 + to_rf) * 2 + k]
 * is the delay (k = 0) or output slew (k = 1) table of the cell's only output;
 * -1 = no arc.  cell_func[c] is the truth table of a combinational cell's
 * output over its inputs (bit s: inputs s); ignored for the flip-flop. */
int32_t gtps_design_create(const gtp_library* lib, const int32_t* arc_tables,
                           const uint64_t* cell_func, const double* cell_pin_cap,
                           const gtps_design_params* p, gtps_design** out);
void gtps_design_free(gtps_design* d);
/* Sizes: n_gate, n_pin, n_net, n_arc (arc_delay length), n_pi. */
void gtps_design_sizes(const gtps_design* d, int64_t* sizes5);
/* A gtp_design view of the generated design (valid while d lives). */
void gtps_design_view(const gtps_design* d, gtp_design* view);
/* Per-net driver gate (-1 for primary inputs and the clock) and per-gate level. */
const int32_t* gtps_net_driver(const gtps_design* d);
const int32_t* gtps_gate_level(const gtps_design* d);

int32_t gtps_trace_create(const gtps_design* d, const gtps_trace_params* p, gtps_trace** out);
void gtps_trace_free(gtps_trace* t);
/* A gtp_trace view (valid while t lives). */
void gtps_trace_view(const gtps_trace* t, gtp_trace* view);
int64_t gtps_trace_events(const gtps_trace* t);
const char* gtps_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
