// FP64 CPU restatement of GTPower's time-based power algorithm.
//
// TEST INFRASTRUCTURE ONLY: the oracle for the HIP engine and the CPU baseline
// of the benchmark.  Every function cites the reference code it follows.  Where
// the reference computes in float (util/Types.hh:3-18) this computes in double;
// where its CPU and GPU paths differ, the CPU path (power/TimeBasedPower.cc) is
// followed.  Two deliberate, documented choices:
//   * time is integer ticks; a cycle is T = ceil(period / time_scale) ticks
//     (power/PowerUtils.cc:27-32) and per-cycle leakage is split by tick
//     overlap / T, which equals the reference's float-second split when
//     period / time_scale is an integer;
//   * the related-pin arrival comparison |t_out - d - t_in| is done exactly, on
//     ticks scaled by 2^16 with the arc delay rounded to that grid, so that the
//     engine can reproduce the same choice bit for bit.
//   * a per-cycle contribution whose cycle index is >= n_cycle (an event at
//     t_max when t_max is a multiple of T) is dropped instead of written one
//     past the end (power/TimeBasedPower.cc:690-699 writes index C).
//
// Built with -DORACLE_REAL=float (liboracle_f32.so) the energy terms, the
// per-gate sums and the per-thread per-cycle waveforms are FP32, as the
// reference's PowerVal / EnergyVal (util/Types.hh:3-18); table interpolation
// stays FP64 inside with an FP32 result, as OpenSTA's findValue.  That build
// measures the FP32-versus-FP64 error budget; the FP64 build is the oracle.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif

namespace {

typedef ORACLE_REAL real;  // energy and power values (FP64 oracle, FP32 diagnostic)
constexpr int RISE = 0, FALL = 1;
constexpr int ARC_FRAC_BITS = 16;

// power/PowerUtils.cc:246-256 getNToggle, times two.
inline int toggle_x2(int prev, int cur) {
  if (prev == cur) return 0;
  return (prev == 2 || cur == 2) ? 1 : 2;
}
// power/PowerUtils.cc:213-236 getNRise, times two.
inline int rise_x2(int prev, int cur) {
  if (prev == cur) return 0;
  if (cur == 2) return prev == 0 ? 1 : 0;
  if (cur == 1) return prev == 2 ? 1 : (prev == 0 ? 2 : 0);
  return 0;
}
// power/PowerUtils.cc:238-244 getRiseFallEdge.
inline int edge(int prev, int cur) { return rise_x2(prev, cur) > 0 ? RISE : FALL; }

// liberty/TableModel.cc:1465-1488 findValueIndex.
int find_index(const double* ax, int n, double x) {
  if (n <= 1 || x <= ax[0]) return 0;
  if (x >= ax[n - 1]) return n - 2;
  int lo = -1, hi = n;
  while (hi - lo > 1) {
    int mid = (hi + lo) >> 1;
    if (x >= ax[mid]) lo = mid; else hi = mid;
  }
  return lo;
}

// liberty/TableModel.cc:858-1062 Table1/Table2::findValue, axes mapped by
// variable as liberty/InternalPower.cc:178-229.  `mag` receives the same
// interpolation of |values| with |weights| (the error-propagation scale).
double table_value(const gtp_library* lib, int t, double slew, double cap, double* mag) {
  const gtp_table& tb = lib->tables[t];
  const double* D = lib->table_data;
  double x1 = tb.var1 == GTP_AXIS_SLEW ? slew : cap;
  double x2 = tb.var2 == GTP_AXIS_SLEW ? slew : cap;
  const double* a1 = D + tb.axis1;
  const double* v = D + tb.values;
  auto lin = [&](const double* ax, int n, double x, const double* y, int stride) {
    if (n == 1) {
      if (mag) *mag = std::fabs(y[0]);
      return y[0];
    }
    int i = find_index(ax, n, x);
    double d = (x - ax[i]) / (ax[i + 1] - ax[i]);
    double y0 = y[(size_t)i * stride], y1 = y[(size_t)(i + 1) * stride];
    if (mag) *mag = std::fabs(1 - d) * std::fabs(y0) + std::fabs(d) * std::fabs(y1);
    return (1 - d) * y0 + d * y1;
  };
  if (tb.n2 == 0) return lin(a1, tb.n1, x1, v, 1);
  const double* a2 = D + tb.axis2;
  const int n1 = tb.n1, n2 = tb.n2;
  if (n1 == 1 && n2 == 1) {
    if (mag) *mag = std::fabs(v[0]);
    return v[0];
  }
  if (n1 == 1) return lin(a2, n2, x2, v, 1);
  if (n2 == 1) return lin(a1, n1, x1, v, 1);
  int i = find_index(a1, n1, x1), j = find_index(a2, n2, x2);
  double d1 = (x1 - a1[i]) / (a1[i + 1] - a1[i]);
  double d2 = (x2 - a2[j]) / (a2[j + 1] - a2[j]);
  double y00 = v[i * n2 + j], y10 = v[(i + 1) * n2 + j];
  double y11 = v[(i + 1) * n2 + j + 1], y01 = v[i * n2 + j + 1];
  double w00 = (1 - d1) * (1 - d2), w10 = d1 * (1 - d2), w11 = d1 * d2, w01 = (1 - d1) * d2;
  if (mag)
    *mag = std::fabs(w00) * std::fabs(y00) + std::fabs(w10) * std::fabs(y10) +
           std::fabs(w11) * std::fabs(y11) + std::fabs(w01) * std::fabs(y01);
  return w00 * y00 + w10 * y10 + w11 * y11 + w01 * y01;
}

// power/TimeBasedPower.cc:589-637 getEventClockCycleBasedGlitchFlags over the
// whole waveform (start 0): event 0 is never grouped; events of one cycle form
// a group; of a group of n >= 2 the first n (n even) or n - 1 (n odd) are
// flagged.
void glitch_flags(const int64_t* t, int64_t n, int64_t T, uint8_t* flags) {
  std::fill(flags, flags + n, 0);
  if (n <= 0) return;
  std::vector<int64_t> group;
  int64_t cur = t[0] / T;
  auto close = [&] {
    size_t m = group.size();
    if (m >= 2) {
      size_t k = (m % 2 == 0) ? m : m - 1;
      for (size_t q = 0; q < k; ++q) flags[group[q]] = 1;
    }
    group.clear();
  };
  for (int64_t pos = 1; pos < n; ++pos) {
    int64_t c = t[pos] / T;
    if (c != cur) {
      close();
      cur = c;
    }
    group.push_back(pos);
  }
  close();
}

// power/PowerUtils.cc:285-294 getTimeBasedGlitchScalingRatioH.
inline double ratio_from_width(double pulse_width_s, double sum_slew) {
  if (sum_slew == 0) return 1.0;
  double r = (pulse_width_s * 2.0) / sum_slew;
  return std::min(1.0, r * r);
}

// power/TimeBasedPower.cc:303-339 getGlitchScalingRatioClockCycleBasedH.
double glitch_ratio(const int64_t* t, int64_t n, const uint8_t* flags, int64_t i, double scale,
                    double slew_r, double slew_f) {
  if (!flags || !flags[i]) return -1.0;
  int64_t prev_w = -1, next_w = -1;
  if (i - 1 >= 0 && flags[i - 1]) prev_w = t[i] - t[i - 1];
  if (i + 1 < n && flags[i + 1]) next_w = t[i + 1] - t[i];
  int64_t w = std::max(prev_w, next_w);
  if (w == -1) return -1.0;
  return ratio_from_width((double)w * scale, slew_r + slew_f);
}

// power/Power.cc:226-386 drives findTimeBasedAllPower per gate; these are the
// per-gate inputs, looked up once.
struct CellInfo {
  int n_pin = 0, n_in = 0;
  uint32_t clock_mask = 0;
  const double* leak = nullptr;  // 2^n_pin values, or null -> default
  double leak_default = 0;
  // internal power groups per pin, Liberty order
  std::vector<std::vector<int>> groups;
};

struct Ctx {
  const gtp_library* lib;
  const gtp_design* d;
  const gtp_trace* tr;
  int64_t T, C;
  double scale, period, t_max_s;
  std::vector<CellInfo> cells;
};

// Cells of more than 6 pins carry unconditional groups only (when_truth all
// ones), so a state index of 64 or more holds exactly when every bit is set.
inline bool when_holds(const gtp_internal_power& ip, uint32_t state) {
  return state < 64 ? ((ip.when_truth >> state) & 1ull) != 0 : ip.when_truth == ~0ull;
}

// power(rf, op_cond, slew, cap) of one internal_power group
// (liberty/InternalPower.cc:112-124); a missing rf table gives 0.
inline double group_power(const Ctx& x, const gtp_internal_power& ip, int rf, double slew, double cap,
                          double* mag) {
  int t = rf == RISE ? ip.rise_table : ip.fall_table;
  if (t < 0) {
    if (mag) *mag = 0;
    return 0.0;
  }
  return table_value(x.lib, t, slew, cap, mag);
}

struct GateScratch {
  std::vector<std::vector<uint8_t>> flags;
  std::vector<int64_t> frontier;
  std::vector<int> prev, cur;
  std::vector<char> trig;
};

struct Acc {
  std::vector<real> cyc;       // [5][C]
  std::vector<double> cyc_abs;  // [5][C]
};

void eval_gate(const Ctx& x, int64_t g, GateScratch& s, Acc& acc, real* gp, double* ga,
               uint64_t* tog, uint64_t* gl) {
  const gtp_design* d = x.d;
  const gtp_trace* tr = x.tr;
  const int c = d->gate_cell[g];
  const CellInfo& ci = x.cells[c];
  const int n = ci.n_pin, nin = ci.n_in;
  const int64_t pb = d->gate_pin_offset[g];
  const int64_t ab = d->gate_arc_offset[g];
  const int64_t T = x.T, C = x.C;
  const double scale = x.scale, period_d = x.period;
  const int64_t t_max = tr->t_max;
  const double denom_d = (double)t_max * scale;  // vcd_.timeMax() * vcd_.timeScale()
  const real period = (real)period_d, denom = (real)denom_d;

  const int64_t* tt[GTP_MAX_CELL_PINS];
  const uint8_t* vv[GTP_MAX_CELL_PINS];
  int64_t ne[GTP_MAX_CELL_PINS];
  for (int p = 0; p < n; ++p) {
    int32_t net = d->pin_net[pb + p];
    if (net >= 0) {
      int64_t o0 = tr->net_event_offset[net], o1 = tr->net_event_offset[net + 1];
      tt[p] = tr->event_time + o0;
      vv[p] = tr->event_value + o0;
      ne[p] = o1 - o0;
    } else {
      tt[p] = nullptr;
      vv[p] = nullptr;
      ne[p] = 0;
    }
  }
  // glitch flags (TimeBasedPower.cc:738-747): none for clock pins
  s.flags.resize(n);
  for (int p = 0; p < n; ++p) {
    s.flags[p].assign(ne[p], 0);
    if (!((ci.clock_mask >> p) & 1u) && ne[p] > 0) glitch_flags(tt[p], ne[p], T, s.flags[p].data());
  }
  // single switching energies C*V^2 of output pins (TimeBasedPower.cc:749-760)
  real sw[GTP_MAX_CELL_PINS];
  for (int p = 0; p < n; ++p)
    sw[p] = (p >= nin) ? (real)(d->pin_load_cap[pb + p] * d->pin_voltage[pb + p] * d->pin_voltage[pb + p]) : (real)0;
  auto leak_of = [&](const std::vector<int>& st) -> real {
    // power/PowerUtils.cc:67-90 findLeakageVal with getPinStatesIndex (:39-48)
    if (!ci.leak) return (real)ci.leak_default;
    uint32_t idx = 0;
    for (int p = 0; p < n; ++p)
      if (st[p] == 1) idx |= 1u << p;
    return (real)ci.leak[idx];
  };
  auto state_index = [&](const std::vector<int>& st) {
    uint32_t idx = 0;
    for (int p = 0; p < n; ++p)
      if (st[p] == 1) idx |= 1u << p;
    return idx;
  };
  // loop initial (TimeBasedPower.cc:769-791)
  s.frontier.assign(n, 0);
  s.prev.assign(n, 0);
  s.trig.assign(n, 0);
  for (int p = 0; p < n; ++p) {
    if (ne[p] == 0) {
      s.prev[p] = d->pin_default_state[pb + p];
      continue;
    }
    if (tt[p][0] == 0) {
      s.prev[p] = vv[p][0];
      s.frontier[p] = 1;
    } else {
      s.prev[p] = 2;
    }
  }
  s.cur = s.prev;
  int64_t prev_time = 0;
  real* cyc = acc.cyc.data();
  double* cab = acc.cyc_abs.data();

  // per-cycle leakage split over [t0, t1) (TimeBasedPower.cc:677-701, integer ticks)
  auto leak_split = [&](real L, int64_t t0, int64_t t1) {
    if (L == (real)0 || t1 <= t0) return;
    int64_t cidx = t0 / T;
    int64_t left = t0;
    while (cidx < C && cidx * T < t1) {
      int64_t right = std::min((cidx + 1) * T, t1);
      real v = L * (real)(right - left) / (real)T;
      cyc[GTP_WAVE_LEAKAGE * C + cidx] += v;
      cab[GTP_WAVE_LEAKAGE * C + cidx] += std::fabs(v);
      left = right;
      ++cidx;
    }
  };

  for (;;) {
    // current min time (TimeBasedPower.cc:796-811)
    int64_t cur_time = INT64_MAX;
    for (int p = 0; p < n; ++p)
      if (s.frontier[p] < ne[p]) cur_time = std::min(cur_time, tt[p][s.frontier[p]]);
    if (cur_time == INT64_MAX) break;
    // new state, advance the frontier (:813-829)
    for (int p = 0; p < n; ++p) {
      s.trig[p] = 0;
      if (s.frontier[p] < ne[p] && tt[p][s.frontier[p]] == cur_time) {
        s.cur[p] = vv[p][s.frontier[p]];
        s.trig[p] = 1;
        ++s.frontier[p];
      }
    }
    // leakage by the previous state (:831-843)
    real L = leak_of(s.prev);
    leak_split(L, prev_time, cur_time);
    gp[GTP_LEAKAGE] += L * (real)(cur_time - prev_time) / (real)t_max;
    ga[GTP_LEAKAGE] += std::fabs(L) * (double)(cur_time - prev_time) / (double)t_max;
    const int64_t cidx = cur_time / T;
    const bool in_range = cidx < C;
    const uint32_t prev_idx = state_index(s.prev);
    // internal and switching (:845-883)
    for (int p = 0; p < n; ++p) {
      if (!s.trig[p]) continue;
      const int pv = s.prev[p], cv = s.cur[p];
      const int tx2 = toggle_x2(pv, cv);
      const real n_toggle = (real)0.5 * (real)tx2;
      const int64_t ei = s.frontier[p] - 1;
      const real ratio = (real)glitch_ratio(tt[p], ne[p], s.flags[p].data(), ei, scale,
                                        d->pin_slew[2 * (pb + p)], d->pin_slew[2 * (pb + p) + 1]);
      *tog += (uint64_t)tx2;
      if (ratio > 0) *gl += 1;
      const int rf = edge(pv, cv);
      real E = 0;
      double M = 0.0;
      if (p < nin) {
        // findInputInternalVal (:341-381)
        double slew = d->pin_slew[2 * (pb + p) + rf];
        if (std::isfinite(slew)) {
          for (int k : ci.groups[p]) {
            const gtp_internal_power& ip = x.lib->ipowers[k];
            if (when_holds(ip, prev_idx)) {
              E = (real)group_power(x, ip, rf, slew, 0.0, &M);
              break;
            }
          }
        }
      } else {
        // findOutputInternalVal (:383-509)
        const int o = p - nin;
        const int to_rf = rf;
        const double cap = d->pin_load_cap[pb + p];
        int64_t min_diff = (int64_t)t_max << ARC_FRAC_BITS;
        int related = -1, related_rf = RISE;
        for (int i = 0; i < nin; ++i) {
          if (ne[i] == 0) continue;
          // largest event index with time < cur_time, at least 0 (:422-425)
          int64_t r = (int64_t)(std::lower_bound(tt[i], tt[i] + ne[i], cur_time) - tt[i]) - 1;
          r = std::max<int64_t>(0, r);
          int64_t pr = std::max<int64_t>(0, r - 1);
          int from_rf = edge(vv[i][pr], vv[i][r]);
          double delay = d->arc_delay[ab + (((int64_t)o * nin + i) * 2 + from_rf) * 2 + to_rf];
          if (std::isnan(delay)) continue;  // no timing arc (:443-445)
          int64_t dq = (int64_t)std::llround(delay / scale * (double)(1 << ARC_FRAC_BITS));
          int64_t diff = ((cur_time - tt[i][r]) << ARC_FRAC_BITS) - dq;
          if (diff < 0) diff = -diff;
          if (diff < min_diff && tt[i][r] / T == cur_time / T) {
            min_diff = diff;
            related = i;
            related_rf = from_rf;
          }
        }
        if (related < 0) {
          // getDefaultOutputPinInternalEnergyVal (:511-555)
          int n_table = 0;
          real tot = 0;
          double totm = 0.0;
          for (int i = 0; i < nin; ++i) {
            for (int frf = 0; frf < 2; ++frf) {
              double slew = d->pin_slew[2 * (pb + i) + frf];
              if (!std::isfinite(slew)) continue;
              for (int k : ci.groups[p]) {
                const gtp_internal_power& ip = x.lib->ipowers[k];
                if (ip.related_pin != i) continue;
                double m = 0;
                tot += (real)group_power(x, ip, to_rf, slew, cap, &m);
                totm += m;
                ++n_table;
              }
            }
          }
          if (n_table) {
            E = tot / (real)n_table;
            M = totm / n_table;
          }
        } else {
          double slew = d->pin_slew[2 * (pb + related) + related_rf];
          if (std::isfinite(slew)) {
            for (int k : ci.groups[p]) {
              const gtp_internal_power& ip = x.lib->ipowers[k];
              if (ip.related_pin == related && when_holds(ip, prev_idx)) {
                E = (real)group_power(x, ip, to_rf, slew, cap, &M);
                break;
              }
            }
          }
        }
        // switching (:856-869)
        int rx2 = rise_x2(pv, cv);
        if (rx2 > 0) {
          real Es = (real)0.5 * (real)rx2 * sw[p];
          if (ratio > 0) {
            real e = ratio * Es;
            gp[GTP_GLITCH_SWITCHING] += e / denom;
            ga[GTP_GLITCH_SWITCHING] += std::fabs((double)e) / denom_d;
            if (in_range) {
              cyc[GTP_WAVE_SWITCHING * C + cidx] += e / period;
              cyc[GTP_WAVE_GLITCH_SWITCHING * C + cidx] += e / period;
              cab[GTP_WAVE_SWITCHING * C + cidx] += std::fabs((double)e) / period_d;
              cab[GTP_WAVE_GLITCH_SWITCHING * C + cidx] += std::fabs((double)e) / period_d;
            }
          } else {
            gp[GTP_SWITCHING] += Es / denom;
            ga[GTP_SWITCHING] += std::fabs((double)Es) / denom_d;
            if (in_range) {
              cyc[GTP_WAVE_SWITCHING * C + cidx] += Es / period;
              cab[GTP_WAVE_SWITCHING * C + cidx] += std::fabs((double)Es) / period_d;
            }
          }
        }
      }
      E *= n_toggle;
      M *= n_toggle;
      if (ratio > 0) {
        real e = ratio * E;
        double m = (double)ratio * M;
        gp[GTP_GLITCH_INTERNAL] += e / denom;
        ga[GTP_GLITCH_INTERNAL] += m / denom_d;
        if (in_range) {
          cyc[GTP_WAVE_INTERNAL * C + cidx] += e / period;
          cyc[GTP_WAVE_GLITCH_INTERNAL * C + cidx] += e / period;
          cab[GTP_WAVE_INTERNAL * C + cidx] += m / period_d;
          cab[GTP_WAVE_GLITCH_INTERNAL * C + cidx] += m / period_d;
        }
      } else {
        gp[GTP_INTERNAL] += E / denom;
        ga[GTP_INTERNAL] += M / denom_d;
        if (in_range) {
          cyc[GTP_WAVE_INTERNAL * C + cidx] += E / period;
          cab[GTP_WAVE_INTERNAL * C + cidx] += M / period_d;
        }
      }
    }
    // wrap up (:885-889)
    prev_time = cur_time;
    s.prev = s.cur;
  }
  // final leakage (:892-897)
  real L = leak_of(s.prev);
  leak_split(L, prev_time, t_max);
  gp[GTP_LEAKAGE] += L * (real)(t_max - prev_time) / (real)t_max;
  ga[GTP_LEAKAGE] += std::fabs(L) * (double)(t_max - prev_time) / (double)t_max;
}

}  // namespace

extern "C" {

int32_t oracle_toggle_x2(int32_t prev, int32_t cur) { return toggle_x2(prev, cur); }
int32_t oracle_rise_x2(int32_t prev, int32_t cur) { return rise_x2(prev, cur); }

void oracle_glitch_flags(const int64_t* t, int64_t n, int64_t T, uint8_t* flags) {
  glitch_flags(t, n, T, flags);
}

double oracle_glitch_ratio(const int64_t* t, int64_t n, const uint8_t* flags, int64_t i,
                           double time_scale, double slew_rise, double slew_fall) {
  return glitch_ratio(t, n, flags, i, time_scale, slew_rise, slew_fall);
}

double oracle_table_value(const gtp_library* lib, int32_t table, double slew, double cap) {
  return table_value(lib, table, slew, cap, nullptr);
}

// power/cuda/Gate.cuh:354-376 setNCyclePerThreadByEventCount.
int64_t oracle_cycles_per_thread(int64_t n_event, int64_t n_cycle, int64_t e_target,
                                 int64_t parallelism_floor) {
  if (n_cycle <= 0 || n_event < 0) return -1;
  int64_t n_cycle_target = (n_cycle + parallelism_floor - 1) / parallelism_floor;
  n_cycle_target = std::max<int64_t>(n_cycle_target, 1);
  double rho = (double)n_event / (double)n_cycle;
  double rho_floor = (double)e_target / (double)n_cycle_target;
  int64_t v = (int64_t)std::llround((double)e_target / (rho + rho_floor));
  v = std::min(v, n_cycle);
  v = std::max<int64_t>(v, 1);
  return v;
}

int32_t oracle_run(const gtp_library* lib, const gtp_design* d, const gtp_trace* tr,
                   int32_t threads, int64_t gate_begin, int64_t gate_end, oracle_result* out) {
  if (!lib || !d || !tr || !out || tr->time_scale <= 0 || tr->clk_period <= 0 || tr->t_max <= 0)
    return GTP_ERR_ARG;
  Ctx x;
  x.lib = lib;
  x.d = d;
  x.tr = tr;
  x.scale = tr->time_scale;
  x.period = tr->clk_period;
  x.T = (int64_t)std::ceil(tr->clk_period / tr->time_scale);  // PowerUtils.cc:30
  // guard against period/scale landing a hair above an integer
  if (std::fabs(tr->clk_period / tr->time_scale - std::round(tr->clk_period / tr->time_scale)) < 1e-9)
    x.T = (int64_t)std::llround(tr->clk_period / tr->time_scale);
  x.C = (tr->t_max + x.T - 1) / x.T;
  x.cells.resize(lib->n_cell);
  for (int c = 0; c < lib->n_cell; ++c) {
    CellInfo& ci = x.cells[c];
    ci.n_pin = lib->cell_n_pin[c];
    ci.n_in = lib->cell_n_in[c];
    ci.clock_mask = lib->cell_clock_mask[c];
    if (lib->cell_leak_offset[c + 1] > lib->cell_leak_offset[c])
      ci.leak = lib->leak_values + lib->cell_leak_offset[c];
    ci.leak_default = lib->cell_leak_default[c];
    ci.groups.assign(ci.n_pin, {});
  }
  for (int k = 0; k < lib->n_ipower; ++k) {
    const gtp_internal_power& ip = lib->ipowers[k];
    if (ip.cell < 0 || ip.cell >= lib->n_cell || ip.pin < 0 || ip.pin >= x.cells[ip.cell].n_pin)
      return GTP_ERR_ARG;
    x.cells[ip.cell].groups[ip.pin].push_back(k);
  }
  const int64_t G = d->n_gate, C = x.C;
  if (gate_end < 0 || gate_end > G) gate_end = G;
  if (gate_begin < 0) gate_begin = 0;
  out->n_cycle = C;
  int nth = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  auto t0 = std::chrono::steady_clock::now();
  std::vector<Acc> accs(nth);
  std::vector<double> gp((size_t)5 * std::max<int64_t>(G, 1), 0.0), ga(gp.size(), 0.0);
  std::vector<uint64_t> tog(std::max<int64_t>(G, 1), 0), gl(tog.size(), 0);
  // static contiguous split per thread (Power.cc:284-294), deterministic order
  const int64_t ng = gate_end - gate_begin;
  std::vector<std::thread> pool;
  for (int t = 0; t < nth; ++t) {
    pool.emplace_back([&, t] {
      Acc& acc = accs[t];
      acc.cyc.assign((size_t)5 * C, (real)0);
      acc.cyc_abs.assign((size_t)5 * C, 0.0);
      GateScratch s;
      int64_t b = gate_begin + ng * t / nth, e = gate_begin + ng * (t + 1) / nth;
      for (int64_t g = b; g < e; ++g) {
        real p5[5] = {0, 0, 0, 0, 0};
        double a5[5] = {0, 0, 0, 0, 0};
        uint64_t tg = 0, gg = 0;
        eval_gate(x, g, s, acc, p5, a5, &tg, &gg);
        for (int k = 0; k < 5; ++k) {
          gp[(size_t)k * G + g] = (double)p5[k];
          ga[(size_t)k * G + g] = a5[k];
        }
        tog[g] = tg;
        gl[g] = gg;
      }
    });
  }
  for (auto& th : pool) th.join();
  // merge per-thread waveforms in thread order (Power.cc:322-329)
  std::vector<real> cyc_r((size_t)5 * C, (real)0);
  std::vector<double> cab((size_t)5 * C, 0.0);
  for (int t = 0; t < nth; ++t)
    for (size_t k = 0; k < cyc_r.size(); ++k) {
      cyc_r[k] += accs[t].cyc[k];
      cab[k] += accs[t].cyc_abs[k];
    }
  std::vector<double> cyc(cyc_r.begin(), cyc_r.end());
  auto t1 = std::chrono::steady_clock::now();
  out->wall_s = std::chrono::duration<double>(t1 - t0).count();
  if (out->gate_power) std::memcpy(out->gate_power, gp.data(), sizeof(double) * 5 * G);
  if (out->gate_abs) std::memcpy(out->gate_abs, ga.data(), sizeof(double) * 5 * G);
  if (out->cycle_power) std::memcpy(out->cycle_power, cyc.data(), sizeof(double) * 5 * C);
  if (out->cycle_abs) std::memcpy(out->cycle_abs, cab.data(), sizeof(double) * 5 * C);
  if (out->gate_toggle_x2) std::memcpy(out->gate_toggle_x2, tog.data(), sizeof(uint64_t) * G);
  if (out->gate_glitch) std::memcpy(out->gate_glitch, gl.data(), sizeof(uint64_t) * G);
  return GTP_OK;
}

}  // extern "C"
