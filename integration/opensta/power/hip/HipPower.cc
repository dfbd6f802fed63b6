// HipPower -- see HipPower.hh.
//
// Each power() call flattens the OpenSTA state the reference's CPU path reads
// per gate (power/TimeBasedPower.cc:703-901 findTimeBasedAllPower) into the
// arrays of include/gtpower.h, runs the engine (gtp_set_design, gtp_set_trace,
// gtp_run; gtp_run_sliced for traces of 2^29 - 1 ticks or more) and fills the
// six PowerResults, the per-instance results behind the cell report and the
// top instance's per-cycle waveforms that PowerResult::printPerCycleResults
// writes (include/sta/PowerClass.hh:123-149).
#include "power/hip/HipPower.hh"

#include <cmath>
#include <limits>
#include <map>
#include <utility>

#include "DcalcAnalysisPt.hh"
#include "Corner.hh"
#include "FuncExpr.hh"
#include "Graph.hh"
#include "GraphDelayCalc.hh"
#include "InternalPower.hh"
#include "LeakagePower.hh"
#include "Liberty.hh"
#include "Network.hh"
#include "PortDirection.hh"
#include "PowerUtils.hh"
#include "Report.hh"
#include "TableModel.hh"
#include "TimingArc.hh"
#include "Transition.hh"
#include "GlobalConfig.hh"

namespace gtpower {

using sta::Corner;
using sta::DcalcAnalysisPt;
using sta::Edge;
using sta::FuncExpr;
using sta::Instance;
using sta::InstancePinIterator;
using sta::InternalPower;
using sta::InternalPowerSeq;
using sta::LeafInstanceIterator;
using sta::LeakagePower;
using sta::LibertyCell;
using sta::LibertyPort;
using sta::MinMax;
using sta::Pin;
using sta::RiseFall;
using sta::Table;
using sta::TableAxis;
using sta::TableAxisVariable;
using sta::TimingArc;
using sta::Vertex;

namespace {

// Index of a port among the gate's ordered pins, by name (the reference keys
// its pin map by port name, power/TimeBasedPower.cc:719-731); -1 when absent.
int32_t portIndex(const std::vector<const LibertyPort *> &ports, const LibertyPort *port)
{
  if (!port)
    return -1;
  for (size_t i = 0; i < ports.size(); ++i)
    if (std::string(ports[i]->name()) == port->name())
      return static_cast<int32_t>(i);
  return -1;
}

// A `when` of an internal_power group on gate state s (bit p = pin p is 1;
// X counts as 0): the full Boolean function, as evalWhen
// (power/PowerUtils.cc:155-187).  Ports outside the gate make it false.
bool whenHolds(const FuncExpr *e, const std::vector<const LibertyPort *> &ports, uint64_t s)
{
  if (!e)
    return true;
  switch (e->op()) {
  case FuncExpr::op_port: {
    const int32_t p = portIndex(ports, e->port());
    return p >= 0 && ((s >> p) & 1u);
  }
  case FuncExpr::op_not:
    return !whenHolds(e->left(), ports, s);
  case FuncExpr::op_and:
    return whenHolds(e->left(), ports, s) && whenHolds(e->right(), ports, s);
  case FuncExpr::op_or:
    return whenHolds(e->left(), ports, s) || whenHolds(e->right(), ports, s);
  case FuncExpr::op_xor:
    return whenHolds(e->left(), ports, s) != whenHolds(e->right(), ports, s);
  case FuncExpr::op_one:
    return true;
  case FuncExpr::op_zero:
    return false;
  }
  return false;
}

// A leakage `when` as the reference reads it (getStateIdx,
// power/TimeBasedPower.cc:557-587, with collectWhenPinStates,
// power/PowerUtils.cc:110-150): a conjunction of pin literals, possibly
// through negations and De Morgan; need[p] = 0 / 1 for a fixed pin, -1 for a
// free one.  false when the expression is not such a conjunction (the
// reference reports it and matches no state).
bool whenConjunction(const FuncExpr *e, bool negated, const std::vector<const LibertyPort *> &ports,
                     std::vector<int> &need)
{
  switch (e->op()) {
  case FuncExpr::op_port: {
    const int32_t p = portIndex(ports, e->port());
    if (p < 0)
      return false;
    const int want = negated ? 0 : 1;
    if (need[p] >= 0 && need[p] != want)
      return false;
    need[p] = want;
    return true;
  }
  case FuncExpr::op_not:
    return whenConjunction(e->left(), !negated, ports, need);
  case FuncExpr::op_one:
    return !negated;
  case FuncExpr::op_zero:
    return negated;
  case FuncExpr::op_and:
  case FuncExpr::op_or:
    if ((e->op() == FuncExpr::op_and) != negated)
      return whenConjunction(e->left(), negated, ports, need) && whenConjunction(e->right(), negated, ports, need);
    return false;
  case FuncExpr::op_xor:
    return false;
  }
  return false;
}

int32_t axisVar(const TableAxis *axis)
{
  if (axis->variable() == TableAxisVariable::input_transition_time)
    return GTP_AXIS_SLEW;
  if (axis->variable() == TableAxisVariable::total_output_net_capacitance)
    return GTP_AXIS_CAP;
  return -1;
}

// VCD value characters as the reference maps them (power/Power.hh:114);
// 'U' is taken as X, as the engine's own VCD reader does.
uint8_t vcdValue(char c)
{
  switch (c) {
  case '0':
    return 0;
  case '1':
    return 1;
  case 'Z':
  case 'z':
    return 3;
  default:
    return 2;
  }
}

void setResult(PowerResult &r, const double *v)
{
  r.internal() = v[GTP_INTERNAL] + v[GTP_GLITCH_INTERNAL];
  r.glitchInternal() = v[GTP_GLITCH_INTERNAL];
  r.switching() = v[GTP_SWITCHING] + v[GTP_GLITCH_SWITCHING];
  r.glitchSwitching() = v[GTP_GLITCH_SWITCHING];
  r.leakage() = v[GTP_LEAKAGE];
}

constexpr int64_t kLongTraceTicks = (int64_t{1} << 29) - 1;

}  // namespace

HipPower::HipPower(sta::StaState *sta) :
  Power(sta)
{
  gtp_options o;
  gtp_options_init(&o);
  if (G_CONFIG.nums.cuda_device_id >= 0)
    o.device = G_CONFIG.nums.cuda_device_id;            // util/GlobalConfig.hh:45
  o.host_threads = G_CONFIG.nums.multi_thread_number;  // :44
  o.max_event_num = G_CONFIG.nums.max_event_num;       // :33
  if (gtp_create(&o, &ctx_) != GTP_OK)
    report_->critical(1700, "HIP power: %s", gtp_create_error());
}

HipPower::~HipPower()
{
  gtp_destroy(ctx_);
}

void HipPower::check(gtp_status st, const char *what) const
{
  if (st != GTP_OK)
    report_->critical(1701, "HIP power: %s: %s", what, gtp_last_error(ctx_));
}

// The report class of an instance, in the order Power::power tests them
// (power/Power.cc:335-346).
int32_t HipPower::instanceClass(const Instance *inst)
{
  const LibertyCell *cell = network_->libertyCell(inst);
  if (cell->isMacro() || cell->isMemory() || cell->interfaceTiming())
    return GTP_CLASS_MACRO;
  if (cell->isPad())
    return GTP_CLASS_PAD;
  if (inClockNetwork(inst))
    return GTP_CLASS_CLOCK;
  if (cell->hasSequentials())
    return GTP_CLASS_SEQUENTIAL;
  return GTP_CLASS_COMBINATIONAL;
}

void HipPower::orderedPins(const Instance *inst, std::vector<const Pin *> &pins, int32_t &n_in) const
{
  pins.clear();
  for (int pass = 0; pass < 2; ++pass) {
    InstancePinIterator *it = network_->pinIterator(inst);
    while (it->hasNext()) {
      const Pin *pin = it->next();
      const LibertyPort *port = network_->libertyPort(pin);
      if (!port)
        continue;
      if (pass == 0 ? port->direction()->isAnyInput() : port->direction()->isAnyOutput())
        pins.push_back(pin);
    }
    delete it;
    if (pass == 0)
      n_in = static_cast<int32_t>(pins.size());
  }
}

// Internal power groups of every pin, in Liberty order per pin
// (LibertyCell::internalPowers(port), as findInputInternalVal and
// findOutputInternalVal take them, power/TimeBasedPower.cc:341-509), with
// each rise/fall table sampled at its own breakpoints through
// InternalPower::power -- the value the reference interpolates there,
// operating-condition scaling included (liberty/InternalPower.cc:112-124).
void HipPower::flattenTables(const LibertyCell *corner_cell, const std::vector<const LibertyPort *> &ports,
                             const DcalcAnalysisPt *dcalc_ap, int32_t cell_index, FlatLibrary &lib)
{
  const int32_t n_pin = static_cast<int32_t>(ports.size());
  for (int32_t p = 0; p < n_pin; ++p) {
    const LibertyPort *corner_port = ports[p]->cornerPort(dcalc_ap);
    if (!corner_port)
      continue;
    InternalPowerSeq pwrs;
    corner_cell->internalPowers(corner_port, pwrs);
    for (const InternalPower *pwr : pwrs) {
      gtp_internal_power ip{};
      ip.cell = cell_index;
      ip.pin = p;
      ip.related_pin = portIndex(ports, pwr->relatedPort());
      // when_truth bit s: the condition on gate state s (6 pins at most)
      uint64_t truth = 0;
      if (n_pin <= 6) {
        const uint64_t n_s = uint64_t{1} << n_pin;
        for (uint64_t s = 0; s < n_s; ++s)
          if (whenHolds(pwr->when(), ports, s))
            truth |= uint64_t{1} << s;
      } else {
        truth = pwr->when() ? 0 : ~uint64_t{0};  // a `when` here is refused by gtp_set_design
      }
      ip.when_truth = truth;
      int32_t tab[2] = {-1, -1};
      for (RiseFall *rf : RiseFall::range()) {
        const auto table = pwr->lookupTable(rf);
        gtp_table t{};
        const TableAxis *a1 = table->order() >= 1 ? table->axis1() : nullptr;
        const TableAxis *a2 = table->order() >= 2 ? table->axis2() : nullptr;
        if (table->order() > 2)
          report_->critical(1702, "HIP power: 3-D internal power tables are not supported");
        t.n1 = a1 ? static_cast<int32_t>(a1->size()) : 1;
        t.n2 = a2 ? static_cast<int32_t>(a2->size()) : 0;
        t.var1 = a1 ? axisVar(a1) : GTP_AXIS_SLEW;
        t.var2 = a2 ? axisVar(a2) : GTP_AXIS_CAP;
        if (t.var1 < 0 || t.var2 < 0)
          report_->critical(1703, "HIP power: internal power table over an unsupported variable");
        t.axis1 = static_cast<int64_t>(lib.table_data.size());
        for (int32_t i = 0; i < t.n1; ++i)
          lib.table_data.push_back(a1 ? a1->axisValue(i) : 0.0);
        t.axis2 = static_cast<int64_t>(lib.table_data.size());
        for (int32_t j = 0; j < t.n2; ++j)
          lib.table_data.push_back(a2->axisValue(j));
        t.values = static_cast<int64_t>(lib.table_data.size());
        for (int32_t i = 0; i < t.n1; ++i) {
          for (int32_t j = 0; j < std::max(t.n2, 1); ++j) {
            float slew = 0.0f, cap = 0.0f;
            const float x1 = a1 ? a1->axisValue(i) : 0.0f;
            (t.var1 == GTP_AXIS_SLEW ? slew : cap) = x1;
            if (a2)
              (t.var2 == GTP_AXIS_SLEW ? slew : cap) = a2->axisValue(j);
            lib.table_data.push_back(pwr->power(rf, dcalc_ap->operatingConditions(), slew, cap));
          }
        }
        tab[rf->index()] = static_cast<int32_t>(lib.tables.size());
        lib.tables.push_back(t);
      }
      ip.rise_table = tab[RiseFall::riseIndex()];
      ip.fall_table = tab[RiseFall::fallIndex()];
      lib.ipowers.push_back(ip);
    }
  }
}

// State-indexed leakage (Power::getLeakagePower, power/TimeBasedPower.cc:258-301):
// 2^n_pin states starting at the cell default (or 0), each leakage_power group
// overwriting the states its conjunction matches; no table for cells without
// groups or above the reference's pin limit (then the default, or the single
// group's value).
void HipPower::flattenLeakage(const LibertyCell *cell, const LibertyCell *corner_cell,
                              const std::vector<const LibertyPort *> &ports, FlatLibrary &lib)
{
  const int32_t n_pin = static_cast<int32_t>(ports.size());
  float dflt = 0.0f;
  bool exists = false;
  cell->leakagePower(dflt, exists);
  const auto &pwrs = corner_cell->leakagePowers();
  if (n_pin > G_CONFIG.nums.max_n_pin_for_leakage_power) {
    if (pwrs.size() == 1) {
      dflt = pwrs[0]->power();
      exists = true;
    }
    lib.leak_default.push_back(exists ? dflt : 0.0);
    lib.leak_off.push_back(lib.leak_off.back());
    return;
  }
  lib.leak_default.push_back(exists ? dflt : 0.0);
  if (pwrs.empty()) {
    lib.leak_off.push_back(lib.leak_off.back());
    return;
  }
  const int64_t base = static_cast<int64_t>(lib.leak_values.size());
  const uint64_t n_s = uint64_t{1} << n_pin;
  lib.leak_values.resize(base + n_s, exists ? dflt : 0.0);
  for (const LeakagePower *pwr : pwrs) {
    std::vector<int> need(n_pin, -1);
    if (pwr->when() && !whenConjunction(pwr->when(), false, ports, need))
      continue;
    for (uint64_t s = 0; s < n_s; ++s) {
      bool match = true;
      for (int32_t p = 0; p < n_pin && match; ++p)
        match = need[p] < 0 || static_cast<int>((s >> p) & 1u) == need[p];
      if (match)
        lib.leak_values[base + s] = pwr->power();
    }
  }
  lib.leak_off.push_back(base + static_cast<int64_t>(n_s));
}

int32_t HipPower::flatCell(const Instance *inst, const std::vector<const Pin *> &pins, int32_t n_in,
                           const Corner *corner, FlatLibrary &lib)
{
  const LibertyCell *cell = network_->libertyCell(inst);
  const CellKey key{cell, instanceClass(inst)};
  const auto found = cell_index_.find(key);
  if (found != cell_index_.end())
    return found->second;
  const int32_t c = static_cast<int32_t>(lib.n_pin.size());
  cell_index_.emplace(key, c);
  std::vector<const LibertyPort *> ports;
  uint32_t clock_mask = 0;
  for (const Pin *pin : pins) {
    const LibertyPort *port = network_->libertyPort(pin);
    if (port->isClock() && ports.size() < 32)
      clock_mask |= 1u << ports.size();
    ports.push_back(port);
  }
  lib.n_pin.push_back(static_cast<int32_t>(ports.size()));
  lib.n_in.push_back(n_in);
  lib.cls.push_back(key.cls);
  lib.clock_mask.push_back(clock_mask);
  const DcalcAnalysisPt *dcalc_ap = corner->findDcalcAnalysisPt(MinMax::max());
  const LibertyCell *corner_cell = const_cast<LibertyCell *>(cell)->cornerCell(corner, MinMax::max());
  flattenLeakage(cell, corner_cell, ports, lib);
  flattenTables(corner_cell, ports, dcalc_ap, c, lib);
  return c;
}

void HipPower::flatten(const Corner *corner, FlatLibrary &lib, FlatDesign &des, FlatTrace &trc)
{
  const DcalcAnalysisPt *dcalc_ap = corner->findDcalcAnalysisPt(MinMax::max());
  cell_index_.clear();
  // one net per VCD variable bit that some pin reads (the reference's
  // (var, bit) pairs, power/cuda/CudaPower.cu:2001-2010)
  std::map<std::pair<const VcdValue *, int>, int32_t> net_of;
  std::vector<std::pair<const VcdValue *, std::pair<int, NEeventVal>>> nets;
  std::vector<const Pin *> pins;
  LeafInstanceIterator *it = network_->leafInstanceIterator();
  while (it->hasNext()) {
    const Instance *inst = it->next();
    const LibertyCell *cell = network_->libertyCell(inst);
    if (!cell)
      continue;
    int32_t n_in = 0;
    orderedPins(inst, pins, n_in);
    const int32_t c = flatCell(inst, pins, n_in, corner, lib);
    const LibertyCell *corner_cell = const_cast<LibertyCell *>(cell)->cornerCell(corner, MinMax::max());
    des.gates.push_back(inst);
    des.gate_cell.push_back(c);
    for (const Pin *pin : pins) {
      const LibertyPort *port = network_->libertyPort(pin);
      const bool is_out = port->direction()->isAnyOutput();
      Vertex *v = graph_->pinLoadVertex(pin);
      // activity of the pin (findActivity, as power/TimeBasedPower.cc:722-731)
      const PwrActivity act = findActivity(pin);
      int32_t net = -1;
      if (act.nEvent() > 0 && act.vcdValues()) {
        const auto k = std::make_pair(act.vcdValues(), act.valueBit());
        const auto ins = net_of.emplace(k, static_cast<int32_t>(nets.size()));
        if (ins.second)
          nets.push_back({act.vcdValues(), {act.valueBit(), act.nEvent()}});
        net = ins.first->second;
      }
      des.pin_net.push_back(net);
      des.pin_slew.push_back(getSlew(v, RiseFall::rise(), corner));  // Power::getSlew, power/Power.cc:1200-1212
      des.pin_slew.push_back(getSlew(v, RiseFall::fall(), corner));
      des.pin_load_cap.push_back(is_out ? graph_delay_calc_->loadCap(pin, dcalc_ap) : 0.0);
      des.pin_voltage.push_back(is_out ? portVoltage(corner_cell, port, dcalc_ap) : 0.0);
      des.pin_default.push_back(static_cast<int8_t>(getVertexDefaultState(v)));  // PowerUtils.cc:50-65
    }
    // arc delays [out][in][from_rf][to_rf] (graph_->gateEdgeArc + arcDelay,
    // as findOutputInternalVal looks them up, power/TimeBasedPower.cc:432-446)
    const int32_t n_out = static_cast<int32_t>(pins.size()) - n_in;
    for (int32_t o = 0; o < n_out; ++o)
      for (int32_t i = 0; i < n_in; ++i)
        for (const RiseFall *from_rf : {RiseFall::rise(), RiseFall::fall()})
          for (const RiseFall *to_rf : {RiseFall::rise(), RiseFall::fall()}) {
            Edge *edge = nullptr;
            const TimingArc *arc = nullptr;
            graph_->gateEdgeArc(pins[i], from_rf, pins[n_in + o], to_rf, edge, arc);
            des.arc_delay.push_back(edge && arc ? static_cast<double>(graph_->arcDelay(edge, arc, dcalc_ap->index()))
                                                : std::numeric_limits<double>::quiet_NaN());
          }
    des.gate_pin_off.push_back(static_cast<int64_t>(des.pin_net.size()));
    des.gate_arc_off.push_back(static_cast<int64_t>(des.arc_delay.size()));
  }
  delete it;

  // the trace: every event of each net's variable, at its bit
  // (VcdValue::value(bit), as the reference's per-pin activity reads it)
  for (const auto &n : nets) {
    const VcdValue *vals = n.first;
    const int bit = n.second.first;
    const NEeventVal cnt = n.second.second;
    for (NEeventVal e = 0; e < cnt; ++e) {
      trc.time.push_back(vals[e].time());
      trc.value.push_back(vcdValue(vals[e].value(bit)));
    }
    trc.net_off.push_back(static_cast<int64_t>(trc.time.size()));
  }

  gtp_library &L = lib.view;
  L.n_cell = static_cast<int32_t>(lib.n_pin.size());
  L.cell_n_pin = lib.n_pin.data();
  L.cell_n_in = lib.n_in.data();
  L.cell_class = lib.cls.data();
  L.cell_clock_mask = lib.clock_mask.data();
  L.cell_leak_offset = lib.leak_off.data();
  L.leak_values = lib.leak_values.data();
  L.cell_leak_default = lib.leak_default.data();
  L.n_table = static_cast<int32_t>(lib.tables.size());
  L.tables = lib.tables.data();
  L.table_data = lib.table_data.data();
  L.n_ipower = static_cast<int32_t>(lib.ipowers.size());
  L.ipowers = lib.ipowers.data();

  gtp_design &D = des.view;
  D.n_gate = static_cast<int64_t>(des.gates.size());
  D.n_pin = static_cast<int64_t>(des.pin_net.size());
  D.n_net = static_cast<int64_t>(nets.size());
  D.gate_cell = des.gate_cell.data();
  D.gate_pin_offset = des.gate_pin_off.data();
  D.gate_arc_offset = des.gate_arc_off.data();
  D.arc_delay = des.arc_delay.data();
  D.pin_net = des.pin_net.data();
  D.pin_slew = des.pin_slew.data();
  D.pin_load_cap = des.pin_load_cap.data();
  D.pin_voltage = des.pin_voltage.data();
  D.pin_default_state = des.pin_default.data();

  gtp_trace &T = trc.view;
  T.n_net = static_cast<int64_t>(nets.size());
  T.net_event_offset = trc.net_off.data();
  T.event_time = trc.time.data();
  T.event_value = trc.value.data();
  T.t_max = vcd_.timeMax();
  T.time_scale = vcd_.timeScale();
  T.clk_period = clk_period_;
}

void HipPower::power(const Corner *corner, PowerResult &total, PowerResult &sequential,
                     PowerResult &combinational, PowerResult &clock, PowerResult &macro, PowerResult &pad)
{
  total.clear();
  sequential.clear();
  combinational.clear();
  clock.clear();
  macro.clear();
  pad.clear();
  cell_power_results_.clear();
  ensureActivities();

  FlatLibrary lib;
  FlatDesign des;
  FlatTrace trc;
  flatten(corner, lib, des, trc);
  check(gtp_set_design(ctx_, &lib.view, &des.view), "gtp_set_design");

  const int64_t G = des.view.n_gate;
  const double ratio = clk_period_ / vcd_.timeScale();
  int64_t T = static_cast<int64_t>(std::ceil(ratio));
  if (std::fabs(ratio - std::round(ratio)) < 1e-9 * std::max(1.0, ratio))
    T = static_cast<int64_t>(std::llround(ratio));
  const int64_t C = (trc.view.t_max + T - 1) / T;
  std::vector<double> gate_power(static_cast<size_t>(GTP_N_COMPONENT) * G);
  std::vector<double> cycle_power(static_cast<size_t>(GTP_N_WAVE) * C);
  double totals[GTP_N_COMPONENT], classes[GTP_N_CLASS * GTP_N_COMPONENT];
  gtp_result r{gate_power.data(), cycle_power.data(), totals, classes, nullptr, nullptr, 0, 0.0, 0.0};
  if (trc.view.t_max >= kLongTraceTicks) {
    // longer than one event word spans: in cycle intervals (any length)
    check(gtp_run_sliced(ctx_, &trc.view, 0, &r), "gtp_run_sliced");
  } else {
    check(gtp_set_trace(ctx_, &trc.view), "gtp_set_trace");
    check(gtp_run(ctx_, &r), "gtp_run");
  }

  // totals by class (Power::power, power/Power.cc:335-347)
  setResult(total, totals);
  setResult(combinational, classes + GTP_N_COMPONENT * GTP_CLASS_COMBINATIONAL);
  setResult(sequential, classes + GTP_N_COMPONENT * GTP_CLASS_SEQUENTIAL);
  setResult(clock, classes + GTP_N_COMPONENT * GTP_CLASS_CLOCK);
  setResult(macro, classes + GTP_N_COMPONENT * GTP_CLASS_MACRO);
  setResult(pad, classes + GTP_N_COMPONENT * GTP_CLASS_PAD);
  // per instance, for the cell report
  cell_power_results_.reserve(G);
  for (int64_t g = 0; g < G; ++g) {
    double v[GTP_N_COMPONENT];
    for (int k = 0; k < GTP_N_COMPONENT; ++k)
      v[k] = gate_power[static_cast<size_t>(k) * G + g];
    PowerResult pr;
    setResult(pr, v);
    cell_power_results_.emplace_back(des.gates[g], pr);
  }
  // the top instance's per-cycle waveforms (Power::power, power/Power.cc:244-250)
  const Instance *top = network_->topInstance();
  const NPeriodVal n_period = static_cast<NPeriodVal>(r.n_cycle);
  total.initLeakagePowerClkedWaveform(top, clk_period_, n_period);
  total.initInternalPowerClkedWaveform(top, clk_period_, n_period);
  total.initGlitchInternalPowerClkedWaveform(top, clk_period_, n_period);
  total.initSwitchingPowerClkedWaveform(top, clk_period_, n_period);
  total.initGlitchSwitchingPowerClkedWaveform(top, clk_period_, n_period);
  PowerClkedWaveform *waves[GTP_N_WAVE] = {
    &total.findLeakagePowerClkedWaveform(top), &total.findInternalPowerClkedWaveform(top),
    &total.findGlitchInternalPowerClkedWaveform(top), &total.findSwitchingPowerClkedWaveform(top),
    &total.findGlitchSwitchingPowerClkedWaveform(top)};
  for (int w = 0; w < GTP_N_WAVE; ++w) {
    std::vector<PowerVal> &wave = waves[w]->waveform();
    wave.resize(static_cast<size_t>(r.n_cycle));
    for (int64_t c = 0; c < r.n_cycle; ++c)
      wave[c] = static_cast<PowerVal>(cycle_power[static_cast<size_t>(w) * C + c]);
  }
}

void HipPower::printCellRes() const
{
  Power::printCellRes();
}

}  // namespace gtpower
