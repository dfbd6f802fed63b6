// HipPower -- the time-based power backend of GTPower on an AMD Instinct GPU
// (gfx950), through the C ABI of this repository (include/gtpower.h).
//
// Drop-in for the reference's CudaPower (power/cuda/CudaPower.hh:24-33): it
// overrides the same virtual entry points of gtpower::Power
// (power/Power.hh:80-87 power(), :109 printCellRes()) and is chosen in
// Sta::makePower (search/Sta.cc:502-509).  It holds no device code: the
// kernels are in libgtpower_hip.so.  What OpenSTA knows implicitly (netlist,
// Liberty tables, slews, load caps, arc delays, VCD waveforms) is flattened
// into the library's explicit arrays on each power() call.
//
// This file belongs in the reference tree (power/hip/); it compiles only
// there, against OpenSTA's headers (power/Power.hh includes the CMake-made
// StaConfig.hh).
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "Power.hh"
#include "gtpower.h"

namespace gtpower {

class HipPower : public Power
{
public:
  explicit HipPower(sta::StaState *sta);
  ~HipPower() override;
  HipPower(const HipPower &) = delete;
  HipPower &operator=(const HipPower &) = delete;

  void power(const sta::Corner *corner,
             // Return values.
             PowerResult &total,
             PowerResult &sequential,
             PowerResult &combinational,
             PowerResult &clock,
             PowerResult &macro,
             PowerResult &pad) override;
  // Writes the cell report of the last power() call in the reference's format
  // (Power::printCellRes, power/Power.cc:388-421) from cell_power_results_.
  void printCellRes() const override;

private:
  // The flattened inputs of one power() call; every pointer in the gtp_*
  // views points into these vectors.
  struct FlatLibrary {
    std::vector<int32_t> n_pin, n_in, cls;
    std::vector<uint32_t> clock_mask;
    std::vector<int64_t> leak_off{0};
    std::vector<double> leak_values, leak_default;
    std::vector<gtp_table> tables;
    std::vector<double> table_data;
    std::vector<gtp_internal_power> ipowers;
    gtp_library view{};
  };
  struct FlatDesign {
    std::vector<const sta::Instance *> gates;  // gate index -> instance
    std::vector<int32_t> gate_cell;
    std::vector<int64_t> gate_pin_off{0}, gate_arc_off{0};
    std::vector<double> arc_delay;
    std::vector<int32_t> pin_net;
    std::vector<double> pin_slew, pin_load_cap, pin_voltage;
    std::vector<int8_t> pin_default;
    gtp_design view{};
  };
  struct FlatTrace {
    std::vector<int64_t> net_off{0}, time;
    std::vector<uint8_t> value;
    gtp_trace view{};
  };
  // A cell type of the flat library: a Liberty cell in one report class
  // (clock-network membership is a property of the instance, so a buffer in
  // the clock tree and one outside it are two flat cells).
  struct CellKey {
    const sta::LibertyCell *cell;
    int32_t cls;
    bool operator==(const CellKey &o) const { return cell == o.cell && cls == o.cls; }
  };
  struct CellKeyHash {
    size_t operator()(const CellKey &k) const
    {
      return std::hash<const void *>()(k.cell) * 31u + static_cast<size_t>(k.cls);
    }
  };

  void check(gtp_status st, const char *what) const;
  int32_t instanceClass(const sta::Instance *inst);
  // Pins of an instance, inputs first, then outputs
  // (Power::getPinInformation, power/TimeBasedPower.cc:223-256).
  void orderedPins(const sta::Instance *inst, std::vector<const sta::Pin *> &pins, int32_t &n_in) const;
  int32_t flatCell(const sta::Instance *inst, const std::vector<const sta::Pin *> &pins, int32_t n_in,
                   const sta::Corner *corner, FlatLibrary &lib);
  void flattenTables(const sta::LibertyCell *corner_cell, const std::vector<const sta::LibertyPort *> &ports,
                     const sta::DcalcAnalysisPt *dcalc_ap, int32_t cell_index, FlatLibrary &lib);
  void flattenLeakage(const sta::LibertyCell *cell, const sta::LibertyCell *corner_cell,
                      const std::vector<const sta::LibertyPort *> &ports, FlatLibrary &lib);
  void flatten(const sta::Corner *corner, FlatLibrary &lib, FlatDesign &des, FlatTrace &trc);

  gtp_ctx *ctx_ = nullptr;
  std::unordered_map<CellKey, int32_t, CellKeyHash> cell_index_;
};

}  // namespace gtpower
