// Internal layouts shared by the host runtime (gtp_api.cpp), the gfx950
// kernels (gtp_kernels.hip) and the device VCD ingest (gtp_vcd.hip).
//
// Device event store: one 32-bit word per event, per-net CSR order,
//   word = (time << 3) | (glitch << 2) | value,
//   time < 2^29 - 1 ticks, value in {0,1,2,3}, glitch set by the glitch
//   annotation when the event's glitch pulse width is nonzero.
// Times compare as words (the low bits are below the time bits), so a
// lower_bound on time is a lower_bound on words.  EV_INF is past every event.
// The array is padded by 4 words so that 16-byte loads of the last quad stay
// inside it.
//
// Results are accumulated as signed 64-bit fixed point: every FP32-evaluated
// term x is added as rint(x / quantum), with a power-of-two quantum such that
// one energy term is at most 2^36 quanta (a term 1e-5 of the design's largest
// is still resolved to 1e-6) and one leakage value at most 2^30 leakage
// quanta.  The sums are then exact integer sums: order-free and
// bit-reproducible.  Per-gate leakage is the exact integer sum of (state
// leakage in quanta) * (ticks in that state).
#pragma once
#include <vector>
#include <stdint.h>
#include <stdlib.h>

#include <hip/hip_runtime.h>

namespace gtp {

// Device array owned by the host runtime.
template <class T>
struct DArr {
  T* p = nullptr;
  size_t n = 0;
  DArr() = default;
  DArr(const DArr&) = delete;
  DArr& operator=(const DArr&) = delete;
  ~DArr() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    if (count == n && p) return hipSuccess;
    release();
    if (count == 0) return hipSuccess;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    n = count;
    return hipSuccess;
  }
  // at least count elements, keeping a larger buffer (no free, so no
  // implicit device synchronisation, when the size does not grow)
  hipError_t reserve(size_t count) {
    if (p && count <= n) return hipSuccess;
    return alloc(count);
  }
  hipError_t upload(const T* src, size_t count, hipStream_t s) {
    hipError_t e = alloc(count);
    if (e != hipSuccess || count == 0) return e;
    return hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s);
  }
};

// Page-locked host memory (hipHostMalloc), grow-only: host -> device copies
// from it run as DMA at PCIe rate, without the runtime's staging through its
// own pinned buffer.
template <typename T>
struct Pinned {
  T* p = nullptr;
  size_t n = 0;
  Pinned() = default;
  Pinned(const Pinned&) = delete;
  Pinned& operator=(const Pinned&) = delete;
  ~Pinned() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t reserve(size_t count) {
    if (p && count <= n) return hipSuccess;
    release();
    if (count == 0) return hipSuccess;
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&p), count * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    n = count;
    return hipSuccess;
  }
};

constexpr uint32_t EV_SHIFT = 3;                // time bits start here
constexpr uint32_t EV_GLITCH = 4u;              // the event carries a glitch pulse width
constexpr uint32_t EV_INF = 0xFFFFFFF8u;       // time field 2^29 - 1, glitch bit and value clear
constexpr uint32_t EV_TIME_INF = 0x1FFFFFFFu;
constexpr int64_t EV_PAD = 4;                   // words of padding after the event array
constexpr int MAXP_SMALL = 4;                   // generic kernel variant for cells with <= 4 pins
constexpr int MAXP_LARGE = 6;                   // ... <= 6 pins (conditional internal power: up to here)
constexpr int MAXP_WIDE = 8;                    // ... <= 8 pins (the widest sky130 cells)
constexpr int MAXP_HUGE = 16;                   // ... <= 16 pins (GTP_MAX_CELL_PINS)
constexpr int LEAK_LDS_MAX = 4096;              // leakage table entries the generic kernel stages in LDS
constexpr int MAX_CELLS_LDS = 64;               // generic kernel: cell types staged in LDS (more: read in place)
constexpr int TABLE_LDS_MAX = 12000;            // energy kernel: table values staged in LDS (more: read in place)
constexpr int ARC_FRAC_BITS = 16;               // related-pin comparison grid: 2^-16 tick
constexpr int ENERGY_FINE_BITS = 16;            // energy quanta per coarse unit of PowerArgs.inv_q_energy
constexpr int ENERGY_TERM_BITS = 36;            // one energy term <= 2^36 quanta
constexpr int32_t ARC_NONE = INT32_MIN;         // arc delays are int32 on that grid (< 2^15 ticks)
// Grid cap of kernels that stride over per-event work: a dispatch counts its
// work-items in 32 bits, so blocks x threads must stay below 2^32.
constexpr int64_t MAX_STRIDE_BLOCKS = 1 << 22;
// The cap in force: GTP_STRIDE_BLOCKS (1 .. MAX_STRIDE_BLOCKS) lowers it, so
// that small inputs run the stride loops too; read at every launch.
inline int64_t stride_blocks(int64_t n_block) {
  const char* e = getenv("GTP_STRIDE_BLOCKS");
  const long long v = e ? atoll(e) : 0;
  const int64_t cap = v > 0 && v < MAX_STRIDE_BLOCKS ? (int64_t)v : MAX_STRIDE_BLOCKS;
  return n_block < cap ? n_block : cap;
}

// Per-cycle accumulator components (LDS bins and partials).  Leakage: BIN_LD
// sums the leakage steps dL (quanta) taken in the cycle and BIN_LX sums
// dL * (t - cycle start) (quanta x ticks); the work item's flush turns them into
// S_c (leakage of the gates' states at the cycle start) and sum dL * (cycle end
// - t), so that the cycle's leakage integral is S_c * len + BIN_LX, exactly.
enum { BIN_INT = 0, BIN_GINT = 1, BIN_SW = 2, BIN_GSW = 3, BIN_LD = 4, BIN_LX = 5, N_BIN = 6 };
// Per-gate partial components (float-float).
enum { PG_INT = 0, PG_GINT = 1, PG_SW = 2, PG_GSW = 3, PG_LEAK = 4, N_PG = 5 };

// One cell type as the power kernel sees it.
struct CellDev {
  int32_t n_pin, n_in;
  uint32_t clock_mask;
  int32_t state_bits;        // n_pin when some internal power is `when`-conditioned, else 0
  int32_t leak_off;          // into the LDS leakage table (2^n_pin floats)
  int32_t n_slot;            // energy slots per gate of this cell
  int32_t slot_off[16];      // energy slot base of each pin
  uint32_t arc_mask;         // bit q: some gate of this cell has an arc from input q (single output)
  int32_t leak_shift;        // per-gate leakage sums of this cell count quanta of 2^-leak_shift common quanta
};

// One energy slot of a cell: which table to interpolate at which pin's slew.
struct SlotDesc {
  int32_t table;             // -1: the slot's energy is 0
  int16_t slew_pin;          // cell pin whose slew is the table's slew axis
  int8_t slew_rf;            // 0 rise, 1 fall
  int8_t cap_pin;            // cell pin whose load cap is the cap axis; -1 = 0.0
  int32_t list_off;          // default-energy slots: mean over list entries
  int32_t list_cnt;          // 0 = plain slot
};
struct ListEntry {
  int32_t table;
  int16_t slew_pin;
  int8_t slew_rf;
  int8_t cap_pin;
};

struct TableDev {             // offsets into the staged float table data
  int32_t n1, n2, var1, var2;
  int32_t axis1, axis2, values;
};

// One trace of a multi-trace run (batched traces, or one interval of a sliced
// run).  Times in its event words, its tiles and cycles are relative to its
// load's origin, tile_base whole tiles into the trace (0 for a whole trace):
// tile k here is tile k + tile_base of the trace, and its per-cycle values go
// to cycle c + cycle_base of a waveform of out_cycles cycles.
struct TraceDesc {
  int64_t net_base;        // first (virtual) net of the trace in the event store
  int64_t tile_idx_base;   // offset of the trace's [N][n_tile + 1] tile index
  int64_t t_max, n_cycle;  // relative to the origin
  int64_t cycle_base, out_cycles;
  int32_t n_tile, tile_base;
};

// Arguments of the power kernel.
struct PowerArgs {
  // gates in schedule order
  const int32_t* order;          // [n_sched] gate index
  int64_t n_sched;
  const int32_t* gate_cell;      // [G]
  const int64_t* gate_pin_off;   // [G+1]
  const int64_t* gate_arc_off;   // [G+1]
  const int64_t* gate_energy_off;// [G]
  // pins
  const int32_t* pin_net;        // [P]
  const float* pin_gk;           // [P] 2 / (slew_rise + slew_fall): glitch ratio factor (1/s)
  float time_scale;              // seconds per tick
  const float* pin_csw;          // [P] C_load * V^2 of output pins
  const int8_t* pin_default;     // [P]
  // per-gate energies and arcs
  const float* energy;           // [sum slots]
  const int32_t* arc_q;          // [sum arcs] delay in 2^-16 ticks, ARC_NONE = no arc
  // event store
  const int64_t* net_off;        // [N+1]
  const uint32_t* ev;            // [E + EV_PAD]
  const int32_t* tile_idx;       // [N][n_tile + 1] first event index >= tile start (net-relative);
                                 // [N][n_work_tiles] by list position when tile_list is set
  int32_t n_tile;
  int32_t tile_cycles;
  int64_t T;                     // ticks per cycle
  int64_t t_max;
  int64_t n_cycle;
  float inv_T;
  // floor(t / T) = (t * cyc_m) >> cyc_sh exactly for t < 2^29: cyc_sh = 29 +
  // ceil(log2 T), cyc_m = ceil(2^cyc_sh / T) (Granlund and Montgomery, 1994,
  // "Division by invariant integers using multiplication", theorem 4.2)
  uint32_t cyc_m;
  int32_t cyc_sh;
  // work decomposition
  int32_t n_range;               // gate ranges per tile (all classes)
  int32_t range_gates;           // gates per range; order[] is padded with -1 to whole ranges
  int32_t xcd_remap;             // 1: remap blockIdx so each XCD takes a contiguous block of work
  const int32_t* tile_list;      // tiles to evaluate (null: all n_tile tiles)
  const int32_t* perm;           // [n_work_tiles][n_sched] per work tile, the schedule positions of each
                                 // range by the gates' events in the tile, most first (full and batched
                                 // runs; null: schedule order)
  int32_t n_work_tiles;          // entries of tile_list, or n_tile
  const int32_t* wt_trace;       // batched runs: trace of each work tile (tile_list: its tile); null otherwise
  const TraceDesc* traces;       // batched runs: [n_trace]
  // cell tables (staged into LDS)
  const CellDev* cells;
  int32_t n_cell;
  const float* leak;             // [n_leak] leakage per gate state (W)
  // [n_leak] the same in leakage quanta: the common quantum (per-cycle sums
  // over gates) << 32 | the cell's own quantum (per-gate sums, leak_shift
  // bits finer; at most 2^30 for the cell's largest state)
  const long long* leak_q;
  int32_t n_leak;
  int32_t leak_lds;              // generic kernel: 1 = leakage tables staged in LDS (n_leak <= LEAK_LDS_MAX)
  int32_t bin_copies;            // lane-interleaved copies of the per-cycle bins (specialised kernel)
  // fixed-point scales (powers of two): term / quantum
  float inv_q_energy;            // coarse energy scale: one term <= 2^20 units, 2^ENERGY_FINE_BITS quanta each
  float inv_q_leak;
  // outputs
  // (partials are indexed by work tile: the tile itself, or its position in tile_list)
  long long* pg;                 // [N_PG][n_work_tiles][n_sched] per-gate sums in quanta (leakage: quanta * ticks)
  uint32_t* pg_tog;              // [n_work_tiles][n_sched]
  uint32_t* pg_glitch;           // [n_work_tiles][n_sched]
  long long* cyc_part;           // [n_work_tiles][n_range][N_BIN][tile_cycles]
};

// Arguments of the energy (table interpolation) kernel.
struct EnergyArgs {
  const int32_t* gate_cell;
  const int64_t* gate_pin_off;
  const int64_t* gate_energy_off;
  int64_t n_gate;
  const float* pin_slew;      // [2P]
  const float* pin_cap;       // [P]
  const CellDev* cells;
  const int32_t* cell_slot_desc_off;  // [n_cell] into descs
  const SlotDesc* descs;
  const ListEntry* list;
  const TableDev* tables;
  int32_t n_table;
  const float* table_data;
  int32_t n_table_data;
  float* energy;
  unsigned int* max_abs_bits;
};

// Launch wrappers, defined with the kernels.
hipError_t launch_pack(const int64_t* t, const uint8_t* v, uint32_t* out, int64_t n, hipStream_t s);
hipError_t launch_tile_index(const int64_t* net_off, const uint32_t* ev, int64_t n_net, int32_t n_tile,
                             int64_t tile_ticks, const int32_t* tile_list, int32_t n_list, int32_t* out,
                             hipStream_t s);
// perm[wt][p] per work tile: within each range of range_gates schedule
// positions, the positions ordered by their gate's pin events in the tile
// (from the tile index), most first, ties by position, padding last, and
// dealt to the merge's waves of `lanes` lanes (a snake over 64-rank groups).
// Full runs: work tile = tile (wt_trace null); batched runs: work tile wt is
// tile wt_tile[wt] of trace wt_trace[wt] (tile index at its tile_idx_base).
constexpr int32_t TILE_ORDER_MAX_RANGE = 4096;
hipError_t launch_tile_order(const int32_t* order, int64_t n_sched, int32_t range_gates, int32_t lanes,
                             const int64_t* gate_pin_off, const int32_t* pin_net, const int32_t* tile_idx,
                             int32_t n_tile, int32_t n_wt, const int32_t* wt_tile, const int32_t* wt_trace,
                             const TraceDesc* traces, int32_t* perm, hipStream_t s);
hipError_t launch_energy(const EnergyArgs& a, hipStream_t s);
hipError_t launch_power(int maxp, int64_t grid, int block, size_t lds, hipStream_t s, const PowerArgs& a,
                        int32_t range_begin, int32_t range_count, unsigned int* wq);
hipError_t launch_cycle_reduce(const long long* part, int32_t n_range, int32_t tile_cycles, int64_t n_cycle,
                               int64_t T, int64_t t_max, double q_energy, double q_leak, double inv_period,
                               double* out, hipStream_t s);
hipError_t launch_cycle_reduce_list(const long long* part, int32_t n_range, int32_t tile_cycles,
                                    const int32_t* tile_list, int32_t n_list, int64_t n_cycle, int64_t T,
                                    int64_t t_max, double q_energy, double q_leak, double inv_period,
                                    double* out, hipStream_t s);
hipError_t launch_cycle_reduce_batch(const long long* part, int32_t n_range, int32_t tile_cycles,
                                     const int32_t* tile_list, const int32_t* wt_trace, int32_t n_wt,
                                     const TraceDesc* traces, const int64_t* out_base, int64_t T, double q_energy,
                                     double q_leak, double inv_period, double* out, hipStream_t s);
hipError_t launch_gate_accumulate(const long long* pg, const uint32_t* tog, const uint32_t* glc,
                                  const int32_t* order, int64_t n_sched, int32_t n_wt, const int32_t* wt_begin,
                                  const int32_t* round_trace, int32_t n_round_trace, int64_t n_gate, long long* acc,
                                  unsigned long long* acc_cnt, hipStream_t s);
hipError_t launch_gate_finalize(const long long* acc, int64_t n_gate, double norm_energy, double norm_leak,
                                const int32_t* gate_cell, const CellDev* cells, double* gate_power, hipStream_t s);
hipError_t launch_zero_u64(unsigned long long* p, int64_t n, hipStream_t s);
hipError_t launch_zero_u32(unsigned int* p, int64_t n, hipStream_t s);
hipError_t launch_gate_reduce(const long long* pg, const uint32_t* tog, const uint32_t* glc,
                              const int32_t* order, int64_t n_sched, int32_t n_tile, int64_t n_gate,
                              double norm_energy, double norm_leak, const int32_t* gate_cell,
                              const CellDev* cells, double* gate_power, unsigned long long* gate_tog,
                              unsigned long long* gate_glitch, hipStream_t s);
constexpr int TOTAL_BLOCKS = 256;
hipError_t launch_chunk_net(const int64_t* chunk_off, int64_t n_net, int32_t* chunk_net, hipStream_t s);
hipError_t launch_glitch_annotate(const int64_t* net_off, uint32_t* ev, const int64_t* chunk_off,
                                  const int32_t* chunk_net, int64_t n_chunk, uint32_t T, hipStream_t s);
hipError_t launch_glitch_annotate_tiles(const int64_t* net_off, uint32_t* ev, const int32_t* idx_begin,
                                        const int32_t* tile_list, int32_t n_list, int64_t n_net, int64_t tile_ticks,
                                        uint32_t T, hipStream_t s);
hipError_t launch_total(const double* gate_power, const int32_t* gate_class, int64_t n_gate, double* part,
                        double* out, hipStream_t s);
// kc >= 0: the cell's only clock pin (np 3, kc 0), whose edges the workgroup
// takes once per work item for the gates on the item's clock net.
hipError_t launch_power_cell(int np, int kc, int64_t grid, int block, size_t lds, hipStream_t s, const PowerArgs& a,
                             int32_t range_begin, int32_t range_count, int32_t cell, unsigned int* wq);
// Longest tile (cycles) whose clock edges the workgroup takes.
constexpr int CLOCK_EDGE_TILE_MAX = 512;

// The multiply-shift pair of PowerArgs.cyc_m / cyc_sh for T ticks per cycle.
inline void cycle_magic(int64_t T, uint32_t* m, int32_t* sh) {
  int l = 0;
  while ((int64_t(1) << l) < T) ++l;
  const int s = 29 + l;
  *sh = s;
  *m = (uint32_t)(((unsigned __int128)1 << s) / (unsigned __int128)T +
                  ((((unsigned __int128)1 << s) % (unsigned __int128)T) ? 1 : 0));
}

// Read-only view of a context's device event store, for the passes that run
// over it besides the power kernels (toggle features, correlation).
struct TraceDev {
  int device;
  hipStream_t stream;
  hipStream_t aux_stream;        // the context's stream for feature work (nullptr until first used)
  int64_t n_net, n_cycle, T, t_max;
  const int64_t* net_off;        // device [N+1]
  const uint32_t* ev;            // device [E]
  const int64_t* chunk_off;      // device [N+1] 256-event blocks per net (prefix)
  const int32_t* chunk_net;      // device [n_chunk] net of each block
  int64_t n_chunk;
  const int64_t* host_net_off;   // host [N+1]
};

}  // namespace gtp

struct gtp_ctx;
namespace gtp {
// Fills *out from a context with a loaded trace; 0 on success, else a
// negative status (the context's last error says why).
int32_t trace_dev(gtp_ctx* ctx, TraceDev* out);
// Device buffers of the context's weighted toggle sums, kept between calls.
struct AuxBuffers {
  DArr<int32_t> sel, wq, idx;
  DArr<unsigned long long> out;
  // what sel and wq hold (uploaded to sel_at), so that a repeated selection
  // is not uploaded again
  std::vector<int32_t> sel_h, wq_h;
  const int32_t* sel_at = nullptr;
};
AuxBuffers* aux_buffers(gtp_ctx* ctx);
// The context's auxiliary stream, created on first use (nullptr on failure).
hipStream_t aux_stream(gtp_ctx* ctx);
// A parsed VCD's per-(id, bit) event store on its device (gtp_vcd.hip), for
// gtp_set_trace_vcd: words as above, padded by EV_PAD words.
struct VcdView {
  int device;
  const uint32_t* ev;       // [n_event + EV_PAD]
  const int64_t* d_off;     // [n_stream + 1] on the device
  const int64_t* h_off;     // the same on the host
  int64_t n_stream, n_event, t_max;
  double time_scale;
};
}  // namespace gtp
struct gtp_vcd;
namespace gtp {
bool vcd_view(const gtp_vcd* v, VcdView* out);

}  // namespace gtp
