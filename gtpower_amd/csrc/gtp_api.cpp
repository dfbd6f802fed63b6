// Host runtime of the power engine behind the C ABI of include/gtpower.h.
//
// Replaces CudaPower's host orchestration (power/cuda/CudaPower.cu:1227-1337):
// flattening of gates, pins, tables and delays (initGateData :1601-1816), the
// LUT construction (:2347-2624), the event store (:1874-2130), the schedule
// (:2718-2849), the launches (:2851-2957) and the result merge (:2959-3132).
#include <emmintrin.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <string>
#include <thread>
#include <sched.h>
#include <vector>

#include "gtp_internal.hpp"
#include "gtpower.h"

#define GTP_API extern "C" __attribute__((visibility("default")))

using namespace gtp;

namespace {

thread_local std::string g_create_error;

// Host worker threads when none are asked for: OMP_NUM_THREADS when set (a
// job's CPU share), else the processors this process may run on.
int default_threads() {
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) return v;
  }
  cpu_set_t cs;
  CPU_ZERO(&cs);
  if (sched_getaffinity(0, sizeof(cs), &cs) == 0 && CPU_COUNT(&cs) > 0) return CPU_COUNT(&cs);
  return (int)std::max(1u, std::thread::hardware_concurrency());
}

// Event words (t - base) << EV_SHIFT | (v & 3) of n events, written with
// stores that bypass the cache (the destination lines are not read first:
// 13 instead of 17 bytes of host memory traffic per event), four at a time;
// the fence makes them visible before the caller hands the words on.
void pack_words_nt(uint32_t* d, const int64_t* tm, const uint8_t* val, int64_t n, int64_t base) {
  int64_t i = 0;
  for (; i < n && ((uintptr_t)(d + i) & 15u) != 0; ++i)
    d[i] = ((uint32_t)(tm[i] - base) << EV_SHIFT) | (uint32_t)(val[i] & 3u);
  const __m128i vb = _mm_set1_epi64x(base), m3 = _mm_set1_epi32(3), z = _mm_setzero_si128();
  for (; i + 4 <= n; i += 4) {
    const __m128i a = _mm_sub_epi64(_mm_loadu_si128(reinterpret_cast<const __m128i*>(tm + i)), vb);
    const __m128i b = _mm_sub_epi64(_mm_loadu_si128(reinterpret_cast<const __m128i*>(tm + i + 2)), vb);
    // the low 32 bits of the four differences
    const __m128i lo = _mm_unpacklo_epi64(_mm_shuffle_epi32(a, _MM_SHUFFLE(2, 0, 2, 0)),
                                          _mm_shuffle_epi32(b, _MM_SHUFFLE(2, 0, 2, 0)));
    int32_t v4;
    std::memcpy(&v4, val + i, 4);
    const __m128i vv = _mm_and_si128(_mm_unpacklo_epi16(_mm_unpacklo_epi8(_mm_cvtsi32_si128(v4), z), z), m3);
    _mm_stream_si128(reinterpret_cast<__m128i*>(d + i), _mm_or_si128(_mm_slli_epi32(lo, EV_SHIFT), vv));
  }
  for (; i < n; ++i) d[i] = ((uint32_t)(tm[i] - base) << EV_SHIFT) | (uint32_t)(val[i] & 3u);
  _mm_sfence();
}

// f(i) for i in [0, n) on up to `threads` host threads, each given at least
// `grain` items (1024 for light items; a few for items that are whole nets)
// pack_words_nt that also checks the events: times strictly increasing from
// t_prev (the time of the net's event before tm[0], -1 when none) and at most
// t_max, values at most 3.  Returns false when one is not.
bool pack_words_checked(uint32_t* d, const int64_t* tm, const uint8_t* val, int64_t n, int64_t base, int64_t t_prev,
                        int64_t t_max) {
  if (n <= 0) return true;
  bool ok = tm[0] > t_prev;
  int64_t i = 0;
  for (; i < n && ((uintptr_t)(d + i) & 15u) != 0; ++i) {
    ok &= (i == 0 || tm[i] > tm[i - 1]) && tm[i] <= t_max && val[i] <= 3;
    d[i] = ((uint32_t)(tm[i] - base) << EV_SHIFT) | (uint32_t)(val[i] & 3u);
  }
  // sign bits: t - t_prev - 1 < 0 (not increasing) or t_max - t < 0
  __m128i bad = _mm_setzero_si128();
  uint32_t vor = 0;
  const __m128i vb = _mm_set1_epi64x(base), one = _mm_set1_epi64x(1), tmx = _mm_set1_epi64x(t_max);
  const __m128i m3 = _mm_set1_epi32(3), z = _mm_setzero_si128();
  if (i == 0 && i + 4 <= n) {  // the first quad has no word before it in tm: checked on its own
    ok &= tm[1] > tm[0] && tm[2] > tm[1] && tm[3] > tm[2];
  }
  for (; i + 4 <= n; i += 4) {
    const __m128i ta = _mm_loadu_si128(reinterpret_cast<const __m128i*>(tm + i));
    const __m128i tb = _mm_loadu_si128(reinterpret_cast<const __m128i*>(tm + i + 2));
    if (i > 0) {
      const __m128i pa = _mm_loadu_si128(reinterpret_cast<const __m128i*>(tm + i - 1));
      const __m128i pb = _mm_loadu_si128(reinterpret_cast<const __m128i*>(tm + i + 1));
      bad = _mm_or_si128(bad, _mm_sub_epi64(_mm_sub_epi64(ta, pa), one));
      bad = _mm_or_si128(bad, _mm_sub_epi64(_mm_sub_epi64(tb, pb), one));
    }
    bad = _mm_or_si128(bad, _mm_sub_epi64(tmx, ta));
    bad = _mm_or_si128(bad, _mm_sub_epi64(tmx, tb));
    const __m128i a = _mm_sub_epi64(ta, vb), b = _mm_sub_epi64(tb, vb);
    const __m128i lo = _mm_unpacklo_epi64(_mm_shuffle_epi32(a, _MM_SHUFFLE(2, 0, 2, 0)),
                                          _mm_shuffle_epi32(b, _MM_SHUFFLE(2, 0, 2, 0)));
    uint32_t v4;
    std::memcpy(&v4, val + i, 4);
    vor |= v4;
    const __m128i vv = _mm_and_si128(_mm_unpacklo_epi16(_mm_unpacklo_epi8(_mm_cvtsi32_si128((int)v4), z), z), m3);
    _mm_stream_si128(reinterpret_cast<__m128i*>(d + i), _mm_or_si128(_mm_slli_epi32(lo, EV_SHIFT), vv));
  }
  for (; i < n; ++i) {
    ok &= (i == 0 || tm[i] > tm[i - 1]) && tm[i] <= t_max && val[i] <= 3;
    d[i] = ((uint32_t)(tm[i] - base) << EV_SHIFT) | (uint32_t)(val[i] & 3u);
  }
  _mm_sfence();
  // the sign bit of each 64-bit lane: bytes 7 and 15
  ok &= (_mm_movemask_epi8(bad) & 0x8080) == 0;
  ok &= (vor & 0xFCFCFCFCu) == 0;
  return ok;
}

template <class F>
void parallel_for(int64_t n, int threads, F f, int64_t grain = 1024) {
  if (n <= 0) return;
  if (threads <= 0) threads = default_threads();
  threads = (int)std::min<int64_t>(threads, std::max<int64_t>(1, n / std::max<int64_t>(grain, 1)));
  if (threads <= 1) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int64_t> next{0};
  const int64_t chunk = std::max<int64_t>(1, n / (threads * 16));
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&] {
      for (;;) {
        int64_t b = next.fetch_add(chunk);
        if (b >= n) break;
        int64_t e = std::min(n, b + chunk);
        for (int64_t i = b; i < e; ++i) f(i);
      }
    });
  for (auto& th : pool) th.join();
}

// Joins a helper thread when the scope ends, on every return path.
struct JoinGuard {
  std::thread& t;
  ~JoinGuard() {
    if (t.joinable()) t.join();
  }
};

struct HostCell {
  int n_pin = 0, n_in = 0, cls = 0, sbits = 0;
  uint32_t clock_mask = 0;
  std::vector<float> leak;   // 2^n_pin
  std::vector<std::vector<int>> groups;  // internal power groups per pin
};

}  // namespace

struct gtp_ctx {
  gtp_options opt{};
  int device = 0;
  hipStream_t stream = nullptr;
  static constexpr int N_SIDE = 3;
  hipStream_t side[N_SIDE] = {nullptr, nullptr, nullptr};  // concurrent per-cell launches
  // feature work (weighted toggle sums) beside the power runs; created on
  // first use, after the run streams, so that it does not take one of the
  // hardware queues the run streams would otherwise get to themselves
  hipStream_t aux = nullptr;
  gtp::AuxBuffers aux_buf;
  hipEvent_t fork = nullptr, join[N_SIDE] = {nullptr, nullptr, nullptr};
  std::string err;
  int threads = 0;

  // library
  std::vector<HostCell> cells;
  std::vector<CellDev> cell_dev;
  std::vector<float> leak_f;     // concatenated per-cell leakage tables
  double leak_max = 0;            // largest |leakage| of the cells the design uses
  int leak_k = 0;                 // common leakage quantum 2^-leak_k W: leak_max <= 2^30 quanta
  DArr<CellDev> d_cells;
  DArr<float> d_leak;
  DArr<long long> d_leak_q;

  // design
  bool have_design = false;
  int64_t G = 0, P = 0, N = 0;
  std::vector<int32_t> gate_cell;
  std::vector<int64_t> gate_pin_off, gate_arc_off;
  std::vector<int32_t> pin_net;
  std::vector<double> arc_delay;
  double csw_max = 0;
  float energy_max = 0;
  // the glitch width (s) beyond which every pin's glitch ratio is 1: the
  // largest finite (s_r + s_f) / 2 (power/PowerUtils.cc:285-294)
  double glitch_sat_s = 0;
  DArr<int32_t> d_gate_cell, d_gate_class;
  DArr<int64_t> d_gate_pin_off, d_gate_arc_off, d_gate_energy_off;
  DArr<int32_t> d_pin_net;
  DArr<float> d_pin_gk, d_pin_csw;
  DArr<int8_t> d_pin_default;
  DArr<float> d_energy;
  DArr<int32_t> d_arc_q;
  double arc_q_scale = 0;  // time scale the arc_q table was built for

  // trace
  bool have_trace = false;
  int64_t E = 0, T = 0, C = 0, t_max = 0, P_ev = 0;
  double scale = 0, period = 0;
  std::vector<int64_t> gate_work;
  int64_t gate_work_max = 0;
  std::vector<int32_t> order_tmp[2];  // build_order's scratch, kept between rebuilds
  Pinned<int32_t> order_pinned;       // staging of gate order uploads
  DArr<int64_t> d_net_off;
  DArr<uint32_t> d_ev;
  DArr<int64_t> d_chunk_off;     // annotation blocks per net (prefix)
  DArr<int32_t> d_chunk_net;     // net of each annotation block
  int64_t n_chunk = 0;
  DArr<int32_t> d_tile_idx;
  int32_t tile_idx_tc = 0;
  // per-tile lane order of the full run's schedule (built with the tile index)
  DArr<int32_t> d_perm;
  int64_t perm_tc = 0;
  int32_t perm_rg = 0, perm_lanes = 0;
  std::vector<int64_t> h_net_off;  // host copy of the event offsets
  bool glitch_valid = false;       // the EV_GLITCH bits hold the annotation of this trace
  DArr<int32_t> d_tile_list;
  DArr<int32_t> d_tile_idx_sub;    // tile index of a subset run's tiles

  // schedule: ranges of range_gates gates; a range holds gates of one
  // specialised cell type, or generic gates
  struct Launch {
    int cell;          // cell type for the specialised kernel, -1 = generic
    int np;            // pins (specialised) or MAXP of the generic kernel
    int32_t range_begin, range_count;
  };
  struct OrderSet {
    std::vector<int32_t> order;
    std::vector<Launch> launches;
    int32_t range_gates = 0, n_range = 0;
    int64_t order_T = 0;
    DArr<int32_t> d_order;
    // per-work-tile partials of this schedule's runs
    DArr<long long> d_pg;
    DArr<uint32_t> d_pg_tog, d_pg_glc;
    DArr<long long> d_cyc_part;
  };
  // two cached schedules: the full run's and the tile-subset run's (whose
  // ranges are sized for its few tiles)
  OrderSet orders[3];              // [2]: batched runs
  int n_cu = 256;                  // compute units of the device
  // autotuner's choice for the loaded trace (0 = none)
  int32_t tuned_tc = 0, tuned_rg = 0;
  double tune_ms = 0;
  DArr<unsigned int> d_wq;         // persistent work-queue counters, 8 per launch
  gtp_schedule_info sched{};

  // scratch and results
  DArr<double> d_cycle, d_gate, d_total, d_total_part;
  DArr<unsigned long long> d_gtog, d_gglc;
  DArr<unsigned int> d_maxbits;

  // page-locked staging of packed event words for loads from host memory
  // (large traces, sliced runs, streamed batches): a ring of buffers, each
  // refilled once its copy to the device has completed, so that packing the
  // next part overlaps the copy of the last one.  The copies go on the
  // auxiliary stream (aux), so that they do not wait behind the kernels of
  // the interval being evaluated on the run stream; no stream is added for
  // them (a stream of their own took one of the process's few hardware queues,
  // and the surrogate's overlapped evaluation, whose base term uses aux, then
  // slowed by ~10%)
  static constexpr int N_STAGE = 4;
  static constexpr int64_t STAGE_WORDS = (int64_t)1 << 25;  // 128 MiB a buffer
  Pinned<uint32_t> stage[N_STAGE];
  hipEvent_t stage_ev[N_STAGE] = {nullptr, nullptr, nullptr, nullptr};
  bool stage_pending[N_STAGE] = {false, false, false, false};
  int stage_next = 0;
  // A run's results come back through this pinned buffer (DMA at PCIe rate;
  // copies into the caller's pageable arrays ran at ~6.5 GB/s), then a
  // threaded host copy moves them into the caller's arrays.
  Pinned<unsigned char> dl;

  gtp_status fail(gtp_status st, const std::string& m) {
    err = m;
    return st;
  }
};

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return (ctx)->fail(e_ == hipErrorOutOfMemory ? GTP_ERR_OOM : GTP_ERR_DEVICE,          \
                         std::string(#expr) + ": " + hipGetErrorString(e_));                \
  } while (0)

namespace {

uint64_t full_truth(int n) { return n >= 6 ? ~0ull : ((1ull << (1 << n)) - 1ull); }

// Resolve the cell library into per-cell slot descriptors.  The choice of
// internal_power group depends on (cell, pin, related pin, rf, state) only, so
// it is made here once; the energy kernel then interpolates per gate.
gtp_status build_library(gtp_ctx* x, const gtp_library* L, std::vector<SlotDesc>& descs,
                         std::vector<int32_t>& desc_off, std::vector<ListEntry>& list,
                         std::vector<TableDev>& tables, std::vector<float>& tdata) {
  if (!L || L->n_cell <= 0 || !L->cell_n_pin || !L->cell_n_in || !L->cell_class ||
      !L->cell_clock_mask || !L->cell_leak_offset || !L->cell_leak_default)
    return x->fail(GTP_ERR_ARG, "gtp_library: missing arrays");
  if (L->n_table < 0 || (L->n_table > 0 && (!L->tables || !L->table_data)))
    return x->fail(GTP_ERR_ARG, "gtp_library: tables");
  // tables -> float data
  size_t data_len = 0;
  for (int t = 0; t < L->n_table; ++t) {
    const gtp_table& tb = L->tables[t];
    if (tb.n1 <= 0 || tb.n2 < 0 || tb.axis1 < 0 || tb.values < 0 || (tb.n2 > 0 && tb.axis2 < 0))
      return x->fail(GTP_ERR_ARG, "gtp_library: malformed table " + std::to_string(t));
    data_len = std::max(data_len, (size_t)tb.values + (size_t)tb.n1 * (tb.n2 ? tb.n2 : 1));
    data_len = std::max(data_len, (size_t)tb.axis1 + tb.n1);
    if (tb.n2) data_len = std::max(data_len, (size_t)tb.axis2 + tb.n2);
  }
  if (data_len > (size_t)INT32_MAX) return x->fail(GTP_ERR_LIMIT, "gtp_library: table data beyond 2^31 values");
  tdata.resize(std::max<size_t>(data_len, 1));
  for (size_t i = 0; i < data_len; ++i) tdata[i] = (float)L->table_data[i];
  tables.resize(L->n_table);
  for (int t = 0; t < L->n_table; ++t) {
    const gtp_table& tb = L->tables[t];
    tables[t] = TableDev{tb.n1, tb.n2, tb.var1, tb.var2, (int32_t)tb.axis1, (int32_t)tb.axis2, (int32_t)tb.values};
  }
  x->cells.assign(L->n_cell, HostCell{});
  for (int c = 0; c < L->n_cell; ++c) {
    HostCell& hc = x->cells[c];
    hc.n_pin = L->cell_n_pin[c];
    hc.n_in = L->cell_n_in[c];
    hc.cls = L->cell_class[c];
    hc.clock_mask = L->cell_clock_mask[c];
    if (hc.n_pin <= 0 || hc.n_pin > GTP_MAX_CELL_PINS || hc.n_in < 0 || hc.n_in > hc.n_pin)
      return x->fail(GTP_ERR_LIMIT, "gtp_library: cell " + std::to_string(c) + " pin count outside 1.." +
                                        std::to_string(GTP_MAX_CELL_PINS));
    if (hc.cls < 0 || hc.cls >= GTP_N_CLASS) return x->fail(GTP_ERR_ARG, "gtp_library: cell class");
    const int64_t o0 = L->cell_leak_offset[c], o1 = L->cell_leak_offset[c + 1];
    const int ns = 1 << hc.n_pin;
    hc.leak.assign(ns, (float)L->cell_leak_default[c]);
    if (o1 != o0) {
      if (o1 - o0 != ns || !L->leak_values) return x->fail(GTP_ERR_ARG, "gtp_library: leakage table size");
      for (int s = 0; s < ns; ++s) hc.leak[s] = (float)L->leak_values[o0 + s];
    }
    hc.groups.assign(hc.n_pin, {});
  }
  for (int k = 0; k < L->n_ipower; ++k) {
    const gtp_internal_power& ip = L->ipowers[k];
    if (ip.cell < 0 || ip.cell >= L->n_cell) return x->fail(GTP_ERR_ARG, "gtp_library: ipower cell");
    HostCell& hc = x->cells[ip.cell];
    if (ip.pin < 0 || ip.pin >= hc.n_pin || ip.related_pin >= hc.n_in || ip.rise_table >= L->n_table ||
        ip.fall_table >= L->n_table)
      return x->fail(GTP_ERR_ARG, "gtp_library: ipower " + std::to_string(k));
    hc.groups[ip.pin].push_back(k);
    const uint64_t full = full_truth(hc.n_pin);
    if ((ip.when_truth & full) != full) {
      if (hc.n_pin > MAXP_LARGE)
        return x->fail(GTP_ERR_UNSUPPORTED, "gtp_library: cell " + std::to_string(ip.cell) +
                                                ": a `when` on internal power of a cell of more than 6 pins");
      hc.sbits = hc.n_pin;
    }
  }
  // slots
  x->cell_dev.assign(L->n_cell, CellDev{});
  x->leak_f.clear();
  desc_off.assign(L->n_cell, 0);
  descs.clear();
  list.clear();
  for (int c = 0; c < L->n_cell; ++c) {
    HostCell& hc = x->cells[c];
    CellDev& cd = x->cell_dev[c];
    cd.n_pin = hc.n_pin;
    cd.n_in = hc.n_in;
    cd.clock_mask = hc.clock_mask;
    cd.state_bits = hc.sbits;
    cd.leak_off = (int32_t)x->leak_f.size();
    for (float v : hc.leak) x->leak_f.push_back(v);
    desc_off[c] = (int32_t)descs.size();
    const int sb = hc.sbits, nst = 1 << sb;
    int off = 0;
    for (int p = 0; p < hc.n_pin; ++p) {
      cd.slot_off[p] = off;
      if (p < hc.n_in) {
        // findInputInternalVal (power/TimeBasedPower.cc:341-381)
        for (int rf = 0; rf < 2; ++rf)
          for (int st = 0; st < nst; ++st) {
            SlotDesc d{-1, (int16_t)p, (int8_t)rf, (int8_t)-1, 0, 0};
            for (int k : hc.groups[p]) {
              const gtp_internal_power& ip = L->ipowers[k];
              // without conditional groups (sb == 0) every group holds in every state
              if ((ip.when_truth >> st) & 1ull) {
                d.table = rf == 0 ? ip.rise_table : ip.fall_table;
                break;
              }
            }
            descs.push_back(d);
          }
        off += 2 * nst;
      } else {
        // findOutputInternalVal (:383-509): related input q, its edge, this edge, state
        for (int q = 0; q < hc.n_in; ++q)
          for (int frf = 0; frf < 2; ++frf)
            for (int trf = 0; trf < 2; ++trf)
              for (int st = 0; st < nst; ++st) {
                SlotDesc d{-1, (int16_t)q, (int8_t)frf, (int8_t)p, 0, 0};
                for (int k : hc.groups[p]) {
                  const gtp_internal_power& ip = L->ipowers[k];
                  if (ip.related_pin != q) continue;
                  if (!((ip.when_truth >> st) & 1ull)) continue;
                  d.table = trf == 0 ? ip.rise_table : ip.fall_table;
                  break;
                }
                descs.push_back(d);
              }
        off += hc.n_in * 4 * nst;
        // getDefaultOutputPinInternalEnergyVal (:511-555): mean over related
        // inputs x {rise, fall} x their groups, `when` ignored
        for (int trf = 0; trf < 2; ++trf) {
          SlotDesc d{-1, 0, 0, (int8_t)p, (int32_t)list.size(), 0};
          for (int i = 0; i < hc.n_in; ++i)
            for (int frf = 0; frf < 2; ++frf)
              for (int k : hc.groups[p]) {
                const gtp_internal_power& ip = L->ipowers[k];
                if (ip.related_pin != i) continue;
                list.push_back(ListEntry{trf == 0 ? ip.rise_table : ip.fall_table, (int16_t)i, (int8_t)frf, (int8_t)p});
                ++d.list_cnt;
              }
          // a default slot with no entries is 0: keep it plain with table -1
          if (d.list_cnt == 0) d.list_off = 0;
          descs.push_back(d);
        }
        off += 2;
      }
    }
    cd.n_slot = off;
  }
  return GTP_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
GTP_API void gtp_options_init(gtp_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->device = -1;
  o->autotune = 0;
}

GTP_API const char* gtp_create_error(void) { return g_create_error.c_str(); }

GTP_API gtp_status gtp_create(const gtp_options* opt, gtp_ctx** out) {
  if (!out) {
    g_create_error = "gtp_create: null output pointer";
    return GTP_ERR_ARG;
  }
  *out = nullptr;
  auto* x = new (std::nothrow) gtp_ctx();
  if (!x) {
    g_create_error = "gtp_create: out of host memory";
    return GTP_ERR_OOM;
  }
  if (opt) x->opt = *opt; else gtp_options_init(&x->opt);
  int n_dev = 0;
  hipError_t e = hipGetDeviceCount(&n_dev);
  if (e != hipSuccess || n_dev <= 0) {
    g_create_error = std::string("gtp_create: no HIP device: ") + hipGetErrorString(e);
    delete x;
    return GTP_ERR_DEVICE;
  }
  int dev = x->opt.device;
  if (dev < 0) {
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  }
  if (dev >= n_dev) {
    g_create_error = "gtp_create: device ordinal out of range";
    delete x;
    return GTP_ERR_ARG;
  }
  x->device = dev;
  if ((e = hipSetDevice(dev)) != hipSuccess || (e = hipStreamCreate(&x->stream)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&x->fork, hipEventDisableTiming)) != hipSuccess) {
    g_create_error = std::string("gtp_create: ") + hipGetErrorString(e);
    delete x;
    return GTP_ERR_DEVICE;
  }
  for (int k = 0; k < gtp_ctx::N_SIDE; ++k) {
    if ((e = hipStreamCreate(&x->side[k])) != hipSuccess ||
        (e = hipEventCreateWithFlags(&x->join[k], hipEventDisableTiming)) != hipSuccess) {
      g_create_error = std::string("gtp_create: ") + hipGetErrorString(e);
      gtp_destroy(x);
      return GTP_ERR_DEVICE;
    }
  }
  x->threads = x->opt.host_threads > 0 ? x->opt.host_threads
                                       : default_threads();
  {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
      x->n_cu = ncu;
  }
  *out = x;
  return GTP_OK;
}

GTP_API void gtp_destroy(gtp_ctx* x) {
  if (!x) return;
  (void)hipSetDevice(x->device);
  if (x->stream) {
    (void)hipStreamSynchronize(x->stream);
    (void)hipStreamDestroy(x->stream);
  }
  if (x->aux) {
    (void)hipStreamSynchronize(x->aux);
    (void)hipStreamDestroy(x->aux);
  }
  for (int k = 0; k < gtp_ctx::N_SIDE; ++k) {
    if (x->side[k]) {
      (void)hipStreamSynchronize(x->side[k]);
      (void)hipStreamDestroy(x->side[k]);
    }
    if (x->join[k]) (void)hipEventDestroy(x->join[k]);
  }
  if (x->fork) (void)hipEventDestroy(x->fork);
  for (int k = 0; k < gtp_ctx::N_STAGE; ++k)
    if (x->stage_ev[k]) (void)hipEventDestroy(x->stage_ev[k]);
  delete x;
}

GTP_API const char* gtp_last_error(const gtp_ctx* x) { return x ? x->err.c_str() : "null context"; }

// ---------------------------------------------------------------------------
GTP_API gtp_status gtp_set_design(gtp_ctx* x, const gtp_library* L, const gtp_design* d) {
  if (!x) return GTP_ERR_ARG;
  HIPCHK(x, hipSetDevice(x->device));
  x->have_design = false;
  x->have_trace = false;
  std::vector<SlotDesc> descs;
  std::vector<int32_t> desc_off;
  std::vector<ListEntry> list;
  std::vector<TableDev> tables;
  std::vector<float> tdata;
  gtp_status st = build_library(x, L, descs, desc_off, list, tables, tdata);
  if (st != GTP_OK) return st;
  if (!d || d->n_gate <= 0 || d->n_pin <= 0 || d->n_net < 0 || !d->gate_cell || !d->gate_pin_offset ||
      !d->gate_arc_offset || !d->pin_net || !d->pin_slew || !d->pin_load_cap || !d->pin_voltage ||
      !d->pin_default_state)
    return x->fail(GTP_ERR_ARG, "gtp_design: missing arrays");
  if (d->n_gate >= INT32_MAX || d->n_pin >= INT32_MAX || d->n_net >= INT32_MAX)
    return x->fail(GTP_ERR_LIMIT, "gtp_design: more than 2^31 gates, pins or nets");
  const int64_t G = d->n_gate, P = d->n_pin, N = d->n_net;
  if (d->gate_pin_offset[0] != 0 || d->gate_arc_offset[0] != 0)
    return x->fail(GTP_ERR_ARG, "gtp_design: offsets must start at 0");
  std::vector<int64_t> eoff(G);
  int64_t eacc = 0;
  for (int64_t g = 0; g < G; ++g) {
    const int c = d->gate_cell[g];
    if (c < 0 || c >= L->n_cell) return x->fail(GTP_ERR_ARG, "gtp_design: gate_cell out of range at " + std::to_string(g));
    const HostCell& hc = x->cells[c];
    if (d->gate_pin_offset[g + 1] - d->gate_pin_offset[g] != hc.n_pin)
      return x->fail(GTP_ERR_ARG, "gtp_design: pin count of gate " + std::to_string(g) + " differs from its cell");
    if (d->gate_arc_offset[g + 1] - d->gate_arc_offset[g] != (int64_t)(hc.n_pin - hc.n_in) * hc.n_in * 4)
      return x->fail(GTP_ERR_ARG, "gtp_design: arc count of gate " + std::to_string(g));
    eoff[g] = eacc;
    eacc += x->cell_dev[c].n_slot;
  }
  if (d->gate_pin_offset[G] != P) return x->fail(GTP_ERR_ARG, "gtp_design: gate_pin_offset[n_gate] != n_pin");
  const int64_t A = d->gate_arc_offset[G];
  if (A > 0 && !d->arc_delay) return x->fail(GTP_ERR_ARG, "gtp_design: arc_delay missing");
  std::vector<float> gk(P), csw(P, 0.f), slew(2 * P), cap(P);
  double glitch_sat = 0;
  std::vector<int32_t> gclass(G);
  double csw_max = 0;
  for (int64_t g = 0; g < G; ++g) {
    const HostCell& hc = x->cells[d->gate_cell[g]];
    gclass[g] = hc.cls;
    for (int p = 0; p < hc.n_pin; ++p) {
      const int64_t q = d->gate_pin_offset[g] + p;
      if (p >= hc.n_in) {
        const double v = d->pin_load_cap[q] * d->pin_voltage[q] * d->pin_voltage[q];
        csw[q] = (float)v;
        csw_max = std::max(csw_max, std::fabs(v));
      }
    }
  }
  for (int64_t q = 0; q < P; ++q) {
    if (d->pin_net[q] < -1 || d->pin_net[q] >= N) return x->fail(GTP_ERR_ARG, "gtp_design: pin_net out of range at " + std::to_string(q));
    if (d->pin_default_state[q] < 0 || d->pin_default_state[q] > 3) return x->fail(GTP_ERR_ARG, "gtp_design: pin_default_state");
    const double sr = d->pin_slew[2 * q], sf = d->pin_slew[2 * q + 1];
    if (std::isnan(sr) || std::isnan(sf)) return x->fail(GTP_ERR_ARG, "gtp_design: NaN slew at pin " + std::to_string(q));
    const double sum = sr + sf;
    // glitch scaling ratio min(1, (2 w / (s_r + s_f))^2) (power/PowerUtils.cc:285-294);
    // a zero slew sum gives ratio 1, an infinite one gives 0 (no glitch).
    gk[q] = sum == 0 ? INFINITY : (float)(2.0 / sum);
    if (std::isfinite(sum)) glitch_sat = std::max(glitch_sat, 0.5 * sum);
    slew[2 * q] = (float)sr;
    slew[2 * q + 1] = (float)sf;
    cap[q] = (float)d->pin_load_cap[q];
  }
  x->G = G;
  x->P = P;
  x->N = N;
  x->gate_cell.assign(d->gate_cell, d->gate_cell + G);
  x->gate_pin_off.assign(d->gate_pin_offset, d->gate_pin_offset + G + 1);
  x->gate_arc_off.assign(d->gate_arc_offset, d->gate_arc_offset + G + 1);
  x->pin_net.assign(d->pin_net, d->pin_net + P);
  x->arc_delay.assign(d->arc_delay, d->arc_delay + A);
  x->csw_max = csw_max;
  x->glitch_sat_s = glitch_sat;
  x->arc_q_scale = 0;
  // inputs with an arc to the output in some gate of each single-output cell
  // (the related-input search skips the others for the whole cell type)
  for (CellDev& cd : x->cell_dev) cd.arc_mask = cd.n_pin - cd.n_in == 1 ? 0u : ~0u;
  for (int64_t g = 0; g < G; ++g) {
    CellDev& cd = x->cell_dev[d->gate_cell[g]];
    if (cd.n_pin - cd.n_in != 1) continue;
    const double* ad = d->arc_delay + d->gate_arc_offset[g];
    for (int i = 0; i < cd.n_in; ++i)
      for (int k = 0; k < 4; ++k)
        if (!std::isnan(ad[i * 4 + k])) cd.arc_mask |= 1u << i;
  }
  // leakage quanta: a common quantum for the per-cycle sums over gates, from
  // the largest state of the cells in use (2^30 quanta at most), and per cell
  // a finer one for its per-gate sums (2^30 quanta for the cell's largest
  // state, at most 31 bits finer), so that a cell of small leakage keeps its
  // own relative precision beside cells of much larger leakage
  std::vector<double> cmax(x->cells.size(), 0.0);
  std::vector<char> used(x->cells.size(), 0);
  for (int64_t g = 0; g < G; ++g) used[d->gate_cell[g]] = 1;
  x->leak_max = 0;
  for (size_t c = 0; c < x->cells.size(); ++c) {
    for (float v : x->cells[c].leak) cmax[c] = std::max(cmax[c], (double)std::fabs(v));
    if (used[c]) x->leak_max = std::max(x->leak_max, cmax[c]);
  }
  auto quantum_exp = [](double m) { return (int)std::floor(30.0 - std::log2(m * 1.000001)); };
  x->leak_k = x->leak_max > 0 ? quantum_exp(x->leak_max) : 0;
  std::vector<long long> lq(x->leak_f.size());
  for (size_t c = 0; c < x->cells.size(); ++c) {
    CellDev& cd = x->cell_dev[c];
    const int kc = cmax[c] > 0 ? std::min(std::max(quantum_exp(cmax[c]), x->leak_k), x->leak_k + 31) : x->leak_k;
    if (kc > 120 || x->leak_k < -120)
      return x->fail(GTP_ERR_UNSUPPORTED, "gtp_design: leakage magnitudes outside the FP32 range");
    cd.leak_shift = kc - x->leak_k;
    const int ns = 1 << cd.n_pin;
    for (int st = 0; st < ns; ++st) {
      const double v = (double)x->leak_f[cd.leak_off + st];
      const long long common = std::llrint(std::ldexp(v, x->leak_k));
      const long long own = std::llrint(std::ldexp(v, kc));
      lq[cd.leak_off + st] = (long long)((unsigned long long)common << 32 | (uint32_t)own);
    }
  }
  hipStream_t s = x->stream;
  HIPCHK(x, x->d_cells.upload(x->cell_dev.data(), x->cell_dev.size(), s));
  HIPCHK(x, x->d_leak.upload(x->leak_f.data(), x->leak_f.size(), s));
  HIPCHK(x, x->d_leak_q.upload(lq.data(), lq.size(), s));
  HIPCHK(x, x->d_gate_cell.upload(d->gate_cell, G, s));
  HIPCHK(x, x->d_gate_class.upload(gclass.data(), G, s));
  HIPCHK(x, x->d_gate_pin_off.upload(d->gate_pin_offset, G + 1, s));
  HIPCHK(x, x->d_gate_arc_off.upload(d->gate_arc_offset, G + 1, s));
  HIPCHK(x, x->d_gate_energy_off.upload(eoff.data(), G, s));
  HIPCHK(x, x->d_pin_net.upload(d->pin_net, P, s));
  HIPCHK(x, x->d_pin_gk.upload(gk.data(), P, s));
  HIPCHK(x, x->d_pin_csw.upload(csw.data(), P, s));
  HIPCHK(x, x->d_pin_default.upload(d->pin_default_state, P, s));
  // energies: LDS-staged table interpolation per gate slot
  DArr<float> d_slew, d_cap, d_tdata;
  DArr<SlotDesc> d_descs;
  DArr<ListEntry> d_list;
  DArr<TableDev> d_tables;
  DArr<int32_t> d_desc_off;
  HIPCHK(x, d_slew.upload(slew.data(), slew.size(), s));
  HIPCHK(x, d_cap.upload(cap.data(), cap.size(), s));
  HIPCHK(x, d_tdata.upload(tdata.data(), tdata.size(), s));
  HIPCHK(x, d_descs.upload(descs.data(), descs.size(), s));
  if (!list.empty()) HIPCHK(x, d_list.upload(list.data(), list.size(), s));
  if (!tables.empty()) HIPCHK(x, d_tables.upload(tables.data(), tables.size(), s));
  HIPCHK(x, d_desc_off.upload(desc_off.data(), desc_off.size(), s));
  HIPCHK(x, x->d_energy.alloc(std::max<int64_t>(eacc, 1)));
  HIPCHK(x, x->d_maxbits.alloc(1));
  HIPCHK(x, hipMemsetAsync(x->d_maxbits.p, 0, sizeof(unsigned int), s));
  EnergyArgs ea{};
  ea.gate_cell = x->d_gate_cell.p;
  ea.gate_pin_off = x->d_gate_pin_off.p;
  ea.gate_energy_off = x->d_gate_energy_off.p;
  ea.n_gate = G;
  ea.pin_slew = d_slew.p;
  ea.pin_cap = d_cap.p;
  ea.cells = x->d_cells.p;
  ea.cell_slot_desc_off = d_desc_off.p;
  ea.descs = d_descs.p;
  ea.list = d_list.p;
  ea.tables = d_tables.p;
  ea.n_table = (int32_t)tables.size();
  ea.table_data = d_tdata.p;
  ea.n_table_data = (int32_t)tdata.size();
  ea.energy = x->d_energy.p;
  ea.max_abs_bits = x->d_maxbits.p;
  HIPCHK(x, launch_energy(ea, s));
  unsigned int bits = 0;
  HIPCHK(x, hipMemcpyAsync(&bits, x->d_maxbits.p, sizeof(bits), hipMemcpyDeviceToHost, s));
  HIPCHK(x, hipStreamSynchronize(s));
  std::memcpy(&x->energy_max, &bits, sizeof(float));
  x->have_design = true;
  return GTP_OK;
}

// ---------------------------------------------------------------------------
namespace {

// Ticks per cycle, ceil(period / time scale) (power/PowerUtils.cc:27-32),
// exact when the ratio is an integer up to rounding.
int64_t ticks_per_cycle(double clk_period, double time_scale) {
  const double ratio = clk_period / time_scale;
  int64_t T = (int64_t)std::ceil(ratio);
  if (std::fabs(ratio - std::round(ratio)) < 1e-9 * std::max(1.0, ratio)) T = (int64_t)std::llround(ratio);
  return T;
}

// Checks one trace against the design; *T_out = ticks per cycle.
// scan_events = false leaves the per-event checks to the caller (a sliced
// run checks each event as it packs it)
gtp_status validate_trace(gtp_ctx* x, const gtp_trace* tr, int64_t* T_out, bool long_ok = false,
                          bool scan_events = true) {
  if (!tr || !tr->net_event_offset || tr->n_net != x->N)
    return x->fail(GTP_ERR_ARG, "gtp_trace: n_net must equal the design's n_net");
  if (!(tr->time_scale > 0) || !(tr->clk_period > 0) || tr->t_max <= 0)
    return x->fail(GTP_ERR_ARG, "gtp_trace: time_scale, clk_period and t_max must be positive");
  if (!long_ok && tr->t_max >= (int64_t)EV_TIME_INF)
    return x->fail(GTP_ERR_UNSUPPORTED,
                   "gtp_trace: t_max must be below 2^29 - 1 ticks here (gtp_run_sliced takes longer traces)");
  const int64_t N = x->N;
  const int64_t* off = tr->net_event_offset;
  if (off[0] != 0) return x->fail(GTP_ERR_ARG, "gtp_trace: net_event_offset[0] != 0");
  for (int64_t n = 0; n < N; ++n)
    if (off[n + 1] < off[n]) return x->fail(GTP_ERR_ARG, "gtp_trace: net_event_offset decreases at " + std::to_string(n));
  const int64_t E = off[N];
  if (E > 0 && (!tr->event_time || !tr->event_value)) return x->fail(GTP_ERR_ARG, "gtp_trace: event arrays missing");
  std::atomic<int64_t> bad{-1};
  if (scan_events) parallel_for(N, x->threads, [&](int64_t n) {
    int64_t prev = -1;
    for (int64_t e = off[n]; e < off[n + 1]; ++e) {
      const int64_t t = tr->event_time[e];
      if (t <= prev || t > tr->t_max || tr->event_value[e] > 3) {
        int64_t expect = -1;
        bad.compare_exchange_strong(expect, n);
        return;
      }
      prev = t;
    }
  }, std::max<int64_t>(1, (int64_t)(65536.0 * (double)N / (double)std::max<int64_t>(E, 1))));  // >= ~64k events a thread
  if (bad.load() >= 0)
    return x->fail(GTP_ERR_ARG, "gtp_trace: events of net " + std::to_string(bad.load()) +
                                    " are not strictly increasing in [0, t_max] or carry a value > 3");
  const int64_t T = ticks_per_cycle(tr->clk_period, tr->time_scale);
  if (T < 1 || T >= (int64_t)EV_TIME_INF) return x->fail(GTP_ERR_UNSUPPORTED, "gtp_trace: ticks per cycle out of range");
  *T_out = T;
  return GTP_OK;
}

// Arc delays on the 2^-16 tick grid of the related-pin comparison.
gtp_status ensure_arc_q(gtp_ctx* x, double time_scale);

}  // namespace

namespace {
gtp_status finish_trace(gtp_ctx* x);
gtp_status staging_init(gtp_ctx* x);
template <class F>
gtp_status staged_upload(gtp_ctx* x, const int64_t* so, int64_t N, uint32_t* dst, F pack_net, hipEvent_t done);
// traces of at least this many events are uploaded through the page-locked
// staging ring (smaller ones from a pageable buffer, without the ring's
// 512 MiB); GTP_STAGED_MIN_EVENTS sets another threshold
int64_t staged_min_events() {
  if (const char* e = std::getenv("GTP_STAGED_MIN_EVENTS")) return std::max<long long>(0, std::atoll(e));
  return (int64_t)1 << 25;
}
}  // namespace

GTP_API gtp_status gtp_set_trace(gtp_ctx* x, const gtp_trace* tr) {
  if (!x) return GTP_ERR_ARG;
  if (!x->have_design) return x->fail(GTP_ERR_STATE, "gtp_set_trace: call gtp_set_design first");
  HIPCHK(x, hipSetDevice(x->device));
  x->have_trace = false;
  int64_t T = 0;
  // a trace for the staging ring has its events checked as they are packed
  const bool staged = tr && tr->net_event_offset && tr->n_net == x->N && tr->net_event_offset[x->N] >= staged_min_events();
  {
    const gtp_status st = validate_trace(x, tr, &T, false, !staged);
    if (st != GTP_OK) return st;
  }
  const int64_t N = x->N;
  const int64_t* off = tr->net_event_offset;
  const int64_t E = off[N];
  if (x->opt.max_event_num > 0 && E > x->opt.max_event_num)
    return x->fail(GTP_ERR_OOM, "gtp_trace: " + std::to_string(E) + " events exceed max_event_num");
  x->T = T;
  x->t_max = tr->t_max;
  x->C = (tr->t_max + T - 1) / T;
  x->scale = tr->time_scale;
  x->period = tr->clk_period;
  x->E = E;
  hipStream_t s = x->stream;
  HIPCHK(x, x->d_net_off.upload(off, N + 1, s));
  x->h_net_off.assign(off, off + N + 1);
  x->glitch_valid = false;
  HIPCHK(x, x->d_ev.alloc(E + EV_PAD));
  const int64_t* tm = tr->event_time;
  const uint8_t* val = tr->event_value;
  if (staged) {
    // packed by the host threads into the page-locked ring, a run of nets
    // per buffer, each buffer copied while the next is packed (4 B per event
    // uploaded instead of 9)
    {
      const gtp_status st = staging_init(x);
      if (st != GTP_OK) return st;
    }
    hipEvent_t done = nullptr;
    HIPCHK(x, hipEventCreateWithFlags(&done, hipEventDisableTiming));
    struct EvG {
      hipEvent_t e;
      ~EvG() { (void)hipEventDestroy(e); }
    } evg{done};
    std::atomic<int64_t> bad{-1};
    const gtp_status st = staged_upload(x, off, N, x->d_ev.p,
                                        [&](int64_t n, uint32_t* d) {
                                          if (!pack_words_checked(d, tm + off[n], val + off[n], off[n + 1] - off[n], 0,
                                                                  -1, tr->t_max)) {
                                            int64_t expect = -1;
                                            bad.compare_exchange_strong(expect, n);
                                          }
                                        },
                                        done);
    if (st != GTP_OK) return st;
    HIPCHK(x, hipEventSynchronize(done));
    if (bad.load() >= 0)
      return x->fail(GTP_ERR_ARG, "gtp_trace: events of net " + std::to_string(bad.load()) +
                                      " are not strictly increasing in [0, t_max] or carry a value > 3");
    return finish_trace(x);
  }
  // The event words are packed on the host a chunk at a time (4 B per event
  // uploaded instead of 9), the next chunk on a helper thread while the
  // current one is copied; two buffers alternate.
  {
    const int64_t CH = 1 << 26;
    std::vector<uint32_t> buf[2];
    auto pack = [&](int which, int64_t b) {
      const int64_t n = std::min(CH, E - b);
      buf[which].resize(n);
      uint32_t* d = buf[which].data();
      parallel_for((n + 65535) / 65536, x->threads, [&](int64_t blk) {
        const int64_t i0 = blk * 65536, i1 = std::min(n, (blk + 1) * 65536);
        pack_words_nt(d + i0, tm + b + i0, val + b + i0, i1 - i0, 0);
      }, 1);
    };
    if (E > 0) pack(0, 0);
    int k = 0;
    for (int64_t b = 0; b < E; b += CH, ++k) {
      const int64_t n = std::min(CH, E - b);
      std::thread next;
      if (b + CH < E) next = std::thread([&, b, k] { pack((k + 1) & 1, b + CH); });
      JoinGuard join{next};
      HIPCHK(x, hipMemcpyAsync(x->d_ev.p + b, buf[k & 1].data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
      // the buffer is packed again two chunks later
      HIPCHK(x, hipStreamSynchronize(s));
    }
  }
  return finish_trace(x);
}

namespace {
// The per-trace structures after the event store (x->d_ev, x->d_net_off,
// x->h_net_off, T, C, t_max, E) is in place.
gtp_status finish_trace(gtp_ctx* x) {
  const int64_t N = x->N;
  const int64_t* off = x->h_net_off.data();
  const int64_t E = x->E;
  hipStream_t s = x->stream;
  x->d_tile_idx.release();
  x->tile_idx_tc = 0;
  x->d_perm.release();
  x->perm_tc = 0;
  {
    // 256-event annotation blocks per net
    std::vector<int64_t> co(N + 1, 0);
    for (int64_t n = 0; n < N; ++n) co[n + 1] = co[n] + (off[n + 1] - off[n] + 255) / 256;
    x->n_chunk = co[N];
    HIPCHK(x, x->d_chunk_off.upload(co.data(), co.size(), s));
    HIPCHK(x, x->d_chunk_net.alloc(std::max<int64_t>(x->n_chunk, 1)));
    HIPCHK(x, launch_chunk_net(x->d_chunk_off.p, N, x->d_chunk_net.p, s));
    HIPCHK(x, hipStreamSynchronize(s));
  }
  {
    gtp_status st = ensure_arc_q(x, x->scale);
    if (st != GTP_OK) return st;
  }
  // per-gate work (pin events) for the schedule
  x->gate_work.assign(x->G, 0);
  x->gate_work_max = 0;
  int64_t pev = 0;
  for (int64_t g = 0; g < x->G; ++g) {
    int64_t w = 0;
    for (int64_t q = x->gate_pin_off[g]; q < x->gate_pin_off[g + 1]; ++q) {
      const int32_t net = x->pin_net[q];
      if (net >= 0) w += off[net + 1] - off[net];
    }
    x->gate_work[g] = w;
    x->gate_work_max = std::max(x->gate_work_max, w);
    pev += w;
  }
  x->P_ev = pev;
  for (auto& o : x->orders) o.order.clear();
  x->tuned_tc = x->tuned_rg = 0;
  x->tune_ms = 0;
  x->have_trace = true;
  return GTP_OK;
}
}  // namespace

GTP_API gtp_status gtp_set_trace_vcd(gtp_ctx* x, const gtp_vcd* v, double clk_period) {
  if (!x) return GTP_ERR_ARG;
  if (!x->have_design) return x->fail(GTP_ERR_STATE, "gtp_set_trace_vcd: call gtp_set_design first");
  VcdView vw{};
  if (!vcd_view(v, &vw)) return x->fail(GTP_ERR_STATE, "gtp_set_trace_vcd: no parsed VCD");
  if (vw.device != x->device) return x->fail(GTP_ERR_ARG, "gtp_set_trace_vcd: the VCD was parsed on another device");
  if (vw.n_stream != x->N)
    return x->fail(GTP_ERR_ARG, "gtp_set_trace_vcd: the VCD has " + std::to_string(vw.n_stream) +
                                    " streams, the design " + std::to_string(x->N) + " nets");
  if (!(vw.time_scale > 0) || !(clk_period > 0) || vw.t_max <= 0)
    return x->fail(GTP_ERR_ARG, "gtp_set_trace_vcd: time scale, clock period and the last time must be positive");
  if (vw.t_max >= (int64_t)EV_TIME_INF)
    return x->fail(GTP_ERR_UNSUPPORTED, "gtp_set_trace_vcd: the last time must be below 2^29 - 1 ticks");
  if (x->opt.max_event_num > 0 && vw.n_event > x->opt.max_event_num)
    return x->fail(GTP_ERR_OOM, "gtp_set_trace_vcd: " + std::to_string(vw.n_event) + " events exceed max_event_num");
  const double ratio = clk_period / vw.time_scale;
  int64_t T = (int64_t)std::ceil(ratio);
  if (std::fabs(ratio - std::round(ratio)) < 1e-9 * std::max(1.0, ratio)) T = (int64_t)std::llround(ratio);
  if (T < 1 || T >= (int64_t)EV_TIME_INF) return x->fail(GTP_ERR_UNSUPPORTED, "gtp_set_trace_vcd: ticks per cycle out of range");
  HIPCHK(x, hipSetDevice(x->device));
  x->have_trace = false;
  hipStream_t s = x->stream;
  const int64_t N = x->N, E = vw.n_event;
  x->T = T;
  x->t_max = vw.t_max;
  x->C = (vw.t_max + T - 1) / T;
  x->scale = vw.time_scale;
  x->period = clk_period;
  x->E = E;
  x->glitch_valid = false;
  HIPCHK(x, x->d_net_off.alloc(N + 1));
  HIPCHK(x, hipMemcpyAsync(x->d_net_off.p, vw.d_off, sizeof(int64_t) * (N + 1), hipMemcpyDeviceToDevice, s));
  x->h_net_off.assign(vw.h_off, vw.h_off + N + 1);
  HIPCHK(x, x->d_ev.alloc(E + EV_PAD));
  HIPCHK(x, hipMemcpyAsync(x->d_ev.p, vw.ev, sizeof(uint32_t) * (E + EV_PAD), hipMemcpyDeviceToDevice, s));
  HIPCHK(x, hipStreamSynchronize(s));
  return finish_trace(x);
}

namespace {

gtp_status ensure_arc_q(gtp_ctx* x, double time_scale) {
  if (x->arc_q_scale == time_scale) return GTP_OK;
  std::vector<int32_t> aq(x->arc_delay.size());
  const double f = (double)(1 << ARC_FRAC_BITS) / time_scale;
  for (size_t i = 0; i < aq.size(); ++i) {
    const double dl = x->arc_delay[i];
    if (std::isnan(dl)) {
      aq[i] = ARC_NONE;
      continue;
    }
    const double q = std::round(dl * f);
    if (!(std::fabs(q) < 2147483000.0))
      return x->fail(GTP_ERR_LIMIT, "gtp_trace: an arc delay exceeds 2^15 ticks (the 2^-16-tick comparison grid)");
    aq[i] = (int32_t)q;
  }
  HIPCHK(x, x->d_arc_q.upload(aq.data(), std::max<size_t>(aq.size(), 1), x->stream));
  HIPCHK(x, hipStreamSynchronize(x->stream));
  x->arc_q_scale = time_scale;
  return GTP_OK;
}

// Schedule: gates grouped by kernel (one group per single-output cell type
// without `when` conditions, one generic group), each group sorted by work (pin
// events), heaviest first, and padded with -1 to whole ranges.
void build_order(gtp_ctx* x, gtp_ctx::OrderSet& os, int32_t range_gates, const std::vector<int64_t>& work,
                 int64_t T, bool coarse = false) {
  const int nc = (int)x->cells.size();
  const int nl = nc + 4;
  // per cell: its own list (specialised kernel) or a generic pin-count class
  // (<= 4, <= 6, <= 8, <= 16 pins): list nc + class
  std::vector<int32_t> list_of(nc);
  for (int c = 0; c < nc; ++c) {
    const HostCell& hc = x->cells[c];
    // the specialised kernel compares arrivals on 32 bits: cycles below 2^15 ticks
    const bool fast = hc.sbits == 0 && hc.n_pin - hc.n_in == 1 && hc.n_pin >= 2 && hc.n_pin <= 6 &&
                      T < 32768;
    list_of[c] = fast ? c
                      : nc + (hc.n_pin <= MAXP_SMALL ? 0 : hc.n_pin <= MAXP_LARGE ? 1 : hc.n_pin <= MAXP_WIDE ? 2 : 3);
  }
  // the lists side by side in one array, gate order within each (the
  // context keeps the arrays, so a rebuild touches no new memory)
  std::vector<int64_t> loff(nl + 1, 0);
  for (int64_t g = 0; g < x->G; ++g) ++loff[list_of[x->gate_cell[g]] + 1];
  for (int l = 0; l < nl; ++l) loff[l + 1] += loff[l];
  std::vector<int32_t>& flat = x->order_tmp[0];
  std::vector<int32_t>& sorted = x->order_tmp[1];
  flat.resize(x->G);
  sorted.resize(x->G);
  {
    std::vector<int64_t> pos(loff.begin(), loff.end() - 1);
    for (int64_t g = 0; g < x->G; ++g) flat[pos[list_of[x->gate_cell[g]]]++] = (int32_t)g;
  }
  auto by_work = [&](int32_t p, int32_t q) {
    if (work[p] != work[q]) return work[p] > work[q];
    return p < q;
  };
  // coarse: gates bucketed by work in quarter octaves (a counting sort, gate
  // order kept within a bucket) -- the order of runs over a few listed tiles,
  // rebuilt for each new trace, where a full sort would cost more host time
  // than the run itself
  auto bucket = [&](int32_t g) {
    if (work[g] <= 0) return 0;
    const uint64_t v = (uint64_t)work[g] + 1;
    const int e = 63 - __builtin_clzll(v);
    const int frac = (int)(e >= 2 ? (v >> (e - 2)) & 3u : (v << (2 - e)) & 3u);
    return std::min(255, 1 + 4 * e + frac);
  };
  os.order.clear();
  os.launches.clear();
  os.range_gates = range_gates;
  auto add = [&](int l, int cell, int np) {
    const int64_t b = loff[l], e = loff[l + 1];
    if (e == b) return;
    int32_t* v = sorted.data() + b;
    if (coarse) {
      int64_t cnt[257] = {0};
      for (int64_t i = b; i < e; ++i) ++cnt[256 - bucket(flat[i])];   // larger work first
      for (int k = 1; k <= 256; ++k) cnt[k] += cnt[k - 1];
      for (int64_t i = b; i < e; ++i) v[cnt[256 - bucket(flat[i]) - 1]++] = flat[i];
    } else {
      std::copy(flat.begin() + b, flat.begin() + e, v);
      std::sort(v, v + (e - b), by_work);
    }
    const int32_t nr = (int32_t)((e - b + range_gates - 1) / range_gates);
    const int32_t rb = (int32_t)(os.order.size() / range_gates);
    os.order.insert(os.order.end(), v, v + (e - b));
    os.order.resize((size_t)(rb + nr) * range_gates, -1);
    os.launches.push_back(gtp_ctx::Launch{cell, np, rb, nr});
  };
  for (int c = 0; c < nc; ++c) add(c, c, x->cells[c].n_pin);
  add(nc + 0, -1, MAXP_SMALL);
  add(nc + 1, -1, MAXP_LARGE);
  add(nc + 2, -1, MAXP_WIDE);
  add(nc + 3, -1, MAXP_HUGE);
  os.n_range = (int32_t)(os.order.size() / range_gates);
}

}  // namespace

namespace {

// Everything a power evaluation needs besides the launches themselves.
struct RunPlan {
  gtp_ctx::OrderSet* os = nullptr;
  int64_t TC = 0;
  int32_t n_tile = 0, RG = 0, n_range = 0;
  int block = 256;
  int64_t n_sched = 0;
  double q_e = 0, q_l = 0;
  size_t lds = 0, lds_cell = 0;
  PowerArgs a{};
};

// Fixed-point quanta, powers of two.  Energy: one term <= 2^36 quanta (the
// kernels take the coarse scale, 2^ENERGY_FINE_BITS quanta per unit, one term
// <= 2^20 units, and convert through two exact int32 steps), so a term 1e-5 of
// the design's largest still carries a rounding error below 1e-6 of itself;
// the int64 sums stay exact while one cycle of one gate range, or one gate
// over the trace, adds up less than 2^26 times the largest term.  Leakage: one
// value <= 2^30 quanta, converted through int32 (a gate's leakage sum would
// need 2^33 ticks of trace to overflow).  Uploads the leakage table in quanta.
struct Quanta {
  float inv_e = 0, inv_l = 0;   // kernel scales: coarse energy units, leakage quanta
  double q_e = 0, q_l = 0;      // the quanta (J, W)
};

gtp_status make_quanta(gtp_ctx* x, int64_t t_max, const std::vector<int64_t>& work, Quanta* q) {
  // a gate's per-trace energy sum: at most its pin events x the largest term
  const int64_t wmax = &work == &x->gate_work ? x->gate_work_max
                                              : (work.empty() ? 0 : *std::max_element(work.begin(), work.end()));
  if (wmax >= ((int64_t)1 << (62 - ENERGY_TERM_BITS)))
    return x->fail(GTP_ERR_LIMIT, "gtp_run: a gate has 2^26 or more pin events in one trace (int64 energy sums)");
  double emax = std::max((double)x->energy_max * 1.000001, x->csw_max);
  if (!(emax > 0)) emax = 1e-30;
  const int ke = (int)std::floor((double)ENERGY_TERM_BITS - std::log2(emax));
  const int kl = x->leak_k;
  q->inv_e = std::ldexp(1.0f, ke - ENERGY_FINE_BITS);
  q->inv_l = std::ldexp(1.0f, kl);
  q->q_e = std::ldexp(1.0, -ke);
  q->q_l = std::ldexp(1.0, -kl);
  if (!std::isfinite(q->inv_e) || !std::isfinite(q->inv_l) || q->inv_e == 0.f || q->inv_l == 0.f)
    return x->fail(GTP_ERR_UNSUPPORTED, "gtp_run: energy or leakage magnitudes outside the FP32 range");
  if ((double)t_max * 2147483648.0 >= 9.2e18)
    return x->fail(GTP_ERR_LIMIT, "gtp_run: trace too long for the int64 leakage sums");
  return GTP_OK;
}

// The kernel arguments that do not depend on the trace.
void fill_args(gtp_ctx* x, const gtp_ctx::OrderSet& os, const Quanta& qt, int64_t TC, int64_t T, double scale,
               PowerArgs& a) {
  a = PowerArgs{};
  a.order = os.d_order.p;
  a.n_sched = (int64_t)os.order.size();
  a.gate_cell = x->d_gate_cell.p;
  a.gate_pin_off = x->d_gate_pin_off.p;
  a.gate_arc_off = x->d_gate_arc_off.p;
  a.gate_energy_off = x->d_gate_energy_off.p;
  a.pin_net = x->d_pin_net.p;
  a.pin_gk = x->d_pin_gk.p;
  a.time_scale = (float)scale;
  a.pin_csw = x->d_pin_csw.p;
  a.pin_default = x->d_pin_default.p;
  a.energy = x->d_energy.p;
  a.arc_q = x->d_arc_q.p;
  a.tile_cycles = (int32_t)TC;
  a.T = T;
  a.inv_T = (float)(1.0 / (double)T);
  cycle_magic(T, &a.cyc_m, &a.cyc_sh);
  a.n_range = os.n_range;
  a.range_gates = os.range_gates;
  a.xcd_remap = 1;
  a.cells = x->d_cells.p;
  a.n_cell = (int32_t)x->cell_dev.size();
  a.leak = x->d_leak.p;
  a.leak_q = x->d_leak_q.p;
  a.n_leak = (int32_t)x->leak_f.size();
  a.inv_q_energy = qt.inv_e;
  a.inv_q_leak = qt.inv_l;
}

gtp_status lds_sizes(gtp_ctx* x, const PowerArgs& a, RunPlan* rp) {
  const int64_t TC = a.tile_cycles;
  // the generic kernel stages the leakage tables when they are small (else it
  // reads them from device memory)
  rp->a.leak_lds = a.n_leak <= LEAK_LDS_MAX ? 1 : 0;
  const size_t n_lds_leak = rp->a.leak_lds ? (size_t)a.n_leak : 0;
  rp->lds = (size_t)N_BIN * TC * 8 + n_lds_leak * 8 + (n_lds_leak + (n_lds_leak & 1)) * 4 +
            (size_t)(a.n_cell <= MAX_CELLS_LDS ? a.n_cell : 0) * sizeof(CellDev);
  // the specialised kernel's bins (one spare column per row): up to 4
  // lane-interleaved copies, as many as fit 32 KB (measured: more copies of
  // longer tiles cost occupancy and gain nothing), at least one
  const size_t one = (size_t)N_BIN * (TC + 1) * 8;
  int copies = (int)std::max<size_t>(1, std::min<size_t>(4, (32 * 1024) / one));
  if (const char* e = std::getenv("GTP_BIN_COPIES")) copies = std::max(1, std::atoi(e));
  rp->a.bin_copies = copies;
  rp->lds_cell = one * copies;
  if (rp->lds > 160 * 1024 || rp->lds_cell > 160 * 1024)
    return x->fail(GTP_ERR_LIMIT, "gtp_run: tile needs more LDS than a workgroup has");
  return GTP_OK;
}

// Schedule parameters, order, tile index, fixed-point quanta and scratch.
// tc_req > 0 asks for exactly that tile length; n_work_tiles > 0 is the
// number of tiles the launches will cover (a subset run), which sizes the
// gate ranges and selects the subset run's cached schedule.
gtp_status prepare_run(gtp_ctx* x, int64_t tc_req, int32_t n_work_tiles, RunPlan* rp, int32_t rg_req = 0) {
  hipStream_t s = x->stream;
  int64_t TC = tc_req > 0 ? tc_req
                          : (x->opt.tile_cycles > 0 ? x->opt.tile_cycles : (x->tuned_tc > 0 ? x->tuned_tc : 128));
  if (x->T < (1 << 24)) TC = std::min<int64_t>(TC, std::max<int64_t>(1, ((1 << 24) - 1) / x->T));
  TC = std::max<int64_t>(1, std::min<int64_t>(TC, x->C));
  TC = std::min<int64_t>(TC, 4096);
  if (tc_req > 0 && TC != tc_req)
    return x->fail(GTP_ERR_ARG, "tile_cycles " + std::to_string(tc_req) + " outside what this trace allows (" +
                                    std::to_string(TC) + ")");
  const int32_t n_tile = (int32_t)((x->C + TC - 1) / TC);
  // gates per range: enough workgroups to fill the GPU several times over
  const bool subset = n_work_tiles > 0;
  const int32_t work_tiles = subset ? n_work_tiles : n_tile;
  int32_t RG = subset ? 0 : (rg_req > 0 ? rg_req : (x->opt.gates_per_wg > 0 ? x->opt.gates_per_wg : x->tuned_rg));
  if (RG <= 0) {
    const double want = (double)x->G * work_tiles / 8192.0;
    RG = subset ? 64 : 256;
    while (RG < 4096 && RG * 2 <= want) RG *= 2;
  }
  // the power kernel is compiled for at most 256 lanes per workgroup; a range
  // narrower than that takes a narrower workgroup
  int block = x->opt.block_lanes > 0 ? std::min(256, std::max(64, (x->opt.block_lanes + 63) / 64 * 64)) : 256;
  block = std::min(block, std::max(64, RG));
  gtp_ctx::OrderSet& os = x->orders[subset ? 1 : 0];
  bool rebuilt = false;
  if (os.order.empty() || os.range_gates != RG || os.order_T != x->T) {
    os.order_T = x->T;
    build_order(x, os, RG, x->gate_work, x->T, subset);
    // through page-locked staging: a pageable copy of a 10^5-gate order
    // costs a quarter of a millisecond
    HIPCHK(x, x->order_pinned.reserve(os.order.size()));
    HIPCHK(x, hipStreamSynchronize(s));  // the staging buffer's last copy is done
    std::memcpy(x->order_pinned.p, os.order.data(), os.order.size() * sizeof(int32_t));
    HIPCHK(x, os.d_order.upload(x->order_pinned.p, os.order.size(), s));
    rebuilt = true;
  }
  const int64_t n_sched = (int64_t)os.order.size();
  const int32_t n_range = os.n_range;
  rp->os = &os;
  if (!subset && (x->tile_idx_tc != TC || !x->d_tile_idx.p)) {
    HIPCHK(x, x->d_tile_idx.alloc((size_t)x->N * (n_tile + 1) + 1));
    HIPCHK(x, launch_tile_index(x->d_net_off.p, x->d_ev.p, x->N, n_tile, TC * x->T, nullptr, 0, x->d_tile_idx.p, s));
    x->tile_idx_tc = (int32_t)TC;
    rebuilt = true;
  }
  // per-tile lane order, like the tile index once per trace and schedule
  // (GTP_TILE_ORDER=0: schedule order)
  bool use_perm = !subset && RG <= TILE_ORDER_MAX_RANGE;
  if (const char* e = std::getenv("GTP_TILE_ORDER"))
    if (std::atoi(e) == 0) use_perm = false;
  if (use_perm &&
      (rebuilt || x->perm_tc != TC || x->perm_rg != RG || x->perm_lanes != block || !x->d_perm.p)) {
    // optional: without the memory, the lanes take the schedule order
    if (x->d_perm.alloc((size_t)n_tile * n_sched) != hipSuccess) {
      (void)hipGetLastError();
      x->d_perm.release();
      x->perm_tc = 0;
      use_perm = false;
    } else {
      HIPCHK(x, launch_tile_order(os.d_order.p, n_sched, RG, block, x->d_gate_pin_off.p, x->d_pin_net.p,
                                  x->d_tile_idx.p, n_tile, n_tile, nullptr, nullptr, nullptr, x->d_perm.p, s));
      x->perm_tc = TC;
      x->perm_rg = RG;
      x->perm_lanes = block;
    }
  }
  Quanta qt;
  {
    const gtp_status st = make_quanta(x, x->t_max, x->gate_work, &qt);
    if (st != GTP_OK) return st;
  }
  // scratch
  HIPCHK(x, os.d_pg.alloc((size_t)N_PG * work_tiles * n_sched));
  HIPCHK(x, os.d_pg_tog.alloc((size_t)work_tiles * n_sched));
  HIPCHK(x, os.d_pg_glc.alloc((size_t)work_tiles * n_sched));
  HIPCHK(x, os.d_cyc_part.alloc((size_t)work_tiles * n_range * N_BIN * TC));
  HIPCHK(x, x->d_cycle.alloc((size_t)GTP_N_WAVE * x->C));
  HIPCHK(x, x->d_gate.alloc((size_t)GTP_N_COMPONENT * x->G));
  HIPCHK(x, x->d_total.alloc(5 + 25));
  HIPCHK(x, x->d_total_part.alloc((size_t)TOTAL_BLOCKS * 30));
  HIPCHK(x, x->d_gtog.alloc(x->G));
  HIPCHK(x, x->d_gglc.alloc(x->G));

  PowerArgs& a = rp->a;
  fill_args(x, os, qt, TC, x->T, x->scale, a);
  a.net_off = x->d_net_off.p;
  a.ev = x->d_ev.p;
  a.tile_idx = x->d_tile_idx.p;
  a.perm = use_perm ? x->d_perm.p : nullptr;
  a.n_tile = n_tile;
  a.t_max = x->t_max;
  a.n_cycle = x->C;
  a.n_work_tiles = n_tile;
  a.pg = os.d_pg.p;
  a.pg_tog = os.d_pg_tog.p;
  a.pg_glitch = os.d_pg_glc.p;
  a.cyc_part = os.d_cyc_part.p;
  rp->TC = TC;
  rp->n_tile = n_tile;
  rp->RG = RG;
  rp->n_range = n_range;
  rp->block = block;
  rp->n_sched = n_sched;
  rp->q_e = qt.q_e;
  rp->q_l = qt.q_l;
  return lds_sizes(x, a, rp);
}

// The per-cell-group power launches over n_work_tiles tiles (rp.a), run
// concurrently on the main and the side streams (all default priority): the
// launches write disjoint outputs.
// The clock pin whose edges the workgroups take once per work item (-1: the
// gates step through their clock events): a 3-pin cell whose one clock pin is
// pin 0, tiles of at most CLOCK_EDGE_TILE_MAX cycles, ranges of fewer than
// 2^16 gates (16-bit state counts); GTP_CLOCK_EDGES=0 turns it off.
int clock_edge_pin(const gtp_ctx* x, const gtp_ctx::Launch& L, const RunPlan& rp) {
  if (const char* e = std::getenv("GTP_CLOCK_EDGES"))
    if (std::atoi(e) == 0) return -1;
  const HostCell& hc = x->cells[L.cell];
  if (hc.n_pin != 3 || hc.clock_mask != 1u) return -1;
  if (rp.a.tile_cycles > CLOCK_EDGE_TILE_MAX || rp.a.range_gates >= 65536) return -1;
  return 0;
}

gtp_status launch_groups(gtp_ctx* x, const RunPlan& rp, bool persistent = false) {
  hipStream_t s = x->stream;
  const std::vector<gtp_ctx::Launch>& all = rp.os->launches;
  if (persistent) {
    HIPCHK(x, x->d_wq.alloc(8 * std::max<size_t>(all.size(), 1)));
    HIPCHK(x, launch_zero_u32(x->d_wq.p, (int64_t)(8 * all.size()), s));
  }
  HIPCHK(x, hipEventRecord(x->fork, s));
  for (int k = 0; k < gtp_ctx::N_SIDE; ++k) HIPCHK(x, hipStreamWaitEvent(x->side[k], x->fork, 0));
  const std::vector<gtp_ctx::Launch>& launches = all;
  std::vector<size_t> idx(launches.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](size_t p, size_t q) {
    return launches[p].range_count > launches[q].range_count;
  });
  int k = 0;
  for (size_t li : idx) {
    const gtp_ctx::Launch& L = launches[li];
    hipStream_t ls = (k == 0) ? s : x->side[k - 1];
    k = (k + 1) % (gtp_ctx::N_SIDE + 1);
    const int64_t nw = (int64_t)rp.a.n_work_tiles * L.range_count;
    // persistent: a few workgroups per compute unit drain the launch's queue
    int64_t grid = (nw + 7) / 8 * 8;
    unsigned int* wq = nullptr;
    if (persistent) {
      grid = std::min<int64_t>(grid, (int64_t)x->n_cu * 4 / 8 * 8 + 8);
      wq = x->d_wq.p + 8 * li;
    }
    if (L.cell >= 0)
      HIPCHK(x, launch_power_cell(L.np, clock_edge_pin(x, L, rp), grid, rp.block, rp.lds_cell, ls, rp.a, L.range_begin,
                                  L.range_count, L.cell, wq));
    else
      HIPCHK(x, launch_power(L.np, grid, rp.block, rp.lds, ls, rp.a, L.range_begin, L.range_count, wq));
  }
  for (int j = 0; j < gtp_ctx::N_SIDE; ++j) {
    HIPCHK(x, hipEventRecord(x->join[j], x->side[j]));
    HIPCHK(x, hipStreamWaitEvent(s, x->join[j], 0));
  }
  return GTP_OK;
}

struct EvSet {
  hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
  ~EvSet() {
    for (hipEvent_t v : e)
      if (v) (void)hipEventDestroy(v);
  }
};

}  // namespace

namespace {
gtp_status run_once(gtp_ctx* x, gtp_result* res, int64_t tc_req, int32_t rg_req);
}

// Autotuner: times whole evaluations of the loaded trace over candidate tile
// lengths (at the default gate ranges), then candidate gate-range sizes at the
// best tile length, and keeps the fastest pair for this trace.
GTP_API gtp_status gtp_autotune(gtp_ctx* x, gtp_schedule_info* best) {
  if (!x) return GTP_ERR_ARG;
  if (!x->have_trace) return x->fail(GTP_ERR_STATE, "gtp_autotune: call gtp_set_trace first");
  HIPCHK(x, hipSetDevice(x->device));
  const auto t0 = std::chrono::steady_clock::now();
  gtp_result r{};
  auto time_of = [&](int64_t tc, int32_t rg, double* ms) -> gtp_status {
    const gtp_status st = run_once(x, &r, tc, rg);
    *ms = r.kernel_ms;
    return st;
  };
  const int64_t tc_cap = std::min<int64_t>(std::min<int64_t>(4096, x->C),
                                           x->T < (1 << 24) ? std::max<int64_t>(1, ((1 << 24) - 1) / x->T) : 1);
  // a first evaluation warms the caches and builds the shared state
  double ms = 0, best_ms = 0;
  gtp_status st = time_of(std::min<int64_t>(128, tc_cap), 0, &ms);
  if (st != GTP_OK) return st;
  int64_t best_tc = 0;
  int32_t best_rg = 0;
  for (int64_t tc : {32, 64, 128, 256, 512}) {
    if (tc > tc_cap) continue;
    if ((st = time_of(tc, 0, &ms)) != GTP_OK) return st;
    if (best_tc == 0 || ms < best_ms) {
      best_ms = ms;
      best_tc = tc;
      best_rg = x->sched.gates_per_wg;
    }
  }
  if (best_tc == 0) best_tc = tc_cap;
  for (int32_t rg : {512, 1024, 2048, 4096, 8192}) {
    if (rg == best_rg) continue;
    if ((st = time_of(best_tc, rg, &ms)) != GTP_OK) return st;
    if (ms < best_ms) {
      best_ms = ms;
      best_rg = rg;
    }
  }
  x->tuned_tc = (int32_t)best_tc;
  x->tuned_rg = best_rg;
  x->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  x->sched.tile_cycles = (int32_t)best_tc;
  x->sched.gates_per_wg = best_rg;
  x->sched.autotune_ms = x->tune_ms;
  if (best) *best = x->sched;
  return GTP_OK;
}

GTP_API gtp_status gtp_run(gtp_ctx* x, gtp_result* res) {
  if (!x) return GTP_ERR_ARG;
  if (!x->have_trace) return x->fail(GTP_ERR_STATE, "gtp_run: call gtp_set_trace first");
  if (!res) return x->fail(GTP_ERR_ARG, "gtp_run: null result");
  if (x->opt.autotune && x->tuned_tc == 0 && x->opt.tile_cycles <= 0) {
    const gtp_status st = gtp_autotune(x, nullptr);
    if (st != GTP_OK) return st;
  }
  return run_once(x, res, 0, 0);
}

namespace {
gtp_status run_once(gtp_ctx* x, gtp_result* res, int64_t tc_req, int32_t rg_req) {
  HIPCHK(x, hipSetDevice(x->device));
  hipStream_t s = x->stream;
  EvSet ev;
  for (hipEvent_t& e : ev.e) HIPCHK(x, hipEventCreate(&e));
  HIPCHK(x, hipEventRecord(ev.e[0], s));
  RunPlan rp;
  gtp_status st = prepare_run(x, tc_req, 0, &rp, rg_req);
  if (st != GTP_OK) return st;
  const PowerArgs& a = rp.a;
  HIPCHK(x, hipEventRecord(ev.e[1], s));
  HIPCHK(x, launch_glitch_annotate(x->d_net_off.p, x->d_ev.p, x->d_chunk_off.p, x->d_chunk_net.p, x->n_chunk,
                                   (uint32_t)x->T, s));
  x->glitch_valid = true;
  if ((st = launch_groups(x, rp)) != GTP_OK) return st;
  HIPCHK(x, hipEventRecord(ev.e[2], s));
  HIPCHK(x, launch_cycle_reduce(rp.os->d_cyc_part.p, rp.n_range, (int32_t)rp.TC, x->C, x->T, x->t_max, rp.q_e, rp.q_l,
                                1.0 / x->period, x->d_cycle.p, s));
  HIPCHK(x, launch_gate_reduce(rp.os->d_pg.p, rp.os->d_pg_tog.p, rp.os->d_pg_glc.p, rp.os->d_order.p, rp.n_sched, rp.n_tile, x->G,
                               rp.q_e / ((double)x->t_max * x->scale), rp.q_l / (double)x->t_max,
                               x->d_gate_cell.p, x->d_cells.p, x->d_gate.p, x->d_gtog.p, x->d_gglc.p, s));
  HIPCHK(x, launch_total(x->d_gate.p, x->d_gate_class.p, x->G, x->d_total_part.p, x->d_total.p, s));
  struct Out {
    void* dst;
    const void* src;
    size_t bytes, off;
  };
  Out outs[5] = {{res->cycle_power, x->d_cycle.p, sizeof(double) * GTP_N_WAVE * x->C, 0},
                 {res->gate_power, x->d_gate.p, sizeof(double) * GTP_N_COMPONENT * x->G, 0},
                 {res->gate_toggle_x2, x->d_gtog.p, sizeof(uint64_t) * x->G, 0},
                 {res->gate_glitch, x->d_gglc.p, sizeof(uint64_t) * x->G, 0},
                 {nullptr, x->d_total.p, sizeof(double) * 30, 0}};
  size_t dl_bytes = 0;
  for (Out& o : outs) {
    o.off = dl_bytes;
    dl_bytes += (o.bytes + 255) & ~(size_t)255;
  }
  HIPCHK(x, x->dl.reserve(dl_bytes));
  for (int k = 0; k < 5; ++k)
    if (outs[k].dst || k == 4)
      HIPCHK(x, hipMemcpyAsync(x->dl.p + outs[k].off, outs[k].src, outs[k].bytes, hipMemcpyDeviceToHost, s));
  HIPCHK(x, hipEventRecord(ev.e[3], s));
  HIPCHK(x, hipStreamSynchronize(s));
  {
    // 1 KiB pieces over up to 8 host threads (parallel_for gives a thread
    // per 1024 items)
    constexpr size_t PIECE = (size_t)1 << 10;
    std::vector<std::pair<int, size_t>> pieces;
    for (int k = 0; k < 4; ++k)
      if (outs[k].dst)
        for (size_t b = 0; b < outs[k].bytes; b += PIECE) pieces.emplace_back(k, b);
    parallel_for((int64_t)pieces.size(), std::min(x->threads, 8), [&](int64_t i) {
      const Out& o = outs[pieces[i].first];
      const size_t b = pieces[i].second, n = std::min(PIECE, o.bytes - b);
      std::memcpy(static_cast<unsigned char*>(o.dst) + b, x->dl.p + o.off + b, n);
    });
  }
  double tot[30];
  std::memcpy(tot, x->dl.p + outs[4].off, sizeof(tot));
  float k_ms = 0, t_ms = 0;
  HIPCHK(x, hipEventElapsedTime(&k_ms, ev.e[1], ev.e[2]));
  HIPCHK(x, hipEventElapsedTime(&t_ms, ev.e[0], ev.e[3]));
  if (res->total) std::memcpy(res->total, tot, sizeof(double) * 5);
  if (res->class_total) std::memcpy(res->class_total, tot + 5, sizeof(double) * 25);
  res->n_cycle = x->C;
  res->kernel_ms = k_ms;
  res->total_ms = t_ms;
  gtp_schedule_info& si = x->sched;
  si.tile_cycles = (int32_t)rp.TC;
  si.block_lanes = rp.block;
  si.gates_per_wg = rp.RG;
  si.n_unit = (int64_t)rp.n_tile * x->G;
  si.n_workgroup = (int64_t)rp.n_tile * rp.n_range;
  si.n_launch = (int32_t)rp.os->launches.size();
  si.n_interval = 1;
  si.n_tile = rp.n_tile;
  si.n_event = x->E;
  si.n_pin_event = x->P_ev;
  si.autotune_ms = x->tune_ms;
  return GTP_OK;
}
}  // namespace

GTP_API gtp_status gtp_net_event_counts(gtp_ctx* x, int64_t* counts) {
  if (!x || !counts) return GTP_ERR_ARG;
  if (!x->have_trace) return x->fail(GTP_ERR_STATE, "gtp_net_event_counts: call gtp_set_trace first");
  for (int64_t n = 0; n < x->N; ++n) counts[n] = x->h_net_off[n + 1] - x->h_net_off[n];
  return GTP_OK;
}

GTP_API gtp_status gtp_run_tiles(gtp_ctx* x, int32_t tile_cycles, int32_t n_list, const int32_t* tiles,
                                 double* cycle_power, double* kernel_ms) {
  if (!x) return GTP_ERR_ARG;
  if (!x->have_trace) return x->fail(GTP_ERR_STATE, "gtp_run_tiles: call gtp_set_trace first");
  if (tile_cycles <= 0 || n_list <= 0 || !tiles || !cycle_power)
    return x->fail(GTP_ERR_ARG, "gtp_run_tiles: bad arguments");
  HIPCHK(x, hipSetDevice(x->device));
  hipStream_t s = x->stream;
  RunPlan rp;
  gtp_status st = prepare_run(x, tile_cycles, n_list, &rp);
  if (st != GTP_OK) return st;
  for (int32_t l = 0; l < n_list; ++l)
    if (tiles[l] < 0 || tiles[l] >= rp.n_tile)
      return x->fail(GTP_ERR_ARG, "gtp_run_tiles: tile " + std::to_string(tiles[l]) + " outside [0, " +
                                      std::to_string(rp.n_tile) + ")");
  HIPCHK(x, x->d_tile_list.upload(tiles, n_list, s));
  EvSet ev;
  for (hipEvent_t& e : ev.e) HIPCHK(x, hipEventCreate(&e));
  HIPCHK(x, hipEventRecord(ev.e[0], s));
  // tile index of the listed tiles only
  HIPCHK(x, x->d_tile_idx_sub.alloc((size_t)x->N * n_list + 1));
  HIPCHK(x, launch_tile_index(x->d_net_off.p, x->d_ev.p, x->N, rp.n_tile, rp.TC * x->T, x->d_tile_list.p, n_list,
                              x->d_tile_idx_sub.p, s));
  // glitch annotation of the listed tiles' events, unless the whole trace is annotated
  if (!x->glitch_valid)
    HIPCHK(x, launch_glitch_annotate_tiles(x->d_net_off.p, x->d_ev.p, x->d_tile_idx_sub.p, x->d_tile_list.p, n_list,
                                           x->N, rp.TC * x->T, (uint32_t)x->T, s));
  rp.a.tile_idx = x->d_tile_idx_sub.p;
  rp.a.tile_list = x->d_tile_list.p;
  rp.a.n_work_tiles = n_list;
  if ((st = launch_groups(x, rp)) != GTP_OK) return st;
  const int64_t n_out = (int64_t)n_list * rp.TC;
  HIPCHK(x, x->d_cycle.alloc((size_t)GTP_N_WAVE * std::max<int64_t>(n_out, x->C)));
  HIPCHK(x, launch_cycle_reduce_list(rp.os->d_cyc_part.p, rp.n_range, (int32_t)rp.TC, x->d_tile_list.p, n_list, x->C,
                                     x->T, x->t_max, rp.q_e, rp.q_l, 1.0 / x->period, x->d_cycle.p, s));
  HIPCHK(x, hipEventRecord(ev.e[1], s));
  HIPCHK(x, hipMemcpyAsync(cycle_power, x->d_cycle.p, sizeof(double) * GTP_N_WAVE * n_out, hipMemcpyDeviceToHost, s));
  HIPCHK(x, hipStreamSynchronize(s));
  float ms = 0;
  HIPCHK(x, hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
  if (kernel_ms) *kernel_ms = ms;
  return GTP_OK;
}

gtp::AuxBuffers* gtp::aux_buffers(gtp_ctx* x) { return x ? &x->aux_buf : nullptr; }

hipStream_t gtp::aux_stream(gtp_ctx* x) {
  if (x && !x->aux && (hipSetDevice(x->device) != hipSuccess || hipStreamCreate(&x->aux) != hipSuccess))
    x->aux = nullptr;
  return x ? x->aux : nullptr;
}

int32_t gtp::trace_dev(gtp_ctx* x, TraceDev* out) {
  if (!x || !out) return GTP_ERR_ARG;
  if (!x->have_trace) return x->fail(GTP_ERR_STATE, "call gtp_set_trace first");
  out->device = x->device;
  out->stream = x->stream;
  out->aux_stream = x->aux;
  out->n_net = x->N;
  out->n_cycle = x->C;
  out->T = x->T;
  out->t_max = x->t_max;
  out->net_off = x->d_net_off.p;
  out->ev = x->d_ev.p;
  out->chunk_off = x->d_chunk_off.p;
  out->chunk_net = x->d_chunk_net.p;
  out->n_chunk = x->n_chunk;
  out->host_net_off = x->h_net_off.data();
  return GTP_OK;
}

// Several traces of the design in one set of persistent launches.  Traces are
// grouped into device loads under the event budget; a load's traces become
// one event store of n_trace_load * N virtual nets (trace k's net n is
// k * N + n), annotated in one pass; its work tiles (trace, tile) are
// processed in rounds bounded by the partials' memory, each round one
// persistent launch per cell group over all the round's traces.  Per-trace
// results are exact integer sums, identical to gtp_run of each trace.
namespace {

// One unit of a multi-trace run: a host event CSR (offsets relative to its own
// time/value arrays), the trace it belongs to (t_max, cycles, tiles: the whole
// trace's, so that the first and last tile keep their meaning), the tiles of
// it to evaluate and the result slot they accumulate into.
struct RunItem {
  const int64_t* off;        // [N + 1], off[0] == 0
  const int64_t* time;
  const uint8_t* value;
  int64_t t_max, n_cycle;    // relative to the load's origin (tile_base whole tiles into the trace)
  int32_t n_tile, tile_begin, tile_end, slot;
  const uint32_t* words = nullptr;  // the events already packed as event words (time / value unused)
  uint32_t* dev_words = nullptr;    // first item only: every item's words already on the device, in item
                                    // order from here (EV_PAD padded)
  int32_t tile_base = 0;     // the origin: tiles of the trace before it
  int64_t out_cycles = 0;    // cycles of the trace's waveform (0: n_cycle)
};

// Schedule, fixed-point scales and scratch of a multi-trace run.
struct MultiPlan {
  RunPlan rp;
  Quanta qt;
  int64_t T = 0, TC = 0;
  int32_t W_round = 1, RG = 0, block = 256, n_range = 0;
  int64_t n_sched = 0, total_tiles = 0, ev_budget = 0;
  float invT = 0;
  double time_scale = 0, clk_period = 0;
  std::vector<int64_t> work;
};

// Tile length of multi-trace runs: the option, else the autotuner's choice
// on the resident trace (gtp_autotune), else 128.
int64_t multi_tile_cycles(const gtp_ctx* x, int64_t T) {
  int64_t TC = x->opt.tile_cycles > 0 ? x->opt.tile_cycles : (x->tuned_tc > 0 ? x->tuned_tc : 128);
  if (T < (1 << 24)) TC = std::min<int64_t>(TC, std::max<int64_t>(1, ((1 << 24) - 1) / T));
  return std::max<int64_t>(1, std::min<int64_t>(TC, 4096));
}

// work[g] = pin events of gate g over the run (orders the gates)
// cycles[k]: cycles of trace k (sizes the rounds)
gtp_status multi_setup(gtp_ctx* x, int64_t T, double time_scale, double clk_period, int64_t tmax_max,
                       const std::vector<int64_t>& cycles, std::vector<int64_t>&& work, MultiPlan* mp) {
  hipStream_t s = x->stream;
  const int64_t G = x->G;
  mp->T = T;
  mp->time_scale = time_scale;
  mp->clk_period = clk_period;
  mp->work = std::move(work);
  mp->invT = (float)(1.0 / (double)T);
  {
    const gtp_status st = ensure_arc_q(x, time_scale);
    if (st != GTP_OK) return st;
  }
  {
    const gtp_status st = make_quanta(x, tmax_max, mp->work, &mp->qt);
    if (st != GTP_OK) return st;
  }
  const int64_t TC = multi_tile_cycles(x, T);
  mp->TC = TC;
  int64_t total_tiles = 0;
  for (int64_t c : cycles) total_tiles += (c + TC - 1) / TC;
  mp->total_tiles = total_tiles;
  // partials per work tile: N_PG int64 + 2 uint32 per gate, bins per range
  const int64_t n_sched_est = G + 4096LL * 16;
  size_t free_b = 0, total_b = 0;
  HIPCHK(x, hipMemGetInfo(&free_b, &total_b));
  const double per_tile_est = (double)n_sched_est * (N_PG * 8 + 8) + 64.0 * N_BIN * TC * 8;
  const int64_t budget_scratch = std::min<int64_t>((int64_t)8 << 30, (int64_t)(free_b / 4));
  int32_t W_round = (int32_t)std::max<int64_t>(1, std::min<int64_t>(std::max<int64_t>(total_tiles, 1),
                                                                     (int64_t)(budget_scratch / per_tile_est)));
  if (x->opt.round_tiles > 0) W_round = (int32_t)std::min<int64_t>(x->opt.round_tiles, std::max<int64_t>(total_tiles, 1));
  int32_t RG = x->opt.gates_per_wg > 0 ? x->opt.gates_per_wg : x->tuned_rg;
  if (RG <= 0) {
    const double want = (double)G * std::min<int64_t>(W_round, total_tiles) / 8192.0;
    RG = 256;
    while (RG < 4096 && RG * 2 <= want) RG *= 2;
  }
  int block = x->opt.block_lanes > 0 ? std::min(256, std::max(64, (x->opt.block_lanes + 63) / 64 * 64)) : 256;
  block = std::min(block, std::max(64, RG));
  gtp_ctx::OrderSet& os = x->orders[2];
  build_order(x, os, RG, mp->work, T);
  os.order_T = -1;  // rebuilt for every multi-trace run
  HIPCHK(x, os.d_order.upload(os.order.data(), os.order.size(), s));
  const int64_t n_sched = (int64_t)os.order.size();
  const int32_t n_range = os.n_range;
  HIPCHK(x, os.d_pg.alloc((size_t)N_PG * W_round * n_sched));
  HIPCHK(x, os.d_pg_tog.alloc((size_t)W_round * n_sched));
  HIPCHK(x, os.d_pg_glc.alloc((size_t)W_round * n_sched));
  HIPCHK(x, os.d_cyc_part.alloc((size_t)W_round * n_range * N_BIN * TC));
  RunPlan& rp = mp->rp;
  rp.os = &os;
  fill_args(x, os, mp->qt, TC, T, time_scale, rp.a);
  rp.a.pg = os.d_pg.p;
  rp.a.pg_tog = os.d_pg_tog.p;
  rp.a.pg_glitch = os.d_pg_glc.p;
  rp.a.cyc_part = os.d_cyc_part.p;
  rp.TC = TC;
  rp.RG = RG;
  rp.n_range = n_range;
  rp.block = block;
  rp.n_sched = n_sched;
  rp.q_e = mp->qt.q_e;
  rp.q_l = mp->qt.q_l;
  {
    const gtp_status st = lds_sizes(x, rp.a, &rp);
    if (st != GTP_OK) return st;
  }
  mp->W_round = W_round;
  mp->RG = RG;
  mp->block = block;
  mp->n_range = n_range;
  mp->n_sched = n_sched;
  // device loads under the event budget (one 32-bit word per event)
  mp->ev_budget = x->opt.max_event_num > 0 ? x->opt.max_event_num : (int64_t)((double)free_b * 0.45 / 4.0);
  return GTP_OK;
}

// One device load: the items' events as one CSR of virtual nets (item k's
// net n is k N + n), annotated, indexed, and their tiles evaluated in rounds
// of W_round work tiles; per-gate sums add into acc/cnt of each item's slot,
// per-cycle values go to cyc at slot_base[slot].
gtp_status run_load(gtp_ctx* x, MultiPlan& mp, const std::vector<RunItem>& items, DArr<unsigned long long>& d_acc,
                    DArr<unsigned long long>& d_cnt, DArr<double>& d_cyc, const std::vector<int64_t>& slot_base,
                    float* ms_out) {
  hipStream_t s = x->stream;
  const int64_t N = x->N, G = x->G, T = mp.T, TC = mp.TC;
  const int32_t nt = (int32_t)items.size();
  RunPlan& rp = mp.rp;
  gtp_ctx::OrderSet& os = *rp.os;
  std::vector<int64_t> voff((size_t)nt * N + 1, 0);
  std::vector<int64_t> E(nt);
  int64_t E_load = 0;
  for (int32_t k = 0; k < nt; ++k) {
    const int64_t* off = items[k].off;
    for (int64_t n = 0; n < N; ++n) voff[(size_t)k * N + n] = E_load + off[n];
    E[k] = off[N];
    E_load += off[N];
  }
  voff[(size_t)nt * N] = E_load;
  DArr<int64_t> d_voff, d_chunk;
  DArr<uint32_t> d_ev;
  HIPCHK(x, d_voff.upload(voff.data(), voff.size(), s));
  uint32_t* ev_p = items[0].dev_words;  // the load's words already on the device, items in order
  if (!ev_p) {
    HIPCHK(x, d_ev.alloc(E_load + EV_PAD));
    ev_p = d_ev.p;
    const int64_t CH = 1 << 26;
    DArr<int64_t> st;
    DArr<uint8_t> sv;
    HIPCHK(x, st.alloc(CH));
    HIPCHK(x, sv.alloc(CH));
    int64_t base = 0;
    for (int32_t k = 0; k < nt; ++k) {
      if (items[k].words) {
        HIPCHK(x, hipMemcpyAsync(ev_p + base, items[k].words, sizeof(uint32_t) * E[k], hipMemcpyHostToDevice, s));
        base += E[k];
        continue;
      }
      for (int64_t b0 = 0; b0 < E[k]; b0 += CH) {
        const int64_t n = std::min(CH, E[k] - b0);
        HIPCHK(x, hipMemcpyAsync(st.p, items[k].time + b0, n * sizeof(int64_t), hipMemcpyHostToDevice, s));
        HIPCHK(x, hipMemcpyAsync(sv.p, items[k].value + b0, n, hipMemcpyHostToDevice, s));
        HIPCHK(x, launch_pack(st.p, sv.p, ev_p + base + b0, n, s));
      }
      base += E[k];
    }
    HIPCHK(x, hipStreamSynchronize(s));
  }
  std::vector<int64_t> co((size_t)nt * N + 1, 0);
  for (size_t v = 0; v < (size_t)nt * N; ++v) co[v + 1] = co[v] + (voff[v + 1] - voff[v] + 255) / 256;
  const int64_t n_chunk = co[(size_t)nt * N];
  HIPCHK(x, d_chunk.upload(co.data(), co.size(), s));
  DArr<int32_t> d_chunk_net;
  HIPCHK(x, d_chunk_net.alloc(std::max<int64_t>(n_chunk, 1)));
  HIPCHK(x, launch_chunk_net(d_chunk.p, (int64_t)nt * N, d_chunk_net.p, s));
  // per-item tile indexes and descriptors; work tiles item-major
  std::vector<TraceDesc> td(nt);
  int64_t tidx_total = 0;
  for (int32_t k = 0; k < nt; ++k) {
    td[k].net_base = (int64_t)k * N;
    td[k].tile_idx_base = tidx_total;
    td[k].t_max = items[k].t_max;
    td[k].n_cycle = items[k].n_cycle;
    td[k].n_tile = items[k].n_tile;
    td[k].tile_base = items[k].tile_base;
    td[k].cycle_base = (int64_t)items[k].tile_base * TC;
    td[k].out_cycles = items[k].out_cycles > 0 ? items[k].out_cycles : items[k].n_cycle;
    tidx_total += N * ((int64_t)items[k].n_tile + 1);
  }
  DArr<int32_t> d_tidx, d_wt_tile, d_wt_trace, d_wt_begin, d_rslot;
  DArr<TraceDesc> d_td;
  HIPCHK(x, d_tidx.alloc(std::max<int64_t>(tidx_total, 1)));
  HIPCHK(x, d_td.upload(td.data(), td.size(), s));
  std::vector<int32_t> wt_tile, wt_trace;
  std::vector<int64_t> out_base(nt);
  for (int32_t k = 0; k < nt; ++k) {
    for (int32_t t = items[k].tile_begin; t < items[k].tile_end; ++t) {
      wt_tile.push_back(t);
      wt_trace.push_back(k);
    }
    out_base[k] = slot_base[items[k].slot];
  }
  const int32_t n_wt = (int32_t)wt_tile.size();
  if (n_wt == 0) {
    *ms_out = 0.f;
    return GTP_OK;
  }
  HIPCHK(x, d_wt_tile.upload(wt_tile.data(), wt_tile.size(), s));
  HIPCHK(x, d_wt_trace.upload(wt_trace.data(), wt_trace.size(), s));
  DArr<int64_t> d_out_base;
  HIPCHK(x, d_out_base.upload(out_base.data(), out_base.size(), s));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(x, hipEventCreate(&e0));
  HIPCHK(x, hipEventCreate(&e1));
  struct EvG {
    hipEvent_t a, b;
    ~EvG() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } evg{e0, e1};
  HIPCHK(x, hipEventRecord(e0, s));
  for (int32_t k = 0; k < nt; ++k)
    HIPCHK(x, launch_tile_index(d_voff.p + (int64_t)k * N, ev_p, N, items[k].n_tile, TC * T, nullptr, 0,
                                d_tidx.p + td[k].tile_idx_base, s));
  HIPCHK(x, launch_glitch_annotate(d_voff.p, ev_p, d_chunk.p, d_chunk_net.p, n_chunk, (uint32_t)T, s));
  // per-work-tile lane order, as in full runs (GTP_TILE_ORDER=0, or when it
  // would take more than half the device memory still free after the load's
  // own buffers: schedule order).  Freed with the load.
  const int64_t n_sched_b = (int64_t)os.order.size();
  int32_t* perm_b = nullptr;
  DArr<int32_t> d_perm;
  {
    bool use = rp.RG <= TILE_ORDER_MAX_RANGE;
    if (const char* e = std::getenv("GTP_TILE_ORDER"))
      if (std::atoi(e) == 0) use = false;
    size_t free_b = 0, total_b = 0;
    if (use && hipMemGetInfo(&free_b, &total_b) != hipSuccess) use = false;
    if (use && (double)n_wt * (double)n_sched_b * 4.0 > 0.5 * (double)free_b) use = false;
    if (use) {
      if (d_perm.alloc((size_t)n_wt * n_sched_b) != hipSuccess) {
        (void)hipGetLastError();
        d_perm.release();
      } else {
        HIPCHK(x, launch_tile_order(os.d_order.p, n_sched_b, rp.RG, rp.block, x->d_gate_pin_off.p, x->d_pin_net.p,
                                    d_tidx.p, 0, n_wt, d_wt_tile.p, d_wt_trace.p, d_td.p, d_perm.p, s));
        perm_b = d_perm.p;
      }
    }
  }
  rp.a.net_off = d_voff.p;
  rp.a.ev = ev_p;
  rp.a.tile_idx = d_tidx.p;
  rp.a.traces = d_td.p;
  const int32_t n_range = mp.n_range;
  for (int32_t r0 = 0; r0 < n_wt; r0 += mp.W_round) {
    const int32_t W = std::min(mp.W_round, n_wt - r0);
    rp.a.tile_list = d_wt_tile.p + r0;
    rp.a.wt_trace = d_wt_trace.p + r0;
    rp.a.perm = perm_b ? perm_b + (int64_t)r0 * n_sched_b : nullptr;
    rp.a.n_work_tiles = W;
    {
      const gtp_status st = launch_groups(x, rp, true);
      if (st != GTP_OK) return st;
    }
    HIPCHK(x, launch_cycle_reduce_batch(os.d_cyc_part.p, n_range, (int32_t)TC, rp.a.tile_list, rp.a.wt_trace, W,
                                        d_td.p, d_out_base.p, T, mp.qt.q_e, mp.qt.q_l, 1.0 / mp.clk_period,
                                        d_cyc.p, s));
    // the round's items, their contiguous work-tile ranges and result slots
    std::vector<int32_t> rb, rs;
    for (int32_t i = 0; i < W; ++i)
      if (i == 0 || wt_trace[r0 + i] != wt_trace[r0 + i - 1]) {
        rb.push_back(i);
        rs.push_back(items[wt_trace[r0 + i]].slot);
      }
    rb.push_back(W);
    HIPCHK(x, d_wt_begin.upload(rb.data(), rb.size(), s));
    HIPCHK(x, d_rslot.upload(rs.data(), rs.size(), s));
    HIPCHK(x, launch_gate_accumulate(os.d_pg.p, os.d_pg_tog.p, os.d_pg_glc.p, os.d_order.p, mp.n_sched, W,
                                     d_wt_begin.p, d_rslot.p, (int32_t)rs.size(), G, (long long*)d_acc.p, d_cnt.p,
                                     s));
    // the host vectors of the uploads above are reused next round
    HIPCHK(x, hipStreamSynchronize(s));
  }
  HIPCHK(x, hipEventRecord(e1, s));
  HIPCHK(x, hipStreamSynchronize(s));
  float ms = 0.f;
  HIPCHK(x, hipEventElapsedTime(&ms, e0, e1));
  *ms_out = ms;
  return GTP_OK;
}

// Results of one slot: per-gate power from its exact sums, totals, per-cycle
// values and counts.
gtp_status finish_slot(gtp_ctx* x, const MultiPlan& mp, int32_t slot, const DArr<unsigned long long>& d_acc,
                       const DArr<unsigned long long>& d_cnt, const DArr<double>& d_cyc, int64_t cyc_base,
                       int64_t t_max, int64_t n_cycle, gtp_result& res) {
  hipStream_t s = x->stream;
  const int64_t G = x->G;
  HIPCHK(x, x->d_gate.alloc((size_t)GTP_N_COMPONENT * G));
  HIPCHK(x, x->d_total.alloc(5 + 25));
  HIPCHK(x, x->d_total_part.alloc((size_t)TOTAL_BLOCKS * 30));
  const double tm = (double)t_max;
  HIPCHK(x, launch_gate_finalize((const long long*)d_acc.p + (int64_t)slot * N_PG * G, G,
                                 mp.qt.q_e / (tm * mp.time_scale), mp.qt.q_l / tm, x->d_gate_cell.p, x->d_cells.p,
                                 x->d_gate.p, s));
  HIPCHK(x, launch_total(x->d_gate.p, x->d_gate_class.p, G, x->d_total_part.p, x->d_total.p, s));
  if (res.gate_power)
    HIPCHK(x, hipMemcpyAsync(res.gate_power, x->d_gate.p, sizeof(double) * GTP_N_COMPONENT * G,
                             hipMemcpyDeviceToHost, s));
  if (res.cycle_power)
    HIPCHK(x, hipMemcpyAsync(res.cycle_power, d_cyc.p + cyc_base, sizeof(double) * GTP_N_WAVE * n_cycle,
                             hipMemcpyDeviceToHost, s));
  if (res.gate_toggle_x2)
    HIPCHK(x, hipMemcpyAsync(res.gate_toggle_x2, d_cnt.p + (int64_t)slot * 2 * G, sizeof(uint64_t) * G,
                             hipMemcpyDeviceToHost, s));
  if (res.gate_glitch)
    HIPCHK(x, hipMemcpyAsync(res.gate_glitch, d_cnt.p + ((int64_t)slot * 2 + 1) * G, sizeof(uint64_t) * G,
                             hipMemcpyDeviceToHost, s));
  double tot[30];
  HIPCHK(x, hipMemcpyAsync(tot, x->d_total.p, sizeof(tot), hipMemcpyDeviceToHost, s));
  HIPCHK(x, hipStreamSynchronize(s));
  if (res.total) std::memcpy(res.total, tot, sizeof(double) * 5);
  if (res.class_total) std::memcpy(res.class_total, tot + 5, sizeof(double) * 25);
  res.n_cycle = n_cycle;
  return GTP_OK;
}

// Creates the staging ring and its stream on first use.
gtp_status staging_init(gtp_ctx* x) {
  for (int k = 0; k < gtp_ctx::N_STAGE; ++k) {
    if (!x->stage_ev[k]) HIPCHK(x, hipEventCreateWithFlags(&x->stage_ev[k], hipEventDisableTiming));
    HIPCHK(x, x->stage[k].reserve((size_t)gtp_ctx::STAGE_WORDS));
  }
  return GTP_OK;
}

// Packs the event words of one load into dst (device) through the staging
// ring: nets [0, N) with word offsets so[N + 1], pack_net(n, out) writing net
// n's so[n + 1] - so[n] words.  Nets are taken in runs that fill a staging
// buffer; each run is packed by the host threads and copied on the staging
// stream while the next run is packed.  done (recorded on the staging stream
// after the last copy) tells a consumer stream when dst is complete.  Runs on
// any host thread (it sets the context's device itself).
template <class F>
gtp_status staged_upload(gtp_ctx* x, const int64_t* so, int64_t N, uint32_t* dst, F pack_net, hipEvent_t done) {
  HIPCHK(x, hipSetDevice(x->device));
  const hipStream_t us = gtp::aux_stream(x);
  if (!us) return x->fail(GTP_ERR_DEVICE, "staging: no upload stream");
  const int64_t cap = gtp_ctx::STAGE_WORDS;
  std::vector<uint32_t> big;
  for (int64_t n0 = 0; n0 < N;) {
    // the longest run of nets from n0 whose words fit one buffer
    int64_t n1 = std::upper_bound(so + n0 + 1, so + N + 1, so[n0] + cap) - so - 1;
    if (n1 <= n0) {
      // one net larger than a buffer: packed whole and copied from pageable memory
      big.resize((size_t)(so[n0 + 1] - so[n0]));
      pack_net(n0, big.data());
      HIPCHK(x, hipMemcpyAsync(dst + so[n0], big.data(), big.size() * sizeof(uint32_t), hipMemcpyHostToDevice, us));
      HIPCHK(x, hipStreamSynchronize(us));
      n0 += 1;
      continue;
    }
    const int b = x->stage_next;
    x->stage_next = (b + 1) % gtp_ctx::N_STAGE;
    if (x->stage_pending[b]) HIPCHK(x, hipEventSynchronize(x->stage_ev[b]));
    x->stage_pending[b] = false;
    uint32_t* h = x->stage[b].p;
    const int64_t base = so[n0];
    // (a net is thousands of events: one per thread is enough)
    parallel_for(n1 - n0, x->threads, [&](int64_t i) { pack_net(n0 + i, h + (so[n0 + i] - base)); }, 1);
    if (so[n1] > base) {
      HIPCHK(x, hipMemcpyAsync(dst + base, h, (size_t)(so[n1] - base) * sizeof(uint32_t), hipMemcpyHostToDevice, us));
      HIPCHK(x, hipEventRecord(x->stage_ev[b], us));
      x->stage_pending[b] = true;
    }
    n0 = n1;
  }
  HIPCHK(x, hipEventRecord(done, us));
  return GTP_OK;
}

void record_multi_schedule(gtp_ctx* x, const MultiPlan& mp, int64_t n_event) {
  gtp_schedule_info& si = x->sched;
  si.tile_cycles = (int32_t)mp.TC;
  si.block_lanes = mp.block;
  si.gates_per_wg = mp.RG;
  si.n_unit = mp.total_tiles * x->G;
  si.n_workgroup = mp.total_tiles * mp.n_range;
  si.n_launch = (int32_t)mp.rp.os->launches.size();
  si.n_interval = 1;
  si.n_tile = mp.total_tiles;
  si.n_event = n_event;
  si.n_pin_event = std::accumulate(mp.work.begin(), mp.work.end(), (int64_t)0);
  si.autotune_ms = 0;
}

}  // namespace

GTP_API gtp_status gtp_run_batch(gtp_ctx* x, int32_t n_trace, const gtp_trace* traces, gtp_result* results) {
  if (!x) return GTP_ERR_ARG;
  if (!x->have_design) return x->fail(GTP_ERR_STATE, "gtp_run_batch: call gtp_set_design first");
  if (n_trace <= 0 || !traces || !results) return x->fail(GTP_ERR_ARG, "gtp_run_batch: no traces");
  HIPCHK(x, hipSetDevice(x->device));
  hipStream_t s = x->stream;
  const int64_t N = x->N, G = x->G;
  // validation: one tick length and time scale for the batch
  int64_t T = 0;
  std::vector<int64_t> C(n_trace), E(n_trace);
  int64_t tmax_max = 0;
  for (int32_t k = 0; k < n_trace; ++k) {
    int64_t Tk = 0;
    const gtp_status st = validate_trace(x, &traces[k], &Tk);
    if (st != GTP_OK) return x->fail(st, "trace " + std::to_string(k) + ": " + x->err);
    if (k == 0) T = Tk;
    if (Tk != T || traces[k].time_scale != traces[0].time_scale || traces[k].clk_period != traces[0].clk_period)
      return x->fail(GTP_ERR_UNSUPPORTED, "gtp_run_batch: traces must share time_scale and clk_period");
    C[k] = (traces[k].t_max + T - 1) / T;
    E[k] = traces[k].net_event_offset[N];
    tmax_max = std::max<int64_t>(tmax_max, traces[k].t_max);
  }
  // gates ordered by their pin events summed over the batch
  std::vector<int64_t> work(G, 0);
  parallel_for(G, x->threads, [&](int64_t g) {
    int64_t w = 0;
    for (int64_t q = x->gate_pin_off[g]; q < x->gate_pin_off[g + 1]; ++q) {
      const int32_t net = x->pin_net[q];
      if (net < 0) continue;
      for (int32_t k = 0; k < n_trace; ++k) w += traces[k].net_event_offset[net + 1] - traces[k].net_event_offset[net];
    }
    work[g] = w;
  });
  // tile count with the schedule's tile length (fixed by multi_setup)
  MultiPlan mp;
  {
    const gtp_status st = multi_setup(x, T, traces[0].time_scale, traces[0].clk_period, tmax_max, C,
                                      std::move(work), &mp);
    if (st != GTP_OK) return st;
  }
  const int64_t TC = mp.TC;
  std::vector<int32_t> n_tile(n_trace);
  for (int32_t k = 0; k < n_trace; ++k) n_tile[k] = (int32_t)((C[k] + TC - 1) / TC);
  // loads under the budget, in 4-byte words: a trace's event words, its tile
  // index (N x (tiles + 1)) and its per-net offsets and block maps (4 words a net)
  std::vector<std::pair<int32_t, int32_t>> loads;
  auto load_words = [&](int32_t k) { return E[k] + N * ((int64_t)n_tile[k] + 1) + 4 * N; };
  for (int32_t k = 0; k < n_trace;) {
    if (load_words(k) > mp.ev_budget)
      return x->fail(GTP_ERR_OOM, "gtp_run_batch: trace " + std::to_string(k) + " alone exceeds the event budget");
    int64_t e = 0;
    int32_t j = k;
    while (j < n_trace && e + load_words(j) <= mp.ev_budget) e += load_words(j++);
    loads.emplace_back(k, j);
    k = j;
  }
  // exact per-trace sums and per-cycle values of every trace
  std::vector<int64_t> slot_base(n_trace + 1, 0);
  for (int32_t k = 0; k < n_trace; ++k) slot_base[k + 1] = slot_base[k] + (int64_t)GTP_N_WAVE * C[k];
  DArr<unsigned long long> d_acc, d_cnt;
  DArr<double> d_cyc;
  HIPCHK(x, d_acc.alloc((size_t)n_trace * N_PG * G + 1));
  HIPCHK(x, d_cnt.alloc((size_t)n_trace * 2 * G + 1));
  HIPCHK(x, d_cyc.alloc(std::max<int64_t>(slot_base[n_trace], 1)));
  HIPCHK(x, launch_zero_u64(d_acc.p, (int64_t)n_trace * N_PG * G, s));
  HIPCHK(x, launch_zero_u64(d_cnt.p, (int64_t)n_trace * 2 * G, s));
  float kernel_ms_total = 0.f;
  for (const auto& ld : loads) {
    std::vector<RunItem> items;
    for (int32_t k = ld.first; k < ld.second; ++k)
      items.push_back(RunItem{traces[k].net_event_offset, traces[k].event_time, traces[k].event_value,
                              traces[k].t_max, C[k], n_tile[k], 0, n_tile[k], k});
    float ms = 0.f;
    const gtp_status st = run_load(x, mp, items, d_acc, d_cnt, d_cyc, slot_base, &ms);
    if (st != GTP_OK) return st;
    kernel_ms_total += ms;
  }
  for (int32_t k = 0; k < n_trace; ++k) {
    const gtp_status st = finish_slot(x, mp, k, d_acc, d_cnt, d_cyc, slot_base[k], traces[k].t_max, C[k], results[k]);
    if (st != GTP_OK) return st;
  }
  for (int32_t k = 0; k < n_trace; ++k) {
    results[k].kernel_ms = kernel_ms_total;
    results[k].total_ms = kernel_ms_total;
  }
  record_multi_schedule(x, mp, std::accumulate(E.begin(), E.end(), (int64_t)0));
  return GTP_OK;
}

GTP_API gtp_status gtp_run_sliced(gtp_ctx* x, const gtp_trace* tr, int64_t max_events, gtp_result* res) {
  if (!x) return GTP_ERR_ARG;
  if (!x->have_design) return x->fail(GTP_ERR_STATE, "gtp_run_sliced: call gtp_set_design first");
  if (!tr || !res) return x->fail(GTP_ERR_ARG, "gtp_run_sliced: no trace or result");
  HIPCHK(x, hipSetDevice(x->device));
  const auto wall0 = std::chrono::steady_clock::now();
  hipStream_t s = x->stream;
  const int64_t N = x->N, G = x->G;
  int64_t T = 0;
  {
    // (each event is checked as it is packed, below)
    const gtp_status st = validate_trace(x, tr, &T, true, false);
    if (st != GTP_OK) return st;
  }
  const int64_t* off = tr->net_event_offset;
  const int64_t* tm = tr->event_time;
  const int64_t C = (tr->t_max + T - 1) / T;
  const int64_t E = off[N];
  std::vector<int64_t> work(G, 0);
  parallel_for(G, x->threads, [&](int64_t g) {
    int64_t w = 0;
    for (int64_t q = x->gate_pin_off[g]; q < x->gate_pin_off[g + 1]; ++q) {
      const int32_t net = x->pin_net[q];
      if (net >= 0) w += off[net + 1] - off[net];
    }
    work[g] = w;
  });
  MultiPlan mp;
  {
    const gtp_status st = multi_setup(x, T, tr->time_scale, tr->clk_period, tr->t_max, std::vector<int64_t>{C},
                                      std::move(work), &mp);
    if (st != GTP_OK) return st;
  }
  const int64_t TC = mp.TC;
  const int32_t n_tile = (int32_t)((C + TC - 1) / TC);
  // event words of one interval: the budget less the interval load's tile
  // index (the whole trace's tiles) and per-net offsets and block maps
  const int64_t fixed_words = N * ((int64_t)n_tile + 1) + 4 * N;
  const int64_t budget = max_events > 0 ? std::min(max_events, mp.ev_budget - fixed_words)
                                        : mp.ev_budget - fixed_words;
  if (budget <= 0)
    return x->fail(GTP_ERR_OOM, "gtp_run_sliced: the tile index alone exceeds the device event budget");
  // Each interval [k0, k1) is loaded with times relative to an origin of kb
  // whole tiles, K cycles and one more before the interval; a trace may then be
  // of any length, an interval (with those cycles) spanning below 2^29 - 1
  // ticks.  K cycles exceed the widest glitch pulse whose scaling ratio is
  // below 1, so the events of the slice before the interval (its last cycle
  // group and the event before that group) can move to K cycles before it, and
  // the two after it to K cycles after it, without changing any result: their
  // cycle groups, their gaps within a cycle and the state they leave are kept,
  // and every gap to the interval's events stays wider than K - 1 cycles
  // (ratio 1, and no arrival in the same cycle).
  const int64_t K = (int64_t)std::ceil(x->glitch_sat_s / tr->time_scale / (double)T) + 2;
  auto origin_tile = [&](int32_t k0) {
    return (int32_t)std::max<int64_t>(0, ((int64_t)k0 * TC - K - 1) / TC);
  };
  // relative span of an interval's load (ticks)
  auto span = [&](int32_t k0, int32_t k1) {
    const int64_t base = (int64_t)origin_tile(k0) * TC * T;
    return (k1 >= n_tile ? tr->t_max : ((int64_t)k1 * TC + K + 2) * T) - base;
  };
  if (span(n_tile - 1, n_tile) >= (int64_t)EV_TIME_INF || span(0, 1) >= (int64_t)EV_TIME_INF)
    return x->fail(GTP_ERR_UNSUPPORTED, "gtp_run_sliced: one tile of " + std::to_string(TC) + " cycles of " +
                                            std::to_string(T) + " ticks spans 2^29 - 1 ticks or more");
  // The slice of net n for tiles [k0, k1): its events from the one before the
  // cycle group of the last event before the interval (so that group's
  // positions and the state before the interval come out as in a whole-trace
  // run; event 0 keeps index 0 when the slice starts there) to the second
  // event at or after the interval's end (the flag of the first one).
  // The first event at or after time t among events [lo, b) of a net (b when
  // none): a guess interpolated in time, then a galloping search around it,
  // so that a lookup costs a few cache lines instead of a binary search over
  // the whole net.  Times are strictly increasing (validated).
  auto seek = [&](int64_t lo, int64_t b, int64_t t) -> int64_t {
    if (lo >= b || tm[lo] >= t) return lo;
    if (tm[b - 1] < t) return b;
    int64_t L = lo, R = b - 1;  // tm[L] < t <= tm[R]
    const double f = (double)(t - tm[lo]) / ((double)(tm[b - 1] - tm[lo]) + 1.0);
    const int64_t g = std::min(b - 1, std::max(lo + 1, lo + (int64_t)(f * (double)(b - 1 - lo))));
    if (tm[g] >= t) {
      R = g;
      for (int64_t step = 1;; step *= 2) {
        const int64_t c = R - step;
        if (c <= L) break;
        if (tm[c] < t) {
          L = c;
          break;
        }
        R = c;
      }
    } else {
      L = g;
      for (int64_t step = 1;; step *= 2) {
        const int64_t c = L + step;
        if (c >= R) break;
        if (tm[c] >= t) {
          R = c;
          break;
        }
        L = c;
      }
    }
    while (R - L > 1) {
      const int64_t m = L + (R - L) / 2;
      if (tm[m] < t) L = m;
      else R = m;
    }
    return R;
  };
  // The slices of tiles [k0, k1): per net the first event at or after the
  // interval start (i0) and end (i1), and the slice [lo, hi) loaded.  `from`,
  // when given, holds each net's i0 (the i1 of the interval before).
  auto slice_bounds = [&](int32_t k0, int32_t k1, const int64_t* from, std::vector<int64_t>& lo,
                          std::vector<int64_t>& hi, std::vector<int64_t>& I0, std::vector<int64_t>& I1) {
    const int64_t t0 = (int64_t)k0 * TC * T;
    const int64_t t1 = k1 >= n_tile ? INT64_MAX : (int64_t)k1 * TC * T;
    std::atomic<int64_t> tot{0};
    parallel_for(N, x->threads, [&](int64_t n) {
      const int64_t a = off[n], b = off[n + 1];
      const int64_t i0 = from ? from[n] : seek(a, b, t0);
      const int64_t i1 = t1 == INT64_MAX ? b : seek(i0, b, t1);
      int64_t st = a;
      if (i0 - a >= 2) {
        const int64_t cs = (tm[i0 - 1] / T) * T;
        int64_t f = i0 - 1;
        while (f - 1 >= a + 1 && tm[f - 1] >= cs) --f;
        st = std::min(i0 - 2, f - 1);
        if (st < a) st = a;
      }
      lo[n] = st;
      hi[n] = std::min(b, i1 + 2);
      I0[n] = i0;
      I1[n] = i1;
      tot += hi[n] - lo[n];
    }, 64);
    return tot.load();
  };
  // intervals: whole tiles, halved until the slice fits the budget
  std::vector<std::pair<int32_t, int32_t>> ivals;
  int64_t max_words = 0;  // events of the largest interval's load
  {
    std::vector<int64_t> sb(N), se(N), i0v(N), i1v(N), next(off, off + N);  // next: i0 of the next interval
    const int32_t guess = (int32_t)std::max<int64_t>(1, std::min<int64_t>(n_tile, (E + budget - 1) / std::max<int64_t>(budget, 1) * 2));
    const int32_t step0 = (n_tile + guess - 1) / guess;
    for (int32_t k0 = 0; k0 < n_tile;) {
      int32_t step = std::max(1, std::min(step0, n_tile - k0));
      while (step > 1 && span(k0, k0 + step) >= (int64_t)EV_TIME_INF) step = (step + 1) / 2;
      if (span(k0, k0 + step) >= (int64_t)EV_TIME_INF)
        return x->fail(GTP_ERR_UNSUPPORTED, "gtp_run_sliced: one tile spans 2^29 - 1 ticks or more");
      while (true) {
        const int64_t need = slice_bounds(k0, k0 + step, next.data(), sb, se, i0v, i1v);
        if (need <= budget) {
          max_words = std::max(max_words, need);
          break;
        }
        if (step == 1)
          return x->fail(GTP_ERR_OOM, "gtp_run_sliced: one tile of " + std::to_string(TC) + " cycles needs " +
                                          std::to_string(need) + " events, over the budget");
        step = (step + 1) / 2;
      }
      ivals.emplace_back(k0, k0 + step);
      next.swap(i1v);
      k0 += step;
    }
  }
  std::vector<int64_t> slot_base{0, (int64_t)GTP_N_WAVE * C};
  DArr<unsigned long long> d_acc, d_cnt;
  DArr<double> d_cyc;
  HIPCHK(x, d_acc.alloc((size_t)N_PG * G + 1));
  HIPCHK(x, d_cnt.alloc((size_t)2 * G + 1));
  HIPCHK(x, d_cyc.alloc(std::max<int64_t>(slot_base[1], 1)));
  HIPCHK(x, launch_zero_u64(d_acc.p, (int64_t)N_PG * G, s));
  HIPCHK(x, launch_zero_u64(d_cnt.p, (int64_t)2 * G, s));
  float kernel_ms_total = 0.f;
  // Each interval's slices are packed on the host straight into event words
  // (4 B per event uploaded) through the page-locked staging ring, into one of
  // two device buffers: a helper thread slices, packs and uploads the next
  // interval while the current one is evaluated (one buffer when two do not
  // fit the device: then one interval after the other).
  struct Slices {
    std::vector<int64_t> sb, se, so, i0, i1;
  };
  Slices sl[2];
  const uint8_t* val = tr->event_value;
  // each net's i1 of the interval prepared last, the i0 of the next one (the
  // intervals are prepared in order)
  std::vector<int64_t> next_i0(off, off + N);
  auto bounds = [&](Slices& P, const std::pair<int32_t, int32_t>& iv) {
    P.sb.resize(N);
    P.se.resize(N);
    P.so.resize(N + 1);
    P.i0.resize(N);
    P.i1.resize(N);
    slice_bounds(iv.first, iv.second, next_i0.data(), P.sb, P.se, P.i0, P.i1);
    std::copy(P.i1.begin(), P.i1.end(), next_i0.begin());
    P.so[0] = 0;
    for (int64_t n = 0; n < N; ++n) P.so[n + 1] = P.so[n] + (P.se[n] - P.sb[n]);
  };
  // the first net found with events out of order, past t_max or above 3
  std::atomic<int64_t> bad_net{-1};
  // net n's slice of interval iv as event words, times relative to the origin
  auto pack_net = [&](const Slices& P, const std::pair<int32_t, int32_t>& iv, int64_t n, uint32_t* d) {
    const int64_t base = (int64_t)origin_tile(iv.first) * TC * T;
    const int64_t c0 = (int64_t)iv.first * TC, t0 = c0 * T;               // interval start
    const bool to_end = iv.second >= n_tile;
    const int64_t c1 = (int64_t)iv.second * TC, t1 = c1 * T;              // interval end (not the last)
    const int64_t a = P.sb[n], b = P.se[n];
    const int64_t i0 = P.i0[n];
    const int64_t i1 = to_end ? b : P.i1[n];
    // before the interval: the event before the group (a) at or after the
    // start of cycle c0 - K - 1, the group (a + 1 .. i0 - 1, one cycle)
    // moved to cycle c0 - K when it lies earlier
    int64_t gshift = 0;
    if (i0 > a) {
      const int64_t gc = tm[i0 - 1] / T;
      if (gc < c0 - K) gshift = (gc - (c0 - K)) * T;
    }
    const int64_t t_before = (c0 - K - 1) * T;
    // after it: the first event (i1) moved to cycle c1 + K when later, the
    // second (i1 + 1) kept in its cycle or put at the start of the next one
    int64_t ashift = 0, ac_new = 0;
    if (!to_end && i1 < b) {
      const int64_t ac = tm[i1] / T;
      if (ac > c1 + K) ashift = (ac - (c1 + K)) * T;
      ac_new = ac - ashift / T;
    }
    for (int64_t i = a; i < i0; ++i) {
      const int64_t t = i > a ? tm[i] - gshift : std::max(tm[i], t_before);
      *d++ = ((uint32_t)(t - base) << EV_SHIFT) | (uint32_t)(val[i] & 3u);
    }
    // the interval's own events (the bulk): times relative to the origin.
    // The bulks of the intervals partition each net's events, so checking
    // each bulk and the event before it checks the whole trace.
    if (!pack_words_checked(d, tm + i0, val + i0, i1 - i0, base, i0 > off[n] ? tm[i0 - 1] : -1, tr->t_max)) {
      int64_t expect = -1;
      bad_net.compare_exchange_strong(expect, n);
    }
    d += i1 - i0;
    for (int64_t i = i1; i < b; ++i) {
      const int64_t t = (i == i1 || tm[i] / T == tm[i1] / T) ? tm[i] - ashift : std::min(tm[i], (ac_new + 1) * T);
      *d++ = ((uint32_t)(t - base) << EV_SHIFT) | (uint32_t)(val[i] & 3u);
    }
  };
  // device buffers sized for the largest interval; the second one only when
  // it takes at most half the memory still free (the loads' tile indexes and
  // block maps need the rest)
  {
    const gtp_status st = staging_init(x);
    if (st != GTP_OK) return st;
  }
  DArr<uint32_t> d_words[2];
  HIPCHK(x, d_words[0].alloc(max_words + EV_PAD));
  int n_buf = 1;
  bool two = ivals.size() > 1;
  if (const char* e = std::getenv("GTP_SLICE_BUFFERS"))
    if (std::atoi(e) == 1) two = false;
  if (two) {
    size_t free_b = 0, total_b = 0;
    HIPCHK(x, hipMemGetInfo(&free_b, &total_b));
    if ((double)(max_words + EV_PAD) * 4.0 <= 0.5 * (double)free_b) {
      if (d_words[1].alloc(max_words + EV_PAD) == hipSuccess) n_buf = 2;
      else (void)hipGetLastError();
    }
  }
  hipEvent_t ready[2] = {nullptr, nullptr};
  HIPCHK(x, hipEventCreateWithFlags(&ready[0], hipEventDisableTiming));
  HIPCHK(x, hipEventCreateWithFlags(&ready[1], hipEventDisableTiming));
  struct EvG2 {
    hipEvent_t* e;
    ~EvG2() {
      (void)hipEventDestroy(e[0]);
      (void)hipEventDestroy(e[1]);
    }
  } evg2{ready};
  auto prepare = [&](size_t i) {
    Slices& P = sl[i & 1];
    bounds(P, ivals[i]);
    const auto iv = ivals[i];
    const gtp_status st = staged_upload(x, P.so.data(), N, d_words[i % n_buf].p,
                                        [&](int64_t n, uint32_t* d) { pack_net(P, iv, n, d); }, ready[i & 1]);
    if (st == GTP_OK && bad_net.load() >= 0)
      return x->fail(GTP_ERR_ARG, "gtp_trace: events of net " + std::to_string(bad_net.load()) +
                                      " are not strictly increasing in [0, t_max] or carry a value > 3");
    return st;
  };
  {
    const gtp_status st = prepare(0);
    if (st != GTP_OK) return st;
  }
  for (size_t i = 0; i < ivals.size(); ++i) {
    Slices& P = sl[i & 1];
    std::thread next;
    gtp_status next_st = GTP_OK;
    if (i + 1 < ivals.size() && n_buf == 2) next = std::thread([&, i] { next_st = prepare(i + 1); });
    {
      JoinGuard join{next};
      // the load as a trace from its origin: kb tiles in, to the trace's end
      // (the last interval) or just past the two events after the interval
      const int32_t kb = origin_tile(ivals[i].first);
      const int64_t base = (int64_t)kb * TC * T;
      const int64_t tm_rel = ivals[i].second >= n_tile ? tr->t_max - base
                                                       : ((int64_t)ivals[i].second * TC + K + 2) * T - base;
      const int64_t c_rel = ivals[i].second >= n_tile ? C - (int64_t)kb * TC : (tm_rel + T - 1) / T;
      RunItem item{P.so.data(), nullptr, nullptr, tm_rel, c_rel, n_tile - kb, ivals[i].first - kb,
                   ivals[i].second - kb, 0};
      item.dev_words = d_words[i % n_buf].p;
      item.tile_base = kb;
      item.out_cycles = C;
      std::vector<RunItem> items{item};
      HIPCHK(x, hipStreamWaitEvent(s, ready[i & 1], 0));
      float ms = 0.f;
      const gtp_status st = run_load(x, mp, items, d_acc, d_cnt, d_cyc, slot_base, &ms);
      if (st != GTP_OK) return st;  // (the helper thread is joined first)
      kernel_ms_total += ms;
    }
    if (next_st != GTP_OK) return next_st;
    if (i + 1 < ivals.size() && n_buf == 1) {
      const gtp_status st = prepare(i + 1);
      if (st != GTP_OK) return st;
    }
  }
  {
    const gtp_status st = finish_slot(x, mp, 0, d_acc, d_cnt, d_cyc, 0, tr->t_max, C, *res);
    if (st != GTP_OK) return st;
  }
  res->kernel_ms = kernel_ms_total;
  res->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
  record_multi_schedule(x, mp, E);
  x->sched.n_interval = (int64_t)ivals.size();
  return GTP_OK;
}

// ---------------------------------------------------------------------------
// Streamed batch: traces handed over one at a time; each is packed into event
// words on the device on arrival (through the staging ring), so the caller
// holds only the trace it is adding.  Loads are evaluated when the next trace
// does not fit the event budget, and at the end.
// ---------------------------------------------------------------------------
struct gtp_batch {
  gtp_ctx* x = nullptr;
  int64_t G = 0, N = 0;            // the design the batch began with
  int64_t T = 0, TC = 0, budget = 0;
  double time_scale = 0, clk_period = 0;
  gtp_status st = GTP_OK;          // first error; later calls return it
  double kernel_ms = 0;
  int32_t n_added = 0, n_load = 0;
  int64_t n_event = 0;
  // the load being filled
  DArr<uint32_t> d_words;
  int64_t words_used = 0, load_words = 0;
  hipEvent_t ready = nullptr;      // the load's last upload done
  struct Pend {
    std::vector<int64_t> off;      // the trace's net offsets
    int64_t t_max = 0, C = 0;
    int32_t n_tile = 0;
    gtp_result* res = nullptr;
  };
  std::vector<Pend> pend;
  std::vector<int64_t> work;       // pin events per gate over the load
};

namespace {

gtp_status batch_flush(gtp_batch* b) {
  gtp_ctx* x = b->x;
  if (b->pend.empty()) return GTP_OK;
  if (x->G != b->G || x->N != b->N)
    return x->fail(GTP_ERR_STATE, "gtp_batch: the design changed during the batch");
  hipStream_t s = x->stream;
  const int64_t G = x->G;
  const int32_t nt = (int32_t)b->pend.size();
  std::vector<int64_t> cyc(nt);
  int64_t tmax_max = 0;
  for (int32_t k = 0; k < nt; ++k) {
    cyc[k] = b->pend[k].C;
    tmax_max = std::max(tmax_max, b->pend[k].t_max);
  }
  MultiPlan mp;
  {
    const gtp_status st = multi_setup(x, b->T, b->time_scale, b->clk_period, tmax_max, cyc, std::move(b->work), &mp);
    if (st != GTP_OK) return st;
  }
  if (mp.TC != b->TC) return x->fail(GTP_ERR_STATE, "gtp_batch: the tile length changed during the batch");
  std::vector<int64_t> slot_base(nt + 1, 0);
  for (int32_t k = 0; k < nt; ++k) slot_base[k + 1] = slot_base[k] + (int64_t)GTP_N_WAVE * cyc[k];
  DArr<unsigned long long> d_acc, d_cnt;
  DArr<double> d_cyc;
  HIPCHK(x, d_acc.alloc((size_t)nt * N_PG * G + 1));
  HIPCHK(x, d_cnt.alloc((size_t)nt * 2 * G + 1));
  HIPCHK(x, d_cyc.alloc(std::max<int64_t>(slot_base[nt], 1)));
  HIPCHK(x, launch_zero_u64(d_acc.p, (int64_t)nt * N_PG * G, s));
  HIPCHK(x, launch_zero_u64(d_cnt.p, (int64_t)nt * 2 * G, s));
  std::vector<RunItem> items;
  for (int32_t k = 0; k < nt; ++k) {
    const auto& P = b->pend[k];
    items.push_back(RunItem{P.off.data(), nullptr, nullptr, P.t_max, P.C, P.n_tile, 0, P.n_tile, k});
  }
  items[0].dev_words = b->d_words.p;
  HIPCHK(x, hipStreamWaitEvent(s, b->ready, 0));
  float ms = 0.f;
  {
    const gtp_status st = run_load(x, mp, items, d_acc, d_cnt, d_cyc, slot_base, &ms);
    if (st != GTP_OK) return st;
  }
  for (int32_t k = 0; k < nt; ++k) {
    gtp_result& r = *b->pend[k].res;
    const gtp_status st = finish_slot(x, mp, k, d_acc, d_cnt, d_cyc, slot_base[k], b->pend[k].t_max, b->pend[k].C, r);
    if (st != GTP_OK) return st;
    r.kernel_ms = ms;  // the load's device time
    r.total_ms = ms;
  }
  record_multi_schedule(x, mp, b->n_event);
  b->kernel_ms += ms;
  b->n_load += 1;
  b->pend.clear();
  b->work.assign(G, 0);
  b->words_used = 0;
  b->load_words = 0;
  return GTP_OK;
}

gtp_status batch_add(gtp_batch* b, const gtp_trace* tr, gtp_result* res) {
  gtp_ctx* x = b->x;
  if (x->G != b->G || x->N != b->N)
    return x->fail(GTP_ERR_STATE, "gtp_batch_add: the design changed during the batch");
  const int64_t N = x->N, G = x->G;
  int64_t Tk = 0;
  {
    // (each event is checked as it is packed, below, or by the sliced run)
    const gtp_status st = validate_trace(x, tr, &Tk, true, false);
    if (st != GTP_OK) return st;
  }
  if (Tk != b->T || tr->time_scale != b->time_scale || tr->clk_period != b->clk_period)
    return x->fail(GTP_ERR_UNSUPPORTED, "gtp_batch_add: traces must share time_scale and clk_period");
  const int64_t* off = tr->net_event_offset;
  const int64_t E = off[N];
  const int64_t C = (tr->t_max + b->T - 1) / b->T;
  const int32_t n_tile = (int32_t)((C + b->TC - 1) / b->TC);
  const int64_t words = E + N * ((int64_t)n_tile + 1) + 4 * N;
  if (tr->t_max >= (int64_t)EV_TIME_INF || words > b->budget) {
    // longer than one event word's time span, or larger than one load: by
    // itself, in cycle intervals
    {
      const gtp_status st = batch_flush(b);
      if (st != GTP_OK) return st;
    }
    const gtp_status st = gtp_run_sliced(x, tr, b->budget, res);
    if (st != GTP_OK) return st;
    b->n_event += E;
    b->kernel_ms += res->kernel_ms;
    b->n_load += (int32_t)x->sched.n_interval;
    return GTP_OK;
  }
  if (b->load_words + words > b->budget) {
    const gtp_status st = batch_flush(b);
    if (st != GTP_OK) return st;
  }
  // room for the words: grown by doubling, the words so far copied over
  const int64_t need = b->words_used + E + EV_PAD;
  if ((int64_t)b->d_words.n < need) {
    HIPCHK(x, hipStreamSynchronize(x->stream));
    if (x->aux) HIPCHK(x, hipStreamSynchronize(x->aux));  // the staged copies into the old buffer
    DArr<uint32_t> nw;
    HIPCHK(x, nw.alloc((size_t)std::max<int64_t>(need, std::min<int64_t>(2 * (int64_t)b->d_words.n, b->budget + EV_PAD))));
    if (b->words_used > 0)
      HIPCHK(x, hipMemcpyAsync(nw.p, b->d_words.p, sizeof(uint32_t) * b->words_used, hipMemcpyDeviceToDevice, x->stream));
    HIPCHK(x, hipStreamSynchronize(x->stream));
    std::swap(nw.p, b->d_words.p);
    std::swap(nw.n, b->d_words.n);
  }
  const int64_t* tm = tr->event_time;
  const uint8_t* val = tr->event_value;
  {
    // a trace with events out of order, past t_max or above 3 is refused; its
    // words lie past words_used, where the next trace's go
    std::atomic<int64_t> bad{-1};
    const gtp_status st = staged_upload(x, off, N, b->d_words.p + b->words_used,
                                        [&](int64_t n, uint32_t* d) {
                                          if (!pack_words_checked(d, tm + off[n], val + off[n], off[n + 1] - off[n], 0,
                                                                  -1, tr->t_max)) {
                                            int64_t expect = -1;
                                            bad.compare_exchange_strong(expect, n);
                                          }
                                        },
                                        b->ready);
    if (st != GTP_OK) return st;
    if (bad.load() >= 0)
      return x->fail(GTP_ERR_ARG, "gtp_trace: events of net " + std::to_string(bad.load()) +
                                      " are not strictly increasing in [0, t_max] or carry a value > 3");
  }
  b->n_event += E;
  // the staged copies read the caller's arrays only through the staging
  // buffers, which are filled before staged_upload returns
  b->words_used += E;
  b->load_words += words;
  gtp_batch::Pend P;
  P.off.assign(off, off + N + 1);
  P.t_max = tr->t_max;
  P.C = C;
  P.n_tile = n_tile;
  P.res = res;
  b->pend.push_back(std::move(P));
  parallel_for(G, x->threads, [&](int64_t g) {
    int64_t w = 0;
    for (int64_t q = x->gate_pin_off[g]; q < x->gate_pin_off[g + 1]; ++q) {
      const int32_t net = x->pin_net[q];
      if (net >= 0) w += off[net + 1] - off[net];
    }
    b->work[g] += w;
  });
  return GTP_OK;
}

}  // namespace

GTP_API gtp_status gtp_batch_begin(gtp_ctx* x, double time_scale, double clk_period, gtp_batch** out) {
  if (!x || !out) return GTP_ERR_ARG;
  *out = nullptr;
  if (!x->have_design) return x->fail(GTP_ERR_STATE, "gtp_batch_begin: call gtp_set_design first");
  if (!(time_scale > 0) || !(clk_period > 0))
    return x->fail(GTP_ERR_ARG, "gtp_batch_begin: time_scale and clk_period must be positive");
  HIPCHK(x, hipSetDevice(x->device));
  auto* b = new gtp_batch();
  b->x = x;
  b->G = x->G;
  b->N = x->N;
  b->time_scale = time_scale;
  b->clk_period = clk_period;
  b->T = ticks_per_cycle(clk_period, time_scale);
  if (b->T < 1 || b->T >= (int64_t)EV_TIME_INF) {
    delete b;
    return x->fail(GTP_ERR_UNSUPPORTED, "gtp_batch_begin: ticks per cycle out of range");
  }
  b->TC = multi_tile_cycles(x, b->T);
  size_t free_b = 0, total_b = 0;
  gtp_status st = GTP_OK;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) st = x->fail(GTP_ERR_DEVICE, "gtp_batch_begin: hipMemGetInfo");
  b->budget = x->opt.max_event_num > 0 ? x->opt.max_event_num : (int64_t)((double)free_b * 0.45 / 4.0);
  if (st == GTP_OK) st = staging_init(x);
  if (st == GTP_OK && hipEventCreateWithFlags(&b->ready, hipEventDisableTiming) != hipSuccess)
    st = x->fail(GTP_ERR_DEVICE, "gtp_batch_begin: hipEventCreate");
  if (st != GTP_OK) {
    if (b->ready) (void)hipEventDestroy(b->ready);
    delete b;
    return st;
  }
  b->work.assign(x->G, 0);
  *out = b;
  return GTP_OK;
}

GTP_API gtp_status gtp_batch_add(gtp_batch* b, const gtp_trace* trace, gtp_result* result) {
  if (!b) return GTP_ERR_ARG;
  if (b->st != GTP_OK) return b->st;
  if (!trace || !result) return b->st = b->x->fail(GTP_ERR_ARG, "gtp_batch_add: no trace or result");
  gtp_ctx* x = b->x;
  if (hipSetDevice(x->device) != hipSuccess) return b->st = x->fail(GTP_ERR_DEVICE, "gtp_batch_add: hipSetDevice");
  b->st = batch_add(b, trace, result);
  if (b->st == GTP_OK) b->n_added += 1;
  return b->st;
}

GTP_API gtp_status gtp_batch_end(gtp_batch* b, double* kernel_ms) {
  if (!b) return GTP_ERR_ARG;
  gtp_ctx* x = b->x;
  gtp_status st = b->st;
  if (st == GTP_OK) st = hipSetDevice(x->device) == hipSuccess ? batch_flush(b)
                                                               : x->fail(GTP_ERR_DEVICE, "gtp_batch_end: hipSetDevice");
  if (st == GTP_OK) {
    if (kernel_ms) *kernel_ms = b->kernel_ms;
    x->sched.n_interval = b->n_load;
    x->sched.n_event = b->n_event;
  }
  (void)hipStreamSynchronize(x->stream);
  if (b->ready) (void)hipEventDestroy(b->ready);
  delete b;
  return st;
}

GTP_API gtp_status gtp_get_schedule(const gtp_ctx* x, gtp_schedule_info* info) {
  if (!x || !info) return GTP_ERR_ARG;
  *info = x->sched;
  return GTP_OK;
}

// power/cuda/Gate.cuh:354-376 setNCyclePerThreadByEventCount, per gate.
GTP_API gtp_status gtp_reference_cycles_per_thread(int64_t n, const int64_t* n_event, int64_t n_cycle,
                                                   int64_t e_target, int64_t parallelism_floor,
                                                   int64_t* out) {
  if (n < 0 || (n > 0 && (!n_event || !out)) || n_cycle <= 0 || e_target <= 0 || parallelism_floor <= 0)
    return GTP_ERR_ARG;
  int64_t n_cycle_target = (n_cycle + parallelism_floor - 1) / parallelism_floor;
  n_cycle_target = std::max<int64_t>(n_cycle_target, 1);
  const double rho_floor = (double)e_target / (double)n_cycle_target;
  for (int64_t i = 0; i < n; ++i) {
    if (n_event[i] < 0) return GTP_ERR_ARG;
    const double rho = (double)n_event[i] / (double)n_cycle;
    int64_t v = (int64_t)std::llround((double)e_target / (rho + rho_floor));
    out[i] = std::max<int64_t>(1, std::min(v, n_cycle));
  }
  return GTP_OK;
}

// ---------------------------------------------------------------------------
// Writers (power/cuda/CudaPower.cu:3134-3165, include/sta/PowerClass.hh:123-149)
// ---------------------------------------------------------------------------
GTP_API gtp_status gtp_write_cell_power(const char* path, int64_t n_gate, const double* gp,
                                        const char* gate_names, const int32_t* gate_cell,
                                        const char* cell_names) {
  if (!path || n_gate < 0 || (n_gate > 0 && !gp)) return GTP_ERR_ARG;
  std::vector<const char*> cells;
  if (cell_names)
    for (const char* c = cell_names; *c; c += std::strlen(c) + 1) cells.push_back(c);
  std::ofstream f(path, std::ios::out);
  if (!f) return GTP_ERR_ARG;
  f << "cell\tInternal_Power\tGlitch_Internal_Power\tSwitching_Power\tGlitch_Switching_Power\tLeakage_Power\tTotal_Power\tStandard_Cell\n";
  const char* name = gate_names;
  for (int64_t g = 0; g < n_gate; ++g) {
    const float v[5] = {(float)gp[0 * n_gate + g], (float)gp[1 * n_gate + g], (float)gp[2 * n_gate + g],
                        (float)gp[3 * n_gate + g], (float)gp[4 * n_gate + g]};
    if (name) {
      f << name;
      name += std::strlen(name) + 1;
    } else {
      f << "g" << g;
    }
    f << '\t' << v[0] << '\t' << v[1] << '\t' << v[2] << '\t' << v[3] << '\t' << v[4] << '\t'
      << v[0] + v[1] + v[2] + v[3] + v[4] << '\t';
    if (gate_cell && gate_cell[g] >= 0 && (size_t)gate_cell[g] < cells.size()) f << cells[gate_cell[g]];
    f << '\n';
  }
  return f ? GTP_OK : GTP_ERR_ARG;
}

GTP_API gtp_status gtp_write_per_cycle(const char* path, const char* top_name, int64_t n_cycle,
                                       const double* cp) {
  if (!path || n_cycle < 0 || (n_cycle > 0 && !cp)) return GTP_ERR_ARG;
  std::ofstream f(path, std::ios::out);
  if (!f) return GTP_ERR_ARG;
  static const char* titles[5] = {"Per cycle leakage power: ", "Per cycle internal power: ",
                                  "Per cycle glitch internal power: ", "Per cycle switching power: ",
                                  "Per cycle glitch switching power: "};
  for (int w = 0; w < 5; ++w) {
    f << titles[w] << std::endl;
    f << (top_name ? top_name : "") << '\t';
    for (int64_t c = 0; c < n_cycle; ++c) f << (float)cp[(int64_t)w * n_cycle + c] << '\t';
    f << std::endl;
  }
  f << "End of per cycle power" << std::endl;
  return f ? GTP_OK : GTP_ERR_ARG;
}
