// gfx950 kernels of the time-based power engine.
//
//   pack_events_kernel     int64 time + u8 value -> one 32-bit event word (device event store)
//   tile_index_kernel      per (net, tile): first event at or after the tile start
//   tile_order_kernel      per (range, tile): the range's gates ordered by their events in
//                          the tile, dealt to the merge's waves (lane balance)
//   energy_kernel          Liberty table interpolation per gate energy slot, tables staged in LDS
//   power_tile_kernel<P>   fused per-gate event merge over one cycle tile: toggle and glitch
//                          filtering, internal / switching / leakage evaluation, per-cycle
//                          binning in LDS and per-gate sums, all int64 fixed point
//   cycle_reduce_kernel    per-cycle sum over gate ranges -> FP64 waveforms
//   gate_reduce_kernel     per-gate sum over tiles (int64) -> FP64 per-gate power
//   glitch_annotate_kernel clock-cycle glitch flags, once per net event (glitch_annotate_tiles_kernel:
//                          the events of listed tiles)
//   power_cell_kernel<NP, KC>  the fused kernel specialised per single-output cell type without
//                          `when`; KC >= 0: flip-flops whose clock edges the workgroup takes once
//   total_*_kernel         FP64 design and class totals
//
// Reference counterparts: kernel1AllPowerCalculationTimeRangePartitionedByCycleUnit
// (power/cuda/CudaPower.cu:317-584), kernel2MergeAllPowerResult (:586-606), the
// CPU algorithm they restate (power/TimeBasedPower.cc:703-901), and the device
// LUTs (util/cuda/OneDimensionalLUT.cuh:62-86, TwoDimensionalLUT.cuh:113-164).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gtp_internal.hpp"

namespace gtp {

__device__ __forceinline__ uint32_t ev_time(uint32_t w) { return w >> EV_SHIFT; }
__device__ __forceinline__ int ev_val(uint32_t w) { return (int)(w & 3u); }

// power/PowerUtils.cc:246-256 getNToggle and :213-236 getNRise, times two.
__device__ __forceinline__ int toggle_x2(int p, int c) {
  return p == c ? 0 : ((p == 2 || c == 2) ? 1 : 2);
}
__device__ __forceinline__ int rise_x2(int p, int c) {
  if (c == 1) return p == 0 ? 2 : (p == 2 ? 1 : 0);
  return (c == 2 && p == 0) ? 1 : 0;
}
// The same two functions as 2-bit entries of a constant, entry 4 p + c
// (checked against them for all 16 pairs by static_asserts below).
constexpr uint32_t TOGGLE_X2_TABLE = 0x1a459298u;
constexpr uint32_t RISE_X2_TABLE = 0x00040018u;
__host__ __device__ constexpr int table_x2(uint32_t tab, int p, int c) { return (int)((tab >> (2 * (4 * p + c))) & 3u); }
constexpr bool tables_match() {
  for (int p = 0; p < 4; ++p)
    for (int c = 0; c < 4; ++c) {
      const int t = p == c ? 0 : ((p == 2 || c == 2) ? 1 : 2);
      const int r = c == 1 ? (p == 0 ? 2 : (p == 2 ? 1 : 0)) : ((c == 2 && p == 0) ? 1 : 0);
      if (table_x2(TOGGLE_X2_TABLE, p, c) != t || table_x2(RISE_X2_TABLE, p, c) != r) return false;
    }
  return true;
}
static_assert(tables_match(), "toggle / rise tables");

// floor(t / T) for t < 2^29: a float estimate, then integer correction.
__device__ __forceinline__ uint32_t cycle_of(uint32_t t, uint32_t T, float invT) {
  uint32_t q = (uint32_t)((float)t * invT);
  int32_t r = (int32_t)(t - q * T);
  while (r < 0) { --q; r += (int32_t)T; }
  while (r >= (int32_t)T) { ++q; r -= (int32_t)T; }
  return q;
}

// floor(t / T) by the exact multiply-shift of PowerArgs (t < 2^29).
__device__ __forceinline__ uint32_t cycle_mul(uint32_t t, uint32_t m, int sh) {
  return (uint32_t)(((uint64_t)m * (uint64_t)t) >> sh);
}

// Compensated accumulation (TwoSum): hi + lo carries ~48 significant bits.
__device__ __forceinline__ void ff_add(float& hi, float& lo, float x) {
  float s = hi + x;
  float bp = s - hi;
  float err = (hi - (s - bp)) + (x - bp);
  hi = s;
  lo += err;
}
// hi + lo += a * b with the product's rounding error kept (fma).
__device__ __forceinline__ void ff_add_prod(float& hi, float& lo, float a, float b) {
  float p = a * b;
  float e = fmaf(a, b, -p);
  ff_add(hi, lo, p);
  lo += e;
}

__device__ __forceinline__ long long quant(float x, float inv_q) {
  return __float2ll_rn(x * inv_q);
}
// One leakage term in leakage quanta: |x * inv_q| <= 2^30 by the choice of
// quantum, so a 32-bit conversion is exact.
__device__ __forceinline__ long long qterm(float x, float inv_q) {
  return (long long)(int)rintf(x * inv_q);
}
// Energy quanta are 2^ENERGY_FINE_BITS times finer than the coarse scale the
// kernels take (PowerArgs.inv_q_energy: one term <= 2^20 coarse units, so one
// term <= 2^36 quanta).  rint(x * inv_q * 2^16) from two exact 32-bit
// conversions: xs - rint(xs) is exact in FP32 and |.| <= 1/2, so its scaled
// rounding is the low part.
__device__ __forceinline__ long long qenergy(float x, float inv_q) {
  // x * inv_q * 2^16 is exact in FP64 (|.| <= 2^36); adding 1.5 * 2^52 rounds
  // it to an integer held in the low mantissa bits (round to nearest even)
  const double y = (double)x * (double)(inv_q * (float)(1 << ENERGY_FINE_BITS));
  const double m = y + 6755399441055744.0;  // 1.5 * 2^52
  return __double_as_longlong(m) - 0x4338000000000000ll;
}

// ---------------------------------------------------------------------------
// Event store
// ---------------------------------------------------------------------------
__global__ void pack_events_kernel(const int64_t* __restrict__ t, const uint8_t* __restrict__ v,
                                   uint32_t* __restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = ((uint32_t)t[i] << EV_SHIFT) | (uint32_t)(v[i] & 3u);
}

// out[n][k] = first event of net n at or after the start of tile k, for
// k < n_tile, and the net's event count for k == n_tile; with a tile list,
// out[n][l] for the tile tile_list[l] (no end column).
__global__ void tile_index_kernel(const int64_t* __restrict__ net_off, const uint32_t* __restrict__ ev,
                                  int64_t n_net, int32_t n_tile, int64_t tile_ticks,
                                  const int32_t* __restrict__ tile_list, int32_t n_list,
                                  int32_t* __restrict__ out) {
  const int64_t per = tile_list ? (int64_t)n_list : (int64_t)n_tile + 1;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n_net * per; i += stride) {
    const int64_t n = i / per;
    const int32_t k = tile_list ? tile_list[i - n * per] : (int32_t)(i - n * per);
    const int64_t b = net_off[n];
    const int32_t cnt = (int32_t)(net_off[n + 1] - b);
    int32_t lo = 0, hi = cnt;
    if (k == n_tile) {
      lo = cnt;
    } else if (cnt > 0) {
      const int64_t tk = (int64_t)k * tile_ticks;
      const uint32_t key = tk >= (int64_t)EV_TIME_INF ? EV_INF : (uint32_t)(tk << EV_SHIFT);
      // lower bound of key: bracket it from an interpolated guess (events are
      // spread over [first, last] time) with doubling steps, then bisect
      const uint32_t t0 = ev_time(ev[b]), t1 = ev_time(ev[b + cnt - 1]);
      int32_t g = 0;
      if ((uint32_t)tk > t0 && t1 > t0)
        g = (int32_t)min((double)(cnt - 1), (double)(cnt - 1) * ((double)((uint32_t)tk - t0) / (double)(t1 - t0)));
      if (ev[b + g] < key) {
        int32_t step = 1;
        lo = g + 1;
        hi = g + 1;
        while (hi < cnt && ev[b + hi] < key) {
          lo = hi + 1;
          hi = min(cnt, hi + step);
          step <<= 1;
        }
      } else {
        int32_t step = 1;
        hi = g;
        lo = g;
        while (lo > 0 && ev[b + lo - 1] >= key) {
          hi = lo - 1;
          lo = max(0, lo - step);
          step <<= 1;
        }
      }
      // now every index < lo is < key and every index >= hi is >= key
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (ev[b + mid] < key) lo = mid + 1; else hi = mid;
      }
    }
    out[i] = lo;
  }
}

// ---------------------------------------------------------------------------
// Per-tile lane order.  A wave of the merge runs until its busiest lane's gate
// is done, so the positions of each range are ordered per tile by the gate's
// pin events in that tile (the tile index difference summed over its pins):
// one workgroup per (range, tile) item sorts the range in LDS (bitonic, on
// key << 32 | ~position, descending).  Groups of 64 consecutive ranks form
// the wave passes; they are dealt to the merge's waves in a snake (pass 0
// waves 0..n-1, pass 1 waves n-1..0, ...) so that the waves of a workgroup,
// which meet at its barriers, carry similar totals.  Only the order of the
// lanes changes: every partial still goes to its schedule position.
// ---------------------------------------------------------------------------
constexpr int TO_MAX = TILE_ORDER_MAX_RANGE;  // longest range sorted

__global__ __launch_bounds__(256) void tile_order_kernel(const int32_t* __restrict__ order, int64_t n_sched,
                                                         int32_t range_gates, int32_t lanes,
                                                         const int64_t* __restrict__ gate_pin_off,
                                                         const int32_t* __restrict__ pin_net,
                                                         const int32_t* __restrict__ tile_idx, int32_t n_tile,
                                                         int32_t n_wt, const int32_t* __restrict__ wt_tile,
                                                         const int32_t* __restrict__ wt_trace,
                                                         const TraceDesc* __restrict__ traces,
                                                         int32_t* __restrict__ perm) {
  __shared__ unsigned long long key[TO_MAX];
  const int64_t n_range = n_sched / range_gates;
  const int64_t n_item = n_range * n_wt;
  int n2 = 1;
  while (n2 < range_gates) n2 <<= 1;
  for (int64_t item = blockIdx.x; item < n_item; item += gridDim.x) {
    const int64_t range = item / n_wt;
    const int wt = (int)(item - range * n_wt);
    // the work tile's tile and tile index (batched runs: of its trace)
    int tile = wt;
    int64_t stride = (int64_t)n_tile + 1;
    const int32_t* tidx = tile_idx;
    if (wt_trace) {
      const TraceDesc td = traces[wt_trace[wt]];
      tile = wt_tile[wt];
      stride = (int64_t)td.n_tile + 1;
      tidx = tile_idx + td.tile_idx_base;
    }
    const int64_t p0 = range * range_gates;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
      long long ev = -1;  // padding: after every gate
      if (i < range_gates) {
        const int g = order[p0 + i];
        if (g >= 0) {
          ev = 0;
          for (int64_t q = gate_pin_off[g]; q < gate_pin_off[g + 1]; ++q) {
            const int32_t n = pin_net[q];
            if (n >= 0) ev += tidx[(int64_t)n * stride + tile + 1] - tidx[(int64_t)n * stride + tile];
          }
        }
      } else {
        ev = -2;  // beyond the range
      }
      // key then the position reversed, so that equal counts keep position order
      key[i] = ((unsigned long long)(ev + 2) << 32) | (unsigned long long)(0xFFFFFFFFu - (uint32_t)i);
    }
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < n2; i += blockDim.x) {
          const int l = i ^ j;
          if (l > i) {
            const unsigned long long a = key[i], b = key[l];
            // descending overall: blocks with (i & k) == 0 descend
            const bool desc = (i & k) == 0;
            if (desc ? (a < b) : (a > b)) {
              key[i] = b;
              key[l] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    // rank r -> the position a lane of the merge takes as its r-th gate
    const bool snake = lanes >= 64 && lanes % 64 == 0 && range_gates % lanes == 0;
    const int nw = lanes / 64;
    for (int r = threadIdx.x; r < range_gates; r += blockDim.x) {
      int pos = r;
      if (snake) {
        const int k = r >> 6, pass = k / nw;
        int wv = k - pass * nw;
        if (pass & 1) wv = nw - 1 - wv;
        pos = pass * lanes + wv * 64 + (r & 63);
      }
      perm[(int64_t)wt * n_sched + p0 + pos] =
          (int32_t)(p0 + (int64_t)(0xFFFFFFFFu - (uint32_t)(key[r] & 0xFFFFFFFFull)));
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Liberty table interpolation (FP32), tables staged in LDS
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_index_f(const float* ax, int n, float x) {
  // liberty/TableModel.cc:1465-1488 findValueIndex
  if (n <= 1 || x <= ax[0]) return 0;
  if (x >= ax[n - 1]) return n - 2;
  int lo = -1, hi = n;
  while (hi - lo > 1) {
    int mid = (hi + lo) >> 1;
    if (x >= ax[mid]) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ float interp_f(const float* D, const TableDev& tb, float slew, float cap) {
  // liberty/TableModel.cc:858-1062 Table1/Table2::findValue
  const float x1 = tb.var1 == 0 ? slew : cap;
  const float x2 = tb.var2 == 0 ? slew : cap;
  const float* a1 = D + tb.axis1;
  const float* v = D + tb.values;
  if (tb.n2 == 0) {
    if (tb.n1 == 1) return v[0];
    int i = find_index_f(a1, tb.n1, x1);
    float d = (x1 - a1[i]) / (a1[i + 1] - a1[i]);
    return (1.f - d) * v[i] + d * v[i + 1];
  }
  const float* a2 = D + tb.axis2;
  const int n1 = tb.n1, n2 = tb.n2;
  if (n1 == 1 && n2 == 1) return v[0];
  if (n1 == 1) {
    int j = find_index_f(a2, n2, x2);
    float d = (x2 - a2[j]) / (a2[j + 1] - a2[j]);
    return (1.f - d) * v[j] + d * v[j + 1];
  }
  if (n2 == 1) {
    int i = find_index_f(a1, n1, x1);
    float d = (x1 - a1[i]) / (a1[i + 1] - a1[i]);
    return (1.f - d) * v[i] + d * v[i + 1];
  }
  int i = find_index_f(a1, n1, x1), j = find_index_f(a2, n2, x2);
  float d1 = (x1 - a1[i]) / (a1[i + 1] - a1[i]);
  float d2 = (x2 - a2[j]) / (a2[j + 1] - a2[j]);
  return (1.f - d1) * (1.f - d2) * v[i * n2 + j] + d1 * (1.f - d2) * v[(i + 1) * n2 + j] +
         d1 * d2 * v[(i + 1) * n2 + j + 1] + (1.f - d1) * d2 * v[i * n2 + j + 1];
}



__global__ void energy_kernel(EnergyArgs a) {
  extern __shared__ float s_tab_lds[];
  const bool staged = a.n_table_data <= TABLE_LDS_MAX;
  if (staged)
    for (int i = threadIdx.x; i < a.n_table_data; i += blockDim.x) s_tab_lds[i] = a.table_data[i];
  __syncthreads();
  const float* s_tab = staged ? s_tab_lds : a.table_data;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float local_max = 0.f;
  if (g < a.n_gate) {
    const int c = a.gate_cell[g];
    const CellDev cd = a.cells[c];
    const int64_t pb = a.gate_pin_off[g];
    const int64_t eb = a.gate_energy_off[g];
    const SlotDesc* ds = a.descs + a.cell_slot_desc_off[c];
    for (int s = 0; s < cd.n_slot; ++s) {
      const SlotDesc d = ds[s];
      float e = 0.f;
      if (d.list_cnt > 0) {
        // mean over (related input, rf, table) entries with a finite slew
        // (power/TimeBasedPower.cc:511-555)
        float tot = 0.f;
        int n = 0;
        for (int k = 0; k < d.list_cnt; ++k) {
          const ListEntry le = a.list[d.list_off + k];
          const float slew = a.pin_slew[2 * (pb + le.slew_pin) + le.slew_rf];
          if (!isfinite(slew)) continue;
          const float cap = le.cap_pin >= 0 ? a.pin_cap[pb + le.cap_pin] : 0.f;
          if (le.table >= 0) tot += interp_f(s_tab, a.tables[le.table], slew, cap);
          ++n;
        }
        e = n ? tot / (float)n : 0.f;
      } else if (d.table >= 0) {
        const float slew = a.pin_slew[2 * (pb + d.slew_pin) + d.slew_rf];
        if (isfinite(slew)) {  // delayInf(slew) -> 0 (TimeBasedPower.cc:367-370, :488)
          const float cap = d.cap_pin >= 0 ? a.pin_cap[pb + d.cap_pin] : 0.f;
          e = interp_f(s_tab, a.tables[d.table], slew, cap);
        }
      }
      a.energy[eb + s] = e;
      local_max = fmaxf(local_max, fabsf(e));
    }
  }
  // wave max, then one atomic per wave (positive floats order like their bits)
  for (int off = 32; off > 0; off >>= 1) local_max = fmaxf(local_max, __shfl_xor(local_max, off));
  if ((threadIdx.x & 63) == 0) atomicMax(a.max_abs_bits, __float_as_uint(local_max));
}

// ---------------------------------------------------------------------------
// Work items of the power kernels: (work tile, gate range).  A work tile is a
// tile of the loaded trace, a listed tile (subset runs), or a tile of one of
// several traces (batched runs).
// ---------------------------------------------------------------------------
struct WorkTile {
  int tile;                // tile within its trace (its load: relative to the load's origin)
  bool trace_start;        // the tile starts the trace (time 0: a time-0 event is the initial state)
  int tile_col;            // column of the tile index
  int32_t n_tile;          // tiles of its trace
  int64_t tile_stride;     // tile index row length
  int64_t t_max, n_cycle;  // of its trace
  int64_t net_base;        // first net of its trace in the event store
  const int32_t* tidx;     // tile index of its trace
};

__device__ __forceinline__ WorkTile resolve_work_tile(const PowerArgs& a, int wt) {
  WorkTile k;
  if (a.wt_trace) {
    const TraceDesc td = a.traces[a.wt_trace[wt]];
    k.tile = a.tile_list[wt];
    k.trace_start = k.tile + td.tile_base == 0;
    k.tile_col = k.tile;
    k.n_tile = td.n_tile;
    k.tile_stride = (int64_t)td.n_tile + 1;
    k.t_max = td.t_max;
    k.n_cycle = td.n_cycle;
    k.net_base = td.net_base;
    k.tidx = a.tile_idx + td.tile_idx_base;
  } else {
    k.tile = a.tile_list ? a.tile_list[wt] : wt;
    k.trace_start = k.tile == 0;
    k.tile_col = a.tile_list ? wt : k.tile;
    k.n_tile = a.n_tile;
    k.tile_stride = a.tile_list ? (int64_t)a.n_work_tiles : (int64_t)a.n_tile + 1;
    k.t_max = a.t_max;
    k.n_cycle = a.n_cycle;
    k.net_base = 0;
    k.tidx = a.tile_idx;
  }
  return k;
}

// The next work item of this workgroup; false when there is none.  Called by
// every thread of the workgroup (the result is uniform).  Without a work
// queue a workgroup takes one item: blockIdx remapped so that each XCD
// (blockIdx % 8) takes a contiguous block of tile-major items and the gates
// of one tile share that XCD's L2.  With a queue (wq[8], zeroed before the
// launch) the workgroups are persistent: each takes items of its XCD's block
// from that XCD's counter until the block is exhausted.
__device__ __forceinline__ bool next_work_item(const PowerArgs& a, unsigned int* wq, int64_t n_work, int64_t it,
                                               int64_t& w, long long* s_item) {
  const int64_t per = (n_work + 7) / 8;
  const int xcd = (int)(blockIdx.x % 8);
  if (!wq) {
    if (it > 0) return false;
    w = a.xcd_remap ? (int64_t)xcd * per + blockIdx.x / 8 : (int64_t)blockIdx.x;
    return w < n_work;
  }
  __syncthreads();  // every thread has read the previous item
  if (threadIdx.x == 0) *s_item = (long long)atomicAdd(&wq[xcd], 1u);
  __syncthreads();
  // the item is the same for the whole workgroup: keep it (and what is derived
  // from it) in scalar registers
  const int64_t k = (int64_t)__builtin_amdgcn_readfirstlane((int)*s_item);
  w = (int64_t)xcd * per + k;
  return k < per && w < n_work;
}

// Leakage bins of one work item, by one thread: on entry ld[ci] = sum of the
// leakage steps dL (quanta) taken in tile cycle ci and lx[ci] = sum of dL * (t -
// cycle start); on exit ld[ci] = S_c, the leakage of the gates' states at the
// cycle start (base + the steps of earlier cycles of the tile), and lx[ci] =
// sum of dL * (cycle end - t), the in-cycle part of the cycle's leakage
// integral (quanta x ticks; the cycle end clamped to t_max).  The cycle's
// integral is then S_c * len + lx[ci], exactly, in integers.
__device__ void leak_flush(long long* ld, long long* lx, int TC, long long base, int64_t tc0, int64_t T, int64_t t_max,
                           int64_t n_cycle) {
  long long run = base;
  for (int ci = 0; ci < TC; ++ci) {
    const int64_t c = tc0 + ci;
    const int64_t len = c < n_cycle ? min((c + 1) * T, t_max) - c * T : 0;
    const long long d = ld[ci];
    lx[ci] = d * (long long)len - lx[ci];
    ld[ci] = run;
    run += d;
  }
}

// ---------------------------------------------------------------------------
// The fused power kernel
// ---------------------------------------------------------------------------
template <int MAXP>
__global__ __launch_bounds__(256) void power_tile_kernel(PowerArgs a, int32_t range_begin,
                                                          int32_t range_count, unsigned int* wq) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int TC = a.tile_cycles;
  long long* bins = reinterpret_cast<long long*>(smem_raw);
  // leakage tables (all cells) in LDS when they are small, else read in place
  const int n_ll = a.leak_lds ? a.n_leak : 0;
  long long* lds_leakq = bins + (size_t)N_BIN * TC;
  float* lds_leak = reinterpret_cast<float*>(lds_leakq + n_ll);
  // cell records in LDS when there are few cell types, else read in place
  const int n_lc = a.n_cell <= MAX_CELLS_LDS ? a.n_cell : 0;
  CellDev* lds_cells = reinterpret_cast<CellDev*>(lds_leak + n_ll + (n_ll & 1));
  __shared__ long long s_base;
  __shared__ long long s_item;

  for (int i = threadIdx.x; i < n_ll; i += blockDim.x) {
    lds_leak[i] = a.leak[i];
    lds_leakq[i] = a.leak_q[i];
  }
  const long long* s_leakq = a.leak_lds ? lds_leakq : a.leak_q;
  const float* s_leak = a.leak_lds ? lds_leak : a.leak;
  for (int i = threadIdx.x; i < n_lc; i += blockDim.x) lds_cells[i] = a.cells[i];
  const CellDev* s_cells = n_lc ? lds_cells : a.cells;

  const int64_t n_work = (int64_t)a.n_work_tiles * range_count;
  int64_t w = 0;
  for (int64_t it = 0; next_work_item(a, wq, n_work, it, w, &s_item); ++it) {
  const int wt = (int)(w / range_count);
  const int range = range_begin + (int)(w - (int64_t)wt * range_count);
  const WorkTile wk = resolve_work_tile(a, wt);
  const int tile = wk.tile;

  for (int i = threadIdx.x; i < N_BIN * TC; i += blockDim.x) bins[i] = 0;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();

  const uint32_t T = (uint32_t)a.T;
  const float invT = a.inv_T;
  const float inv_qf = a.inv_q_energy * (float)(1 << ENERGY_FINE_BITS);  // energy quanta (fine)
  const int64_t tc0 = (int64_t)tile * TC;
  const int tcn = (int)min((int64_t)TC, wk.n_cycle - tc0);
  const bool last_tile = tile == wk.n_tile - 1;
  const uint32_t tstart = (uint32_t)(tc0 * (int64_t)T);
  const uint32_t tend = last_tile ? EV_TIME_INF : (uint32_t)((tc0 + TC) * (int64_t)T);
  const uint32_t lend = last_tile ? (uint32_t)wk.t_max : tend;
  const uint32_t t_max = (uint32_t)wk.t_max;

  long long base_acc = 0;
  const int64_t gi_begin = (int64_t)range * a.range_gates;
  const int64_t gi_end = min(a.n_sched, gi_begin + a.range_gates);
  for (int64_t gi = gi_begin + threadIdx.x; gi < gi_end; gi += blockDim.x) {
    const int g = a.order[gi];
    if (g < 0) continue;  // padding to a whole range
    const CellDev cd = s_cells[a.gate_cell[g]];
    const int n = cd.n_pin, nin = cd.n_in;
    const int64_t pb = a.gate_pin_off[g];
    const int64_t eb = a.gate_energy_off[g];
    const int64_t ab = a.gate_arc_off[g];
    const int sbits = cd.state_bits;

    const uint32_t* evp[MAXP];
    int32_t cnt[MAXP], idx[MAXP];
    uint32_t w0[MAXP], w1[MAXP], w2[MAXP];
    uint32_t c0[MAXP], c1[MAXP];
    uint32_t lt[MAXP], lc[MAXP];
    int cv[MAXP], jl[MAXP], rfl[MAXP];
    bool lF[MAXP];
    float gk[MAXP], csw[MAXP];
    uint32_t sb = 0;
#pragma unroll
    for (int p = 0; p < MAXP; ++p) {
      evp[p] = a.ev;
      cnt[p] = 0;
      idx[p] = 0;
      w0[p] = w1[p] = w2[p] = EV_INF;
      c0[p] = c1[p] = 0;
      lt[p] = 0;
      lc[p] = 0xFFFFFFFFu;
      cv[p] = 2;
      jl[p] = -1;
      rfl[p] = 1;
      lF[p] = false;
      gk[p] = 0.f;
      csw[p] = 0.f;
      if (p < n) {
        const int32_t net = a.pin_net[pb + p];
        gk[p] = a.pin_gk[pb + p] * a.time_scale;
        csw[p] = a.pin_csw[pb + p];
        int64_t nb = 0;
        if (net >= 0) {
          nb = a.net_off[net + wk.net_base];
          cnt[p] = (int32_t)(a.net_off[net + wk.net_base + 1] - nb);
        }
        if (cnt[p] == 0) {
          cv[p] = a.pin_default[pb + p];
        } else {
          evp[p] = a.ev + nb;
          const int32_t i0 = wk.tidx[(int64_t)net * wk.tile_stride + wk.tile_col];
          const bool isclk = (cd.clock_mask >> p) & 1u;
          if (i0 == 0) {
            const uint32_t e0 = evp[p][0];
            if (wk.trace_start && ev_time(e0) == 0) {
              // the reference consumes a time-0 event as the initial state
              // (power/TimeBasedPower.cc:782-787)
              cv[p] = ev_val(e0);
              idx[p] = 1;
              lt[p] = 0;
              lc[p] = 0;
            } else {
              cv[p] = 2;  // unknown initial state
              idx[p] = 0;
              lt[p] = ev_time(e0);
              lc[p] = cycle_of(lt[p], T, invT);
            }
          } else {
            const uint32_t el = evp[p][i0 - 1];
            cv[p] = ev_val(el);
            lt[p] = ev_time(el);
            lc[p] = cycle_of(lt[p], T, invT);
            idx[p] = i0;
            if (i0 - 1 >= 1) {
              rfl[p] = rise_x2(ev_val(evp[p][i0 - 2]), cv[p]) > 0 ? 0 : 1;
              // position of event i0-1 in its cycle group; it is the group's
              // last event because event i0 lies in a later cycle
              int j = 0;
              const uint32_t cstart = lc[p] * T;
              for (int e = i0 - 2; e >= 1 && ev_time(evp[p][e]) >= cstart; --e) ++j;
              jl[p] = j;
              lF[p] = !isclk && (j & 1);
            }
          }
          const int32_t i = idx[p];
          w0[p] = i < cnt[p] ? evp[p][i] : EV_INF;
          w1[p] = i + 1 < cnt[p] ? evp[p][i + 1] : EV_INF;
          w2[p] = i + 2 < cnt[p] ? evp[p][i + 2] : EV_INF;
          if (w0[p] != EV_INF) c0[p] = cycle_of(ev_time(w0[p]), T, invT);
          if (w1[p] != EV_INF) c1[p] = cycle_of(ev_time(w1[p]), T, invT);
        }
      }
      if (p < n && cv[p] == 1) sb |= 1u << p;
    }
    const int loff = cd.leak_off;
    base_acc += s_leakq[loff + sb] >> 32;

    float acc_hi[N_PG], acc_lo[N_PG];
#pragma unroll
    for (int k = 0; k < N_PG; ++k) acc_hi[k] = acc_lo[k] = 0.f;
    uint32_t tog = 0, glc = 0;
    int pc = -1;  // pending cycle (tile-relative), -1 none
    float p_int = 0.f, p_gint = 0.f, p_sw = 0.f, p_gsw = 0.f;
    long long p_ld = 0, p_lx = 0;
    uint32_t prev_t = tstart;

    auto flush_pending = [&]() {
      if (pc >= 0 && pc < tcn) {
        if (p_int != 0.f) atomicAdd((unsigned long long*)&bins[BIN_INT * TC + pc], (unsigned long long)quant(p_int, inv_qf));
        if (p_gint != 0.f) atomicAdd((unsigned long long*)&bins[BIN_GINT * TC + pc], (unsigned long long)quant(p_gint, inv_qf));
        if (p_sw != 0.f) atomicAdd((unsigned long long*)&bins[BIN_SW * TC + pc], (unsigned long long)quant(p_sw, inv_qf));
        if (p_gsw != 0.f) atomicAdd((unsigned long long*)&bins[BIN_GSW * TC + pc], (unsigned long long)quant(p_gsw, inv_qf));
        if (p_ld != 0) atomicAdd((unsigned long long*)&bins[BIN_LD * TC + pc], (unsigned long long)p_ld);
        if (p_lx != 0) atomicAdd((unsigned long long*)&bins[BIN_LX * TC + pc], (unsigned long long)p_lx);
      }
      p_int = p_gint = p_sw = p_gsw = 0.f;
      p_ld = p_lx = 0;
    };

    for (;;) {
      uint32_t wm = w0[0];
#pragma unroll
      for (int p = 1; p < MAXP; ++p) wm = min(wm, w0[p]);
      const uint32_t t = ev_time(wm);
      if (t >= tend) break;
      uint32_t c = 0;
      bool trig[MAXP];
      uint32_t sa = sb;
#pragma unroll
      for (int p = 0; p < MAXP; ++p) {
        trig[p] = ev_time(w0[p]) == t;
        if (trig[p]) {
          c = c0[p];
          sa = (sa & ~(1u << p)) | ((ev_val(w0[p]) == 1 ? 1u : 0u) << p);
        }
      }
      const int cr = (int)((int64_t)c - tc0);
      if (cr != pc) {
        flush_pending();
        pc = cr;
      }
      // leakage of the previous state over [prev_t, t) (TimeBasedPower.cc:831-843)
      const float Lb = s_leak[loff + sb];
      ff_add_prod(acc_hi[PG_LEAK], acc_lo[PG_LEAK], Lb, (float)(t - prev_t));
      if (sa != sb) {
        const long long dq = (s_leakq[loff + sa] >> 32) - (s_leakq[loff + sb] >> 32);
        p_ld += dq;
        p_lx += dq * (long long)(t - c * T);
      }
      // internal and switching energy of every pin changing at t (:845-883)
      float ri = 0.f, rg = 0.f, rs = 0.f, gs = 0.f;
      bool Fi[MAXP];
      int ji[MAXP];
      // fully unrolled up to 8 pins; the 16-pin variant steps its pins in a
      // loop (the compiler cannot unroll it fully, and its partial unrolling
      // of this loop gave wrong glitch flags against the oracle)
      constexpr int UNROLL_PINS = MAXP <= MAXP_WIDE ? MAXP : 1;
#pragma unroll UNROLL_PINS
      for (int p = 0; p < MAXP; ++p) {
        Fi[p] = false;
        ji[p] = -1;
        if (!trig[p]) continue;
        const int pv = cv[p], nv = ev_val(w0[p]);
        const int tx2 = toggle_x2(pv, nv);
        const int r2 = rise_x2(pv, nv);
        const int rf = r2 > 0 ? 0 : 1;
        tog += (uint32_t)tx2;
        // clock-cycle glitch flag of this event (TimeBasedPower.cc:589-637)
        // and its scaling ratio from flagged neighbours (:303-339)
        float ratio = -1.f;
        const int i = idx[p];
        if (i >= 1) {
          const int j = (jl[p] >= 0 && lc[p] == c) ? jl[p] + 1 : 0;
          ji[p] = j;
          const uint32_t cend = (c + 1) * T;
          const bool same1 = w1[p] != EV_INF && ev_time(w1[p]) < cend;
          const bool isclk = (cd.clock_mask >> p) & 1u;
          const bool F = !isclk && (same1 || (j & 1));
          Fi[p] = F;
          if (F) {
            int64_t pw = lF[p] ? (int64_t)(t - lt[p]) : -1;
            if (w1[p] != EV_INF) {
              const int j1 = same1 ? j + 1 : 0;
              const bool same2 = w2[p] != EV_INF && ev_time(w2[p]) < (c1[p] + 1) * T;
              if (same2 || (j1 & 1)) pw = max(pw, (int64_t)(ev_time(w1[p]) - t));
            }
            if (pw > 0) {
              const float r = (float)pw * gk[p];
              ratio = fminf(1.f, r * r);
            }
          }
        }
        if (ratio > 0.f) ++glc;
        float E;
        if (p < nin) {
          E = a.energy[eb + cd.slot_off[p] + ((rf << sbits) | (sbits ? (int)sb : 0))];
        } else {
          // related input: the latest event before t whose arc-delayed arrival is
          // closest to t, within the same cycle (TimeBasedPower.cc:412-473)
          const int o = p - nin;
          int rel = -1, relrf = 0;
          int64_t best = (int64_t)wk.t_max << ARC_FRAC_BITS;
#pragma unroll
          for (int q = 0; q < MAXP; ++q) {
            if (q >= nin || cnt[q] == 0) continue;
            const int32_t d = a.arc_q[ab + (((int64_t)o * nin + q) * 2 + rfl[q]) * 2 + rf];
            if (d == ARC_NONE) continue;
            int64_t diff = (((int64_t)t - (int64_t)lt[q]) << ARC_FRAC_BITS) - (int64_t)d;
            diff = diff < 0 ? -diff : diff;
            if (diff < best && lc[q] == c) {
              best = diff;
              rel = q;
              relrf = rfl[q];
            }
          }
          const int64_t so = eb + cd.slot_off[p];
          if (rel >= 0)
            E = a.energy[so + ((((rel * 2 + relrf) * 2 + rf) << sbits) | (sbits ? (int)sb : 0))];
          else
            E = a.energy[so + ((nin * 4) << sbits) + rf];
          if (r2 > 0) {
            const float es = 0.5f * (float)r2 * csw[p];
            if (ratio > 0.f) gs += ratio * es; else rs += es;
          }
        }
        E *= 0.5f * (float)tx2;
        if (ratio > 0.f) rg += ratio * E; else ri += E;
      }
      ff_add(acc_hi[PG_INT], acc_lo[PG_INT], ri);
      ff_add(acc_hi[PG_GINT], acc_lo[PG_GINT], rg);
      ff_add(acc_hi[PG_SW], acc_lo[PG_SW], rs);
      ff_add(acc_hi[PG_GSW], acc_lo[PG_GSW], gs);
      p_int += ri;
      p_gint += rg;
      p_sw += rs;
      p_gsw += gs;
      // consume the events at t
#pragma unroll
      for (int p = 0; p < MAXP; ++p) {
        if (!trig[p]) continue;
        const int nv = ev_val(w0[p]);
        rfl[p] = (idx[p] == 0) ? 1 : (rise_x2(cv[p], nv) > 0 ? 0 : 1);
        lF[p] = Fi[p];
        jl[p] = ji[p];
        lt[p] = t;
        lc[p] = c;
        cv[p] = nv;
        const int32_t i = ++idx[p];
        w0[p] = w1[p];
        c0[p] = c1[p];
        w1[p] = w2[p];
        if (w1[p] != EV_INF) c1[p] = cycle_of(ev_time(w1[p]), T, invT);
        w2[p] = i + 2 < cnt[p] ? evp[p][i + 2] : EV_INF;
      }
      sb = sa;
      prev_t = t;
    }
    flush_pending();
    if (lend > prev_t) ff_add_prod(acc_hi[PG_LEAK], acc_lo[PG_LEAK], s_leak[loff + sb], (float)(lend - prev_t));
    const int64_t slot = (int64_t)wt * a.n_sched + gi;
    const int64_t plane = (int64_t)a.n_work_tiles * a.n_sched;
    const float iql = ldexpf(a.inv_q_leak, cd.leak_shift);  // the cell's own leakage quanta
#pragma unroll
    for (int k = 0; k < N_PG; ++k) {
      const float iq = k == PG_LEAK ? iql : inv_qf;  // powers of two: exact scaling
      a.pg[k * plane + slot] = __float2ll_rn(acc_hi[k] * iq) + __float2ll_rn(acc_lo[k] * iq);
    }
    a.pg_tog[slot] = tog;
    a.pg_glitch[slot] = glc;
  }
  // leakage at the tile start, summed over this workgroup's gates
  for (int off = 32; off > 0; off >>= 1) base_acc += __shfl_xor(base_acc, off);
  if ((threadIdx.x & 63) == 0) atomicAdd((unsigned long long*)&s_base, (unsigned long long)base_acc);
  __syncthreads();

  // flush: S_c = base + sum of leakage steps in earlier cycles of the tile,
  // and the in-cycle part of the leakage integral
  if (threadIdx.x == 0)
    leak_flush(bins + BIN_LD * TC, bins + BIN_LX * TC, TC, s_base, tc0, (int64_t)T, wk.t_max, wk.n_cycle);
  __syncthreads();
  long long* out = a.cyc_part + ((int64_t)wt * a.n_range + range) * N_BIN * TC;
  for (int i = threadIdx.x; i < N_BIN * TC; i += blockDim.x) out[i] = bins[i];
  __syncthreads();  // the bins are cleared for the next item
  }
}


// ---------------------------------------------------------------------------
// Glitch filtering, once per net event (power/TimeBasedPower.cc:589-637 and
// :303-339): the EV_GLITCH bit of event e is set when its clock-cycle glitch
// pulse width (ticks) is nonzero.  The power kernels recompute that width from
// the flags and times of e's neighbours.  A chunk is 256 consecutive events of
// one net (chunk_off[n] is the first chunk of net n).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int group_pos(const uint32_t* ev, int64_t i, uint32_t cstart) {
  // number of events before i in the same cycle group (event 0 is never grouped)
  int j = 0;
  for (int64_t k = i - 1; k >= 1 && ev_time(ev[k]) >= cstart; --k) ++j;
  return j;
}

// glitch pulse width of event i (of cnt) of one net's event list e (0: none)
__device__ __forceinline__ uint32_t annotate_one(const uint32_t* e, int64_t i, int64_t cnt, uint32_t T, uint32_t cyc_m,
                                                 int cyc_sh) {
  if (i < 1) return 0;
  const uint32_t t = ev_time(e[i]);
  const uint32_t c = cycle_mul(t, cyc_m, cyc_sh);
  const uint32_t cs = c * T, ce = cs + T;
  const int j = group_pos(e, i, cs);
  const bool has1 = i + 1 < cnt;
  const uint32_t t1 = has1 ? ev_time(e[i + 1]) : 0;
  const bool last = !has1 || t1 >= ce;
  if (!(!last || (j & 1))) return 0;
  int64_t pw = -1;
  if (i - 1 >= 1) {
    bool fprev;
    if (j > 0) {
      fprev = true;  // same group and not its last event
    } else {
      const uint32_t tp = ev_time(e[i - 1]);
      const uint32_t cps = cycle_mul(tp, cyc_m, cyc_sh) * T;
      fprev = group_pos(e, i - 1, cps) & 1;  // last of its group: flagged iff at an odd position
    }
    if (fprev) pw = t - ev_time(e[i - 1]);
  }
  if (has1) {
    bool fnext;
    const bool has2 = i + 2 < cnt;
    const uint32_t t2 = has2 ? ev_time(e[i + 2]) : 0;
    if (!last) {
      const bool last1 = !has2 || t2 >= ce;
      fnext = !last1 || ((j + 1) & 1);
    } else {
      const uint32_t c1e = (cycle_mul(t1, cyc_m, cyc_sh) + 1) * T;
      fnext = has2 && t2 < c1e;
    }
    if (fnext) pw = max(pw, (int64_t)(t1 - t));
  }
  return pw > 0 ? (uint32_t)pw : 0u;
}

// chunk_net[c] = the net whose events block c covers (blocks of 256 events,
// chunk_off[n] = first block of net n); one thread per net.
__global__ void chunk_net_kernel(const int64_t* __restrict__ chunk_off, int64_t n_net, int32_t* __restrict__ chunk_net) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_net) return;
  for (int64_t c = chunk_off[n]; c < chunk_off[n + 1]; ++c) chunk_net[c] = (int32_t)n;
}

// Whole-store annotation, the rule of annotate_one restated for one wave per
// chunk range.  With j(m) the position of event m in its cycle group (events
// from index 1 on) and last(m) whether it ends the group, event m is a
// candidate when m >= 1 and (!last(m) || j(m) odd); event k then gets
// pw = max(t_k - t_{k-1} if k-1 is a candidate, t_{k+1} - t_k if k+1 is one)
// when k itself is a candidate, else 0.
//
// Every wave takes a contiguous range of chunks and walks it in order; lane l
// holds the events at chunk positions 4l .. 4l + 3 in registers, the events
// next to the chunk (positions -1, 256, 257) are loaded by lanes 0 and 63,
// and neighbours across lanes come from shuffles.  Group positions come from the
// start of each event's group: a 64-lane max-scan of the last group start in
// each lane, seeded with the group start of the event at position -1, which
// is carried from the previous chunk of the same net (found in memory only
// for a wave's first chunk).  Chunks are loaded two ahead of the one being
// evaluated.
constexpr uint32_t AN_NONE = 0xFFFFFFFEu;  // cycle of index <= 0 (never grouped)
constexpr uint32_t AN_END = 0xFFFFFFFFu;   // cycle past the net's last event

// A wave-uniform value moved to scalar registers, so that branches on it are
// scalar branches.
__device__ __forceinline__ int64_t an_uniform(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Chunk c of a range, after the chunk at e/left/next_first (same net: the
// next 256 events; else the first chunk of a later net).
__device__ __forceinline__ void an_next_chunk(const int64_t* __restrict__ net_off, uint32_t* ev,
                                              const int64_t* __restrict__ chunk_off,
                                              const int32_t* __restrict__ chunk_net, int64_t c, uint32_t*& e,
                                              int64_t& left, int64_t& next_first, bool& first) {
  if (c < next_first) {
    e += 256;
    left -= 256;
    first = false;
    return;
  }
  const int64_t n = an_uniform(chunk_net[c]);
  const int64_t b = an_uniform(net_off[n]), i0 = (c - an_uniform(chunk_off[n])) * 256;
  e = ev + b + i0;
  left = an_uniform(net_off[n + 1]) - b - i0;
  next_first = an_uniform(chunk_off[n + 1]);
  first = i0 == 0;
}

// Raw words of one chunk for lane `lane`: positions 4 lane + r; lane 0 also
// -1 and lane 63 also 256, 257 (the other lanes load position 0 in their
// place: per-lane addresses keep these loads from being treated as uniform
// values, which would wait for them at once).  Every position is clamped into
// the net (a chunk holds at least one event), so every load is issued
// unconditionally; absent positions are masked to EV_INF only when the chunk
// is evaluated, which leaves the loads in flight until then.  Words of
// neighbouring chunks are read for their time bits only, which other waves
// never change (they may rewrite the glitch bit meanwhile).
__device__ __forceinline__ void an_fetch(const uint32_t* e, int64_t left, bool first, int lane, uint32_t (&x)[7]) {
  const int p0 = 4 * lane;
  const int lim = left >= 256 ? 256 : (int)left;  // valid positions, >= 1
#pragma unroll
  for (int r = 0; r < 4; ++r) x[r] = e[min(p0 + r, lim - 1)];
  x[4] = e[lane == 0 && !first ? -1 : 0];
  x[5] = e[lane == 63 && left > 256 ? 256 : 0];
  x[6] = e[lane == 63 && left > 257 ? 257 : 0];
}

// One chunk's state in a wave's pipeline.
#define AN_SLOT(X) uint32_t x##X[7]; uint32_t* e##X = nullptr; int64_t left##X = 0, nf##X = -1; bool first##X = false

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void glitch_annotate_kernel(
    const int64_t* __restrict__ net_off,
                                                              uint32_t* __restrict__ ev,
                                                              const int64_t* __restrict__ chunk_off,
                                                              const int32_t* __restrict__ chunk_net,
                                                              int64_t n_chunk, uint32_t cyc_m, int cyc_sh) {
  const int lane = threadIdx.x & 63;
  // the wave index in scalar registers: the chunk loop is then a scalar loop
  const int64_t wave = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t n_wave = (int64_t)gridDim.x * 4;
  const int64_t per = (n_chunk + n_wave - 1) / n_wave;
  const int64_t c_begin = wave * per;
  const int64_t c_end = min(n_chunk, c_begin + per);
  if (c_begin >= c_end) return;  // the whole wave: no barriers follow
  // Three chunk slots take turns: the chunk being evaluated, the next one
  // (loaded) and the one after (being loaded); the loop is unrolled three
  // times so that no slot's registers are copied while its loads are in
  // flight.  Past the range's end the last chunk is loaded again (never
  // evaluated).
  AN_SLOT(A);
  AN_SLOT(B);
  AN_SLOT(C);
  an_next_chunk(net_off, ev, chunk_off, chunk_net, c_begin, eA, leftA, nfA, firstA);
  an_fetch(eA, leftA, firstA, lane, xA);
  eB = eA, leftB = leftA, nfB = nfA, firstB = firstA;
  if (c_begin + 1 < c_end) an_next_chunk(net_off, ev, chunk_off, chunk_net, c_begin + 1, eB, leftB, nfB, firstB);
  an_fetch(eB, leftB, firstB, lane, xB);
  // group start of the event at position -1, relative to the chunk: found in
  // memory for the wave's first chunk, carried afterwards (a net's first chunk
  // has no event at -1 and ignores it: its position 1 always starts a group)
  int s_rel = -1;
  if (!firstA) {
    const int64_t i0 = (c_begin - an_uniform(chunk_off[an_uniform(chunk_net[c_begin])])) * 256;  // >= 256
    const uint32_t cc = __builtin_amdgcn_readlane(cycle_mul(ev_time(xA[4]), cyc_m, cyc_sh), 0);
    int64_t q = 0;  // events of the group before position -1 (index i0 - 2 - q >= 1)
    while (i0 - 2 - q >= 1 &&
           __builtin_amdgcn_readfirstlane(cycle_mul(ev_time(eA[-2 - q]), cyc_m, cyc_sh)) == cc)
      ++q;
    s_rel = (int)(-1 - q);
  }
  const int p0 = 4 * lane;
  // Evaluate chunk c held in slot (x0, e0, left0, first0) while loading chunk
  // c + 2 into slot z from slot y's state (chunk c + 1).
  auto step = [&](int64_t c, const uint32_t (&x0)[7], uint32_t* e0, int64_t left0, bool first0, uint32_t* ey,
                  int64_t lefty, int64_t nfy, bool firsty, uint32_t (&xz)[7], uint32_t*& ez, int64_t& leftz,
                  int64_t& nfz, bool& firstz) {
    ez = ey, leftz = lefty, nfz = nfy, firstz = firsty;
    if (c + 2 < c_end) an_next_chunk(net_off, ev, chunk_off, chunk_net, c + 2, ez, leftz, nfz, firstz);
    an_fetch(ez, leftz, firstz, lane, xz);
    // ---- chunk c: positions p = 4 lane + r, rem of them valid
    const int rem = left0 >= 256 ? 256 : (int)left0;
    uint32_t w[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) w[r] = p0 + r < rem ? x0[r] : EV_INF;
    const uint32_t hp = first0 ? EV_INF : x0[4];         // meaningful in lane 0
    const uint32_t hn0 = left0 > 256 ? x0[5] : EV_INF;  // meaningful in lane 63
    const uint32_t hn1 = left0 > 257 ? x0[6] : EV_INF;
    uint32_t cy[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
      cy[r] = p0 + r >= rem ? AN_END
                            : ((first0 && p0 + r == 0) ? AN_NONE : cycle_mul(ev_time(w[r]), cyc_m, cyc_sh));
    // position -1 is index i0 - 1 >= 255, absent when the chunk starts the net
    const uint32_t c_m1 = first0 ? AN_NONE : cycle_mul(ev_time(hp), cyc_m, cyc_sh);
    const uint32_t c_256 = left0 > 256 ? cycle_mul(ev_time(hn0), cyc_m, cyc_sh) : AN_END;
    const uint32_t c_257 = left0 > 257 ? cycle_mul(ev_time(hn1), cyc_m, cyc_sh) : AN_END;
    const uint32_t c_up = __shfl_up(cy[3], 1);
    const uint32_t c_dn = __shfl_down(cy[0], 1);
    const uint32_t cp0 = lane == 0 ? c_m1 : c_up;
    const uint32_t cn3 = lane == 63 ? c_256 : c_dn;
    // group starts (bit r), the last one in this lane, then the scan
    const uint32_t st = (cy[0] != cp0 ? 1u : 0u) | (cy[1] != cy[0] ? 2u : 0u) | (cy[2] != cy[1] ? 4u : 0u) |
                        (cy[3] != cy[2] ? 8u : 0u);
    int v = st ? p0 + 31 - __builtin_clz(st) : INT32_MIN;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int o = __shfl_up(v, d);
      if (lane >= d) v = max(v, o);
    }
    int run = __shfl_up(v, 1);
    if (lane == 0) run = INT32_MIN;
    run = max(run, s_rel);
    int j[4];
    uint32_t cand = 0;  // bit r: candidate
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (st & (1u << r)) run = p0 + r;
      j[r] = p0 + r - run;
      const uint32_t cn = r < 3 ? cy[r + 1] : cn3;
      // a real cycle: index >= 1 and inside the net
      if (cy[r] < AN_NONE && (cn != cy[r] ? (j[r] & 1) != 0 : true)) cand |= 1u << r;
    }
    // candidates next to the lane's events
    bool cprev0 = (__shfl_up(cand, 1) & 8u) != 0;
    bool cnext3 = (__shfl_down(cand, 1) & 1u) != 0;
    if (lane == 0)  // position -1: its group ends at position 0 when cycles differ
      cprev0 = c_m1 < AN_NONE && (cy[0] != c_m1 ? ((-1 - s_rel) & 1) != 0 : true);
    if (lane == 63) {  // position 256
      const int jn = c_256 != cy[3] ? 0 : j[3] + 1;
      cnext3 = c_256 < AN_NONE && (c_257 != c_256 ? (jn & 1) != 0 : true);
    }
    const uint32_t t_up = __shfl_up(ev_time(w[3]), 1), t_dn = __shfl_down(ev_time(w[0]), 1);
    const int32_t tp0 = (int32_t)(lane == 0 ? ev_time(hp) : t_up);
    const int32_t tn3 = (int32_t)(lane == 63 ? ev_time(hn0) : t_dn);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int32_t t = (int32_t)ev_time(w[r]);
      const bool cp = r == 0 ? cprev0 : ((cand >> (r - 1)) & 1u) != 0;
      const bool cn = r == 3 ? cnext3 : ((cand >> (r + 1)) & 1u) != 0;
      const int32_t tp = r == 0 ? tp0 : (int32_t)ev_time(w[r - 1]);
      const int32_t tn = r == 3 ? tn3 : (int32_t)ev_time(w[r + 1]);
      int32_t pw = 0;
      if ((cand >> r) & 1u) pw = max(cp ? t - tp : 0, cn ? tn - t : 0);
      // the word is rewritten only when its flag changes
      const uint32_t nw = pw > 0 ? (w[r] | EV_GLITCH) : (w[r] & ~EV_GLITCH);
      if (p0 + r < rem && nw != w[r]) e0[p0 + r] = nw;
    }
    // group start of position 255, relative to the next chunk (of this net)
    s_rel = (int)__builtin_amdgcn_readlane(run, 63) - 256;
  };
  for (int64_t c = c_begin;;) {
    step(c, xA, eA, leftA, firstA, eB, leftB, nfB, firstB, xC, eC, leftC, nfC, firstC);
    if (++c >= c_end) break;
    step(c, xB, eB, leftB, firstB, eC, leftC, nfC, firstC, xA, eA, leftA, nfA, firstA);
    if (++c >= c_end) break;
    step(c, xC, eC, leftC, firstC, eA, leftA, nfA, firstA, xB, eB, leftB, nfB, firstB);
    if (++c >= c_end) break;
  }
}
#undef AN_SLOT

// The same for the events of listed tiles only: one thread per (net n,
// listed tile l) walks the events of n from the one before idx_begin[n][l] up
// to the first event at or after the end of tile tile_list[l], both included
// (the power kernels read the glitch flags of the events next to a tile's
// own).  annotate_one's rule, with cand(m) = m >= 1 && (!last(m) || j(m) odd):
// event m is flagged when cand(m) and the gap to a candidate neighbour is
// positive.  The thread carries the group position j along its events and
// reads each word once; only its first event and the one before it take
// group_pos's walk back.
constexpr int AN_PASS = 8;  // events per pass of a (net, tile) thread

__device__ __forceinline__ bool cand_at(const uint32_t* e, int64_t i, int64_t cnt, uint32_t T, uint32_t cyc_m,
                                        int cyc_sh) {
  if (i < 1) return false;
  const uint32_t cs = cycle_mul(ev_time(e[i]), cyc_m, cyc_sh) * T;
  const bool last = i + 1 >= cnt || ev_time(e[i + 1]) >= cs + T;
  return !last || (group_pos(e, i, cs) & 1);
}

__global__ __launch_bounds__(256) void glitch_annotate_tiles_kernel(const int64_t* __restrict__ net_off,
                                                                    uint32_t* __restrict__ ev,
                                                                    const int32_t* __restrict__ idx_begin,
                                                                    const int32_t* __restrict__ tile_list,
                                                                    int32_t n_list, int64_t n_net, int64_t tile_ticks,
                                                                    uint32_t T, uint32_t cyc_m, int cyc_sh) {
  // threads stride over the (net, tile) pairs under a capped grid
  const int64_t n_thr = n_net * n_list;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_thr; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = i / n_list;
    const int64_t b = net_off[n];
    const int64_t cnt = net_off[n + 1] - b;
    const int64_t tend = ((int64_t)tile_list[i - n * n_list] + 1) * tile_ticks;
    int64_t m = max(idx_begin[i] - 1, 0);
    if (m >= cnt) continue;
    uint32_t* e = ev + b;
    // words of m and m + 1; tp = time of m - 1; j = j(m)
    uint32_t w0 = e[m];
    uint32_t w1 = m + 1 < cnt ? e[m + 1] : 0u;
    int j = group_pos(e, m, cycle_mul(ev_time(w0), cyc_m, cyc_sh) * T);
    bool cprev = cand_at(e, m - 1, cnt, T, cyc_m, cyc_sh);
    uint32_t tp = m >= 1 ? ev_time(e[m - 1]) : 0u;
    // AN_PASS events per pass, their next words loaded together
    for (bool done = false; !done; m += AN_PASS) {
      uint32_t win[AN_PASS + 2];
      win[0] = w0;
      win[1] = w1;
#pragma unroll
      for (int q = 0; q < AN_PASS; ++q) win[2 + q] = m + 2 + q < cnt ? e[m + 2 + q] : 0u;
#pragma unroll
      for (int q = 0; q < AN_PASS; ++q) {
        const int64_t mq = m + q;
        const uint32_t t0 = ev_time(win[q]), t1 = ev_time(win[q + 1]), t2 = ev_time(win[q + 2]);
        const bool has1 = mq + 1 < cnt, has2 = mq + 2 < cnt;
        const uint32_t ce = cycle_mul(t0, cyc_m, cyc_sh) * T + T;
        const bool last = !has1 || t1 >= ce;
        const bool cand = mq >= 1 && (!last || (j & 1));
        // the next event's group position and candidacy
        const int j1 = (mq >= 1 && !last) ? j + 1 : 0;
        bool cnext = false;
        if (has1) {
          const uint32_t ce1 = cycle_mul(t1, cyc_m, cyc_sh) * T + T;
          const bool last1 = !has2 || t2 >= ce1;
          cnext = !last1 || (j1 & 1);
        }
        int64_t pw = -1;
        if (cand) {
          if (mq >= 2 && cprev) pw = (int64_t)(t0 - tp);
          if (has1 && cnext) pw = max(pw, (int64_t)(t1 - t0));
        }
        const uint32_t nw = pw > 0 ? (win[q] | EV_GLITCH) : (win[q] & ~EV_GLITCH);
        if (nw != win[q]) e[mq] = nw;
        if ((int64_t)t0 >= tend || !has1) {
          done = true;
          break;
        }
        cprev = cand;
        tp = t0;
        j = j1;
      }
      w0 = win[AN_PASS];
      w1 = win[AN_PASS + 1];
    }
  }
}

// ---------------------------------------------------------------------------
// Fused power kernel specialised for one single-output cell type with NP pins
// (NP - 1 inputs, output = pin NP - 1) whose internal power has no `when`
// condition.  Every workgroup takes gates of one cell type, so the pin loops
// are compile-time, the per-gate energies and arc delays live in registers and
// the cell's leakage table in LDS.  Glitch flags come from the annotation
// pass.  Per-cycle sums go to a.bin_copies lane-interleaved copies of the LDS
// bins so that lanes working on the same cycle do not serialise on one address.
// Semantics are those of power_tile_kernel.
// ---------------------------------------------------------------------------
// Occupancy targets by pin count (waves per SIMD): 5 for 2 pins (a few spills
// outside the step loop; 12% faster than 4), 4 for 3 pins (at 5 it spills in
// the step loop and runs 1.8x slower), 3 for 4 pins, 2 above.
// Clock edges taken per workgroup (power_cell_kernel<3, KC>, KC >= 0): the
// edges of one tile of the clock net sit in LDS slots 1..n (slot 0: the state
// before the tile at the tile start, slots n + 1 and n + 2: EV_INF), so a
// tile holds at most CK_SLOTS - 3 edges and CK_TILE_MAX cycles.
constexpr int CK_SLOTS = 1040;
constexpr int CK_TILE_MAX = 512;

// Inclusive prefix sums of x[0..n] in place (n < blockDim.x * 64), by the
// whole workgroup; tmp holds one value per wave.
template <typename V>
__device__ void block_scan_inplace(V* x, int n, V* tmp) {
  const int nt = (int)blockDim.x, j = (int)threadIdx.x;
  const int seg = (n + nt) / nt;
  const int b = j * seg;
  V loc = 0;
  for (int k = 0; k < seg; ++k)
    if (b + k <= n) loc += x[b + k];
  V v = loc;
  const int lane = j & 63, wv = j >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const V y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) tmp[wv] = v;
  __syncthreads();
  V run = v - loc;
  for (int k = 0; k < wv; ++k) run += tmp[k];
  for (int k = 0; k < seg; ++k)
    if (b + k <= n) {
      run += x[b + k];
      x[b + k] = run;
    }
  __syncthreads();
}

// State without the clock bit KC (the histogram index) and back.
template <int KC>
__device__ __forceinline__ uint32_t ck_pack(uint32_t st) {
  return ((st >> (KC + 1)) << KC) | (st & ((1u << KC) - 1u));
}
template <int KC>
__device__ __forceinline__ uint32_t ck_unpack(uint32_t j, uint32_t b) {
  return ((j >> KC) << (KC + 1)) | (b << KC) | (j & ((1u << KC) - 1u));
}

// (KC >= 0: the clock-edge tables hold the workgroups to 3 per compute unit,
// so the registers of 3 waves per SIMD)
template <int NP, int KC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    KC >= 0 ? 3 : (NP <= 2 ? 5 : (NP == 3 ? 4 : (NP == 4 ? 3 : 2))))))
void power_cell_kernel(PowerArgs a, int32_t range_begin, int32_t range_count, int32_t cell, unsigned int* wq) {
  // KC >= 0: pin KC is the cell's only clock pin; the gates whose clock pin is
  // on the work item's clock net leave the clock events out of their merge and
  // the workgroup takes them once per item.  The per-gate effects of a clock
  // edge (internal energy, toggles, the leakage state) and the per-cycle bins
  // are the same integers as when each gate steps through the edges: a gate's
  // leakage splits its intervals between D/Q events by the clock-high ticks
  // (prefix over the edges), and the bins get, per edge, the leakage steps of
  // the gates in each non-clock state (counted through a difference histogram
  // over the edges) and the summed edge energies.
  constexpr bool CLKB = KC >= 0;
  static_assert(!CLKB || NP == 3, "the histogram packs 4 non-clock states into 16-bit fields");
  constexpr int NIN = NP - 1;
  constexpr int OUT = NP - 1;
  constexpr int NS = 6 * NIN + 2;
  constexpr int NSTATE = 1 << NP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ long long s_leakq[NSTATE];
  __shared__ long long s_base;
  const int TC = a.tile_cycles;
  // one copy of the bins: N_BIN rows of TC + 1 columns; the last column takes
  // the steps of events in cycles past the tile (the last tile's end), so the
  // bin updates need no test
  const int TCP = TC + 1;
  const int BW = N_BIN * TCP;
  long long* bins = reinterpret_cast<long long*>(smem_raw);

  __shared__ long long s_item;
  constexpr int CKE = CLKB ? CK_SLOTS : 1;
  constexpr int CKC = CLKB ? CK_TILE_MAX + 2 : 1;
  // per slot: clock-high ticks from the tile start to the edge << 32 | the
  // word (time << 3 | not-a-rise << 2 | value)
  __shared__ unsigned long long ck_wh[CKE];
  __shared__ unsigned long long ck_d[CKE];    // gates per non-clock state (16 bits each): differences, then prefix
  __shared__ uint16_t ck_cf[CKC];             // first slot in each tile cycle (the clamped column)
  __shared__ long long ck_q[4];               // summed edge energies of the item's gates, per edge kind
  __shared__ int ck_n[4];                     // edges per kind: 2 * not-a-rise + (toggle_x2 == 2)
  __shared__ long long ck_l[2];               // clock net: first event, events
  __shared__ int ck_s[8];                     // clock net, first edge index, first-event flag, edges, overflow
  __shared__ unsigned long long ck_t[4];      // scan scratch
  __shared__ longlong2 ck_lk[CLKB ? 4 : 1];   // per non-clock state: leakage quanta, clock low and high
  const CellDev cd = a.cells[cell];
  const uint32_t clk = cd.clock_mask;
  const uint32_t arcm = cd.arc_mask;  // inputs that can be the related pin of an output event
  for (int i = threadIdx.x; i < NSTATE; i += blockDim.x) s_leakq[i] = a.leak_q[cd.leak_off + i];
  if constexpr (CLKB) {
    for (int j = threadIdx.x; j < 4; j += blockDim.x) {
      const uint32_t s0 = ck_unpack<KC>((uint32_t)j, 0u), s1 = ck_unpack<KC>((uint32_t)j, 1u);
      ck_lk[j] = make_longlong2(a.leak_q[cd.leak_off + s0], a.leak_q[cd.leak_off + s1]);
    }
  }
  const int n_copy = a.bin_copies;
  long long* mybins = bins + (threadIdx.x % n_copy) * BW;

  const int64_t n_work = (int64_t)a.n_work_tiles * range_count;
  int64_t w = 0;
  for (int64_t it = 0; next_work_item(a, wq, n_work, it, w, &s_item); ++it) {
  const int wt = (int)(w / range_count);
  const int range = range_begin + (int)(w - (int64_t)wt * range_count);
  const WorkTile wk = resolve_work_tile(a, wt);
  const int tile = wk.tile;
  for (int i = threadIdx.x; i < BW * n_copy; i += blockDim.x) bins[i] = 0;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();

  const uint32_t T = (uint32_t)a.T;
  const float invT = a.inv_T;
  const int64_t tc0 = (int64_t)tile * TC;
  const bool last_tile = tile == wk.n_tile - 1;
  const uint32_t tstart = (uint32_t)(tc0 * (int64_t)T);
  const uint32_t tend = last_tile ? EV_TIME_INF : (uint32_t)((tc0 + TC) * (int64_t)T);
  const uint32_t lend = last_tile ? (uint32_t)wk.t_max : tend;
  const uint32_t t_max = (uint32_t)wk.t_max;
  const float inv_qe = a.inv_q_energy, inv_ql = a.inv_q_leak;

  long long base_acc = 0;
  const int64_t gi_begin = (int64_t)range * a.range_gates;
  const int64_t gi_end = min(a.n_sched, gi_begin + a.range_gates);

  // The item's clock net (that of its first gate) and its edges in the tile:
  // the events a gate's clock pin would consume, from the same start as the
  // per-pin set-up below.  -1 edges: the gates step through their clock
  // events (no clock net, more edges than slots, a longer tile, or more than
  // two edges in a cycle).
  int ck_ne = -1, ck_net = -1;
  if constexpr (CLKB) {
    if (threadIdx.x == 0) {
      int cnet = -1;
      int64_t cnb = 0, ccn = 0;
      if (TC <= CK_TILE_MAX) {
        for (int64_t gi = gi_begin; gi < gi_end; ++gi) {
          const int g = a.order[gi];
          if (g >= 0) {
            cnet = a.pin_net[a.gate_pin_off[g] + KC];
            break;
          }
        }
      }
      if (cnet >= 0) {
        const int64_t ni = cnet + wk.net_base;
        cnb = a.net_off[ni];
        ccn = a.net_off[ni + 1] - cnb;
      }
      int32_t fi = 0, fx = 0;
      if (ccn > 0) {
        const int32_t i0 = wk.tidx[(int64_t)cnet * wk.tile_stride + wk.tile_col];
        const uint32_t w1 = a.ev[cnb + max(i0 - 1, 0)];
        uint32_t l = w1;
        fi = i0;
        if (i0 == 0) {
          if (wk.trace_start && ev_time(w1) == 0) {
            fi = 1;  // the time-0 event is the initial state
          } else {
            l = 2u;  // unknown state before the net's first event
            fx = 1;
          }
        }
        ck_wh[0] = (tstart << EV_SHIFT) | (l & 3u);
        ck_d[0] = 0;
      } else {
        cnet = -1;
      }
      ck_l[0] = cnb;
      ck_l[1] = ccn;
      ck_s[0] = cnet;
      ck_s[1] = fi;
      ck_s[2] = fx;
      ck_s[3] = 0;
      ck_s[4] = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        ck_q[k] = 0;
        ck_n[k] = 0;
      }
    }
    __syncthreads();
    const int cnet = ck_s[0];
    if (cnet >= 0) {
      const int64_t cnb = ck_l[0], ccn = ck_l[1];
      const int32_t fi = ck_s[1];
      const bool fx = ck_s[2] != 0;
      const uint32_t w0 = (uint32_t)ck_wh[0];
      constexpr int NE_MAX = CK_SLOTS - 3;
      // edge k in slot k + 1; k = NE_MAX probes for one edge too many
      for (int k = threadIdx.x; k <= NE_MAX; k += blockDim.x) {
        const int64_t idx = (int64_t)fi + k;
        if (idx >= ccn) break;
        const uint32_t w = a.ev[cnb + idx];
        const uint32_t t = ev_time(w);
        if (t >= tend) break;
        if (k == NE_MAX) {
          ck_s[4] = 1;
          break;
        }
        const uint32_t pw = k == 0 ? w0 : a.ev[cnb + idx - 1];
        const uint32_t pt = k == 0 ? tstart : ev_time(pw);
        const int pv = ev_val(pw), nv = ev_val(w);
        const int r2 = rise_x2(pv, nv), tx2 = toggle_x2(pv, nv);
        // TimeBasedPower.cc:420-431: after the net's first event (not at
        // time 0) the edge counts as not a rise
        const uint32_t nr = (k == 0 && fx) ? 1u : (r2 > 0 ? 0u : 1u);
        ck_wh[k + 1] = (t << EV_SHIFT) | (nr << 2) | (uint32_t)nv;
        ck_d[k + 1] = pv == 1 ? t - pt : 0u;  // clock-high ticks since the slot before, scanned below
        if (tx2 > 0) atomicAdd(&ck_n[(r2 > 0 ? 0 : 2) + (tx2 == 2 ? 1 : 0)], 1);
        atomicAdd(&ck_s[3], 1);
      }
    }
    __syncthreads();
    int ne = (cnet >= 0 && ck_s[4] == 0) ? ck_s[3] : -1;
    if (ne >= 0) {
      block_scan_inplace(ck_d, ne, ck_t);
      for (int sl = threadIdx.x; sl <= ne; sl += blockDim.x) ck_wh[sl] |= ck_d[sl] << 32;
      __syncthreads();
      // first slot of each tile cycle; more than two edges in a cycle fall back
      for (int sl = 1 + (int)threadIdx.x; sl <= ne; sl += blockDim.x) {
        const auto col = [&](int q) {
          return min((int)((int64_t)cycle_mul(ev_time((uint32_t)ck_wh[q]), a.cyc_m, a.cyc_sh) - tc0), TC);
        };
        const int ce = col(sl);
        const int cp = sl == 1 ? -1 : col(sl - 1);
        for (int ci = cp + 1; ci <= ce; ++ci) ck_cf[ci] = (uint16_t)sl;
        if (sl + 2 <= ne && col(sl + 2) == ce) ck_s[4] = 1;
      }
      const int last =
          ne >= 1 ? min((int)((int64_t)cycle_mul(ev_time((uint32_t)ck_wh[ne]), a.cyc_m, a.cyc_sh) - tc0), TC) : -1;
      for (int ci = last + 1 + (int)threadIdx.x; ci <= TC + 1; ci += blockDim.x) ck_cf[ci] = (uint16_t)(ne + 1);
      for (int sl = threadIdx.x; sl <= ne + 1; sl += blockDim.x) ck_d[sl] = 0;
      if (threadIdx.x == 0) {
        ck_wh[ne + 1] = EV_INF;
        ck_wh[ne + 2] = EV_INF;
      }
      __syncthreads();
      if (ck_s[4] != 0) ne = -1;
    }
    ck_ne = ne;
    ck_net = cnet;
  }

  for (int64_t gs = gi_begin + threadIdx.x; gs < gi_end; gs += blockDim.x) {
    // the lane's schedule position: by this tile's order when there is one
    const int64_t gi = a.perm ? (int64_t)a.perm[(int64_t)wt * a.n_sched + gs] : gs;
    const int g = a.order[gi];
    if (g < 0) continue;
    const int64_t pb = a.gate_pin_off[g];
    const int64_t eb = a.gate_energy_off[g];
    const int64_t ab = a.gate_arc_off[g];
    // per-gate energies and arc delays, kept in registers (the empty asm
    // statements before the step loop stop the compiler from turning the
    // selections below into a dynamically indexed scratch array)
    float E[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) E[k] = a.energy[eb + k];
    int32_t D[4 * NIN];
#pragma unroll
    for (int k = 0; k < 4 * NIN; ++k) D[k] = a.arc_q[ab + k];

    // Per pin: the current and next event words cw/nw (EV_INF past the net's
    // events), the last consumed event lw (its time, value and glitch flag;
    // before the first event of the tile the event before it, or a word
    // holding the initial state), the number of events from the current one
    // on, and two buffers fed by aligned 16-byte loads: words 2 and 3 of the
    // quad holding the current event (q2, q3) and the whole quad after it
    // (n0..n3).  The quad after is loaded at the end of the step that moves a
    // pin into a new quad and first read in the next step, so its load
    // overlaps a whole step.  Glitch widths come from the flags of the
    // neighbouring events (lw and nw), so no width array is read.
    // The set-up loads go out level by level for all pins at once (pin ->
    // net -> net range -> tile start -> event words): every address is kept
    // inside the arrays and unused results are discarded.
    const uint32_t* cp[NP];    // the current event
    int32_t rem[NP];           // events from the current one to the net's end
    uint32_t q2[NP], q3[NP], n0[NP], n1[NP], n2[NP], n3[NP];
    uint32_t cw[NP], nw[NP], lw[NP];
    float gk[NP];
    uint32_t rflm = (1u << NP) - 1;  // bit p: pin p's last consumed event was not a rise (TimeBasedPower.cc:420-431)
    uint32_t firstm = 0;             // bit p: pin p's next event is the net's event 0, after time 0
    uint32_t hasm = 0;               // bit p: pin p's net has events
    uint32_t sb = 0;
    const float csw = a.pin_csw[pb + OUT];
    int32_t net[NP];
    uint32_t pdef[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      net[p] = a.pin_net[pb + p];
      gk[p] = a.pin_gk[pb + p] * a.time_scale;
      pdef[p] = (uint32_t)a.pin_default[pb + p];
    }
    int64_t nb[NP], cn[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int64_t ni = (net[p] >= 0 ? net[p] : 0) + wk.net_base;
      nb[p] = a.net_off[ni];
      const int64_t ne = a.net_off[ni + 1];
      cn[p] = net[p] >= 0 ? ne - nb[p] : 0;
    }
    int32_t i0[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int32_t v = wk.tidx[(int64_t)(net[p] >= 0 ? net[p] : 0) * wk.tile_stride + wk.tile_col];
      i0[p] = cn[p] > 0 ? v : 0;
    }
    // the event before the tile (event 0 when there is none) and the one before that
    uint32_t wp1[NP], wp2[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      wp1[p] = a.ev[nb[p] + max(i0[p] - 1, 0)];
      wp2[p] = a.ev[nb[p] + max(i0[p] - 2, 0)];
    }
    int64_t g0[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      uint32_t l = pdef[p];  // time 0, no flag, the default state
      int32_t i = i0[p];
      if (cn[p] > 0) {
        hasm |= 1u << p;
        if (i0[p] == 0) {
          if (wk.trace_start && ev_time(wp1[p]) == 0) {
            l = wp1[p];  // the time-0 event is the initial state (TimeBasedPower.cc:782-787)
            i = 1;
          } else {
            l = (ev_time(wp1[p]) << EV_SHIFT) | 2u;  // unknown state before the net's first event
            firstm |= 1u << p;
          }
        } else {
          l = wp1[p];
          if (i0[p] - 1 >= 1 && rise_x2(ev_val(wp2[p]), ev_val(l)) > 0) rflm &= ~(1u << p);
        }
      }
      lw[p] = l;
      if (ev_val(l) == 1) sb |= 1u << p;
      rem[p] = (int32_t)(cn[p] - i);
      g0[p] = nb[p] + i;
      cp[p] = a.ev + g0[p];
    }
    // the quad holding the current event and, when the net has events past
    // it, the quad after (the store is padded by a quad)
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int64_t qa = g0[p] & ~(int64_t)3;
      const int kk = (int)(g0[p] - qa);
      const bool more = rem[p] > 4 - kk;
      const uint4 v = *reinterpret_cast<const uint4*>(a.ev + qa);
      const uint4 u = *reinterpret_cast<const uint4*>(a.ev + (more ? qa + 4 : qa));
      q2[p] = v.z;
      q3[p] = v.w;
      n0[p] = more ? u.x : EV_INF;
      n1[p] = more ? u.y : EV_INF;
      n2[p] = more ? u.z : EV_INF;
      n3[p] = more ? u.w : EV_INF;
      const uint32_t c0 = kk == 0 ? v.x : kk == 1 ? v.y : kk == 2 ? v.z : v.w;
      const uint32_t nx = kk == 0 ? v.y : kk == 1 ? v.z : kk == 2 ? v.w : u.x;
      cw[p] = rem[p] >= 1 ? c0 : EV_INF;
      nw[p] = rem[p] >= 2 ? nx : EV_INF;
    }
    // an input whose net has no events is never the related input: no arcs
#pragma unroll
    for (int q = 0; q < NIN; ++q)
      if (!((hasm >> q) & 1u)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) D[q * 4 + k] = ARC_NONE;
      }
#pragma unroll
    for (int k = 0; k < NS; ++k) asm volatile("" : "+v"(E[k]));
#pragma unroll
    for (int k = 0; k < 4 * NIN; ++k) asm volatile("" : "+v"(D[k]));
    base_acc += s_leakq[sb] >> 32;

    long long acc[N_PG];     // per-gate sums in quanta (leakage: quanta * ticks)
#pragma unroll
    for (int k = 0; k < N_PG; ++k) acc[k] = 0;
    uint32_t tog = 0, glc = 0;
    uint32_t prev_t = tstart;
    // leakage of the current state: common quanta << 32 | the cell's own quanta
    long long lq_cur = s_leakq[sb];
    // CLKB: the leakage of the current non-clock state with the clock low and
    // high, the clock-high ticks from the tile start to prev_t, and whether
    // the clock events come from the item's edges (bulk)
    long long lq_lo = 0, lq_hi = 0;
    uint32_t h_prev = 0;
    bool bulk = false;
    if constexpr (CLKB) {
      bulk = ck_ne >= 0 && net[KC] == ck_net;
      if (bulk) {
        cw[KC] = EV_INF;
        nw[KC] = EV_INF;
        atomicAdd(&ck_d[0], 1ull << (16 * ck_pack<KC>(sb)));
      }
      const uint32_t s0 = sb & ~(1u << KC), s1 = sb | (1u << KC);
      lq_lo = s_leakq[s0];
      lq_hi = s_leakq[s1];
    }

    // One pin event per step: the earliest pending event, the output pin's
    // first when several pins change at the same time (so that the related
    // input search sees only input events strictly before it).  Pins changing
    // together thus take successive steps of zero duration, which leaves the
    // leakage integral unchanged.  Everything per pin is picked by selects, so
    // the step has no branch per pin.
    for (;;) {
      int ps = OUT;
      uint32_t ws = cw[OUT];
      uint32_t t = ev_time(ws);
#pragma unroll
      for (int p = 0; p < NIN; ++p) {
        const uint32_t tp = ev_time(cw[p]);
        if (tp < t) {
          t = tp;
          ws = cw[p];
          ps = p;
        }
      }
      if (t >= tend) break;
      uint32_t lws = lw[0], nws = nw[0];
      float gks = gk[0];
#pragma unroll
      for (int p = 1; p < NP; ++p)
        if (p == ps) {
          lws = lw[p];
          nws = nw[p];
          gks = gk[p];
        }
      const bool fs = (firstm >> ps) & 1u;
      const int pv = ev_val(lws);
      const bool is_out = ps == OUT;
      const int nv = ev_val(ws);
      const uint32_t c = cycle_mul(t, a.cyc_m, a.cyc_sh);
      const uint32_t cstart = c * T, cend = cstart + T;
      const int cr = min((int)((int64_t)c - tc0), TC);
      long long* cb = mybins + cr;

      // leakage: the previous state over [prev_t, t), exact in quanta * ticks;
      // the step to the new state enters this cycle's bins
      if constexpr (!CLKB) {
        acc[PG_LEAK] += (long long)(int)lq_cur * (long long)(t - prev_t);
        const uint32_t sa = (sb & ~(1u << ps)) | ((nv == 1 ? 1u : 0u) << ps);
        const long long lq_new = s_leakq[sa];
        // (zero when the state does not change)
        const long long dq = (lq_new >> 32) - (lq_cur >> 32);
        atomicAdd((unsigned long long*)&cb[BIN_LD * TCP], (unsigned long long)dq);
        atomicAdd((unsigned long long*)&cb[BIN_LX * TCP], (unsigned long long)(dq * (long long)(t - cstart)));
        sb = sa;
        lq_cur = lq_new;
      } else {
        // the clock: its own events (not bulk), or the item's edges consumed
        // before this event: those before t, and those at t when the clock
        // pin precedes pin ps (the output goes first at equal times, then the
        // inputs in pin order); at most two edges share a cycle
        uint32_t bit, H;
        int el = 0;
        if (bulk) {
          const bool tie = ps != OUT && KC < ps;
          // the slots before, at and after the cycle's first: loaded together
          const int e = ck_cf[cr];
          const unsigned long long s0 = ck_wh[e - 1], s1 = ck_wh[e], s2 = ck_wh[e + 1];
          const uint32_t we0 = (uint32_t)s1, we1 = (uint32_t)s2;
          const bool c0 = ev_time(we0) < t || (ev_time(we0) == t && tie);
          const bool c1 = ev_time(we1) < t || (ev_time(we1) == t && tie);
          el = e - 1 + (c0 ? 1 : 0) + (c1 ? 1 : 0);
          const unsigned long long sl = c1 ? s2 : (c0 ? s1 : s0);
          const uint32_t wl = (uint32_t)sl;
          bit = (wl & 3u) == 1u ? 1u : 0u;
          H = (uint32_t)(sl >> 32) + (bit ? t - ev_time(wl) : 0u);
          if (el > 0) {
            lw[KC] = wl;  // the latest consumed clock event, for the related-input search
            rflm = (rflm & ~(1u << KC)) | (((wl >> 2) & 1u) << KC);
          }
        } else {
          bit = (sb >> KC) & 1u;
          H = h_prev + (bit ? t - prev_t : 0u);
        }
        const uint32_t dH = H - h_prev, dt = t - prev_t;
        acc[PG_LEAK] += (long long)(int)lq_hi * (long long)dH + (long long)(int)lq_lo * (long long)(dt - dH);
        h_prev = H;
        const uint32_t sbe = (sb & ~(1u << KC)) | (bit << KC);
        const uint32_t sa = (sbe & ~(1u << ps)) | ((nv == 1 ? 1u : 0u) << ps);
        const uint32_t sa0 = sa & ~(1u << KC);
        const longlong2 lk = ck_lk[ck_pack<KC>(sa)];
        const long long nlo = lk.x, nhi = lk.y;
        const bool abit = (sa >> KC) & 1u;
        const long long lq_c = bit ? lq_hi : lq_lo, lq_new = abit ? nhi : nlo;
        const long long dq = (lq_new >> 32) - (lq_c >> 32);
        atomicAdd((unsigned long long*)&cb[BIN_LD * TCP], (unsigned long long)dq);
        atomicAdd((unsigned long long*)&cb[BIN_LX * TCP], (unsigned long long)(dq * (long long)(t - cstart)));
        // a new non-clock state counts from the next unconsumed edge on
        if (bulk && sa0 != (sbe & ~(1u << KC)))
          atomicAdd(&ck_d[el + 1], (1ull << (16 * ck_pack<KC>(sa))) - (1ull << (16 * ck_pack<KC>(sbe))));
        sb = sa;
        lq_lo = nlo;
        lq_hi = nhi;
      }

      const int tx2 = table_x2(TOGGLE_X2_TABLE, pv, nv);
      const int r2 = table_x2(RISE_X2_TABLE, pv, nv);
      const int rf = r2 > 0 ? 0 : 1;
      tog += (uint32_t)tx2;
      // glitch scaling (TimeBasedPower.cc:303-339): the width to the flagged
      // neighbours of a flagged event (event 0 is never flagged)
      const uint32_t pwp = (lws & EV_GLITCH) ? t - ev_time(lws) : 0u;
      const uint32_t pwn = (nws & EV_GLITCH) ? ev_time(nws) - t : 0u;
      const float r = (float)max(pwp, pwn) * gks;
      const float ratio = fminf(1.f, r * r);
      const bool gl = (ws & EV_GLITCH) && !((clk >> ps) & 1u) && ratio > 0.f;
      glc += gl ? 1u : 0u;

      // input pin: its table at its own edge
      float En_in = rf ? E[1] : E[0];
#pragma unroll
      for (int p = 1; p < NIN; ++p)
        if (p == ps) En_in = rf ? E[2 * p + 1] : E[2 * p];
      // output pin: the related input's table, the related input being the
      // latest event before t in the same cycle whose arc-delayed arrival is
      // closest to t (TimeBasedPower.cc:412-473).  Evaluated on every step
      // (no branch) and selected.
      int32_t best = INT32_MAX;  // every same-cycle candidate is closer than t_max
      float En_out = rf ? E[6 * NIN + 1] : E[6 * NIN];  // default when none qualifies
#pragma unroll
      for (int q = 0; q < NIN; ++q) {
        if (!((arcm >> q) & 1u)) continue;  // no gate of this cell has an arc from q
        const bool hi_rf = (rflm >> q) & 1u, lo_rf = rf != 0;
        const int32_t d01 = lo_rf ? D[q * 4 + 1] : D[q * 4 + 0];
        const int32_t d23 = lo_rf ? D[q * 4 + 3] : D[q * 4 + 2];
        const int32_t d = hi_rf ? d23 : d01;
        const float e01 = lo_rf ? E[2 * NIN + q * 4 + 1] : E[2 * NIN + q * 4 + 0];
        const float e23 = lo_rf ? E[2 * NIN + q * 4 + 3] : E[2 * NIN + q * 4 + 2];
        const float eq = hi_rf ? e23 : e01;
        const uint32_t ltq = ev_time(lw[q]);
        // (an input without events has no arcs here; before its net's first
        // event lw carries that event's time, a candidate only in this cycle)
        const bool ok = d != ARC_NONE && ltq >= cstart && ltq < cend;
        // within one cycle t - lt < T < 2^15, so the grid difference fits 32 bits
        int32_t diff = (int32_t)((t - ltq) << ARC_FRAC_BITS) - d;
        diff = diff < 0 ? -diff : diff;
        const bool take = ok && diff < best;
        best = take ? diff : best;
        En_out = take ? eq : En_out;
      }
      const float En = is_out ? En_out : En_in;
      // switching: rising output edges
      {
        const float es = is_out ? 0.5f * (float)r2 * csw : 0.f;
        const long long qs = qenergy(gl ? ratio * es : es, inv_qe);
        acc[PG_SW] += qs;  // all switching; the glitch part is taken out after the loop
        acc[PG_GSW] += gl ? qs : 0;
        atomicAdd((unsigned long long*)&cb[(gl ? BIN_GSW : BIN_SW) * TCP], (unsigned long long)qs);
      }
      const float e = En * (0.5f * (float)tx2);
      const long long qi = qenergy(gl ? ratio * e : e, inv_qe);
      acc[PG_INT] += qi;  // all internal; the glitch part is taken out after the loop
      acc[PG_GINT] += gl ? qi : 0;
      atomicAdd((unsigned long long*)&cb[(gl ? BIN_GINT : BIN_INT) * TCP], (unsigned long long)qi);

      // consume the event of pin ps: the next word becomes current, and the
      // word after it comes from q2/q3 or the quad after; entering a new quad
      // moves that quad's words 2 and 3 into q2/q3
      const uint32_t* cps = cp[0];
      int32_t rs = rem[0];
      uint32_t sq2 = q2[0], sq3 = q3[0], sn0 = n0[0], sn1 = n1[0];
#pragma unroll
      for (int p = 1; p < NP; ++p)
        if (p == ps) {
          cps = cp[p];
          rs = rem[p];
          sq2 = q2[p];
          sq3 = q3[p];
          sn0 = n0[p];
          sn1 = n1[p];
        }
      const uint32_t* cn1 = cps + 1;
      const int32_t rn = rs - 1;
      const int kn = (int)(((uintptr_t)cn1 >> 2) & 3u);  // position of the new current event in its quad
      const bool enter = kn == 0;
      const uint32_t nx = kn == 1 ? sq2 : kn == 2 ? sq3 : kn == 3 ? sn0 : sn1;
      const uint32_t nwn = rn >= 2 ? nx : EV_INF;
#pragma unroll
      for (int p = 0; p < NP; ++p)
        if (p == ps) {
          cp[p] = cn1;
          rem[p] = rn;
          cw[p] = nws;
          nw[p] = nwn;
          lw[p] = ws;
          if (enter) {
            q2[p] = n2[p];
            q3[p] = n3[p];
          }
        }
      rflm = (rflm & ~(1u << ps)) | ((fs ? 1u : (uint32_t)rf) << ps);
      firstm &= ~(1u << ps);
      prev_t = t;
      // the quad after a newly entered quad, when the net has events there;
      // every read of n0..n3 above precedes this load
      asm volatile("" ::: "memory");
      if (enter && rn > 4) {
#pragma unroll
        for (int p = 0; p < NP; ++p)
          if (p == ps) {
            const uint4 v = *reinterpret_cast<const uint4*>(cp[p] + 4);
            n0[p] = v.x;
            n1[p] = v.y;
            n2[p] = v.z;
            n3[p] = v.w;
          }
      }
    }
    if constexpr (!CLKB) {
      if (lend > prev_t) acc[PG_LEAK] += (long long)(int)lq_cur * (long long)(lend - prev_t);
    } else {
      // to the tile end, or to the latest event consumed (an edge included)
      uint32_t endp = max(lend, prev_t), H;
      if (bulk) {
        const unsigned long long sn = ck_wh[ck_ne];
        const uint32_t wn = (uint32_t)sn;
        endp = max(endp, ev_time(wn));
        H = (uint32_t)(sn >> 32) + ((wn & 3u) == 1u ? endp - ev_time(wn) : 0u);
        // the edges' internal energy and toggles, per edge kind
        // (the clock pin's own table at its edge, as for any input event)
        const float Er = E[2 * KC], Ef = E[2 * KC + 1];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int tx = (k & 1) ? 2 : 1;
          const float En = (k >> 1) ? Ef : Er;
          const long long q = qenergy(En * (0.5f * (float)tx), inv_qe);
          const int n = ck_n[k];
          acc[PG_INT] += (long long)n * q;
          tog += (uint32_t)(n * tx);
          atomicAdd((unsigned long long*)&ck_q[k], (unsigned long long)q);
        }
      } else {
        H = h_prev + (((sb >> KC) & 1u) ? endp - prev_t : 0u);
      }
      const uint32_t dH = H - h_prev, dt = endp - prev_t;
      acc[PG_LEAK] += (long long)(int)lq_hi * (long long)dH + (long long)(int)lq_lo * (long long)(dt - dH);
    }
    acc[PG_SW] -= acc[PG_GSW];
    acc[PG_INT] -= acc[PG_GINT];
    const int64_t slot = (int64_t)wt * a.n_sched + gi;
    const int64_t plane = (int64_t)a.n_work_tiles * a.n_sched;
#pragma unroll
    for (int k = 0; k < N_PG; ++k) a.pg[k * plane + slot] = acc[k];
    a.pg_tog[slot] = tog;
    a.pg_glitch[slot] = glc;
  }
  for (int off = 32; off > 0; off >>= 1) base_acc += __shfl_xor(base_acc, off);
  if ((threadIdx.x & 63) == 0) atomicAdd((unsigned long long*)&s_base, (unsigned long long)base_acc);
  __syncthreads();
  if constexpr (CLKB) {
    // the item's clock edges into copy 0 of the bins: per edge, the leakage
    // steps of the bulk gates in each non-clock state and the summed edge
    // energies (the terms each gate would add at the edge)
    if (ck_ne >= 0) {
      block_scan_inplace(ck_d, ck_ne, ck_t);
      for (int sl = 1 + (int)threadIdx.x; sl <= ck_ne; sl += blockDim.x) {
        const uint32_t w = (uint32_t)ck_wh[sl];
        const int pv = ev_val((uint32_t)ck_wh[sl - 1]), nv = ev_val(w);
        const uint32_t t = ev_time(w);
        const uint32_t c = cycle_mul(t, a.cyc_m, a.cyc_sh);
        const int cr = min((int)((int64_t)c - tc0), TC);
        const int tx2 = toggle_x2(pv, nv), r2 = rise_x2(pv, nv);
        if (tx2 > 0)
          atomicAdd((unsigned long long*)&bins[BIN_INT * TCP + cr],
                    (unsigned long long)ck_q[(r2 > 0 ? 0 : 2) + (tx2 == 2 ? 1 : 0)]);
        const uint32_t bo = pv == 1 ? 1u : 0u, bn = nv == 1 ? 1u : 0u;
        if (bo != bn) {
          const unsigned long long cnt = ck_d[sl];
          long long ld = 0;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const long long n = (long long)((cnt >> (16 * j)) & 0xFFFFull);
            const uint32_t so = ck_unpack<KC>((uint32_t)j, bo), sn = ck_unpack<KC>((uint32_t)j, bn);
            ld += n * ((s_leakq[sn] >> 32) - (s_leakq[so] >> 32));
          }
          // the gates' steps dL each, and dL * (t - cycle start) in total
          atomicAdd((unsigned long long*)&bins[BIN_LD * TCP + cr], (unsigned long long)ld);
          atomicAdd((unsigned long long*)&bins[BIN_LX * TCP + cr],
                    (unsigned long long)(ld * (long long)(t - c * T)));
        }
      }
      __syncthreads();
    }
  }
  // fold the bin copies into copy 0, then S_c = base + earlier leakage steps
  for (int i = threadIdx.x; i < BW; i += blockDim.x) {
    long long v = bins[i];
    for (int k = 1; k < n_copy; ++k) v += bins[k * BW + i];
    bins[i] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    leak_flush(bins + BIN_LD * TCP, bins + BIN_LX * TCP, TC, s_base, tc0, (int64_t)T, wk.t_max, wk.n_cycle);
  __syncthreads();
  long long* out = a.cyc_part + ((int64_t)wt * a.n_range + range) * N_BIN * TC;
  for (int i = threadIdx.x; i < N_BIN * TC; i += blockDim.x) {
    const int b = i / TC;
    out[i] = bins[b * TCP + (i - b * TC)];
  }
  __syncthreads();  // the bins are cleared for the next item
  }
}

template __global__ void power_tile_kernel<MAXP_SMALL>(PowerArgs, int32_t, int32_t, unsigned int*);
template __global__ void power_tile_kernel<MAXP_LARGE>(PowerArgs, int32_t, int32_t, unsigned int*);
template __global__ void power_tile_kernel<MAXP_WIDE>(PowerArgs, int32_t, int32_t, unsigned int*);
template __global__ void power_tile_kernel<MAXP_HUGE>(PowerArgs, int32_t, int32_t, unsigned int*);

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
__global__ void cycle_reduce_kernel(const long long* __restrict__ part, int32_t n_range,
                                    int32_t tile_cycles, int64_t n_cycle, int64_t T, int64_t t_max,
                                    double q_energy, double q_leak, double inv_period,
                                    double* __restrict__ out /* [5][C] */) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cycle) return;
  const int64_t tile = c / tile_cycles;
  const int ci = (int)(c - tile * tile_cycles);
  long long s[N_BIN] = {0, 0, 0, 0, 0, 0};
  for (int r = 0; r < n_range; ++r) {
    const long long* p = part + ((tile * n_range + r) * N_BIN) * tile_cycles + ci;
#pragma unroll
    for (int k = 0; k < N_BIN; ++k) s[k] += p[(int64_t)k * tile_cycles];
  }
  const int64_t cend = min((c + 1) * T, t_max);
  const int64_t len = cend - c * T;
  out[0 * n_cycle + c] = ((double)s[BIN_LD] * (double)len + (double)s[BIN_LX]) * (q_leak / (double)T);
  out[1 * n_cycle + c] = (double)(s[BIN_INT] + s[BIN_GINT]) * q_energy * inv_period;
  out[2 * n_cycle + c] = (double)s[BIN_GINT] * q_energy * inv_period;
  out[3 * n_cycle + c] = (double)(s[BIN_SW] + s[BIN_GSW]) * q_energy * inv_period;
  out[4 * n_cycle + c] = (double)s[BIN_GSW] * q_energy * inv_period;
}

// cycle_reduce_kernel over the tiles of a list only (partials indexed by list
// position): output column l * tile_cycles + i is cycle
// tile_list[l] * tile_cycles + i (0 past the end)
// The tile lists of a judge are short (about a thousand output cycles), so
// each output cycle's sum over the gate ranges is split over RL_GROUPS
// threads (exact int64 partial sums, combined in LDS).
constexpr int RL_CYCLES = 32, RL_GROUPS = 8;  // a block: 32 output cycles x 8 range groups

__global__ __launch_bounds__(RL_CYCLES * RL_GROUPS) void cycle_reduce_list_kernel(
    const long long* __restrict__ part, int32_t n_range, int32_t tile_cycles, const int32_t* __restrict__ tile_list,
    int32_t n_list, int64_t n_cycle, int64_t T, int64_t t_max, double q_energy, double q_leak, double inv_period,
    double* __restrict__ out /* [5][n_list * tile_cycles] */) {
  __shared__ long long s_part[RL_GROUPS][N_BIN][RL_CYCLES];
  const int cl = threadIdx.x % RL_CYCLES, grp = threadIdx.x / RL_CYCLES;
  const int64_t j = (int64_t)blockIdx.x * RL_CYCLES + cl;
  const int64_t n_out = (int64_t)n_list * tile_cycles;
  const bool live = j < n_out;
  const int64_t l = live ? j / tile_cycles : 0;
  const int ci = live ? (int)(j % tile_cycles) : 0;
  const int64_t c = (int64_t)tile_list[l] * tile_cycles + ci;
  long long sk[N_BIN] = {0, 0, 0, 0, 0, 0};
  if (live && c < n_cycle) {
    for (int r = grp; r < n_range; r += RL_GROUPS) {
      const long long* p = part + ((l * n_range + r) * N_BIN) * tile_cycles + ci;
#pragma unroll
      for (int k = 0; k < N_BIN; ++k) sk[k] += p[(int64_t)k * tile_cycles];
    }
  }
#pragma unroll
  for (int k = 0; k < N_BIN; ++k) s_part[grp][k][cl] = sk[k];
  __syncthreads();
  if (grp != 0 || !live) return;
  if (c >= n_cycle) {
#pragma unroll
    for (int k = 0; k < 5; ++k) out[k * n_out + j] = 0.0;
    return;
  }
  long long sum[N_BIN];
#pragma unroll
  for (int k = 0; k < N_BIN; ++k) {
    sum[k] = 0;
#pragma unroll
    for (int g = 0; g < RL_GROUPS; ++g) sum[k] += s_part[g][k][cl];
  }
  const int64_t len = min((c + 1) * T, t_max) - c * T;
  out[0 * n_out + j] = ((double)sum[BIN_LD] * (double)len + (double)sum[BIN_LX]) * (q_leak / (double)T);
  out[1 * n_out + j] = (double)(sum[BIN_INT] + sum[BIN_GINT]) * q_energy * inv_period;
  out[2 * n_out + j] = (double)sum[BIN_GINT] * q_energy * inv_period;
  out[3 * n_out + j] = (double)(sum[BIN_SW] + sum[BIN_GSW]) * q_energy * inv_period;
  out[4 * n_out + j] = (double)sum[BIN_GSW] * q_energy * inv_period;
}

// Batched runs: per-cycle outputs of the work tiles of one round.  Column
// j = wt * tile_cycles + i is cycle c = tile_list[wt] * tile_cycles + i of trace
// wt_trace[wt] (relative to its load's origin), written to out + out_base[trace]
// as [5][out_cycles], at column c + cycle_base.
__global__ void cycle_reduce_batch_kernel(const long long* __restrict__ part, int32_t n_range,
                                          int32_t tile_cycles, const int32_t* __restrict__ tile_list,
                                          const int32_t* __restrict__ wt_trace, int32_t n_wt,
                                          const TraceDesc* __restrict__ traces,
                                          const int64_t* __restrict__ out_base, int64_t T, double q_energy,
                                          double q_leak, double inv_period, double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= (int64_t)n_wt * tile_cycles) return;
  const int64_t l = j / tile_cycles;
  const int ci = (int)(j % tile_cycles);
  const int tr = wt_trace[l];
  const TraceDesc td = traces[tr];
  const int64_t c = (int64_t)tile_list[l] * tile_cycles + ci;
  if (c >= td.n_cycle) return;  // (relative cycle)
  long long s[N_BIN] = {0, 0, 0, 0, 0, 0};
  for (int r = 0; r < n_range; ++r) {
    const long long* p = part + ((l * n_range + r) * N_BIN) * tile_cycles + ci;
#pragma unroll
    for (int k = 0; k < N_BIN; ++k) s[k] += p[(int64_t)k * tile_cycles];
  }
  const int64_t len = min((c + 1) * T, td.t_max) - c * T;
  double* o = out + out_base[tr] + td.cycle_base;
  const int64_t C = td.out_cycles;
  o[0 * C + c] = ((double)s[BIN_LD] * (double)len + (double)s[BIN_LX]) * (q_leak / (double)T);
  o[1 * C + c] = (double)(s[BIN_INT] + s[BIN_GINT]) * q_energy * inv_period;
  o[2 * C + c] = (double)s[BIN_GINT] * q_energy * inv_period;
  o[3 * C + c] = (double)(s[BIN_SW] + s[BIN_GSW]) * q_energy * inv_period;
  o[4 * C + c] = (double)s[BIN_GSW] * q_energy * inv_period;
}

// Batched runs: add one round's per-gate partials into exact per-trace sums.
// Work tiles of a trace are contiguous in the round: [wt_begin[k], wt_begin[k+1])
// for the round's k-th trace round_trace[k]; grid.y = traces in the round.
__global__ void gate_accumulate_kernel(const long long* __restrict__ pg, const uint32_t* __restrict__ tog,
                                       const uint32_t* __restrict__ glc, const int32_t* __restrict__ order,
                                       int64_t n_sched, int32_t n_wt, const int32_t* __restrict__ wt_begin,
                                       const int32_t* __restrict__ round_trace, int64_t n_gate,
                                       long long* __restrict__ acc /* [trace][N_PG][G] */,
                                       unsigned long long* __restrict__ acc_cnt /* [trace][2][G] */) {
  const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= n_sched) return;
  const int g = order[gi];
  if (g < 0) return;
  const int k = blockIdx.y;
  const int tr = round_trace[k];
  const int64_t plane = (int64_t)n_wt * n_sched;
  long long v[N_PG] = {0, 0, 0, 0, 0};
  unsigned long long tg = 0, gg = 0;
  for (int t = wt_begin[k]; t < wt_begin[k + 1]; ++t) {
#pragma unroll
    for (int c = 0; c < N_PG; ++c) v[c] += pg[c * plane + (int64_t)t * n_sched + gi];
    tg += tog[(int64_t)t * n_sched + gi];
    gg += glc[(int64_t)t * n_sched + gi];
  }
  long long* A = acc + (int64_t)tr * N_PG * n_gate;
#pragma unroll
  for (int c = 0; c < N_PG; ++c) A[(int64_t)c * n_gate + g] += v[c];
  acc_cnt[((int64_t)tr * 2 + 0) * n_gate + g] += tg;
  acc_cnt[((int64_t)tr * 2 + 1) * n_gate + g] += gg;
}

// Batched runs: per-gate power of one trace from its exact sums.
__global__ void gate_finalize_kernel(const long long* __restrict__ acc, int64_t n_gate, double norm_energy,
                                     double norm_leak, const int32_t* __restrict__ gate_cell,
                                     const CellDev* __restrict__ cells, double* __restrict__ gate_power) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)N_PG * n_gate) return;
  // leakage sums count the gate's cell's own quanta
  const double norm = i / n_gate == PG_LEAK ? ldexp(norm_leak, -cells[gate_cell[i % n_gate]].leak_shift) : norm_energy;
  gate_power[i] = (double)acc[i] * norm;
}

__global__ void fill_u64_kernel(unsigned long long* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void fill_u32_zero_kernel(unsigned int* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void gate_reduce_kernel(const long long* __restrict__ pg, const uint32_t* __restrict__ tog,
                                   const uint32_t* __restrict__ glc, const int32_t* __restrict__ order,
                                   int64_t n_sched, int32_t n_tile, int64_t n_gate, double norm_energy,
                                   double norm_leak, const int32_t* __restrict__ gate_cell,
                                   const CellDev* __restrict__ cells, double* __restrict__ gate_power,
                                   unsigned long long* __restrict__ gate_tog,
                                   unsigned long long* __restrict__ gate_glitch) {
  // one thread per (schedule position, array): the N_PG sums, then toggles and
  // glitches (blockIdx.y), each over the tiles, 16 loads in flight into four
  // independent accumulators (integer sums: the order does not matter)
  const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= n_sched) return;
  const int g = order[gi];
  if (g < 0) return;
  const int k = (int)blockIdx.y;
  const int64_t plane = (int64_t)n_tile * n_sched;
  if (k < N_PG) {
    const long long* p = pg + k * plane + gi;
    long long v0 = 0, v1 = 0, v2 = 0, v3 = 0;
    int t = 0;
    for (; t + 16 <= n_tile; t += 16) {
      long long x[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) x[u] = p[(int64_t)(t + u) * n_sched];
#pragma unroll
      for (int u = 0; u < 16; u += 4) {
        v0 += x[u];
        v1 += x[u + 1];
        v2 += x[u + 2];
        v3 += x[u + 3];
      }
    }
    for (; t < n_tile; ++t) v0 += p[(int64_t)t * n_sched];
    const long long v = (v0 + v1) + (v2 + v3);
    // leakage sums count the gate's cell's own quanta
    const double norm = k == PG_LEAK ? ldexp(norm_leak, -cells[gate_cell[g]].leak_shift) : norm_energy;
    gate_power[(int64_t)k * n_gate + g] = (double)v * norm;
  } else {
    const uint32_t* p = (k == N_PG ? tog : glc) + gi;
    unsigned long long v0 = 0, v1 = 0, v2 = 0, v3 = 0;
    int t = 0;
    for (; t + 16 <= n_tile; t += 16) {
      uint32_t x[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) x[u] = p[(int64_t)(t + u) * n_sched];
#pragma unroll
      for (int u = 0; u < 16; u += 4) {
        v0 += x[u];
        v1 += x[u + 1];
        v2 += x[u + 2];
        v3 += x[u + 3];
      }
    }
    for (; t < n_tile; ++t) v0 += p[(int64_t)t * n_sched];
    (k == N_PG ? gate_tog : gate_glitch)[g] = (v0 + v1) + (v2 + v3);
  }
}

// FP64 design and class totals: per-block partials, then one fixed-order pass.
__global__ void total_partial_kernel(const double* __restrict__ gate_power, const int32_t* __restrict__ gate_class,
                                     int64_t n_gate, double* __restrict__ part /* [blocks][30] */) {
  __shared__ double s[256];
  double v[30];
#pragma unroll
  for (int k = 0; k < 30; ++k) v[k] = 0.0;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n_gate; g += (int64_t)gridDim.x * blockDim.x) {
    const int cls = gate_class[g];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double x = gate_power[(int64_t)k * n_gate + g];
      v[k] += x;
#pragma unroll
      for (int c = 0; c < 5; ++c)
        if (cls == c) v[5 + c * 5 + k] += x;
    }
  }
  for (int k = 0; k < 30; ++k) {
    s[threadIdx.x] = v[k];
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x * 30 + k] = s[0];
    __syncthreads();
  }
}

__global__ void total_final_kernel(const double* __restrict__ part, int n_part, double* __restrict__ out) {
  if (threadIdx.x >= 30) return;
  double v = 0.0;
  for (int b = 0; b < n_part; ++b) v += part[b * 30 + threadIdx.x];
  out[threadIdx.x] = v;
}

// ---------------------------------------------------------------------------
// Host-side launch wrappers (the API translation unit launches through these)
// ---------------------------------------------------------------------------
hipError_t launch_pack(const int64_t* t, const uint8_t* v, uint32_t* out, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  pack_events_kernel<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(t, v, out, n);
  return hipGetLastError();
}

hipError_t launch_tile_index(const int64_t* net_off, const uint32_t* ev, int64_t n_net, int32_t n_tile,
                             int64_t tile_ticks, const int32_t* tile_list, int32_t n_list, int32_t* out,
                             hipStream_t s) {
  const int64_t n = n_net * (tile_list ? (int64_t)n_list : (int64_t)n_tile + 1);
  if (n <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  tile_index_kernel<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(net_off, ev, n_net, n_tile, tile_ticks, tile_list,
                                                                n_list, out);
  return hipGetLastError();
}

hipError_t launch_tile_order(const int32_t* order, int64_t n_sched, int32_t range_gates, int32_t lanes,
                             const int64_t* gate_pin_off, const int32_t* pin_net, const int32_t* tile_idx,
                             int32_t n_tile, int32_t n_wt, const int32_t* wt_tile, const int32_t* wt_trace,
                             const TraceDesc* traces, int32_t* perm, hipStream_t s) {
  if (n_sched <= 0 || n_wt <= 0) return hipSuccess;
  if (range_gates <= 0 || range_gates > TO_MAX || n_sched % range_gates != 0) return hipErrorInvalidValue;
  const int64_t n_range = n_sched / range_gates;
  const int64_t n_item = n_range * n_wt;
  tile_order_kernel<<<dim3((unsigned)std::min<int64_t>(n_item, 1 << 20)), dim3(256), 0, s>>>(
      order, n_sched, range_gates, lanes, gate_pin_off, pin_net, tile_idx, n_tile, n_wt, wt_tile, wt_trace, traces,
      perm);
  return hipGetLastError();
}

hipError_t launch_energy(const EnergyArgs& a, hipStream_t s) {
  if (a.n_gate <= 0) return hipSuccess;
  const int64_t blocks = (a.n_gate + 255) / 256;
  const size_t lds = a.n_table_data <= TABLE_LDS_MAX ? sizeof(float) * a.n_table_data : 0;
  energy_kernel<<<dim3((unsigned)blocks), dim3(256), lds, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_power(int maxp, int64_t grid, int block, size_t lds, hipStream_t s, const PowerArgs& a,
                        int32_t range_begin, int32_t range_count, unsigned int* wq) {
  if (grid <= 0) return hipSuccess;
  if (maxp <= MAXP_SMALL)
    power_tile_kernel<MAXP_SMALL><<<dim3((unsigned)grid), dim3(block), lds, s>>>(a, range_begin, range_count, wq);
  else if (maxp <= MAXP_LARGE)
    power_tile_kernel<MAXP_LARGE><<<dim3((unsigned)grid), dim3(block), lds, s>>>(a, range_begin, range_count, wq);
  else if (maxp <= MAXP_WIDE)
    power_tile_kernel<MAXP_WIDE><<<dim3((unsigned)grid), dim3(block), lds, s>>>(a, range_begin, range_count, wq);
  else if (maxp <= MAXP_HUGE)
    power_tile_kernel<MAXP_HUGE><<<dim3((unsigned)grid), dim3(block), lds, s>>>(a, range_begin, range_count, wq);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_cycle_reduce(const long long* part, int32_t n_range, int32_t tile_cycles, int64_t n_cycle,
                               int64_t T, int64_t t_max, double q_energy, double q_leak, double inv_period,
                               double* out, hipStream_t s) {
  if (n_cycle <= 0) return hipSuccess;
  cycle_reduce_kernel<<<dim3((unsigned)((n_cycle + 255) / 256)), dim3(256), 0, s>>>(
      part, n_range, tile_cycles, n_cycle, T, t_max, q_energy, q_leak, inv_period, out);
  return hipGetLastError();
}

hipError_t launch_cycle_reduce_list(const long long* part, int32_t n_range, int32_t tile_cycles,
                                    const int32_t* tile_list, int32_t n_list, int64_t n_cycle, int64_t T,
                                    int64_t t_max, double q_energy, double q_leak, double inv_period,
                                    double* out, hipStream_t s) {
  const int64_t n = (int64_t)n_list * tile_cycles;
  if (n <= 0) return hipSuccess;
  cycle_reduce_list_kernel<<<dim3((unsigned)((n + RL_CYCLES - 1) / RL_CYCLES)), dim3(RL_CYCLES * RL_GROUPS), 0, s>>>(
      part, n_range, tile_cycles, tile_list, n_list, n_cycle, T, t_max, q_energy, q_leak, inv_period, out);
  return hipGetLastError();
}

hipError_t launch_cycle_reduce_batch(const long long* part, int32_t n_range, int32_t tile_cycles,
                                     const int32_t* tile_list, const int32_t* wt_trace, int32_t n_wt,
                                     const TraceDesc* traces, const int64_t* out_base, int64_t T, double q_energy,
                                     double q_leak, double inv_period, double* out, hipStream_t s) {
  const int64_t n = (int64_t)n_wt * tile_cycles;
  if (n <= 0) return hipSuccess;
  cycle_reduce_batch_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(
      part, n_range, tile_cycles, tile_list, wt_trace, n_wt, traces, out_base, T, q_energy, q_leak, inv_period, out);
  return hipGetLastError();
}

hipError_t launch_gate_accumulate(const long long* pg, const uint32_t* tog, const uint32_t* glc,
                                  const int32_t* order, int64_t n_sched, int32_t n_wt, const int32_t* wt_begin,
                                  const int32_t* round_trace, int32_t n_round_trace, int64_t n_gate, long long* acc,
                                  unsigned long long* acc_cnt, hipStream_t s) {
  if (n_sched <= 0 || n_round_trace <= 0) return hipSuccess;
  gate_accumulate_kernel<<<dim3((unsigned)((n_sched + 255) / 256), (unsigned)n_round_trace), dim3(256), 0, s>>>(
      pg, tog, glc, order, n_sched, n_wt, wt_begin, round_trace, n_gate, acc, acc_cnt);
  return hipGetLastError();
}

hipError_t launch_gate_finalize(const long long* acc, int64_t n_gate, double norm_energy, double norm_leak,
                                const int32_t* gate_cell, const CellDev* cells, double* gate_power, hipStream_t s) {
  const int64_t n = (int64_t)N_PG * n_gate;
  if (n <= 0) return hipSuccess;
  gate_finalize_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(acc, n_gate, norm_energy, norm_leak,
                                                                              gate_cell, cells, gate_power);
  return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long* p, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  fill_u64_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(p, n);
  return hipGetLastError();
}

hipError_t launch_zero_u32(unsigned int* p, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  fill_u32_zero_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(p, n);
  return hipGetLastError();
}

hipError_t launch_gate_reduce(const long long* pg, const uint32_t* tog, const uint32_t* glc,
                              const int32_t* order, int64_t n_sched, int32_t n_tile, int64_t n_gate,
                              double norm_energy, double norm_leak, const int32_t* gate_cell,
                              const CellDev* cells, double* gate_power, unsigned long long* gate_tog,
                              unsigned long long* gate_glitch, hipStream_t s) {
  if (n_sched <= 0) return hipSuccess;
  gate_reduce_kernel<<<dim3((unsigned)((n_sched + 255) / 256), N_PG + 2), dim3(256), 0, s>>>(
      pg, tog, glc, order, n_sched, n_tile, n_gate, norm_energy, norm_leak, gate_cell, cells, gate_power,
      gate_tog, gate_glitch);
  return hipGetLastError();
}

hipError_t launch_total(const double* gate_power, const int32_t* gate_class, int64_t n_gate, double* part,
                        double* out, hipStream_t s) {
  total_partial_kernel<<<dim3(TOTAL_BLOCKS), dim3(256), 0, s>>>(gate_power, gate_class, n_gate, part);
  total_final_kernel<<<dim3(1), dim3(64), 0, s>>>(part, TOTAL_BLOCKS, out);
  return hipGetLastError();
}

hipError_t launch_chunk_net(const int64_t* chunk_off, int64_t n_net, int32_t* chunk_net, hipStream_t s) {
  if (n_net <= 0) return hipSuccess;
  chunk_net_kernel<<<dim3((unsigned)((n_net + 255) / 256)), dim3(256), 0, s>>>(chunk_off, n_net, chunk_net);
  return hipGetLastError();
}

hipError_t launch_glitch_annotate(const int64_t* net_off, uint32_t* ev, const int64_t* chunk_off,
                                  const int32_t* chunk_net, int64_t n_chunk, uint32_t T, hipStream_t s) {
  if (n_chunk <= 0) return hipSuccess;
  // the persistent grid: as many blocks as the device holds at once
  int dev = 0, n_cu = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, glitch_annotate_kernel, 256, 0);
  if (e != hipSuccess) return e;
  const int64_t resident = std::max<int64_t>(1, (int64_t)n_cu * std::max(per_cu, 1));
  const int64_t blocks = stride_blocks(std::min<int64_t>((n_chunk + 3) / 4, resident));
  uint32_t cyc_m = 0;
  int32_t cyc_sh = 0;
  cycle_magic(T, &cyc_m, &cyc_sh);
  glitch_annotate_kernel<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(net_off, ev, chunk_off, chunk_net, n_chunk, cyc_m,
                                                                     cyc_sh);
  return hipGetLastError();
}

hipError_t launch_glitch_annotate_tiles(const int64_t* net_off, uint32_t* ev, const int32_t* idx_begin,
                                        const int32_t* tile_list, int32_t n_list, int64_t n_net, int64_t tile_ticks,
                                        uint32_t T, hipStream_t s) {
  const int64_t n = n_net * n_list;
  if (n <= 0) return hipSuccess;
  uint32_t cyc_m = 0;
  int32_t cyc_sh = 0;
  cycle_magic(T, &cyc_m, &cyc_sh);
  glitch_annotate_tiles_kernel<<<dim3((unsigned)stride_blocks((n + 255) / 256)), dim3(256), 0, s>>>(
      net_off, ev, idx_begin, tile_list, n_list, n_net, tile_ticks, T, cyc_m, cyc_sh);
  return hipGetLastError();
}

hipError_t launch_power_cell(int np, int kc, int64_t grid, int block, size_t lds, hipStream_t s, const PowerArgs& a,
                             int32_t range_begin, int32_t range_count, int32_t cell, unsigned int* wq) {
  if (grid <= 0) return hipSuccess;
  const dim3 g((unsigned)grid), b(block);
  if (kc >= 0) {
    if (np != 3 || kc != 0) return hipErrorInvalidValue;
    power_cell_kernel<3, 0><<<g, b, lds, s>>>(a, range_begin, range_count, cell, wq);
    return hipGetLastError();
  }
  switch (np) {
    case 2: power_cell_kernel<2, -1><<<g, b, lds, s>>>(a, range_begin, range_count, cell, wq); break;
    case 3: power_cell_kernel<3, -1><<<g, b, lds, s>>>(a, range_begin, range_count, cell, wq); break;
    case 4: power_cell_kernel<4, -1><<<g, b, lds, s>>>(a, range_begin, range_count, cell, wq); break;
    case 5: power_cell_kernel<5, -1><<<g, b, lds, s>>>(a, range_begin, range_count, cell, wq); break;
    case 6: power_cell_kernel<6, -1><<<g, b, lds, s>>>(a, range_begin, range_count, cell, wq); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace gtp
