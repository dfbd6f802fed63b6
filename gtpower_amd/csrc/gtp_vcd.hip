// Device VCD ingest: VCD text -> the per-(id, bit) event store of the power
// engine, built on the GPU.
//
// Replaces the reference's single-threaded reader VcdReader::read /
// parseVarValues (power/VcdReader.cc:80-131, :238-285) with
// Vcd::varAppendValue / varAppendBusValue (power/Vcd.cc:203-232), and the host
// phases that cut per-bit event ranges out of it and copy them into the
// device event store (power/cuda/CudaPower.cu:1874-2086, value table
// :1129-1145).  The header (declarations, a few KB to MB) is read on the host
// with the reference's token rules; the value-change section, which is nearly
// all of the text, is parsed on the device:
//
//   token starts     whitespace -> non-whitespace transitions, counted per
//                    4 KB block, block offsets scanned, starts written
//   token parse      kind (time / scalar / bus / other), time and bus values
//                    (stoll / strtol base 2 semantics, saturating)
//   bus ids          a bus token takes the next token as its id, so which
//                    tokens are ids follows from the parity of each run of
//                    bus-looking tokens (a max-scan gives run starts)
//   times            each change takes the latest time token before it that
//                    is not an id (max-scan over token indices)
//   changes          compacted in text order; ids looked up in a device hash
//                    table of the header's ids (exact byte compare)
//   grouping         stable radix sort of changes by id (text order kept),
//                    carrying each change's packed time, kind and value
//   dedupe           a change equal to the previous change of its id is
//                    dropped (ids mixing scalar and bus values replay the
//                    reference's last-kept rule sequentially); an id whose
//                    first change is after 0 gets X (scalar) or 0 (bus) at 0
//   validation       kept times of an id strictly increasing and below
//                    2^29 - 1 ticks (max-scan of (id, time) keys)
//   store            per-stream counts scanned into offsets, every kept change
//                    scattered to the streams of its id's bits as 32-bit event
//                    words (time << 3 | value)
//
// A text longer than the part size is parsed in parts cut at whitespace (the
// reference reads its text as one stream of tokens, so the result does not
// depend on the cuts): each part runs the whole pipeline with the time in
// force and every id's last kept change carried in, and the parts' stores are
// appended per stream at the end.
//
// Streams are numbered per unique id in declaration order, bits 0..w-1 (w =
// the widest declaration of the id); bit b of a bus value is (value >> b) & 1.
// Value codes: 0 -> 0, 1 -> 1, X -> 2, Z -> 3 and U -> 2 (the reference's
// table leaves U unmapped).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gtp_internal.hpp"
#include "gtpower.h"

#define GTP_API extern "C" __attribute__((visibility("default")))

namespace gtp {
namespace vcd {

enum : uint8_t { TK_OTHER = 0, TK_TIME = 1, TK_SCALAR = 2, TK_BUS = 3, TK_TIME_BAD = 4 };
enum : int { GTP_VCD_ERR_TIME = 1, GTP_VCD_ERR_ID = 2, GTP_VCD_ERR_ORDER = 3, GTP_VCD_ERR_RANGE = 4 };

// the first error in text order: (item << 8) | code, ~0 when none
struct DevErr {
  unsigned long long packed;
};

__device__ __forceinline__ bool is_ws(uint8_t c) { return c == ' ' || (c >= 9 && c <= 13); }
__device__ __forceinline__ uint8_t upper(uint8_t c) { return (c >= 'a' && c <= 'z') ? (uint8_t)(c - 32) : c; }

__device__ void set_err(DevErr* e, int code, int64_t where) {
  atomicMin(&e->packed, ((unsigned long long)where << 8) | (unsigned long long)code);
}

constexpr int TB = 256;             // threads per block
constexpr int BPT = 16;             // bytes per thread
constexpr int BPB = TB * BPT;       // bytes per block

// token starts among the 16 bytes of one thread: bit j set when byte
// base + j starts a token (it is not whitespace, the byte before it is, and
// it lies in [b0, L)); the byte before b0 is whitespace by construction
__device__ __forceinline__ uint32_t start_bits(const uint8_t* t, int64_t base, int64_t b0, int64_t L) {
  const uint4 q = *reinterpret_cast<const uint4*>(t + base);
  const uint32_t w[4] = {q.x, q.y, q.z, q.w};
  uint8_t prev = base > 0 ? t[base - 1] : (uint8_t)' ';
  uint32_t bits = 0;
#pragma unroll
  for (int j = 0; j < BPT; ++j) {
    const uint8_t c = (uint8_t)(w[j >> 2] >> (8 * (j & 3)));
    const int64_t i = base + j;
    if (i >= b0 && i < L && !is_ws(c) && is_ws(prev)) bits |= 1u << j;
    prev = c;
  }
  return bits;
}

__device__ __forceinline__ int block_excl_sum(int v, int* lds, int* total) {
  // 256 threads, 4 waves of 64
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wv] = x;
  __syncthreads();
  int wo = 0;
  for (int k = 0; k < wv; ++k) wo += lds[k];
  *total = lds[0] + lds[1] + lds[2] + lds[3];
  __syncthreads();
  return wo + x - v;
}

__global__ __launch_bounds__(TB) void count_starts_kernel(const uint8_t* __restrict__ t, int64_t base0, int64_t b0,
                                                          int64_t L, uint64_t* __restrict__ cnt) {
  __shared__ int lds[4];
  const int64_t base = base0 + (int64_t)blockIdx.x * BPB + threadIdx.x * BPT;
  const int c = __popc(start_bits(t, base, b0, L));
  int total;
  (void)block_excl_sum(c, lds, &total);
  if (threadIdx.x == 0) cnt[blockIdx.x] = (uint64_t)total;
}

__global__ __launch_bounds__(TB) void write_starts_kernel(const uint8_t* __restrict__ t, int64_t base0, int64_t b0,
                                                          int64_t L, const uint64_t* __restrict__ blk_off,
                                                          uint32_t* __restrict__ tok, int64_t K) {
  __shared__ int lds[4];
  const int64_t base = base0 + (int64_t)blockIdx.x * BPB + threadIdx.x * BPT;
  uint32_t bits = start_bits(t, base, b0, L);
  int total;
  int64_t o = (int64_t)blk_off[blockIdx.x] + block_excl_sum(__popc(bits), lds, &total);
  while (bits) {
    const int j = __ffs(bits) - 1;
    bits &= bits - 1;
    if (o < K) tok[o] = (uint32_t)(base + j);  // the last token is dropped when the text ends inside it
    ++o;
  }
}

// Kind, length and value of every token (power/VcdReader.cc:245-278).
// Ids of 1 to 3 characters in [33, 126] index a direct table: key =
// sum of (c_i - 32) * 95^i (digits >= 1, so lengths do not collide).
constexpr int32_t SHORT_ID_KEYS = 95 * 95 * 95;
template <typename At>
__device__ __forceinline__ int32_t short_id_key(At at, int64_t s, int64_t n) {
  if (n < 1 || n > 3) return -1;
  int32_t k = 0, m = 1;
  for (int64_t i = 0; i < n; ++i) {
    const uint8_t c = at(s + i);
    if (c < 33 || c > 126) return -1;
    k += (c - 32) * m;
    m *= 95;
  }
  return k;
}

// One token [s, e), bytes read through at(p).
template <typename At>
__device__ __forceinline__ void parse_one(At at, int64_t k, int64_t s, int64_t e, uint8_t* __restrict__ kind,
                                          uint32_t* __restrict__ len, int64_t* __restrict__ val,
                                          int32_t* __restrict__ skey) {
  const int64_t n = e - s;
  const uint8_t c0 = upper(at(s));
  uint8_t kd = TK_OTHER;
  int64_t v = 0;
  if (c0 == '#' && n > 1) {
    // std::stoll: optional sign, decimal digits; no digit or overflow = error
    int64_t p = s + 1;
    bool neg = false;
    const uint8_t c1 = at(p);
    if (c1 == '+' || c1 == '-') {
      neg = c1 == '-';
      ++p;
    }
    uint64_t acc = 0;
    bool any = false, ovf = false;
    if (e - p <= 9) {
      // at most 9 digits: no overflow, 32-bit arithmetic
      uint32_t a32 = 0;
      for (; p < e; ++p) {
        const uint8_t c = at(p);
        if (c < '0' || c > '9') break;
        any = true;
        a32 = a32 * 10u + (uint32_t)(c - '0');
      }
      acc = a32;
    } else {
      // 10 acc + d > lim  <=>  acc > lim / 10 or (acc == lim / 10 and d > lim % 10)
      const uint64_t q = 922337203685477580ull;  // (2^63 - 1) / 10 = 2^63 / 10
      const uint64_t r = neg ? 8u : 7u;
      for (; p < e; ++p) {
        const uint8_t c = at(p);
        if (c < '0' || c > '9') break;
        any = true;
        const uint64_t d = c - '0';
        if (acc > q || (acc == q && d > r)) ovf = true;
        else acc = acc * 10 + d;
      }
    }
    if (!any || ovf) {
      kd = TK_TIME_BAD;
    } else {
      kd = TK_TIME;
      v = neg ? (int64_t)(0 - acc) : (int64_t)acc;
    }
  } else if (c0 == '0' || c0 == '1' || c0 == 'X' || c0 == 'U' || c0 == 'Z') {
    kd = TK_SCALAR;
    v = c0;
  } else if (c0 == 'B') {
    kd = TK_BUS;
    const uint8_t c1 = n > 1 ? upper(at(s + 1)) : 0;
    if (!(c1 == 'X' || c1 == 'U' || c1 == 'Z')) {
      // strtol(..., 2): optional sign, binary digits, saturating
      int64_t p = s + 1;
      bool neg = false;
      if (p < e && (at(p) == '+' || at(p) == '-')) {
        neg = at(p) == '-';
        ++p;
      }
      uint64_t acc = 0;
      bool ovf = false;
      if (e - p <= 31) {
        // at most 31 digits: no overflow, 32-bit arithmetic
        uint32_t a32 = 0;
        for (; p < e; ++p) {
          const uint8_t c = at(p);
          if (c != '0' && c != '1') break;
          a32 = (a32 << 1) | (uint32_t)(c - '0');
        }
        acc = a32;
      } else {
        // 2 acc + d > lim, as above: lim / 2 = 2^62 - 1 or 2^62, lim % 2 = 1 or 0
        const uint64_t q = neg ? (1ull << 62) : (1ull << 62) - 1;
        const uint64_t r = neg ? 0u : 1u;
        for (; p < e; ++p) {
          const uint8_t c = at(p);
          if (c != '0' && c != '1') break;
          const uint64_t d = c - '0';
          if (ovf || acc > q || (acc == q && d > r)) ovf = true;
          else acc = acc * 2 + d;
        }
      }
      if (ovf) v = neg ? INT64_MIN : INT64_MAX;
      else v = neg ? (int64_t)(0 - acc) : (int64_t)acc;
    }
  }
  kind[k] = kd;
  len[k] = (uint32_t)n;
  val[k] = v;
  // a scalar's id, else the whole token (the id when it follows a bus value)
  skey[k] = kd == TK_SCALAR ? short_id_key(at, s + 1, n - 1) : short_id_key(at, s, n);
}

// One workgroup per text block (the blocks of the token-start count): the
// block's BPB bytes and PT_HALO bytes after it are staged in LDS with 16-byte
// loads, and the tokens that start in the block are parsed from there (a byte
// outside the staged range is read from memory).  A token ends where the
// whitespace before the next token starts (only whitespace separates tokens):
// step back from the next start, usually over one separator; the last token
// scans forward.
constexpr int PT_HALO = 256;
constexpr int PT_STAGE = BPB + PT_HALO;

__global__ __launch_bounds__(TB) void parse_tokens_kernel(const uint8_t* __restrict__ t, int64_t L, int64_t t_bytes,
                                                          int64_t base0, const uint64_t* __restrict__ boff,
                                                          int64_t n_blk, const uint32_t* __restrict__ tok, int64_t K,
                                                          uint8_t* __restrict__ kind, uint32_t* __restrict__ len,
                                                          int64_t* __restrict__ val, int32_t* __restrict__ skey) {
  __shared__ __align__(16) uint8_t st[PT_STAGE];
  const int64_t b = blockIdx.x;
  const int64_t base = base0 + b * BPB;
  for (int i = threadIdx.x; i < PT_STAGE / 16; i += blockDim.x) {
    const int64_t p = base + 16 * (int64_t)i;
    // past the buffer: blanks, as the buffer's own padding
    const uint4 q = p + 16 <= t_bytes ? *reinterpret_cast<const uint4*>(t + p)
                                      : make_uint4(0x20202020u, 0x20202020u, 0x20202020u, 0x20202020u);
    reinterpret_cast<uint4*>(st)[i] = q;
  }
  __syncthreads();
  const int64_t k0 = (int64_t)boff[b];
  const int64_t k1 = min(b + 1 < n_blk ? (int64_t)boff[b + 1] : K, K);
  // two readers, so that the staged bytes are read as LDS (one reader that
  // picks LDS or memory per byte compiles to generic loads)
  const auto in_stage = [base](int64_t p) -> uint8_t { return st[p - base]; };
  const auto in_memory = [t](int64_t p) -> uint8_t { return t[p]; };
  for (int64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
    const int64_t s = tok[k];
    int64_t e;
    if (k + 1 < K) {
      e = tok[k + 1];
      if (e - base <= PT_STAGE) {
        while (e > s + 1 && is_ws(in_stage(e - 1))) --e;
      } else {
        while (e > s + 1 && is_ws(in_memory(e - 1))) --e;
      }
    } else {
      e = s + 1;
      while (e < L && !is_ws(in_memory(e))) ++e;
    }
    if (e - base <= PT_STAGE) parse_one(in_stage, k, s, e, kind, len, val, skey);
    else parse_one(in_memory, k, s, e, kind, len, val, skey);
  }
}

// start of the run of bus-looking tokens holding token k (k itself when the
// token before is not one), as max-scan input
__global__ void run_start_in_kernel(const uint8_t* __restrict__ kind, int64_t K, int32_t* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  out[k] = (k == 0 || kind[k - 1] != TK_BUS) ? (int32_t)k : -1;
}

// token k is the id of the bus token before it
__device__ __forceinline__ bool consumed_at(const uint8_t* kind, const int32_t* rs, int64_t k) {
  return k > 0 && kind[k - 1] == TK_BUS && (((k - 1 - rs[k - 1]) & 1) == 0);
}

// time tokens (not ids) as max-scan input; change flags
__global__ void time_in_kernel(const uint8_t* __restrict__ kind, const int32_t* __restrict__ rs, int64_t K,
                               int32_t* __restrict__ ti_in, uint32_t* __restrict__ is_change, DevErr* err) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const bool cons = consumed_at(kind, rs, k);
  const uint8_t kd = kind[k];
  ti_in[k] = (kd == TK_TIME && !cons) ? (int32_t)k : -1;
  if (kd == TK_TIME_BAD && !cons) set_err(err, GTP_VCD_ERR_TIME, k);
  bool ch = false;
  if (!cons && kd == TK_SCALAR) ch = true;
  if (!cons && kd == TK_BUS) {
    if (k + 1 < K) ch = true;
    else set_err(err, GTP_VCD_ERR_ID, k);  // the bus value has no id before the end
  }
  is_change[k] = ch ? 1u : 0u;
}

__device__ __forceinline__ uint64_t fnv1a(const uint8_t* s, int64_t n) {
  uint64_t h = 1469598103934665603ull;
  for (int64_t i = 0; i < n; ++i) h = (h ^ s[i]) * 1099511628211ull;
  return h;
}

__device__ int32_t lookup_id(const uint8_t* s, int64_t n, int32_t key, const int32_t* direct, const int32_t* tab,
                             uint64_t mask, const uint8_t* idc, const int64_t* ido) {
  if (key >= 0) return direct[key];
  uint64_t i = fnv1a(s, n) & mask;
  for (;;) {
    const int32_t u = tab[i];
    if (u < 0) return -1;
    const int64_t b = ido[u], m = ido[u + 1] - b;
    if (m == n) {
      bool eq = true;
      for (int64_t j = 0; j < n && eq; ++j) eq = idc[b + j] == s[j];
      if (eq) return u;
    }
    i = (i + 1) & mask;
  }
}

struct ChangeArgs {
  const uint8_t* t;
  const uint32_t* tok;
  const uint8_t* kind;
  const uint32_t* len;
  const int64_t* val;
  const int32_t* skey;      // short id key of a scalar's id part, else of the token
  const int32_t* ti_in;     // k for a time token that is not a bus value's id, else -1
  const int32_t* ti;        // inclusive max-scan of time tokens
  const uint32_t* chpos;    // exclusive scan of is_change
  const uint32_t* is_change;
  int64_t K;
  const int32_t* direct;    // [SHORT_ID_KEYS] id of a short key, -1 = none
  const int32_t* tab;
  uint64_t mask;
  const uint8_t* idc;
  const int64_t* ido;
  int64_t t_carry;          // time of the changes before the first time token
  uint32_t* ch_uid;
  uint64_t* ch_pack;   // bus << 63 | time << 32 | (scalar: value letter, bus: change index)
  uint64_t* ch_val;    // bus value of bus changes
  uint32_t* flags;     // [0]: set when a time token is below the time before it
  DevErr* err;
};

// one change per scalar token or (bus token, id token) pair, in text order
__global__ void emit_changes_kernel(ChangeArgs a) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.K) return;
  if (a.ti_in[k] == k) {
    const int32_t pk = k > 0 ? a.ti[k - 1] : -1;
    if (a.val[k] < (pk >= 0 ? a.val[pk] : a.t_carry)) a.flags[0] = 1u;
  }
  if (!a.is_change[k]) return;
  const uint32_t c = a.chpos[k];
  const bool bus = a.kind[k] == TK_BUS;
  const uint8_t* id;
  int64_t n;
  int32_t key;
  if (bus) {
    // the next token's own key, unless it looks like a scalar change (its
    // key is then that of its id part)
    key = a.kind[k + 1] != TK_SCALAR ? a.skey[k + 1] : -1;
    id = nullptr;
    n = 0;
    if (key < 0) {
      id = a.t + a.tok[k + 1];
      n = a.len[k + 1];
      key = short_id_key([id](int64_t p) { return id[p]; }, 0, n);
    }
  } else {
    id = a.t + a.tok[k] + 1;
    n = (int64_t)a.len[k] - 1;
    key = a.skey[k];
  }
  const int32_t u = lookup_id(id, n, key, a.direct, a.tab, a.mask, a.idc, a.ido);
  if (u < 0) set_err(a.err, GTP_VCD_ERR_ID, k);
  const int32_t tk = a.ti[k];
  const int64_t time = tk >= 0 ? a.val[tk] : a.t_carry;
  if (time < 0 || time >= (int64_t)EV_TIME_INF) set_err(a.err, GTP_VCD_ERR_RANGE, k);
  a.ch_uid[c] = u < 0 ? 0u : (uint32_t)u;
  const uint64_t tt = (uint64_t)(time < 0 ? 0 : min(time, (int64_t)EV_TIME_INF));
  a.ch_pack[c] = ((uint64_t)(bus ? 1 : 0) << 63) | (tt << 32) | (bus ? (uint64_t)c : (uint64_t)a.val[k]);
  if (bus) a.ch_val[c] = (uint64_t)a.val[k];
}

// the sorted payload unpacked; bus values are the only gather
__global__ void unpack_kernel(const uint64_t* __restrict__ s_pack, int64_t n, const uint64_t* __restrict__ ch_val,
                              uint32_t* __restrict__ s_time, uint64_t* __restrict__ s_val,
                              uint8_t* __restrict__ s_bus) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint64_t w = s_pack[p];
  const bool bus = (w >> 63) != 0;
  s_time[p] = (uint32_t)((w >> 32) & 0x7fffffffu);
  s_bus[p] = bus ? 1 : 0;
  s_val[p] = bus ? ch_val[(uint32_t)w] : (w & 0xffffffffu);
}

// The last kept change of every id in the text parts read so far (a text is
// parsed in parts cut at whitespace): whether there was one, its kind, value
// and time.  All zero before the first part.
struct Carry {
  uint8_t* has;
  uint8_t* bus;
  uint64_t* val;
  uint32_t* time;
};

// the reference's repeat rule against the last kept change (power/Vcd.cc:
// 203-232): a scalar compares its character with the last value's ('\0' for
// a bus value), a bus value compares with the last one's (0 for a scalar)
__device__ __forceinline__ bool differs_from(bool bus, uint64_t v, bool last_bus, uint64_t last_v) {
  return bus ? v != (last_bus ? last_v : 0ull) : v != (last_bus ? 0ull : last_v);
}

// segment bounds per id and the parallel keep rule: a change equal to the
// previous change of its id is dropped (exact when the id's changes are all
// scalar or all bus); ids mixing both are flagged for the sequential pass.
// The first change of an id in a part compares with the carried one.
// flags[1] is set when two changes of an id have the same time, or the
// first has the carried time or an earlier one: only then (or when a time
// token goes back, flags[0]) can kept times fail to increase.
__global__ void keep_kernel(const uint32_t* __restrict__ key, int64_t n, const uint64_t* __restrict__ s_val,
                            const uint8_t* __restrict__ s_bus, const uint32_t* __restrict__ s_time, Carry c,
                            int64_t* __restrict__ seg_start, int64_t* __restrict__ seg_end,
                            uint32_t* __restrict__ mixed, uint32_t* __restrict__ keep, uint32_t* __restrict__ flags) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t u = key[p];
  const bool first = p == 0 || key[p - 1] != u;
  const bool last = p == n - 1 || key[p + 1] != u;
  if (first) seg_start[u] = p;
  if (last) seg_end[u] = p + 1;
  bool k = true;
  if (!first) {
    if (s_bus[p] != s_bus[p - 1]) mixed[u] = 1;
    k = s_val[p] != s_val[p - 1];
    if (s_time[p] == s_time[p - 1]) flags[1] = 1u;
  } else if (c.has[u]) {
    k = differs_from(s_bus[p] != 0, s_val[p], c.bus[u] != 0, c.val[u]);
    if (s_time[p] <= c.time[u]) flags[1] = 1u;
  }
  keep[p] = k ? 1u : 0u;
}

// ids with both scalar and bus changes: the reference's rule with the last
// kept value (a scalar compares its character with the last value's, '\0'
// for a bus value; a bus compares its value with the last one's, 0 for a
// scalar) (power/Vcd.cc:203-232), one thread per id
__global__ void keep_mixed_kernel(const uint32_t* __restrict__ mixed, int64_t n_id,
                                  const int64_t* __restrict__ seg_start, const int64_t* __restrict__ seg_end,
                                  const uint64_t* __restrict__ s_val, const uint8_t* __restrict__ s_bus, Carry c,
                                  uint32_t* __restrict__ keep) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n_id || !mixed[u]) return;
  const bool had = c.has[u] != 0;
  bool last_is_bus = had && c.bus[u] != 0;
  uint64_t last_v = had ? c.val[u] : 0;
  for (int64_t p = seg_start[u]; p < seg_end[u]; ++p) {
    const bool bus = s_bus[p] != 0;
    const uint64_t v = s_val[p];
    const bool k = (p == seg_start[u] && !had) || differs_from(bus, v, last_is_bus, last_v);
    keep[p] = k ? 1u : 0u;
    if (k) {
      last_is_bus = bus;
      last_v = v;
    }
  }
}

// (id, time + 1) keys of kept changes for the order check
__global__ void order_key_kernel(const uint32_t* __restrict__ key, int64_t n, const uint32_t* __restrict__ keep,
                                 const uint32_t* __restrict__ s_time, uint64_t* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  out[p] = keep[p] ? (((uint64_t)key[p] << 32) | ((uint64_t)s_time[p] + 1)) : 0ull;
}

// a kept change no later than an earlier kept change of its id (in this
// part, or the carried one of the parts before)
__global__ void order_check_kernel(const uint32_t* __restrict__ key, int64_t n, const uint32_t* __restrict__ keep,
                                   const uint64_t* __restrict__ okey, const uint64_t* __restrict__ omax,
                                   const uint32_t* __restrict__ s_time, Carry c, DevErr* err) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || !keep[p]) return;
  const uint32_t u = key[p];
  // exclusive max of earlier keys; 0 (the scan's identity; kept keys hold
  // time + 1 >= 1) when the part has none
  const uint64_t prev = omax[p];
  if ((prev >> 32) == u && (prev & 0xffffffffull) != 0) {
    if ((prev & 0xffffffffull) >= (okey[p] & 0xffffffffull)) set_err(err, GTP_VCD_ERR_ORDER, p);
  } else if (c.has[u] && s_time[p] <= c.time[u]) {
    set_err(err, GTP_VCD_ERR_ORDER, p);
  }
}

// the initial value goes in when an id's first change in the text is after 0
__device__ __forceinline__ int initial_event(const Carry& c, uint32_t u, uint32_t first_time) {
  return (!c.has[u] && first_time != 0) ? 1 : 0;
}

// events per stream in this part: the id's kept changes, plus the initial
// value
__global__ void stream_count_kernel(const int32_t* __restrict__ stream_uid, int64_t S,
                                    const int64_t* __restrict__ seg_start, const int64_t* __restrict__ seg_end,
                                    const uint32_t* __restrict__ ks, const uint32_t* __restrict__ s_time, Carry cr,
                                    int64_t* __restrict__ cnt) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const int32_t u = stream_uid[s];
  const int64_t a = seg_start[u];
  int64_t c = 0;
  if (a >= 0) c = (int64_t)(ks[seg_end[u]] - ks[a]) + initial_event(cr, (uint32_t)u, s_time[a]);
  cnt[s] = c;
}

struct ScatterArgs {
  const uint32_t* key;
  int64_t n;
  const uint32_t* keep;
  const uint32_t* ks;
  const int64_t* seg_start;
  const uint32_t* s_time;
  const uint64_t* s_val;
  const uint8_t* s_bus;
  const int32_t* uid_width;
  const int64_t* stream_base;
  const int64_t* stream_off;
  Carry carry;
  uint32_t* ev;
};

__device__ __forceinline__ uint32_t code_of(uint64_t ch) {
  return ch == '0' ? 0u : ch == '1' ? 1u : ch == 'Z' ? 3u : 2u;
}

// every kept change to the streams of its id's bits; the initial value of an
// id whose first change is after 0 at rank 0
__global__ void scatter_kernel(ScatterArgs a) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n || !a.keep[p]) return;
  const uint32_t u = a.key[p];
  const int64_t s0 = a.seg_start[u];
  const int init = initial_event(a.carry, u, a.s_time[s0]);
  const int64_t rank = (int64_t)(a.ks[p] - a.ks[s0]) + init;
  const uint32_t tw = a.s_time[p] << EV_SHIFT;
  const bool bus = a.s_bus[p] != 0;
  const uint64_t v = a.s_val[p];
  const int w = a.uid_width[u];
  const int64_t sb = a.stream_base[u];
  for (int b = 0; b < w; ++b) {
    const uint32_t code = bus ? (uint32_t)((v >> b) & 1u) : code_of(v);
    a.ev[a.stream_off[sb + b] + rank] = tw | code;
  }
  if (p == s0 && init) {
    const uint32_t code0 = bus ? 0u : 2u;  // bus 0 or X at time 0
    for (int b = 0; b < w; ++b) a.ev[a.stream_off[sb + b]] = code0;
  }
}

// the carried change of every id kept in this part: its last kept change
__global__ void carry_kernel(const uint32_t* __restrict__ key, int64_t n, const uint32_t* __restrict__ keep,
                             const uint32_t* __restrict__ ks, const int64_t* __restrict__ seg_end,
                             const uint8_t* __restrict__ s_bus, const uint64_t* __restrict__ s_val,
                             const uint32_t* __restrict__ s_time, Carry c) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || !keep[p]) return;
  const uint32_t u = key[p];
  if (ks[p] + 1 != ks[seg_end[u]]) return;
  c.has[u] = 1;
  c.bus[u] = s_bus[p];
  c.val[u] = s_val[p];
  c.time[u] = s_time[p];
}

// per-stream event counts of one part added up
__global__ void add_counts_kernel(const int64_t* __restrict__ off, int64_t S, int64_t* __restrict__ total) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < S) total[s] += off[s + 1] - off[s];
}

// one part's store appended to the final one: a wave per stream copies the
// stream's words behind those of the parts before (done[s] of them)
__global__ __launch_bounds__(256) void append_part_kernel(const uint32_t* __restrict__ part_ev,
                                                          const int64_t* __restrict__ part_off, int64_t S,
                                                          const int64_t* __restrict__ off, int64_t* __restrict__ done,
                                                          uint32_t* __restrict__ ev) {
  const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (s >= S) return;
  const int64_t a = part_off[s], n = part_off[s + 1] - a;
  const int64_t d = off[s] + done[s];
  for (int64_t i = lane; i < n; i += 64) ev[d + i] = part_ev[a + i];
  if (lane == 0) done[s] += n;
}

__global__ void fill_words_kernel(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_i64_kernel(int64_t* __restrict__ p, int64_t n, int64_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_u8_kernel(uint8_t* __restrict__ p, int64_t n, uint8_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

struct MaxI32 {
  __device__ __host__ int32_t operator()(int32_t a, int32_t b) const { return a > b ? a : b; }
};
struct MaxU64 {
  __device__ __host__ uint64_t operator()(uint64_t a, uint64_t b) const { return a > b ? a : b; }
};

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace vcd
}  // namespace gtp

using namespace gtp;
using namespace gtp::vcd;

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct gtp_vcd {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  // text uploads: on a stream of their own through two page-locked staging
  // buffers, into two device text buffers that parts take in turn (part
  // i + 1 is uploaded while part i is parsed)
  hipStream_t up = nullptr;
  Pinned<uint8_t> stage[2];
  hipEvent_t stage_ev[2] = {nullptr, nullptr};  // the last copy out of a staging buffer
  hipEvent_t text_ev[2] = {nullptr, nullptr};   // a text buffer's upload
  DArr<uint8_t> textb[2];
  // header
  struct Var {
    std::string name, id;
    int32_t width;
    int32_t uid;
  };
  std::vector<Var> vars;
  std::vector<std::string> ids;     // unique ids in declaration order
  std::vector<int32_t> uid_width;   // widest declaration
  std::vector<int64_t> stream_base; // [n_id + 1]
  double time_scale = 1.0;
  // result
  DArr<uint32_t> d_ev;              // [n_event + EV_PAD] event words
  DArr<int64_t> d_off;              // [n_stream + 1]
  // scratch, kept between parses (grown, never shrunk)
  struct Scratch {
    DArr<uint8_t> idc, kind, s_bus, tmp;
    DArr<int32_t> tab, direct, stream_uid, uid_width, rs_in, rs, ti_in, ti, skey;
    DArr<int64_t> ido, stream_base, val, seg_start, seg_end, scnt;
    DArr<uint64_t> bcnt, boff, ch_val, ch_pack, s_pack, s_val, okey, omax;
    DArr<uint32_t> tok, len, is_change, chpos, ch_uid, key, s_time, keep, ks, mixed;
    DArr<unsigned long long> err;
    // carried between the parts of one text
    DArr<uint8_t> c_has, c_bus;
    DArr<uint64_t> c_val;
    DArr<uint32_t> c_time;
    DArr<int64_t> tot, done;
    DArr<uint32_t> flags;  // a time token went back; two changes of an id at one time
  } sc;
  std::vector<int64_t> h_off;
  gtp_vcd_info info{};
  bool parsed = false;

  gtp_status fail(gtp_status st, const std::string& m) {
    error = m;
    return st;
  }
};

namespace {

bool host_ws(int c) { return c == ' ' || (c >= 9 && c <= 13); }

// The reference's token reader over the header (power/VcdReader.cc:287-310):
// the character after a token is consumed with it; a token that the text's
// end cuts off is dropped.
struct HeaderReader {
  const char* t;
  int64_t L, p = 0;
  bool tok(std::string& out) {
    out.clear();
    while (p < L && host_ws((unsigned char)t[p])) ++p;
    while (p < L && !host_ws((unsigned char)t[p])) out.push_back(t[p++]);
    if (p >= L) {
      out.clear();
      return false;
    }
    ++p;  // the terminating whitespace
    return true;
  }
  std::vector<std::string> stmt() {
    std::vector<std::string> v;
    std::string s;
    while (tok(s) && s != "$end") v.push_back(s);
    return v;
  }
};

// power/VcdReader.cc:95-122, :145-236: declarations up to the token that
// starts the value changes; returns the offset right after that token
gtp_status read_header(gtp_vcd* v, const char* text, int64_t L, int64_t* body) {
  HeaderReader r{text, L};
  std::vector<std::string> scope;
  std::unordered_map<std::string, int32_t> uid_of;
  v->vars.clear();
  v->ids.clear();
  v->uid_width.clear();
  v->time_scale = 1.0;
  double unit = 1.0;
  static const char* types[] = {"wire", "reg", "parameter", "integer", "real", "supply0", "supply1", "time",
                                "tri", "triand", "trior", "trireg", "tri0", "tri1", "wand", "wor"};
  auto set_unit = [&](const std::string& u) {
    if (u == "fs") unit = 1e-15;
    else if (u == "ps") unit = 1e-12;
    else if (u == "ns") unit = 1e-9;
    else return false;
    return true;
  };
  std::string t;
  *body = L;
  while (r.tok(t)) {
    if (t == "$date" || t == "$comment" || t == "$version" || t == "$enddefinitions") {
      r.stmt();
    } else if (t == "$timescale") {
      std::vector<std::string> ts = r.stmt();
      try {
        if (ts.size() == 1) {
          size_t last = 0;
          const double s = std::stod(ts[0], &last);
          if (!set_unit(ts[0].substr(last))) return v->fail(GTP_ERR_ARG, "vcd: unknown timescale unit");
          v->time_scale = s * unit;
        } else if (ts.size() == 2) {
          if (!set_unit(ts[1])) return v->fail(GTP_ERR_ARG, "vcd: unknown timescale unit");
          v->time_scale = std::stod(ts[0]) * unit;
        } else {
          return v->fail(GTP_ERR_ARG, "vcd: timescale syntax error");
        }
      } catch (...) {
        return v->fail(GTP_ERR_ARG, "vcd: timescale syntax error");
      }
    } else if (t == "$var") {
      std::vector<std::string> ts = r.stmt();
      if (ts.size() != 4 && ts.size() != 5) return v->fail(GTP_ERR_ARG, "vcd: variable syntax error");
      bool known = false;
      for (const char* ty : types) known = known || ts[0] == ty;
      if (!known) continue;  // the reference warns and skips the variable
      int width = 0;
      try {
        width = std::stoi(ts[1]);
      } catch (...) {
        return v->fail(GTP_ERR_ARG, "vcd: variable width syntax error");
      }
      if (width < 1 || width > 64) return v->fail(GTP_ERR_UNSUPPORTED, "vcd: variable width outside [1, 64]");
      std::string name;
      for (const std::string& c : scope) name += c + "/";
      for (char c : ts[3]) {
        if (c == '/') name += "\\/";
        else name += c;
      }
      if (ts.size() == 5) {
        if (name[0] == '\\') name += ' ';
        name += ts[4];
      }
      auto it = uid_of.find(ts[2]);
      int32_t uid;
      if (it == uid_of.end()) {
        uid = (int32_t)v->ids.size();
        uid_of[ts[2]] = uid;
        v->ids.push_back(ts[2]);
        v->uid_width.push_back(width);
      } else {
        uid = it->second;
        v->uid_width[uid] = std::max(v->uid_width[uid], (int32_t)width);
      }
      v->vars.push_back(gtp_vcd::Var{name, ts[2], (int32_t)width, uid});
    } else if (t == "$scope") {
      std::vector<std::string> ts = r.stmt();
      if (ts.size() < 2) return v->fail(GTP_ERR_ARG, "vcd: scope syntax error");
      scope.push_back(ts[1]);
    } else if (t == "$upscope") {
      r.stmt();
      if (scope.empty()) return v->fail(GTP_ERR_ARG, "vcd: upscope without scope");
      scope.pop_back();
    } else if (t == "$dumpall" || t == "$dumpvars" || t[0] != '$') {
      *body = r.p;  // the value changes start after this token
      return GTP_OK;
    } else {
      return v->fail(GTP_ERR_ARG, "vcd: unhandled vcd command " + t);
    }
  }
  return GTP_OK;
}

uint64_t host_fnv1a(const std::string& s) {
  uint64_t h = 1469598103934665603ull;
  for (unsigned char c : s) h = (h ^ c) * 1099511628211ull;
  return h;
}

const char* err_text(int code) {
  switch (code) {
    case GTP_VCD_ERR_TIME: return "time syntax error";
    case GTP_VCD_ERR_ID: return "unknown variable id";
    case GTP_VCD_ERR_ORDER: return "value changes of an id not increasing in time";
    case GTP_VCD_ERR_RANGE: return "value change time outside [0, 2^29 - 1) ticks";
    default: return "error";
  }
}

}  // namespace

#define VCHK(v, expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return (v)->fail(GTP_ERR_DEVICE, std::string("vcd: ") + #expr + ": " + \
                                               hipGetErrorString(e_));                      \
  } while (0)

GTP_API gtp_status gtp_vcd_create(int32_t device, gtp_vcd** out) {
  if (!out) return GTP_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return GTP_ERR_DEVICE;
  gtp_vcd* v = new gtp_vcd();
  v->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&v->stream) != hipSuccess) {
    delete v;
    return GTP_ERR_DEVICE;
  }
  *out = v;
  return GTP_OK;
}

GTP_API void gtp_vcd_destroy(gtp_vcd* v) {
  if (!v) return;
  (void)hipSetDevice(v->device);
  if (v->stream) (void)hipStreamSynchronize(v->stream);
  if (v->up) (void)hipStreamSynchronize(v->up);
  v->d_ev.release();
  v->d_off.release();
  for (int k = 0; k < 2; ++k) {
    if (v->stage_ev[k]) (void)hipEventDestroy(v->stage_ev[k]);
    if (v->text_ev[k]) (void)hipEventDestroy(v->text_ev[k]);
    v->stage[k].release();
    v->textb[k].release();
  }
  if (v->up) (void)hipStreamDestroy(v->up);
  if (v->stream) (void)hipStreamDestroy(v->stream);
  delete v;
}

GTP_API const char* gtp_vcd_last_error(const gtp_vcd* v) { return v ? v->error.c_str() : "null handle"; }

namespace {
template <class T>
hipError_t grow(DArr<T>& a, size_t n) {
  return a.n >= n && a.p ? hipSuccess : a.alloc(n);
}
template <class T>
hipError_t put(DArr<T>& a, const T* src, size_t n, hipStream_t s) {
  hipError_t e = grow(a, std::max<size_t>(n, 1));
  if (e != hipSuccess || n == 0) return e;
  return hipMemcpyAsync(a.p, src, n * sizeof(T), hipMemcpyHostToDevice, s);
}

constexpr int64_t VCD_STAGE_BYTES = int64_t(64) << 20;

// n bytes from src to dst, by up to 8 host threads
void host_copy(uint8_t* dst, const char* src, int64_t n) {
  const int th = n >= (int64_t(8) << 20) ? 8 : 1;
  if (th == 1) {
    std::memcpy(dst, src, (size_t)n);
    return;
  }
  const int64_t per = (n + th - 1) / th;
  std::vector<std::thread> pool;
  for (int k = 0; k < th; ++k) {
    const int64_t a = k * per, b = std::min(n, a + per);
    if (a < b) pool.emplace_back([=] { std::memcpy(dst + a, src + a, (size_t)(b - a)); });
  }
  for (std::thread& t : pool) t.join();
}

// The device bytes a part takes: whole 4 KB blocks from its body's 16-byte
// aligned start, and 16 more, blank past its text.
size_t part_text_bytes(int64_t len, int64_t b0) {
  const int64_t b0al = b0 & ~(int64_t)15;
  const int64_t n_blk = std::max<int64_t>(1, (len - b0al + BPB - 1) / BPB);
  return (size_t)(b0al + n_blk * BPB + 16);
}

// A part's text [0, len) into the device buffer dst (text_bytes long, blank
// past len) on the upload stream: 64 MB pieces copied by the host threads
// into the two staging buffers in turn (a buffer is refilled once its last
// copy to the device has ended), each copied on to the device; `done` is
// recorded after the last.
hipError_t upload_text(gtp_vcd* v, uint8_t* dst, const char* src, int64_t len, size_t text_bytes, hipEvent_t done) {
  hipError_t e;
  for (int k = 0; k < 2; ++k)
    if ((e = v->stage[k].reserve((size_t)VCD_STAGE_BYTES)) != hipSuccess) return e;
  int64_t k = 0;
  for (int64_t off = 0; off < len; off += VCD_STAGE_BYTES, ++k) {
    const int j = (int)(k & 1);
    if ((e = hipEventSynchronize(v->stage_ev[j])) != hipSuccess) return e;
    const int64_t n = std::min(VCD_STAGE_BYTES, len - off);
    host_copy(v->stage[j].p, src + off, n);
    if ((e = hipMemcpyAsync(dst + off, v->stage[j].p, (size_t)n, hipMemcpyHostToDevice, v->up)) != hipSuccess)
      return e;
    if ((e = hipEventRecord(v->stage_ev[j], v->up)) != hipSuccess) return e;
  }
  if ((e = hipMemsetAsync(dst + len, ' ', text_bytes - (size_t)len, v->up)) != hipSuccess) return e;
  return hipEventRecord(done, v->up);
}
}  // namespace

namespace {

// Text parts: a text longer than the part size is parsed in parts cut at
// whitespace, each through the whole device pipeline, with the time, the
// last kept change of every id and the error positions carried from part to
// part; each part's event store is appended per stream at the end.
constexpr int64_t VCD_PART_BYTES = int64_t(1) << 30;

int64_t part_bytes() {
  int64_t n = VCD_PART_BYTES;
  if (const char* e = std::getenv("GTP_VCD_PART_BYTES")) {
    const long long x = std::atoll(e);
    if (x > 0) n = std::min<int64_t>(x, int64_t(1) << 31);
  }
  return n;
}

// A part may end before the whitespace byte w (w > b0) unless the body
// token before w looks like a bus value: that token takes the next one as
// its id, and which tokens are ids follows from runs of such tokens.
bool good_cut(const char* t, int64_t b0, int64_t w) {
  if (!host_ws((unsigned char)t[w])) return false;
  int64_t q = w - 1;
  while (q >= b0 && host_ws((unsigned char)t[q])) --q;
  if (q < b0) return true;
  int64_t s0 = q;
  while (s0 > b0 && !host_ws((unsigned char)t[s0 - 1])) --s0;
  return t[s0] != 'b' && t[s0] != 'B';
}

// part bounds [0 = c_0, c_1, ..., c_n = len): each cut at the good cut
// nearest below start + size, else the first one above it
std::vector<int64_t> plan_parts(const char* t, int64_t len, int64_t b0, int64_t size) {
  std::vector<int64_t> c{0};
  int64_t start = 0;
  while (len - start > size) {
    const int64_t lo = std::max(start, b0) + 1;
    const int64_t w = std::min(std::max(start + size, lo), len - 1);
    int64_t cut = -1;
    for (int64_t x = w; x >= lo; --x)
      if (good_cut(t, b0, x)) {
        cut = x;
        break;
      }
    for (int64_t x = w + 1; cut < 0 && x < len; ++x)
      if (good_cut(t, b0, x)) cut = x;
    if (cut < 0) break;
    c.push_back(cut);
    start = cut;
  }
  c.push_back(len);
  return c;
}

struct PartIn {
  const char* t;
  int64_t len, b0;      // the part's bytes; its body starts at b0
  const uint8_t* dev_text;  // the part on the device (part_text_bytes long), uploaded when text_ready is
  hipEvent_t text_ready;
  bool cut;             // the text ends inside the part's last token
  int64_t t_carry;      // the time in force at the part's start
  bool carry_out;       // record the part's last kept change of every id
  int64_t tok_before, chg_before;  // tokens and changes of the parts before
  DArr<uint32_t>* ev;   // the part's event store
  DArr<int64_t>* off;   // [n_stream + 1]
  bool pad;             // EV_PAD words of EV_INF after the store
};
struct PartOut {
  int64_t K = 0, n_change = 0, E = 0;
  bool has_time = false;
  int64_t t_last = 0;   // the value of the part's last time token
  bool order_checked = false;  // the kept times needed the full order check
  float ms = 0;
};

const unsigned long long kNoErr = ~0ull;

gtp_status parse_part(gtp_vcd* v, const PartIn& in, PartOut* out, uint64_t tsz, const Carry& carry) {
  hipStream_t s = v->stream;
  gtp_vcd::Scratch& S_ = v->sc;
  const int64_t U = (int64_t)v->ids.size();
  const int64_t S = v->stream_base[U];
  const int64_t len = in.len, b0 = in.b0;
  const int64_t b0al = b0 & ~(int64_t)15;
  const int64_t n_blk = std::max<int64_t>(1, (len - b0al + BPB - 1) / BPB);
  const size_t text_bytes = part_text_bytes(len, b0);
  const uint8_t* const text_d = in.dev_text;
  VCHK(v, hipStreamWaitEvent(s, in.text_ready, 0));
  VCHK(v, put(S_.err, &kNoErr, 1, s));
  DevErr* d_err = reinterpret_cast<DevErr*>(S_.err.p);
  // upper bounds of tokens and changes: a token and its whitespace take 2 bytes
  const int64_t K_max = (len - b0) / 2 + 2;
  VCHK(v, grow(S_.bcnt, n_blk));
  VCHK(v, grow(S_.boff, n_blk));
  VCHK(v, grow(S_.tok, K_max));
  VCHK(v, grow(S_.kind, K_max));
  VCHK(v, grow(S_.len, K_max));
  VCHK(v, grow(S_.val, K_max));
  VCHK(v, grow(S_.skey, K_max));
  VCHK(v, grow(S_.rs_in, K_max));
  VCHK(v, grow(S_.rs, K_max));
  VCHK(v, grow(S_.ti_in, K_max));
  VCHK(v, grow(S_.ti, K_max));
  VCHK(v, grow(S_.is_change, K_max + 1));
  VCHK(v, grow(S_.chpos, K_max + 1));
  VCHK(v, grow(S_.seg_start, std::max<int64_t>(U, 1)));
  VCHK(v, grow(S_.seg_end, std::max<int64_t>(U, 1)));
  VCHK(v, grow(S_.mixed, std::max<int64_t>(U, 1)));
  VCHK(v, grow(S_.scnt, S + 1));
  VCHK(v, in.off->alloc(S + 1));
  auto tmp = [&](size_t need) -> hipError_t { return grow(S_.tmp, std::max<size_t>(need, 1)); };
  hipEvent_t ev0, ev1;
  VCHK(v, hipEventCreate(&ev0));
  VCHK(v, hipEventCreate(&ev1));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } evg{ev0, ev1};
  VCHK(v, hipEventRecord(ev0, s));

  // token starts
  count_starts_kernel<<<dim3((unsigned)n_blk), dim3(TB), 0, s>>>(text_d, b0al, b0, len, S_.bcnt.p);
  VCHK(v, hipGetLastError());
  {
    size_t need = 0;
    VCHK(v, hipcub::DeviceScan::ExclusiveSum(nullptr, need, S_.bcnt.p, S_.boff.p, n_blk, s));
    VCHK(v, tmp(need));
    VCHK(v, hipcub::DeviceScan::ExclusiveSum(S_.tmp.p, need, S_.bcnt.p, S_.boff.p, n_blk, s));
  }
  uint64_t last_off = 0, last_cnt = 0;
  VCHK(v, hipMemcpyAsync(&last_off, S_.boff.p + n_blk - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  VCHK(v, hipMemcpyAsync(&last_cnt, S_.bcnt.p + n_blk - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  VCHK(v, hipStreamSynchronize(s));
  int64_t K = (int64_t)(last_off + last_cnt);
  if (in.cut && K > 0) --K;
  if (K > K_max) return v->fail(GTP_ERR_DEVICE, "gtp_vcd_parse: token count above its bound");
  write_starts_kernel<<<dim3((unsigned)n_blk), dim3(TB), 0, s>>>(text_d, b0al, b0, len, S_.boff.p, S_.tok.p, K);
  VCHK(v, hipGetLastError());

  // tokens
  int64_t n_change = 0;
  if (K > 0) {
    parse_tokens_kernel<<<dim3((unsigned)n_blk), dim3(TB), 0, s>>>(text_d, len, (int64_t)text_bytes, b0al,
                                                                  S_.boff.p, n_blk, S_.tok.p, K, S_.kind.p,
                                                                  S_.len.p, S_.val.p, S_.skey.p);
    VCHK(v, hipGetLastError());
    run_start_in_kernel<<<nblk(K), 256, 0, s>>>(S_.kind.p, K, S_.rs_in.p);
    VCHK(v, hipGetLastError());
    size_t need = 0;
    VCHK(v, hipcub::DeviceScan::InclusiveScan(nullptr, need, S_.rs_in.p, S_.rs.p, MaxI32(), K, s));
    VCHK(v, tmp(need));
    VCHK(v, hipcub::DeviceScan::InclusiveScan(S_.tmp.p, need, S_.rs_in.p, S_.rs.p, MaxI32(), K, s));
    time_in_kernel<<<nblk(K), 256, 0, s>>>(S_.kind.p, S_.rs.p, K, S_.ti_in.p, S_.is_change.p, d_err);
    VCHK(v, hipGetLastError());
    VCHK(v, hipcub::DeviceScan::InclusiveScan(nullptr, need, S_.ti_in.p, S_.ti.p, MaxI32(), K, s));
    VCHK(v, tmp(need));
    VCHK(v, hipcub::DeviceScan::InclusiveScan(S_.tmp.p, need, S_.ti_in.p, S_.ti.p, MaxI32(), K, s));
    fill_u32_kernel<<<1, 1, 0, s>>>(S_.is_change.p + K, 1, 0u);
    VCHK(v, hipcub::DeviceScan::ExclusiveSum(nullptr, need, S_.is_change.p, S_.chpos.p, K + 1, s));
    VCHK(v, tmp(need));
    VCHK(v, hipcub::DeviceScan::ExclusiveSum(S_.tmp.p, need, S_.is_change.p, S_.chpos.p, K + 1, s));
    uint32_t nc = 0;
    int32_t tlast = -1;
    VCHK(v, hipMemcpyAsync(&nc, S_.chpos.p + K, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    VCHK(v, hipMemcpyAsync(&tlast, S_.ti.p + K - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    VCHK(v, hipStreamSynchronize(s));
    n_change = nc;
    if (tlast >= 0) {
      VCHK(v, hipMemcpyAsync(&out->t_last, S_.val.p + tlast, sizeof(int64_t), hipMemcpyDeviceToHost, s));
      VCHK(v, hipStreamSynchronize(s));
      out->has_time = true;
    }
  }

  // changes in text order, grouped by id (stable), payload gathered once
  const int64_t NC1 = std::max<int64_t>(n_change, 1);
  VCHK(v, grow(S_.ch_uid, NC1));
  VCHK(v, grow(S_.key, NC1));
  VCHK(v, grow(S_.ch_pack, NC1));
  VCHK(v, grow(S_.s_pack, NC1));
  VCHK(v, grow(S_.ch_val, NC1));
  VCHK(v, grow(S_.s_time, NC1));
  VCHK(v, grow(S_.s_val, NC1));
  VCHK(v, grow(S_.s_bus, NC1));
  VCHK(v, grow(S_.keep, NC1 + 1));
  VCHK(v, grow(S_.ks, NC1 + 1));
  VCHK(v, grow(S_.okey, NC1));
  VCHK(v, grow(S_.omax, NC1));
  const int64_t U1 = std::max<int64_t>(U, 1);
  fill_i64_kernel<<<nblk(U1), 256, 0, s>>>(S_.seg_start.p, U1, -1);
  fill_i64_kernel<<<nblk(U1), 256, 0, s>>>(S_.seg_end.p, U1, 0);
  fill_u32_kernel<<<nblk(U1), 256, 0, s>>>(S_.mixed.p, U1, 0u);
  VCHK(v, grow(S_.flags, 2));
  fill_u32_kernel<<<1, 2, 0, s>>>(S_.flags.p, 2, 0u);
  VCHK(v, hipGetLastError());
  if (n_change > 0) {
    ChangeArgs ca{text_d, S_.tok.p, S_.kind.p, S_.len.p, S_.val.p, S_.skey.p, S_.ti_in.p, S_.ti.p,
                  S_.chpos.p, S_.is_change.p, K, S_.direct.p, S_.tab.p, tsz - 1, S_.idc.p, S_.ido.p,
                  in.t_carry, S_.ch_uid.p, S_.ch_pack.p, S_.ch_val.p, S_.flags.p, d_err};
    emit_changes_kernel<<<nblk(K), 256, 0, s>>>(ca);
    VCHK(v, hipGetLastError());
    int bits = 1;
    while (bits < 32 && (1ll << bits) < U) ++bits;
    size_t need = 0;
    VCHK(v, hipcub::DeviceRadixSort::SortPairs(nullptr, need, S_.ch_uid.p, S_.key.p, S_.ch_pack.p, S_.s_pack.p,
                                               n_change, 0, bits, s));
    VCHK(v, tmp(need));
    VCHK(v, hipcub::DeviceRadixSort::SortPairs(S_.tmp.p, need, S_.ch_uid.p, S_.key.p, S_.ch_pack.p, S_.s_pack.p,
                                               n_change, 0, bits, s));
    unpack_kernel<<<nblk(n_change), 256, 0, s>>>(S_.s_pack.p, n_change, S_.ch_val.p, S_.s_time.p, S_.s_val.p,
                                                 S_.s_bus.p);
    VCHK(v, hipGetLastError());
    keep_kernel<<<nblk(n_change), 256, 0, s>>>(S_.key.p, n_change, S_.s_val.p, S_.s_bus.p, S_.s_time.p, carry,
                                               S_.seg_start.p, S_.seg_end.p, S_.mixed.p, S_.keep.p, S_.flags.p);
    VCHK(v, hipGetLastError());
    if (U > 0) {
      keep_mixed_kernel<<<nblk(U), 256, 0, s>>>(S_.mixed.p, U, S_.seg_start.p, S_.seg_end.p, S_.s_val.p,
                                                S_.s_bus.p, carry, S_.keep.p);
      VCHK(v, hipGetLastError());
    }
  }
  fill_u32_kernel<<<1, 1, 0, s>>>(S_.keep.p + n_change, 1, 0u);
  {
    size_t need = 0;
    VCHK(v, hipcub::DeviceScan::ExclusiveSum(nullptr, need, S_.keep.p, S_.ks.p, n_change + 1, s));
    VCHK(v, tmp(need));
    VCHK(v, hipcub::DeviceScan::ExclusiveSum(S_.tmp.p, need, S_.keep.p, S_.ks.p, n_change + 1, s));
  }
  // streams
  if (S > 0) {
    stream_count_kernel<<<nblk(S), 256, 0, s>>>(S_.stream_uid.p, S, S_.seg_start.p, S_.seg_end.p, S_.ks.p,
                                                S_.s_time.p, carry, S_.scnt.p);
    VCHK(v, hipGetLastError());
  }
  fill_i64_kernel<<<1, 1, 0, s>>>(S_.scnt.p + S, 1, 0);
  {
    size_t need = 0;
    VCHK(v, hipcub::DeviceScan::ExclusiveSum(nullptr, need, S_.scnt.p, in.off->p, S + 1, s));
    VCHK(v, tmp(need));
    VCHK(v, hipcub::DeviceScan::ExclusiveSum(S_.tmp.p, need, S_.scnt.p, in.off->p, S + 1, s));
  }
  unsigned long long he = 0;
  int64_t E = 0;
  uint32_t fl[2] = {0, 0};
  VCHK(v, hipMemcpyAsync(&he, S_.err.p, sizeof(he), hipMemcpyDeviceToHost, s));
  VCHK(v, hipMemcpyAsync(&E, in.off->p + S, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  VCHK(v, hipMemcpyAsync(fl, S_.flags.p, sizeof(fl), hipMemcpyDeviceToHost, s));
  VCHK(v, hipStreamSynchronize(s));
  // Kept times of an id strictly increasing: when the time tokens never go
  // back (from the carried time on), the changes of an id are in time order,
  // so only two changes at one time (or a first one at the carried time) can
  // break it; otherwise every kept change is checked against the largest
  // earlier kept (id, time) key.
  out->order_checked = n_change > 0 && (fl[0] || fl[1]);
  if (out->order_checked) {
    size_t need = 0;
    order_key_kernel<<<nblk(n_change), 256, 0, s>>>(S_.key.p, n_change, S_.keep.p, S_.s_time.p, S_.okey.p);
    VCHK(v, hipGetLastError());
    VCHK(v, hipcub::DeviceScan::ExclusiveScan(nullptr, need, S_.okey.p, S_.omax.p, MaxU64(), (uint64_t)0,
                                              n_change, s));
    VCHK(v, tmp(need));
    VCHK(v, hipcub::DeviceScan::ExclusiveScan(S_.tmp.p, need, S_.okey.p, S_.omax.p, MaxU64(), (uint64_t)0,
                                              n_change, s));
    order_check_kernel<<<nblk(n_change), 256, 0, s>>>(S_.key.p, n_change, S_.keep.p, S_.okey.p, S_.omax.p,
                                                      S_.s_time.p, carry, d_err);
    VCHK(v, hipGetLastError());
    VCHK(v, hipMemcpyAsync(&he, S_.err.p, sizeof(he), hipMemcpyDeviceToHost, s));
    VCHK(v, hipStreamSynchronize(s));
  }
  if (he != ~0ull) {
    const int code = (int)(he & 0xff);
    const int64_t item = (int64_t)(he >> 8) + (code == GTP_VCD_ERR_ORDER ? in.chg_before : in.tok_before);
    return v->fail(GTP_ERR_ARG, std::string("vcd: ") + err_text(code) + " (item " + std::to_string(item) +
                                    " of the value-change section)");
  }
  const int64_t pad = in.pad ? EV_PAD : 0;
  VCHK(v, grow(*in.ev, std::max<int64_t>(E + pad, 1)));
  if (pad) fill_words_kernel<<<1, 64, 0, s>>>(in.ev->p + E, EV_PAD, EV_INF);
  if (n_change > 0) {
    ScatterArgs sa{S_.key.p, n_change, S_.keep.p, S_.ks.p, S_.seg_start.p, S_.s_time.p, S_.s_val.p, S_.s_bus.p,
                   S_.uid_width.p, S_.stream_base.p, in.off->p, carry, in.ev->p};
    scatter_kernel<<<nblk(n_change), 256, 0, s>>>(sa);
    VCHK(v, hipGetLastError());
    if (in.carry_out) {
      carry_kernel<<<nblk(n_change), 256, 0, s>>>(S_.key.p, n_change, S_.keep.p, S_.ks.p, S_.seg_end.p, S_.s_bus.p,
                                                  S_.s_val.p, S_.s_time.p, carry);
      VCHK(v, hipGetLastError());
    }
  }
  VCHK(v, hipEventRecord(ev1, s));
  VCHK(v, hipStreamSynchronize(s));
  VCHK(v, hipEventElapsedTime(&out->ms, ev0, ev1));
  out->K = K;
  out->n_change = n_change;
  out->E = E;
  return GTP_OK;
}

}  // namespace

GTP_API gtp_status gtp_vcd_parse(gtp_vcd* v, const char* text, int64_t len, gtp_vcd_info* info) {
  if (!v) return GTP_ERR_ARG;
  if (!text || len < 0) return v->fail(GTP_ERR_ARG, "gtp_vcd_parse: bad text");
  v->parsed = false;
  v->error.clear();
  VCHK(v, hipSetDevice(v->device));
  hipStream_t s = v->stream;
  gtp_vcd::Scratch& S_ = v->sc;
  int64_t b0 = 0;
  {
    const gtp_status st = read_header(v, text, len, &b0);
    if (st != GTP_OK) return st;
  }
  const int64_t U = (int64_t)v->ids.size();
  if (U >= (int64_t)INT32_MAX) return v->fail(GTP_ERR_LIMIT, "gtp_vcd_parse: too many ids");
  const std::vector<int64_t> bounds = plan_parts(text, len, b0, part_bytes());
  const int64_t n_part = (int64_t)bounds.size() - 1;
  for (int64_t i = 0; i < n_part; ++i)
    if (bounds[i + 1] - bounds[i] >= (int64_t)UINT32_MAX - 64 - BPB)
      return v->fail(GTP_ERR_UNSUPPORTED, "gtp_vcd_parse: no place to cut the text into parts under 4 GiB");
  v->stream_base.assign(U + 1, 0);
  for (int64_t u = 0; u < U; ++u) v->stream_base[u + 1] = v->stream_base[u] + v->uid_width[u];
  const int64_t S = v->stream_base[U];
  std::vector<int32_t> stream_uid(std::max<int64_t>(S, 1), 0);
  for (int64_t u = 0; u < U; ++u)
    for (int64_t q = v->stream_base[u]; q < v->stream_base[u + 1]; ++q) stream_uid[q] = (int32_t)u;
  // ids of 1-3 printable characters: a direct table; longer ones: open
  // addressing, exact compare on the device
  std::vector<int32_t> direct(SHORT_ID_KEYS, -1);
  uint64_t tsz = 16;
  while (tsz < 2 * (uint64_t)std::max<int64_t>(U, 1)) tsz <<= 1;
  std::vector<int32_t> tab(tsz, -1);
  std::vector<int64_t> ido(U + 1, 0);
  std::string idc;
  for (int64_t u = 0; u < U; ++u) {
    const std::string& id = v->ids[u];
    int32_t key = -1;
    if (!id.empty() && id.size() <= 3) {
      key = 0;
      int32_t m = 1;
      for (unsigned char c : id) {
        if (c < 33 || c > 126) {
          key = -1;
          break;
        }
        key += (c - 32) * m;
        m *= 95;
      }
    }
    if (key >= 0) direct[key] = (int32_t)u;
    uint64_t i = host_fnv1a(id) & (tsz - 1);
    while (tab[i] >= 0) i = (i + 1) & (tsz - 1);
    tab[i] = (int32_t)u;
    idc += id;
    ido[u + 1] = (int64_t)idc.size();
  }

  const auto t0 = std::chrono::steady_clock::now();
  VCHK(v, put(S_.tab, tab.data(), tab.size(), s));
  VCHK(v, put(S_.direct, direct.data(), direct.size(), s));
  VCHK(v, put(S_.ido, ido.data(), ido.size(), s));
  VCHK(v, put(S_.idc, reinterpret_cast<const uint8_t*>(idc.data()), idc.size(), s));
  VCHK(v, put(S_.stream_uid, stream_uid.data(), stream_uid.size(), s));
  VCHK(v, put(S_.uid_width, v->uid_width.data(), v->uid_width.size(), s));
  VCHK(v, put(S_.stream_base, v->stream_base.data(), v->stream_base.size(), s));
  const int64_t U1 = std::max<int64_t>(U, 1);
  VCHK(v, grow(S_.c_has, U1));
  VCHK(v, grow(S_.c_bus, U1));
  VCHK(v, grow(S_.c_val, U1));
  VCHK(v, grow(S_.c_time, U1));
  fill_u8_kernel<<<nblk(U1), 256, 0, s>>>(S_.c_has.p, U1, 0);
  VCHK(v, hipGetLastError());
  const Carry carry{S_.c_has.p, S_.c_bus.p, S_.c_val.p, S_.c_time.p};

  // one part: its store is the result; several: each part's store is kept
  // until the per-stream totals are known
  std::vector<std::unique_ptr<DArr<uint32_t>>> part_ev;
  std::vector<std::unique_ptr<DArr<int64_t>>> part_off;
  int64_t t_now = 0, n_token = 0, n_change = 0;
  double ms = 0;
  if (n_part > 1) {
    VCHK(v, grow(S_.tot, S + 1));
    VCHK(v, grow(S_.done, S + 1));
    fill_i64_kernel<<<nblk(S + 1), 256, 0, s>>>(S_.tot.p, S + 1, 0);
    fill_i64_kernel<<<nblk(S + 1), 256, 0, s>>>(S_.done.p, S + 1, 0);
    VCHK(v, hipGetLastError());
  }
  // part 0's text is uploaded here, part i + 1's by a helper thread while
  // part i is parsed
  if (!v->up) VCHK(v, hipStreamCreate(&v->up));
  for (int k = 0; k < 2; ++k) {
    if (!v->stage_ev[k]) VCHK(v, hipEventCreateWithFlags(&v->stage_ev[k], hipEventDisableTiming));
    if (!v->text_ev[k]) VCHK(v, hipEventCreateWithFlags(&v->text_ev[k], hipEventDisableTiming));
  }
  const auto part_len = [&](int64_t i) { return bounds[i + 1] - bounds[i]; };
  {
    const size_t tb0 = part_text_bytes(part_len(0), b0);
    VCHK(v, grow(v->textb[0], tb0));
    VCHK(v, upload_text(v, v->textb[0].p, text, part_len(0), tb0, v->text_ev[0]));
  }
  for (int64_t i = 0; i < n_part; ++i) {
    const bool last = i + 1 == n_part;
    const int j = (int)(i & 1);
    std::thread helper;
    hipError_t up_err = hipSuccess;
    if (!last) {
      // buffer 1 - j held part i - 1, whose parse has ended
      const int jn = 1 - j;
      const size_t tbn = part_text_bytes(part_len(i + 1), 0);
      VCHK(v, grow(v->textb[jn], tbn));
      helper = std::thread([v, jn, tbn, src = text + bounds[i + 1], n = part_len(i + 1), &up_err] {
        up_err = hipSetDevice(v->device);
        if (up_err == hipSuccess) up_err = upload_text(v, v->textb[jn].p, src, n, tbn, v->text_ev[jn]);
      });
    }
    PartIn in{};
    in.dev_text = v->textb[j].p;
    in.text_ready = v->text_ev[j];
    in.t = text + bounds[i];
    in.len = bounds[i + 1] - bounds[i];
    in.b0 = i == 0 ? b0 : 0;
    in.cut = last && in.len > in.b0 && !host_ws((unsigned char)in.t[in.len - 1]);
    in.t_carry = t_now;
    in.carry_out = !last;
    in.tok_before = n_token;
    in.chg_before = n_change;
    if (n_part == 1) {
      in.ev = &v->d_ev;
      in.off = &v->d_off;
      in.pad = true;
    } else {
      part_ev.emplace_back(new DArr<uint32_t>());
      part_off.emplace_back(new DArr<int64_t>());
      in.ev = part_ev.back().get();
      in.off = part_off.back().get();
      in.pad = false;
    }
    PartOut o;
    const gtp_status st = parse_part(v, in, &o, tsz, carry);
    if (helper.joinable()) helper.join();
    if (st != GTP_OK) return st;
    if (up_err != hipSuccess)
      return v->fail(GTP_ERR_DEVICE, std::string("vcd: text upload: ") + hipGetErrorString(up_err));
    if (o.has_time) t_now = o.t_last;
    n_token += o.K;
    n_change += o.n_change;
    ms += o.ms;
    if (n_part > 1 && S > 0) {
      add_counts_kernel<<<nblk(S), 256, 0, s>>>(in.off->p, S, S_.tot.p);
      VCHK(v, hipGetLastError());
    }
  }
  if (n_part > 1) {
    hipEvent_t ea, eb;
    VCHK(v, hipEventCreate(&ea));
    VCHK(v, hipEventCreate(&eb));
    struct EvGuard {
      hipEvent_t a, b;
      ~EvGuard() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
      }
    } evg{ea, eb};
    VCHK(v, v->d_off.alloc(S + 1));
    VCHK(v, hipEventRecord(ea, s));
    size_t need = 0;
    VCHK(v, hipcub::DeviceScan::ExclusiveSum(nullptr, need, S_.tot.p, v->d_off.p, S + 1, s));
    VCHK(v, grow(S_.tmp, std::max<size_t>(need, 1)));
    VCHK(v, hipcub::DeviceScan::ExclusiveSum(S_.tmp.p, need, S_.tot.p, v->d_off.p, S + 1, s));
    int64_t E = 0;
    VCHK(v, hipMemcpyAsync(&E, v->d_off.p + S, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    VCHK(v, hipStreamSynchronize(s));
    VCHK(v, grow(v->d_ev, E + EV_PAD));
    fill_words_kernel<<<1, 64, 0, s>>>(v->d_ev.p + E, EV_PAD, EV_INF);
    if (S > 0)
      for (int64_t i = 0; i < n_part; ++i) {
        append_part_kernel<<<nblk(S, 4), 256, 0, s>>>(part_ev[i]->p, part_off[i]->p, S, v->d_off.p, S_.done.p,
                                                      v->d_ev.p);
        VCHK(v, hipGetLastError());
      }
    VCHK(v, hipEventRecord(eb, s));
    VCHK(v, hipStreamSynchronize(s));
    float am = 0;
    VCHK(v, hipEventElapsedTime(&am, ea, eb));
    ms += am;
  }
  v->h_off.resize(S + 1);
  VCHK(v, hipMemcpyAsync(v->h_off.data(), v->d_off.p, sizeof(int64_t) * (S + 1), hipMemcpyDeviceToHost, s));
  VCHK(v, hipStreamSynchronize(s));
  const auto t1 = std::chrono::steady_clock::now();
  v->info = gtp_vcd_info{};
  v->info.n_var = (int64_t)v->vars.size();
  v->info.n_id = U;
  v->info.n_stream = S;
  v->info.n_event = v->h_off[S];
  v->info.t_max = t_now;
  v->info.time_scale = v->time_scale;
  v->info.n_token = n_token;
  v->info.n_change = n_change;
  v->info.body_bytes = len - b0;
  v->info.device_ms = ms;
  v->info.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  v->info.n_part = n_part;
  v->parsed = true;
  if (info) *info = v->info;
  return GTP_OK;
}

GTP_API gtp_status gtp_vcd_var(const gtp_vcd* v, int64_t i, gtp_vcd_var_info* out) {
  if (!v || !out || i < 0 || i >= (int64_t)v->vars.size()) return GTP_ERR_ARG;
  const gtp_vcd::Var& x = v->vars[i];
  out->name = x.name.c_str();
  out->id = x.id.c_str();
  out->width = x.width;
  out->id_index = x.uid;
  out->first_stream = v->stream_base[x.uid];
  return GTP_OK;
}

GTP_API gtp_status gtp_vcd_events(gtp_vcd* v, int64_t* offsets, int64_t* times, uint8_t* values) {
  if (!v || !v->parsed) return GTP_ERR_STATE;
  if (!offsets) return GTP_ERR_ARG;
  const int64_t S = v->info.n_stream, E = v->info.n_event;
  std::memcpy(offsets, v->h_off.data(), sizeof(int64_t) * (S + 1));
  if (E > 0 && (!times || !values)) return GTP_ERR_ARG;
  std::vector<uint32_t> w(E);
  if (E > 0) {
    VCHK(v, hipSetDevice(v->device));
    VCHK(v, hipMemcpy(w.data(), v->d_ev.p, sizeof(uint32_t) * E, hipMemcpyDeviceToHost));
  }
  for (int64_t e = 0; e < E; ++e) {
    times[e] = w[e] >> EV_SHIFT;
    values[e] = (uint8_t)(w[e] & 3u);
  }
  return GTP_OK;
}

namespace gtp {
bool vcd_view(const gtp_vcd* v, VcdView* out) {
  if (!v || !v->parsed) return false;
  out->device = v->device;
  out->ev = v->d_ev.p;
  out->d_off = v->d_off.p;
  out->h_off = v->h_off.data();
  out->n_stream = v->info.n_stream;
  out->n_event = v->info.n_event;
  out->t_max = v->info.t_max;
  out->time_scale = v->info.time_scale;
  return true;
}
}  // namespace gtp
