// Learned per-cycle power surrogate: a small MLP over per-cycle toggle bits,
// trained and run by hand-written gfx950 BF16 MFMA kernels (FP32 accumulate).
//
// No counterpart in the reference (a new method on top of its path).  The
// exact engine (gtp_run) produces the training targets and judges the
// surrogate on sampled cycles.
//
//   features : bit k of cycle c = net sel[k] has an event (index >= 1) in cycle c
//   model    : z1 = X W1^T + b1, h1 = GELU(z1), z2 = h1 W2^T + b2, h2 = GELU(z2),
//              y = h2 . w3 + b3     (W1 [H][K], W2 [H][H], GELU(x) = x Phi(x))
//   training : MSE on standardised targets, Adam, FP32 master weights, BF16
//              GEMM operands (weights, activations, gated gradients)
//
// GEMM: C[M][N] = A[M][K] * B[K][N] on v_mfma_f32_16x16x32_bf16.  A workgroup
// of 4 waves computes a 128 x 128 tile; each wave a 64 x 64 block (4 x 4 MFMA
// tiles, 64 accumulators).  Operands are staged in LDS as A[m][k] and B[n][k]
// (k contiguous, 16-byte fragment reads).  Operand modes cover row-major and
// transposed BF16 matrices and bit-packed features, so no transpose is ever
// materialised.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "gtp_internal.hpp"
#include "gtpower.h"

#define GTP_API extern "C" __attribute__((visibility("default")))

namespace gtp {
namespace mlp {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 32, LDSK = BK + 8;  // LDS row: 40 bf16 = 80 B

enum { A_ROW = 0, A_COL = 1, A_BITS_ROW = 2 };
enum { B_ROW = 0, B_TRANS = 1, B_BITS_ROW = 2 };
enum { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_GELU_BACK = 2 };

struct GemmArgs {
  int M, N, K;
  const void* A;
  int lda;          // row stride of A as stored (elements; words for bits)
  const void* B;
  int ldb;
  float* C;         // EPI_F32: result; EPI_BIAS_GELU: pre-activation (may be null)
  int ldc;
  bf16* Cb;         // bf16 result (activation or gated gradient)
  int ldcb;
  const float* bias;
  const float* zin; // EPI_GELU_BACK: pre-activation
  int ldz;
  // training-time input dropout on bit operands: a bit of row r, word w is
  // kept when bit b of keep_mask(drop_seed, r, w) is set (probability 1/2) and
  // then counts 2 (inverted dropout); drop_seed == 0: no dropout
  uint32_t drop_seed;
};

// 32 pseudo-random keep bits for (seed, row, word): a murmur3-style finaliser
__host__ __device__ __forceinline__ uint32_t keep_mask(uint32_t seed, uint32_t row, uint32_t word) {
  uint32_t h = seed ^ (row * 0x9E3779B1u) ^ (word * 0x85EBCA77u);
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return h;
}

__device__ __forceinline__ bf16 f2bf(float x) { return (bf16)x; }
__device__ __forceinline__ float bf2f(bf16 x) { return (float)x; }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// 8 bits -> 8 bf16 zeros and `one`s (bf16 bits of 1.0 = 0x3F80, of 2.0 = 0x4000),
// as one 16-byte vector
__device__ __forceinline__ uint4 expand8(uint32_t b, uint32_t one = 0x3F80u) {
  const uint32_t hi = one << 16;
  uint4 v;
  v.x = ((b & 1u) ? one : 0u) | ((b & 2u) ? hi : 0u);
  v.y = ((b & 4u) ? one : 0u) | ((b & 8u) ? hi : 0u);
  v.z = ((b & 16u) ? one : 0u) | ((b & 32u) ? hi : 0u);
  v.w = ((b & 64u) ? one : 0u) | ((b & 128u) ? hi : 0u);
  return v;
}

// The operand words one thread moves from global memory into the LDS tiles
// for one BK step (which fields are used depends on the layouts).
struct TileRegs {
  uint4 a0, a1, b0, b1;
  uint32_t aw, bw;
};

template <int AM, int BMODE, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
  // two LDS buffers per operand: the step being multiplied reads one while the
  // next step's operands, loaded into registers before the multiply, go into
  // the other (one barrier per step)
  __shared__ __align__(16) bf16 As[2][BM * LDSK];
  __shared__ __align__(16) bf16 Bs[2][BN * LDSK];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  floatx4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  // global -> registers for the step at k0
  const auto fetch = [&](int k0) {
    TileRegs r{};
    if (AM == A_ROW) {
      const bf16* A = (const bf16*)g.A;
      const int row = tid >> 1, ks = (tid & 1) * 16;
      if (m0 + row < g.M) {
        const uint4* src = (const uint4*)(A + (int64_t)(m0 + row) * g.lda + k0 + ks);
        r.a0 = src[0];
        r.a1 = src[1];
      }
    } else if (AM == A_COL) {
      // A[m][k] = S[k][m], S row-major with stride lda
      // (M is a multiple of 16 in this mode: whole 16-row groups or none)
      const bf16* S = (const bf16*)g.A;
      const int kk = tid >> 3, ms = (tid & 7) * 16;
      if (m0 + ms < g.M) {
        const uint4* src = (const uint4*)(S + (int64_t)(k0 + kk) * g.lda + m0 + ms);
        r.a0 = src[0];
        r.a1 = src[1];
      }
    } else {
      // bit-packed rows: A[m][k] = bit k of row m
      const uint32_t* X = (const uint32_t*)g.A;
      const int row = tid >> 1;
      r.aw = (m0 + row < g.M) ? X[(int64_t)(m0 + row) * g.lda + (k0 >> 5)] : 0u;
    }
    if (BMODE == B_TRANS) {
      // B[k][n] = W[n][k], W row-major with stride ldb
      const bf16* W = (const bf16*)g.B;
      const int row = tid >> 1, ks = (tid & 1) * 16;
      const uint4* src = (const uint4*)(W + (int64_t)(n0 + row) * g.ldb + k0 + ks);
      r.b0 = src[0];
      r.b1 = src[1];
    } else if (BMODE == B_ROW) {
      const bf16* Bm = (const bf16*)g.B;
      const int kk = tid >> 3, ns = (tid & 7) * 16;
      const uint4* src = (const uint4*)(Bm + (int64_t)(k0 + kk) * g.ldb + n0 + ns);
      r.b0 = src[0];
      r.b1 = src[1];
    } else {
      // B[k][n] = bit n of bit-packed row k
      const uint32_t* X = (const uint32_t*)g.B;
      if (tid < BN) r.bw = X[(int64_t)(k0 + (tid >> 2)) * g.ldb + (n0 >> 5) + (tid & 3)];
    }
    return r;
  };
  // registers -> LDS buffer buf (As[m][k], Bs[n][k])
  const auto stash = [&](const TileRegs& r, int k0, int buf) {
    bf16* as = As[buf];
    bf16* bs = Bs[buf];
    if (AM == A_ROW) {
      const int row = tid >> 1, ks = (tid & 1) * 16;
      uint4* dst = (uint4*)(as + row * LDSK + ks);
      dst[0] = r.a0;
      dst[1] = r.a1;
    } else if (AM == A_COL) {
      const int kk = tid >> 3, ms = (tid & 7) * 16;
      union {
        uint4 u[2];
        bf16 h[16];
      } v;
      v.u[0] = r.a0;
      v.u[1] = r.a1;
#pragma unroll
      for (int j = 0; j < 16; ++j) as[(ms + j) * LDSK + kk] = v.h[j];
    } else {
      const int row = tid >> 1, h = tid & 1;
      uint32_t w = r.aw, one = 0x3F80u;
      if (g.drop_seed) {
        w &= keep_mask(g.drop_seed, (uint32_t)(m0 + row), (uint32_t)(k0 >> 5));
        one = 0x4000u;
      }
      uint4* dst = (uint4*)(as + row * LDSK + h * 16);
      dst[0] = expand8((w >> (h * 16)) & 0xFFu, one);
      dst[1] = expand8((w >> (h * 16 + 8)) & 0xFFu, one);
    }
    if (BMODE == B_TRANS) {
      const int row = tid >> 1, ks = (tid & 1) * 16;
      uint4* dst = (uint4*)(bs + row * LDSK + ks);
      dst[0] = r.b0;
      dst[1] = r.b1;
    } else if (BMODE == B_ROW) {
      const int kk = tid >> 3, ns = (tid & 7) * 16;
      union {
        uint4 u[2];
        bf16 h[16];
      } v;
      v.u[0] = r.b0;
      v.u[1] = r.b1;
#pragma unroll
      for (int j = 0; j < 16; ++j) bs[(ns + j) * LDSK + kk] = v.h[j];
    } else {
      if (tid < BN) {
        const int kk = tid >> 2, wi = tid & 3;
        uint32_t w = r.bw;
        bf16 one = (bf16)1.0f;
        if (g.drop_seed) {
          w &= keep_mask(g.drop_seed, (uint32_t)(k0 + kk), (uint32_t)((n0 >> 5) + wi));
          one = (bf16)2.0f;
        }
#pragma unroll
        for (int b = 0; b < 32; ++b) bs[(wi * 32 + b) * LDSK + kk] = ((w >> b) & 1u) ? one : (bf16)0.0f;
      }
    }
  };

  const int nk = g.K / BK;
  stash(fetch(0), 0, 0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    // the next step's operands, in flight during the multiply (the last step
    // loads its own again, unused)
    const TileRegs r = fetch((more ? kt + 1 : kt) * BK);
    // ---- 4 x 4 MFMA 16x16x32 per wave
    bf16x8 af[4], bfr[4];
    const int fr = lane & 15, fk = (lane >> 4) * 8;
#pragma unroll
    for (int i = 0; i < 4; ++i) af[i] = *(const bf16x8*)(As[cur] + (wm * 64 + i * 16 + fr) * LDSK + fk);
#pragma unroll
    for (int j = 0; j < 4; ++j) bfr[j] = *(const bf16x8*)(Bs[cur] + (wn * 64 + j * 16 + fr) * LDSK + fk);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    // the other buffer was last read in the previous step, before its barrier
    if (more) stash(r, (kt + 1) * BK, cur ^ 1);
    __syncthreads();
  }
  // ---- epilogue: lane holds D[(lane/16)*4 + e][lane%16] of each 16x16 tile
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = m0 + wm * 64 + i * 16 + (lane >> 4) * 4 + e;
        const int col = n0 + wn * 64 + j * 16 + (lane & 15);
        if (row >= g.M) continue;
        const float v = acc[i][j][e];
        if (EPI == EPI_F32) {
          g.C[(int64_t)row * g.ldc + col] = v;
        } else if (EPI == EPI_BIAS_GELU) {
          const float z = v + g.bias[col];
          if (g.C) g.C[(int64_t)row * g.ldc + col] = z;
          g.Cb[(int64_t)row * g.ldcb + col] = f2bf(gelu(z));
        } else {
          const float z = g.zin[(int64_t)row * g.ldz + col];
          g.Cb[(int64_t)row * g.ldcb + col] = f2bf(v * gelu_grad(z));
        }
      }
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
// y[m] = h2[m] . w3 + b3 (one wave per row)
__global__ void head_forward_kernel(const bf16* __restrict__ h2, const float* __restrict__ w3,
                                    const float* __restrict__ b3, int M, int H, float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= M) return;
  float s = 0.f;
  for (int n = lane; n < H; n += 64) s += bf2f(h2[row * H + n]) * w3[n];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) out[row] = s + b3[0];
}

// d_out = 2 (out - y) / M and per-block loss partial sums
__global__ void loss_kernel(const float* __restrict__ out, const float* __restrict__ y, int M,
                            float* __restrict__ dout, float* __restrict__ loss_part) {
  __shared__ float s[256];
  const int m = blockIdx.x * 256 + threadIdx.x;
  float l = 0.f;
  if (m < M) {
    const float d = out[m] - y[m];
    dout[m] = 2.f * d / (float)M;
    l = d * d;
  }
  s[threadIdx.x] = l;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = s[0];
}

// dW3[n] = sum_m h2[m][n] dout[m]; db3 = sum_m dout[m]   (thread per n)
// Column sums over the rows of a [M][H] bf16 matrix, each row weighted by
// w[m] (w null: 1): out[n] = sum_m x[m][n] w[m].  A block takes 32 columns
// with 8 row groups (thread (g, c) sums rows g, g + 8, ... of column c, the
// loads of a row group coalesced over 32 columns), then adds the 8 partials
// in a fixed order: deterministic, and the M loads of a column no longer run
// one after another in one thread.  With sum1 != null, the extra block
// (blockIdx.x == H / 32) also writes sum_m w[m] there.
constexpr int CR_COLS = 32, CR_GROUPS = 8;
__global__ __launch_bounds__(CR_COLS * CR_GROUPS) void col_reduce_kernel(const bf16* __restrict__ x,
                                                                          const float* __restrict__ w, int M,
                                                                          int H, float* __restrict__ out,
                                                                          float* __restrict__ sum1) {
  __shared__ float part[CR_GROUPS][CR_COLS];
  const int c = threadIdx.x % CR_COLS, g = threadIdx.x / CR_COLS;
  const int n = blockIdx.x * CR_COLS + c;
  const bool extra = sum1 && (int)blockIdx.x == H / CR_COLS;
  float s = 0.f;
  if (extra) {
    for (int m = threadIdx.x; m < M; m += CR_COLS * CR_GROUPS) s += w[m];
  } else if (n < H) {
#pragma unroll 4
    for (int m = g; m < M; m += CR_GROUPS) s += bf2f(x[(int64_t)m * H + n]) * (w ? w[m] : 1.f);
  }
  if (extra) {
    // a block sum of the per-thread partials, in a fixed order
    __shared__ float all[CR_COLS * CR_GROUPS];
    all[threadIdx.x] = s;
    __syncthreads();
    for (int o = CR_COLS * CR_GROUPS / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) all[threadIdx.x] += all[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) sum1[0] = all[0];
    return;
  }
  part[g][c] = s;
  __syncthreads();
  if (g == 0 && n < H) {
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < CR_GROUPS; ++k) t += part[k][c];
    out[n] = t;
  }
}

// dz2[m][n] = bf16(dout[m] w3[n] GELU'(z2[m][n]))
__global__ void head_back_kernel(const float* __restrict__ dout, const float* __restrict__ w3,
                                 const float* __restrict__ z2, int M, int H, bf16* __restrict__ dz2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * H) return;
  const int m = (int)(i / H), n = (int)(i - (int64_t)m * H);
  dz2[i] = f2bf(dout[m] * w3[n] * gelu_grad(z2[i]));
}

// Adam with decoupled weight decay (AdamW): p -= lr (m^ / (sqrt(v^) + eps) + wd p)
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                            float* __restrict__ m2, int64_t n, float lr, float b1, float b2, float eps,
                            float c1, float c2, float wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  const float pi = p[i];
  p[i] = pi - lr * ((a / c1) / (sqrtf(v / c2) + eps) + wd * pi);
}

__global__ void cast_bf16_kernel(const float* __restrict__ src, bf16* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = f2bf(src[i]);
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void gather_rows_kernel(const uint32_t* __restrict__ X, const float* __restrict__ y,
                                   const int32_t* __restrict__ idx, int rows, int words,
                                   uint32_t* __restrict__ Xb, float* __restrict__ yb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows * words) return;
  const int r = (int)(i / words), w = (int)(i - (int64_t)r * words);
  const int src = idx[r];
  Xb[i] = X[(int64_t)src * words + w];
  if (w == 0) yb[r] = y[src];
}

__device__ __forceinline__ int64_t cycle_of_tick(uint32_t t, uint32_t T, float invT) {
  uint32_t q = (uint32_t)((float)t * invT);
  int32_t r = (int32_t)(t - q * T);
  while (r < 0) { --q; r += (int32_t)T; }
  while (r >= (int32_t)T) { ++q; r -= (int32_t)T; }
  return (int64_t)q;
}

// First index i in [0, cnt) with e[i] >= key (cnt if none): bracket from an
// interpolated guess with doubling steps, then bisect.
__device__ __forceinline__ int64_t lower_bound_ev(const uint32_t* e, int64_t cnt, uint32_t key) {
  if (cnt <= 0) return 0;
  const uint32_t t0 = e[0] >> EV_SHIFT, t1 = e[cnt - 1] >> EV_SHIFT, tk = key >> EV_SHIFT;
  int64_t g = 0;
  if (tk > t0 && t1 > t0) g = (int64_t)min((double)(cnt - 1), (double)(cnt - 1) * ((double)(tk - t0) / (double)(t1 - t0)));
  int64_t lo, hi, step = 1;
  if (e[g] < key) {
    lo = hi = g + 1;
    while (hi < cnt && e[hi] < key) {
      lo = hi + 1;
      hi = min(cnt, hi + step);
      step <<= 1;
    }
  } else {
    lo = hi = g;
    while (lo > 0 && e[lo - 1] >= key) {
      hi = lo - 1;
      lo = max((int64_t)0, lo - step);
      step <<= 1;
    }
  }
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (e[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

constexpr int FB_CYCLES = 2048;  // cycles per feature-bit block

// Bits of the selected features: feature k = (net, kind) with
// feat[k] = kind * n_net + net; bit k of row c is set when the net has an
// event of index >= 1 (a value change; index 0 is the initial value) in cycle
// c0 + c, of any value (kind 0) or to value 1 (kind 1, a rise).  Block (w, cb)
// builds word w of rows [cb * FB_CYCLES, +FB_CYCLES) in LDS: its 32 features'
// events in those cycles are walked together (coalesced), the bits set by LDS
// atomics, and the words written out; no global atomics.
__global__ __launch_bounds__(256) void feature_bits_kernel(const int64_t* __restrict__ net_off,
                                                           const uint32_t* __restrict__ ev,
                                                           const int32_t* __restrict__ feat, int64_t n_net,
                                                           int64_t T, uint32_t cyc_m, int cyc_sh, int64_t c0,
                                                           int64_t n_cyc, int words, uint32_t* __restrict__ X) {
  __shared__ uint32_t bits[FB_CYCLES];
  __shared__ int64_t s_beg[32];
  __shared__ int32_t s_off[33];
  __shared__ uint32_t s_rise;
  const int w = blockIdx.x;
  const int64_t cs = (int64_t)blockIdx.y * FB_CYCLES;
  const int nr = (int)min((int64_t)FB_CYCLES, n_cyc - cs);
  for (int i = threadIdx.x; i < nr; i += blockDim.x) bits[i] = 0;
  if (threadIdx.x == 0) s_rise = 0;
  __syncthreads();
  if (threadIdx.x < 32) {
    const int j = threadIdx.x;
    const int64_t f = feat[(int64_t)w * 32 + j];
    const bool rise = f >= n_net;
    const int64_t n = rise ? f - n_net : f;
    const int64_t b = net_off[n], cnt = net_off[n + 1] - b;
    const int64_t tlo = (c0 + cs) * T, thi = (c0 + cs + nr) * T;
    const uint32_t klo = tlo >= (int64_t)EV_TIME_INF ? EV_INF : (uint32_t)(tlo << EV_SHIFT);
    const uint32_t khi = thi >= (int64_t)EV_TIME_INF ? EV_INF : (uint32_t)(thi << EV_SHIFT);
    const int64_t i0 = max((int64_t)1, lower_bound_ev(ev + b, cnt, klo));
    const int64_t i1 = lower_bound_ev(ev + b, cnt, khi);
    s_beg[j] = b + i0;
    s_off[j + 1] = (int32_t)max((int64_t)0, i1 - i0);
    if (rise) atomicOr(&s_rise, 1u << j);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    s_off[0] = 0;
    for (int j = 0; j < 32; ++j) s_off[j + 1] += s_off[j];
  }
  __syncthreads();
  const uint32_t rise_mask = s_rise;
  const int64_t cbase = c0 + cs;
  // one feature per wave at a time, its events in the window read by the 64
  // lanes together, four loads in flight per lane
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, n_wave = blockDim.x >> 6;
  for (int j = wave; j < 32; j += n_wave) {
    const uint32_t* e = ev + s_beg[j];
    const int cnt = s_off[j + 1] - s_off[j];
    const bool rise = (rise_mask >> j) & 1u;
    for (int i = lane; i < cnt; i += 4 * 64) {
      uint32_t wd[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) wd[u] = i + u * 64 < cnt ? e[i + u * 64] : 0u;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (i + u * 64 >= cnt) break;
        if (rise && (wd[u] & 3u) != 1u) continue;
        const uint32_t t = wd[u] >> EV_SHIFT;
        const int64_t c = (int64_t)(((uint64_t)cyc_m * (uint64_t)t) >> cyc_sh) - cbase;
        atomicOr(&bits[c], 1u << j);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nr; i += blockDim.x) X[(cs + i) * words + w] = bits[i];
}

constexpr int WS_CYCLES = 2048;  // cycles per weighted-sum window
constexpr int WS_NETS = 256;     // selected nets per weighted-sum block

// First event index (>= 1) of each selected net at or after the start of
// each window: idx[k * (n_win + 1) + j] for window j of net sel[k], j = n_win
// the end of the range; one thread per (net, window), so the searches run in
// parallel rather than one after another in the summing waves.
__global__ void net_window_start_kernel(const int64_t* __restrict__ net_off, const uint32_t* __restrict__ ev,
                                        const int32_t* __restrict__ sel, int32_t n_sel, int32_t n_win, int64_t T,
                                        int64_t c0, int64_t n_cyc, int32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_sel * (n_win + 1)) return;
  const int64_t k = i / (n_win + 1);
  const int j = (int)(i - k * (n_win + 1));
  const int32_t n = sel[k];
  const int64_t b = net_off[n], cnt = net_off[n + 1] - b;
  const int64_t t = (c0 + min((int64_t)j * WS_CYCLES, n_cyc)) * T;
  const uint32_t key = t >= (int64_t)EV_TIME_INF ? EV_INF : (uint32_t)(t << EV_SHIFT);
  idx[i] = (int32_t)max((int64_t)1, lower_bound_ev(ev + b, cnt, key));
}

// Weighted toggle sums: out[c] += sum over the selected nets k of wq[k] x
// (value changes of sel[k] in cycle c0 + c; event index >= 1).  Block (g, j):
// nets [g * WS_NETS, +WS_NETS) over window j, one net per wave at a time (its
// events in the window, from the window index, read by the 64 lanes
// together), int64 LDS bins, one global atomic per nonzero bin.
__global__ __launch_bounds__(256) void net_weight_sum_kernel(const int64_t* __restrict__ net_off,
                                                             const uint32_t* __restrict__ ev,
                                                             const int32_t* __restrict__ sel,
                                                             const int32_t* __restrict__ wq,
                                                             const int32_t* __restrict__ idx, int32_t n_sel,
                                                             int32_t n_win, uint32_t cyc_m, int cyc_sh, int64_t c0,
                                                             int64_t n_cyc, unsigned long long* __restrict__ out) {
  __shared__ long long bins[WS_CYCLES];
  // the block's nets: weight, first event and the window's event range,
  // loaded by all threads at once (one dependent chain for the block instead
  // of one per net and wave)
  __shared__ long long s_w[WS_NETS];
  __shared__ int64_t s_base[WS_NETS];
  __shared__ int32_t s_i0[WS_NETS], s_i1[WS_NETS];
  const int j = blockIdx.y;
  const int64_t cs = (int64_t)j * WS_CYCLES;
  const int nr = (int)min((int64_t)WS_CYCLES, n_cyc - cs);
  const int k0 = (int)blockIdx.x * WS_NETS;
  const int nk = min(n_sel - k0, WS_NETS);
  for (int q = threadIdx.x; q < nk; q += blockDim.x) {
    const int k = k0 + q;
    s_w[q] = wq[k];
    s_base[q] = net_off[sel[k]];
    s_i0[q] = idx[(int64_t)k * (n_win + 1) + j];
    s_i1[q] = idx[(int64_t)k * (n_win + 1) + j + 1];
  }
  for (int i = threadIdx.x; i < nr; i += blockDim.x) bins[i] = 0;
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, n_wave = blockDim.x >> 6;
  const int64_t cbase = c0 + cs;
  for (int q = wave; q < nk; q += n_wave) {
    const long long w = s_w[q];
    const uint32_t* e = ev + s_base[q];
    const int32_t i0 = s_i0[q], i1 = s_i1[q];
    // four loads in flight per lane before their events are binned (the
    // loop is bound by load latency, not by the LDS atomics)
    for (int32_t i = i0 + lane; i < i1; i += 4 * 64) {
      uint32_t wd[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) wd[u] = i + u * 64 < i1 ? e[i + u * 64] : 0u;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (i + u * 64 >= i1) break;
        const uint32_t t = wd[u] >> EV_SHIFT;
        const int64_t c = (int64_t)(((uint64_t)cyc_m * (uint64_t)t) >> cyc_sh) - cbase;
        atomicAdd((unsigned long long*)&bins[c], (unsigned long long)w);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nr; i += blockDim.x)
    if (bins[i] != 0) atomicAdd(&out[cs + i], (unsigned long long)bins[i]);
}

// Per net over the cycles with mask[c] != 0 (all when mask is null): the
// number of cycles with a value change (kind 0) and with a change to 1 (kind
// 1), and the sums of the quantised target over those cycles (int64: exact
// and order-free); outputs are [2][n_net].  One block-iteration per 256
// events (chunk_off as for the glitch annotation).
__global__ __launch_bounds__(256) void net_corr_kernel(const int64_t* __restrict__ net_off,
                                                       const uint32_t* __restrict__ ev,
                                                       const int64_t* __restrict__ chunk_off,
                                                       const int32_t* __restrict__ chunk_net, int64_t n_net,
                                                       uint32_t T, float invT, int64_t n_cyc,
                                                       const uint8_t* __restrict__ mask,
                                                       const long long* __restrict__ yq,
                                                       unsigned long long* __restrict__ cnt_out,
                                                       unsigned long long* __restrict__ sum_out, int64_t n_chunk) {
  // blocks stride over the chunks (a dispatch counts work-items in 32 bits);
  // every lane of a block runs the same iterations, so the wave sums below
  // see all 64 lanes
  for (int64_t chunk = blockIdx.x; chunk < n_chunk; chunk += gridDim.x) {
    const int64_t n = chunk_net[chunk];
    const int64_t b = net_off[n], cnt = net_off[n + 1] - b;
    const int64_t i = (chunk - chunk_off[n]) * 256 + threadIdx.x;
    bool take = false, take_rise = false;
    int64_t c = 0;
    if (i >= 1 && i < cnt) {
      const uint32_t w = ev[b + i];
      c = cycle_of_tick(w >> EV_SHIFT, T, invT);
      if (c < n_cyc && (!mask || mask[c])) {
        take = i == 1 || cycle_of_tick(ev[b + i - 1] >> EV_SHIFT, T, invT) != c;
        if ((w & 3u) == 1u) {
          // the first change to 1 of this net in cycle c
          take_rise = true;
          for (int64_t j = i - 1; j >= 1; --j) {
            const uint32_t wj = ev[b + j];
            if (cycle_of_tick(wj >> EV_SHIFT, T, invT) != c) break;
            if ((wj & 3u) == 1u) {
              take_rise = false;
              break;
            }
          }
        }
      }
    }
    unsigned long long k = take ? 1ull : 0ull, kr = take_rise ? 1ull : 0ull;
    long long s = take ? yq[c] : 0, sr = take_rise ? yq[c] : 0;
    for (int off = 32; off > 0; off >>= 1) {
      k += __shfl_xor(k, off);
      s += __shfl_xor(s, off);
      kr += __shfl_xor(kr, off);
      sr += __shfl_xor(sr, off);
    }
    if ((threadIdx.x & 63) == 0) {
      if (k) {
        atomicAdd(&cnt_out[n], k);
        atomicAdd(&sum_out[n], (unsigned long long)s);
      }
      if (kr) {
        atomicAdd(&cnt_out[n_net + n], kr);
        atomicAdd(&sum_out[n_net + n], (unsigned long long)sr);
      }
    }
  }
}

template <int AM, int BMODE, int EPI>
hipError_t gemm(const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0) return hipSuccess;
  if (g.N % BN || g.K % BK || (AM == A_COL && g.M % 16)) return hipErrorInvalidValue;
  dim3 grid(g.N / BN, (g.M + BM - 1) / BM);
  gemm_kernel<AM, BMODE, EPI><<<grid, dim3(256), 0, s>>>(g);
  return hipGetLastError();
}

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace mlp
}  // namespace gtp

using namespace gtp;
using namespace gtp::mlp;

// ---------------------------------------------------------------------------
// Scratch of the allocation-free inference (gtp_mlp_predict_noalloc): feature
// rows, the two hidden layers' activations and the outputs of one chunk of
// rows.  Every GPU the library runs on holds its own copy until the process
// ends.  Host code never passes `scratch` itself to a kernel or a copy (in host
// code the name is a host-side stand-in, not a device address): it uses the
// address hipGetSymbolAddress returns on each device, kept in noalloc_base[].
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
unsigned char* noalloc_base[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on device dev, which must be current;
// resolved once per device.
hipError_t noalloc_resolve(int dev, unsigned char** out) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (!noalloc_base[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    noalloc_base[dev] = static_cast<unsigned char*>(p);
  }
  *out = noalloc_base[dev];
  return hipSuccess;
}
}  // namespace

struct gtp_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  int K = 0, H = 0, words = 0;
  std::string err;
  // parameters (FP32 master) in one buffer: W1[H][K] b1[H] W2[H][H] b2[H] w3[H] b3[1]
  int64_t n_param = 0;
  float *P = nullptr, *G = nullptr, *M1 = nullptr, *M2 = nullptr;
  bf16 *W1b = nullptr, *W2b = nullptr;
  float y_mean = 0.f, y_std = 1.f;
  int64_t adam_t = 0;
  // batch workspace
  int cap = 0;
  uint32_t* Xb = nullptr;
  float *yb = nullptr, *z1 = nullptr, *z2 = nullptr, *out = nullptr, *dout = nullptr, *lossp = nullptr;
  bf16 *h1 = nullptr, *h2 = nullptr, *dz1 = nullptr, *dz2 = nullptr;
  int32_t* idx = nullptr;
  // predictions over a loaded trace: kept between calls (no free, so no
  // implicit device synchronisation while another thread's runs are queued)
  uint32_t* tr_X = nullptr;
  float* tr_out = nullptr;
  int32_t* tr_net = nullptr;
  std::vector<int32_t> tr_net_h;  // what tr_net holds
  int64_t tr_X_n = 0, tr_out_n = 0;

  float* W1() { return P; }
  float* b1() { return P + (int64_t)H * K; }
  float* W2() { return b1() + H; }
  float* b2() { return W2() + (int64_t)H * H; }
  float* w3() { return b2() + H; }
  float* b3() { return w3() + H; }
  float* g(int which) {
    float* base = G;
    const int64_t off[6] = {0, (int64_t)H * K, (int64_t)H * K + H, (int64_t)H * K + H + (int64_t)H * H,
                            (int64_t)H * K + 2 * H + (int64_t)H * H, (int64_t)H * K + 3 * H + (int64_t)H * H};
    return base + off[which];
  }
  gtp_status fail(gtp_status s, const std::string& m) {
    err = m;
    return s;
  }
  void release_batch() {
    for (void* p : {(void*)Xb, (void*)yb, (void*)z1, (void*)z2, (void*)out, (void*)dout, (void*)lossp,
                    (void*)h1, (void*)h2, (void*)dz1, (void*)dz2, (void*)idx})
      if (p) (void)hipFree(p);
    Xb = nullptr; yb = z1 = z2 = out = dout = lossp = nullptr;
    h1 = h2 = dz1 = dz2 = nullptr; idx = nullptr;
    cap = 0;
  }
  hipError_t reserve(int n) {
    if (n <= cap) return hipSuccess;
    release_batch();
    const int c = (n + 127) / 128 * 128;
    hipError_t e;
#define ALLOC(p, bytes) if ((e = hipMalloc((void**)&p, (bytes))) != hipSuccess) return e;
    ALLOC(Xb, (size_t)c * words * 4);
    ALLOC(yb, (size_t)c * 4);
    ALLOC(z1, (size_t)c * H * 4);
    ALLOC(z2, (size_t)c * H * 4);
    ALLOC(out, (size_t)c * 4);
    ALLOC(dout, (size_t)c * 4);
    ALLOC(lossp, (size_t)(c / 256 + 2) * 4);
    ALLOC(h1, (size_t)c * H * 2);
    ALLOC(h2, (size_t)c * H * 2);
    ALLOC(dz1, (size_t)c * H * 2);
    ALLOC(dz2, (size_t)c * H * 2);
    ALLOC(idx, (size_t)c * 4);
#undef ALLOC
    cap = c;
    return hipSuccess;
  }
};

#define MCHK(m, expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return (m)->fail(e_ == hipErrorOutOfMemory ? GTP_ERR_OOM : GTP_ERR_DEVICE,             \
                       std::string(#expr) + ": " + hipGetErrorString(e_));                   \
  } while (0)

namespace {

hipError_t refresh_bf16(gtp_mlp* m) {
  const int64_t n1 = (int64_t)m->H * m->K, n2 = (int64_t)m->H * m->H;
  cast_bf16_kernel<<<nblk(n1), 256, 0, m->stream>>>(m->W1(), m->W1b, n1);
  cast_bf16_kernel<<<nblk(n2), 256, 0, m->stream>>>(m->W2(), m->W2b, n2);
  return hipGetLastError();
}

// forward pass over n rows of bit features already on the device
// (drop_seed != 0: training-time input dropout)
// (h1, h2: the bf16 activations of the two hidden layers, [n][H]; null: the
// model's batch workspace)
hipError_t forward(gtp_mlp* m, const uint32_t* X, int n, bool keep_z, float* out, uint32_t drop_seed = 0,
                   bf16* h1 = nullptr, bf16* h2 = nullptr) {
  const int H = m->H;
  if (!h1) h1 = m->h1;
  if (!h2) h2 = m->h2;
  GemmArgs g{};
  g.M = n; g.N = H; g.K = m->K;
  g.A = X; g.lda = m->words;
  g.drop_seed = drop_seed;
  g.B = m->W1b; g.ldb = m->K;
  g.C = keep_z ? m->z1 : nullptr; g.ldc = H;
  g.Cb = h1; g.ldcb = H;
  g.bias = m->b1();
  hipError_t e = gemm<A_BITS_ROW, B_TRANS, EPI_BIAS_GELU>(g, m->stream);
  if (e != hipSuccess) return e;
  GemmArgs g2{};
  g2.M = n; g2.N = H; g2.K = H;
  g2.A = h1; g2.lda = H;
  g2.B = m->W2b; g2.ldb = H;
  g2.C = keep_z ? m->z2 : nullptr; g2.ldc = H;
  g2.Cb = h2; g2.ldcb = H;
  g2.bias = m->b2();
  if ((e = gemm<A_ROW, B_TRANS, EPI_BIAS_GELU>(g2, m->stream)) != hipSuccess) return e;
  head_forward_kernel<<<nblk(n, 4), 256, 0, m->stream>>>(h2, m->w3(), m->b3(), n, H, out);
  return hipGetLastError();
}

// forward + backward of one batch held in m->Xb / m->yb (standardised y);
// gradients into m->G; loss partials into m->lossp
hipError_t backward(gtp_mlp* m, int n, uint32_t drop_seed = 0) {
  const int H = m->H;
  hipError_t e = forward(m, m->Xb, n, true, m->out, drop_seed);
  if (e != hipSuccess) return e;
  loss_kernel<<<nblk(n), 256, 0, m->stream>>>(m->out, m->yb, n, m->dout, m->lossp);
  // dw3 = h2^T dout, db3 = sum dout
  col_reduce_kernel<<<(H + CR_COLS - 1) / CR_COLS + 1, CR_COLS * CR_GROUPS, 0, m->stream>>>(m->h2, m->dout, n, H,
                                                                                          m->g(4), m->g(5));
  head_back_kernel<<<nblk((int64_t)n * H), 256, 0, m->stream>>>(m->dout, m->w3(), m->z2, n, H, m->dz2);
  // dW2 = dz2^T h1
  GemmArgs a{};
  a.M = H; a.N = H; a.K = n;
  a.A = m->dz2; a.lda = H;
  a.B = m->h1; a.ldb = H;
  a.C = m->g(2); a.ldc = H;
  if ((e = gemm<A_COL, B_ROW, EPI_F32>(a, m->stream)) != hipSuccess) return e;
  col_reduce_kernel<<<(H + CR_COLS - 1) / CR_COLS, CR_COLS * CR_GROUPS, 0, m->stream>>>(m->dz2, nullptr, n, H,
                                                                                      m->g(3), nullptr);
  // dz1 = (dz2 W2) * GELU'(z1)
  GemmArgs b{};
  b.M = n; b.N = H; b.K = H;
  b.A = m->dz2; b.lda = H;
  b.B = m->W2b; b.ldb = H;
  b.Cb = m->dz1; b.ldcb = H;
  b.zin = m->z1; b.ldz = H;
  if ((e = gemm<A_ROW, B_ROW, EPI_GELU_BACK>(b, m->stream)) != hipSuccess) return e;
  col_reduce_kernel<<<(H + CR_COLS - 1) / CR_COLS, CR_COLS * CR_GROUPS, 0, m->stream>>>(m->dz1, nullptr, n, H,
                                                                                      m->g(1), nullptr);
  // dW1 = dz1^T X
  GemmArgs c{};
  c.M = H; c.N = m->K; c.K = n;
  c.A = m->dz1; c.lda = H;
  c.B = m->Xb; c.ldb = m->words;
  c.C = m->g(0); c.ldc = m->K;
  c.drop_seed = drop_seed;
  if ((e = gemm<A_COL, B_BITS_ROW, EPI_F32>(c, m->stream)) != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace

GTP_API gtp_status gtp_mlp_create(const gtp_mlp_config* cfg, gtp_mlp** out) {
  if (!cfg || !out || cfg->n_in <= 0 || cfg->n_in % 128 || cfg->n_hidden <= 0 || cfg->n_hidden % 128)
    return GTP_ERR_ARG;
  *out = nullptr;
  auto* m = new (std::nothrow) gtp_mlp();
  if (!m) return GTP_ERR_OOM;
  m->K = cfg->n_in;
  m->H = cfg->n_hidden;
  m->words = m->K / 32;
  int dev = cfg->device;
  if (dev < 0 && hipGetDevice(&dev) != hipSuccess) dev = 0;
  m->device = dev;
  if (hipSetDevice(dev) != hipSuccess || hipStreamCreate(&m->stream) != hipSuccess) {
    delete m;
    return GTP_ERR_DEVICE;
  }
  const int K = m->K, H = m->H;
  m->n_param = (int64_t)H * K + H + (int64_t)H * H + H + H + 1;
  size_t pb = (size_t)m->n_param * 4;
  if (hipMalloc((void**)&m->P, pb) != hipSuccess || hipMalloc((void**)&m->G, pb) != hipSuccess ||
      hipMalloc((void**)&m->M1, pb) != hipSuccess || hipMalloc((void**)&m->M2, pb) != hipSuccess ||
      hipMalloc((void**)&m->W1b, (size_t)H * K * 2) != hipSuccess ||
      hipMalloc((void**)&m->W2b, (size_t)H * H * 2) != hipSuccess) {
    gtp_mlp_destroy(m);
    return GTP_ERR_OOM;
  }
  // He-style initialisation from a seeded generator
  std::vector<float> p(m->n_param, 0.f);
  std::mt19937_64 rng(cfg->seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  const float s1 = std::sqrt(2.f / K), s2 = std::sqrt(2.f / H), s3 = std::sqrt(1.f / H);
  for (int64_t i = 0; i < (int64_t)H * K; ++i) p[i] = nd(rng) * s1;
  float* w2 = p.data() + (int64_t)H * K + H;
  for (int64_t i = 0; i < (int64_t)H * H; ++i) w2[i] = nd(rng) * s2;
  float* w3 = w2 + (int64_t)H * H + H;
  for (int i = 0; i < H; ++i) w3[i] = nd(rng) * s3;
  if (hipMemcpy(m->P, p.data(), pb, hipMemcpyHostToDevice) != hipSuccess) {
    gtp_mlp_destroy(m);
    return GTP_ERR_DEVICE;
  }
  std::vector<float> zero(m->n_param, 0.f);
  (void)hipMemcpy(m->M1, zero.data(), pb, hipMemcpyHostToDevice);
  (void)hipMemcpy(m->M2, zero.data(), pb, hipMemcpyHostToDevice);
  (void)refresh_bf16(m);
  (void)hipStreamSynchronize(m->stream);
  // the allocation-free inference's scratch on this device (a failure here is
  // reported by that entry point, which resolves it again)
  {
    unsigned char* base = nullptr;
    if (noalloc_resolve(dev, &base) != hipSuccess) (void)hipGetLastError();
  }
  *out = m;
  return GTP_OK;
}

GTP_API void gtp_mlp_destroy(gtp_mlp* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  m->release_batch();
  for (void* p : {(void*)m->P, (void*)m->G, (void*)m->M1, (void*)m->M2, (void*)m->W1b, (void*)m->W2b,
                  (void*)m->tr_X, (void*)m->tr_out, (void*)m->tr_net})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

GTP_API uint32_t gtp_keep_mask(uint32_t seed, uint32_t row, uint32_t word) { return keep_mask(seed, row, word); }

GTP_API const char* gtp_mlp_last_error(const gtp_mlp* m) { return m ? m->err.c_str() : "null model"; }

GTP_API int64_t gtp_mlp_param_count(const gtp_mlp* m) { return m ? m->n_param : 0; }

GTP_API gtp_status gtp_mlp_get_params(gtp_mlp* m, float* params, float* y_mean_std) {
  if (!m || !params) return GTP_ERR_ARG;
  MCHK(m, hipSetDevice(m->device));
  MCHK(m, hipMemcpy(params, m->P, (size_t)m->n_param * 4, hipMemcpyDeviceToHost));
  if (y_mean_std) {
    y_mean_std[0] = m->y_mean;
    y_mean_std[1] = m->y_std;
  }
  return GTP_OK;
}

GTP_API gtp_status gtp_mlp_set_params(gtp_mlp* m, const float* params, const float* y_mean_std) {
  if (!m || !params) return GTP_ERR_ARG;
  MCHK(m, hipSetDevice(m->device));
  MCHK(m, hipMemcpy(m->P, params, (size_t)m->n_param * 4, hipMemcpyHostToDevice));
  if (y_mean_std) {
    m->y_mean = y_mean_std[0];
    m->y_std = y_mean_std[1];
  }
  MCHK(m, refresh_bf16(m));
  MCHK(m, hipStreamSynchronize(m->stream));
  return GTP_OK;
}

// Gradients of the mean squared error of one batch (targets already
// standardised by the caller), without updating the model.
GTP_API gtp_status gtp_mlp_gradients(gtp_mlp* m, const uint32_t* bits, const float* y, int32_t n,
                                     uint32_t drop_seed, float* grads, float* loss) {
  if (!m || !bits || !y || !grads || n <= 0 || n % 32) return m ? m->fail(GTP_ERR_ARG, "n must be a positive multiple of 32") : GTP_ERR_ARG;
  MCHK(m, hipSetDevice(m->device));
  MCHK(m, m->reserve(n));
  MCHK(m, hipMemcpyAsync(m->Xb, bits, (size_t)n * m->words * 4, hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemcpyAsync(m->yb, y, (size_t)n * 4, hipMemcpyHostToDevice, m->stream));
  MCHK(m, backward(m, n, drop_seed));
  MCHK(m, hipMemcpyAsync(grads, m->G, (size_t)m->n_param * 4, hipMemcpyDeviceToHost, m->stream));
  std::vector<float> lp(nblk(n));
  MCHK(m, hipMemcpyAsync(lp.data(), m->lossp, lp.size() * 4, hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  if (loss) *loss = std::accumulate(lp.begin(), lp.end(), 0.f) / (float)n;
  return GTP_OK;
}

// Train on n rows; the last tc->val_rows of them are held out: after every
// epoch their mean squared error is measured and the parameters of the best
// epoch are kept.  epoch_loss[2 ep] = mean training loss of epoch ep,
// epoch_loss[2 ep + 1] = its validation loss (NaN without validation rows),
// both on standardised targets.
GTP_API gtp_status gtp_mlp_train(gtp_mlp* m, const uint32_t* bits, const float* y, int64_t n,
                                 const gtp_train_config* tc, float* epoch_loss) {
  if (!m) return GTP_ERR_ARG;
  if (!bits || !y || !tc || tc->batch <= 0 || tc->batch % 32 || tc->epochs <= 0 || tc->val_rows < 0 ||
      n - tc->val_rows < tc->batch || !(tc->weight_decay >= 0.f) || (tc->input_dropout != 0 && tc->input_dropout != 1))
    return m->fail(GTP_ERR_ARG, "gtp_mlp_train: bad arguments (batch a multiple of 32, at least one batch of "
                                "training rows besides the validation rows, weight_decay >= 0)");
  MCHK(m, hipSetDevice(m->device));
  const int64_t n_val = tc->val_rows, n_tr = n - n_val;
  // standardise the targets with the training rows' statistics
  double mu = 0, var = 0;
  for (int64_t i = 0; i < n_tr; ++i) mu += y[i];
  mu /= (double)n_tr;
  for (int64_t i = 0; i < n_tr; ++i) var += (y[i] - mu) * (y[i] - mu);
  const double sd = std::sqrt(var / (double)n_tr) > 0 ? std::sqrt(var / (double)n_tr) : 1.0;
  m->y_mean = (float)mu;
  m->y_std = (float)sd;
  std::vector<float> ys(n);
  for (int64_t i = 0; i < n; ++i) ys[i] = (float)((y[i] - mu) / sd);
  uint32_t* dX = nullptr;
  float *dy = nullptr, *best = nullptr;
  struct Free3 {
    void **a, **b, **c;
    ~Free3() {
      for (void** p : {a, b, c})
        if (*p) (void)hipFree(*p);
    }
  } fr{(void**)&dX, (void**)&dy, (void**)&best};
  MCHK(m, hipMalloc((void**)&dX, (size_t)n * m->words * 4));
  MCHK(m, hipMalloc((void**)&dy, (size_t)n * 4));
  if (n_val > 0) MCHK(m, hipMalloc((void**)&best, (size_t)m->n_param * 4));
  MCHK(m, hipMemcpyAsync(dX, bits, (size_t)n * m->words * 4, hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemcpyAsync(dy, ys.data(), (size_t)n * 4, hipMemcpyHostToDevice, m->stream));
  const int B = tc->batch;
  MCHK(m, m->reserve(B));
  std::mt19937_64 rng(tc->seed);
  std::vector<int32_t> perm(n_tr);
  std::iota(perm.begin(), perm.end(), 0);
  const int64_t steps_per_epoch = n_tr / B;
  const int64_t total_steps = steps_per_epoch * tc->epochs;
  int64_t step = 0;
  std::vector<float> lp(nblk(B)), vp;
  double best_val = INFINITY;
  for (int ep = 0; ep < tc->epochs; ++ep) {
    std::shuffle(perm.begin(), perm.end(), rng);
    double eloss = 0;
    int64_t n_loss = 0;
    for (int64_t s = 0; s < steps_per_epoch; ++s, ++step) {
      MCHK(m, hipMemcpyAsync(m->idx, perm.data() + s * B, (size_t)B * 4, hipMemcpyHostToDevice, m->stream));
      gather_rows_kernel<<<nblk((int64_t)B * m->words), 256, 0, m->stream>>>(dX, dy, m->idx, B, m->words, m->Xb, m->yb);
      // a fresh, never-zero dropout seed per step
      const uint32_t ds = tc->input_dropout ? (uint32_t)(rng() | 1u) : 0u;
      MCHK(m, backward(m, B, ds));
      // cosine learning-rate decay
      const float lr = (float)(tc->lr * 0.5 * (1.0 + std::cos(M_PI * (double)step / (double)total_steps)));
      ++m->adam_t;
      const float c1 = 1.f - std::pow(0.9f, (float)m->adam_t), c2 = 1.f - std::pow(0.999f, (float)m->adam_t);
      adam_kernel<<<nblk(m->n_param), 256, 0, m->stream>>>(m->P, m->G, m->M1, m->M2, m->n_param, lr, 0.9f,
                                                            0.999f, 1e-8f, c1, c2, tc->weight_decay);
      MCHK(m, refresh_bf16(m));
      if (epoch_loss && (s == steps_per_epoch - 1 || s % 16 == 0)) {
        MCHK(m, hipMemcpyAsync(lp.data(), m->lossp, lp.size() * 4, hipMemcpyDeviceToHost, m->stream));
        MCHK(m, hipStreamSynchronize(m->stream));
        eloss += std::accumulate(lp.begin(), lp.end(), 0.0) / B;
        ++n_loss;
      }
    }
    // the host permutation is rewritten next epoch: let the copies finish
    MCHK(m, hipStreamSynchronize(m->stream));
    double vloss = NAN;
    if (n_val > 0) {
      double se = 0;
      for (int64_t b0 = 0; b0 < n_val; b0 += m->cap) {
        const int c = (int)std::min<int64_t>(m->cap, n_val - b0);
        MCHK(m, forward(m, dX + (n_tr + b0) * m->words, c, false, m->out));
        vp.resize(c);
        MCHK(m, hipMemcpyAsync(vp.data(), m->out, (size_t)c * 4, hipMemcpyDeviceToHost, m->stream));
        MCHK(m, hipStreamSynchronize(m->stream));
        for (int i = 0; i < c; ++i) {
          const double d = (double)vp[i] - (double)ys[n_tr + b0 + i];
          se += d * d;
        }
      }
      vloss = se / (double)n_val;
      if (vloss < best_val) {
        best_val = vloss;
        MCHK(m, hipMemcpyAsync(best, m->P, (size_t)m->n_param * 4, hipMemcpyDeviceToDevice, m->stream));
      }
    }
    if (epoch_loss) {
      epoch_loss[2 * ep] = (float)(n_loss ? eloss / (double)n_loss : NAN);
      epoch_loss[2 * ep + 1] = (float)vloss;
    }
  }
  if (n_val > 0) {
    MCHK(m, hipMemcpyAsync(m->P, best, (size_t)m->n_param * 4, hipMemcpyDeviceToDevice, m->stream));
    MCHK(m, refresh_bf16(m));
  }
  MCHK(m, hipStreamSynchronize(m->stream));
  return GTP_OK;
}

GTP_API gtp_status gtp_mlp_predict(gtp_mlp* m, const uint32_t* bits, int64_t n, float* out) {
  if (!m || !bits || !out || n <= 0) return GTP_ERR_ARG;
  MCHK(m, hipSetDevice(m->device));
  const int64_t chunk = 1 << 18;
  MCHK(m, m->reserve((int)std::min<int64_t>(n, chunk)));
  for (int64_t b = 0; b < n; b += chunk) {
    const int c = (int)std::min<int64_t>(chunk, n - b);
    MCHK(m, hipMemcpyAsync(m->Xb, bits + b * m->words, (size_t)c * m->words * 4, hipMemcpyHostToDevice, m->stream));
    MCHK(m, forward(m, m->Xb, c, false, m->out));
    MCHK(m, hipMemcpyAsync(out + b, m->out, (size_t)c * 4, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
  }
  for (int64_t i = 0; i < n; ++i) out[i] = out[i] * m->y_std + m->y_mean;
  return GTP_OK;
}

// Rows of one chunk of gtp_mlp_predict_noalloc and the scratch offsets of its
// parts (256-byte aligned): feature words, layer-1 and layer-2 activations,
// outputs.  0 rows: a 128-row chunk of this model does not fit the scratch.
namespace {
struct NoallocLayout {
  int64_t rows = 0;
  size_t x = 0, h1 = 0, h2 = 0, out = 0;
};

NoallocLayout noalloc_layout(const gtp_mlp* m) {
  const auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t per_row = (size_t)m->words * 4 + (size_t)m->H * 4 + 4;
  NoallocLayout L;
  int64_t r = (int64_t)((NOALLOC_SCRATCH_BYTES - 4 * 256) / per_row) / 128 * 128;
  for (; r >= 128; r -= 128) {
    L.x = 0;
    L.h1 = al(L.x + (size_t)r * m->words * 4);
    L.h2 = al(L.h1 + (size_t)r * m->H * 2);
    L.out = al(L.h2 + (size_t)r * m->H * 2);
    if (L.out + (size_t)r * 4 <= NOALLOC_SCRATCH_BYTES) break;
  }
  L.rows = r >= 128 ? r : 0;
  return L;
}
}  // namespace

GTP_API gtp_status gtp_mlp_predict_noalloc(gtp_mlp* m, const uint32_t* bits, int64_t n, float* out) {
  if (!m || !bits || !out || n <= 0) return GTP_ERR_ARG;
  const NoallocLayout L = noalloc_layout(m);
  if (L.rows == 0)
    return m->fail(GTP_ERR_LIMIT, "gtp_mlp_predict_noalloc: n_in " + std::to_string(m->K) + " and n_hidden " +
                                      std::to_string(m->H) + " need more than the " +
                                      std::to_string(NOALLOC_SCRATCH_BYTES >> 20) +
                                      " MiB scratch for a 128-row chunk");
  MCHK(m, hipSetDevice(m->device));
  unsigned char* base = nullptr;
  MCHK(m, noalloc_resolve(m->device, &base));
  uint32_t* X = reinterpret_cast<uint32_t*>(base + L.x);
  bf16* h1 = reinterpret_cast<bf16*>(base + L.h1);
  bf16* h2 = reinterpret_cast<bf16*>(base + L.h2);
  float* o = reinterpret_cast<float*>(base + L.out);
  for (int64_t b = 0; b < n; b += L.rows) {
    const int c = (int)std::min<int64_t>(L.rows, n - b);
    MCHK(m, hipMemcpyAsync(X, bits + b * m->words, (size_t)c * m->words * 4, hipMemcpyHostToDevice, m->stream));
    MCHK(m, forward(m, X, c, false, o, 0, h1, h2));
    MCHK(m, hipMemcpyAsync(out + b, o, (size_t)c * 4, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
  }
  for (int64_t i = 0; i < n; ++i) out[i] = out[i] * m->y_std + m->y_mean;
  return GTP_OK;
}

// ---------------------------------------------------------------------------
// device-resident paths over the context's event store
// ---------------------------------------------------------------------------
namespace {

// Feature selection upload: net of each feature and the first 256-event block
// of each feature's events.
struct FeatureSel {
  int32_t* d_net = nullptr;
  bool own = false;
  ~FeatureSel() {
    if (own && d_net) (void)hipFree(d_net);
  }
  hipError_t build(const TraceDev& t, int K, const int32_t* nets, hipStream_t s) {
    (void)t;
    hipError_t e = hipMalloc((void**)&d_net, (size_t)K * 4);
    own = e == hipSuccess;
    if (e == hipSuccess) e = hipMemcpyAsync(d_net, nets, (size_t)K * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
  }
  // the same into a caller-kept device buffer of K entries
  hipError_t use(int32_t* buf, int K, const int32_t* nets, hipStream_t s) {
    d_net = buf;
    own = false;
    hipError_t e = hipMemcpyAsync(d_net, nets, (size_t)K * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
  }
  void launch(const TraceDev& t, int K, int64_t c0, int64_t n_cyc, int words, uint32_t* X, hipStream_t s) const {
    (void)K;
    uint32_t m = 0;
    int32_t sh = 0;
    cycle_magic(t.T, &m, &sh);
    const dim3 grid((unsigned)words, (unsigned)((n_cyc + FB_CYCLES - 1) / FB_CYCLES));
    feature_bits_kernel<<<grid, 256, 0, s>>>(t.net_off, t.ev, d_net, t.n_net, t.T, m, sh, c0, n_cyc, words, X);
  }
};

// feature ids are kind * n_net + net, kind 0 (any change) or 1 (rise)
bool check_nets(const TraceDev& t, int K, const int32_t* nets) {
  for (int k = 0; k < K; ++k)
    if (nets[k] < 0 || nets[k] >= 2 * t.n_net) return false;
  return true;
}

}  // namespace

// Pearson correlation of each feature's per-cycle bit with y over the cycles
// where mask[c] != 0 (all cycles when mask is null); corr is [2][n_net]:
// kind 0 (any change), kind 1 (rise).
GTP_API gtp_status gtp_feature_correlation(gtp_ctx* ctx, const double* y, const uint8_t* mask, float* corr) {
  TraceDev t;
  gtp_status st = gtp::trace_dev(ctx, &t);
  if (st != GTP_OK) return st;
  if (!y || !corr) return GTP_ERR_ARG;
  if (hipSetDevice(t.device) != hipSuccess) return GTP_ERR_DEVICE;
  const int64_t C = t.n_cycle, N = t.n_net;
  double mu = 0, amax = 0;
  int64_t nv = 0;
  for (int64_t c = 0; c < C; ++c)
    if (!mask || mask[c]) {
      mu += y[c];
      amax = std::max(amax, std::fabs(y[c]));
      ++nv;
    }
  if (nv == 0) return GTP_ERR_ARG;
  mu /= (double)nv;
  double var = 0;
  for (int64_t c = 0; c < C; ++c)
    if (!mask || mask[c]) var += (y[c] - mu) * (y[c] - mu);
  const double sy = std::sqrt(var / (double)nv);
  if (!(amax > 0)) amax = 1.0;
  const double qs = std::ldexp(1.0, 30) / amax;  // |yq| <= 2^30: sums over 2^33 cycles fit
  std::vector<long long> yq(C);
  for (int64_t c = 0; c < C; ++c) yq[c] = std::llrint(y[c] * qs);
  long long* d_yq = nullptr;
  unsigned long long *d_cnt = nullptr, *d_sum = nullptr;
  uint8_t* d_mask = nullptr;
  hipError_t e = hipMalloc((void**)&d_yq, (size_t)C * 8);
  if (e == hipSuccess) e = hipMalloc((void**)&d_cnt, (size_t)2 * N * 8);
  if (e == hipSuccess) e = hipMalloc((void**)&d_sum, (size_t)2 * N * 8);
  if (e == hipSuccess && mask) e = hipMalloc((void**)&d_mask, (size_t)C);
  if (e == hipSuccess) e = hipMemcpyAsync(d_yq, yq.data(), (size_t)C * 8, hipMemcpyHostToDevice, t.stream);
  if (e == hipSuccess && mask) e = hipMemcpyAsync(d_mask, mask, (size_t)C, hipMemcpyHostToDevice, t.stream);
  std::vector<unsigned long long> cnt(2 * N), sum(2 * N);
  if (e == hipSuccess) {
    fill_u32_kernel<<<nblk(4 * N), 256, 0, t.stream>>>((uint32_t*)d_cnt, 4 * N, 0u);
    fill_u32_kernel<<<nblk(4 * N), 256, 0, t.stream>>>((uint32_t*)d_sum, 4 * N, 0u);
    if (t.n_chunk > 0)
      net_corr_kernel<<<(unsigned)stride_blocks(t.n_chunk), 256, 0, t.stream>>>(
          t.net_off, t.ev, t.chunk_off, t.chunk_net, N, (uint32_t)t.T, (float)(1.0 / (double)t.T), C, d_mask, d_yq,
          d_cnt, d_sum, t.n_chunk);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), d_cnt, (size_t)2 * N * 8, hipMemcpyDeviceToHost, t.stream);
  if (e == hipSuccess) e = hipMemcpyAsync(sum.data(), d_sum, (size_t)2 * N * 8, hipMemcpyDeviceToHost, t.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(t.stream);
  for (void* p : {(void*)d_yq, (void*)d_cnt, (void*)d_sum, (void*)d_mask})
    if (p) (void)hipFree(p);
  if (e != hipSuccess) return e == hipErrorOutOfMemory ? GTP_ERR_OOM : GTP_ERR_DEVICE;
  for (int64_t n = 0; n < 2 * N; ++n) {
    const double p = (double)cnt[n] / (double)nv;
    const double mean_on = cnt[n] ? (double)(long long)sum[n] / qs / (double)cnt[n] : 0.0;
    const double cov = p * (mean_on - mu);
    const double sb = std::sqrt(p * (1 - p));
    corr[n] = (sb > 0 && sy > 0) ? (float)(cov / (sb * sy)) : 0.f;
  }
  return GTP_OK;
}

GTP_API gtp_status gtp_toggle_features(gtp_ctx* ctx, int32_t n_sel, const int32_t* nets, int64_t cycle_begin,
                                       int64_t n_cycle, uint32_t* bits) {
  TraceDev t;
  gtp_status st = gtp::trace_dev(ctx, &t);
  if (st != GTP_OK) return st;
  if (n_sel <= 0 || n_sel % 32 || !nets || !bits || cycle_begin < 0 || n_cycle <= 0 || !check_nets(t, n_sel, nets))
    return GTP_ERR_ARG;
  if (hipSetDevice(t.device) != hipSuccess) return GTP_ERR_DEVICE;
  const int words = n_sel / 32;
  FeatureSel fs;
  hipError_t e = fs.build(t, n_sel, nets, t.stream);
  uint32_t* d_X = nullptr;
  if (e == hipSuccess) e = hipMalloc((void**)&d_X, (size_t)n_cycle * words * 4);
  if (e == hipSuccess) {
    fs.launch(t, n_sel, cycle_begin, n_cycle, words, d_X, t.stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(bits, d_X, (size_t)n_cycle * words * 4, hipMemcpyDeviceToHost, t.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(t.stream);
  if (d_X) (void)hipFree(d_X);
  return e == hipSuccess ? GTP_OK : (e == hipErrorOutOfMemory ? GTP_ERR_OOM : GTP_ERR_DEVICE);
}

// Weighted toggle sums of the selected nets over cycles [cycle_begin,
// +n_cycle) of the loaded trace, on the context's auxiliary stream.  The
// weights are held as int32 multiples of a power-of-two quantum (the largest
// |weight| is 2^20 quanta), summed exactly in int64.
GTP_API gtp_status gtp_net_weight_sum(gtp_ctx* ctx, int32_t n_sel, const int32_t* nets, const double* weights,
                                      int64_t cycle_begin, int64_t n_cycle, double* out, double* ms) {
  TraceDev t;
  gtp_status st = gtp::trace_dev(ctx, &t);
  if (st != GTP_OK) return st;
  if (n_sel <= 0 || !nets || !weights || !out || cycle_begin < 0 || n_cycle <= 0 ||
      cycle_begin + n_cycle > t.n_cycle)
    return GTP_ERR_ARG;
  double wmax = 0;
  for (int32_t k = 0; k < n_sel; ++k) {
    if (nets[k] < 0 || nets[k] >= t.n_net || !std::isfinite(weights[k])) return GTP_ERR_ARG;
    // the window index holds event positions within a net in 32 bits
    if (t.host_net_off[nets[k] + 1] - t.host_net_off[nets[k]] >= INT32_MAX) return GTP_ERR_LIMIT;
    wmax = std::max(wmax, std::fabs(weights[k]));
  }
  if ((n_cycle + WS_CYCLES - 1) / WS_CYCLES > 65535) return GTP_ERR_LIMIT;  // windows on the grid's y axis
  if (!(wmax > 0)) {
    for (int64_t c = 0; c < n_cycle; ++c) out[c] = 0.0;
    if (ms) *ms = 0.0;
    return GTP_OK;
  }
  const int kq = (int)std::floor(20.0 - std::log2(wmax * 1.000001));
  std::vector<int32_t> wq(n_sel);
  for (int32_t k = 0; k < n_sel; ++k) wq[k] = (int32_t)std::llrint(std::ldexp(weights[k], kq));
  if (hipSetDevice(t.device) != hipSuccess) return GTP_ERR_DEVICE;
  hipStream_t s = gtp::aux_stream(ctx);
  if (!s) return GTP_ERR_DEVICE;
  const int32_t n_win = (int32_t)((n_cycle + WS_CYCLES - 1) / WS_CYCLES);
  gtp::AuxBuffers& ab = *gtp::aux_buffers(ctx);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = ab.sel.reserve(n_sel);
  if (e == hipSuccess) e = ab.wq.reserve(n_sel);
  if (e == hipSuccess) e = ab.idx.reserve((size_t)n_sel * (n_win + 1));
  if (e == hipSuccess) e = ab.out.reserve(n_cycle);
  int32_t *d_sel = ab.sel.p, *d_wq = ab.wq.p, *d_idx = ab.idx.p;
  unsigned long long* d_out = ab.out.p;
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  // the selection and its weights: uploaded unless the buffers hold them
  const bool same = e == hipSuccess && ab.sel_at == d_sel && ab.sel_h.size() == (size_t)n_sel &&
                    std::equal(nets, nets + n_sel, ab.sel_h.begin()) && std::equal(wq.begin(), wq.end(), ab.wq_h.begin());
  if (!same) {
    ab.sel_at = nullptr;
    if (e == hipSuccess) e = hipMemcpyAsync(d_sel, nets, (size_t)n_sel * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_wq, wq.data(), (size_t)n_sel * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
      ab.sel_h.assign(nets, nets + n_sel);
      ab.wq_h = wq;
      ab.sel_at = d_sel;
    }
  }
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  if (e == hipSuccess) {
    fill_u32_kernel<<<dim3((unsigned)((2 * n_cycle + 255) / 256)), 256, 0, s>>>((uint32_t*)d_out, 2 * n_cycle, 0u);
    uint32_t m = 0;
    int32_t sh = 0;
    cycle_magic(t.T, &m, &sh);
    const int64_t n_idx = (int64_t)n_sel * (n_win + 1);
    net_window_start_kernel<<<dim3((unsigned)((n_idx + 255) / 256)), 256, 0, s>>>(t.net_off, t.ev, d_sel, n_sel, n_win,
                                                                                  t.T, cycle_begin, n_cycle, d_idx);
    const dim3 grid((unsigned)((n_sel + WS_NETS - 1) / WS_NETS), (unsigned)n_win);
    net_weight_sum_kernel<<<grid, 256, 0, s>>>(t.net_off, t.ev, d_sel, d_wq, d_idx, n_sel, n_win, m, sh, cycle_begin,
                                              n_cycle, d_out);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  std::vector<long long> acc;
  if (e == hipSuccess) {
    acc.resize(n_cycle);
    e = hipMemcpyAsync(acc.data(), d_out, (size_t)n_cycle * 8, hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  float tms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&tms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (e != hipSuccess) return e == hipErrorOutOfMemory ? GTP_ERR_OOM : GTP_ERR_DEVICE;
  const double q = std::ldexp(1.0, -kq);
  for (int64_t c = 0; c < n_cycle; ++c) out[c] = (double)acc[c] * q;
  if (ms) *ms = tms;
  return GTP_OK;
}

// Surrogate inference straight from the resident event store: toggle bits of
// the selected nets (device), then the MLP.  `ms` receives the device time of
// features + inference.
GTP_API gtp_status gtp_mlp_predict_trace(gtp_mlp* m, gtp_ctx* ctx, const int32_t* nets, int64_t cycle_begin,
                                         int64_t n_cycle, float* out, double* ms) {
  if (!m) return GTP_ERR_ARG;
  TraceDev t;
  gtp_status st = gtp::trace_dev(ctx, &t);
  if (st != GTP_OK) return m->fail(st, "gtp_mlp_predict_trace: the context has no trace");
  if (!nets || !out || n_cycle <= 0 || cycle_begin < 0 || !check_nets(t, m->K, nets))
    return m->fail(GTP_ERR_ARG, "gtp_mlp_predict_trace: bad arguments");
  MCHK(m, hipSetDevice(m->device));
  const int words = m->words;
  auto grow = [](auto*& p, int64_t& have, int64_t need, size_t elem) {
    if (p && need <= have) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    have = 0;
    hipError_t e = hipMalloc((void**)&p, (size_t)need * elem);
    if (e == hipSuccess) have = need;
    return e;
  };
  int64_t net_n = m->tr_net ? m->K : 0;
  if (!m->tr_net) MCHK(m, grow(m->tr_net, net_n, m->K, 4));
  FeatureSel fs;
  if (m->tr_net_h.size() == (size_t)m->K && std::equal(nets, nets + m->K, m->tr_net_h.begin())) {
    fs.d_net = m->tr_net;  // the same selection as the last call: already on the device
  } else {
    m->tr_net_h.clear();
    MCHK(m, fs.use(m->tr_net, m->K, nets, m->stream));
    m->tr_net_h.assign(nets, nets + m->K);
  }
  hipError_t e = grow(m->tr_X, m->tr_X_n, n_cycle * words, 4);
  if (e == hipSuccess) e = grow(m->tr_out, m->tr_out_n, n_cycle, 4);
  uint32_t* d_X = m->tr_X;
  float* d_out = m->tr_out;
  const int64_t chunk = 1 << 18;
  if (e == hipSuccess) e = m->reserve((int)std::min<int64_t>(n_cycle, chunk));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, m->stream);
  if (e == hipSuccess) {
    fs.launch(t, m->K, cycle_begin, n_cycle, words, d_X, m->stream);
    e = hipGetLastError();
    for (int64_t b = 0; b < n_cycle && e == hipSuccess; b += chunk) {
      const int c = (int)std::min<int64_t>(chunk, n_cycle - b);
      e = forward(m, d_X + b * words, c, false, d_out + b);
    }
  }
  if (e == hipSuccess) e = hipEventRecord(e1, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t)n_cycle * 4, hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  float tms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&tms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (e != hipSuccess) return m->fail(GTP_ERR_DEVICE, std::string("gtp_mlp_predict_trace: ") + hipGetErrorString(e));
  for (int64_t i = 0; i < n_cycle; ++i) out[i] = out[i] * m->y_std + m->y_mean;
  if (ms) *ms = tms;
  return GTP_OK;
}

// Device GEMM entry for layout tests: C[M][N] (fp32) = A[M][K] * B[K][N] with
// A given as row-major bf16 bits (a_mode 0), bf16 stored [K][M] (1) or packed
// bits [M][K/32] (2), B as row-major [K][N] (0), stored [N][K] (1) or packed
// bits [K][N/32] (2).  All pointers are host arrays.
GTP_API gtp_status gtp_mlp_gemm_test(int32_t a_mode, int32_t b_mode, int32_t M, int32_t N, int32_t K,
                                     const void* A, int64_t a_bytes, const void* B, int64_t b_bytes, float* C) {
  if (M <= 0 || N % 128 || K % 32 || !A || !B || !C) return GTP_ERR_ARG;
  void *dA = nullptr, *dB = nullptr;
  float* dC = nullptr;
  hipStream_t s = nullptr;
  if (hipMalloc(&dA, a_bytes) != hipSuccess || hipMalloc(&dB, b_bytes) != hipSuccess ||
      hipMalloc((void**)&dC, (size_t)M * N * 4) != hipSuccess)
    return GTP_ERR_OOM;
  hipError_t e = hipMemcpy(dA, A, a_bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dB, B, b_bytes, hipMemcpyHostToDevice);
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K;
  g.A = dA; g.B = dB; g.C = dC; g.ldc = N;
  g.lda = a_mode == 0 ? K : (a_mode == 1 ? M : K / 32);
  g.ldb = b_mode == 0 ? N : (b_mode == 1 ? K : N / 32);
  if (e == hipSuccess) {
    if (a_mode == 0 && b_mode == 0) e = gemm<A_ROW, B_ROW, EPI_F32>(g, s);
    else if (a_mode == 0 && b_mode == 1) e = gemm<A_ROW, B_TRANS, EPI_F32>(g, s);
    else if (a_mode == 1 && b_mode == 0) e = gemm<A_COL, B_ROW, EPI_F32>(g, s);
    else if (a_mode == 1 && b_mode == 2) e = gemm<A_COL, B_BITS_ROW, EPI_F32>(g, s);
    else if (a_mode == 2 && b_mode == 1) e = gemm<A_BITS_ROW, B_TRANS, EPI_F32>(g, s);
    else e = hipErrorInvalidValue;
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  return e == hipSuccess ? GTP_OK : (e == hipErrorInvalidValue ? GTP_ERR_UNSUPPORTED : GTP_ERR_DEVICE);
}
