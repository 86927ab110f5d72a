// Learned proposal network for the reuse search (no counterpart in the
// reference): a 48 -> 256 -> 256 -> 262 MLP whose 256 logits score the
// offsets of a 16x16 block window and whose 6 further outputs seed the block's
// two endpoints (RGB / 255).  BF16 weights and activations, FP32 accumulation
// on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16); trained on the GPU by
// the kernels below.  Two networks of this shape are kept: one scores the
// inter window (previous frame), one the causal intra window (current frame).
//
// Logit index of a window offset (16x16 grid, the search window of S = 8):
//   inter  L = (dy + 8) * 16 + (dx + 8),  dy = j - by in [-8, 7], dx = i - bx in [-8, 7]
//   intra  L = dy * 16 + (dx + 8),        dy = by - j in [0, 15], dx = i - bx in [-8, 7]
//          (row dy = 0 holds only dx < 0: the window is causal)
//
// Training kernels: k_train_rows, k_train_wgrad, k_grad_reduce (with the
// Adam step), k_train_copies, k_collect_masks; k_net_input (the inputs of a
// forward pass).  Search: k_frame_net_input (the network input of a batch)
// and k_lintra_pass (learned intra fixpoint); the inference kernels and the
// top-k proposal kernels are in net_infer.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "batch_kernels.h"

namespace mptc {

constexpr int kNetIn = 48, kNetHid = 256, kNetLogits = 256, kNetSeeds = 6;
constexpr int kNetOut = kNetLogits + kNetSeeds;  // 262
constexpr int kNetWin = 16;                      // logit grid side
constexpr int kMaxTopK = 16;
// flat fp32 parameter vector: W1 [48][256], b1, W2 [256][256], b2, W3 [256][262], b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kNetIn * kNetHid;
constexpr int kOffW2 = kOffB1 + kNetHid;
constexpr int kOffB2 = kOffW2 + kNetHid * kNetHid;
constexpr int kOffW3 = kOffB2 + kNetHid;
constexpr int kOffB3 = kOffW3 + kNetHid * kNetOut;
constexpr int kNetParams = kOffB3 + kNetOut;  // 145,670

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

// fp32 -> bf16, round to nearest even (a NaN stays a quiet NaN)
__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) return (uint16_t)((u >> 16) | 0x40);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Two finite fp32 values -> packed bf16 (lo in bits 0-15), round to nearest
// even as f2bf, in one v_cvt_pk_bf16_f32.  For the per-frame kernels'
// activations and pixel inputs (never NaN or infinite).
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
  const bf16x2_t b = __builtin_convertvector((f32x2_t){lo, hi}, bf16x2_t);
  uint32_t u;
  __builtin_memcpy(&u, &b, 4);
  return u;
}

// Network input: x = pixel / 255 (pixel-major RGB), bf16.  Row r of X takes
// pixel row idx[r] (or r when idx is null).
__global__ void __launch_bounds__(256) k_net_input(const uint8_t* __restrict__ px, const int32_t* __restrict__ idx,
                                                   int n, uint16_t* __restrict__ X) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)n * kNetIn) return;
  const long row = i / kNetIn;
  const int c = (int)(i % kNetIn);
  const long src = idx ? (long)idx[row] : row;
  X[i] = f2bf((float)px[src * kNetIn + c] * (1.0f / 255.0f));
}

// ---------------------------------------------------------------------------
// Training step: k_train_rows, k_train_wgrad, k_grad_reduce.
//
// MFMA v_mfma_f32_16x16x32_bf16 operand layout (lane l of a wave): A row
// l % 16, k (l / 16) * 8 .. +7; B column l % 16, k (l / 16) * 8 .. +7; D rows
// (l / 16) * 4 .. +3, column l % 16.  Every operand fragment is one 16-byte
// load of 8 consecutive k, so each operand is kept with k contiguous:
//  - bf16 weight copies (written by the Adam step / k_train_copies; pads stay zero):
//      W1T [256][64]  (W1 transposed, k 48..63 zero)      forward, layer 1
//      W2T [256][256] (W2 transposed)                     forward, layer 2
//      W3T [272][256] (W3 transposed, rows 262..271 zero)  forward, layer 3
//      W2R [256][256] (W2 as stored)                      dH1 = dH2 W2^T
//      W3R [256][288] (W3, columns 262..287 zero)          dH2 = dY W3^T
//  - the batch's activations and output gradients, feature-major (row f
//    holds feature f of every batch row, row stride ld = the batch rounded up
//    to 32), bf16: X (64 rows, 48..63 zero), H1, H2, dY (320 rows, 262..319
//    zero), dH2, dH1 -- the operands of the weight gradients, whose k is the
//    batch row.
// ---------------------------------------------------------------------------
constexpr int kTrW1T = 0;
constexpr int kTrW2T = kTrW1T + kNetHid * 64;
constexpr int kTrW3T = kTrW2T + kNetHid * kNetHid;
constexpr int kTrW2R = kTrW3T + 272 * kNetHid;
constexpr int kTrW3R = kTrW2R + kNetHid * kNetHid;
constexpr int kTrSize = kTrW3R + kNetHid * 288;  // bf16 elements
constexpr int kActX = 0, kActH1 = 64, kActH2 = kActH1 + kNetHid, kActDY = kActH2 + kNetHid;
constexpr int kActDH2 = kActDY + 320, kActDH1 = kActDH2 + kNetHid, kActRows = kActDH1 + kNetHid;  // 1408
constexpr int kBiasCols = 2 * kNetHid + kNetOut;           // 774: db1, db2, db3 in parameter order
constexpr long kWgPart = 64L * 256 + 256L * 256 + 256L * 320;  // floats of one split's weight gradients
constexpr int kWgTiles = 4 + 16 + 20;                      // 64 x 64 tiles of dW1, dW2, dW3

// the bf16 training copies of parameter i (a weight; biases have none)
__device__ __forceinline__ void store_train_copies(int i, float wi, uint16_t* __restrict__ wt) {
  const uint16_t h = f2bf(wi);
  if (i < kOffB1) {
    const int k = i / kNetHid, n = i % kNetHid;
    wt[kTrW1T + n * 64 + k] = h;
  } else if (i >= kOffW2 && i < kOffB2) {
    const int j = i - kOffW2, k = j / kNetHid, n = j % kNetHid;
    wt[kTrW2T + n * kNetHid + k] = h;
    wt[kTrW2R + j] = h;
  } else if (i >= kOffW3 && i < kOffB3) {
    const int j = i - kOffW3, k = j / kNetOut, n = j % kNetOut;
    wt[kTrW3T + n * kNetHid + k] = h;
    wt[kTrW3R + k * 288 + n] = h;
  }
}

__device__ __forceinline__ bf16x8_t frag(const uint16_t* p) { return *reinterpret_cast<const bf16x8_t*>(p); }

// The weight fragments of KS k-steps (from k0) of NJ column tiles of a
// weight copy gB (tile j: output columns col0 + cstep * j .. +15, one row of
// gB per output column, k contiguous, row stride ldb).  nj (wave-uniform)
// tiles are loaded.  Loading a chunk before the work that precedes its
// MFMAs keeps the L2 round trip off the critical path.
template <int NJ, int KS>
struct WFrags {
  bf16x8_t b[KS][NJ];
};

template <int NJ, int KS>
__device__ __forceinline__ void load_frags(WFrags<NJ, KS>& f, const uint16_t* __restrict__ gB, int ldb, int col0,
                                           int cstep, int nj, int k0, int lane) {
  const int r = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
      if (j < nj) f.b[ks][j] = frag(gB + (long)(col0 + cstep * j + r) * ldb + k0 + ks * 32 + kq);
}

// acc[i][j] += A x B: A = rows 16 i .. 16 i + 15 of the LDS matrix sA (row
// stride lda, k contiguous), k from k0 over the chunk's KS steps.
template <int RT, int NJ, int KS>
__device__ __forceinline__ void mfma_frags(f32x4_t (&acc)[RT][NJ], const WFrags<NJ, KS>& f, const uint16_t* sA,
                                           int lda, int k0, int nj, int lane) {
  const int r = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    bf16x8_t a[RT];
#pragma unroll
    for (int i = 0; i < RT; ++i) a[i] = frag(sA + (16 * i + r) * lda + k0 + ks * 32 + kq);
#pragma unroll
    for (int j = 0; j < NJ; ++j)
      if (j < nj)
#pragma unroll
        for (int i = 0; i < RT; ++i)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], f.b[ks][j], acc[i][j], 0, 0, 0);
  }
}

template <int RT, int NJ>
__device__ __forceinline__ void zero_acc(f32x4_t (&acc)[RT][NJ]) {
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
}

// Epilogue of a 256-wide layer: wave columns c0 + 16 j + (lane % 16), rows
// 16 i + (lane / 16) * 4 + t.  H = relu(acc + bias) (forward, mask null) or
// acc * [mask > 0] (backward); bf16 to the LDS rows sO (when non-null) and as
// 4 consecutive rows of one feature to gO (feature-major, row stride ld);
// with cpo, the fp32 column sums over the workgroup's rows go to cpo[col].
template <int RT>
__device__ __forceinline__ void layer_out(const f32x4_t (&acc)[RT][4], int c0, int lane, const float* __restrict__ bias,
                                          const uint16_t* mask, uint16_t* sO, uint16_t* __restrict__ gO, int ld,
                                          float* __restrict__ cpo) {
  constexpr int LH = kNetHid + 8;
  const int col_in = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int col = c0 + 16 * j + col_in;
    const float b = bias ? bias[col] : 0.0f;
    float cs = 0.0f;
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      float v[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int r = 16 * i + rq + t;
        v[t] = mask ? (bf2f(mask[r * LH + col]) > 0.0f ? acc[i][j][t] : 0.0f) : fmaxf(acc[i][j][t] + b, 0.0f);
        cs += v[t];
      }
      const uint32_t p01 = pack_bf16x2(v[0], v[1]), p23 = pack_bf16x2(v[2], v[3]);
      if (sO) {
        const int r = 16 * i + rq;
        sO[r * LH + col] = (uint16_t)p01;
        sO[(r + 1) * LH + col] = (uint16_t)(p01 >> 16);
        sO[(r + 2) * LH + col] = (uint16_t)p23;
        sO[(r + 3) * LH + col] = (uint16_t)(p23 >> 16);
      }
      *reinterpret_cast<uint2*>(gO + (long)col * ld + 16 * i + rq) = make_uint2(p01, p23);
    }
    if (cpo) {
      cs += __shfl_xor(cs, 16);
      cs += __shfl_xor(cs, 32);
      if (lane < 16) cpo[col] = cs;
    }
  }
}

// One workgroup (4 waves) per 16 RT batch rows: the forward pass (H1, H2,
// logits and seeds), the loss and dY = d(mean loss)/dY, and the backward
// products dH2 = (dY W3^T) * [H2 > 0], dH1 = (dH2 W2^T) * [H1 > 0], all
// held in LDS; the weight-gradient operands go out feature-major (act) and
// the bias gradients as this workgroup's column sums of dH1, dH2, dY in fp32
// (colpart[wg][774]).  Wave w computes columns 64 w .. 64 w + 63 of each
// 256-wide layer (RT x 4 MFMA tiles) and column tiles w, w + 4, ... of the 17
// of the output layer.  Batch row r (r < n) takes pair idx[r] (or r); rows
// n .. ld - 1 are zero in the inputs and gradients.  The weight fragments of
// each layer are requested in chunks of 4 or 5 k-steps one chunk ahead of
// their MFMAs, the backward's during the loss.
//
// Loss: -log of the probability the softmax of the 256 logits gives the row's
// target set (tmask: 8 words, bit L of word L / 32; one bit = the usual
// cross-entropy), plus seed_w * the mean squared error of the 6 seeds against
// the endpoints / 255.  d/dy_c of the set term is p_c - [c in set] exp(y_c) /
// sum_{set} exp(y).  A row with an empty set adds only its seed term.  A wave
// takes its 4 RT rows at once (64 / (4 RT) lanes per row).  The batch's mean
// loss is added to *loss (one atomic per workgroup).
template <int RT>
__global__ void __launch_bounds__(256) k_train_rows(int n, int ld, const uint8_t* __restrict__ px,
                                                    const int32_t* __restrict__ idx, const uint32_t* __restrict__ tmask,
                                                    const uint8_t* __restrict__ eps, float seed_w,
                                                    const uint16_t* __restrict__ wt, const float* __restrict__ w,
                                                    const float* __restrict__ base, uint16_t* __restrict__ act,
                                                    float* __restrict__ colpart, float* __restrict__ loss) {
  constexpr int R = 16 * RT;                                                // rows of the workgroup
  constexpr int LX = 64 + 8, LH = kNetHid + 8, LY = 288 + 4, LD = 288 + 8;  // LDS row strides (skewed banks)
  __shared__ __attribute__((aligned(16))) uint16_t sX[R * LX];
  __shared__ __attribute__((aligned(16))) uint16_t sH1[R * LH];
  __shared__ __attribute__((aligned(16))) uint16_t sH2[R * LH];
  __shared__ __attribute__((aligned(16))) float sY[R * LY];       // logits, then dY (fp32)
  __shared__ __attribute__((aligned(16))) uint16_t sDY[R * LD];   // dY (bf16), columns 262..287 zero
  __shared__ __attribute__((aligned(16))) uint16_t sDH2[R * LH];  // dH2 (bf16)
  __shared__ float s_l[4];
  __shared__ long s_src[R];
  __shared__ uint32_t s_tm[R * 8];
  __shared__ uint8_t s_eps[R * kNetSeeds];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int r0 = blockIdx.x * R, c0 = wave * 64;
  const int col_in = lane & 15, rq = (lane >> 4) * 4;  // D layout: rows rq .. rq + 3 of a tile, column col_in
  const int nj3 = wave == 0 ? 5 : 4;                   // output-layer column tiles of this wave
  float* const cp = colpart + (long)blockIdx.x * kBiasCols;
  WFrags<4, 2> fA;
  WFrags<4, 4> fB0, fB1;
  load_frags(fA, wt + kTrW1T, 64, c0, 16, 4, 0, lane);
  load_frags(fB0, wt + kTrW2T, kNetHid, c0, 16, 4, 0, lane);
  if (tid < R) s_src[tid] = r0 + tid < n ? (idx ? (long)idx[r0 + tid] : (long)(r0 + tid)) : -1;
  __syncthreads();
  // the rows' pixels, target sets and seeds, all requested at once; inputs
  // x = pixel / 255 (pixel-major RGB), columns 48..63 zero
  if (tid < R * 8) {
    const long src = s_src[tid >> 3];
    s_tm[tid] = src >= 0 ? tmask[src * 8 + (tid & 7)] : 0u;
  }
  if (tid < R * kNetSeeds) {
    const long src = s_src[tid / kNetSeeds];
    s_eps[tid] = src >= 0 ? eps[src * kNetSeeds + tid % kNetSeeds] : (uint8_t)0;
  }
  for (int e = tid; e < R * 64; e += 256) {
    const int r = e >> 6, cc = e & 63;
    const long src = s_src[r];
    const float x = (src >= 0 && cc < kNetIn) ? (float)px[src * kNetIn + cc] * (1.0f / 255.0f) : 0.0f;
    sX[r * LX + cc] = f2bf(x);
  }
  __syncthreads();
  for (int e = tid; e < 64 * (R / 8); e += 256) {  // X feature-major: (feature, 8 rows)
    const int f = e / (R / 8), rr = (e % (R / 8)) * 8;
    uint16_t h[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) h[j] = sX[(rr + j) * LX + f];
    uint4 v;
    __builtin_memcpy(&v, h, 16);
    *reinterpret_cast<uint4*>(act + (long)(kActX + f) * ld + r0 + rr) = v;
  }
  {  // layer 1: H1 = relu(X W1 + b1)
    f32x4_t acc[RT][4];
    zero_acc(acc);
    mfma_frags(acc, fA, sX, LX, 0, 4, lane);
    load_frags(fB1, wt + kTrW2T, kNetHid, c0, 16, 4, 128, lane);
    layer_out<RT>(acc, c0, lane, w + kOffB1, nullptr, sH1, act + (long)kActH1 * ld + r0, ld, nullptr);
  }
  __syncthreads();
  WFrags<5, 4> fC0, fC1;
  {  // layer 2: H2 = relu(H1 W2 + b2)
    f32x4_t acc[RT][4];
    zero_acc(acc);
    mfma_frags(acc, fB0, sH1, LH, 0, 4, lane);
    load_frags(fC0, wt + kTrW3T, kNetHid, 16 * wave, 64, nj3, 0, lane);
    mfma_frags(acc, fB1, sH1, LH, 128, 4, lane);
    load_frags(fC1, wt + kTrW3T, kNetHid, 16 * wave, 64, nj3, 128, lane);
    layer_out<RT>(acc, c0, lane, w + kOffB2, nullptr, sH2, act + (long)kActH2 * ld + r0, ld, nullptr);
  }
  __syncthreads();
  WFrags<4, 4> fD0;
  WFrags<4, 5> fD1;
  {  // layer 3: Y = H2 W3 + b3 (+ base), column tiles wave, wave + 4, ... of 17
    f32x4_t acc[RT][5];
    zero_acc(acc);
    mfma_frags(acc, fC0, sH2, LH, 0, nj3, lane);
    load_frags(fD0, wt + kTrW3R, 288, c0, 16, 4, 0, lane);
    mfma_frags(acc, fC1, sH2, LH, 128, nj3, lane);
    load_frags(fD1, wt + kTrW3R, 288, c0, 16, 4, 128, lane);
#pragma unroll
    for (int j = 0; j < 5; ++j)
      if (j < nj3) {
        const int col = 16 * (wave + 4 * j) + col_in;
        const float b = col < kNetOut ? w[kOffB3 + col] + (base ? base[col] : 0.0f) : 0.0f;
#pragma unroll
        for (int i = 0; i < RT; ++i)
#pragma unroll
          for (int t = 0; t < 4; ++t) sY[(16 * i + rq + t) * LY + col] = col < kNetOut ? acc[i][j][t] + b : 0.0f;
      }
  }
  __syncthreads();
  const float inv_n = 1.0f / (float)n;
  float l = 0.0f;  // this lane's share of the wave's rows' losses
  {  // loss and dY: the wave's 4 RT rows at once, LPR lanes per row
    constexpr int LPR = 16 / RT, CPL = kNetLogits / LPR;
    const int r = wave * 4 * RT + lane / LPR, sub = lane % LPR;
    const bool valid = r0 + r < n;
    float* y = sY + r * LY;
    uint16_t* dyb = sDY + r * LD;
    uint32_t tm[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) tm[q] = s_tm[r * 8 + q];
    float e[CPL];
    float mx = -3.0e38f;
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      e[q] = y[sub + LPR * q];
      mx = fmaxf(mx, e[q]);
    }
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    float se = 0.0f, ss = 0.0f;  // sum over all logits, over the set
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int cc = sub + LPR * q;
      e[q] = expf(e[q] - mx);
      se += e[q];
      if ((tm[cc >> 5] >> (cc & 31)) & 1u) ss += e[q];
    }
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) {
      se += __shfl_xor(se, o);
      ss += __shfl_xor(ss, o);
    }
    const bool has_set = ss > 0.0f;
    if (valid && has_set && sub == 0) l += logf(se) - logf(ss);
    const float gs = valid && has_set ? inv_n / se : 0.0f, gq = valid && has_set ? inv_n / ss : 0.0f;
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int cc = sub + LPR * q;
      const float g = e[q] * gs - (((tm[cc >> 5] >> (cc & 31)) & 1u) ? e[q] * gq : 0.0f);
      y[cc] = g;
      dyb[cc] = f2bf(g);
    }
    for (int cc = kNetLogits + sub; cc < 288; cc += LPR) {  // the seeds, then zero columns
      float g = 0.0f;
      if (valid && cc < kNetOut) {
        const float t = (float)s_eps[r * kNetSeeds + (cc - kNetLogits)] * (1.0f / 255.0f);
        const float d = y[cc] - t;
        l += seed_w * d * d * (1.0f / kNetSeeds);
        g = 2.0f * seed_w * d * (1.0f / kNetSeeds) * inv_n;
      }
      y[cc] = g;
      dyb[cc] = f2bf(g);
    }
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if (lane == 0) s_l[wave] = l;
  __syncthreads();
  if (tid == 0) atomicAdd(loss, (s_l[0] + s_l[1] + s_l[2] + s_l[3]) * inv_n);
  // db3 column sums (fp32, rows in order); dY feature-major (rows 262..319 zero)
  for (int cc = tid; cc < kNetOut; cc += 256) {
    float s = 0.0f;
    for (int r = 0; r < R; ++r) s += sY[r * LY + cc];
    cp[2 * kNetHid + cc] = s;
  }
  for (int e = tid; e < 320 * (R / 8); e += 256) {
    const int f = e / (R / 8), rr = (e % (R / 8)) * 8;
    uint16_t h[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) h[j] = f < 288 ? sDY[(rr + j) * LD + f] : (uint16_t)0;
    uint4 v;
    __builtin_memcpy(&v, h, 16);
    *reinterpret_cast<uint4*>(act + (long)(kActDY + f) * ld + r0 + rr) = v;
  }
  WFrags<4, 4> fE0, fE1;
  {  // dH2 = (dY W3^T) * [H2 > 0]
    f32x4_t acc[RT][4];
    zero_acc(acc);
    mfma_frags(acc, fD0, sDY, LD, 0, 4, lane);
    load_frags(fE0, wt + kTrW2R, kNetHid, c0, 16, 4, 0, lane);
    mfma_frags(acc, fD1, sDY, LD, 128, 4, lane);
    load_frags(fE1, wt + kTrW2R, kNetHid, c0, 16, 4, 128, lane);
    layer_out<RT>(acc, c0, lane, nullptr, sH2, sDH2, act + (long)kActDH2 * ld + r0, ld, cp + kNetHid);
  }
  __syncthreads();
  {  // dH1 = (dH2 W2^T) * [H1 > 0]
    f32x4_t acc[RT][4];
    zero_acc(acc);
    mfma_frags(acc, fE0, sDH2, LH, 0, 4, lane);
    mfma_frags(acc, fE1, sDH2, LH, 128, 4, lane);
    layer_out<RT>(acc, c0, lane, nullptr, sH1, nullptr, act + (long)kActDH1 * ld + r0, ld, cp);
  }
}

// Weight gradients dW1 = X^T dH1, dW2 = H1^T dH2, dW3 = H2^T dY as 40 tiles
// of 64 x 64 (blockIdx.x) with the batch split into gridDim.y ranges of kspan
// rows (a multiple of 32); split s stores its partial products to part + s *
// kWgPart ([64][256], [256][256], [256][320]; k_grad_reduce adds them in a
// fixed order).  Wave w owns the 32 x 32 quadrant (w / 2, w % 2) as 2 x 2
// MFMA tiles; both operands are feature-major rows of act, loaded straight
// into registers.
__global__ void __launch_bounds__(256) k_train_wgrad(const uint16_t* __restrict__ act, int ld, int kspan,
                                                     float* __restrict__ part) {
  const int t = blockIdx.x, s = blockIdx.y;
  int arow, brow, m0, n0, ncols;
  long poff;
  if (t < 4) {
    arow = kActX, brow = kActDH1, m0 = 0, n0 = t * 64, ncols = kNetHid, poff = 0;
  } else if (t < 20) {
    const int u = t - 4;
    arow = kActH1, brow = kActDH2, m0 = (u >> 2) * 64, n0 = (u & 3) * 64, ncols = kNetHid, poff = 64L * 256;
  } else {
    const int u = t - 20;
    arow = kActH2, brow = kActDY, m0 = (u / 5) * 64, n0 = (u % 5) * 64, ncols = 320, poff = 64L * 256 + 256L * 256;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32, r = lane & 15, kq = (lane >> 4) * 8;
  const uint16_t* A0 = act + (long)(arow + m0 + wm + r) * ld + kq;
  const uint16_t* B0 = act + (long)(brow + n0 + wn + r) * ld + kq;
  const long l16 = 16L * ld;
  f32x4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const int kb = s * kspan, ke = min(ld, kb + kspan);
  int k = kb;
  for (; k + 256 <= ke; k += 256) {  // 8 steps' fragments requested before their MFMAs
    bf16x8_t a[8][2], b[8][2];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      a[q][0] = frag(A0 + k + 32 * q);
      a[q][1] = frag(A0 + l16 + k + 32 * q);
      b[q][0] = frag(B0 + k + 32 * q);
      b[q][1] = frag(B0 + l16 + k + 32 * q);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[q][i], b[q][j], acc[i][j], 0, 0, 0);
  }
  for (; k < ke; k += 32) {
    const bf16x8_t a0 = frag(A0 + k), a1 = frag(A0 + l16 + k), b0 = frag(B0 + k), b1 = frag(B0 + l16 + k);
    acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
  }
  float* P = part + (long)s * kWgPart + poff;
  const int col_in = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        P[(long)(m0 + wm + 16 * i + rq + q) * ncols + n0 + wn + 16 * j + col_in] = acc[i][j][q];
}

// One Adam step (on when lr > 0); c1, c2 = 1 - beta1^t, 1 - beta2^t.
struct AdamStep {
  float lr, b1, b2, c1, c2, eps;
};

// Adam on fp32 master parameter i with gradient gi; refreshes its bf16 copies.
__device__ __forceinline__ void adam_update(int i, float gi, float* __restrict__ w, float* __restrict__ m,
                                            float* __restrict__ v, uint16_t* __restrict__ wt, const AdamStep& a) {
  const float mi = a.b1 * m[i] + (1.0f - a.b1) * gi;
  const float vi = a.b2 * v[i] + (1.0f - a.b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - a.lr * (mi / a.c1) / (sqrtf(vi / a.c2) + a.eps);
  w[i] = wi;
  store_train_copies(i, wi, wt);
}

// g = d(mean loss)/d(parameters), then (a.lr > 0) the Adam step: a weight's
// gradient is the sum of the splits' partials in split order; a bias's the
// sum of the k_train_rows workgroups' column sums (a wave per bias: lane l
// adds workgroups l, l + 64, ... in order, then the lanes in a fixed tree).
// Blocks 0 .. nwb - 1 take the parameters by index, the others the 774
// biases four at a time.
__global__ void __launch_bounds__(256) k_grad_reduce(const float* __restrict__ part, int splits,
                                                     const float* __restrict__ colpart, int nwg, int nwb,
                                                     float* __restrict__ g, float* __restrict__ w,
                                                     float* __restrict__ m, float* __restrict__ v,
                                                     uint16_t* __restrict__ wt, AdamStep a) {
  if ((int)blockIdx.x < nwb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    long p;
    if (i < kOffB1)
      p = i;
    else if (i >= kOffW2 && i < kOffB2)
      p = 64L * 256 + (i - kOffW2);
    else if (i >= kOffW3 && i < kOffB3)
      p = 64L * 256 + 256L * 256 + (long)((i - kOffW3) / kNetOut) * 320 + (i - kOffW3) % kNetOut;
    else
      return;  // a bias, or past the end
    float s = 0.0f;
#pragma unroll 4
    for (int z = 0; z < splits; ++z) s += part[(long)z * kWgPart + p];
    g[i] = s;
    if (a.lr > 0.0f) adam_update(i, s, w, m, v, wt, a);
    return;
  }
  const int bc = ((int)blockIdx.x - nwb) * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (bc >= kBiasCols) return;  // wave-uniform
  float s = 0.0f;
  for (int q = lane; q < nwg; q += 64) s += colpart[(long)q * kBiasCols + bc];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) {
    const int i = bc < kNetHid ? kOffB1 + bc : bc < 2 * kNetHid ? kOffB2 + bc - kNetHid : kOffB3 + bc - 2 * kNetHid;
    g[i] = s;
    if (a.lr > 0.0f) adam_update(i, s, w, m, v, wt, a);
  }
}

// the bf16 training copies of all n parameters (after they are set)
__global__ void __launch_bounds__(256) k_train_copies(const float* __restrict__ w, int n, uint16_t* __restrict__ wt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) store_train_copies(i, w[i], wt);
}

// Training pairs of the last encoded step: every block whose search window
// (net 0: inter, blocks of inter frames; net 1: intra, blocks that searched
// it) holds at least one offset that passes the accept test gives (its 48
// pixels, the 256-bit set of passing offsets as 8 words, bit L of word L / 32,
// its endpoints after the encode -- the frame's eps output, or its own fit's
// when the frame has none).  amask: per block 16 rows of 16 offset bits, as
// k_inter / k_intra_amask record them.  Pairs are appended in no particular
// order; pairs beyond capacity are dropped (count still advances).
__global__ void __launch_bounds__(256) k_collect_masks(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid,
                                                       int total, int net, int S, const uint8_t* __restrict__ need,
                                                       const uint16_t* __restrict__ amask,
                                                       const uint64_t* __restrict__ own, uint8_t* __restrict__ px_out,
                                                       uint32_t* __restrict__ tgt_out, uint8_t* __restrict__ eps_out,
                                                       unsigned int* __restrict__ count, int capacity) {
  __shared__ unsigned int s_base;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool want = false;
  if (g < total) {
    const FrameDev& d = fr[sid[g]];
    if (net == 0 ? !d.is_intra : need[g] != 0) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        // inter rows outside the window (search areas below 8) are not recorded
        const bool row_ok = net == 1 || (r - 8 >= -S && r - 8 < S);
        const uint32_t bits = row_ok ? (uint32_t)amask[(size_t)g * 16 + r] : 0u;
        m[r >> 1] |= bits << (16 * (r & 1));
        want = want || bits != 0;
      }
    }
  }
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  const unsigned long long bal = __ballot(want);
  const int lane = (int)__lane_id();
  unsigned int off = 0;
  if (bal && lane == 0) off = atomicAdd(&s_base, (unsigned int)__popcll(bal));
  off = __shfl(off, 0);
  __syncthreads();
  if (threadIdx.x == 0 && s_base) s_base = atomicAdd(count, s_base);
  __syncthreads();
  if (!want) return;
  const unsigned slot = s_base + off + (unsigned)__popcll(bal & ((1ull << lane) - 1));
  if (slot >= (unsigned)capacity) return;
  const FrameDev& d = fr[sid[g]];
  const int k = g - d.base;
  BlockPx b;
  load_block(d.rgb, d.width, k % d.bw, k / d.bw, b);
  uint32_t* o = reinterpret_cast<uint32_t*>(px_out + (size_t)slot * kNetIn);
  for (int w = 0; w < 12; ++w) o[w] = b.w[w];
  for (int w = 0; w < 8; ++w) tgt_out[(size_t)slot * 8 + w] = m[w];
  uint8_t* e = eps_out + (size_t)slot * kNetSeeds;
  if (d.eps) {
    for (int c = 0; c < kNetSeeds; ++c) e[c] = d.eps[(size_t)kNetSeeds * k + c];
  } else {
    uint32_t r1, g1, b1, r2, g2, b2;
    decode565((uint32_t)(own[g] & 0xFFFF), r1, g1, b1);
    decode565((uint32_t)((own[g] >> 16) & 0xFFFF), r2, g2, b2);
    e[0] = (uint8_t)r1; e[1] = (uint8_t)g1; e[2] = (uint8_t)b1;
    e[3] = (uint8_t)r2; e[4] = (uint8_t)g2; e[5] = (uint8_t)b2;
  }
}

// Network input row of one block: 48 bf16 (pixel bytes / 255, in memory order).
__device__ __forceinline__ void net_input_row(const BlockPx& b, uint16_t* __restrict__ row) {
  uint4* o = reinterpret_cast<uint4*>(row);
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    uint32_t h[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint32_t w = b.w[2 * q + (e >> 1)];
      const uint32_t lo = (w >> (16 * (e & 1))) & 0xFFu, hi = (w >> (16 * (e & 1) + 8)) & 0xFFu;
      h[e] = pack_bf16x2((float)lo * (1.0f / 255.0f), (float)hi * (1.0f / 255.0f));
    }
    o[q] = make_uint4(h[0], h[1], h[2], h[3]);
  }
}

// Network input of every block of a batch: total x 48 bf16.
__global__ void __launch_bounds__(256) k_frame_net_input(const FrameDev* __restrict__ fr,
                                                         const uint8_t* __restrict__ sid, int total,
                                                         uint16_t* __restrict__ X) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const FrameDev& d = fr[sid[g]];
  const int k = g - d.base;
  BlockPx b;
  load_block(d.rgb, d.width, k % d.bw, k / d.bw, b);
  net_input_row(b, X + (size_t)g * kNetIn);
}

// The step's first kernel: clears the step's counters (z1[0..n1), z2[0..n2))
// and does the per-block preparation: the candidate table of the reference
// frame's index word (blocks of inter frames), and from one load of the
// block's pixels the stb fit (own, when not null; as k_fit) and the network
// input row (X, when not null).  The grid covers max(total, n1 + n2) items.
__global__ void __launch_bounds__(256) k_block_prep(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid,
                                                    int total, uint64_t* __restrict__ own, uint32_t* __restrict__ rw,
                                                    float4* __restrict__ rtab, uint16_t* __restrict__ X,
                                                    uint32_t* __restrict__ z1, int n1, uint32_t* __restrict__ z2,
                                                    int n2, BoundTab* __restrict__ rbt) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n1) z1[g] = 0u;
  else if (g - n1 < n2) z2[g - n1] = 0u;
  if (g >= total) return;
  const FrameDev& d = fr[sid[g]];
  const int k = g - d.base;
  if (!d.is_intra) {
    const uint32_t c = (uint32_t)(d.ref[k] >> 32);
    const CandTab t = cand_table(c);
    rw[g] = c;
    rtab[g] = make_float4(t.asq, t.bsq, t.ab, t.f);
    if (rbt) rbt[g] = bound_table(c);  // the pruned inter search's bound table
  }
  if (!own && !X) return;
  BlockPx px;
  load_block(d.rgb, d.width, k % d.bw, k / d.bw, px);
  if (own) own[g] = fit_block(px);
  if (X) net_input_row(px, X + (size_t)g * kNetIn);
}

// Learned intra search, one fixpoint pass.  A wave owns a tile of
// 2^ltw_log x 64 / 2^ltw_log blocks (lane = row * width + column in the tile;
// wider tiles settle longer horizontal chains within a pass, taller ones
// vertical chains).  Every block that needs the
// intra search evaluates its proposed offsets against the current words --
// words of its own tile come from the lanes that own them, others from
// memory -- and once all its evaluations are current it takes the valid
// proposal with the smallest error delta (first in rank order on ties),
// accepts it if the delta passes the threshold, and its resulting word
// becomes current.  Each round a lane redoes one evaluation whose word
// changed (seen/errs keep the others).  Proposals only point at earlier blocks
// in raster order, which inside a tile are lower lanes, so the rounds settle
// the tile's chains within the pass.  Words from other tiles are read once per
// pass; a change there is caught by the next pass.  Each block's window holds
// only earlier blocks, so a pass that changes no word ends at the unique
// fixpoint.  A tile not settled after kMaxRounds rounds stores, for every
// lane, the decision its current evaluations give; what is stale is redone in
// the next pass (never-done evaluations are stored as kNotEvaluated), and the
// tile counts as one change, so that a next pass follows even when no word
// changed.  Pass counters: this pass adds its changed words (and unsettled
// tiles) to n_changed[par] (zero at its start) and reads the previous pass's,
// n_changed[par ^ 1], which its last workgroup clears for the next pass
// (pass_end); the count also goes to *conv.
constexpr int kMaxRounds = 512;
constexpr int kNotEvaluated = (int)0x80000000;

template <int K>
__global__ void __launch_bounds__(256) k_lintra_pass(const FrameDev* __restrict__ fr, int n_frames, int n_itiles,
                                                     int ltw_log, int S, int thr, int kk, const int16_t* __restrict__ prop,
                                                     const uint64_t* __restrict__ own, const uint8_t* __restrict__ need,
                                                     uint32_t* __restrict__ cw, uint32_t* __restrict__ seen,
                                                     int* __restrict__ errs, uint64_t* __restrict__ dec,
                                                     DevStats* __restrict__ st, int par, int first,
                                                     unsigned int* __restrict__ conv, uint8_t* __restrict__ tflag,
                                                     long tf_stride) {
  __shared__ uint8_t s_snap[512];
  __shared__ float2 s_ab[4];
  __shared__ unsigned long long s_cnt[2][4];  // per wave: first evaluations, re-evaluations
  __shared__ unsigned int s_chg, s_prev;
  const int wv = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int lane = (int)(threadIdx.x & 63);
  // The previous pass's count, the evaluation tables and this wave's tile
  // flag are requested together, before the one barrier.  After a pass that
  // changed no word every evaluation is current: nothing to do.  The
  // previous pass's count is read once, before this workgroup reports its end
  // (the last workgroup to end clears it).  It is the same for every
  // workgroup (only the last one to report changes it), so either all return
  // here or none does; an empty pass reports nothing: its counter stays zero
  // and the previous pass already wrote 0 to *conv.
  // Tiles to do: every tile in the first pass; later, the tiles the previous
  // pass flagged (a word some block of the tile may propose changed, or the
  // tile did not settle).  The other tiles' evaluations are all current.  A
  // pass reads and clears the flags of parity par ^ 1 and sets those of
  // parity par; stale set flags only cost a redo.
  if (threadIdx.x == 0) {
    s_prev = first ? 1u : st->n_changed[par ^ 1];
    s_chg = 0;
  }
  for (int i = threadIdx.x; i < 512; i += blockDim.x) s_snap[i] = c_snap[i];
  init_ab_table(s_ab);
  uint8_t* const fl = wv < n_itiles ? tflag + (size_t)(par ^ 1) * tf_stride + wv : nullptr;
  const bool todo = fl && (first || *fl != 0);
  __syncthreads();
  if (s_prev == 0) {
    report_fixpoint(conv);
    return;
  }
  if (fl && lane == 0) *fl = 0;
  unsigned long long n_cand = 0, n_re = 0;  // first evaluations, re-evaluations
  if (todo) {
    int f = 0;
    while (f + 1 < n_frames && fr[f + 1].itile_base <= wv) ++f;
    const FrameDev& d = fr[f];
    // the wave's tile: ltw x lth blocks, lane = row * ltw + column
    const int ltw = 1 << ltw_log, lth_log = 6 - ltw_log;
    const int lt = wv - d.itile_base, tx = lt % d.itiles_x, ty = lt / d.itiles_x;
    const int bx = (tx << ltw_log) + (lane & (ltw - 1)), by = (ty << lth_log) + (lane >> ltw_log);
    const bool inside = bx < d.bw && by < d.bh;
    const int g = inside ? d.base + by * d.bw + bx : 0;
    const bool act = inside && need[g];
    uint32_t cur = inside ? cw[g] : 0u;
    const uint32_t start = cur;
    const uint32_t own_w = act ? (uint32_t)(own[g] >> 32) : 0u;  // loaded once, not per decision
    int src[K];  // lane holding the proposal's block, or -1 (other tile / none)
    uint32_t w[K], sn[K];
    int er[K], Ls[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int L = (act && i < kk) ? (int)prop[(size_t)g * kk + i] : -1;
      Ls[i] = L;
      src[i] = -1;
      w[i] = 0;
      sn[i] = 0;
      er[i] = kNotEvaluated;
      if (L >= 0) {
        const int qx = bx + L % kNetWin - 8, qy = by - L / kNetWin;
        if ((qx >> ltw_log) == tx && (qy >> lth_log) == ty)
          src[i] = ((qy & ((1 << lth_log) - 1)) << ltw_log) | (qx & (ltw - 1));
        else
          w[i] = cw[d.base + qy * d.bw + qx];
        if (!first) {
          sn[i] = seen[(size_t)g * kk + i];
          er[i] = errs[(size_t)g * kk + i];
        }
      }
    }
    BlockEval be;
    bool have_be = false, published = false, clean_once = first == 0;
    uint64_t dc = 0;
    // the decision the current evaluations give: dc, and the word it makes current
    auto decide = [&]() {
      int best = 0x7fffffff, bi = -1;
#pragma unroll
      for (int i = 0; i < K; ++i)
        if (Ls[i] >= 0 && er[i] != kNotEvaluated && er[i] < best) {
          best = er[i];
          bi = i;
        }
      dc = 0;
      uint32_t nw = own_w;
#pragma unroll
      for (int i = 0; i < K; ++i)
        if (i == bi && best <= thr) {
          const int dy = Ls[i] / kNetWin, dx = Ls[i] % kNetWin - 8;
          dc = pack_dec(sn[i], (uint32_t)(dx + S), (uint32_t)(2 * S - 1 - dy), true);
          nw = sn[i];
        }
      cur = nw;
      published = true;
      clean_once = true;
    };
    if (first) {
      // The first pass evaluates every proposal against the words the pass
      // starts from in one go (a word already evaluated in an earlier slot is
      // taken from it); the rounds below then only follow changes.  The
      // fixpoint is unique, so the order of evaluations does not change it.
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const uint32_t v = __shfl(cur, src[i] >= 0 ? src[i] : lane);
        if (src[i] >= 0) w[i] = v;
      }
      if (act) {
        BlockPx px;
        load_block(d.rgb, d.width, bx, by, px);
        make_block_eval(px, own[g], be);
        have_be = true;
#pragma unroll
        for (int i = 0; i < K; ++i) {
          if (Ls[i] < 0) continue;
          int hit = kNotEvaluated;
#pragma unroll
          for (int k2 = 0; k2 < i; ++k2)
            if (Ls[k2] >= 0 && sn[k2] == w[i]) hit = er[k2];
          if (hit == kNotEvaluated) {
            const CandResult res = eval_candidate<false, true>(be, w[i], cand_table(w[i]), s_snap, s_ab);
            hit = res.valid ? res.err : 0x7fffffff;
            ++n_cand;
          }
          sn[i] = w[i];
          er[i] = hit;
        }
      }
    }
    bool settled = false;
    for (int round = 0; round < kMaxRounds; ++round) {
      // current words of the proposals inside the tile
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const uint32_t v = __shfl(cur, src[i] >= 0 ? src[i] : lane);
        if (src[i] >= 0) w[i] = v;
      }
      // A word some slot has evaluated for this block (sn, er: the error of
      // that word, whether or not the slot's own proposal still holds it)
      // needs no new evaluation: neighbouring blocks often share index words,
      // and a word moving along a chain of reuse arrives in several slots.
#pragma unroll
      for (int i = 0; i < K; ++i) {
        if (Ls[i] >= 0 && (er[i] == kNotEvaluated || w[i] != sn[i])) {
          int hit = kNotEvaluated;
#pragma unroll
          for (int k2 = 0; k2 < K; ++k2)
            if (k2 != i && er[k2] != kNotEvaluated && sn[k2] == w[i]) hit = er[k2];
          if (hit != kNotEvaluated) {
            sn[i] = w[i];
            er[i] = hit;
            clean_once = false;  // decide again when clean
          }
        }
      }
      int j = -1;  // first proposal whose evaluation is stale
#pragma unroll
      for (int i = K - 1; i >= 0; --i)
        if (Ls[i] >= 0 && (er[i] == kNotEvaluated || w[i] != sn[i])) j = i;
      const bool dirty = act && j >= 0;
      // a lane publishes once nothing of it is stale (first pass: after its
      // first full set of evaluations even if nothing changed)
      const bool publish = act && j < 0 && !clean_once;
      if (__ballot(dirty || publish) == 0ull) {
        settled = true;
        break;
      }
      if (dirty) {
        if (!have_be) {
          BlockPx px;
          load_block(d.rgb, d.width, bx, by, px);
          make_block_eval(px, own[g], be);
          have_be = true;
        }
        uint32_t wj = 0;
        bool fresh = false;
#pragma unroll
        for (int i = 0; i < K; ++i)
          if (i == j) {
            wj = w[i];
            fresh = er[i] == kNotEvaluated;
          }
        const CandResult res = eval_candidate<false, true>(be, wj, cand_table(wj), s_snap, s_ab);
        const int e = res.valid ? res.err : 0x7fffffff;
        if (fresh)
          ++n_cand;
        else
          ++n_re;
#pragma unroll
        for (int i = 0; i < K; ++i)
          if (i == j) {
            sn[i] = wj;
            er[i] = e;
          }
        clean_once = false;  // something was re-evaluated: decide again when clean
      } else if (publish) {
        decide();
      }
    }
    if (act && !clean_once) decide();  // round limit reached: store a consistent state
    if (!settled && lane == 0) {  // redo what is stale next pass
      tflag[(size_t)par * tf_stride + wv] = 1;
      atomicAdd(&s_chg, 1u);
    }
    if (published) {
#pragma unroll
      for (int i = 0; i < K; ++i)
        if (Ls[i] >= 0) {
          seen[(size_t)g * kk + i] = sn[i];
          errs[(size_t)g * kk + i] = er[i];
        }
      dec[g] = dc;
      if (cur != start) {
        cw[g] = cur;
        atomicAdd(&s_chg, 1u);
        // the tiles of the blocks that may propose this one: rows by .. by + 15,
        // columns bx - 7 .. bx + 8 (offsets dx in -8..7, dy in 0..15)
        const int c0 = max(bx - 7, 0) >> ltw_log, c1 = min(bx + 8, d.bw - 1) >> ltw_log;
        const int r0 = by >> lth_log, r1 = min(by + 15, d.bh - 1) >> lth_log;
        uint8_t* fl = tflag + (size_t)par * tf_stride + d.itile_base;
        for (int ty2 = r0; ty2 <= r1; ++ty2)
          for (int tx2 = c0; tx2 <= c1; ++tx2) fl[ty2 * d.itiles_x + tx2] = 1;
      }
    }
  }
  // the workgroup's counts through one barrier: wave sums, then thread 0
  for (int o = 32; o > 0; o >>= 1) {
    n_cand += __shfl_xor(n_cand, o);
    n_re += __shfl_xor(n_re, o);
  }
  if (lane == 0) {
    s_cnt[0][threadIdx.x >> 6] = n_cand;
    s_cnt[1][threadIdx.x >> 6] = n_re;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_chg) {
      const unsigned old = atomicAdd(&st->n_changed[par], s_chg);
      if (old == 0) atomicAdd(&st->active_passes, 1u);
    }
    unsigned long long c0 = 0, c1 = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      c0 += s_cnt[0][w];
      c1 += s_cnt[1][w];
    }
    if (c0) atomicAdd(&st->intra_cands[(blockIdx.x % kStatSlots) * kSlot64], c0);
    if (c1) atomicAdd(&st->reeval_cands[(blockIdx.x % kStatSlots) * kSlot64], c1);
  }
  pass_end(st, par, &st->n_changed[par ^ 1], conv);
}

}  // namespace mptc
