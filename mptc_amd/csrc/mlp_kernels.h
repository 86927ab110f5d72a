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
// Training kernels: k_gemm_bf16, k_net_input, k_net_loss, k_colsum_part/final,
// k_adam, k_to_bf16, k_collect_pairs.  Search kernels: k_gather_blocks,
// k_propose_inter, k_top_intra, k_lintra_pass.
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

// ---------------------------------------------------------------------------
// GEMM.  Element (m, k) of A is A[m * sam + k * sak]; (k, n) of B is
// B[k * sbk + n * sbn]; elements outside M x K / K x N read as 0.  A workgroup
// of 256 threads computes a 64 x 64 tile of C; wave w owns the 32 x 32
// quadrant (w / 2, w % 2) as 2 x 2 MFMA 16x16x32 tiles.  MFMA operand layout
// (lane l of the wave): A row l % 16, k (l / 16) * 8 .. +7; B column l % 16,
// k (l / 16) * 8 .. +7; D rows (l / 16) * 4 .. +3, column l % 16.
// ---------------------------------------------------------------------------
struct GemmEpi {
  const float* bias;     // [N] or null
  int relu;              // max(0, .) before storing
  const uint16_t* mask;  // bf16 [M x N] (row stride ldm): result zeroed where mask <= 0; or null
  long ldm;
  float* Cf;             // fp32 output or null
  uint16_t* Cb;          // bf16 output or null
  long ldc, ldcb;        // row strides of Cf, Cb
  int trans_f;           // Cf stored transposed: Cf[n * ldc + m]
};

__global__ void __launch_bounds__(256) k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A, long sam,
                                                   long sak, const uint16_t* __restrict__ B, long sbk, long sbn,
                                                   GemmEpi ep) {
  // k-contiguous rows (80 bytes: 16-byte aligned fragments, skewed banks)
  __shared__ __attribute__((aligned(16))) uint16_t sA[64][32 + 8];
  __shared__ __attribute__((aligned(16))) uint16_t sB[64][32 + 8];  // B transposed: sB[n][k]
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4_t acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    // consecutive threads walk the contiguous dimension of each operand
    for (int e = tid; e < 64 * 32; e += 256) {
      const int mm = sak == 1 ? e / 32 : e % 64, kk = sak == 1 ? e % 32 : e / 64;
      const int m = m0 + mm, k = k0 + kk;
      sA[mm][kk] = (m < M && k < K) ? A[(long)m * sam + (long)k * sak] : (uint16_t)0;
    }
    for (int e = tid; e < 64 * 32; e += 256) {
      const int nn = sbn == 1 ? e % 64 : e / 32, kk = sbn == 1 ? e / 64 : e % 32;
      const int k = k0 + kk, n = n0 + nn;
      sB[nn][kk] = (k < K && n < N) ? B[(long)k * sbk + (long)n * sbn] : (uint16_t)0;
    }
    __syncthreads();
    const int r = lane & 15, kq = (lane >> 4) * 8;
    bf16x8_t a[2], b[2];
    for (int i = 0; i < 2; ++i) {
      a[i] = *reinterpret_cast<const bf16x8_t*>(&sA[wm + i * 16 + r][kq]);
      b[i] = *reinterpret_cast<const bf16x8_t*>(&sB[wn + i * 16 + r][kq]);
    }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  const int col_in = lane & 15, row_q = (lane >> 4) * 4;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int t = 0; t < 4; ++t) {
        const int m = m0 + wm + i * 16 + row_q + t, n = n0 + wn + j * 16 + col_in;
        if (m >= M || n >= N) continue;
        float v = acc[i][j][t];
        if (ep.bias) v += ep.bias[n];
        if (ep.relu) v = fmaxf(v, 0.0f);
        if (ep.mask && !(bf2f(ep.mask[(long)m * ep.ldm + n]) > 0.0f)) v = 0.0f;
        if (ep.Cf) ep.Cf[ep.trans_f ? (long)n * ep.ldc + m : (long)m * ep.ldc + n] = v;
        if (ep.Cb) ep.Cb[(long)m * ep.ldcb + n] = f2bf(v);
      }
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

// Loss of a batch: softmax cross-entropy of the 256 logits against the chosen
// offset, plus seed_w * the mean squared error of the 6 seeds against the
// re-fit endpoints / 255.  Writes dY = d(mean loss)/dY in fp32 and bf16 and
// adds the batch's mean loss to *loss.  One wave per row; target and eps rows
// are taken through idx like the inputs.
__global__ void __launch_bounds__(256) k_net_loss(const float* __restrict__ Y, int n, const int32_t* __restrict__ target,
                                                  const uint8_t* __restrict__ eps, const int32_t* __restrict__ idx,
                                                  float seed_w, float* __restrict__ dYf, uint16_t* __restrict__ dYb,
                                                  float* __restrict__ loss) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  const long src = idx ? (long)idx[row] : row;
  const float* y = Y + (size_t)row * kNetOut;
  float mx = -3.0e38f;
  for (int c = lane; c < kNetLogits; c += 64) mx = fmaxf(mx, y[c]);
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  float se = 0.0f;
  for (int c = lane; c < kNetLogits; c += 64) se += expf(y[c] - mx);
  for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o);
  const int tgt = target[src];
  const float inv_n = 1.0f / (float)n, lse = logf(se);
  float l = 0.0f;
  for (int c = lane; c < kNetOut; c += 64) {
    float g;
    if (c < kNetLogits) {
      const float p = expf(y[c] - mx - lse);
      g = (p - (c == tgt ? 1.0f : 0.0f)) * inv_n;
      if (c == tgt) l += lse - (y[c] - mx);
    } else {
      const float t = (float)eps[src * kNetSeeds + (c - kNetLogits)] * (1.0f / 255.0f);
      const float d = y[c] - t;
      l += seed_w * d * d * (1.0f / kNetSeeds);
      g = 2.0f * seed_w * d * (1.0f / kNetSeeds) * inv_n;
    }
    dYf[(size_t)row * kNetOut + c] = g;
    dYb[(size_t)row * kNetOut + c] = f2bf(g);
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if (lane == 0) atomicAdd(loss, l * inv_n);
}

// Column sums in two deterministic stages: part[p][c] = sum of rows
// p*256 .. p*256+255, then out[c] = sum over p in order.
__global__ void __launch_bounds__(256) k_colsum_part(const float* __restrict__ X, int rows, int cols, long ld,
                                                     float* __restrict__ part) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, p = blockIdx.y;
  if (c >= cols) return;
  float s = 0.0f;
  const int r1 = min(rows, (p + 1) * 256);
  for (int r = p * 256; r < r1; ++r) s += X[(long)r * ld + c];
  part[(long)p * cols + c] = s;
}

__global__ void __launch_bounds__(256) k_colsum_final(const float* __restrict__ part, int parts, int cols,
                                                      float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.0f;
  for (int p = 0; p < parts; ++p) s += part[(long)p * cols + c];
  out[c] = s;
}

// Adam on the fp32 master parameters; refreshes their bf16 copy.
// c1, c2 = 1 - beta1^t, 1 - beta2^t.
__global__ void __launch_bounds__(256) k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, uint16_t* __restrict__ wb, int n, float lr,
                                              float b1, float b2, float c1, float c2, float epsv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + epsv);
  w[i] = wi;
  wb[i] = f2bf(wi);
}

__global__ void __launch_bounds__(256) k_to_bf16(const float* __restrict__ x, int n, uint16_t* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// Logit index of a decision's motion bytes (codec/dxt_image.cpp:760-764 for
// inter, :683-687 for intra), or -1 when the offset lies outside the 16x16
// grid (search areas above 8).
__device__ __forceinline__ int logit_of_motion(int net, uint32_t mx, uint32_t my, int S) {
  const int dx = (int)(mx & 0x7F) - S;
  const int dy = net == 0 ? (int)(my & 0x7F) - S : 2 * S - 1 - (int)my;
  if (dx < -8 || dx > 7) return -1;
  if (net == 0) return (dy < -8 || dy > 7) ? -1 : (dy + 8) * kNetWin + (dx + 8);
  return (dy < 0 || dy > 15) ? -1 : dy * kNetWin + (dx + 8);
}

// Training pairs of one encoded frame: every block the search reused from the
// window this network scores (net 0: inter, motion x has bit 7; net 1:
// intra) gives (its 48 pixels, the logit index of its offset, its re-fit
// endpoints).  Pairs are appended in no particular order.
__global__ void __launch_bounds__(256) k_collect_pairs(const uint8_t* __restrict__ rgb, int width, int bw, int nb,
                                                       const uint8_t* __restrict__ motion,
                                                       const uint8_t* __restrict__ eps, int S, int net,
                                                       uint8_t* __restrict__ px_out, int32_t* __restrict__ tgt_out,
                                                       uint8_t* __restrict__ eps_out, unsigned int* __restrict__ count,
                                                       int capacity) {
  __shared__ unsigned int s_base;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  int L = -1;
  if (k < nb) {
    const uint32_t mx = motion[2 * k], my = motion[2 * k + 1];
    const bool reused = mx != 255u;
    const bool inter = (mx & 0x80u) != 0;
    if (reused && inter == (net == 0)) L = logit_of_motion(net, mx, my, S);
  }
  const bool want = L >= 0;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  const unsigned long long m = __ballot(want);
  const int lane = (int)__lane_id();
  unsigned int off = 0;
  if (m && lane == 0) off = atomicAdd(&s_base, (unsigned int)__popcll(m));
  off = __shfl(off, 0);
  __syncthreads();
  if (threadIdx.x == 0 && s_base) s_base = atomicAdd(count, s_base);
  __syncthreads();
  if (!want) return;
  const unsigned slot = s_base + off + (unsigned)__popcll(m & ((1ull << lane) - 1));
  if (slot >= (unsigned)capacity) return;
  BlockPx b;
  load_block(rgb, width, k % bw, k / bw, b);
  uint32_t* o = reinterpret_cast<uint32_t*>(px_out + (size_t)slot * kNetIn);
  for (int w = 0; w < 12; ++w) o[w] = b.w[w];
  tgt_out[slot] = L;
  for (int c = 0; c < kNetSeeds; ++c) eps_out[(size_t)slot * kNetSeeds + c] = eps[(size_t)kNetSeeds * k + c];
}

// Block pixels of every block of a batch -> total x 48 bytes.
__global__ void __launch_bounds__(256) k_gather_blocks(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid,
                                                       int total, uint8_t* __restrict__ px) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const FrameDev& d = fr[sid[g]];
  const int k = g - d.base;
  BlockPx b;
  load_block(d.rgb, d.width, k % d.bw, k / d.bw, b);
  uint4* o = reinterpret_cast<uint4*>(px + (size_t)g * kNetIn);
  o[0] = make_uint4(b.w[0], b.w[1], b.w[2], b.w[3]);
  o[1] = make_uint4(b.w[4], b.w[5], b.w[6], b.w[7]);
  o[2] = make_uint4(b.w[8], b.w[9], b.w[10], b.w[11]);
}

// Network input of every block of a batch: total x 48 bf16 (pixel / 255).
__global__ void __launch_bounds__(256) k_frame_net_input(const FrameDev* __restrict__ fr,
                                                         const uint8_t* __restrict__ sid, int total,
                                                         uint16_t* __restrict__ X) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const FrameDev& d = fr[sid[g]];
  const int k = g - d.base;
  BlockPx b;
  load_block(d.rgb, d.width, k % d.bw, k / d.bw, b);
  uint4* o = reinterpret_cast<uint4*>(X + (size_t)g * kNetIn);
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    uint32_t h[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint32_t w = b.w[2 * q + (e >> 1)];
      const uint32_t lo = (w >> (16 * (e & 1))) & 0xFFu, hi = (w >> (16 * (e & 1) + 8)) & 0xFFu;
      h[e] = (uint32_t)f2bf((float)lo * (1.0f / 255.0f)) | ((uint32_t)f2bf((float)hi * (1.0f / 255.0f)) << 16);
    }
    o[q] = make_uint4(h[0], h[1], h[2], h[3]);
  }
}

// Top-k logit indices (descending score, ties to the lower index) among the
// offsets `valid` admits.  Scores are Ys[L * ys_l + g * ys_g]: the logits of
// the network (transposed, ys_l = rows) or one shared prior (ys_g = 0).
template <typename Valid>
__device__ __forceinline__ void top_k(const float* __restrict__ Ys, long ys_l, long ys_g, int g, int kk,
                                      Valid&& valid, int (&ti)[kMaxTopK]) {
  float tv[kMaxTopK];
#pragma unroll
  for (int i = 0; i < kMaxTopK; ++i) {
    tv[i] = -3.0e38f;
    ti[i] = -1;
  }
  const float* y = Ys + (long)g * ys_g;
  for (int L = 0; L < kNetLogits; ++L) {
    if (!valid(L)) continue;
    float v = y[(long)L * ys_l];
    if (v != v) v = -3.0e38f;
    int cand = L;
    // insertion into the sorted list (static indices keep it in registers)
#pragma unroll
    for (int i = 0; i < kMaxTopK; ++i) {
      if (i < kk && cand >= 0 && (ti[i] < 0 || v > tv[i])) {
        const float t = tv[i];
        const int u = ti[i];
        tv[i] = v;
        ti[i] = cand;
        v = t;
        cand = u;
      }
    }
  }
}

// Learned inter search: the top_k offsets are evaluated exactly (the same
// re-fit and error as InterBlockSearch, codec/dxt_image.cpp:715-774); the
// valid one with the smallest error delta (first in score order on ties) is
// reused if the delta passes the threshold (:890).
__global__ void __launch_bounds__(256) k_propose_inter(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid,
                                                       int total, int S, int thr, int kk,
                                                       const float* __restrict__ Ys, long ys_l, long ys_g,
                                                       const uint64_t* __restrict__ own,
                                                       const uint32_t* __restrict__ rw,
                                                       const float4* __restrict__ rtab,
                                                       uint64_t* __restrict__ dec_out, DevStats* __restrict__ st) {
  __shared__ uint8_t s_snap[512];
  __shared__ float2 s_ab[4];
  __shared__ unsigned long long s_acc;
  load_lds_tables(s_snap, s_ab);
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n_cand = 0;
  if (g < total && !fr[sid[g]].is_intra) {
    const FrameDev& d = fr[sid[g]];
    const int k = g - d.base, bx = k % d.bw, by = k / d.bw, bw = d.bw, bh = d.bh;
    const int lo = -min(S, 8), hi = min(S, 8) - 1;
    auto valid = [&](int L) {
      const int dy = L / kNetWin - 8, dx = L % kNetWin - 8;
      return dy >= lo && dy <= hi && dx >= lo && dx <= hi && by + dy >= 0 && by + dy < bh && bx + dx >= 0 &&
             bx + dx < bw;
    };
    int ti[kMaxTopK];
    top_k(Ys, ys_l, ys_g, g, kk, valid, ti);
    BlockPx px;
    load_block(d.rgb, d.width, bx, by, px);
    BlockEval be;
    make_block_eval(px, own[g], be);
    int best = 0x7fffffff, best_L = -1;
    uint32_t best_c = 0;
#pragma unroll
    for (int i = 0; i < kMaxTopK; ++i) {
      const int L = i < kk ? ti[i] : -1;
      if (L >= 0) {
        const int q = d.base + (by + L / kNetWin - 8) * bw + (bx + L % kNetWin - 8);
        const float4 t4 = rtab[q];
        const uint32_t c = rw[q];
        const CandResult res = eval_candidate<false>(be, c, CandTab{t4.x, t4.y, t4.z, t4.w}, s_snap, s_ab);
        ++n_cand;
        if (res.valid && res.err < best) {
          best = res.err;
          best_L = L;
          best_c = c;
        }
      }
    }
    uint64_t dc = 0;
    if (best_L >= 0 && best <= thr)
      dc = pack_dec(best_c, (uint32_t)(best_L % kNetWin - 8 + S) | 0x80u, (uint32_t)(best_L / kNetWin - 8 + S) | 0x80u,
                    true);
    dec_out[g] = dc;
  }
  wg_count(&s_acc, n_cand, st->inter_cands);
}

// Learned intra search, step 1: the top_k causal-window offsets of every
// block that needs the intra search (prop[g * kk + i], -1 past the end).
__global__ void __launch_bounds__(256) k_top_intra(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid,
                                                   int total, int S, int kk, const float* __restrict__ Ys, long ys_l,
                                                   long ys_g, const uint8_t* __restrict__ need,
                                                   int16_t* __restrict__ prop) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total || !need[g]) return;
  const FrameDev& d = fr[sid[g]];
  const int k = g - d.base, bx = k % d.bw, by = k / d.bw, bw = d.bw;
  const int lo = -min(S, 8), hi = min(S, 8) - 1, dmax = min(2 * S - 1, 15);
  auto valid = [&](int L) {
    const int dy = L / kNetWin, dx = L % kNetWin - 8;
    return dy <= dmax && dx >= lo && dx <= hi && by - dy >= 0 && bx + dx >= 0 && bx + dx < bw && (dy > 0 || dx < 0);
  };
  int ti[kMaxTopK];
  top_k(Ys, ys_l, ys_g, g, kk, valid, ti);
#pragma unroll
  for (int i = 0; i < kMaxTopK; ++i)
    if (i < kk) prop[(size_t)g * kk + i] = (int16_t)ti[i];
}

// Learned intra search, one fixpoint pass.  Every block that needs the intra
// search compares the current words at its proposed offsets with the words it
// last evaluated; where one differs (or on the first pass) it re-evaluates the
// changed ones, takes the valid proposal with the smallest error delta (first
// in score order on ties), accepts it if the delta passes the threshold, and
// writes its resulting word as current at once.  The candidate table of a
// word is derived from the word itself, so a word read while another
// workgroup rewrites it is still evaluated consistently; a stale read is
// caught by the next pass's comparison.  Each block's window holds only
// earlier blocks, so a pass that changes no word ends at the unique fixpoint.
// Pass counters: this pass adds its changed words to n_changed[par] and
// clears n_changed[par ^ 1] for the next pass.
__global__ void __launch_bounds__(256) k_lintra_pass(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid,
                                                     int total, int S, int thr, int kk,
                                                     const int16_t* __restrict__ prop,
                                                     const uint64_t* __restrict__ own, const uint8_t* __restrict__ need,
                                                     uint32_t* __restrict__ cw, uint32_t* __restrict__ seen,
                                                     int* __restrict__ errs, uint64_t* __restrict__ dec,
                                                     DevStats* __restrict__ st, int par, int first) {
  __shared__ uint8_t s_snap[512];
  __shared__ float2 s_ab[4];
  __shared__ unsigned long long s_acc;
  __shared__ unsigned int s_chg;
  load_lds_tables(s_snap, s_ab);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->n_changed[par ^ 1] = 0;
  if (threadIdx.x == 0) s_chg = 0;
  __syncthreads();
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n_cand = 0;
  if (g < total && need[g]) {
    const FrameDev& d = fr[sid[g]];
    const int k = g - d.base, bx = k % d.bw, by = k / d.bw;
    uint32_t w[kMaxTopK];
    bool dirty = first != 0;
#pragma unroll
    for (int i = 0; i < kMaxTopK; ++i) {
      w[i] = 0;
      if (i < kk) {
        const int L = prop[(size_t)g * kk + i];
        if (L >= 0) {
          const int q = d.base + (by - L / kNetWin) * d.bw + (bx + L % kNetWin - 8);
          w[i] = __builtin_nontemporal_load(&cw[q]);
          if (!first && w[i] != seen[(size_t)g * kk + i]) dirty = true;
        }
      }
    }
    if (dirty) {
      BlockPx px;
      load_block(d.rgb, d.width, bx, by, px);
      BlockEval be;
      const uint64_t o = own[g];
      make_block_eval(px, o, be);
      int best = 0x7fffffff, bi = -1, bL = -1;
#pragma unroll
      for (int i = 0; i < kMaxTopK; ++i) {
        const int L = i < kk ? (int)prop[(size_t)g * kk + i] : -1;
        if (L >= 0) {
          const size_t s = (size_t)g * kk + i;
          int e;
          if (first || w[i] != seen[s]) {
            const CandResult res = eval_candidate<false>(be, w[i], cand_table(w[i]), s_snap, s_ab);
            ++n_cand;
            e = res.valid ? res.err : 0x7fffffff;
            seen[s] = w[i];
            errs[s] = e;
          } else {
            e = errs[s];
          }
          if (e < best) {
            best = e;
            bi = i;
            bL = L;
          }
        }
      }
      uint64_t dc = 0;
      uint32_t nw = (uint32_t)(o >> 32);
      if (bi >= 0 && best <= thr) {
        const int dy = bL / kNetWin, dx = bL % kNetWin - 8;
        dc = pack_dec(w[bi], (uint32_t)(dx + S), (uint32_t)(2 * S - 1 - dy), true);
        nw = w[bi];
      }
      dec[g] = dc;
      if (nw != cw[g]) {
        cw[g] = nw;
        atomicAdd(&s_chg, 1u);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_chg) {
    const unsigned old = atomicAdd(&st->n_changed[par], s_chg);
    if (old == 0) atomicAdd(&st->active_passes, 1u);
  }
  wg_count(&s_acc, n_cand, first ? st->intra_cands : st->reeval_cands);
}

}  // namespace mptc
