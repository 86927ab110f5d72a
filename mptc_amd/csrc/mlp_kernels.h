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
// k_adam, k_to_bf16, k_collect_masks.  Search: k_frame_net_input (the
// network input of a batch) and k_lintra_pass (learned intra fixpoint); the
// inference kernels and the top-k proposal kernels are in net_infer.h.
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

// ---------------------------------------------------------------------------
// GEMM.  Element (m, k) of A is A[m * sam + k * sak]; (k, n) of B is
// B[k * sbk + n * sbn]; elements outside M x K / K x N read as 0.  A workgroup
// of 256 threads computes a 64 x 64 tile of C; wave w owns the 32 x 32
// quadrant (w / 2, w % 2) as 2 x 2 MFMA 16x16x32 tiles.  MFMA operand layout
// (lane l of the wave): A row l % 16, k (l / 16) * 8 .. +7; B column l % 16,
// k (l / 16) * 8 .. +7; D rows (l / 16) * 4 .. +3, column l % 16.
// Split K (gridDim.z > 1, for the weight gradients, whose K is the batch):
// workgroup z takes a 32-aligned range of k and stores its partial product,
// without epilogue, to Cf + z * split_stride; k_split_sum adds the partials
// in a fixed order.
// ---------------------------------------------------------------------------
struct GemmEpi {
  const float* bias;     // [N] or null
  const float* bias2;    // [N] or null: a second, constant bias added after `bias`
  int relu;              // max(0, .) before storing
  const uint16_t* mask;  // bf16 [M x N] (row stride ldm): result zeroed where mask <= 0; or null
  long ldm;
  float* Cf;             // fp32 output or null
  uint16_t* Cb;          // bf16 output or null
  long ldc, ldcb;        // row strides of Cf, Cb
  int trans_f;           // Cf stored transposed: Cf[n * ldc + m]
  long split_stride;     // split K: floats between the partial products
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
  const int kspan = ((K + (int)gridDim.z - 1) / (int)gridDim.z + 31) & ~31;
  const int kb = (int)blockIdx.z * kspan, ke = min(K, kb + kspan);
  // 16-byte loads of 8 elements along an operand's contiguous dimension where
  // the other stride keeps them aligned (k0, m0, n0 are multiples of 8)
  const bool a16 = ((uintptr_t)A & 15u) == 0, b16 = ((uintptr_t)B & 15u) == 0;
  const bool a_k = a16 && sak == 1 && (sam & 7) == 0, a_m = a16 && sam == 1 && (sak & 7) == 0;
  const bool b_k = b16 && sbk == 1 && (sbn & 7) == 0, b_n = b16 && sbn == 1 && (sbk & 7) == 0;
  for (int k0 = kb; k0 < ke; k0 += 32) {
    const bool kfull = k0 + 32 <= ke;
    if (kfull && m0 + 64 <= M && a_k) {  // a row's 8 consecutive k
      const int mm = tid >> 2, kq = (tid & 3) * 8;
      *reinterpret_cast<uint4*>(&sA[mm][kq]) =
          *reinterpret_cast<const uint4*>(A + (long)(m0 + mm) * sam + (k0 + kq));
    } else if (kfull && m0 + 64 <= M && a_m) {  // 8 consecutive m at one k
      const int kk = tid >> 3, mq = (tid & 7) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(A + (long)(m0 + mq) + (long)(k0 + kk) * sak);
      uint16_t h[8];
      __builtin_memcpy(h, &v, 16);
#pragma unroll
      for (int j = 0; j < 8; ++j) sA[mq + j][kk] = h[j];
    } else {
      // consecutive threads walk the contiguous dimension of each operand
      for (int e = tid; e < 64 * 32; e += 256) {
        const int mm = sak == 1 ? e / 32 : e % 64, kk = sak == 1 ? e % 32 : e / 64;
        const int m = m0 + mm, k = k0 + kk;
        sA[mm][kk] = (m < M && k < ke) ? A[(long)m * sam + (long)k * sak] : (uint16_t)0;
      }
    }
    if (kfull && n0 + 64 <= N && b_k) {  // a column's 8 consecutive k
      const int nn = tid >> 2, kq = (tid & 3) * 8;
      *reinterpret_cast<uint4*>(&sB[nn][kq]) =
          *reinterpret_cast<const uint4*>(B + (long)(n0 + nn) * sbn + (k0 + kq));
    } else if (kfull && n0 + 64 <= N && b_n) {  // 8 consecutive n at one k
      const int kk = tid >> 3, nq = (tid & 7) * 8;
      const uint4 v = *reinterpret_cast<const uint4*>(B + (long)(k0 + kk) * sbk + (n0 + nq));
      uint16_t h[8];
      __builtin_memcpy(h, &v, 16);
#pragma unroll
      for (int j = 0; j < 8; ++j) sB[nq + j][kk] = h[j];
    } else {
      for (int e = tid; e < 64 * 32; e += 256) {
        const int nn = sbn == 1 ? e % 64 : e / 32, kk = sbn == 1 ? e / 64 : e % 32;
        const int k = k0 + kk, n = n0 + nn;
        sB[nn][kk] = (k < ke && n < N) ? B[(long)k * sbk + (long)n * sbn] : (uint16_t)0;
      }
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
  float* const Cf = ep.Cf ? ep.Cf + (long)blockIdx.z * ep.split_stride : nullptr;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int t = 0; t < 4; ++t) {
        const int m = m0 + wm + i * 16 + row_q + t, n = n0 + wn + j * 16 + col_in;
        if (m >= M || n >= N) continue;
        float v = acc[i][j][t];
        if (ep.bias) v += ep.bias[n];
        if (ep.bias2) v += ep.bias2[n];
        if (ep.relu) v = fmaxf(v, 0.0f);
        if (ep.mask && !(bf2f(ep.mask[(long)m * ep.ldm + n]) > 0.0f)) v = 0.0f;
        if (Cf) Cf[ep.trans_f ? (long)n * ep.ldc + m : (long)m * ep.ldc + n] = v;
        if (ep.Cb) ep.Cb[(long)m * ep.ldcb + n] = f2bf(v);
      }
}

// C[m * ldc + n] = sum over z = 0 .. splits-1 of P[z * stride + m * N + n].
__global__ void __launch_bounds__(256) k_split_sum(const float* __restrict__ P, int splits, long stride, int M, int N,
                                                   float* __restrict__ C, long ldc) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)M * N) return;
  const int m = (int)(i / N), n = (int)(i % N);
  float s = 0.0f;
  for (int z = 0; z < splits; ++z) s += P[(long)z * stride + i];
  C[(long)m * ldc + n] = s;
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

// Loss of a batch: -log of the probability the softmax of the 256 logits
// gives the row's target set (tmask: 8 words, bit L of word L / 32; one bit =
// the usual cross-entropy), plus seed_w * the mean squared error of the 6
// seeds against the endpoints / 255.  d/dy_c of the set term is
// p_c - [c in set] exp(y_c) / sum_{set} exp(y).  Writes dY = d(mean loss)/dY in
// fp32 and bf16 and adds the batch's mean loss to *loss.  A wave per row (rows
// strided over the grid); target and eps rows are taken through idx like the
// inputs.  A row with an empty set adds only its seed term.
__global__ void __launch_bounds__(256) k_net_loss(const float* __restrict__ Y, int n, const uint32_t* __restrict__ tmask,
                                                  const uint8_t* __restrict__ eps, const int32_t* __restrict__ idx,
                                                  float seed_w, float* __restrict__ dYf, uint16_t* __restrict__ dYb,
                                                  float* __restrict__ loss) {
  __shared__ float s_l[4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float inv_n = 1.0f / (float)n;
  float l = 0.0f;  // this lane's share of the wave's rows' losses
  for (int row = blockIdx.x * 4 + wave; row < n; row += gridDim.x * 4) {
    const long src = idx ? (long)idx[row] : row;
    const float* y = Y + (size_t)row * kNetOut;
    uint32_t tm[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) tm[w] = tmask[src * 8 + w];
    float mx = -3.0e38f;
    for (int c = lane; c < kNetLogits; c += 64) mx = fmaxf(mx, y[c]);
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    float se = 0.0f, ss = 0.0f;  // sum over all logits, over the set
    for (int c = lane; c < kNetLogits; c += 64) {
      const float e = expf(y[c] - mx);
      se += e;
      if ((tm[c >> 5] >> (c & 31)) & 1u) ss += e;
    }
    for (int o = 32; o > 0; o >>= 1) {
      se += __shfl_xor(se, o);
      ss += __shfl_xor(ss, o);
    }
    const float lse = logf(se);
    const bool has_set = ss > 0.0f;
    if (has_set && lane == 0) l += lse - logf(ss);
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const int c = lane + 64 * j;
      if (c >= kNetOut) break;
      float g;
      if (c < kNetLogits) {
        const float e = expf(y[c] - mx);
        const float q = (has_set && ((tm[c >> 5] >> (c & 31)) & 1u)) ? e / ss : 0.0f;
        g = has_set ? (e / se - q) * inv_n : 0.0f;
      } else {
        const float t = (float)eps[src * kNetSeeds + (c - kNetLogits)] * (1.0f / 255.0f);
        const float d = y[c] - t;
        l += seed_w * d * d * (1.0f / kNetSeeds);
        g = 2.0f * seed_w * d * (1.0f / kNetSeeds) * inv_n;
      }
      dYf[(size_t)row * kNetOut + c] = g;
      dYb[(size_t)row * kNetOut + c] = f2bf(g);
    }
  }
  // one atomic per workgroup: same-address atomics serialise
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if (lane == 0) s_l[wave] = l;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (s_l[0] + s_l[1] + s_l[2] + s_l[3]) * inv_n);
}

// Column sums in two deterministic stages: part[p][c] = sum of rows
// p*kColsumRows .. (p+1)*kColsumRows - 1, then out[c] = sum over p in a fixed
// order (k_colsum_final).
constexpr int kColsumRows = 32;  // a 4096-row batch: 128 partials per column
__global__ void __launch_bounds__(256) k_colsum_part(const float* __restrict__ X, int rows, int cols, long ld,
                                                     float* __restrict__ part) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, p = blockIdx.y;
  if (c >= cols) return;
  float s = 0.0f;
  const int r1 = min(rows, (p + 1) * kColsumRows);
  for (int r = p * kColsumRows; r < r1; ++r) s += X[(long)r * ld + c];
  part[(long)p * cols + c] = s;
}

// out[c] = sum over p of part[p][c]: a wave per column; lane l adds parts
// l, l + 64, ... in order, then the wave adds its lanes in a fixed tree (the
// same order every run).
__global__ void __launch_bounds__(256) k_colsum_final(const float* __restrict__ part, int parts, int cols,
                                                      float* __restrict__ out) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (c >= cols) return;  // wave-uniform
  float s = 0.0f;
  for (int p = lane; p < parts; p += 64) s += part[(long)p * cols + c];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) out[c] = s;
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
// the next pass (never-done evaluations are stored as kNotEvaluated).  Pass
// counters: this pass adds its changed words to n_changed[par] (zero at its
// start) and reads the previous pass's, n_changed[par ^ 1], which its last
// workgroup clears for the next pass (pass_end); the count also goes to *conv.
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
  __shared__ unsigned long long s_acc;
  __shared__ unsigned int s_chg, s_prev;
  // After a pass that changed no word every evaluation is current: nothing
  // to do.  The previous pass's count is read once, before this workgroup
  // reports its end (the last workgroup to end clears it).  It is the same
  // for every workgroup (only the last one to report changes it), so either
  // all return here or none does; an empty pass reports nothing: its counter
  // stays zero and the previous pass already wrote 0 to *conv.
  if (threadIdx.x == 0) s_prev = first ? 1u : st->n_changed[par ^ 1];
  __syncthreads();
  if (s_prev == 0) {
    report_fixpoint(conv);
    return;
  }
  load_lds_tables(s_snap, s_ab);
  if (threadIdx.x == 0) s_chg = 0;
  __syncthreads();
  const int wv = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int lane = (int)(threadIdx.x & 63);
  unsigned long long n_cand = 0, n_re = 0;  // first evaluations, re-evaluations
  // Tiles to do: every tile in the first pass; later, the tiles the previous
  // pass flagged (a word some block of the tile may propose changed, or the
  // tile did not settle).  The other tiles' evaluations are all current.  A
  // pass reads and clears the flags of parity par ^ 1 and sets those of
  // parity par; stale set flags only cost a redo.
  bool todo = false;
  if (wv < n_itiles) {
    uint8_t* fl = tflag + (size_t)(par ^ 1) * tf_stride + wv;
    todo = first || *fl != 0;
    if (lane == 0) *fl = 0;
  }
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
    if (!settled && lane == 0) tflag[(size_t)par * tf_stride + wv] = 1;  // redo what is stale next pass
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
  __syncthreads();
  if (threadIdx.x == 0 && s_chg) {
    const unsigned old = atomicAdd(&st->n_changed[par], s_chg);
    if (old == 0) atomicAdd(&st->active_passes, 1u);
  }
  wg_count(&s_acc, n_cand, st->intra_cands);
  wg_count(&s_acc, n_re, st->reeval_cands);
  pass_end(st, par, &st->n_changed[par ^ 1], conv);
}

}  // namespace mptc
