// Inference kernels of the proposal network (mlp_kernels.h): two fused
// kernels whose weights stay resident in LDS for the whole launch.
//
//   k_net_l12   x -> h1 -> h2: layers 1 and 2 in registers, h2 stored to HBM
//               in the operand order layer 3 reads
//   k_net_l3    h2 -> y: layer 3, then either the 262 outputs (row-major
//               fp32) or, fused in the epilogue, each block's top-k window
//               offsets among those its search admits
//
// MFMA v_mfma_f32_32x32x16_bf16 computes D[n][m] = sum_k A[n][k] B[k][m] with
// A = W^T (32 outputs n x 16 inputs k), B = activations^T (16 k x 32 blocks m).
// Lane l (h = l / 32) holds A row l % 32, k = 8h .. 8h+7; B column l % 32,
// k = 8h .. 8h+7; D column l % 32, rows 8(i/4) + 4h + i%4, i = 0..15.
// A layer's D tiles become the next layer's B operands without leaving the
// registers: for k-step s, lane half h takes the 8 D values i = 8(s%2) .. +7
// of tile s/2, which hold hidden unit f(s, h, j) = 16s + 8(j/4) + 4h + j%4
// (j = 0..7).  The next layer's weights are stored with their input index
// permuted the same way (position p = 16s + 8h + j holds input f(s, h, j)),
// so its A fragments are contiguous 16-byte LDS reads.  Rows of the LDS
// weight images are split in 16-byte chunks stored at chunk ^ (row % 8), so
// the 8 lanes of an LDS read phase hit 8 different bank groups.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp_kernels.h"

namespace mptc {

typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;

constexpr int kL3Rows = 288;  // 262 outputs padded to 9 tiles of 32
// packed inference image (bf16 units): W1^T [256][48], W2^T [256][256] and
// W3^T [288][256] with permuted inputs; then fp32 biases b1, b2, b3 (288)
constexpr int kPkW1 = 0;
constexpr int kPkW2 = kPkW1 + kNetHid * kNetIn;
constexpr int kPkW3 = kPkW2 + kNetHid * kNetHid;
constexpr int kPkBias = kPkW3 + kL3Rows * kNetHid;         // bf16 offset of the fp32 biases
constexpr int kPkSize = kPkBias + 2 * (2 * kNetHid + kL3Rows);  // bf16 units
constexpr int kH2TileBytes = 16 * 64 * 16;                // one 32-block tile of h2 operands

// hidden unit held at permuted input position p
__host__ __device__ __forceinline__ int perm_unit(int p) {
  const int s = p >> 4, h = (p >> 3) & 1, j = p & 7;
  return 16 * s + 8 * (j >> 2) + 4 * h + (j & 3);
}

// fp32 master parameters -> packed inference image; base (kNetOut or null)
// is added to the layer-3 biases (fixed scores under the logits).
__global__ void __launch_bounds__(256) k_pack_infer(const float* __restrict__ w, const float* __restrict__ base,
                                                    uint16_t* __restrict__ pk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kNetHid * kNetIn) {  // W1^T[n][k] = W1[k][n]
    const int n = i / kNetIn, k = i % kNetIn;
    pk[kPkW1 + i] = f2bf(w[kOffW1 + k * kNetHid + n]);
  }
  if (i < kNetHid * kNetHid) {  // W2^T[n][p] = W2[f(p)][n]
    const int n = i / kNetHid, p = i % kNetHid;
    pk[kPkW2 + i] = f2bf(w[kOffW2 + perm_unit(p) * kNetHid + n]);
  }
  if (i < kL3Rows * kNetHid) {  // W3^T[n][p] = W3[f(p)][n], rows >= 262 zero
    const int n = i / kNetHid, p = i % kNetHid;
    pk[kPkW3 + i] = n < kNetOut ? f2bf(w[kOffW3 + perm_unit(p) * kNetOut + n]) : (uint16_t)0;
  }
  float* bias = reinterpret_cast<float*>(pk + kPkBias);
  if (i < kNetHid) {
    bias[i] = w[kOffB1 + i];
    bias[kNetHid + i] = w[kOffB2 + i];
  }
  if (i < kL3Rows) bias[2 * kNetHid + i] = i < kNetOut ? w[kOffB3 + i] + (base ? base[i] : 0.0f) : 0.0f;
}

// LDS swizzle of the weight images: chunk c of row r sits at c ^ (r & kSwz).
// An MFMA operand read has 16 consecutive rows read the same chunk index
// together; with kSwz = 15 they land on 16 different 16-byte bank groups.
constexpr int kSwz = 15;

// Copies `rows` rows of `chunks` 16-byte chunks into LDS, chunk c of row r at
// r * chunks + (c ^ (r & kSwz)) (swizzled; chunks >= 16) or r * chunks + c.
__device__ __forceinline__ void stage_rows(uint4* __restrict__ dst, const uint4* __restrict__ src, int rows,
                                           int chunks, bool swz) {
  for (int e = threadIdx.x; e < rows * chunks; e += blockDim.x) {
    const int r = e / chunks, c = e % chunks;
    dst[r * chunks + (swz ? (c ^ (r & kSwz)) : c)] = src[e];
  }
}

__device__ __forceinline__ bf16x8_t as_bf16x8(uint4 v) {
  bf16x8_t r;
  __builtin_memcpy(&r, &v, 16);
  return r;
}

// D values i = 8*half .. +7 of a tile (bias included), ReLU, as a bf16 B operand.
__device__ __forceinline__ uint4 relu_pack(const f32x16_t& d, int half) {
  uint32_t w[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i0 = 8 * half + 2 * q, i1 = i0 + 1;
    const float v0 = fmaxf(d[i0], 0.0f), v1 = fmaxf(d[i1], 0.0f);
    w[q] = pack_bf16x2(v0, v1);
  }
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// Accumulator tile nb of a lane (half h) starting at its outputs' biases:
// D value i is output nb * 32 + 8 (i / 4) + 4 h + i % 4.
__device__ __forceinline__ f32x16_t bias_tile(const float* __restrict__ sb, int nb, int h) {
  f32x16_t a;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float4 b4 = *reinterpret_cast<const float4*>(&sb[nb * 32 + 8 * q + 4 * h]);
    a[4 * q] = b4.x;
    a[4 * q + 1] = b4.y;
    a[4 * q + 2] = b4.z;
    a[4 * q + 3] = b4.w;
  }
  return a;
}

constexpr int kInferWaves = 8;  // waves per workgroup of k_net_l3 (255 VGPRs: two per SIMD)
constexpr int kL12Waves = 16;   // waves per workgroup of k_net_l12 (<= 128 VGPRs: four per SIMD)

// Layers 1 and 2.  X: rows x 48 bf16; H2p: per 32-block tile, per k-step s,
// per lane, the 8 bf16 layer-3 operand values (kH2TileBytes per tile).
// One workgroup per CU; each wave loops over tiles.
__global__ void __launch_bounds__(64 * kL12Waves) k_net_l12(const uint16_t* __restrict__ X, int rows,
                                                 const uint16_t* __restrict__ pk, uint4* __restrict__ H2p) {
  __shared__ uint4 sW1[kNetHid * kNetIn / 8];   // 24 KiB, rows of 6 chunks
  __shared__ uint4 sW2[kNetHid * kNetHid / 8];  // 128 KiB, rows of 32 chunks, swizzled
  __shared__ __attribute__((aligned(16))) float sb[2 * kNetHid];
  stage_rows(sW1, reinterpret_cast<const uint4*>(pk + kPkW1), kNetHid, kNetIn / 8, false);
  stage_rows(sW2, reinterpret_cast<const uint4*>(pk + kPkW2), kNetHid, kNetHid / 8, true);
  const float* bias = reinterpret_cast<const float*>(pk + kPkBias);
  for (int i = threadIdx.x; i < 2 * kNetHid; i += blockDim.x) sb[i] = bias[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, m = lane & 31, h = lane >> 5;
  const int n_tiles = (rows + 31) / 32;
  const int wave_g = blockIdx.x * kL12Waves + (threadIdx.x >> 6);
  for (int tile = wave_g; tile < n_tiles; tile += gridDim.x * kL12Waves) {
    const int row = tile * 32 + m;
    // layer 1: B operands straight from X (k = 16s + 8h .. +7)
    bf16x8_t xb[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const uint4 v = row < rows ? *reinterpret_cast<const uint4*>(X + (size_t)row * kNetIn + 16 * s + 8 * h)
                                 : make_uint4(0, 0, 0, 0);
      xb[s] = as_bf16x8(v);
    }
    bf16x8_t hb[16];
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
      f32x16_t acc = bias_tile(sb, nb, h);
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const bf16x8_t a = as_bf16x8(sW1[(nb * 32 + m) * 6 + 2 * s + h]);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, xb[s], acc, 0, 0, 0);
      }
      hb[2 * nb] = as_bf16x8(relu_pack(acc, 0));
      hb[2 * nb + 1] = as_bf16x8(relu_pack(acc, 1));
      __builtin_amdgcn_sched_barrier(0);  // one accumulator live at a time
    }
    // layer 2 (one output tile at a time keeps the live registers to h1 + one accumulator)
    uint4* out = H2p + (size_t)tile * (kH2TileBytes / 16);
#pragma unroll 1
    for (int nb = 0; nb < 8; ++nb) {
      f32x16_t acc = bias_tile(sb + kNetHid, nb, h);
      const int r = nb * 32 + m;
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const bf16x8_t a = as_bf16x8(sW2[r * 32 + ((2 * s + h) ^ (r & kSwz))]);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, hb[s], acc, 0, 0, 0);
      }
      out[(2 * nb) * 64 + lane] = relu_pack(acc, 0);
      out[(2 * nb + 1) * 64 + lane] = relu_pack(acc, 1);
    }
  }
}

// Which window offsets a block's search admits: the offsets of the 16x16
// grid within the search area and the frame, and for the intra network only
// causal ones (rows dy > 0, or dy = 0 with dx < 0).  Bit r of `rows`, bit c
// of `cols` (row 0: `cols0`) admit logit L = 16 r + c.
struct WinMask {
  uint32_t rows, cols, cols0;
  __device__ __forceinline__ bool operator()(int L) const {
    const int r = L >> 4, c = L & 15;
    return ((rows >> r) & ((r == 0 ? cols0 : cols) >> c) & 1u) != 0;
  }
};

__device__ __forceinline__ WinMask win_mask(int net, int S, int bx, int by, int bw, int bh) {
  const int lo = -min(S, 8), hi = min(S, 8) - 1;
  WinMask m{0u, 0u, 0u};
  for (int c = 0; c < kNetWin; ++c) {
    const int dx = c - 8;
    if (dx >= lo && dx <= hi && bx + dx >= 0 && bx + dx < bw) m.cols |= 1u << c;
  }
  for (int r = 0; r < kNetWin; ++r) {
    bool ok;
    if (net == 0) {
      const int dy = r - 8;
      ok = dy >= lo && dy <= hi && by + dy >= 0 && by + dy < bh;
    } else {
      ok = r <= min(2 * S - 1, 15) && by - r >= 0;
    }
    if (ok) m.rows |= 1u << r;
  }
  m.cols0 = net == 0 ? m.cols : (m.cols & 0xFFu);  // intra row 0: dx < 0
  return m;
}

// Index similarity of a candidate: the number of pixels (0..16) whose 2-bit
// index differs between the block's own index word and the candidate word.
constexpr int kSimBins = 17;
__device__ __forceinline__ int index_diff(uint32_t a, uint32_t b) {
  const uint32_t x = a ^ b;
  return __popc((x | (x >> 1)) & 0x55555555u);
}

// The candidate block of offset L (logit index) for block (bx, by) of frame d:
// inter (net 0) rows dy = L / 16 - 8, intra (net 1) rows by - L / 16; columns
// dx = L % 16 - 8.  L must be admitted by win_mask.
__device__ __forceinline__ int offset_block(const FrameDev& d, int net, int bx, int by, int L) {
  const int dx = (L & 15) - 8, r = L >> 4;
  const int y = net == 0 ? by + r - 8 : by - r;
  return d.base + y * d.bw + bx + dx;
}

// Rank key of an admitted offset's score: the fp32 score (a NaN counts as
// -FLT_MAX) mapped to an order-preserving unsigned integer, its low 8 bits
// replaced by 255 - L.  A larger key ranks first; scores equal to 24 bits rank
// by the lower index; 0 means no offset.  Branch-free.
__device__ __forceinline__ uint32_t rank_key(float v, int L) {
  const uint32_t u = __float_as_uint(fmaxf(v, -3.402823466e38f));
  const uint32_t m = (uint32_t)((int32_t)u >> 31);
  return ((u ^ (m | 0x80000000u)) & 0xFFFFFF00u) | (uint32_t)(255 - L);
}

// Bit L of the 256-bit set of admitted offsets, as words: bit b of word w is
// offset 32 w + b (rows 2w and 2w + 1 of the 16x16 grid).
__device__ __forceinline__ void admitted_words(const WinMask& m, uint32_t (&w)[8]) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int r0 = 2 * q, r1 = 2 * q + 1;
    const uint32_t c0 = ((m.rows >> r0) & 1u) ? (r0 == 0 ? m.cols0 : m.cols) : 0u;
    const uint32_t c1 = ((m.rows >> r1) & 1u) ? m.cols : 0u;
    w[q] = (c0 & 0xFFFFu) | ((c1 & 0xFFFFu) << 16);
  }
}

// top := the K largest of top and b, both sorted in descending order: the
// elementwise maximum of top and reversed b is bitonic and holds them; a
// bitonic merge sorts it.
template <int K>
__device__ __forceinline__ void merge_top(uint32_t (&top)[K], const uint32_t (&b)[K]) {
#pragma unroll
  for (int i = 0; i < K; ++i) top[i] = max(top[i], b[K - 1 - i]);
#pragma unroll
  for (int j = K >> 1; j >= 1; j >>= 1)
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int l = i ^ j;
      if (l > i) {
        const uint32_t x = top[i], y = top[l];
        top[i] = max(x, y);
        top[l] = min(x, y);
      }
    }
}

__device__ __forceinline__ int key_offset(uint32_t key) { return key ? 255 - (int)(key & 0xFFu) : -1; }

// Inserts a key into a list sorted in descending order: entry i becomes the
// median of t[i - 1], t[i] and the key (t[i - 1] >= t[i], so the key clamped
// into [t[i], t[i - 1]]), entry 0 the maximum of t[0] and the key.  Every
// entry depends only on the old list: K independent v_max_u32 / v_med3_u32,
// branch-free.
__device__ __forceinline__ uint32_t med3_u32(uint32_t a, uint32_t b, uint32_t c) {
  return max(min(a, b), min(max(a, b), c));  // selected as v_med3_u32
}
template <int K>
__device__ __forceinline__ void key_insert(uint32_t (&t)[K], uint32_t key) {
  uint32_t n[K];
  n[0] = max(t[0], key);
#pragma unroll
  for (int i = 1; i < K; ++i) n[i] = med3_u32(t[i - 1], t[i], key);
#pragma unroll
  for (int i = 0; i < K; ++i) t[i] = n[i];
}

// Layer 3.  NT = 9: writes the 262 outputs of each row to Y (row-major).
// With sim (kSimBins scores) the top-k adds, to each admitted offset's logit,
// sim[index_diff(own word, candidate word)] (words: the candidates' current
// index words, own: the blocks' fits).
// NT = 8 (the 256 logit columns only; the endpoint-seed outputs are not
// computed): for each block g of the batch (rows = total) that `need`
// admits (need null: every block of an inter frame), the top kk <= K
// admitted offsets by score go to prop[g * kk + i] (-1 past the last).
template <int K, int NT>
__global__ void __launch_bounds__(64 * kInferWaves) k_net_l3(const uint4* __restrict__ H2p, int rows,
                                                             const uint16_t* __restrict__ pk, float* __restrict__ Y,
                                                             const FrameDev* __restrict__ fr,
                                                             const uint8_t* __restrict__ sid, int net, int S, int kk,
                                                             const uint8_t* __restrict__ need,
                                                             int16_t* __restrict__ prop, const float* __restrict__ sim,
                                                             const uint32_t* __restrict__ words,
                                                             const uint64_t* __restrict__ own) {
  static_assert(NT == 8 || NT == 9, "8 logit tiles, or all 9 output tiles");
  __shared__ uint4 sW3[NT * 32 * kNetHid / 8];  // 128 or 144 KiB, rows of 32 chunks, swizzled
  __shared__ __attribute__((aligned(16))) float sb[NT * 32];
  __shared__ float s_sim[kSimBins];
  // per wave: the candidate words of a tile of 32 blocks of one frame row,
  // 16 window rows x (32 + 15) columns (similarity term, search mode only)
  constexpr int kWinCols = 32 + kNetWin - 1;
  __shared__ uint32_t s_win[NT == 8 ? kInferWaves : 1][NT == 8 ? kNetWin * kWinCols : 1];
  stage_rows(sW3, reinterpret_cast<const uint4*>(pk + kPkW3), NT * 32, kNetHid / 8, true);
  const float* bias = reinterpret_cast<const float*>(pk + kPkBias) + 2 * kNetHid;
  for (int i = threadIdx.x; i < NT * 32; i += blockDim.x) sb[i] = bias[i];
  if (threadIdx.x < kSimBins) s_sim[threadIdx.x] = sim ? sim[threadIdx.x] : 0.0f;
  __syncthreads();
  const int lane = threadIdx.x & 63, m = lane & 31, h = lane >> 5;
  const int n_tiles = (rows + 31) / 32;
  const int wave_g = blockIdx.x * kInferWaves + (threadIdx.x >> 6);
  for (int tile = wave_g; tile < n_tiles; tile += gridDim.x * kInferWaves) {
    const int row = tile * 32 + m;
    // accumulators start at the biases of the lane's outputs
    f32x16_t acc[NT];
#pragma unroll
    for (int nb = 0; nb < NT; ++nb) acc[nb] = bias_tile(sb, nb, h);
    const uint4* in = H2p + (size_t)tile * (kH2TileBytes / 16);
#pragma unroll 2
    for (int s = 0; s < 16; ++s) {
      const bf16x8_t b = as_bf16x8(in[s * 64 + lane]);
#pragma unroll
      for (int nb = 0; nb < NT; ++nb) {
        const int r = nb * 32 + m;
        const bf16x8_t a = as_bf16x8(sW3[r * 32 + ((2 * s + h) ^ (r & kSwz))]);
        acc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[nb], 0, 0, 0);
      }
    }
    if (Y) {  // launched with NT = 9 (a runtime test: the compiler then keeps the top-k path free of spills)
      if (row < rows) {
#pragma unroll
        for (int nb = 0; nb < NT; ++nb)
#pragma unroll
          for (int i = 0; i < 16; ++i) {
            const int u = nb * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
            if (u < kNetOut) Y[(size_t)row * kNetOut + u] = acc[nb][i];
          }
      }
      continue;
    }
    // fused top-k: each lane inserts its 128 logits' keys into a running top
    // K; lane h = 0 merges the partner lane's (l + 32) list and writes
    bool act = row < rows;
    uint32_t vw[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
    int fbx = 0, fby = 0;
    const FrameDev* dp = fr;
    if (act) {
      dp = &fr[sid[row]];
      const int k = row - dp->base;
      fbx = k % dp->bw;
      fby = k / dp->bw;
      act = need ? need[row] != 0 : !dp->is_intra;
      if (act) admitted_words(win_mask(net, S, fbx, fby, dp->bw, dp->bh), vw);
    }
    if (NT == 8 && sim) {
      // Index similarity: each offset's score gains sim[number of pixels
      // whose index differs between the own word and the candidate's].  The
      // tile's rows are taken in segments that lie in one row of one frame
      // (normally the whole tile): the segment's window words, 16 window rows
      // x 47 columns, are staged in the wave's LDS (coalesced loads, all
      // requested at once) and its lanes add the scores branch-free.  An
      // offset that is not admitted reads a staged 0 or a neighbour's word
      // and gains a score too, but its key is masked out of the ranking.
      const uint32_t ow = act ? (uint32_t)(own[row] >> 32) : 0u;
      uint32_t* wv = s_win[threadIdx.x >> 6];
      const int g_end = min(rows, tile * 32 + 32);
      for (int g_lo = tile * 32; g_lo < g_end;) {  // wave-uniform
        const FrameDev& d0 = fr[sid[g_lo]];
        const int k0 = g_lo - d0.base, by0 = k0 / d0.bw, bx_lo = k0 % d0.bw;
        const int seg_end = g_lo + min(g_end - g_lo, d0.bw - bx_lo);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // the last segment's reads are done
        __builtin_amdgcn_wave_barrier();
        constexpr int kStage = (kNetWin * kWinCols + 63) / 64;  // 12 words per lane
        uint32_t sw[kStage];
#pragma unroll
        for (int q = 0; q < kStage; ++q) {
          const int idx = lane + 64 * q;
          const int r = idx / kWinCols, col = idx - r * kWinCols;
          const int y = net == 0 ? by0 + r - 8 : by0 - r, x = bx_lo - 8 + col;
          sw[q] = (idx < kNetWin * kWinCols && y >= 0 && y < d0.bh && x >= 0 && x < d0.bw)
                      ? words[d0.base + y * d0.bw + x]
                      : 0u;
        }
#pragma unroll
        for (int q = 0; q < kStage; ++q)
          if (lane + 64 * q < kNetWin * kWinCols) wv[lane + 64 * q] = sw[q];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (act && row >= g_lo && row < seg_end) {
          const uint32_t* wl = wv + (fbx - bx_lo);
#pragma unroll
          for (int nb = 0; nb < 8; ++nb) {
            uint32_t wd[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) {
              const int L = nb * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
              wd[i] = wl[(L >> 4) * kWinCols + (L & 15)];
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[nb][i] += s_sim[index_diff(ow, wd[i])];
          }
        }
        g_lo = seg_end;
      }
    }
    uint32_t top[K];
#pragma unroll
    for (int i = 0; i < K; ++i) top[i] = 0u;
#pragma unroll
    for (int nb = 0; nb < 8; ++nb)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int bit = 8 * (i >> 2) + 4 * h + (i & 3);  // L = nb * 32 + bit
        const uint32_t ok = (uint32_t)((int32_t)(vw[nb] << (31 - bit)) >> 31);
        key_insert<K>(top, rank_key(acc[nb][i], nb * 32 + bit) & ok);
      }
    uint32_t partner[K];
#pragma unroll
    for (int i = 0; i < K; ++i) partner[i] = __shfl_xor(top[i], 32);
    merge_top<K>(top, partner);
    if (h == 0 && act) {
#pragma unroll
      for (int i = 0; i < K; ++i)
        if (i < kk) prop[(size_t)row * kk + i] = (int16_t)key_offset(top[i]);
    }
  }
}

// Top-k admitted offsets from fixed scores (one shared prior, ys_g = 0) or
// transposed logits (ys_l = rows, ys_g = 1), ranked as k_net_l3's fused top-k
// (with the same index-similarity term when sim is given).
__global__ void __launch_bounds__(256) k_top_scores(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid,
                                                    int total, int net, int S, int kk, const float* __restrict__ Ys,
                                                    long ys_l, long ys_g, const uint8_t* __restrict__ need,
                                                    int16_t* __restrict__ prop, const float* __restrict__ sim,
                                                    const uint32_t* __restrict__ words,
                                                    const uint64_t* __restrict__ own) {
  __shared__ float s_sim[kSimBins];
  if (threadIdx.x < kSimBins) s_sim[threadIdx.x] = sim ? sim[threadIdx.x] : 0.0f;
  __syncthreads();
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const FrameDev& d = fr[sid[g]];
  if (need ? !need[g] : d.is_intra) return;
  const int k = g - d.base, bx = k % d.bw, by = k / d.bw;
  const WinMask ok = win_mask(net, S, bx, by, d.bw, d.bh);
  uint32_t t[kMaxTopK];
#pragma unroll
  for (int i = 0; i < kMaxTopK; ++i) t[i] = 0u;
  const float* y = Ys + (long)g * ys_g;
  const uint32_t ow = sim ? (uint32_t)(own[g] >> 32) : 0u;
  for (int L = 0; L < kNetLogits; ++L)
    if (ok(L)) {
      float v = y[(long)L * ys_l];
      if (sim) v += s_sim[index_diff(ow, words[offset_block(d, net, bx, by, L)])];
      key_insert<kMaxTopK>(t, rank_key(v, L));
    }
#pragma unroll
  for (int i = 0; i < kMaxTopK; ++i)
    if (i < kk) prop[(size_t)g * kk + i] = (int16_t)key_offset(t[i]);
}

// Learned inter search: each proposed offset is evaluated exactly (the same
// re-fit and error as InterBlockSearch, codec/dxt_image.cpp:715-774); the
// valid one with the smallest error delta (first in rank order on ties) is
// reused if the delta passes the threshold (:890).  With need and cw given
// (the learned intra search follows), it also sets them for every block of
// the batch as k_intra_prep does, which then is not launched.
__global__ void __launch_bounds__(256) k_eval_inter(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid,
                                                    int total, int S, int thr, int kk,
                                                    const int16_t* __restrict__ prop,
                                                    const uint64_t* __restrict__ own, const uint32_t* __restrict__ rw,
                                                    const float4* __restrict__ rtab, uint64_t* __restrict__ dec_out,
                                                    uint8_t* __restrict__ need, uint32_t* __restrict__ cw,
                                                    DevStats* __restrict__ st) {
  __shared__ uint8_t s_snap[512];
  __shared__ float2 s_ab[4];
  __shared__ unsigned long long s_acc;
  load_lds_tables(s_snap, s_ab);
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n_cand = 0;
  if (need && g < total && fr[sid[g]].is_intra) {  // (as k_intra_prep) an intra frame's block searches
    need[g] = 1;
    cw[g] = (uint32_t)(own[g] >> 32);
  }
  if (g < total && !fr[sid[g]].is_intra) {
    const FrameDev& d = fr[sid[g]];
    const int k = g - d.base, bx = k % d.bw, by = k / d.bw, bw = d.bw;
    BlockPx px;
    load_block(d.rgb, d.width, bx, by, px);
    BlockEval be;
    make_block_eval(px, own[g], be);
    int best = 0x7fffffff, best_L = -1;
    uint32_t best_c = 0;
#pragma unroll
    for (int i = 0; i < kMaxTopK; ++i) {
      const int L = i < kk ? (int)prop[(size_t)g * kk + i] : -1;
      if (L >= 0) {
        const int q = d.base + (by + L / kNetWin - 8) * bw + (bx + L % kNetWin - 8);
        const float4 t4 = rtab[q];
        const uint32_t c = rw[q];
        const CandResult res = eval_candidate<false, true>(be, c, CandTab{t4.x, t4.y, t4.z, t4.w}, s_snap, s_ab);
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
    if (need) {  // (as k_intra_prep) a block without an inter reuse searches the intra window
      need[g] = dec_acc(dc) ? 0 : 1;
      cw[g] = dec_acc(dc) ? dec_word(dc) : (uint32_t)(own[g] >> 32);
    }
  }
  wg_count(&s_acc, n_cand, st->inter_cands);
}

// Index-similarity statistics of a recorded step (the exhaustive search of
// network `net`'s window): for every block that searched that window and
// every admitted offset, h = index_diff(own word, candidate word) is counted
// in cnt[h] and, when the offset passes the accept test (amask), also in
// cnt[kSimBins + h].  The candidates are the words the learned search sees
// when it proposes: inter, the previous frame's words (rw); intra, the words
// the intra search starts from (the accepted inter word where a block reuses
// one, else the block's own word).
__global__ void __launch_bounds__(256) k_sim_counts(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid,
                                                    int total, int net, int S, const uint8_t* __restrict__ need,
                                                    const uint16_t* __restrict__ amask,
                                                    const uint64_t* __restrict__ own, const uint32_t* __restrict__ rw,
                                                    const uint64_t* __restrict__ inter_dec,
                                                    unsigned long long* __restrict__ cnt) {
  __shared__ unsigned int s_c[2 * kSimBins];
  for (int i = threadIdx.x; i < 2 * kSimBins; i += blockDim.x) s_c[i] = 0u;
  __syncthreads();
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < total) {
    const FrameDev& d = fr[sid[g]];
    if (net == 0 ? !d.is_intra : need[g] != 0) {
      const int k = g - d.base, bx = k % d.bw, by = k / d.bw;
      const WinMask ok = win_mask(net, S, bx, by, d.bw, d.bh);
      const uint32_t ow = (uint32_t)(own[g] >> 32);
      for (int L = 0; L < kNetLogits; ++L) {
        if (!ok(L)) continue;
        const int q = offset_block(d, net, bx, by, L);
        const uint32_t c = net == 0 ? rw[q] : need[q] ? (uint32_t)(own[q] >> 32) : dec_word(inter_dec[q]);
        const int hd = index_diff(ow, c);
        atomicAdd(&s_c[hd], 1u);
        if ((amask[(size_t)g * 16 + (L >> 4)] >> (L & 15)) & 1u) atomicAdd(&s_c[kSimBins + hd], 1u);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 2 * kSimBins; i += blockDim.x)
    if (s_c[i]) atomicAdd(&cnt[i], (unsigned long long)s_c[i]);
}

}  // namespace mptc
