// libmptc_hip.so: HIP kernels for gfx950 and the C-ABI of include/mptc.h.
//
// Per frame (device-resident):
//   k_fit          stb DXT1 fit of every block (lane per block)
//   k_table        per-position candidate table of the reference frame's words
//   k_inter        InterBlockSearch + accept test, lane per block, exact scan order
//   k_intra_prep   which blocks need the intra search; initial current words
//   k_intra_dense  every intra candidate of every such block (lane per block);
//                  stores the error deltas E[pos][block] and folds the selection
//   k_intra_update words of blocks whose decision changed -> change list
//   k_intra_reeval re-evaluates the (block, position) pairs that see a changed word
//   k_intra_select re-folds the selection of blocks with a re-evaluated position
//   (update / reeval / select repeat until no word changes: the intra words are
//    the unique fixpoint of a raster-triangular system, which equals the
//    reference's sequential raster-order result)
//   k_commit       final blocks, motion bytes, logical endpoints, unique flags
//   k_unique       unique palette in raster order
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mptc.h"
#include "dxt_device.h"
#include "framegen.h"

using namespace mptc;

namespace {

constexpr int kWG = 256;          // threads per workgroup (4 wave64)
constexpr int kUniqueTile = 1024; // blocks per unique-compaction tile
constexpr int16_t kNoCand = 32767;// E sentinel: invalid or absent candidate
constexpr int kMaxS = 63;

// Per-block search decision, packed into 64 bits:
// [0,32) index word, [32,40) motion x, [40,48) motion y, bit 48 accepted.
__device__ __forceinline__ uint64_t pack_dec(uint32_t c, uint32_t x, uint32_t y, bool acc) {
  return (uint64_t)c | ((uint64_t)x << 32) | ((uint64_t)y << 40) | ((uint64_t)(acc ? 1 : 0) << 48);
}
__device__ __forceinline__ bool dec_acc(uint64_t d) { return (d >> 48) & 1; }
__device__ __forceinline__ uint32_t dec_word(uint64_t d) { return (uint32_t)d; }

struct DevStats {
  unsigned long long inter_cands, intra_cands, reeval_cands;
  unsigned int n_inter, n_intra, n_unique, n_changed;
};

// Adds a per-lane count to a global counter with one atomic per workgroup.
// Every thread of the workgroup must call it.
__device__ __forceinline__ void wg_count(unsigned long long* s_acc, unsigned long long v,
                                         unsigned long long* global) {
  if (threadIdx.x == 0) *s_acc = 0;
  __syncthreads();
  if (v) atomicAdd(s_acc, v);
  __syncthreads();
  if (threadIdx.x == 0 && *s_acc) atomicAdd(global, *s_acc);
}

__device__ __forceinline__ void load_snap_lds(uint8_t* s_snap) {
  for (int i = threadIdx.x; i < 512; i += blockDim.x) s_snap[i] = c_snap[i];
  __syncthreads();
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kWG) k_fit(const uint8_t* __restrict__ rgb, int width, int bw,
                                             int nb, uint64_t* __restrict__ own) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  BlockPx px;
  load_block(rgb, width, k % bw, k / bw, px);
  own[k] = fit_block(px);
}

// Candidate table of a word array (stride in 32-bit words: 2 reads the index
// word of a uint64_t physical block array).
__global__ void __launch_bounds__(kWG) k_table(const uint32_t* __restrict__ words, int stride, int nb,
                                               uint32_t* __restrict__ tw, float4* __restrict__ tab) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb) return;
  const uint32_t c = words[(size_t)q * stride];
  const CandTab t = cand_table(c);
  tw[q] = c;
  tab[q] = make_float4(t.asq, t.bsq, t.ab, t.f);
}

// InterBlockSearch (codec/dxt_image.cpp:715-774) + the accept test (:890).
// Rows j = by-S .. by+S-1 ascending, columns i = bx-S .. bx+S-1 ascending;
// the first error delta <= 0 clamps the running minimum to 0 and leaves the
// column loop only.
__global__ void __launch_bounds__(kWG) k_inter(const uint8_t* __restrict__ rgb, int width, int bw, int bh,
                                               int S, int thr, const uint64_t* __restrict__ own,
                                               const uint32_t* __restrict__ rw, const float4* __restrict__ rtab,
                                               uint64_t* __restrict__ dec_out, DevStats* __restrict__ st) {
  __shared__ uint8_t s_snap[512];
  __shared__ unsigned long long s_acc;
  load_snap_lds(s_snap);
  const int nb = bw * bh;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n_cand = 0;
  if (k < nb) {
    const int bx = k % bw, by = k / bw;
    BlockPx px;
    load_block(rgb, width, bx, by, px);
    BlockEval be;
    make_block_eval(px, own[k], be);
    int min_err = 0x7fffffff;
    uint32_t sel_c = 0, sel_x = 0, sel_y = 0;
    const int i_lo = max(bx - S, 0), i_hi = min(bx + S, bw);
    const int j_lo = max(by - S, 0), j_hi = min(by + S, bh);
    for (int j = j_lo; j < j_hi; ++j) {
      for (int i = i_lo; i < i_hi; ++i) {
        const int q = j * bw + i;
        const uint32_t c = rw[q];
        const float4 t4 = rtab[q];
        const CandResult r = eval_candidate<false>(be, c, CandTab{t4.x, t4.y, t4.z, t4.w}, s_snap);
        if (r.valid && r.err < min_err) {
          min_err = r.err;
          sel_c = c;
          sel_x = (uint32_t)(i - bx + S);
          sel_y = (uint32_t)(j - by + S);
          if (r.err <= 0) {
            min_err = 0;
            break;  // leaves the column loop only
          }
        }
      }
    }
    dec_out[k] = pack_dec(sel_c, sel_x | 0x80u, sel_y | 0x80u, min_err <= thr);
    n_cand = (unsigned long long)((i_hi - i_lo) * (j_hi - j_lo));
  }
  wg_count(&s_acc, n_cand, &st->inter_cands);
}

// Blocks that need the intra search, and the current words the search sees
// before any intra decision: inter-reused words are final (they read only the
// previous frame), every other block starts from its own word.
__global__ void __launch_bounds__(kWG) k_intra_prep(int nb, int is_intra, const uint64_t* __restrict__ own,
                                                    const uint64_t* __restrict__ inter_dec,
                                                    uint8_t* __restrict__ need, uint32_t* __restrict__ cw,
                                                    float4* __restrict__ ctab, uint8_t* __restrict__ dirty) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb) return;
  bool inter_ok = false;
  uint32_t w = (uint32_t)(own[q] >> 32);
  if (!is_intra) {
    const uint64_t d = inter_dec[q];
    if (dec_acc(d)) {
      inter_ok = true;
      w = dec_word(d);
    }
  }
  need[q] = inter_ok ? 0 : 1;
  dirty[q] = 0;
  cw[q] = w;
  const CandTab t = cand_table(w);
  ctab[q] = make_float4(t.asq, t.bsq, t.ab, t.f);
}

// Intra scan order (codec/dxt_image.cpp:672-675): rows j = by .. by-2S+1
// (r = by - j ascending), columns i = bx+S-1 .. bx-S (t = bx+S-1-i ascending),
// skipping (j == by, i >= bx).  Position pos = r*2S + t.
__device__ __forceinline__ uint64_t fold_intra(const int16_t* __restrict__ E, size_t nbs, int k, int S) {
  const int W2 = 2 * S;
  int min_err = 0x7fffffff, sel = -1;
  for (int r = 0; r < W2; ++r) {
    for (int t = (r == 0 ? S : 0); t < W2; ++t) {
      const int e = E[(size_t)(r * W2 + t) * nbs + k];
      if (e == kNoCand) continue;
      if (e < min_err) {
        min_err = e;
        sel = r * W2 + t;
        if (e <= 0) {
          min_err = 0;
          break;
        }
      }
    }
  }
  // the word is filled in by the caller from the current words
  return (uint64_t)(uint32_t)sel | ((uint64_t)(uint32_t)min_err << 32);
}

__device__ __forceinline__ uint64_t intra_decision(const uint32_t* __restrict__ cw, int bw, int k, int S,
                                                   int thr, uint64_t folded) {
  const int sel = (int)(uint32_t)folded;
  const int min_err = (int)(uint32_t)(folded >> 32);
  if (sel < 0 || min_err > thr) return 0;
  const int W2 = 2 * S;
  const int r = sel / W2, t = sel % W2;
  const int bx = k % bw, by = k / bw;
  const int j = by - r, i = bx + S - 1 - t;
  return pack_dec(cw[j * bw + i], (uint32_t)(i - bx + S), (uint32_t)(j - by + 2 * S - 1), true);
}

__global__ void __launch_bounds__(kWG) k_intra_dense(const uint8_t* __restrict__ rgb, int width, int bw, int bh,
                                                     int S, int thr, const uint64_t* __restrict__ own,
                                                     const uint8_t* __restrict__ need,
                                                     const uint32_t* __restrict__ cw,
                                                     const float4* __restrict__ ctab, int16_t* __restrict__ E,
                                                     size_t nbs, uint64_t* __restrict__ dec,
                                                     DevStats* __restrict__ st) {
  __shared__ uint8_t s_snap[512];
  __shared__ unsigned long long s_acc;
  load_snap_lds(s_snap);
  const int nb = bw * bh;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n_cand = 0;
  if (k < nb && need[k]) {
    const int bx = k % bw, by = k / bw;
    BlockPx px;
    load_block(rgb, width, bx, by, px);
    BlockEval be;
    make_block_eval(px, own[k], be);
    const int W2 = 2 * S;
    for (int r = 0; r < W2; ++r) {
      const int j = by - r;
      for (int t = 0; t < W2; ++t) {
        const int i = bx + S - 1 - t;
        int16_t ev = kNoCand;
        if (j >= 0 && i >= 0 && i < bw && !(r == 0 && t < S)) {
          const int q = j * bw + i;
          const float4 t4 = ctab[q];
          const CandResult res = eval_candidate<false>(be, cw[q], CandTab{t4.x, t4.y, t4.z, t4.w}, s_snap);
          ++n_cand;
          if (res.valid) ev = (int16_t)min(max(res.err, -32768), 32766);
        }
        E[(size_t)(r * W2 + t) * nbs + k] = ev;
      }
    }
    dec[k] = intra_decision(cw, bw, k, S, thr, fold_intra(E, nbs, k, S));
  }
  wg_count(&s_acc, n_cand, &st->intra_cands);
}

// New current words from the decisions; changed positions go to the list.
__global__ void __launch_bounds__(kWG) k_intra_update(int nb, const uint64_t* __restrict__ own,
                                                      const uint8_t* __restrict__ need,
                                                      const uint64_t* __restrict__ dec, uint32_t* __restrict__ cw,
                                                      float4* __restrict__ ctab, int* __restrict__ chg_list,
                                                      DevStats* __restrict__ st) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb || !need[q]) return;
  const uint64_t d = dec[q];
  const uint32_t w = dec_acc(d) ? dec_word(d) : (uint32_t)(own[q] >> 32);
  if (w != cw[q]) {
    cw[q] = w;
    const CandTab t = cand_table(w);
    ctab[q] = make_float4(t.asq, t.bsq, t.ab, t.f);
    const unsigned slot = atomicAdd(&st->n_changed, 1u);
    chg_list[slot] = q;
  }
}

// Every (block, position) pair whose intra window holds a changed position is
// re-evaluated with the new word.  Work item = (changed position, slot).
__global__ void __launch_bounds__(kWG) k_intra_reeval(const uint8_t* __restrict__ rgb, int width, int bw, int bh,
                                                      int S, const uint64_t* __restrict__ own,
                                                      const uint8_t* __restrict__ need,
                                                      const uint32_t* __restrict__ cw,
                                                      const float4* __restrict__ ctab,
                                                      const int* __restrict__ chg_list, int16_t* __restrict__ E,
                                                      size_t nbs, uint8_t* __restrict__ dirty,
                                                      DevStats* __restrict__ st) {
  __shared__ uint8_t s_snap[512];
  __shared__ unsigned long long s_acc;
  load_snap_lds(s_snap);
  const int W2 = 2 * S, slots = W2 * W2;
  const unsigned n_chg = st->n_changed;
  const size_t total = (size_t)n_chg * slots;
  unsigned long long n_eval = 0;
  for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (size_t)gridDim.x * blockDim.x) {
    const int q = chg_list[w / slots];
    const int s = (int)(w % slots);
    const int r = s / W2, t = s % W2;
    if (r == 0 && t < S) continue;
    const int qx = q % bw, qy = q / bw;
    const int by = qy + r, bx = qx - S + 1 + t;
    if (by >= bh || bx < 0 || bx >= bw) continue;
    const int k = by * bw + bx;
    if (!need[k]) continue;
    BlockPx px;
    load_block(rgb, width, bx, by, px);
    BlockEval be;
    make_block_eval(px, own[k], be);
    const float4 t4 = ctab[q];
    const CandResult res = eval_candidate<false>(be, cw[q], CandTab{t4.x, t4.y, t4.z, t4.w}, s_snap);
    E[(size_t)s * nbs + k] = res.valid ? (int16_t)min(max(res.err, -32768), 32766) : kNoCand;
    dirty[k] = 1;
    ++n_eval;
  }
  wg_count(&s_acc, n_eval, &st->reeval_cands);
}

__global__ void __launch_bounds__(kWG) k_intra_select(int bw, int bh, int S, int thr,
                                                      const uint32_t* __restrict__ cw,
                                                      const int16_t* __restrict__ E, size_t nbs,
                                                      uint8_t* __restrict__ dirty, uint64_t* __restrict__ dec) {
  const int nb = bw * bh;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb || !dirty[k]) return;
  dirty[k] = 0;
  dec[k] = intra_decision(cw, bw, k, S, thr, fold_intra(E, nbs, k, S));
}

__global__ void k_zero_u32(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Commit (codec/dxt_image.cpp:900-905, 922-926, 953-954).
__global__ void __launch_bounds__(kWG) k_commit(const uint8_t* __restrict__ rgb, int width, int bw, int nb,
                                                int is_intra, const uint64_t* __restrict__ own,
                                                const uint64_t* __restrict__ inter_dec,
                                                const uint8_t* __restrict__ need,
                                                const uint64_t* __restrict__ intra_dec,
                                                uint64_t* __restrict__ phys_out, uint8_t* __restrict__ motion,
                                                uint8_t* __restrict__ eps, uint8_t* __restrict__ kind,
                                                uint32_t* __restrict__ tile_count, DevStats* __restrict__ st) {
  __shared__ uint8_t s_snap[512];
  __shared__ unsigned long long s_acc;
  load_snap_lds(s_snap);
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  int kd = -1;
  if (k < nb) {
    const uint64_t o = own[k];
    uint64_t d = 0;
    kd = 0;
    if (!is_intra && dec_acc(inter_dec[k])) {
      d = inter_dec[k];
      kd = 2;
    } else if (need[k] && dec_acc(intra_dec[k])) {
      d = intra_dec[k];
      kd = 1;
    }
    uint64_t p = o;
    uint32_t e1, e2;
    {
      uint32_t r1, g1, b1, r2, g2, b2;
      decode565((uint32_t)(o & 0xFFFF), r1, g1, b1);
      decode565((uint32_t)((o >> 16) & 0xFFFF), r2, g2, b2);
      e1 = r1 | (g1 << 8) | (b1 << 16);
      e2 = r2 | (g2 << 8) | (b2 << 16);
    }
    uint8_t mx = 255, my = 255;
    if (kd) {
      const uint32_t c = dec_word(d);
      mx = (uint8_t)(d >> 32);
      my = (uint8_t)(d >> 40);
      if (c != (uint32_t)(o >> 32)) {
        // re-assigned indices: RecalculateEndpoints (codec/dxt_image.cpp:901-905, 923-926)
        BlockPx px;
        load_block(rgb, width, k % bw, k / bw, px);
        BlockEval be;
        make_block_eval(px, o, be);
        const CandResult r = eval_candidate<true>(be, c, cand_table(c), s_snap);
        e1 = r.e1;
        e2 = r.e2;
        p = (uint64_t)pack565(e1 & 0xFF, (e1 >> 8) & 0xFF, e1 >> 16) |
            ((uint64_t)pack565(e2 & 0xFF, (e2 >> 8) & 0xFF, e2 >> 16) << 16) | ((uint64_t)c << 32);
      }
    }
    phys_out[k] = p;
    motion[2 * k] = mx;
    motion[2 * k + 1] = my;
    if (eps) {
      eps[6 * k + 0] = (uint8_t)e1; eps[6 * k + 1] = (uint8_t)(e1 >> 8); eps[6 * k + 2] = (uint8_t)(e1 >> 16);
      eps[6 * k + 3] = (uint8_t)e2; eps[6 * k + 4] = (uint8_t)(e2 >> 8); eps[6 * k + 5] = (uint8_t)(e2 >> 16);
    }
    if (kind) kind[k] = (uint8_t)kd;
  }
  // one tile of kUniqueTile blocks spans kUniqueTile / kWG whole workgroups
  __shared__ unsigned s_cnt[3];
  if (threadIdx.x < 3) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  if (kd >= 0) atomicAdd(&s_cnt[kd], 1u);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_cnt[0]) atomicAdd(&tile_count[(blockIdx.x * blockDim.x) / kUniqueTile], s_cnt[0]);
    if (s_cnt[1]) atomicAdd(&st->n_intra, s_cnt[1]);
    if (s_cnt[2]) atomicAdd(&st->n_inter, s_cnt[2]);
  }
  (void)s_acc;
}

// Unique palette in raster order: workgroup per tile of kUniqueTile blocks.
__global__ void __launch_bounds__(kWG) k_unique(int nb, const uint64_t* __restrict__ own,
                                                const uint8_t* __restrict__ motion,
                                                const uint32_t* __restrict__ tile_count, int n_tiles,
                                                uint32_t* __restrict__ unique_out, uint32_t* __restrict__ n_unique,
                                                DevStats* __restrict__ st) {
  __shared__ uint32_t s_sum[kWG];
  __shared__ uint32_t s_wave[kWG / 64];
  const int tile = blockIdx.x;
  // prefix of the preceding tiles
  uint32_t acc = 0;
  for (int t = threadIdx.x; t < tile; t += blockDim.x) acc += tile_count[t];
  s_sum[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kWG / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s_sum[threadIdx.x] += s_sum[threadIdx.x + off];
    __syncthreads();
  }
  uint32_t base = s_sum[0];
  __syncthreads();
  // each thread handles kUniqueTile / kWG consecutive blocks
  constexpr int per = kUniqueTile / kWG;
  const int k0 = tile * kUniqueTile + threadIdx.x * per;
  uint32_t flags = 0, cnt = 0;
#pragma unroll
  for (int u = 0; u < per; ++u) {
    const int k = k0 + u;
    const bool uq = k < nb && motion[2 * k] == 255 && motion[2 * k + 1] == 255;
    flags |= (uq ? 1u : 0u) << u;
    cnt += uq ? 1u : 0u;
  }
  // exclusive scan of cnt across the workgroup
  s_sum[threadIdx.x] = cnt;
  __syncthreads();
  for (int off = 1; off < kWG; off <<= 1) {
    const uint32_t v = (int)threadIdx.x >= off ? s_sum[threadIdx.x - off] : 0u;
    __syncthreads();
    s_sum[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t pos = base + s_sum[threadIdx.x] - cnt;
#pragma unroll
  for (int u = 0; u < per; ++u)
    if (flags & (1u << u)) unique_out[pos++] = (uint32_t)(own[k0 + u] >> 32);
  (void)s_wave;
  if (tile == n_tiles - 1 && threadIdx.x == kWG - 1) {
    *n_unique = pos;
    st->n_unique = pos;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(MPTC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

// stb__PrepareOptTable (Include/stb_dxt.h:111-137)
void prepare_opt_table(uint8_t table[256][2], const uint8_t* expand, int size) {
  for (int i = 0; i < 256; ++i) {
    int best = 256;
    for (int mn = 0; mn < size; ++mn)
      for (int mx = 0; mx < size; ++mx) {
        const int mine = expand[mn], maxe = expand[mx];
        const int e = std::abs((2 * maxe + mine) / 3 - i) + std::abs(maxe - mine) * 3 / 100;
        if (e < best) {
          table[i][0] = (uint8_t)mx;
          table[i][1] = (uint8_t)mn;
          best = e;
        }
      }
  }
}

// ToFiveBits / ToSixBits (codec/dxt_image.cpp:72-121)
uint8_t snap_bits(uint8_t x, uint8_t keep, uint8_t half, int shift) {
  uint8_t base = (uint8_t)(x & keep);
  uint8_t high = (uint8_t)(base + half);
  uint8_t low = base == 0 ? base : (uint8_t)(base - half);
  base = (uint8_t)(base | (base >> shift));
  high = (uint8_t)(high | (high >> shift));
  low = (uint8_t)(low | (low >> shift));
  auto ad = [](uint8_t a, uint8_t b) { return (uint8_t)(a > b ? a - b : b - a); };
  const uint8_t db = ad(x, base), dh = ad(x, high), dl = ad(x, low);
  if (db <= dh) return db < dl ? base : low;
  return high;
}

int init_tables(int device) {
  static std::mutex mu;
  static std::vector<int> done;
  std::lock_guard<std::mutex> lk(mu);
  for (int d : done)
    if (d == device) return MPTC_OK;
  uint8_t e5[32], e6[64], om5[256][2], om6[256][2], snap[512];
  for (int i = 0; i < 32; ++i) e5[i] = (uint8_t)((i << 3) | (i >> 2));
  for (int i = 0; i < 64; ++i) e6[i] = (uint8_t)((i << 2) | (i >> 4));
  prepare_opt_table(om5, e5, 32);
  prepare_opt_table(om6, e6, 64);
  for (int x = 0; x < 256; ++x) {
    snap[x] = snap_bits((uint8_t)x, 0xF8, 4, 5);
    snap[256 + x] = snap_bits((uint8_t)x, 0xFC, 2, 6);
  }
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_omatch5), om5, sizeof(om5)));
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_omatch6), om6, sizeof(om6)));
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_snap), snap, sizeof(snap)));
  done.push_back(device);
  return MPTC_OK;
}

enum KernelId { KF_ZERO, KF_FIT, KF_TABLE, KF_INTER, KF_DENSE, KF_UPDATE, KF_REEVAL, KF_SELECT, KF_COMMIT, KF_UNIQUE, KF_N };
const char* kKernelNames[KF_N] = {"zero", "fit", "table", "inter", "intra_dense", "intra_update",
                                  "intra_reeval", "intra_select", "commit", "unique"};

}  // namespace

struct mptc_ctx {
  mptc_config cfg;
  int alloc_S = 0;  // search area the E buffer is sized for
  hipStream_t stream = nullptr;
  int max_nb = 0;
  size_t nbs = 0;  // E row stride (blocks)
  // device buffers
  uint8_t* d_rgb = nullptr;
  uint64_t *d_own = nullptr, *d_ref = nullptr, *d_phys = nullptr;
  uint64_t *d_inter = nullptr, *d_intra = nullptr;
  uint32_t *d_rw = nullptr, *d_cw = nullptr;
  float4 *d_rtab = nullptr, *d_ctab = nullptr;
  uint8_t *d_need = nullptr, *d_dirty = nullptr;
  int16_t* d_E = nullptr;
  int* d_chg = nullptr;
  uint32_t* d_tiles = nullptr;
  uint8_t *d_motion = nullptr, *d_eps = nullptr, *d_kind = nullptr;
  uint32_t *d_unique = nullptr, *d_nunique = nullptr;
  DevStats* d_stats = nullptr;
  DevStats* h_stats = nullptr;  // pinned
  int last_iters = 0, last_nb = 0, last_intra = 0;
  // profiling
  bool prof = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> spare;
  int64_t prof_n[KF_N] = {};
  double prof_ms[KF_N] = {};
};

namespace {

template <typename F>
int launch(mptc_ctx* c, hipStream_t s, int kid, F&& f) {
  hipEvent_t a = nullptr, b = nullptr;
  if (c->prof) {
    if (!c->spare.empty()) {
      a = c->spare.back().first;
      b = c->spare.back().second;
      c->spare.pop_back();
    } else {
      HIP_TRY(hipEventCreate(&a));
      HIP_TRY(hipEventCreate(&b));
    }
    HIP_TRY(hipEventRecord(a, s));
  }
  f();
  HIP_TRY(hipGetLastError());
  if (c->prof) {
    HIP_TRY(hipEventRecord(b, s));
    c->pending.push_back({kid, {a, b}});
  }
  return MPTC_OK;
}

int collect_profile(mptc_ctx* c) {
  for (auto& p : c->pending) {
    HIP_TRY(hipEventSynchronize(p.second.second));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, p.second.first, p.second.second));
    c->prof_n[p.first] += 1;
    c->prof_ms[p.first] += ms;
    c->spare.push_back(p.second);
  }
  c->pending.clear();
  return MPTC_OK;
}

int check_params(int S, int thr) {
  if (S < 1 || S > kMaxS) return fail(MPTC_EINVAL, "search_area must be in 1..63");
  if (thr < -32768 || thr > 30000) return fail(MPTC_EINVAL, "err_threshold must be in -32768..30000");
  return MPTC_OK;
}

template <typename T>
int dalloc(T** p, size_t n) {
  if (hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (n ? n : 1)) != hipSuccess)
    return fail(MPTC_ENOMEM, "hipMalloc failed");
  return MPTC_OK;
}

inline unsigned grid_for(int n) { return (unsigned)((n + kWG - 1) / kWG); }

// One frame, all device pointers.  own_in: fitted blocks to start from
// (nullptr -> fit on device from rgb).
int run_frame(mptc_ctx* c, hipStream_t s, const uint8_t* rgb, int w, int h, int is_intra,
              const uint64_t* ref, const uint64_t* own_in, uint64_t* phys_out, uint8_t* motion,
              uint32_t* unique, uint32_t* n_unique, uint8_t* eps, uint8_t* kind) {
  const int bw = w / 4, bh = h / 4, nb = bw * bh;
  const int S = c->cfg.search_area, thr = c->cfg.err_threshold;
  const size_t nbs = c->nbs;
  const int n_tiles = (nb + kUniqueTile - 1) / kUniqueTile;
  int rc;
  // zero counters
  if ((rc = launch(c, s, KF_ZERO, [&] {
         hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kWG), 0, s, reinterpret_cast<uint32_t*>(c->d_stats),
                            (int)(sizeof(DevStats) / 4));
       })))
    return rc;
  if ((rc = launch(c, s, KF_ZERO, [&] {
         hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(n_tiles)), dim3(kWG), 0, s, c->d_tiles, n_tiles);
       })))
    return rc;
  const uint64_t* own = own_in;
  if (!own) {
    if ((rc = launch(c, s, KF_FIT, [&] {
           hipLaunchKernelGGL(k_fit, dim3(grid_for(nb)), dim3(kWG), 0, s, rgb, w, bw, nb, c->d_own);
         })))
      return rc;
    own = c->d_own;
  }
  if (!is_intra) {
    if ((rc = launch(c, s, KF_TABLE, [&] {
           hipLaunchKernelGGL(k_table, dim3(grid_for(nb)), dim3(kWG), 0, s,
                              reinterpret_cast<const uint32_t*>(ref) + 1, 2, nb, c->d_rw, c->d_rtab);
         })))
      return rc;
    if ((rc = launch(c, s, KF_INTER, [&] {
           hipLaunchKernelGGL(k_inter, dim3(grid_for(nb)), dim3(kWG), 0, s, rgb, w, bw, bh, S, thr, own,
                              c->d_rw, c->d_rtab, c->d_inter, c->d_stats);
         })))
      return rc;
  }
  if ((rc = launch(c, s, KF_TABLE, [&] {
         hipLaunchKernelGGL(k_intra_prep, dim3(grid_for(nb)), dim3(kWG), 0, s, nb, is_intra, own, c->d_inter,
                            c->d_need, c->d_cw, c->d_ctab, c->d_dirty);
       })))
    return rc;
  if ((rc = launch(c, s, KF_DENSE, [&] {
         hipLaunchKernelGGL(k_intra_dense, dim3(grid_for(nb)), dim3(kWG), 0, s, rgb, w, bw, bh, S, thr, own,
                            c->d_need, c->d_cw, c->d_ctab, c->d_E, nbs, c->d_intra, c->d_stats);
       })))
    return rc;
  int iters = 0;
  for (;;) {
    // the system is raster-triangular: it reaches its fixpoint in at most nb + 1 passes
    if (++iters > nb + 2) return fail(MPTC_ESTATE, "intra fixpoint did not converge");
    if ((rc = launch(c, s, KF_UPDATE, [&] {
           hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, &c->d_stats->n_changed, 1);
           hipLaunchKernelGGL(k_intra_update, dim3(grid_for(nb)), dim3(kWG), 0, s, nb, own, c->d_need,
                              c->d_intra, c->d_cw, c->d_ctab, c->d_chg, c->d_stats);
         })))
      return rc;
    HIP_TRY(hipMemcpyAsync(c->h_stats, c->d_stats, sizeof(DevStats), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const unsigned n_chg = c->h_stats->n_changed;
    if (n_chg == 0) break;
    const size_t work = (size_t)n_chg * (size_t)(4 * S * S);
    const unsigned grid = (unsigned)std::min<size_t>((work + kWG - 1) / kWG, 8192);
    if ((rc = launch(c, s, KF_REEVAL, [&] {
           hipLaunchKernelGGL(k_intra_reeval, dim3(grid), dim3(kWG), 0, s, rgb, w, bw, bh, S, own, c->d_need,
                              c->d_cw, c->d_ctab, c->d_chg, c->d_E, nbs, c->d_dirty, c->d_stats);
         })))
      return rc;
    if ((rc = launch(c, s, KF_SELECT, [&] {
           hipLaunchKernelGGL(k_intra_select, dim3(grid_for(nb)), dim3(kWG), 0, s, bw, bh, S, thr, c->d_cw,
                              c->d_E, nbs, c->d_dirty, c->d_intra);
         })))
      return rc;
  }
  if ((rc = launch(c, s, KF_COMMIT, [&] {
         hipLaunchKernelGGL(k_commit, dim3(grid_for(nb)), dim3(kWG), 0, s, rgb, w, bw, nb, is_intra, own,
                            c->d_inter, c->d_need, c->d_intra, phys_out, motion, eps, kind, c->d_tiles,
                            c->d_stats);
       })))
    return rc;
  if ((rc = launch(c, s, KF_UNIQUE, [&] {
         hipLaunchKernelGGL(k_unique, dim3(n_tiles), dim3(kWG), 0, s, nb, own, motion, c->d_tiles, n_tiles,
                            unique, n_unique, c->d_stats);
       })))
    return rc;
  c->last_iters = iters;
  c->last_nb = nb;
  c->last_intra = is_intra;
  return MPTC_OK;
}

int check_dims(mptc_ctx* c, int w, int h) {
  if (w <= 0 || h <= 0 || (w & 3) || (h & 3)) return fail(MPTC_EINVAL, "width and height must be positive multiples of 4");
  if ((w / 4) * (h / 4) > c->max_nb) return fail(MPTC_EINVAL, "frame larger than the context's maximum");
  return MPTC_OK;
}

}  // namespace

extern "C" {

const char* mptc_last_error(void) { return g_err.c_str(); }

int32_t mptc_create(const mptc_config* cfg, mptc_ctx** out) {
  if (!cfg || !out) return fail(MPTC_EINVAL, "null argument");
  *out = nullptr;
  if (cfg->max_width <= 0 || cfg->max_height <= 0 || (cfg->max_width & 3) || (cfg->max_height & 3))
    return fail(MPTC_EINVAL, "max_width/max_height must be positive multiples of 4");
  int rc = check_params(cfg->search_area, cfg->err_threshold);
  if (rc) return rc;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (cfg->device < 0 || cfg->device >= ndev) return fail(MPTC_EINVAL, "no such HIP device");
  HIP_TRY(hipSetDevice(cfg->device));
  if ((rc = init_tables(cfg->device))) return rc;
  mptc_ctx* c = new mptc_ctx();
  c->cfg = *cfg;
  c->alloc_S = cfg->search_area;
  const int nb = (cfg->max_width / 4) * (cfg->max_height / 4);
  c->max_nb = nb;
  c->nbs = (size_t)((nb + 63) / 64) * 64;
  auto cleanup = [&](int code) {
    mptc_destroy(c);
    return code;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(MPTC_EHIP, "hipStreamCreate failed"));
  const int n_tiles = (nb + kUniqueTile - 1) / kUniqueTile;
  const size_t e_slots = 4 * (size_t)cfg->search_area * cfg->search_area;
  if ((rc = dalloc(&c->d_rgb, (size_t)nb * 48)) || (rc = dalloc(&c->d_own, nb)) || (rc = dalloc(&c->d_ref, nb)) ||
      (rc = dalloc(&c->d_phys, nb)) || (rc = dalloc(&c->d_inter, nb)) || (rc = dalloc(&c->d_intra, nb)) ||
      (rc = dalloc(&c->d_rw, nb)) || (rc = dalloc(&c->d_cw, nb)) || (rc = dalloc(&c->d_rtab, nb)) ||
      (rc = dalloc(&c->d_ctab, nb)) || (rc = dalloc(&c->d_need, nb)) || (rc = dalloc(&c->d_dirty, nb)) ||
      (rc = dalloc(&c->d_E, e_slots * c->nbs)) || (rc = dalloc(&c->d_chg, nb)) ||
      (rc = dalloc(&c->d_tiles, n_tiles)) || (rc = dalloc(&c->d_motion, (size_t)nb * 2)) ||
      (rc = dalloc(&c->d_eps, (size_t)nb * 6)) || (rc = dalloc(&c->d_kind, nb)) ||
      (rc = dalloc(&c->d_unique, nb)) || (rc = dalloc(&c->d_nunique, 1)) || (rc = dalloc(&c->d_stats, 1)))
    return cleanup(rc);
  if (hipHostMalloc(reinterpret_cast<void**>(&c->h_stats), sizeof(DevStats), hipHostMallocDefault) != hipSuccess)
    return cleanup(fail(MPTC_ENOMEM, "hipHostMalloc failed"));
  std::memset(c->h_stats, 0, sizeof(DevStats));
  *out = c;
  return MPTC_OK;
}

void mptc_destroy(mptc_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->cfg.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& p : c->pending) { (void)hipEventDestroy(p.second.first); (void)hipEventDestroy(p.second.second); }
  for (auto& p : c->spare) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
  void* bufs[] = {c->d_rgb, c->d_own, c->d_ref, c->d_phys, c->d_inter, c->d_intra, c->d_rw, c->d_cw,
                  c->d_rtab, c->d_ctab, c->d_need, c->d_dirty, c->d_E, c->d_chg, c->d_tiles, c->d_motion,
                  c->d_eps, c->d_kind, c->d_unique, c->d_nunique, c->d_stats};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (c->h_stats) (void)hipHostFree(c->h_stats);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int32_t mptc_set_params(mptc_ctx* c, int32_t S, int32_t thr) {
  if (!c) return fail(MPTC_EINVAL, "null context");
  int rc = check_params(S, thr);
  if (rc) return rc;
  if (S > c->alloc_S) return fail(MPTC_EINVAL, "search_area above the one the context was created with");
  c->cfg.search_area = S;
  c->cfg.err_threshold = thr;
  return MPTC_OK;
}

int32_t mptc_fit(mptc_ctx* c, const uint8_t* rgb, int32_t w, int32_t h, uint64_t* phys_out) {
  if (!c || !rgb || !phys_out) return fail(MPTC_EINVAL, "null argument");
  int rc = check_dims(c, w, h);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(c->cfg.device));
  const int bw = w / 4, nb = bw * (h / 4);
  hipStream_t s = c->stream;
  HIP_TRY(hipMemcpyAsync(c->d_rgb, rgb, (size_t)w * h * 3, hipMemcpyHostToDevice, s));
  if ((rc = launch(c, s, KF_FIT, [&] {
         hipLaunchKernelGGL(k_fit, dim3(grid_for(nb)), dim3(kWG), 0, s, c->d_rgb, w, bw, nb, c->d_own);
       })))
    return rc;
  HIP_TRY(hipMemcpyAsync(phys_out, c->d_own, (size_t)nb * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return MPTC_OK;
}

int32_t mptc_reencode(mptc_ctx* c, const uint8_t* rgb, int32_t w, int32_t h, int32_t is_intra,
                      const uint64_t* ref_phys, uint64_t* phys_inout, uint8_t* motion_out,
                      uint32_t* unique_out, uint32_t* n_unique_out, uint8_t* eps_out, uint8_t* kind_out) {
  if (!c || !rgb || !phys_inout || !motion_out || !unique_out || !n_unique_out)
    return fail(MPTC_EINVAL, "null argument");
  if (!is_intra && !ref_phys) return fail(MPTC_EINVAL, "inter frame needs ref_phys");
  int rc = check_dims(c, w, h);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(c->cfg.device));
  const int nb = (w / 4) * (h / 4);
  hipStream_t s = c->stream;
  HIP_TRY(hipMemcpyAsync(c->d_rgb, rgb, (size_t)w * h * 3, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(c->d_own, phys_inout, (size_t)nb * 8, hipMemcpyHostToDevice, s));
  if (!is_intra) HIP_TRY(hipMemcpyAsync(c->d_ref, ref_phys, (size_t)nb * 8, hipMemcpyHostToDevice, s));
  if ((rc = run_frame(c, s, c->d_rgb, w, h, is_intra, c->d_ref, c->d_own, c->d_phys, c->d_motion, c->d_unique,
                      c->d_nunique, eps_out ? c->d_eps : nullptr, kind_out ? c->d_kind : nullptr)))
    return rc;
  uint32_t n_unique = 0;
  HIP_TRY(hipMemcpyAsync(&n_unique, c->d_nunique, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(phys_inout, c->d_phys, (size_t)nb * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(motion_out, c->d_motion, (size_t)nb * 2, hipMemcpyDeviceToHost, s));
  if (eps_out) HIP_TRY(hipMemcpyAsync(eps_out, c->d_eps, (size_t)nb * 6, hipMemcpyDeviceToHost, s));
  if (kind_out) HIP_TRY(hipMemcpyAsync(kind_out, c->d_kind, (size_t)nb, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(hipMemcpy(unique_out, c->d_unique, (size_t)n_unique * 4, hipMemcpyDeviceToHost));
  *n_unique_out = n_unique;
  return MPTC_OK;
}

int32_t mptc_encode_frame_device(mptc_ctx* c, const uint8_t* d_rgb, int32_t w, int32_t h, int32_t is_intra,
                                 const uint64_t* d_ref_phys, uint64_t* d_phys_out, uint8_t* d_motion_out,
                                 uint32_t* d_unique_out, uint32_t* d_n_unique_out, uint8_t* d_eps_out,
                                 uint8_t* d_kind_out, void* hip_stream) {
  if (!c || !d_rgb || !d_phys_out || !d_motion_out || !d_unique_out || !d_n_unique_out)
    return fail(MPTC_EINVAL, "null argument");
  if (!is_intra && !d_ref_phys) return fail(MPTC_EINVAL, "inter frame needs d_ref_phys");
  int rc = check_dims(c, w, h);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(c->cfg.device));
  hipStream_t s = hip_stream ? static_cast<hipStream_t>(hip_stream) : c->stream;
  return run_frame(c, s, d_rgb, w, h, is_intra, d_ref_phys, nullptr, d_phys_out, d_motion_out, d_unique_out,
                   d_n_unique_out, d_eps_out, d_kind_out);
}

int32_t mptc_get_stats(mptc_ctx* c, mptc_frame_stats* out) {
  if (!c || !out) return fail(MPTC_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(c->cfg.device));
  HIP_TRY(hipDeviceSynchronize());
  DevStats d;
  HIP_TRY(hipMemcpy(&d, c->d_stats, sizeof(d), hipMemcpyDeviceToHost));
  out->n_blocks = c->last_nb;
  out->n_inter = d.n_inter;
  out->n_intra = d.n_intra;
  out->n_unique = d.n_unique;
  out->inter_candidates = (int64_t)d.inter_cands;
  out->intra_candidates = (int64_t)d.intra_cands;
  out->reeval_candidates = (int64_t)d.reeval_cands;
  out->intra_iterations = c->last_iters;
  out->is_intra = c->last_intra;
  return MPTC_OK;
}

int32_t mptc_profile_enable(mptc_ctx* c, int32_t enable) {
  if (!c) return fail(MPTC_EINVAL, "null context");
  HIP_TRY(hipSetDevice(c->cfg.device));
  int rc = collect_profile(c);
  if (rc) return rc;
  for (int i = 0; i < KF_N; ++i) {
    c->prof_n[i] = 0;
    c->prof_ms[i] = 0.0;
  }
  c->prof = enable != 0;
  return MPTC_OK;
}

int32_t mptc_profile_read(mptc_ctx* c, const char* name, int64_t* launches, double* total_ms) {
  if (!c || !name || !launches || !total_ms) return fail(MPTC_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(c->cfg.device));
  int rc = collect_profile(c);
  if (rc) return rc;
  for (int i = 0; i < KF_N; ++i)
    if (std::strcmp(name, kKernelNames[i]) == 0) {
      *launches = c->prof_n[i];
      *total_ms = c->prof_ms[i];
      return MPTC_OK;
    }
  return fail(MPTC_EINVAL, std::string("unknown kernel name: ") + name);
}

int32_t mptc_synth_frames(const void* cfg, int32_t first, int32_t n, uint8_t* rgb_out) {
  if (!cfg || !rgb_out || n < 0) return fail(MPTC_EINVAL, "bad argument");
  const mptc_synth_cfg* sc = static_cast<const mptc_synth_cfg*>(cfg);
  if (sc->width <= 0 || sc->height <= 0) return fail(MPTC_EINVAL, "bad frame size");
  const size_t fb = (size_t)sc->width * sc->height * 3;
  for (int f = 0; f < n; ++f) mptc_synth_frame(sc, first + f, rgb_out + fb * f);
  return MPTC_OK;
}

}  // extern "C"
