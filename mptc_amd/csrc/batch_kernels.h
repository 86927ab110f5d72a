// Kernels of the per-frame encode step, over a batch of frames described in
// device memory (FrameDev).  Every per-block buffer is the concatenation of
// the frames' blocks: global block id g = frame base + by*bw + bx.
//
//   k_fill_sid     block -> frame index
//   k_fit          stb DXT1 fit of every block (lane per block); in the step
//                  k_block_prep (mlp_kernels.h) does it together with the
//                  candidate tables of the reference frames' index words
//   k_inter        InterBlockSearch + accept test (lane per block, wave tile)
//   k_intra_prep   which blocks need the intra search; initial current words
//                  (pruned search: their bound tables, empty passing sets)
//   k_intra_dense  every intra candidate of every such block: stores the error
//                  deltas E, folds the selection, lists the words that change
//   k_intra_apply  the listed words become current (first pass only)
//   fixpoint passes, each proportional to the work left:
//     k_intra_reeval    (block, position) pairs that see a changed word
//     k_intra_select    re-folds those blocks; changed words become current
//     or k_intra_tile_pass, which also settles the reuse chains inside each
//                       64-block tile (tuning itile_w)
//   The intra words are the unique fixpoint of a raster-triangular system
//   (each block's window holds only earlier blocks); that fixpoint is the
//   reference's sequential raster-order result.  Passes stop when no word
//   changes.
//   k_commit       blocks, motion bytes, logical endpoints, unique counts
//   k_unique       unique palette of every frame in raster order
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dxt_device.h"

namespace mptc {

constexpr int kUniqueTile = 1024;   // blocks per unique-compaction tile
constexpr int16_t kNoCand = 32767;  // E sentinel: invalid or absent candidate
constexpr int kMaxFrames = 64;      // frames per batch

struct FrameDev {
  const uint8_t* rgb;      // width*height*3 bytes
  const uint64_t* ref;     // previous frame's blocks (inter frames)
  uint64_t* phys_out;
  uint8_t* motion;
  uint32_t* unique;
  uint32_t* n_unique;
  uint8_t* eps;            // optional
  uint8_t* kind;           // optional
  int32_t width, bw, bh, nb;
  int32_t base;            // first global block id
  int32_t is_intra;
  int32_t tile_base;       // first unique-compaction tile
  int32_t n_tiles;
  int32_t wtile_base;      // first wave tile of the search kernels
  int32_t tiles_x;         // wave tiles per row of wave tiles
  int32_t itile_base;      // first tile of the intra fixpoint passes (64 blocks each)
  int32_t itiles_x;        // those tiles per row
};

// Statistics counters are spread over kStatSlots copies (slot = workgroup
// index mod kStatSlots), each on a 128-byte line of its own, so that
// thousands of workgroups do not queue on one address or line; the host sums
// the slots.  Slot i of a 64-bit counter array is element i * kSlot64, of a
// 32-bit one element i * kSlot32.
constexpr int kStatSlots = 32;
constexpr int kSlot64 = 16;
constexpr int kSlot32 = 32;

// pass_end counts finished workgroups in two levels: groups of kPassGroup
// workgroups count into their own counter (one 256-byte line each), and the
// last workgroup of each group into pass_done.
constexpr int kPassGroup = 64;
constexpr int kPassGroups = 64;  // grids of up to 4,096 workgroups; larger ones count flat
constexpr int kPassStride = 64;  // words between the group counters

struct DevStats {
  unsigned long long inter_cands[kStatSlots * kSlot64], intra_cands[kStatSlots * kSlot64];
  unsigned long long reeval_cands[kStatSlots * kSlot64];
  unsigned long long exact_inter[kStatSlots * kSlot64], exact_intra[kStatSlots * kSlot64];  // pruned search: exact evaluations
  unsigned int n_inter[kStatSlots * kSlot32], n_intra[kStatSlots * kSlot32];
  unsigned int n_unique, active_passes;  // passes that changed a word
  unsigned int n_changed[2];             // length of the changed-word list, by pass parity
  unsigned int pass_done;                // workgroups of the running fixpoint pass that finished
  unsigned int skip_tile;                // the last pass was tile-resolved (k_intra_tile_pass)
  unsigned int pass_grp[kPassGroups * kPassStride];  // pass_end's group counters
};

// The pruning bound's table of an index word (prune_kernels.h).
// Per index word: byte masks of groups 0..2 over the four 4-pixel chunks and
// 1 / (pixels per group) for groups 0..3 (0 for an empty group).
struct BoundTab {
  uint4 m0, m1, m2;
  float4 inv;
};

__device__ __forceinline__ BoundTab bound_table(uint32_t c) {
  uint32_t m[3][4];
  uint32_t n[3] = {0u, 0u, 0u};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t x = spread_indices(c, j);
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      m[g][j] = byte_gather(1u << (8 * g), x);  // 1 in byte n where pixel 4j+n has index g
      n[g] = __builtin_amdgcn_udot4(m[g][j], 0x01010101u, n[g], false);
    }
  }
  const uint32_t n3 = 16u - n[0] - n[1] - n[2];
  BoundTab t;
  t.m0 = make_uint4(m[0][0], m[0][1], m[0][2], m[0][3]);
  t.m1 = make_uint4(m[1][0], m[1][1], m[1][2], m[1][3]);
  t.m2 = make_uint4(m[2][0], m[2][1], m[2][2], m[2][3]);
  t.inv = make_float4(n[0] ? 1.0f / (float)n[0] : 0.0f, n[1] ? 1.0f / (float)n[1] : 0.0f,
                      n[2] ? 1.0f / (float)n[2] : 0.0f, n3 ? 1.0f / (float)n3 : 0.0f);
  return t;
}

// Per-block search decision, packed into 64 bits:
// [0,32) index word, [32,40) motion x, [40,48) motion y, bit 48 accepted.
__device__ __forceinline__ uint64_t pack_dec(uint32_t c, uint32_t x, uint32_t y, bool acc) {
  return (uint64_t)c | ((uint64_t)x << 32) | ((uint64_t)y << 40) | ((uint64_t)(acc ? 1 : 0) << 48);
}
__device__ __forceinline__ bool dec_acc(uint64_t d) { return (d >> 48) & 1; }
__device__ __forceinline__ uint32_t dec_word(uint64_t d) { return (uint32_t)d; }

// Error deltas as stored: saturated to int16, kNoCand reserved.  Values above
// 30000 cannot pass any threshold and only the sign of a negative delta is
// ever used, so saturation changes no decision.
__device__ __forceinline__ int16_t sat16(int e) { return (int16_t)min(max(e, -32768), 32766); }

// Adds a per-lane count to a statistics counter with one atomic per
// workgroup.  Every thread of the workgroup must call it.
__device__ __forceinline__ void wg_count(unsigned long long* s_acc, unsigned long long v,
                                         unsigned long long* slots) {
  if (threadIdx.x == 0) *s_acc = 0;
  __syncthreads();
  if (v) atomicAdd(s_acc, v);
  __syncthreads();
  if (threadIdx.x == 0 && *s_acc) atomicAdd(&slots[(blockIdx.x % kStatSlots) * kSlot64], *s_acc);
}

// End of a fixpoint pass's workgroup (thread 0 acts; call it from every
// workgroup exactly once, after the workgroup's counter updates).  The last
// workgroup to finish knows that every workgroup has read its inputs and added
// its changes: it clears `clear` (the counter the next pass adds to, or null),
// rearms pass_done and writes the pass's count n_changed[par] to `conv`
// (pinned host memory the host polls for convergence; null: not written).
// The workgroups' counter updates are atomics whose results they consume
// before they get here, so they are complete before their count (no
// device-scope fence, which would write back the XCD's L2).  The count is
// two-level (kPassGroup workgroups per group counter, then the groups), so
// that no one address takes every workgroup's atomic; a group's last
// workgroup has its group counter's result before it counts the group.
__device__ __forceinline__ void pass_end(DevStats* st, int par, unsigned int* clear, unsigned int* conv) {
  if (threadIdx.x != 0) return;
  const unsigned n_wg = gridDim.x, ng = (n_wg + kPassGroup - 1) / kPassGroup;
  if (ng > 1 && ng <= (unsigned)kPassGroups) {
    const unsigned grp = blockIdx.x / kPassGroup;
    const unsigned members = min((unsigned)kPassGroup, n_wg - grp * kPassGroup);
    unsigned int* gc = &st->pass_grp[grp * kPassStride];
    if (atomicAdd(gc, 1u) + 1 != members) return;
    atomicExch(gc, 0u);  // rearmed for the next pass
    if (atomicAdd(&st->pass_done, 1u) + 1 != ng) return;
  } else if (atomicAdd(&st->pass_done, 1u) + 1 != n_wg) {
    return;
  }
  const unsigned n = atomicAdd(&st->n_changed[par], 0u);
  if (clear) atomicExch(clear, 0u);
  atomicExch(&st->pass_done, 0u);
  if (conv) {
    *reinterpret_cast<volatile unsigned int*>(conv) = n;
    __threadfence_system();
  }
}

// An empty pass (the previous one changed no word) that reports to the host
// (conv non-null: the last pass of a batch) writes the count 0.
__device__ __forceinline__ void report_fixpoint(unsigned int* conv) {
  if (conv && blockIdx.x == 0 && threadIdx.x == 0) {
    *reinterpret_cast<volatile unsigned int*>(conv) = 0u;
    __threadfence_system();
  }
}

// Appends v (where want) to a global list with one atomic per workgroup.
// Every thread of the workgroup must call it; s_base is a __shared__ word.
__device__ __forceinline__ void wg_append(bool want, int v, int* list, unsigned int* count, unsigned int* s_base) {
  if (threadIdx.x == 0) *s_base = 0;
  __syncthreads();
  const unsigned long long m = __ballot(want);
  const int lane = (int)__lane_id();
  unsigned int off = 0;
  if (m && lane == 0) off = atomicAdd(s_base, (unsigned int)__popcll(m));  // offset of this wave in the workgroup
  off = __shfl(off, 0);
  __syncthreads();
  if (threadIdx.x == 0 && *s_base) *s_base = atomicAdd(count, *s_base);  // workgroup range in the list
  __syncthreads();
  if (want) list[*s_base + off + __popcll(m & ((1ull << lane) - 1))] = v;
}

// Per-workgroup LDS tables: the 5/6-bit snap table and the LSQ weights.
__device__ __forceinline__ void load_lds_tables(uint8_t* s_snap, float2* s_ab) {
  for (int i = threadIdx.x; i < 512; i += blockDim.x) s_snap[i] = c_snap[i];
  init_ab_table(s_ab);
  __syncthreads();
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  This renumbering gives each XCD a contiguous range of workgroup ids (a
// bijection on 0..gridDim.x-1: XCD x takes q + (x < r) ids from base
// x*q + min(x, r), q = n / 8, r = n % 8), so that with raster-ordered tiles
// an XCD works on one band of the frame and the window data it reads stays in
// its L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block_id() {
  const int n = (int)gridDim.x, i = (int)blockIdx.x;
  const int x = i % kXcds, s = i / kXcds, q = n / kXcds, r = n % kXcds;
  return x * q + min(x, r) + s;
}

// Work-item mapping of the search kernels: wave w covers chunk (w % n_chunks)
// of the window of wave tile (w / n_chunks); a wave tile is tw x (64/tw)
// blocks of one frame.  Workgroup ids are XCD-banded (xcd_block_id).
// Returns false for lanes beyond the frame edge.
__device__ __forceinline__ bool tile_block(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ tsid,
                                           int n_wtiles, int tw, int n_chunks, int& f, int& bx, int& by,
                                           int& chunk) {
  const int wi = xcd_block_id() * blockDim.x + threadIdx.x;
  // a wave is one chunk of one tile of one frame: keep the frame index (and so
  // the frame's fields) in scalar registers
  const int w = __builtin_amdgcn_readfirstlane(wi >> 6), l = wi & 63;
  const int t = w / n_chunks;
  chunk = w - t * n_chunks;
  if (t >= n_wtiles) return false;
  f = __builtin_amdgcn_readfirstlane(tsid[t]);
  const int lt = t - fr[f].wtile_base;
  const int tx = lt % fr[f].tiles_x, ty = lt / fr[f].tiles_x;
  bx = tx * tw + (l & (tw - 1));  // tw is a power of two
  by = ty * (64 / tw) + (l >> __builtin_ctz((unsigned)tw));
  return bx < fr[f].bw && by < fr[f].bh;
}

// The selection rule (codec/dxt_image.cpp:697-708) over a run of whole window
// rows, summarised so that consecutive runs combine in scan order without
// knowing the state the earlier rows leave:
//   first_le0  first position with delta <= 0 (it becomes the selection, the
//              running minimum clamps to 0 and its row ends);
//   after_sel  for the rows after first_le0's row: the first negative delta of
//              the last row that has one (each such row selects it and ends);
//   last_neg   the same over every row of the run (used when an earlier run
//              already clamped the minimum to 0);
//   min_val/min_pos  earliest strict minimum (used when no delta <= 0 exists).
struct ChunkFold {
  int first_le0 = -1, le0_row = -1, after_sel = -1, last_neg = -1;
  int min_val = 0x7fffffff, min_pos = -1;
  int row_neg = -1;
  __device__ __forceinline__ void begin_row() { row_neg = -1; }
  __device__ __forceinline__ void step(int e, int pos, int row) {
    if (e == kNoCand) return;
    if (e < min_val) {
      min_val = e;
      min_pos = pos;
    }
    if (e <= 0 && first_le0 < 0) {
      first_le0 = pos;
      le0_row = row;
    }
    if (e < 0 && row_neg < 0) row_neg = pos;
  }
  __device__ __forceinline__ void end_row(int row) {
    if (row_neg >= 0) {
      last_neg = row_neg;
      if (first_le0 >= 0 && row > le0_row) after_sel = row_neg;
    }
  }
  __device__ __forceinline__ uint4 pack() const {
    return make_uint4(((uint32_t)first_le0 & 0xFFFFu) | ((uint32_t)after_sel << 16),
                      ((uint32_t)last_neg & 0xFFFFu) | ((uint32_t)min_pos << 16), (uint32_t)min_val, 0u);
  }
};

__device__ __forceinline__ int lo16(uint32_t v) { return (v & 0xFFFFu) == 0xFFFFu ? -1 : (int)(v & 0xFFFFu); }
__device__ __forceinline__ int hi16(uint32_t v) { return (v >> 16) == 0xFFFFu ? -1 : (int)(v >> 16); }

// Folds the chunk summaries of block g in scan order: (selected position or
// -1, returned error delta).
__device__ __forceinline__ void combine_chunks(const uint4* __restrict__ summ, size_t stride, int g, int n_chunks,
                                               int& sel, int& ret) {
  int m = 0x7fffffff;
  sel = -1;
  for (int k = 0; k < n_chunks; ++k) {
    const uint4 c = summ[(size_t)k * stride + g];
    const int first_le0 = lo16(c.x), after_sel = hi16(c.x), last_neg = lo16(c.y), min_pos = hi16(c.y);
    const int min_val = (int)c.z;
    if (m == 0) {
      if (last_neg >= 0) sel = last_neg;
    } else if (first_le0 >= 0) {
      m = 0;
      sel = after_sel >= 0 ? after_sel : first_le0;
    } else if (min_pos >= 0 && min_val < m) {
      m = min_val;
      sel = min_pos;
    }
  }
  ret = m;
}

__global__ void __launch_bounds__(256) k_fill_sid(const FrameDev* __restrict__ fr, int n_frames, uint8_t* __restrict__ sid) {
  const int f = blockIdx.y;
  if (f >= n_frames) return;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < fr[f].nb; k += gridDim.x * blockDim.x)
    sid[fr[f].base + k] = (uint8_t)f;
}

__global__ void __launch_bounds__(256) k_fit(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid, int total,
                      uint64_t* __restrict__ own) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const FrameDev& d = fr[sid[g]];
  const int k = g - d.base;
  BlockPx px;
  load_block(d.rgb, d.width, k % d.bw, k / d.bw, px);
  own[g] = fit_block(px);
}

// InterBlockSearch (codec/dxt_image.cpp:715-774): window rows j = by-S+r
// (r ascending), columns i = bx-S+t (t ascending), position r*2S + t.  Each
// work item scans `rows` rows of the window and leaves a ChunkFold summary;
// inter_decision (in k_intra_prep) folds them and applies the accept test
// (:890).
//
// With amask non-null, also records which offsets pass the accept test:
// amask[g * 16 + dy + 8] bit dx + 8 for the offsets (dx, dy) of the 16x16
// grid around the block (the training targets of the inter proposal network).
__global__ void __launch_bounds__(256) k_inter(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ tsid,
                                               int n_wtiles, int tw, int n_chunks, int S,
                                               const uint64_t* __restrict__ own, const uint32_t* __restrict__ rw,
                                               const float4* __restrict__ rtab, uint4* __restrict__ summ,
                                               size_t sstride, DevStats* __restrict__ st, int thr,
                                               uint16_t* __restrict__ amask) {
  __shared__ uint8_t s_snap[512];
  __shared__ float2 s_ab[4];
  __shared__ unsigned long long s_acc;
  load_lds_tables(s_snap, s_ab);
  int f = 0, bx = 0, by = 0, chunk = 0;
  unsigned long long n_cand = 0;
  if (tile_block(fr, tsid, n_wtiles, tw, n_chunks, f, bx, by, chunk) && !fr[f].is_intra) {
    const FrameDev& d = fr[f];
    const int bw = d.bw, bh = d.bh, base = d.base;
    const int g = base + by * bw + bx;
    BlockPx px;
    load_block(d.rgb, d.width, bx, by, px);
    BlockEval be;
    make_block_eval(px, own[g], be);
    const int W2 = 2 * S, rows = W2 / n_chunks;
    ChunkFold cf;
    const int t_lo = max(S - bx, 0), t_hi = min(bw - bx + S, W2);
    for (int r = chunk * rows; r < (chunk + 1) * rows; ++r) {
      const int j = by - S + r;
      uint32_t rm = 0;
      if (j >= 0 && j < bh) {
        cf.begin_row();
        for (int t = t_lo; t < t_hi; ++t) {
          const int q = base + j * bw + (bx - S + t);
          const float4 t4 = rtab[q];
          const CandResult res = eval_candidate<false>(be, rw[q], CandTab{t4.x, t4.y, t4.z, t4.w}, s_snap, s_ab);
          cf.step(res.valid ? (int)sat16(res.err) : (int)kNoCand, r * W2 + t, r);
          ++n_cand;
          const int col = t - S + 8;
          if (amask && res.valid && res.err <= thr && col >= 0 && col < 16) rm |= 1u << col;
        }
        cf.end_row(r);
      }
      const int lrow = r - S + 8;
      if (amask && lrow >= 0 && lrow < 16) amask[(size_t)g * 16 + lrow] = (uint16_t)rm;
    }
    summ[(size_t)chunk * sstride + g] = cf.pack();
  }
  wg_count(&s_acc, n_cand, st->inter_cands);
}

// The inter decision of block g of an inter frame from its chunk summaries
// (selected position and returned delta) and the previous frame's words rw.
__device__ __forceinline__ uint64_t inter_decision(const FrameDev& d, int g, int n_chunks, int S, int thr,
                                                   const uint32_t* __restrict__ rw, const uint4* __restrict__ summ,
                                                   size_t sstride) {
  int sel, ret;
  combine_chunks(summ, sstride, g, n_chunks, sel, ret);
  uint64_t dc = 0;
  if (sel >= 0 && ret <= thr) {
    const int W2 = 2 * S, r = sel / W2, t = sel % W2;
    const int k = g - d.base, bx = k % d.bw, by = k / d.bw;
    const uint32_t c = rw[d.base + (by - S + r) * d.bw + (bx - S + t)];
    dc = pack_dec(c, (uint32_t)t | 0x80u, (uint32_t)r | 0x80u, true);
  }
  return dc;
}

// Blocks that need the intra search, the current words the search starts
// from (inter-reused words are final: they read only the previous frame;
// every other block starts from its own word), and for the exact intra
// search (orig not null) each block's own error, the current words' tables
// and cleared pass flags.  With summ (not null: after an exhaustive inter
// search) the blocks of inter frames first take their inter decisions from
// the chunk summaries and write them to inter_dec; otherwise inter_dec holds
// them already (learned proposals).  With orig_ready the blocks of inter
// frames already have their own error in orig (the pruned inter search).
__global__ void __launch_bounds__(256) k_intra_prep(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid, int total,
                             const uint64_t* __restrict__ own, uint64_t* __restrict__ inter_dec,
                             uint8_t* __restrict__ need, uint32_t* __restrict__ cw, float4* __restrict__ ctab,
                             uint32_t* __restrict__ dirty, int* __restrict__ orig, BoundTab* __restrict__ cbt,
                             uint4* __restrict__ pm4, int n_chunks, int S, int thr,
                             const uint32_t* __restrict__ rw, const uint4* __restrict__ summ, size_t sstride,
                             int orig_ready) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const FrameDev& d = fr[sid[g]];
  const int k = g - d.base;
  bool inter_ok = false;
  const uint64_t o = own[g];
  uint32_t w = (uint32_t)(o >> 32);
  if (!d.is_intra) {
    uint64_t dc;
    if (summ) {
      dc = inter_decision(d, g, n_chunks, S, thr, rw, summ, sstride);
      inter_dec[g] = dc;
    } else {
      dc = inter_dec[g];
    }
    if (dec_acc(dc)) {
      inter_ok = true;
      w = dec_word(dc);
    }
  }
  need[g] = inter_ok ? 0 : 1;
  cw[g] = w;
  if (!orig) return;  // the learned intra search builds its own tables and errors
  dirty[g] = 0;
  const CandTab t = cand_table(w);
  ctab[g] = make_float4(t.asq, t.bsq, t.ab, t.f);
  if (cbt) {  // the pruned intra search: bound table of the word, passing set empty
    cbt[g] = bound_table(w);
    pm4[2 * g] = make_uint4(0u, 0u, 0u, 0u);
    pm4[2 * g + 1] = make_uint4(0u, 0u, 0u, 0u);
  }
  if (orig_ready && !d.is_intra) return;  // the pruned inter search wrote the own error
  BlockPx px;
  load_block(d.rgb, d.width, k % d.bw, k / d.bw, px);
  orig[g] = own_error(px, o);
}

// Intra scan order (codec/dxt_image.cpp:672-675): rows j = by .. by-2S+1
// (r = by - j ascending), columns i = bx+S-1 .. bx-S (t = bx+S-1-i ascending),
// skipping (j == by, i >= bx).  Position pos = r*2S + t.  E holds the error
// deltas of every position, four positions (int16) per 8-byte group:
// E4[(pos / 4) * nbs + g].

// One step of the selection rule (codec/dxt_image.cpp:697-708): a strictly
// smaller delta becomes the selection; the first delta <= 0 clamps the
// running minimum to 0 and ends the row.
struct Fold {
  int min_err = 0x7fffffff;
  int sel = -1;
  bool row_done = false;
  __device__ __forceinline__ void step(int e, int pos) {
    if (!row_done && e != kNoCand && e < min_err) {
      min_err = e;
      sel = pos;
      if (e <= 0) {
        min_err = 0;
        row_done = true;
      }
    }
  }
};

__device__ __forceinline__ uint64_t intra_decision(const uint32_t* __restrict__ cw, int base, int bw, int bx,
                                                   int by, int S, int thr, const Fold& fd) {
  if (fd.sel < 0 || fd.min_err > thr) return 0;
  const int W2 = 2 * S;
  const int r = fd.sel / W2, t = fd.sel % W2;
  const int j = by - r, i = bx + S - 1 - t;
  return pack_dec(cw[base + j * bw + i], (uint32_t)(i - bx + S), (uint32_t)(j - by + 2 * S - 1), true);
}


__device__ __forceinline__ int e_at(uint2 g, int n) {
  const uint32_t w = (n & 2) ? g.y : g.x;
  return (int)(int16_t)((n & 1) ? (w >> 16) : (w & 0xFFFFu));
}

// The selection rule (codec/dxt_image.cpp:697-708) over the passing deltas of
// block g's E (for thr >= 0 that is the full rule's result: before the first
// passing delta the others only raise the running minimum, which that delta
// then undercuts; after it they can never update it).  The S^2 groups of four
// are loaded 32 at a time.
__device__ __forceinline__ Fold fold_passing(const uint2* __restrict__ E4, size_t nbs, int g, int S, int thr) {
  const int W2 = 2 * S, groups = S * S;
  Fold fd;
  int row = -1;
  for (int g0 = 0; g0 < groups; g0 += 32) {
    uint2 buf[32];
#pragma unroll
    for (int b = 0; b < 32; ++b)  // past the last group: absent candidates
      buf[b] = g0 + b < groups ? E4[(size_t)(g0 + b) * nbs + g] : make_uint2(0x7FFF7FFFu, 0x7FFF7FFFu);
#pragma unroll
    for (int b = 0; b < 32; ++b) {
      const uint2 gg = buf[b];
      bool any = false;
#pragma unroll
      for (int nn = 0; nn < 4; ++nn) {
        const int e = e_at(gg, nn);
        any = any || (e != kNoCand && e <= thr);
      }
      if (!any) continue;
#pragma unroll
      for (int nn = 0; nn < 4; ++nn) {
        const int e = e_at(gg, nn);
        if (e == kNoCand || e > thr) continue;
        const int pos = 4 * (g0 + b) + nn;
        const int r = pos / W2;
        if (r != row) {
          row = r;
          fd.row_done = false;
        }
        fd.step(e, pos);
      }
    }
  }
  return fd;
}

// Passing sets: for windows of at most 256 positions (S <= 8), bit p % 32 of
// word p / 32 of pm[8 g .. 8 g + 7] is set exactly when block g's delta at
// position p passes (valid, <= thr).  Folding a block then reads its 32 bytes
// of set and the few passing deltas instead of all of its E.  Where a block
// keeps a passing set, E holds only its passing deltas (the pruned search
// writes only groups of four with a passing one, the passes only passing
// values); an entry whose bit is clear is not read.
constexpr int kPassWords = 8;

__device__ __forceinline__ Fold fold_passing_set(const uint32_t* __restrict__ pm, const uint2* __restrict__ E4,
                                                 size_t nbs, int g, int S) {
  const int W2 = 2 * S;
  const uint4* p4 = reinterpret_cast<const uint4*>(pm + (size_t)g * kPassWords);
  const uint4 a = p4[0], b = p4[1];
  const uint32_t m[kPassWords] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  Fold fd;
  int row = -1;
#pragma unroll
  for (int w = 0; w < kPassWords; ++w) {
    uint32_t mw = m[w];
    while (mw != 0u) {  // set bits in increasing position: scan order
      const int pos = 32 * w + __builtin_ctz(mw);
      mw &= mw - 1u;
      const int e = e_at(E4[(size_t)(pos >> 2) * nbs + g], pos & 3);
      const int r = pos / W2;
      if (r != row) {
        row = r;
        fd.row_done = false;
      }
      fd.step(e, pos);
    }
  }
  return fd;
}

__global__ void __launch_bounds__(256) k_intra_dense(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ tsid,
                                                     int n_wtiles, int tw, int n_chunks, int S,
                                                     const uint64_t* __restrict__ own, const uint8_t* __restrict__ need,
                                                     const uint32_t* __restrict__ cw, const float4* __restrict__ ctab,
                                                     const int* __restrict__ orig, uint2* __restrict__ E4, size_t nbs,
                                                     uint4* __restrict__ summ, size_t sstride,
                                                     DevStats* __restrict__ st) {
  __shared__ uint8_t s_snap[512];
  __shared__ float2 s_ab[4];
  __shared__ unsigned long long s_acc;
  load_lds_tables(s_snap, s_ab);
  int f = 0, bx = 0, by = 0, g = 0, chunk = 0;
  unsigned long long n_cand = 0;
  const bool in = tile_block(fr, tsid, n_wtiles, tw, n_chunks, f, bx, by, chunk);
  if (in) g = fr[f].base + by * fr[f].bw + bx;
  if (in && need[g]) {
    const FrameDev& d = fr[f];
    const int bw = d.bw, base = d.base;
    BlockPx px;
    load_block(d.rgb, d.width, bx, by, px);
    BlockEval be;
    make_block_eval(px, own[g], orig[g], be);
    const int W2 = 2 * S, rows = W2 / n_chunks;
    ChunkFold cf;
    uint32_t g0 = 0, g1 = 0;
    for (int r = chunk * rows; r < (chunk + 1) * rows; ++r) {
      const int j = by - r;
      cf.begin_row();
      for (int t = 0; t < W2; ++t) {
        const int pos = r * W2 + t;
        const int i = bx + S - 1 - t;
        int ev = kNoCand;
        if (j >= 0 && i >= 0 && i < bw && !(r == 0 && t < S)) {
          const int q = base + j * bw + i;
          const float4 t4 = ctab[q];
          const CandResult res =
              eval_candidate<false>(be, cw[q], CandTab{t4.x, t4.y, t4.z, t4.w}, s_snap, s_ab);
          ++n_cand;
          if (res.valid) ev = sat16(res.err);
        }
        cf.step(ev, pos, r);
        const uint32_t h = (uint32_t)(uint16_t)(int16_t)ev;
        switch (pos & 3) {
          case 0: g0 = h; break;
          case 1: g0 |= h << 16; break;
          case 2: g1 = h; break;
          default: g1 |= h << 16; E4[(size_t)(pos >> 2) * nbs + g] = make_uint2(g0, g1); break;
        }
      }
      cf.end_row(r);
    }
    summ[(size_t)chunk * sstride + g] = cf.pack();
  }
  wg_count(&s_acc, n_cand, st->intra_cands);
}

// Intra decisions from the dense pass's chunk summaries; blocks whose word
// changes go to the first changed list.
__global__ void __launch_bounds__(256) k_intra_combine(const FrameDev* __restrict__ fr,
                                                       const uint8_t* __restrict__ sid, int total, int n_chunks,
                                                       int S, int thr, const uint64_t* __restrict__ own,
                                                       const uint8_t* __restrict__ need,
                                                       const uint32_t* __restrict__ cw, const uint4* __restrict__ summ,
                                                       size_t sstride, uint64_t* __restrict__ dec,
                                                       uint32_t* __restrict__ cw_new, int* __restrict__ chg_list,
                                                       DevStats* __restrict__ st) {
  __shared__ unsigned int s_base;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  bool changed = false;
  if (g < total && need[g]) {
    const FrameDev& d = fr[sid[g]];
    int sel, ret;
    combine_chunks(summ, sstride, g, n_chunks, sel, ret);
    Fold fd;
    fd.sel = sel;
    fd.min_err = ret;
    const int k = g - d.base;
    const uint64_t dc = intra_decision(cw, d.base, d.bw, k % d.bw, k / d.bw, S, thr, fd);
    dec[g] = dc;
    const uint32_t w = dec_acc(dc) ? dec_word(dc) : (uint32_t)(own[g] >> 32);
    if (w != cw[g]) {
      cw_new[g] = w;
      changed = true;
    }
  }
  wg_append(changed, g, chg_list, &st->n_changed[0], &s_base);
}

// The dense pass's changed words become current; clears the lists of pass 0.
__global__ void __launch_bounds__(256) k_intra_apply(const uint32_t* __restrict__ cw_new, uint32_t* __restrict__ cw,
                              float4* __restrict__ ctab, const int* __restrict__ chg_list,
                              DevStats* __restrict__ st) {
  if (blockIdx.x == 0 && threadIdx.x == 0) st->n_changed[1] = 0;
  const unsigned n = st->n_changed[0];
  for (unsigned idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
    const int q = chg_list[idx];
    const uint32_t w = cw_new[q];
    cw[q] = w;
    const CandTab t = cand_table(w);
    ctab[q] = make_float4(t.asq, t.bsq, t.ab, t.f);
  }
}

// Fixpoint pass, part 1, over work items w = w0, w0 + stride, ... < n_chg *
// 4S^2: every (block, position) pair whose intra window holds a word changed
// in the previous step is re-evaluated with the new word, except (skip_tw_log
// >= 0) pairs inside one tile of 2^skip_tw_log x 64 / 2^skip_tw_log blocks,
// which the previous pass, tile-resolved, evaluated with the final word.  Work item =
// (changed position, slot); when 64 divides the 4S^2 slots every wave serves
// one changed position.  Blocks whose fold may change are flagged for part 2
// (plain stores: racing writers all store 1), and the passing sets pm (when
// kept) follow E.  Returns the evaluations done.
__device__ __forceinline__ unsigned long long reeval_items(
    const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid, int S, int thr,
    const uint64_t* __restrict__ own, const uint8_t* __restrict__ need, const uint32_t* __restrict__ cw,
    const float4* __restrict__ ctab, const int* __restrict__ chg_list, unsigned n_chg, uint2* __restrict__ E4,
    size_t nbs, uint32_t* __restrict__ dirty, const int* __restrict__ orig, const uint64_t* __restrict__ dec,
    uint32_t* __restrict__ pm, const uint8_t* s_snap, const float2* s_ab, size_t w0, size_t stride,
    int skip_tw_log) {
  const int W2 = 2 * S, slots = W2 * W2;
  const size_t total = (size_t)n_chg * slots;
  unsigned long long n_eval = 0;
  int16_t* E = reinterpret_cast<int16_t*>(E4);
  const bool wave_q = (slots & 63) == 0;
  for (size_t w = w0; w < total; w += stride) {
    int qi = (int)(w / slots);
    if (wave_q) qi = __builtin_amdgcn_readfirstlane(qi);
    const int q = chg_list[qi];
    const int s = (int)(w % slots);
    const int r = s / W2, t = s % W2;
    const int f = wave_q ? __builtin_amdgcn_readfirstlane((int)sid[q]) : (int)sid[q];
    const FrameDev& d = fr[f];
    const int ql = q - d.base;
    const int qx = ql % d.bw, qy = ql / d.bw;
    const int by = qy + r, bx = qx - S + 1 + t;
    if ((r == 0 && t < S) || by >= d.bh || bx < 0 || bx >= d.bw) continue;
    // pairs inside one tile of the last (tile-resolved) pass are current
    if (skip_tw_log >= 0 && (qx >> skip_tw_log) == (bx >> skip_tw_log) &&
        (qy >> (6 - skip_tw_log)) == (by >> (6 - skip_tw_log)))
      continue;
    const int k = d.base + by * d.bw + bx;
    if (!need[k]) continue;
    BlockPx px;
    load_block(d.rgb, d.width, bx, by, px);
    BlockEval be;
    make_block_eval(px, own[k], orig[k], be);
    const float4 t4 = ctab[q];
    const CandResult res = eval_candidate<false>(be, cw[q], CandTab{t4.x, t4.y, t4.z, t4.w}, s_snap, s_ab);
    const int16_t ev = res.valid ? sat16(res.err) : kNoCand;
    int16_t* slot = &E[((size_t)(s >> 2) * nbs + k) * 4 + (s & 3)];
    ++n_eval;
    // The decision depends only on the deltas that pass the threshold (for
    // thr >= 0: before the first passing delta the others only raise the
    // running minimum, which that delta then undercuts; after it they can
    // never update it; for thr < 0 nothing is ever accepted).  So the fold
    // needs redoing if a passing delta appeared, vanished or changed, and the
    // decision a new word if this position is the one the block reuses.
    const uint64_t dc = dec[k];
    const bool is_sel =
        dec_acc(dc) && (W2 - 1 - (int)((dc >> 40) & 0xFF)) * W2 + (W2 - 1 - (int)((dc >> 32) & 0xFF)) == s;
    const bool passes_new = ev != kNoCand && ev <= thr;
    bool passes_old;
    int16_t old;
    if (pm) {  // E is kept only where the passing set has the position
      passes_old = (pm[(size_t)k * kPassWords + (s >> 5)] >> (s & 31)) & 1u;
      old = passes_old ? *slot : kNoCand;
      if (passes_new && old != ev) *slot = ev;
    } else {
      old = *slot;
      passes_old = old != kNoCand && old <= thr;
      if (old != ev) *slot = ev;
    }
    if ((old != ev && (passes_old || passes_new)) || is_sel) dirty[k] = 1u;
    // one work item per (block, position) in a pass: bits of a word flip
    // concurrently only for different positions
    if (pm && passes_old != passes_new) atomicXor(&pm[(size_t)k * kPassWords + (s >> 5)], 1u << (s & 31));
  }
  return n_eval;
}

// Fixpoint pass, part 2, for block k (lane per block, raster order): re-folds
// a flagged block; a block whose word changes writes it (and its candidate
// table) as current at once.  A concurrent reader of that word in this pass
// may see the old or the new value: either way the next pass re-evaluates
// every pair that sees the word, so the passes still end at the same
// fixpoint.  pm: the passing sets (null: fold from E alone).  Returns whether
// the word changed.
// The decision of a flagged block k from its folded deltas (dirty cleared);
// returns false for a block that is not flagged.
__device__ __forceinline__ bool decide_block(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid, int k,
                                             int S, int thr, const uint32_t* __restrict__ cw,
                                             const uint2* __restrict__ E4, size_t nbs, uint32_t* __restrict__ dirty,
                                             uint64_t* __restrict__ dec, const uint32_t* __restrict__ pm,
                                             uint64_t& dc) {
  if (dirty[k] == 0u) return false;
  dirty[k] = 0;
  // Only deltas that pass the threshold can change the decision (see
  // reeval_items): fold those, from the passing set where one is kept.
  const Fold fd = pm ? fold_passing_set(pm, E4, nbs, k, S) : fold_passing(E4, nbs, k, S, thr);
  const FrameDev& d = fr[sid[k]];
  const int kl = k - d.base;
  dc = intra_decision(cw, d.base, d.bw, kl % d.bw, kl / d.bw, S, thr, fd);
  dec[k] = dc;
  return true;
}

__device__ __forceinline__ bool select_block(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid, int k,
                                             int S, int thr, const uint64_t* __restrict__ own,
                                             uint32_t* __restrict__ cw, float4* __restrict__ ctab,
                                             const uint2* __restrict__ E4, size_t nbs, uint32_t* __restrict__ dirty,
                                             uint64_t* __restrict__ dec, const uint32_t* __restrict__ pm) {
  uint64_t dc = 0;
  if (!decide_block(fr, sid, k, S, thr, cw, E4, nbs, dirty, dec, pm, dc)) return false;
  const uint32_t w = dec_acc(dc) ? dec_word(dc) : (uint32_t)(own[k] >> 32);
  if (w == cw[k]) return false;
  cw[k] = w;
  const CandTab t = cand_table(w);
  ctab[k] = make_float4(t.asq, t.bsq, t.ab, t.f);
  return true;
}

// Fixpoint pass, part 1, one launch: reads the changed list of parity `par`
// and clears the one of the next pass.
__global__ void __launch_bounds__(256) k_intra_reeval(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid,
                                                      int S, int thr, const uint64_t* __restrict__ own,
                                                      const uint8_t* __restrict__ need, const uint32_t* __restrict__ cw,
                                                      const float4* __restrict__ ctab, const int* __restrict__ chg_list,
                                                      uint2* __restrict__ E4, size_t nbs, uint32_t* __restrict__ dirty,
                                                      const int* __restrict__ orig, const uint64_t* __restrict__ dec,
                                                      DevStats* __restrict__ st, int par, uint32_t* __restrict__ pm,
                                                      int tw_log) {
  const int slots = 4 * S * S;
  const unsigned n_chg = st->n_changed[par];
  const int skip = st->skip_tile ? tw_log : -1;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->n_changed[par ^ 1] = 0;
  if ((size_t)blockIdx.x * blockDim.x >= (size_t)n_chg * slots) return;  // workgroup-uniform
  __shared__ uint8_t s_snap[512];
  __shared__ float2 s_ab[4];
  __shared__ unsigned long long s_acc;
  load_lds_tables(s_snap, s_ab);
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&st->active_passes, 1u);
  const unsigned long long n_eval =
      reeval_items(fr, sid, S, thr, own, need, cw, ctab, chg_list, n_chg, E4, nbs, dirty, orig, dec, pm, s_snap, s_ab,
                   (size_t)blockIdx.x * blockDim.x + threadIdx.x, (size_t)gridDim.x * blockDim.x, skip);
  wg_count(&s_acc, n_eval, st->reeval_cands);
}

// Fixpoint pass, part 2, one launch: writes the changed list of parity `par`
// (cleared by the preceding k_intra_reeval); its length goes to *conv when
// the pass ends.
__global__ void __launch_bounds__(256) k_intra_select(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid,
                                                      int total, int S, int thr, const uint64_t* __restrict__ own,
                                                      uint32_t* __restrict__ cw, float4* __restrict__ ctab,
                                                      const uint2* __restrict__ E4, size_t nbs,
                                                      uint32_t* __restrict__ dirty, uint64_t* __restrict__ dec,
                                                      int* __restrict__ chg_list, DevStats* __restrict__ st, int par,
                                                      unsigned int* __restrict__ conv,
                                                      const uint32_t* __restrict__ pm) {
  // The preceding k_intra_reeval flags blocks only from the changed list of
  // parity par ^ 1; when that list was empty there is nothing to re-fold.
  // No workgroup of this pass writes that counter, so either every workgroup
  // returns here or none does; an empty pass reports nothing (pass_end): the
  // pass that produced the empty list already wrote 0 to *conv.
  if (st->n_changed[par ^ 1] == 0) {
    report_fixpoint(conv);
    return;
  }
  __shared__ unsigned int s_base;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool changed = k < total && select_block(fr, sid, k, S, thr, own, cw, ctab, E4, nbs, dirty, dec, pm);
  wg_append(changed, k, chg_list, &st->n_changed[par], &s_base);
  pass_end(st, par, nullptr, conv);
}

// Fixpoint pass, part 2, tile-resolved: a wave
// owns a tile of 2^tw_log x 64 / 2^tw_log blocks of one frame (lane = row *
// width + column).  Round 0 re-folds the blocks k_intra_reeval flagged.
// Then, while a word of the tile changed in the last round, every block whose
// window holds such a word re-evaluates that position with it (taken from the
// owning lane), updates E and its passing set, and re-folds when a passing
// delta appeared, vanished or changed or the position is the one it reuses.
// A window holds only earlier blocks in raster order, which inside a tile are
// lower lanes: lane l is final after round l, so the rounds end by round 64.
// The tile's words are written once, after the rounds (other tiles read the
// word a block started the pass with or the one it ends with), and those that
// changed are listed for the next pass, whose k_intra_reeval skips the pairs
// inside one tile (settled here).  Writes the changed list of parity `par`;
// its length goes to *conv when the pass ends.
constexpr int kTileRounds = 66;

__global__ void __launch_bounds__(256) k_intra_tile_pass(
    const FrameDev* __restrict__ fr, int n_frames, int n_itiles, int tw_log, int S, int thr,
    const uint64_t* __restrict__ own, const uint8_t* __restrict__ need, uint32_t* __restrict__ cw,
    float4* __restrict__ ctab, uint2* __restrict__ E4, size_t nbs, uint32_t* __restrict__ dirty,
    const int* __restrict__ orig, uint64_t* __restrict__ dec, int* __restrict__ chg_list, DevStats* __restrict__ st,
    int par, unsigned int* __restrict__ conv, uint32_t* __restrict__ pm) {
  // The preceding k_intra_reeval flags blocks only from the changed list of
  // parity par ^ 1; when that list was empty there is nothing to re-fold.
  // No workgroup of this pass writes that counter, so either every workgroup
  // returns here or none does; an empty pass reports nothing (pass_end): the
  // pass that produced the empty list already wrote 0 to *conv.
  if (st->n_changed[par ^ 1] == 0) {
    report_fixpoint(conv);
    return;
  }
  __shared__ uint8_t s_snap[512];
  __shared__ float2 s_ab[4];
  __shared__ unsigned long long s_acc;
  __shared__ unsigned int s_base;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->skip_tile = 1u;  // read by the next pass's k_intra_reeval
  const int wv = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int lane = (int)(threadIdx.x & 63);
  const int W2 = 2 * S;
  int16_t* E = reinterpret_cast<int16_t*>(E4);
  unsigned long long n_eval = 0;
  bool changed = false;
  int g = 0;
  // the tile's geometry and its flagged blocks (k_intra_reeval flags only
  // blocks that search the intra window)
  const FrameDev* dp = nullptr;
  int x0 = 0, y0 = 0, bx = 0, by = 0;
  bool inside = false, redo = false;
  if (wv < n_itiles) {
    int f = 0;
    while (f + 1 < n_frames && fr[f + 1].itile_base <= wv) ++f;
    dp = &fr[f];
    const int lt = wv - dp->itile_base, tx = lt % dp->itiles_x, ty = lt / dp->itiles_x;
    x0 = tx << tw_log;
    y0 = ty << (6 - tw_log);
    bx = x0 + (lane & ((1 << tw_log) - 1));
    by = y0 + (lane >> tw_log);
    inside = bx < dp->bw && by < dp->bh;
    g = inside ? dp->base + by * dp->bw + bx : 0;
    redo = inside && dirty[g] != 0u;
  }
  // the evaluation tables, only where some block of the workgroup re-folds
  if (__syncthreads_or(redo)) load_lds_tables(s_snap, s_ab);
  if (__ballot(redo) != 0ull) {
    const FrameDev& d = *dp;
    const int tw = 1 << tw_log;
    const bool act = inside && need[g];
    const uint32_t start = inside ? cw[g] : 0u;
    uint32_t cur = start;
    uint64_t dc = act ? dec[g] : 0;
    bool have_be = false;
    // the lanes of the tile whose blocks lie in this block's window
    unsigned long long wmask = 0ull;
    if (act)
      for (int L = 0; L < 64; ++L) {
        const int r = by - (y0 + (L >> tw_log)), t = bx + S - 1 - (x0 + (L & (tw - 1)));
        if (r >= 0 && r < W2 && t >= 0 && t < W2 && !(r == 0 && t < S)) wmask |= 1ull << L;
      }
    BlockEval be;
    for (int round = 0; round < kTileRounds; ++round) {
      // fold the flagged blocks; a selected position inside the tile takes
      // its word from the owning lane
      int src = -1, sel_i = 0, sel_j = 0;
      bool acc = false;
      if (redo) {
        dirty[g] = 0u;
        __asm__ volatile("" ::: "memory");  // E and pm: this lane's stores before the fold reads them
        const Fold fd = pm ? fold_passing_set(pm, E4, nbs, g, S) : fold_passing(E4, nbs, g, S, thr);
        if (fd.sel >= 0 && fd.min_err <= thr) {
          acc = true;
          sel_j = by - fd.sel / W2;
          sel_i = bx + S - 1 - fd.sel % W2;
          if (sel_j >= y0 && sel_i >= x0 && sel_i < x0 + tw) src = ((sel_j - y0) << tw_log) | (sel_i - x0);
        }
      }
      const uint32_t v = __shfl(cur, src >= 0 ? src : lane);
      bool ch = false;
      if (redo) {
        const uint32_t w = !acc ? (uint32_t)(own[g] >> 32) : src >= 0 ? v : cw[d.base + sel_j * d.bw + sel_i];
        dc = acc ? pack_dec(w, (uint32_t)(sel_i - bx + S), (uint32_t)(sel_j - by + W2 - 1), true) : 0;
        dec[g] = dc;
        ch = w != cur;
        cur = w;
      }
      const unsigned long long mask = __ballot(ch);
      if (mask == 0ull) break;
      // re-evaluate the window positions that hold a word changed this round;
      // each lane walks its own such positions (a lane reads the word from
      // the owning lane), so lanes evaluate side by side
      redo = false;
      unsigned long long pend = mask & wmask;
      while (__ballot(pend != 0ull) != 0ull) {
        const int L = pend != 0ull ? __builtin_ctzll(pend) : lane;
        const uint32_t w = __shfl(cur, L);
        if (pend == 0ull) continue;
        pend &= pend - 1ull;
        const int r = by - (y0 + (L >> tw_log)), t = bx + S - 1 - (x0 + (L & (tw - 1)));
        if (!have_be) {
          BlockPx px;
          load_block(d.rgb, d.width, bx, by, px);
          make_block_eval(px, own[g], orig[g], be);
          have_be = true;
        }
        const CandResult res = eval_candidate<false>(be, w, cand_table(w), s_snap, s_ab);
        const int16_t ev = res.valid ? sat16(res.err) : kNoCand;
        const int s = r * W2 + t;
        int16_t* slot = &E[((size_t)(s >> 2) * nbs + g) * 4 + (s & 3)];
        ++n_eval;
        const bool passes_new = ev != kNoCand && ev <= thr;
        bool passes_old;
        int16_t old;
        if (pm) {  // E is kept only where the passing set has the position
          passes_old = (pm[(size_t)g * kPassWords + (s >> 5)] >> (s & 31)) & 1u;
          old = passes_old ? *slot : kNoCand;
          if (passes_new && old != ev) *slot = ev;
        } else {
          old = *slot;
          passes_old = old != kNoCand && old <= thr;
          if (old != ev) *slot = ev;
        }
        const bool is_sel = dec_acc(dc) && (W2 - 1 - (int)((dc >> 40) & 0xFF)) * W2 +
                                                   (W2 - 1 - (int)((dc >> 32) & 0xFF)) == s;
        if ((old != ev && (passes_old || passes_new)) || is_sel) redo = true;
        if (pm && passes_old != passes_new) pm[(size_t)g * kPassWords + (s >> 5)] ^= 1u << (s & 31);
      }
    }
    changed = inside && cur != start;
    if (changed) {
      cw[g] = cur;
      const CandTab t = cand_table(cur);
      ctab[g] = make_float4(t.asq, t.bsq, t.ab, t.f);
    }
  }
  wg_count(&s_acc, n_eval, st->reeval_cands);
  wg_append(changed, g, chg_list, &st->n_changed[par], &s_base);
  pass_end(st, par, nullptr, conv);
}

// Which intra offsets pass the accept test at the fixpoint (the training
// targets of the intra proposal network): amask[g * 16 + dy] bit dx + 8 for
// the offsets (dx, dy = by - j) of the 16x16 causal grid, for every block
// that searched the intra window.
__global__ void __launch_bounds__(256) k_intra_amask(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid,
                                                     int total, int S, int thr, const uint8_t* __restrict__ need,
                                                     const uint2* __restrict__ E4, size_t nbs,
                                                     uint16_t* __restrict__ amask) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total || !need[g]) return;
  const int W2 = 2 * S;
  for (int dy = 0; dy < 16; ++dy) {
    uint32_t rm = 0;
    if (dy < W2) {
      for (int t = 0; t < W2; ++t) {
        const int col = S - 1 - t + 8;
        if (col < 0 || col >= 16) continue;
        const int pos = dy * W2 + t;
        const int e = e_at(E4[(size_t)(pos >> 2) * nbs + g], pos & 3);
        if (e != kNoCand && e <= thr) rm |= 1u << col;
      }
    }
    amask[(size_t)g * 16 + dy] = (uint16_t)rm;
  }
}

// Commit (codec/dxt_image.cpp:900-905, 922-926, 953-954).
__global__ void __launch_bounds__(256) k_commit(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid, int total,
                         const uint64_t* __restrict__ own, const uint64_t* __restrict__ inter_dec,
                         const uint8_t* __restrict__ need, const uint64_t* __restrict__ intra_dec,
                         uint32_t* __restrict__ tile_count, DevStats* __restrict__ st) {
  __shared__ uint8_t s_snap[512];
  __shared__ float2 s_ab[4];
  __shared__ unsigned s_cnt[2];
  load_lds_tables(s_snap, s_ab);
  if (threadIdx.x < 2) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  int tile = -1;
  if (g < total) {
    const FrameDev& d = fr[sid[g]];
    const int k = g - d.base;
    const uint64_t o = own[g];
    uint64_t dc = 0;
    int kd = 0;
    if (!d.is_intra && dec_acc(inter_dec[g])) {
      dc = inter_dec[g];
      kd = 2;
    } else if (need[g] && dec_acc(intra_dec[g])) {
      dc = intra_dec[g];
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
      const uint32_t c = dec_word(dc);
      mx = (uint8_t)(dc >> 32);
      my = (uint8_t)(dc >> 40);
      if (c != (uint32_t)(o >> 32)) {
        // re-assigned indices: RecalculateEndpoints (codec/dxt_image.cpp:901-905, 923-926)
        BlockPx px;
        load_block(d.rgb, d.width, k % d.bw, k / d.bw, px);
        BlockEval be;
        make_block_eval(px, o, be);
        const CandResult r = eval_candidate<true>(be, c, cand_table(c), s_snap, s_ab);
        e1 = r.e1;
        e2 = r.e2;
        p = (uint64_t)pack565(e1 & 0xFF, (e1 >> 8) & 0xFF, e1 >> 16) |
            ((uint64_t)pack565(e2 & 0xFF, (e2 >> 8) & 0xFF, e2 >> 16) << 16) | ((uint64_t)c << 32);
      }
    }
    d.phys_out[k] = p;
    d.motion[2 * k] = mx;
    d.motion[2 * k + 1] = my;
    if (d.eps) {
      uint8_t* e = d.eps + 6 * k;
      e[0] = (uint8_t)e1; e[1] = (uint8_t)(e1 >> 8); e[2] = (uint8_t)(e1 >> 16);
      e[3] = (uint8_t)e2; e[4] = (uint8_t)(e2 >> 8); e[5] = (uint8_t)(e2 >> 16);
    }
    if (d.kind) d.kind[k] = (uint8_t)kd;
    if (kd) atomicAdd(&s_cnt[kd - 1], 1u);
    tile = kd == 0 ? d.tile_base + k / kUniqueTile : -1;
  }
  // unique counts per tile, one atomic per wave when the wave's unique blocks share a tile
  {
    const unsigned long long m = __ballot(tile >= 0);
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      const int t0 = __shfl(tile, leader);
      const unsigned long long same = __ballot(tile == t0);
      if (same == m) {
        if ((int)__lane_id() == leader) atomicAdd(&tile_count[t0], (unsigned)__popcll(m));
      } else if (tile >= 0) {
        atomicAdd(&tile_count[tile], 1u);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_cnt[0]) atomicAdd(&st->n_intra[(blockIdx.x % kStatSlots) * kSlot32], s_cnt[0]);
    if (s_cnt[1]) atomicAdd(&st->n_inter[(blockIdx.x % kStatSlots) * kSlot32], s_cnt[1]);
  }
}

// Unique palette of each frame in raster order: one workgroup (256 threads)
// per tile of kUniqueTile blocks.
__global__ void __launch_bounds__(256) k_unique(const FrameDev* __restrict__ fr, int n_frames,
                                                const uint64_t* __restrict__ own,
                                                const uint32_t* __restrict__ tile_count,
                                                DevStats* __restrict__ st) {
  constexpr int kT = 256;
  __shared__ uint32_t s_sum[kT];
  __shared__ int s_f;
  const int tile = blockIdx.x;
  if (threadIdx.x == 0) {
    int f = 0;
    while (f + 1 < n_frames && fr[f + 1].tile_base <= tile) ++f;
    s_f = f;
  }
  __syncthreads();
  const FrameDev& d = fr[s_f];
  const int lt = tile - d.tile_base;
  // prefix of the frame's preceding tiles
  uint32_t acc = 0;
  for (int t = threadIdx.x; t < lt; t += kT) acc += tile_count[d.tile_base + t];
  s_sum[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kT / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s_sum[threadIdx.x] += s_sum[threadIdx.x + off];
    __syncthreads();
  }
  const uint32_t base = s_sum[0];
  __syncthreads();
  constexpr int per = kUniqueTile / kT;
  const int k0 = lt * kUniqueTile + threadIdx.x * per;
  uint32_t flags = 0, cnt = 0;
#pragma unroll
  for (int u = 0; u < per; ++u) {
    const int k = k0 + u;
    const bool uq = k < d.nb && d.motion[2 * k] == 255 && d.motion[2 * k + 1] == 255;
    flags |= (uq ? 1u : 0u) << u;
    cnt += uq ? 1u : 0u;
  }
  s_sum[threadIdx.x] = cnt;
  __syncthreads();
  for (int off = 1; off < kT; off <<= 1) {
    const uint32_t v = (int)threadIdx.x >= off ? s_sum[threadIdx.x - off] : 0u;
    __syncthreads();
    s_sum[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t pos = base + s_sum[threadIdx.x] - cnt;
#pragma unroll
  for (int u = 0; u < per; ++u)
    if (flags & (1u << u)) d.unique[pos++] = (uint32_t)(own[d.base + k0 + u] >> 32);
  if (lt == d.n_tiles - 1 && threadIdx.x == kT - 1) {
    *d.n_unique = pos;
    atomicAdd(&st->n_unique, pos);
  }
}

}  // namespace mptc
