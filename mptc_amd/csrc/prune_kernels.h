// Pruned exhaustive search: the same decisions as k_inter / k_intra_dense
// (bit-exact), with most candidate evaluations replaced by a cheap lower
// bound on their error.
//
// Why it is exact.  For a threshold thr >= 0 a block's decision depends only
// on the candidates whose error delta passes (valid, delta <= thr): before the
// first passing candidate (scan order) the others can only raise the running
// minimum, which that candidate then undercuts; after it they can never update
// it (codec/dxt_image.cpp:697-708).  So a candidate proved non-passing may be
// recorded as any non-passing value (kPruned) without changing the decision.
//
// The bound.  Any palette gives one colour per index value, so the squared
// error of a block under index word c is at least the scatter of its pixels
// around their group means, W_c = sum |p|^2 - sum_g |S_g|^2 / n_g (S_g, n_g:
// sum and count of the pixels with index g).  Error() = floor(SSE / 48)
// (:258-276), so delta >= floor(W_c / 48) - err_own, and the candidate cannot
// pass when W_c >= 48 (thr + err_own + 1).  W_c is computed from integer group
// sums (v_dot4 over byte masks of the word) and one fp32 division per group,
// with a 64 margin covering the fp32 rounding (< 48).
//
//   k_pruned_search wave tiles as the search kernels; per position the bound,
//                   non-passing -> kPruned; the wave's survivors are compacted
//                   in LDS and evaluated exactly, 64 per round.  Inter: the
//                   chunk summaries of the selection rule (as k_inter, then
//                   inter_decision in k_intra_prep); intra: E for the
//                   fixpoint passes
//   k_fold_intra    per block, the reference's selection over the passing
//                   deltas of E: the intra decision with the changed-word list
//                   (as k_intra_combine).  With 1, 2 or 4 chunks a workgroup
//                   holds every chunk of its tiles and the intra search takes
//                   these decisions itself, after a workgroup barrier.
//
// The words' bound tables (BoundTab, batch_kernels.h) are built by
// k_block_prep (the previous frame's words) and k_intra_prep (the current
// words, which also clears the passing sets).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "batch_kernels.h"

namespace mptc {

constexpr int16_t kPruned = 32766;  // E value of a candidate proved not to pass
constexpr int kPruneMaxS = 8;       // window positions of one batch fit the LDS list
static_assert(4 * kPruneMaxS * kPruneMaxS <= 32 * kPassWords, "a passing set holds every window position");

// Pixel side of the bound for one block.
struct BoundBlk {
  uint32_t pc[12];  // pc[3j + ch]: channel ch of pixels 4j..4j+3
  int tot[3];       // channel sums
  uint32_t sq;      // sum of squared channel values
  float q;          // the same as fp32 (exact)
  float lim;        // prune when W_c >= lim = 48 (thr + err_own + 1) + 64
};

// Prune when W_c >= 48 (thr + err_own + 1) + 64 (64: the fp32 rounding margin).
__device__ __forceinline__ float bound_limit(int thr, int err_own) { return (float)(48 * (thr + err_own + 1) + 64); }

__device__ __forceinline__ void make_bound_blk(const BlockPx& px, int err_own, int thr, BoundBlk& b) {
  uint32_t sq = 0;
  int tot[3] = {0, 0, 0};
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const uint32_t v = px.ch(4 * j, ch) | (px.ch(4 * j + 1, ch) << 8) | (px.ch(4 * j + 2, ch) << 16) |
                         (px.ch(4 * j + 3, ch) << 24);
      b.pc[3 * j + ch] = v;
      sq = __builtin_amdgcn_udot4(v, v, sq, false);
      tot[ch] = (int)__builtin_amdgcn_udot4(v, 0x01010101u, (uint32_t)tot[ch], false);
    }
  b.tot[0] = tot[0];
  b.tot[1] = tot[1];
  b.tot[2] = tot[2];
  b.sq = sq;
  b.q = (float)sq;  // <= 16 * 3 * 255^2 < 2^24: exact
  b.lim = bound_limit(thr, err_own);
}

__device__ __forceinline__ int group_sum(const uint4& m, const BoundBlk& b, int ch) {
  uint32_t s = __builtin_amdgcn_udot4(m.x, b.pc[ch], 0u, false);
  s = __builtin_amdgcn_udot4(m.y, b.pc[3 + ch], s, false);
  s = __builtin_amdgcn_udot4(m.z, b.pc[6 + ch], s, false);
  s = __builtin_amdgcn_udot4(m.w, b.pc[9 + ch], s, false);
  return (int)s;
}

// True when the candidate word of `t` provably does not pass the threshold.
__device__ __forceinline__ bool bound_prunes(const BoundBlk& b, const BoundTab& t) {
  int s0[3], s1[3], s2[3], s3[3];
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    s0[ch] = group_sum(t.m0, b, ch);
    s1[ch] = group_sum(t.m1, b, ch);
    s2[ch] = group_sum(t.m2, b, ch);
    s3[ch] = b.tot[ch] - s0[ch] - s1[ch] - s2[ch];
  }
  // 0 <= S_g <= 4080 (12 bits): 24-bit multiplies are exact; |S_g|^2 <= 3 * 4080^2 < 2^26 fits
  // int32, and the fp32 products and sum err by < 48 in total
  auto sq3 = [](const int* v) {
    return (int)(__umul24((uint32_t)v[0], (uint32_t)v[0]) + __umul24((uint32_t)v[1], (uint32_t)v[1]) +
                 __umul24((uint32_t)v[2], (uint32_t)v[2]));
  };
  const int q0 = sq3(s0), q1 = sq3(s1), q2 = sq3(s2), q3 = sq3(s3);
  const float T = (float)q0 * t.inv.x + (float)q1 * t.inv.y + (float)q2 * t.inv.z + (float)q3 * t.inv.w;
  return b.q - T >= b.lim;
}

// Window position (row r, column t; position r * 2S + t) of block (bx, by) ->
// the candidate block (j, i); false outside the frame or (intra) outside the
// causal window.
template <bool kIntra>
__device__ __forceinline__ bool window_block(int r, int t, int S, int bx, int by, int bw, int bh, int& j, int& i) {
  if (kIntra) {
    j = by - r;
    i = bx + S - 1 - t;
    return j >= 0 && i >= 0 && i < bw && !(r == 0 && t < S);
  }
  j = by - S + r;
  i = bx - S + t;
  return j >= 0 && j < bh && i >= 0 && i < bw;
}

constexpr int kPruneBatch = 16;  // window positions per bound / evaluate round of a wave
constexpr int kPackedBlk = 15;   // LDS words per block: pc[12], sum of squares, own word, own error
constexpr int kFoldWords = 8;    // LDS words of an inter lane's fold: ChunkFold's 7 fields and its row

// An inter lane's ChunkFold lives in LDS between the (rare) batches that
// have passing deltas, so that it holds no registers through the rest.  Its
// fields fit 16 bits: positions < 1024, rows < 64, and min_val is a passing
// delta (<= thr) or the initial "none" (stored as kNoCand).
__device__ __forceinline__ void cf_load(int16_t (*s)[64], int lane, ChunkFold& cf, int& row) {
  cf.first_le0 = s[0][lane];
  cf.le0_row = s[1][lane];
  cf.after_sel = s[2][lane];
  cf.last_neg = s[3][lane];
  cf.min_val = s[4][lane] == kNoCand ? 0x7fffffff : (int)s[4][lane];
  cf.min_pos = s[5][lane];
  cf.row_neg = s[6][lane];
  row = s[7][lane];
}
__device__ __forceinline__ void cf_store(int16_t (*s)[64], int lane, const ChunkFold& cf, int row) {
  s[0][lane] = (int16_t)cf.first_le0;
  s[1][lane] = (int16_t)cf.le0_row;
  s[2][lane] = (int16_t)cf.after_sel;
  s[3][lane] = (int16_t)cf.last_neg;
  s[4][lane] = (int16_t)min(cf.min_val, (int)kNoCand);
  s[5][lane] = (int16_t)cf.min_pos;
  s[6][lane] = (int16_t)cf.row_neg;
  s[7][lane] = (int16_t)row;
}

// Pruned search over the window of every searching block (inter: blocks of
// inter frames against the previous frame's words; intra: blocks with need[]
// set against the current words), in the wave-tile mapping of the dense
// kernels.  Per batch of kPruneBatch positions each lane bounds its block's
// candidates; the survivors of the whole wave are compacted in LDS and
// evaluated exactly 64 at a time, a lane taking any block's survivor (the
// pixel side of the wave's 64 blocks is kept in LDS).  Each position's value
// is the exact delta of a survivor, kPruned for a pruned and kNoCand for an
// absent position.  Intra: the values go to E.  Inter: each lane folds its
// chunk's passing values in scan order into the chunk summary summ[chunk][g]
// of the dense inter kernel (for thr >= 0 the selection depends only on the
// passing deltas, and a row without one leaves a summary as it is).
template <bool kIntra>
__global__ void __launch_bounds__(256) k_pruned_search(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ tsid,
                                                       int n_wtiles, int tw, int n_chunks, int S, int thr,
                                                       const uint64_t* __restrict__ own,
                                                       const uint8_t* __restrict__ need,
                                                       const uint32_t* __restrict__ words,
                                                       const float4* __restrict__ tabs,
                                                       const BoundTab* __restrict__ btab, int* __restrict__ orig,
                                                       uint2* __restrict__ E4, size_t nbs, uint32_t* __restrict__ pm,
                                                       uint4* __restrict__ summ, DevStats* __restrict__ st,
                                                       uint64_t* __restrict__ dec, uint32_t* __restrict__ cw_new,
                                                       int* __restrict__ chg_list, int fold) {
  __shared__ uint8_t s_snap[512];
  __shared__ float2 s_ab[4];
  __shared__ uint32_t s_blk[4][kPackedBlk][64];      // [wave][field][lane]
  __shared__ int16_t s_e[4][kPruneBatch][64];        // the batch's E values, [wave][position][lane]
  __shared__ uint16_t s_list[4][kPruneBatch * 64 + 64];  // survivors: lane | position << 6; + per-lane dummies
  __shared__ uint8_t s_pass[4][64];                      // inter: the lane's batch holds a passing delta
  __shared__ int16_t s_cf[4][kFoldWords][64];            // inter: each lane's ChunkFold state and row
  __shared__ unsigned long long s_acc;
  __shared__ unsigned int s_base;
  load_lds_tables(s_snap, s_ab);
  int f = 0, bx = 0, by = 0, chunk = 0;
  const bool in = tile_block(fr, tsid, n_wtiles, tw, n_chunks, f, bx, by, chunk);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const FrameDev& d = fr[f];
  const int g = in ? d.base + by * d.bw + bx : 0;
  const bool act = in && (kIntra ? need[g] != 0 : !d.is_intra);
  // origin of the wave tile (a wave with survivors is a tile inside the batch);
  // tw is a power of two
  const int tw_log = __builtin_ctz((unsigned)tw);
  const int bx0 = bx - (lane & (tw - 1)), by0 = by - (lane >> tw_log);
  if (act) {
    BlockPx px;
    load_block(d.rgb, d.width, bx, by, px);
    const int e_own = kIntra ? orig[g] : own_error(px, own[g]);
    if (!kIntra && orig && chunk == 0) orig[g] = e_own;  // inter: k_intra_prep then skips the block's own error
    BoundBlk b;
    make_bound_blk(px, e_own, thr, b);
#pragma unroll
    for (int k = 0; k < 12; ++k) s_blk[wave][k][lane] = b.pc[k];
    s_blk[wave][12][lane] = b.sq;
    s_blk[wave][13][lane] = (uint32_t)(own[g] >> 32);
    s_blk[wave][14][lane] = (uint32_t)e_own;
  }
  unsigned long long n_bound = 0, n_exact = 0;
  const int W2 = 2 * S, rows = W2 / n_chunks, p0 = chunk * rows * W2, p1 = p0 + rows * W2;
  // (rb, tb): window row and column of the batch's first position, stepped
  // without divisions
  int rb = chunk * rows, tb = 0;
  // A wave none of whose blocks searches has nothing to do.  Otherwise every
  // lane computes every bound (a vector instruction costs the same for any
  // number of active lanes) and the lanes without a candidate ignore theirs:
  // no divergent branches in the loop.
  const bool wave_act = __ballot(act) != 0ull;
  // the frame's fields in registers (the loops' stores could otherwise alias them)
  const int fbw = d.bw, fbh = d.bh, fbase = d.base;
  // inter: the selection rule over the chunk's positions in scan order
  if (!kIntra) {
    ChunkFold cf;
    cf.begin_row();
    cf_store(s_cf[wave], lane, cf, rb);
  }
  for (int pb = p0; wave_act && pb < p1; pb += kPruneBatch) {
    const int nb = min(kPruneBatch, p1 - pb);  // a multiple of 4
    uint32_t cnt = 0;                          // wave-uniform
    int r = rb, t = tb;
    if (!kIntra) s_pass[wave][lane] = 0;
    // the block's bound constants, rebuilt from LDS for each batch so that they
    // hold no registers through the evaluations (zero for lanes that do not
    // search: their bounds are computed and ignored)
    BoundBlk b{};
    if (act) {
#pragma unroll
      for (int w = 0; w < 12; ++w) b.pc[w] = s_blk[wave][w][lane];
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        uint32_t tot = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) tot = __builtin_amdgcn_udot4(b.pc[3 * j + ch], 0x01010101u, tot, false);
        b.tot[ch] = (int)tot;
      }
      b.sq = s_blk[wave][12][lane];
      b.q = (float)b.sq;
      b.lim = bound_limit(thr, (int)s_blk[wave][14][lane]);
    }
    for (int k = 0; k < nb; ++k) {
      int j = 0, i = 0;
      const bool there = act & window_block<kIntra>(r, t, S, bx, by, fbw, fbh, j, i);
      t = t + 1 == W2 ? 0 : t + 1;
      r += t == 0 ? 1 : 0;
      // clamped into the frame: the load is always in bounds
      const int q = fbase + min(max(j, 0), fbh - 1) * fbw + min(max(i, 0), fbw - 1);
      const bool pruned = bound_prunes(b, btab[q]);
      const bool surv = there && !pruned;
      n_bound += there ? 1u : 0u;
      const unsigned long long bal = __ballot(surv);
      // survivors take the next list slots; the others write their lane's dummy slot
      const uint32_t slot = surv ? cnt + (uint32_t)__popcll(bal & ((1ull << lane) - 1))
                                 : (uint32_t)(kPruneBatch * 64 + lane);
      s_list[wave][slot] = (uint16_t)(lane | (k << 6));
      s_e[wave][k][lane] = (int16_t)(there ? kPruned : kNoCand);  // survivors' entries are replaced below
      cnt += (uint32_t)__popcll(bal);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t e0 = 0; e0 < cnt; e0 += 64) {
      if (e0 + lane < cnt) {
        const uint32_t e = s_list[wave][e0 + lane];
        const int lb = (int)(e & 63u), k = (int)(e >> 6);
        int re = rb, te = tb + k;  // k < kPruneBatch: a few steps at most
        while (te >= W2) {
          te -= W2;
          ++re;
        }
        int j = 0, i = 0;
        window_block<kIntra>(re, te, S, bx0 + (lb & (tw - 1)), by0 + (lb >> tw_log), fbw, fbh, j, i);
        const int q = fbase + j * fbw + i;
        BlockEval be;  // the pixel side from LDS; eval_candidate converts the bytes itself
#pragma unroll
        for (int w = 0; w < 12; ++w) be.pc[w] = s_blk[wave][w][lb];
        be.sq = s_blk[wave][12][lb];
        be.own = s_blk[wave][13][lb];
        be.orig = (int)s_blk[wave][14][lb];
        const float4 t4 = tabs[q];
        const CandResult res =
            eval_candidate<false, true>(be, words[q], CandTab{t4.x, t4.y, t4.z, t4.w}, s_snap, s_ab);
        s_e[wave][k][lb] = res.valid ? sat16(res.err) : kNoCand;
        if (!kIntra && res.valid && res.err <= thr) s_pass[wave][lb] = 1;  // (equal writes may race)
        ++n_exact;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (act && kIntra) {  // intra: E and the passing set, for the fixpoint passes
      uint32_t bits = 0;  // bit k: position pb + k passes
      for (int k = 0; k < nb; k += 4) {
        const int16_t e0 = s_e[wave][k][lane], e1 = s_e[wave][k + 1][lane], e2 = s_e[wave][k + 2][lane],
                      e3 = s_e[wave][k + 3][lane];
        const uint32_t gb = (uint32_t)((e0 != kNoCand && e0 <= thr) | (e1 != kNoCand && e1 <= thr) << 1 |
                                       (e2 != kNoCand && e2 <= thr) << 2 | (e3 != kNoCand && e3 <= thr) << 3);
        if (gb) {  // E is kept only where a delta passes (the passing set says where)
          const uint32_t lo = (uint32_t)(uint16_t)e0 | ((uint32_t)(uint16_t)e1 << 16);
          const uint32_t hi = (uint32_t)(uint16_t)e2 | ((uint32_t)(uint16_t)e3 << 16);
          E4[(size_t)((pb + k) >> 2) * nbs + g] = make_uint2(lo, hi);
        }
        bits |= gb << k;
      }
      if (bits) {  // rare; the batch's 16 bits may straddle two words, which other chunks' waves share
        const int sh = pb & 31;
        uint32_t* wp = pm + (size_t)g * kPassWords + (pb >> 5);
        atomicOr(wp, bits << sh);
        if (sh > 16 && (bits >> (32 - sh)) != 0u) atomicOr(wp + 1, bits >> (32 - sh));
      }
    } else if (!kIntra && __ballot(act && s_pass[wave][lane] != 0) != 0ull) {
      // inter: fold the batch's passing deltas in scan order.  The selection
      // depends only on them (thr >= 0), and a row without one leaves the
      // summary as it is, so the other positions are skipped.
      if (act && s_pass[wave][lane] != 0) {
        ChunkFold cf;
        int cf_row;
        cf_load(s_cf[wave], lane, cf, cf_row);
        int rr = rb, tt = tb;
        for (int k = 0; k < nb; ++k) {
          const int v = s_e[wave][k][lane];
          if (v != kNoCand && v <= thr) {
            if (rr != cf_row) {
              cf.end_row(cf_row);
              cf.begin_row();
              cf_row = rr;
            }
            cf.step(v, pb + k, rr);
          }
          tt = tt + 1 == W2 ? 0 : tt + 1;
          rr += tt == 0 ? 1 : 0;
        }
        cf_store(s_cf[wave], lane, cf, cf_row);
      }
    }
    // the next batch overwrites s_e and s_list
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    rb = r;
    tb = t;
  }
  if (act && !kIntra) {
    ChunkFold cf;
    int cf_row;
    cf_load(s_cf[wave], lane, cf, cf_row);
    cf.end_row(cf_row);
    summ[(size_t)chunk * nbs + g] = cf.pack();
  }
  if (kIntra && fold) {
    // Intra decisions (as k_fold_intra), when every tile's chunks are waves of
    // this workgroup (n_chunks divides 4): once all of them have written
    // their passing sets and E, the chunk-0 wave of each tile folds its
    // blocks.  The workgroup's waves share the CU's L1, so the barrier
    // orders those writes before the reads.
    __syncthreads();
    bool changed = false;
    if (chunk == 0 && act) {
      const Fold fd = fold_passing_set(pm, E4, nbs, g, S);
      const uint64_t dc = intra_decision(words, fbase, fbw, bx, by, S, thr, fd);
      dec[g] = dc;
      const uint32_t w = dec_acc(dc) ? dec_word(dc) : (uint32_t)(own[g] >> 32);
      if (w != words[g]) {
        cw_new[g] = w;
        changed = true;
      }
    }
    wg_append(changed, g, chg_list, &st->n_changed[0], &s_base);
  }
  wg_count(&s_acc, n_bound, kIntra ? st->intra_cands : st->inter_cands);
  wg_count(&s_acc, n_exact, kIntra ? st->exact_intra : st->exact_inter);
}

// Intra decisions from the passing sets and E (as k_intra_combine): dec, the
// new current word and the first changed-word list.
__global__ void __launch_bounds__(256) k_fold_intra(const FrameDev* __restrict__ fr, const uint8_t* __restrict__ sid,
                                                    int total, int S, int thr, const uint64_t* __restrict__ own,
                                                    const uint8_t* __restrict__ need,
                                                    const uint32_t* __restrict__ words,
                                                    const uint2* __restrict__ E4, size_t nbs,
                                                    const uint32_t* __restrict__ pm,
                                                    uint64_t* __restrict__ dec, uint32_t* __restrict__ cw_new,
                                                    int* __restrict__ chg_list, DevStats* __restrict__ st) {
  __shared__ unsigned int s_base;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  bool changed = false;
  if (g < total && need[g]) {
    const FrameDev& d = fr[sid[g]];
    const int k = g - d.base, bx = k % d.bw, by = k / d.bw;
    const Fold fd = fold_passing_set(pm, E4, nbs, g, S);
    const uint64_t dc = intra_decision(words, d.base, d.bw, bx, by, S, thr, fd);
    dec[g] = dc;
    const uint32_t w = dec_acc(dc) ? dec_word(dc) : (uint32_t)(own[g] >> 32);
    if (w != words[g]) {
      cw_new[g] = w;
      changed = true;
    }
  }
  wg_append(changed, g, chg_list, &st->n_changed[0], &s_base);
}

}  // namespace mptc
