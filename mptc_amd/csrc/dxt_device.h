// Device-side DXT1 primitives for gfx950.  Every routine reproduces the
// reference's integer and IEEE-754 binary32/binary64 arithmetic exactly: the
// translation unit is compiled with -ffp-contract=off (no fused multiply-add)
// and correctly rounded fp32 division, so each multiply, add and divide is
// rounded separately and in the reference's order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mptc {

// ---------------------------------------------------------------------------
// Tables (host-computed once per device by init_tables in mptc_hip.hip)
// ---------------------------------------------------------------------------
// stb__OMatch5 / stb__OMatch6 (Include/stb_dxt.h:111-137): [value][0] = mx, [1] = mn
__constant__ uint8_t c_omatch5[256][2];
__constant__ uint8_t c_omatch6[256][2];
// ToFiveBits / ToSixBits (codec/dxt_image.cpp:72-121) as 256-entry tables:
// [0..255] five-bit snap, [256..511] six-bit snap.
__constant__ uint8_t c_snap[512];

// 2/3 and 1/3 as float: the values of (3-order)/3.0f and order/3.0f at
// codec/dxt_image.cpp:307-308.  3 * (1/3 rounded) rounds back to 1.0f, so
// weight w in {0,1,2,3} times kThird gives exactly {0, 1/3, 2/3, 1}.
constexpr float kThird = 1.0f / 3.0f;

__device__ __forceinline__ uint32_t expand5(uint32_t v) { return (v << 3) | (v >> 2); }
__device__ __forceinline__ uint32_t expand6(uint32_t v) { return (v << 2) | (v >> 4); }

// Pack565 (codec/dxt_image.cpp:63-69) of an (r, g, b) triple.
__device__ __forceinline__ uint32_t pack565(uint32_t r, uint32_t g, uint32_t b) {
  return ((r & 0xF8u) << 8) | ((g & 0xFCu) << 3) | (b >> 3);
}

// Unsigned x / 3 for x < 98304 (palette lerps of 8-bit values).
__device__ __forceinline__ uint32_t div3(uint32_t x) { return (x * 0xAAABu) >> 17; }

// ---------------------------------------------------------------------------
// One 4x4 block held by one lane.
// ---------------------------------------------------------------------------
struct BlockPx {
  uint32_t w[12];  // 48 RGB bytes, row-major, little-endian words
  __device__ __forceinline__ uint32_t ch(int i, int c) const {  // pixel i, channel c
    const int byte = 3 * i + c;
    return (w[byte >> 2] >> (8 * (byte & 3))) & 0xFFu;
  }
};

// Loads block (bx, by) of an RGB8 frame of width `width` (multiple of 4, so
// each 12-byte block row is 4-byte aligned).
__device__ __forceinline__ void load_block(const uint8_t* __restrict__ rgb, int width, int bx, int by,
                                           BlockPx& p) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint32_t* row = reinterpret_cast<const uint32_t*>(rgb + ((size_t)(4 * by + r) * width + 4 * bx) * 3);
    p.w[3 * r + 0] = row[0];
    p.w[3 * r + 1] = row[1];
    p.w[3 * r + 2] = row[2];
  }
}

// ---------------------------------------------------------------------------
// stb_dxt v1.04 HIGHQUAL color-block fit (Include/stb_dxt.h:273-538), as
// called by CompressRGB (codec/dxt_image.cpp:123-149).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int mul8bit(int a, int b) {
  const int t = a * b + 128;
  return (t + (t >> 8)) >> 8;
}
__device__ __forceinline__ uint32_t as16bit(int r, int g, int b) {
  return (uint32_t)((mul8bit(r, 31) << 11) + (mul8bit(g, 63) << 5) + mul8bit(b, 31));
}
__device__ __forceinline__ int lerp13(int a, int b) { return (2 * a + b) / 3; }

struct Pal4 {
  int c[4][3];
};

__device__ __forceinline__ void eval_colors(Pal4& p, uint32_t c0, uint32_t c1) {
  p.c[0][0] = (int)expand5((c0 >> 11) & 31); p.c[0][1] = (int)expand6((c0 >> 5) & 63); p.c[0][2] = (int)expand5(c0 & 31);
  p.c[1][0] = (int)expand5((c1 >> 11) & 31); p.c[1][1] = (int)expand6((c1 >> 5) & 63); p.c[1][2] = (int)expand5(c1 & 31);
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    p.c[2][ch] = lerp13(p.c[0][ch], p.c[1][ch]);
    p.c[3][ch] = lerp13(p.c[1][ch], p.c[0][ch]);
  }
}

// stb__MatchColorsBlock without dithering (stb_dxt.h:176-215)
__device__ __forceinline__ uint32_t match_colors(const BlockPx& px, const Pal4& p) {
  const int dr = p.c[0][0] - p.c[1][0], dg = p.c[0][1] - p.c[1][1], db = p.c[0][2] - p.c[1][2];
  int stops[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) stops[k] = p.c[k][0] * dr + p.c[k][1] * dg + p.c[k][2] * db;
  const int c0pt = (stops[1] + stops[3]) >> 1;
  const int half = (stops[3] + stops[2]) >> 1;
  const int c3pt = (stops[2] + stops[0]) >> 1;
  uint32_t mask = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int dot = (int)px.ch(i, 0) * dr + (int)px.ch(i, 1) * dg + (int)px.ch(i, 2) * db;
    const uint32_t v = dot < half ? (dot < c0pt ? 1u : 3u) : (dot < c3pt ? 2u : 0u);
    mask |= v << (2 * i);
  }
  return mask;
}

// stb__OptimizeColorsBlock (stb_dxt.h:273-375)
__device__ __forceinline__ void optimize_colors(const BlockPx& px, uint32_t& max16, uint32_t& min16) {
  int mu[3], mn[3], mx[3];
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    int s = (int)px.ch(0, ch), lo = s, hi = s;
#pragma unroll
    for (int i = 1; i < 16; ++i) {
      const int v = (int)px.ch(i, ch);
      s += v;
      lo = v < lo ? v : lo;
      hi = v > hi ? v : hi;
    }
    mu[ch] = (s + 8) >> 4;
    mn[ch] = lo;
    mx[ch] = hi;
  }
  int cov[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int r = (int)px.ch(i, 0) - mu[0], g = (int)px.ch(i, 1) - mu[1], b = (int)px.ch(i, 2) - mu[2];
    cov[0] += r * r; cov[1] += r * g; cov[2] += r * b;
    cov[3] += g * g; cov[4] += g * b; cov[5] += b * b;
  }
  float covf[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) covf[k] = (float)cov[k] / 255.0f;
  float vr = (float)(mx[0] - mn[0]);
  float vg = (float)(mx[1] - mn[1]);
  float vb = (float)(mx[2] - mn[2]);
#pragma unroll
  for (int it = 0; it < 4; ++it) {
    const float r = vr * covf[0] + vg * covf[1] + vb * covf[2];
    const float g = vr * covf[1] + vg * covf[3] + vb * covf[4];
    const float b = vr * covf[2] + vg * covf[4] + vb * covf[5];
    vr = r; vg = g; vb = b;
  }
  double magn = fabs((double)vr);
  if (fabs((double)vg) > magn) magn = fabs((double)vg);
  if (fabs((double)vb) > magn) magn = fabs((double)vb);
  int v_r, v_g, v_b;
  if (magn < 4.0) {
    v_r = 299; v_g = 587; v_b = 114;
  } else {
    magn = 512.0 / magn;
    v_r = (int)((double)vr * magn);
    v_g = (int)((double)vg * magn);
    v_b = (int)((double)vb * magn);
  }
  // extreme points along the axis; the first pixel wins ties (strict < / >)
  int mind = 0x7fffffff, maxd = -0x7fffffff;
  uint32_t minc = 0, maxc = 0;  // packed r | g<<8 | b<<16 of the extreme pixels
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int r = (int)px.ch(i, 0), g = (int)px.ch(i, 1), b = (int)px.ch(i, 2);
    const int dot = r * v_r + g * v_g + b * v_b;
    const uint32_t pc = (uint32_t)r | ((uint32_t)g << 8) | ((uint32_t)b << 16);
    if (dot < mind) { mind = dot; minc = pc; }
    if (dot > maxd) { maxd = dot; maxc = pc; }
  }
  max16 = as16bit((int)(maxc & 0xFF), (int)((maxc >> 8) & 0xFF), (int)(maxc >> 16));
  min16 = as16bit((int)(minc & 0xFF), (int)((minc >> 8) & 0xFF), (int)(minc >> 16));
}

__device__ __forceinline__ int sclamp(float y, int p0, int p1) {
  const int x = (int)y;  // |y| is bounded by a few thousand here
  return x < p0 ? p0 : (x > p1 ? p1 : x);
}

// stb__RefineBlock (stb_dxt.h:388-464); returns true if the endpoints changed
__device__ __forceinline__ bool refine_block(const BlockPx& px, uint32_t& max16, uint32_t& min16, uint32_t mask) {
  const uint32_t old_min = min16, old_max = max16;
  uint32_t mn16, mx16;
  if ((mask ^ (mask << 2)) < 4u) {
    int r = 8, g = 8, b = 8;
#pragma unroll
    for (int i = 0; i < 16; ++i) { r += (int)px.ch(i, 0); g += (int)px.ch(i, 1); b += (int)px.ch(i, 2); }
    r >>= 4; g >>= 4; b >>= 4;
    mx16 = ((uint32_t)c_omatch5[r][0] << 11) | ((uint32_t)c_omatch6[g][0] << 5) | c_omatch5[b][0];
    mn16 = ((uint32_t)c_omatch5[r][1] << 11) | ((uint32_t)c_omatch6[g][1] << 5) | c_omatch5[b][1];
  } else {
    int akku = 0, a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const uint32_t step = (mask >> (2 * i)) & 3u;
      // w1Tab = {3,0,2,1}; prods = {0x090000,0x000900,0x040102,0x010402}
      const int w1 = step == 0 ? 3 : step == 1 ? 0 : step == 2 ? 2 : 1;
      const int pr = step == 0 ? 0x090000 : step == 1 ? 0x000900 : step == 2 ? 0x040102 : 0x010402;
      akku += pr;
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        const int v = (int)px.ch(i, ch);
        a1[ch] += w1 * v;
        a2[ch] += v;
      }
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) a2[ch] = 3 * a2[ch] - a1[ch];
    const int xx = akku >> 16, yy = (akku >> 8) & 0xff, xy = akku & 0xff;
    const float frb = 3.0f * 31.0f / 255.0f / (float)(xx * yy - xy * xy);
    const float fg = frb * 63.0f / 31.0f;
    mx16 = (uint32_t)sclamp((float)(a1[0] * yy - a2[0] * xy) * frb + 0.5f, 0, 31) << 11;
    mx16 |= (uint32_t)sclamp((float)(a1[1] * yy - a2[1] * xy) * fg + 0.5f, 0, 63) << 5;
    mx16 |= (uint32_t)sclamp((float)(a1[2] * yy - a2[2] * xy) * frb + 0.5f, 0, 31);
    mn16 = (uint32_t)sclamp((float)(a2[0] * xx - a1[0] * xy) * frb + 0.5f, 0, 31) << 11;
    mn16 |= (uint32_t)sclamp((float)(a2[1] * xx - a1[1] * xy) * fg + 0.5f, 0, 63) << 5;
    mn16 |= (uint32_t)sclamp((float)(a2[2] * xx - a1[2] * xy) * frb + 0.5f, 0, 31);
  }
  min16 = mn16;
  max16 = mx16;
  return old_min != mn16 || old_max != mx16;
}

// stb__CompressColorBlock (stb_dxt.h:467-538), HIGHQUAL: returns the physical
// block ep1 | ep2 << 16 | interp << 32 (PhysicalDXTBlock, codec/dxt_image.h:16-23).
__device__ __forceinline__ uint64_t fit_block(const BlockPx& px) {
  uint32_t max16, min16, mask;
  bool constant = true;
  {
    const uint32_t p0 = px.ch(0, 0) | (px.ch(0, 1) << 8) | (px.ch(0, 2) << 16);
#pragma unroll
    for (int i = 1; i < 16; ++i)
      constant = constant && ((px.ch(i, 0) | (px.ch(i, 1) << 8) | (px.ch(i, 2) << 16)) == p0);
  }
  if (constant) {
    const int r = (int)px.ch(0, 0), g = (int)px.ch(0, 1), b = (int)px.ch(0, 2);
    mask = 0xaaaaaaaau;
    max16 = ((uint32_t)c_omatch5[r][0] << 11) | ((uint32_t)c_omatch6[g][0] << 5) | c_omatch5[b][0];
    min16 = ((uint32_t)c_omatch5[r][1] << 11) | ((uint32_t)c_omatch6[g][1] << 5) | c_omatch5[b][1];
  } else {
    Pal4 pal;
    optimize_colors(px, max16, min16);
    if (max16 != min16) {
      eval_colors(pal, max16, min16);
      mask = match_colors(px, pal);
    } else {
      mask = 0;
    }
    for (int pass = 0; pass < 2; ++pass) {
      const uint32_t last = mask;
      if (refine_block(px, max16, min16, mask)) {
        if (max16 != min16) {
          eval_colors(pal, max16, min16);
          mask = match_colors(px, pal);
        } else {
          mask = 0;
          break;
        }
      }
      if (mask == last) break;
    }
  }
  if (max16 < min16) {
    const uint32_t t = min16;
    min16 = max16;
    max16 = t;
    mask ^= 0x55555555u;
  }
  return (uint64_t)max16 | ((uint64_t)min16 << 16) | ((uint64_t)mask << 32);
}

// ---------------------------------------------------------------------------
// Error of a physical block as logical block (PhysicalToLogical,
// codec/dxt_image.cpp:198-227, then CompressedBlock::Error :258-276).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void decode565(uint32_t x, uint32_t& r, uint32_t& g, uint32_t& b) {
  r = expand5(x >> 11);
  g = expand6((x >> 5) & 0x3F);
  b = expand5(x & 0x1F);
}

// Palette of a physical block packed as r | g<<8 | b<<16 per entry.
__device__ __forceinline__ void phys_palette(uint64_t phys, uint32_t pal[4]) {
  const uint32_t e1 = (uint32_t)(phys & 0xFFFF), e2 = (uint32_t)((phys >> 16) & 0xFFFF);
  uint32_t r1, g1, b1, r2, g2, b2;
  decode565(e1, r1, g1, b1);
  decode565(e2, r2, g2, b2);
  pal[0] = r1 | (g1 << 8) | (b1 << 16);
  pal[1] = r2 | (g2 << 8) | (b2 << 16);
  if (e1 <= e2) {
    pal[2] = ((r1 + r2) >> 1) | (((g1 + g2) >> 1) << 8) | (((b1 + b2) >> 1) << 16);
    pal[3] = 0;
  } else {
    pal[2] = div3(2 * r1 + r2) | (div3(2 * g1 + g2) << 8) | (div3(2 * b1 + b2) << 16);
    pal[3] = div3(r1 + 2 * r2) | (div3(g1 + 2 * g2) << 8) | (div3(b1 + 2 * b2) << 16);
  }
}

// Sum of squared RGB differences of a block against a 4-entry packed palette
// selected by a 2-bit index word.
__device__ __forceinline__ uint32_t block_sse(const BlockPx& px, uint32_t interp, const uint32_t pal[4]) {
  uint32_t sse = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const uint32_t k = (interp >> (2 * i)) & 3u;
    const uint32_t q = k == 0 ? pal[0] : k == 1 ? pal[1] : k == 2 ? pal[2] : pal[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const int d = (int)px.ch(i, ch) - (int)((q >> (8 * ch)) & 0xFF);
      sse += (uint32_t)(d * d);
    }
  }
  return sse;
}

// ---------------------------------------------------------------------------
// Candidate table entry: everything about an index word that does not
// depend on the block it is tried on (codec/dxt_image.cpp:295-320 index part).
// ---------------------------------------------------------------------------
struct CandTab {
  float asq, bsq, ab, f;  // the three index sums, f = 1/(asq*bsq - ab*ab)
};

__device__ __forceinline__ CandTab cand_table(uint32_t c) {
  float asq = 0.0f, bsq = 0.0f, ab = 0.0f;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const uint32_t k = (c >> (2 * i)) & 3u;
    // idx -> order {0,3,1,2}; a = (3-order)/3, b = order/3
    const uint32_t wa = k == 0 ? 3u : k == 1 ? 0u : k == 2 ? 2u : 1u;
    const float a = (float)wa * kThird;
    const float b = (float)(3u - wa) * kThird;
    asq += a * a;
    bsq += b * b;
    ab += a * b;
  }
  CandTab t;
  t.asq = asq;
  t.bsq = bsq;
  t.ab = ab;
  t.f = 1.0f / (asq * bsq - ab * ab);
  return t;
}

// Quantise p + 0.5f to [0, 255] with the reference's x86 truncation
// (cvttss2si returns INT32_MIN for NaN and |v| >= 2^31, which the clamp at
// codec/dxt_image.cpp:330-331 then maps to 0).
__device__ __forceinline__ uint32_t quant255(float p) {
  const float v = p + 0.5f;
  const float w = (v < 2147483648.0f) ? v : 0.0f;  // NaN and >= 2^31 -> 0
  const float cl = fminf(fmaxf(w, 0.0f), 255.0f);
  return (uint32_t)cl;
}

// quant255 when |p| < 2^31 is known: clamp then truncate.
__device__ __forceinline__ uint32_t quant255_small(float p) {
  return (uint32_t)__builtin_amdgcn_fmed3f(p + 0.5f, 0.0f, 255.0f);
}

// Candidates whose |f| is at most this produce |p| < 2^31 for any block
// (|ax|, |bx| <= 16*255 and asq, bsq, |ab| <= 16), so the x86 overflow case
// of quant255 cannot occur.
constexpr float kSmallF = 16384.0f;

// Per-block constants used by the candidate evaluation.
struct BlockEval {
  float f[48];       // pixel channels as float, pixel-major
  uint32_t pc[12];   // pc[3j + ch]: channel ch of pixels 4j..4j+3, one byte each
  uint32_t sq;       // sum of squared channel values
  uint32_t own;      // own (stb) index word
  int orig;          // Error() of the own block
};

// Error() of the block's own (fitted) logical block (codec/dxt_image.cpp:664-668).
__device__ __forceinline__ int own_error(const BlockPx& px, uint64_t own_phys) {
  uint32_t pal[4];
  phys_palette(own_phys, pal);
  return (int)(block_sse(px, (uint32_t)(own_phys >> 32), pal) / 48u);
}

// Pixel-side constants of a block; `orig` is own_error(px, own_phys).
__device__ __forceinline__ void make_block_eval(const BlockPx& px, uint64_t own_phys, int orig, BlockEval& be) {
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    be.f[3 * i + 0] = (float)px.ch(i, 0);
    be.f[3 * i + 1] = (float)px.ch(i, 1);
    be.f[3 * i + 2] = (float)px.ch(i, 2);
  }
  uint32_t sq = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const uint32_t v = px.ch(4 * j, ch) | (px.ch(4 * j + 1, ch) << 8) | (px.ch(4 * j + 2, ch) << 16) |
                         (px.ch(4 * j + 3, ch) << 24);
      be.pc[3 * j + ch] = v;
      sq = __builtin_amdgcn_udot4(v, v, sq, false);
    }
  be.sq = sq;
  be.own = (uint32_t)(own_phys >> 32);
  be.orig = orig;
}

__device__ __forceinline__ void make_block_eval(const BlockPx& px, uint64_t own_phys, BlockEval& be) {
  make_block_eval(px, own_phys, own_error(px, own_phys), be);
}

// Result of one candidate: valid flag, error delta, quirky re-fit endpoints.
struct CandResult {
  bool valid;
  int err;           // err(candidate) - err(own) in floor(SSE/48) units
  uint32_t e1, e2;   // packed r | g<<8 | b<<16 quirky logical endpoints
};

// Spreads the 2-bit indices of pixels 4j..4j+3 into one byte each.
__device__ __forceinline__ uint32_t spread_indices(uint32_t c, int j) {
  const uint32_t b = (c >> (8 * j)) & 0xFFu;
  const uint32_t x = (b | (b << 12)) & 0x000F000Fu;
  return (x | (x << 6)) & 0x03030303u;
}

// Byte gather: byte n of the result is byte sel.byte[n] (0..3) of `table`.
__device__ __forceinline__ uint32_t byte_gather(uint32_t table, uint32_t sel) {
  return __builtin_amdgcn_perm(0u, table, sel);
}

// The body of the search loops (codec/dxt_image.cpp:677-698): AssignIndices,
// RecalculateEndpoints (:290-351), the validity check (:689-694) and Error.
// `snap` is the 512-byte snap table (LDS copy).
// LSQ weights per 2-bit index, addressed by index*8 in LDS: (a, b) =
// ((3-order)/3, order/3) with order = {0,3,1,2}[idx] (codec/dxt_image.cpp:301-308).
__device__ __forceinline__ void init_ab_table(float2* s_ab) {
  if (threadIdx.x < 4) {
    const uint32_t k = threadIdx.x;
    const uint32_t wa = k == 0 ? 3u : k == 1 ? 0u : k == 2 ? 2u : 1u;
    s_ab[k] = make_float2((float)wa * kThird, (float)(3u - wa) * kThird);
  }
}

// kFromBytes: the pixel channels are converted from be.pc (the same exact
// float values) instead of read from be.f, which then needs no registers.
template <bool kWantEps, bool kFromBytes = false>
__device__ __forceinline__ CandResult eval_candidate(const BlockEval& be, uint32_t c, const CandTab& t,
                                                     const uint8_t* __restrict__ snap,
                                                     const float2* __restrict__ s_ab) {
  CandResult res;
  res.e1 = res.e2 = 0;
  if (c == be.own) {  // AssignIndices leaves the logical block unchanged
    res.valid = true;
    res.err = 0;
    return res;
  }
  uint32_t sel[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) sel[j] = spread_indices(c, j);
  // ---- least squares sums over the 16 pixels, in pixel order ----
  // 0 + x == x for the non-negative products, so the sums start at pixel 0.
  const char* ab_base = reinterpret_cast<const char*>(s_ab);
  float ax0, ax1, ax2, bx0, bx1, bx2;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t s8 = sel[j] << 3;  // byte offsets of the (a, b) entries
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int i = 4 * j + n;
      const float2 w = *reinterpret_cast<const float2*>(ab_base + ((s8 >> (8 * n)) & 0xFFu));
      const float p0 = kFromBytes ? (float)((be.pc[3 * j + 0] >> (8 * n)) & 0xFFu) : be.f[3 * i + 0];
      const float p1 = kFromBytes ? (float)((be.pc[3 * j + 1] >> (8 * n)) & 0xFFu) : be.f[3 * i + 1];
      const float p2 = kFromBytes ? (float)((be.pc[3 * j + 2] >> (8 * n)) & 0xFFu) : be.f[3 * i + 2];
      if (i == 0) {
        ax0 = p0 * w.x; ax1 = p1 * w.x; ax2 = p2 * w.x;
        bx0 = p0 * w.y; bx1 = p1 * w.y; bx2 = p2 * w.y;
      } else {
        ax0 += p0 * w.x; ax1 += p1 * w.x; ax2 += p2 * w.x;
        bx0 += p0 * w.y; bx1 += p1 * w.y; bx2 += p2 * w.y;
      }
    }
  }
  // ---- solve, quantise, snap ----
  const float s1r = t.f * (ax0 * t.bsq - bx0 * t.ab), s1g = t.f * (ax1 * t.bsq - bx1 * t.ab);
  const float s1b = t.f * (ax2 * t.bsq - bx2 * t.ab), s2r = t.f * (bx0 * t.asq - ax0 * t.ab);
  const float s2g = t.f * (bx1 * t.asq - ax1 * t.ab), s2b = t.f * (bx2 * t.asq - ax2 * t.ab);
  uint32_t r1, g1, b1, r2, g2, b2;
  if (fabsf(t.f) <= kSmallF) {
    r1 = snap[quant255_small(s1r)]; g1 = snap[256 + quant255_small(s1g)]; b1 = snap[quant255_small(s1b)];
    r2 = snap[quant255_small(s2r)]; g2 = snap[256 + quant255_small(s2g)]; b2 = snap[quant255_small(s2b)];
  } else {
    r1 = snap[quant255(s1r)]; g1 = snap[256 + quant255(s1g)]; b1 = snap[quant255(s1b)];
    r2 = snap[quant255(s2r)]; g2 = snap[256 + quant255(s2g)]; b2 = snap[quant255(s2b)];
  }
  // ---- validity: LogicalToPhysical keeps the indices iff Pack565(e1) > Pack565(e2) ----
  res.valid = pack565(r1, g1, b1) > pack565(r2, g2, b2);
  if (!res.valid) {
    res.err = 0;
    return res;
  }
  // ---- palette (LerpChannels 1/3, 2/3), one byte per entry per channel ----
  const uint32_t pr = r1 | (r2 << 8) | (div3(2 * r1 + r2) << 16) | (div3(r1 + 2 * r2) << 24);
  const uint32_t pg = g1 | (g2 << 8) | (div3(2 * g1 + g2) << 16) | (div3(g1 + 2 * g2) << 24);
  const uint32_t pb = b1 | (b2 << 8) | (div3(2 * b1 + b2) << 16) | (div3(b1 + 2 * b2) << 24);
  // SSE = sum|p|^2 - 2 sum p.q + sum|q|^2, all exact in 32-bit integers
  uint32_t dot = 0, qq = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t qr = byte_gather(pr, sel[j]);
    const uint32_t qg = byte_gather(pg, sel[j]);
    const uint32_t qb = byte_gather(pb, sel[j]);
    dot = __builtin_amdgcn_udot4(be.pc[3 * j + 0], qr, dot, false);
    dot = __builtin_amdgcn_udot4(be.pc[3 * j + 1], qg, dot, false);
    dot = __builtin_amdgcn_udot4(be.pc[3 * j + 2], qb, dot, false);
    qq = __builtin_amdgcn_udot4(qr, qr, qq, false);
    qq = __builtin_amdgcn_udot4(qg, qg, qq, false);
    qq = __builtin_amdgcn_udot4(qb, qb, qq, false);
  }
  const uint32_t sse = be.sq - 2u * dot + qq;
  res.err = (int)(sse / 48u) - be.orig;
  if (kWantEps) {
    res.e1 = r1 | (g1 << 8) | (b1 << 16);
    res.e2 = r2 | (g2 << 8) | (b2 << 16);
  }
  return res;
}

}  // namespace mptc
