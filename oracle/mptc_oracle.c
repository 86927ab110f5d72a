/*
 * CPU oracle for the MPTC encode path -- TEST INFRASTRUCTURE ONLY (see
 * mptc_oracle.h).  Build with -O2 -ffp-contract=off and without -ffast-math:
 * the float re-fit must round every multiply and add separately, in the
 * reference's order, exactly as the x86-64 SSE build of the reference does.
 */
#include "mptc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* stb_dxt v1.04 restatement (Include/stb_dxt.h)                             */
/* ------------------------------------------------------------------------ */

static uint8_t g_expand5[32], g_expand6[64];
static uint8_t g_omatch5[256][2], g_omatch6[256][2];
static int g_init_done = 0;

/* stb_dxt.h:64-68 */
static int mul8bit(int a, int b) {
  int t = a * b + 128;
  return (t + (t >> 8)) >> 8;
}

/* stb_dxt.h:82-85 */
static uint16_t as16bit(int r, int g, int b) {
  return (uint16_t)((mul8bit(r, 31) << 11) + (mul8bit(g, 63) << 5) + mul8bit(b, 31));
}

/* stb_dxt.h:88-98, no rounding bias */
static int lerp13(int a, int b) { return (2 * a + b) / 3; }

/* stb_dxt.h:111-137: for every 8-bit target value, the first (mn, mx) pair in
 * mn-major order minimising |lerp13(e[mx], e[mn]) - i| + |e[mx] - e[mn]| * 3 / 100. */
static void prepare_opt_table(uint8_t table[256][2], const uint8_t* expand, int size) {
  for (int i = 0; i < 256; ++i) {
    int best = 256;
    for (int mn = 0; mn < size; ++mn)
      for (int mx = 0; mx < size; ++mx) {
        int e = abs(lerp13(expand[mx], expand[mn]) - i) + abs(expand[mx] - expand[mn]) * 3 / 100;
        if (e < best) {
          table[i][0] = (uint8_t)mx;
          table[i][1] = (uint8_t)mn;
          best = e;
        }
      }
  }
}

/* stb_dxt.h:593-611 (the dithering quantisers are unused: alpha=0, no dither) */
void orc_init(void) {
  if (g_init_done) return;
  for (int i = 0; i < 32; ++i) g_expand5[i] = (uint8_t)((i << 3) | (i >> 2));
  for (int i = 0; i < 64; ++i) g_expand6[i] = (uint8_t)((i << 2) | (i >> 4));
  prepare_opt_table(g_omatch5, g_expand5, 32);
  prepare_opt_table(g_omatch6, g_expand6, 64);
  g_init_done = 1;
}

/* stb_dxt.h:70-80 + 139-145: colour[k] = palette entry k as RGB */
static void eval_colors(int color[4][3], uint16_t c0, uint16_t c1) {
  color[0][0] = g_expand5[(c0 >> 11) & 31]; color[0][1] = g_expand6[(c0 >> 5) & 63]; color[0][2] = g_expand5[c0 & 31];
  color[1][0] = g_expand5[(c1 >> 11) & 31]; color[1][1] = g_expand6[(c1 >> 5) & 63]; color[1][2] = g_expand5[c1 & 31];
  for (int ch = 0; ch < 3; ++ch) {
    color[2][ch] = lerp13(color[0][ch], color[1][ch]);
    color[3][ch] = lerp13(color[1][ch], color[0][ch]);
  }
}

/* stb_dxt.h:176-215, dither = 0 */
static uint32_t match_colors(const uint8_t px[16][3], int color[4][3]) {
  const int dr = color[0][0] - color[1][0], dg = color[0][1] - color[1][1], db = color[0][2] - color[1][2];
  int stops[4];
  for (int k = 0; k < 4; ++k) stops[k] = color[k][0] * dr + color[k][1] * dg + color[k][2] * db;
  const int c0pt = (stops[1] + stops[3]) >> 1;
  const int half = (stops[3] + stops[2]) >> 1;
  const int c3pt = (stops[2] + stops[0]) >> 1;
  uint32_t mask = 0;
  for (int i = 15; i >= 0; --i) {
    const int dot = px[i][0] * dr + px[i][1] * dg + px[i][2] * db;
    mask <<= 2;
    if (dot < half)
      mask |= (dot < c0pt) ? 1u : 3u;
    else
      mask |= (dot < c3pt) ? 2u : 0u;
  }
  return mask;
}

/* stb_dxt.h:273-375 */
static void optimize_colors(const uint8_t px[16][3], uint16_t* pmax16, uint16_t* pmin16) {
  int mu[3], mn[3], mx[3];
  for (int ch = 0; ch < 3; ++ch) {
    int s = px[0][ch], lo = px[0][ch], hi = px[0][ch];
    for (int i = 1; i < 16; ++i) {
      s += px[i][ch];
      if (px[i][ch] < lo) lo = px[i][ch];
      else if (px[i][ch] > hi) hi = px[i][ch];
    }
    mu[ch] = (s + 8) >> 4;
    mn[ch] = lo;
    mx[ch] = hi;
  }
  int cov[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 16; ++i) {
    const int r = px[i][0] - mu[0], g = px[i][1] - mu[1], b = px[i][2] - mu[2];
    cov[0] += r * r; cov[1] += r * g; cov[2] += r * b;
    cov[3] += g * g; cov[4] += g * b; cov[5] += b * b;
  }
  float covf[6];
  for (int k = 0; k < 6; ++k) covf[k] = (float)cov[k] / 255.0f;
  float vr = (float)(mx[0] - mn[0]);
  float vg = (float)(mx[1] - mn[1]);
  float vb = (float)(mx[2] - mn[2]);
  for (int it = 0; it < 4; ++it) {
    /* each product rounded, left-to-right sums, no fused multiply-add */
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
  int mind = 0x7fffffff, maxd = -0x7fffffff, imin = 0, imax = 0;
  for (int i = 0; i < 16; ++i) {
    const int dot = px[i][0] * v_r + px[i][1] * v_g + px[i][2] * v_b;
    if (dot < mind) { mind = dot; imin = i; }
    if (dot > maxd) { maxd = dot; imax = i; }
  }
  *pmax16 = as16bit(px[imax][0], px[imax][1], px[imax][2]);
  *pmin16 = as16bit(px[imin][0], px[imin][1], px[imin][2]);
}

/* stb_dxt.h:377-383 (in-range truncation only: the argument is bounded here) */
static int sclamp(float y, int p0, int p1) {
  const int x = (int)y;
  return x < p0 ? p0 : x > p1 ? p1 : x;
}

/* stb_dxt.h:388-464 */
static int refine_block(const uint8_t px[16][3], uint16_t* pmax16, uint16_t* pmin16, uint32_t mask) {
  const uint16_t old_min = *pmin16, old_max = *pmax16;
  uint16_t min16, max16;
  if ((mask ^ (mask << 2)) < 4) {
    int r = 8, g = 8, b = 8;
    for (int i = 0; i < 16; ++i) { r += px[i][0]; g += px[i][1]; b += px[i][2]; }
    r >>= 4; g >>= 4; b >>= 4;
    max16 = (uint16_t)((g_omatch5[r][0] << 11) | (g_omatch6[g][0] << 5) | g_omatch5[b][0]);
    min16 = (uint16_t)((g_omatch5[r][1] << 11) | (g_omatch6[g][1] << 5) | g_omatch5[b][1]);
  } else {
    /* weight of endpoint 0 (x3) per 2-bit index and the packed normal-matrix terms */
    static const int w1tab[4] = {3, 0, 2, 1};
    static const int prods[4] = {0x090000, 0x000900, 0x040102, 0x010402};
    int akku = 0, a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
    uint32_t cm = mask;
    for (int i = 0; i < 16; ++i, cm >>= 2) {
      const int step = (int)(cm & 3);
      akku += prods[step];
      for (int ch = 0; ch < 3; ++ch) {
        a1[ch] += w1tab[step] * px[i][ch];
        a2[ch] += px[i][ch];
      }
    }
    for (int ch = 0; ch < 3; ++ch) a2[ch] = 3 * a2[ch] - a1[ch];
    const int xx = akku >> 16, yy = (akku >> 8) & 0xff, xy = akku & 0xff;
    const float frb = 3.0f * 31.0f / 255.0f / (float)(xx * yy - xy * xy);
    const float fg = frb * 63.0f / 31.0f;
    max16 = (uint16_t)(sclamp((float)(a1[0] * yy - a2[0] * xy) * frb + 0.5f, 0, 31) << 11);
    max16 |= (uint16_t)(sclamp((float)(a1[1] * yy - a2[1] * xy) * fg + 0.5f, 0, 63) << 5);
    max16 |= (uint16_t)(sclamp((float)(a1[2] * yy - a2[2] * xy) * frb + 0.5f, 0, 31));
    min16 = (uint16_t)(sclamp((float)(a2[0] * xx - a1[0] * xy) * frb + 0.5f, 0, 31) << 11);
    min16 |= (uint16_t)(sclamp((float)(a2[1] * xx - a1[1] * xy) * fg + 0.5f, 0, 63) << 5);
    min16 |= (uint16_t)(sclamp((float)(a2[2] * xx - a1[2] * xy) * frb + 0.5f, 0, 31));
  }
  *pmin16 = min16;
  *pmax16 = max16;
  return old_min != min16 || old_max != max16;
}

/* stb_dxt.h:467-538 with mode = HIGHQUAL (two refinement passes), alpha = 0,
 * reached through CompressRGB (codec/dxt_image.cpp:123-149), which sets A = 0xFF. */
uint64_t orc_fit_block(const uint8_t in[48]) {
  uint8_t px[16][3];
  memcpy(px, in, 48);
  uint16_t max16, min16;
  uint32_t mask;
  int constant = 1;
  for (int i = 1; i < 16 && constant; ++i)
    constant = px[i][0] == px[0][0] && px[i][1] == px[0][1] && px[i][2] == px[0][2];
  if (constant) {
    const int r = px[0][0], g = px[0][1], b = px[0][2];
    mask = 0xaaaaaaaau;
    max16 = (uint16_t)((g_omatch5[r][0] << 11) | (g_omatch6[g][0] << 5) | g_omatch5[b][0]);
    min16 = (uint16_t)((g_omatch5[r][1] << 11) | (g_omatch6[g][1] << 5) | g_omatch5[b][1]);
  } else {
    int color[4][3];
    optimize_colors(px, &max16, &min16);
    if (max16 != min16) {
      eval_colors(color, max16, min16);
      mask = match_colors(px, color);
    } else {
      mask = 0;
    }
    for (int pass = 0; pass < 2; ++pass) {
      const uint32_t last = mask;
      if (refine_block(px, &max16, &min16, mask)) {
        if (max16 != min16) {
          eval_colors(color, max16, min16);
          mask = match_colors(px, color);
        } else {
          mask = 0;
          break;
        }
      }
      if (mask == last) break;
    }
  }
  if (max16 < min16) {
    const uint16_t t = min16;
    min16 = max16;
    max16 = t;
    mask ^= 0x55555555u;
  }
  /* dest bytes: max16 LE, min16 LE, mask LE == PhysicalDXTBlock{ep1, ep2, interp} */
  return (uint64_t)max16 | ((uint64_t)min16 << 16) | ((uint64_t)mask << 32);
}

/* codec/dxt_image.cpp:636-650 */
void orc_gather(const uint8_t* rgb, int w, int bx, int by, uint8_t px[48]) {
  for (int r = 0; r < 4; ++r)
    memcpy(px + 12 * r, rgb + ((size_t)(4 * by + r) * w + 4 * bx) * 3, 12);
}

/* codec/dxt_image.cpp:419-431 */
void orc_fit_frame(const uint8_t* rgb, int w, int h, uint64_t* phys) {
  orc_init();
  const int bw = (w + 3) >> 2, bh = (h + 3) >> 2;
  uint8_t px[48];
  for (int by = 0; by < bh; ++by)
    for (int bx = 0; bx < bw; ++bx) {
      orc_gather(rgb, w, bx, by, px);
      phys[(size_t)by * bw + bx] = orc_fit_block(px);
    }
}

/* ------------------------------------------------------------------------ */
/* Block conversions (codec/dxt_image.cpp:34-227)                            */
/* ------------------------------------------------------------------------ */

/* :34-41 */
static void lerp_channels(const uint8_t a[4], const uint8_t b[4], uint8_t out[4], int num, int div) {
  for (int i = 0; i < 3; ++i) out[i] = (uint8_t)(((int)a[i] * (div - num) + (int)b[i] * num) / div);
  out[3] = 0xFF;
}

/* :43-61 */
static void decode565(uint16_t x, uint8_t out[4]) {
  const uint32_t r = x >> 11, g = (x >> 5) & 0x3F, b = x & 0x1F;
  out[0] = (uint8_t)((r << 3) | (r >> 2));
  out[1] = (uint8_t)((g << 2) | (g >> 4));
  out[2] = (uint8_t)((b << 3) | (b >> 2));
  out[3] = 255;
}

/* :63-69 */
static uint16_t pack565(const uint8_t in[3]) {
  return (uint16_t)(((uint16_t)(in[0] & 0xF8) << 8) | ((uint16_t)(in[1] & 0xFC) << 3) |
                    ((uint16_t)(in[2] & 0xF8) >> 3));
}

static uint8_t absdiff8(uint8_t a, uint8_t b) { return (uint8_t)(a > b ? a - b : b - a); }

/* :72-95 and :98-121: snap to a 5-/6-bit grid value, kept as the 8-bit
 * "base | base >> shift" form (not necessarily the 565 expansion). */
static uint8_t snap_bits(uint8_t x, uint8_t keep, uint8_t half, int shift) {
  uint8_t base = (uint8_t)(x & keep);
  uint8_t high = (uint8_t)(base + half);
  uint8_t low = base == 0 ? base : (uint8_t)(base - half);
  base = (uint8_t)(base | (base >> shift));
  high = (uint8_t)(high | (high >> shift));
  low = (uint8_t)(low | (low >> shift));
  const uint8_t db = absdiff8(x, base), dh = absdiff8(x, high), dl = absdiff8(x, low);
  if (db <= dh) return db < dl ? base : low;
  return high;
}
uint8_t orc_to_five_bits(uint8_t x) { return snap_bits(x, 0xF8, 4, 5); }
uint8_t orc_to_six_bits(uint8_t x) { return snap_bits(x, 0xFC, 2, 6); }

/* :154-196 */
uint64_t orc_l2p(const orc_logical* b) {
  uint16_t p0 = pack565(b->ep1), p1 = pack565(b->ep2);
  const int swap = (p0 > p1 && b->pal[3][3] == 0) || (p0 <= p1 && b->pal[3][3] == 255);
  if (swap) { const uint16_t t = p0; p0 = p1; p1 = t; }
  uint32_t interp = 0;
  for (int i = 15; i >= 0; --i) interp = (interp << 2) | (b->idx[i] & 3u);
  if (swap) interp ^= 0x55555555u;
  return (uint64_t)p0 | ((uint64_t)p1 << 16) | ((uint64_t)interp << 32);
}

/* :198-227 */
void orc_p2l(uint64_t phys, orc_logical* out) {
  const uint16_t e1 = (uint16_t)(phys & 0xFFFF), e2 = (uint16_t)((phys >> 16) & 0xFFFF);
  const uint32_t interp = (uint32_t)(phys >> 32);
  decode565(e1, out->ep1);
  decode565(e2, out->ep2);
  memcpy(out->pal[0], out->ep1, 4);
  memcpy(out->pal[1], out->ep2, 4);
  if (e1 <= e2) {
    lerp_channels(out->ep1, out->ep2, out->pal[2], 1, 2);
    memset(out->pal[3], 0, 4);
  } else {
    lerp_channels(out->ep1, out->ep2, out->pal[2], 1, 3);
    lerp_channels(out->ep1, out->ep2, out->pal[3], 2, 3);
  }
  for (int i = 0; i < 16; ++i) out->idx[i] = (uint8_t)((interp >> (2 * i)) & 3u);
}

/* :258-276 */
int orc_error(const orc_logical* b, const uint8_t px[48]) {
  size_t err = 0;
  for (int i = 0; i < 16; ++i) {
    const uint8_t* q = b->pal[b->idx[i]];
    for (int ch = 0; ch < 3; ++ch) {
      const size_t d = absdiff8(px[3 * i + ch], q[ch]);
      err += d * d;
    }
  }
  return (int)(err / 48);
}

/* x86 cvttss2si: NaN and out-of-range values give INT32_MIN */
static int32_t cvtt_x86(float v) {
  if (!(v >= -2147483648.0f && v < 2147483648.0f)) return INT32_MIN;
  return (int32_t)v;
}

static int clamp255(int32_t v) { return v < 0 ? 0 : v > 255 ? 255 : v; }

/* :290-351 */
void orc_refit(orc_logical* b, const uint8_t px[48]) {
  static const float order_of[4] = {0.f, 3.f, 1.f, 2.f};
  float asq = 0.0f, bsq = 0.0f, ab = 0.0f;
  float ax[3] = {0.0f, 0.0f, 0.0f}, bx[3] = {0.0f, 0.0f, 0.0f};
  for (int i = 0; i < 16; ++i) {
    const float order = order_of[b->idx[i]];
    const float a = (3.0f - order) / 3.0f;
    const float bb = order / 3.0f;
    asq += a * a;
    bsq += bb * bb;
    ab += a * bb;
    for (int ch = 0; ch < 3; ++ch) {
      ax[ch] += (float)px[3 * i + ch] * a;
      bx[ch] += (float)px[3 * i + ch] * bb;
    }
  }
  const float f = 1.0f / (asq * bsq - ab * ab);
  for (int ch = 0; ch < 3; ++ch) {
    const float p1 = f * (ax[ch] * bsq - bx[ch] * ab);
    const float p2 = f * (bx[ch] * asq - ax[ch] * ab);
    b->ep1[ch] = (uint8_t)clamp255(cvtt_x86(p1 + 0.5f));
    b->ep2[ch] = (uint8_t)clamp255(cvtt_x86(p2 + 0.5f));
  }
  b->ep1[0] = orc_to_five_bits(b->ep1[0]);
  b->ep2[0] = orc_to_five_bits(b->ep2[0]);
  b->ep1[1] = orc_to_six_bits(b->ep1[1]);
  b->ep2[1] = orc_to_six_bits(b->ep2[1]);
  b->ep1[2] = orc_to_five_bits(b->ep1[2]);
  b->ep2[2] = orc_to_five_bits(b->ep2[2]);
  memcpy(b->pal[0], b->ep1, 4);
  memcpy(b->pal[1], b->ep2, 4);
  lerp_channels(b->ep1, b->ep2, b->pal[2], 1, 3);
  lerp_channels(b->ep1, b->ep2, b->pal[3], 2, 3);
}

/* AssignIndices (:285-289): logical of (L2P(own) with interp := c) */
static void assign_indices(orc_logical* b, uint32_t c) {
  const uint64_t p = orc_l2p(b);
  orc_p2l((p & 0xFFFFFFFFull) | ((uint64_t)c << 32), b);
}

/* The per-candidate body shared by both searches (:677-698, :734-758), given
 * the block's own logical block and its error. */
static int eval_with_own(const orc_logical* own, int orig, const uint8_t px[48], uint32_t cand,
                         int* err_diff, int* re_assigned, uint8_t eps_out[6]) {
  orc_logical b2 = *own;
  assign_indices(&b2, cand);
  int reset = 0;
  if (memcmp(own, &b2, sizeof(b2)) != 0) {
    orc_refit(&b2, px);
    reset = 1;
    const uint64_t maybe = orc_l2p(&b2);
    if (!((uint32_t)(maybe >> 32) == cand && b2.pal[3][3] == 0xFF)) return 0;
  }
  *err_diff = orc_error(&b2, px) - orig;
  *re_assigned = reset;
  if (eps_out) {
    memcpy(eps_out, b2.ep1, 3);
    memcpy(eps_out + 3, b2.ep2, 3);
  }
  return 1;
}

int orc_eval_candidate(uint64_t own_phys, const uint8_t px[48], uint32_t cand, int* err_diff,
                       int* re_assigned, uint8_t eps_out[6]) {
  orc_logical own;
  orc_p2l(own_phys, &own);
  return eval_with_own(&own, orc_error(&own, px), px, cand, err_diff, re_assigned, eps_out);
}

typedef struct {
  int ret;
  int sel_x, sel_y; /* offsets as written to the motion bytes (before the inter flag) */
  uint32_t word;
  int re_assigned;
} search_result;

/* IntraSearch (:652-713) when intra != 0, InterBlockSearch (:715-774) otherwise. */
static search_result search_block(const uint8_t* rgb, int w, int bw, int bh, int S, int intra,
                                  const uint64_t* words_src, const uint64_t* own_phys, int bx,
                                  int by) {
  search_result r;
  r.ret = 0x7fffffff;
  r.sel_x = r.sel_y = 0;
  r.word = 0;
  r.re_assigned = 0;
  uint8_t px[48];
  orc_gather(rgb, w, bx, by, px);
  orc_logical own;
  orc_p2l(own_phys[(size_t)by * bw + bx], &own);
  const int orig = orc_error(&own, px);
  int j0, j1, dj, i0, i1, di;
  if (intra) {
    j0 = by; j1 = by - 2 * S; dj = -1;          /* j from by down to by-2S+1 */
    i0 = bx + S - 1; i1 = bx - S - 1; di = -1;  /* i from bx+S-1 down to bx-S */
  } else {
    j0 = by - S; j1 = by + S; dj = 1;           /* j from by-S up to by+S-1 */
    i0 = bx - S; i1 = bx + S; di = 1;           /* i from bx-S up to bx+S-1 */
  }
  for (int j = j0; j != j1; j += dj) {
    for (int i = i0; i != i1; i += di) {
      if (i >= bw || j >= bh || j < 0 || i < 0) continue;
      if (intra && j == by && i >= bx) continue;
      const uint32_t cand = (uint32_t)(words_src[(size_t)j * bw + i] >> 32);
      int e, re;
      if (!eval_with_own(&own, orig, px, cand, &e, &re, NULL)) continue;
      if (e < r.ret) {
        r.ret = e;
        r.sel_x = (i - bx) + S;
        r.sel_y = intra ? (j - by) + 2 * S - 1 : (j - by) + S;
        r.word = cand;
        r.re_assigned = re;
        if (e <= 0) {
          r.ret = 0;
          break; /* leaves the column loop only */
        }
      }
    }
  }
  return r;
}

/* Reencode (:868-957); the commit mirrors :900-905 / :922-926 / :953-954. */
int orc_reencode(const uint8_t* rgb, int w, int h, int S, int thr, int is_intra,
                 const uint64_t* ref_phys, uint64_t* phys, uint8_t* motion, uint32_t* unique,
                 uint8_t* kind, uint8_t* eps) {
  orc_init();
  const int bw = (w + 3) >> 2, bh = (h + 3) >> 2, nb = bw * bh;
  /* the own (stb) blocks stay readable while phys is re-written in place */
  uint64_t* own = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nb);
  memcpy(own, phys, sizeof(uint64_t) * (size_t)nb);
  int n_unique = 0;
  uint8_t px[48];
  for (int k = 0; k < nb; ++k) {
    const int bx = k % bw, by = k / bw;
    search_result r;
    int accepted = 0, k_kind = 0, mx = 255, my = 255;
    if (!is_intra) {
      r = search_block(rgb, w, bw, bh, S, 0, ref_phys, own, bx, by);
      if (r.ret <= thr) {
        accepted = 1;
        k_kind = 2;
        mx = r.sel_x | 0x80;
        my = r.sel_y | 0x80;
      }
    }
    if (!accepted) {
      /* intra candidates are this frame's current (already re-encoded) words */
      r = search_block(rgb, w, bw, bh, S, 1, phys, own, bx, by);
      if (r.ret <= thr) {
        accepted = 1;
        k_kind = 1;
        mx = r.sel_x;
        my = r.sel_y;
      }
    }
    orc_logical lb;
    orc_p2l(own[k], &lb);
    if (accepted) {
      assign_indices(&lb, r.word);
      if (r.re_assigned) {
        orc_gather(rgb, w, bx, by, px);
        orc_refit(&lb, px);
      }
      phys[k] = orc_l2p(&lb);
    } else {
      unique[n_unique++] = (uint32_t)(own[k] >> 32);
    }
    motion[2 * k] = (uint8_t)mx;
    motion[2 * k + 1] = (uint8_t)my;
    if (kind) kind[k] = (uint8_t)k_kind;
    if (eps) {
      memcpy(eps + 6 * k, lb.ep1, 3);
      memcpy(eps + 6 * k + 3, lb.ep2, 3);
    }
  }
  free(own);
  return n_unique;
}

static double psnr_from_sse(double sse, int w, int h) {
  const double mse = sse / (double)(w * h);
  return 10.0 * log10((3.0 * 255.0 * 255.0) / mse);
}

double orc_psnr_logical(const uint8_t* rgb, int w, int h, const uint64_t* phys, const uint8_t* eps) {
  const int bw = (w + 3) >> 2, bh = (h + 3) >> 2;
  double sse = 0.0;
  for (int by = 0; by < bh; ++by)
    for (int bx = 0; bx < bw; ++bx) {
      const size_t k = (size_t)by * bw + bx;
      orc_logical lb;
      orc_p2l(phys[k], &lb);
      if ((phys[k] & 0xFFFF) > ((phys[k] >> 16) & 0xFFFF)) {
        memcpy(lb.ep1, eps + 6 * k, 3);
        memcpy(lb.ep2, eps + 6 * k + 3, 3);
        memcpy(lb.pal[0], lb.ep1, 3);
        memcpy(lb.pal[1], lb.ep2, 3);
        lerp_channels(lb.ep1, lb.ep2, lb.pal[2], 1, 3);
        lerp_channels(lb.ep1, lb.ep2, lb.pal[3], 2, 3);
      }
      uint8_t px[48];
      orc_gather(rgb, w, bx, by, px);
      for (int i = 0; i < 16; ++i)
        for (int ch = 0; ch < 3; ++ch) {
          const double d = (double)px[3 * i + ch] - (double)lb.pal[lb.idx[i]][ch];
          sse += d * d;
        }
    }
  return psnr_from_sse(sse, w, h);
}

double orc_psnr_physical(const uint8_t* rgb, int w, int h, const uint64_t* phys) {
  const int bw = (w + 3) >> 2, bh = (h + 3) >> 2;
  double sse = 0.0;
  for (int by = 0; by < bh; ++by)
    for (int bx = 0; bx < bw; ++bx) {
      orc_logical lb;
      orc_p2l(phys[(size_t)by * bw + bx], &lb);
      uint8_t px[48];
      orc_gather(rgb, w, bx, by, px);
      for (int i = 0; i < 16; ++i)
        for (int ch = 0; ch < 3; ++ch) {
          const double d = (double)px[3 * i + ch] - (double)lb.pal[lb.idx[i]][ch];
          sse += d * d;
        }
    }
  return psnr_from_sse(sse, w, h);
}

/* The shared seeded frame generator, exported for the tests. */
#include "../mptc_amd/csrc/framegen.h"
void orc_synth_frame(const mptc_synth_cfg* c, int t, uint8_t* rgb) { mptc_synth_frame(c, t, rgb); }
