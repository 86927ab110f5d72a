/*
 * Seeded synthetic video frames for the MPTC encode path.
 *
 * A frame is W x H interleaved RGB8 (row-major, 3 bytes per pixel, the
 * layout stbi_load(..., 3) hands to DXTImage at codec/dxt_image.cpp:396).
 * Content: a smooth 2-D colour gradient, axis-aligned rectangles of random
 * colour that translate every frame (wrapping at the frame edges), and
 * uniform integer noise in [-amp, amp], clamped to [0, 255].  A scene cut
 * (probability cut_permille/1000 per frame, never on frame 0) regenerates
 * the gradient; rectangles persist across cuts.
 *
 * Integer-only and counter-based (SplitMix64 finaliser over fixed keys), so
 * any frame can be produced independently, in any order, bit-identically on
 * every machine and compiler.  Header-only C so that the library, the CPU
 * restatement and the reference harness all share exactly one definition.
 */
#ifndef MPTC_FRAMEGEN_H
#define MPTC_FRAMEGEN_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t width;         /* pixels, multiple of 4 */
  int32_t height;        /* pixels, multiple of 4 */
  uint64_t seed;
  int32_t n_rects;       /* number of moving rectangles */
  int32_t vel_mode;      /* 0: fixed speed `vel` along one random axis direction;
                            1: each component uniform in [-vel, vel] */
  int32_t vel;
  int32_t noise_amp;     /* noise uniform in [-noise_amp, noise_amp] */
  int32_t cut_permille;  /* per-frame scene-cut probability, 1/1000 units */
} mptc_synth_cfg;

static inline uint64_t mptc_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Keyed hash of (seed, domain, a, b). */
static inline uint64_t mptc_hash4(uint64_t seed, uint64_t domain, uint64_t a, uint64_t b) {
  uint64_t h = mptc_mix64(seed ^ (domain * 0xD6E8FEB86659FD93ull));
  h = mptc_mix64(h ^ a);
  return mptc_mix64(h ^ (b + 0x632BE59BD9B4E019ull));
}

static inline uint32_t mptc_uniform(uint64_t h, uint32_t n) { return (uint32_t)(h % n); }

enum { MPTC_DOM_GRAD = 1, MPTC_DOM_RECT = 2, MPTC_DOM_CUT = 3, MPTC_DOM_NOISE = 4 };

/* Frame index of the gradient epoch that frame t belongs to. */
static inline int32_t mptc_synth_epoch(const mptc_synth_cfg* c, int32_t t) {
  for (int32_t f = t; f > 0; --f)
    if (mptc_uniform(mptc_hash4(c->seed, MPTC_DOM_CUT, (uint64_t)f, 0), 1000) <
        (uint32_t)c->cut_permille)
      return f;
  return 0;
}

static inline uint8_t mptc_clamp255(int32_t v) { return (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v); }

/* Writes frame t into rgb (width*height*3 bytes). */
static inline void mptc_synth_frame(const mptc_synth_cfg* c, int32_t t, uint8_t* rgb) {
  const int32_t W = c->width, H = c->height;
  const int32_t epoch = mptc_synth_epoch(c, t);
  int32_t base[3], gx[3], gy[3];
  for (int ch = 0; ch < 3; ++ch) {
    base[ch] = (int32_t)mptc_uniform(mptc_hash4(c->seed, MPTC_DOM_GRAD, (uint64_t)epoch, (uint64_t)(3 * ch + 0)), 256);
    gx[ch] = (int32_t)mptc_uniform(mptc_hash4(c->seed, MPTC_DOM_GRAD, (uint64_t)epoch, (uint64_t)(3 * ch + 1)), 257) - 128;
    gy[ch] = (int32_t)mptc_uniform(mptc_hash4(c->seed, MPTC_DOM_GRAD, (uint64_t)epoch, (uint64_t)(3 * ch + 2)), 257) - 128;
  }
  /* gradient */
  for (int32_t y = 0; y < H; ++y)
    for (int32_t x = 0; x < W; ++x) {
      uint8_t* p = rgb + ((size_t)y * W + x) * 3;
      for (int ch = 0; ch < 3; ++ch)
        p[ch] = mptc_clamp255(base[ch] + (gx[ch] * x) / W + (gy[ch] * y) / H);
    }
  /* rectangles, drawn in index order (later ones on top) */
  for (int32_t r = 0; r < c->n_rects; ++r) {
#define MPTC_RH(k) mptc_hash4(c->seed, MPTC_DOM_RECT, (uint64_t)r, (uint64_t)(k))
    const int32_t wmin = W / 16 > 1 ? W / 16 : 1, wmax = W / 4 > wmin ? W / 4 : wmin;
    const int32_t hmin = H / 16 > 1 ? H / 16 : 1, hmax = H / 4 > hmin ? H / 4 : hmin;
    const int32_t rw = wmin + (int32_t)mptc_uniform(MPTC_RH(0), (uint32_t)(wmax - wmin + 1));
    const int32_t rh = hmin + (int32_t)mptc_uniform(MPTC_RH(1), (uint32_t)(hmax - hmin + 1));
    const int32_t x0 = (int32_t)mptc_uniform(MPTC_RH(2), (uint32_t)W);
    const int32_t y0 = (int32_t)mptc_uniform(MPTC_RH(3), (uint32_t)H);
    uint8_t col[3];
    for (int ch = 0; ch < 3; ++ch) col[ch] = (uint8_t)mptc_uniform(MPTC_RH(4 + ch), 256);
    int32_t vx, vy;
    if (c->vel_mode == 0) {
      const uint32_t dir = mptc_uniform(MPTC_RH(7), 4);
      vx = dir == 0 ? c->vel : dir == 1 ? -c->vel : 0;
      vy = dir == 2 ? c->vel : dir == 3 ? -c->vel : 0;
    } else {
      vx = (int32_t)mptc_uniform(MPTC_RH(7), (uint32_t)(2 * c->vel + 1)) - c->vel;
      vy = (int32_t)mptc_uniform(MPTC_RH(8), (uint32_t)(2 * c->vel + 1)) - c->vel;
    }
#undef MPTC_RH
    const int64_t px = (((int64_t)x0 + (int64_t)vx * t) % W + W) % W;
    const int64_t py = (((int64_t)y0 + (int64_t)vy * t) % H + H) % H;
    for (int32_t dy = 0; dy < rh; ++dy) {
      const int64_t y = (py + dy) % H;
      for (int32_t dx = 0; dx < rw; ++dx) {
        const int64_t x = (px + dx) % W;
        uint8_t* p = rgb + ((size_t)y * W + (size_t)x) * 3;
        p[0] = col[0]; p[1] = col[1]; p[2] = col[2];
      }
    }
  }
  /* noise */
  if (c->noise_amp > 0) {
    const uint64_t key = mptc_mix64(c->seed ^ 0xA0761D6478BD642Full) ^ ((uint64_t)MPTC_DOM_NOISE << 56);
    const uint64_t n3 = (uint64_t)W * (uint64_t)H * 3u;
    const uint32_t span = (uint32_t)(2 * c->noise_amp + 1);
    for (uint64_t i = 0; i < n3; ++i) {
      const int32_t nz = (int32_t)mptc_uniform(mptc_mix64(key + (uint64_t)t * n3 + i), span) - c->noise_amp;
      rgb[i] = mptc_clamp255((int32_t)rgb[i] + nz);
    }
  }
}

/* Named workloads. */
static inline mptc_synth_cfg mptc_synth_preset(int32_t width, int32_t height, uint64_t seed,
                                               int32_t n_rects, int32_t vel_mode, int32_t vel,
                                               int32_t cut_permille) {
  mptc_synth_cfg c;
  c.width = width; c.height = height; c.seed = seed; c.n_rects = n_rects;
  c.vel_mode = vel_mode; c.vel = vel; c.noise_amp = 8; c.cut_permille = cut_permille;
  return c;
}

#ifdef __cplusplus
}
#endif
#endif
