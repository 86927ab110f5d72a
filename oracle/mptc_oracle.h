/*
 * CPU oracle for the MPTC per-frame encode path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain, sequential C restatement of the reference algorithm:
 *   - stb_dxt v1.04 DXT1 HIGHQUAL block fit (Include/stb_dxt.h:57-627) as
 *     driven by CompressRGB (codec/dxt_image.cpp:123-149);
 *   - the logical/physical block conversions (codec/dxt_image.cpp:34-227);
 *   - CompressedBlock::{Error, AssignIndices, RecalculateEndpoints}
 *     (codec/dxt_image.cpp:244-353);
 *   - IntraSearch / InterBlockSearch (codec/dxt_image.cpp:652-774) and the
 *     raster-order Reencode driver (codec/dxt_image.cpp:868-957).
 *
 * Only tests/, the smoke check and the bench's CPU-baseline leg may load this
 * library, and only as the checker.  The product path never calls it.
 *
 * Parity pinning: every function here is checked against fixtures produced
 * by the compiled reference (oracle/_ref), against exhaustive or closed-form
 * checks, or both (tests/test_oracle.py).
 */
#ifndef MPTC_ORACLE_H
#define MPTC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint8_t ep1[4];
  uint8_t ep2[4];
  uint8_t pal[4][4];
  uint8_t idx[16];
} orc_logical;

void orc_init(void);
/* one 4x4 block, px = 16 pixels x RGB, row-major; returns ep1 | ep2<<16 | interp<<32 */
uint64_t orc_fit_block(const uint8_t px[48]);
void orc_fit_frame(const uint8_t* rgb, int w, int h, uint64_t* phys);

uint8_t orc_to_five_bits(uint8_t x);
uint8_t orc_to_six_bits(uint8_t x);
void orc_p2l(uint64_t phys, orc_logical* out);
uint64_t orc_l2p(const orc_logical* b);
int orc_error(const orc_logical* b, const uint8_t px[48]);
void orc_refit(orc_logical* b, const uint8_t px[48]);

/* One candidate of a search: returns 1 if valid and writes the error delta
 * err(candidate) - err(own) to *err_diff, the re-assign flag and the
 * (quirky) re-fit logical endpoints. */
int orc_eval_candidate(uint64_t own_phys, const uint8_t px[48], uint32_t cand,
                       int* err_diff, int* re_assigned, uint8_t eps_out[6]);

/* Gathers block (bx, by) of an RGB8 frame into px (codec/dxt_image.cpp:636-650). */
void orc_gather(const uint8_t* rgb, int w, int bx, int by, uint8_t px[48]);

/*
 * DXTImage::Reencode for one frame.
 *  phys      [nb]   in: stb fit of this frame; out: re-encoded blocks
 *  ref_phys  [nb]   previous frame's re-encoded blocks (ignored if is_intra)
 *  motion    [2nb]  (x, y) bytes per block, (255, 255) for unique blocks
 *  unique    [nb]   unique palette (stb index words of unique blocks, raster order)
 *  kind      [nb]   0 unique, 1 intra, 2 inter (may be NULL)
 *  eps       [6nb]  logical ep1 rgb, ep2 rgb after re-encode (may be NULL)
 * returns the number of unique words.
 */
int orc_reencode(const uint8_t* rgb, int w, int h, int search_area, int threshold,
                 int is_intra, const uint64_t* ref_phys, uint64_t* phys, uint8_t* motion,
                 uint32_t* unique, uint8_t* kind, uint8_t* eps);

/* Reference CPU-side PSNR over logical blocks (codec/dxt_image.cpp:363-383) given
 * the re-encoded physical blocks and the quirky logical endpoints. */
double orc_psnr_logical(const uint8_t* rgb, int w, int h, const uint64_t* phys, const uint8_t* eps);
/* Decoder-side PSNR over the physical blocks (palette from P2L). */
double orc_psnr_physical(const uint8_t* rgb, int w, int h, const uint64_t* phys);

#ifdef __cplusplus
}
#endif
#endif
