/*
 * mptc_codec.h -- C-ABI of libmptc_codec.so, the CPU side of the MPTC file
 * format around the encode path: the entropy stage that turns re-encoded
 * frames into the .mptc stream (codec/codec.cpp:1307-1532 CompressMultiUnique,
 * :1115-1158 EntropyEncode(frame), :804-899 CompressEndpoint), the stream
 * reader, and the decoder's index-word reconstruction
 * (codec/decoder.cpp:203-254 ReconstructDXTFrame).  Plain C types; every call
 * returns 0 or a negative status (the MPTC_* codes of mptc.h) and
 * mptc_codec_last_error() gives the message for the calling thread.  No GPU.
 *
 * Stream layout (little-endian), as the reference writes it:
 *   header, 34 bytes: height u32, width u32, unique_interval u8, search_area
 *     u8, groups u32 (frames / unique_interval), then the maxima over the
 *     stream of: palette bytes per group u32, compressed palette bytes u32,
 *     compressed motion bytes per frame u32, compressed Y bytes u32,
 *     compressed CoCg bytes u32;
 *   per group of unique_interval frames: compressed palette size u32, the
 *     palette (the groups' unique index words as bytes, adaptive arithmetic
 *     code over 257 symbols), palette byte count u32, then per frame: unique
 *     word count u32, compressed motion size u32, the motion bytes (x, y per
 *     block, one adaptive model), and for endpoint 1 then 2: compressed Y size
 *     u32, Y code, compressed CoCg size u32, CoCg code -- the 565 endpoints
 *     as YCoCg 6-6-7 planes, 64x64-block 5/3 wavelet, coefficients + 128 as
 *     bytes in raster order, Co plane then Cg plane.
 * Frames after the last whole group are not written (they still count in
 * the maxima), as in the reference.
 */
#ifndef MPTC_CODEC_H
#define MPTC_CODEC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mptc_writer mptc_writer;
typedef struct mptc_reader mptc_reader;

typedef struct {
  int32_t width, height;          /* pixels */
  int32_t unique_interval;        /* frames per palette group */
  int32_t search_area;            /* S of the motion bytes */
  int32_t groups;                 /* groups in the stream */
  uint32_t max_palette_bytes;     /* the header's maxima */
  uint32_t max_palette_code;
  uint32_t max_motion_code;
  uint32_t max_y_code;
  uint32_t max_c_code;
} mptc_stream_info;

const char* mptc_codec_last_error(void);

/* Stream writer, CompressMultiUnique's file (codec/codec.cpp:1307-1532) built
 * in memory.  total_frames: the number of frames the caller will add (the
 * header's group count is total_frames / unique_interval).  The frame's block
 * grid must be a multiple of 64 blocks in each direction (the endpoint
 * wavelet's block, codec/codec.cpp:79). */
int32_t mptc_writer_create(int32_t width, int32_t height, int32_t search_area, int32_t unique_interval,
                           int32_t total_frames, mptc_writer** out);
/* One re-encoded frame, in order: blocks (u64, layout of mptc.h), motion
 * bytes (2 per block), its unique palette (n_unique words). */
int32_t mptc_writer_add_frame(mptc_writer* w, const uint64_t* phys, const uint8_t* motion,
                              const uint32_t* unique, uint32_t n_unique);
/* Patches the header's maxima (codec/codec.cpp:1514-1520) and returns the
 * stream; the buffer stays owned by the writer until mptc_writer_destroy. */
int32_t mptc_writer_finish(mptc_writer* w, const uint8_t** data, uint64_t* size);
void mptc_writer_destroy(mptc_writer* w);

/* Stream reader: entropy-decodes the stream frame by frame and rebuilds each
 * frame's blocks: index words from the motion bytes and the group's palette
 * (mptc_reconstruct_indices), endpoints from the inverse wavelet.  The
 * stream is copied. */
int32_t mptc_reader_open(const uint8_t* data, uint64_t size, mptc_reader** out, mptc_stream_info* info);
/* Next frame: blocks (width/4 x height/4 u64), motion bytes (2 per block),
 * and the frame's unique palette (capacity one word per block; *n_unique
 * receives the count).  Returns 1 when the stream holds no more frames. */
int32_t mptc_reader_next_frame(mptc_reader* r, uint64_t* phys_out, uint8_t* motion_out, uint32_t* unique_out,
                               uint32_t* n_unique_out);
void mptc_reader_destroy(mptc_reader* r);

/* The byte coder of the stream: EntropyEncode / EntropyDecode of a byte
 * vector (codec/codec.cpp:156-213, 560-577), one adaptive model over 257
 * symbols.  Encode writes at most `capacity` bytes (*code_bytes receives the
 * length; MPTC_EINVAL if it does not fit); decode reads n symbols. */
int32_t mptc_entropy_encode(const uint8_t* symbols, uint64_t n, uint8_t* code_out, uint64_t capacity,
                            uint64_t* code_bytes);
int32_t mptc_entropy_decode(const uint8_t* code, uint64_t code_bytes, uint64_t n, uint8_t* symbols_out);

/* ReconstructDXTFrame (codec/decoder.cpp:203-254): the index words of a
 * frame from its motion bytes: (255, 255) takes the next word of `unique`;
 * both bits 7 set: the previous frame's word at offset (x - S, y - S);
 * otherwise this frame's already rebuilt word at (x - S, y - 2S + 1).  Writes
 * bits 32-63 of phys_inout, keeps its endpoints.  Errors (the reference
 * asserts): more unique blocks than n_unique, an offset outside the frame, an
 * intra source not earlier in raster order, an inter block without prev. */
int32_t mptc_reconstruct_indices(const uint8_t* motion, int32_t blocks_x, int32_t blocks_y, int32_t search_area,
                                 const uint32_t* unique, uint32_t n_unique, const uint64_t* prev_phys,
                                 uint64_t* phys_inout);

#ifdef __cplusplus
}
#endif
#endif
