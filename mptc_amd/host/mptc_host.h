// C++ host side of the MPTC encode path, above the C-ABI of libmptc_hip.so
// (include/mptc.h).  It mirrors the reference's interface for this path:
//
//   EncodedFrame     the DXTImage data members the encoder's consumers read
//                    (codec/dxt_image.h:166-194): _physical_blocks,
//                    _motion_indices, _unique_palette and the logical
//                    endpoints behind _logical_blocks
//   Encoder          the DXTImage constructor's fit loop
//                    (codec/dxt_image.cpp:419-431) followed by
//                    DXTImage::Reencode (codec/dxt_image.cpp:868-957)
//   SequenceEncoder  CompressMultiUnique's frame schedule: intra frames at
//                    0, N, 2N, ...; each inter frame re-encodes against the
//                    previous re-encoded frame (codec/codec.cpp:1379-1469,
//                    1508)
//   CompressDirectory  CompressMultiUnique itself (codec/codec.cpp:1307-1532):
//                    the PNG frames of a directory in name order, encoded on
//                    the GPU with that schedule, written as the .mptc stream
//                    by the CPU entropy stage (include/mptc_codec.h), palette
//                    groups of unique_interval frames
//   LoadPNG/SavePNG  frame input as the reference's stbi_load(..., 3)
//                    (codec/dxt_image.cpp:396) gives it, and output
//
// Errors: the C-ABI returns status codes; this layer turns a failed call into
// mptc::host::Error carrying the code and mptc_last_error()'s message (the
// reference asserts or continues silently, codec/dxt_image.cpp:397-401).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mptc.h"

namespace mptc {
namespace host {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

struct Options {
  int device = 0;
  int max_width = 1920;
  int max_height = 1080;
  int search_area = 8;    // S: window side 2S blocks (DXTImage's search_area)
  int err_threshold = 0;  // vErrThreshold, floor(SSE/48) units
  int intra_interval = 8; // CompressMultiUnique's intra interval
};

struct EncodedFrame {
  int width = 0, height = 0;
  int blocks_x = 0, blocks_y = 0;  // _blocks_width, _blocks_height
  bool is_intra = false;
  std::vector<uint64_t> physical_blocks;   // ep1 | ep2 << 16 | indices << 32, raster order
  std::vector<uint8_t> motion_indices;     // (x, y) per block; (255, 255) = unique
  std::vector<uint32_t> unique_palette;    // own index words of the unique blocks
  std::vector<uint8_t> logical_endpoints;  // ep1 rgb, ep2 rgb per block after re-encode
  std::vector<uint8_t> kinds;              // 0 unique, 1 intra reuse, 2 inter reuse
  int num_blocks() const { return blocks_x * blocks_y; }
};

class Encoder {
 public:
  explicit Encoder(const Options& opt);
  ~Encoder();
  Encoder(const Encoder&) = delete;
  Encoder& operator=(const Encoder&) = delete;

  // The constructor's fit loop: stb DXT1 HIGHQUAL of every block.
  std::vector<uint64_t> Fit(const uint8_t* rgb, int width, int height);
  // Fit + Reencode of one frame (reference = previous re-encoded frame,
  // ignored for intra frames).
  EncodedFrame Encode(const uint8_t* rgb, int width, int height, bool is_intra,
                      const EncodedFrame* reference);
  void SetParams(int search_area, int err_threshold);
  mptc_ctx* context() { return ctx_; }
  const Options& options() const { return opt_; }

 private:
  Options opt_;
  mptc_ctx* ctx_ = nullptr;
};

class SequenceEncoder {
 public:
  explicit SequenceEncoder(const Options& opt) : enc_(opt) {}
  // Encodes the next frame of the sequence; the returned frame stays valid
  // until the next call.
  const EncodedFrame& Add(const uint8_t* rgb, int width, int height);
  int frames() const { return count_; }
  Encoder& encoder() { return enc_; }

 private:
  Encoder enc_;
  EncodedFrame prev_;
  bool set_intra_ = true;
  int num_ = 1;
  int count_ = 0;
};

// RGB8 pixels of a PNG file (see png.cpp for the formats read).
std::vector<uint8_t> LoadPNG(const std::string& path, int* width, int* height);
void SavePNG(const std::string& path, const uint8_t* rgb, int width, int height);

struct DirectoryResult {
  int frames = 0;              // frames encoded
  int frames_written = 0;      // frames in whole palette groups (the rest are dropped)
  uint64_t bytes = 0;          // size of the .mptc file
};

// CompressMultiUnique (codec/codec.cpp:1307-1532): every file of `dir` in
// name order as a PNG frame (all the same size, a block grid that is a
// multiple of 64 in each direction), encoded with opt's search area,
// threshold and intra interval, written to out_file with palette groups of
// unique_interval frames.
DirectoryResult CompressDirectory(const std::string& dir, const std::string& out_file, const Options& opt,
                                  int unique_interval);

}  // namespace host
}  // namespace mptc
