// Harness around the reference codec's file writer, CompressMultiUnique
// (codec/codec.cpp:1307-1532), and its frame decoder, compiled from the
// reference's own sources where they lie (codec/codec.cpp, dxt_image.cpp,
// image.cpp, wavelet.cpp, image_processing.cpp, entropy/arithmetic_codec.cpp).
// TEST INFRASTRUCTURE ONLY: it writes the golden bitstreams; the product never
// links it.
//
// Build: oracle/build_ref.sh (outputs only into oracle/_ref/).
#include <string>

#include "codec.h"
#include "stb_image_write.h"

extern "C" {

// Encodes the PNG frames of `dir` (sorted by name) into `out_file` exactly as
// the reference's encoder does (codec/encoder.cpp:482-483).  The reference
// also writes out.txt into the working directory.
int ref_compress_dir(const char* dir, const char* out_file, unsigned search_area, int threshold,
                     unsigned intra_interval, unsigned unique_interval) {
  MPTC::CompressMultiUnique(std::string(dir), std::string(out_file), search_area, threshold, intra_interval,
                            unique_interval, std::string(dir));
  return 0;
}

// An RGB8 frame as a PNG file (stb_image_write, compiled in with
// codec/dxt_image.cpp); PNG is lossless, so the encoder reads back these pixels.
int ref_write_png(const char* path, const unsigned char* rgb, int w, int h) {
  return stbi_write_png(path, w, h, 3, rgb, w * 3) ? 0 : -1;
}

}  // extern "C"
