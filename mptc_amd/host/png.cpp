// PNG frame input and output for the host layer (mptc_host.h), on zlib.
//
// The reference loads each frame with stbi_load(path, ..., 3)
// (codec/dxt_image.cpp:396): any 8-bit PNG becomes RGB8.  This reader does
// the same conversions for non-interlaced PNGs: greyscale replicated, alpha
// dropped, palette expanded, 16-bit samples reduced to their high byte, and
// 1/2/4-bit greyscale and palette images expanded.  Adam7-interlaced files
// are rejected with an error.  The writer stores RGB8 with filter 0.
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mptc_host.h"

namespace mptc {
namespace host {

namespace {

uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
  for (int i = 3; i >= 0; --i) v.push_back((uint8_t)(x >> (8 * i)));
}

const uint8_t kSignature[8] = {137, 80, 78, 71, 13, 10, 26, 10};

[[noreturn]] void bad(const std::string& path, const std::string& why) {
  throw Error(MPTC_EINVAL, "png " + path + ": " + why);
}

int paeth(int a, int b, int c) {
  const int p = a + b - c, pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
  if (pa <= pb && pa <= pc) return a;
  return pb <= pc ? b : c;
}

}  // namespace

std::vector<uint8_t> LoadPNG(const std::string& path, int* width, int* height) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) bad(path, "cannot open");
  std::vector<uint8_t> file;
  uint8_t buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) file.insert(file.end(), buf, buf + n);
  std::fclose(f);
  if (file.size() < 8 || std::memcmp(file.data(), kSignature, 8) != 0) bad(path, "not a PNG file");
  uint32_t w = 0, h = 0;
  int depth = 0, ctype = -1, interlace = 0;
  std::vector<uint8_t> idat, plte;
  size_t pos = 8;
  bool end = false;
  while (!end) {
    if (pos + 12 > file.size()) bad(path, "truncated chunk");
    const uint32_t len = be32(&file[pos]);
    const char* type = reinterpret_cast<const char*>(&file[pos + 4]);
    if (pos + 12 + (size_t)len > file.size()) bad(path, "truncated chunk");
    const uint8_t* d = &file[pos + 8];
    if (!std::memcmp(type, "IHDR", 4)) {
      if (len < 13) bad(path, "short IHDR");
      w = be32(d);
      h = be32(d + 4);
      depth = d[8];
      ctype = d[9];
      interlace = d[12];
    } else if (!std::memcmp(type, "PLTE", 4)) {
      plte.assign(d, d + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), d, d + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      end = true;
    }
    pos += 12 + (size_t)len;
  }
  if (w == 0 || h == 0 || w > (1u << 16) || h > (1u << 16)) bad(path, "bad size");
  if (interlace) bad(path, "interlaced PNG is not supported");
  int channels;
  switch (ctype) {
    case 0: channels = 1; break;
    case 2: channels = 3; break;
    case 3: channels = 1; break;
    case 4: channels = 2; break;
    case 6: channels = 4; break;
    default: bad(path, "unknown colour type");
  }
  const bool sub_byte = depth < 8;
  if (!(depth == 8 || depth == 16 || (sub_byte && (ctype == 0 || ctype == 3) && (depth == 1 || depth == 2 || depth == 4))))
    bad(path, "unsupported bit depth");
  if (ctype == 3 && plte.size() < 3) bad(path, "palette image without PLTE");
  const size_t bits_px = (size_t)channels * depth;
  const size_t row_bytes = (bits_px * w + 7) / 8;
  const size_t bpp = std::max<size_t>(1, bits_px / 8);  // filter unit
  std::vector<uint8_t> raw((row_bytes + 1) * h);
  uLongf raw_len = raw.size();
  if (uncompress(raw.data(), &raw_len, idat.data(), idat.size()) != Z_OK || raw_len != raw.size())
    bad(path, "corrupt image data");
  // undo the per-row filters
  std::vector<uint8_t> img(row_bytes * h);
  for (uint32_t y = 0; y < h; ++y) {
    const uint8_t ft = raw[y * (row_bytes + 1)];
    const uint8_t* src = &raw[y * (row_bytes + 1) + 1];
    uint8_t* dst = &img[y * row_bytes];
    const uint8_t* up = y ? &img[(y - 1) * row_bytes] : nullptr;
    for (size_t i = 0; i < row_bytes; ++i) {
      const int a = i >= bpp ? dst[i - bpp] : 0, b = up ? up[i] : 0, c = (up && i >= bpp) ? up[i - bpp] : 0;
      int v = src[i];
      switch (ft) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += b; break;
        case 3: v += (a + b) >> 1; break;
        case 4: v += paeth(a, b, c); break;
        default: bad(path, "unknown filter");
      }
      dst[i] = (uint8_t)v;
    }
  }
  // to RGB8
  std::vector<uint8_t> rgb((size_t)w * h * 3);
  for (uint32_t y = 0; y < h; ++y) {
    const uint8_t* row = &img[y * row_bytes];
    for (uint32_t x = 0; x < w; ++x) {
      uint8_t* o = &rgb[((size_t)y * w + x) * 3];
      if (sub_byte) {
        const size_t bit = (size_t)x * depth;
        const int v = (row[bit / 8] >> (8 - depth - (int)(bit % 8))) & ((1 << depth) - 1);
        if (ctype == 3) {
          if ((size_t)v * 3 + 2 >= plte.size()) bad(path, "palette index out of range");
          std::memcpy(o, &plte[(size_t)v * 3], 3);
        } else {
          const uint8_t g = (uint8_t)(v * 255 / ((1 << depth) - 1));
          o[0] = o[1] = o[2] = g;
        }
        continue;
      }
      const size_t step = (size_t)channels * (depth / 8);
      const uint8_t* p = row + x * step;
      auto sample = [&](int ch) { return p[ch * (depth / 8)]; };  // 16-bit: the high byte
      switch (ctype) {
        case 0: case 4: o[0] = o[1] = o[2] = sample(0); break;
        case 2: case 6: o[0] = sample(0); o[1] = sample(1); o[2] = sample(2); break;
        case 3: {
          const size_t v = sample(0);
          if (v * 3 + 2 >= plte.size()) bad(path, "palette index out of range");
          std::memcpy(o, &plte[v * 3], 3);
          break;
        }
      }
    }
  }
  *width = (int)w;
  *height = (int)h;
  return rgb;
}

void SavePNG(const std::string& path, const uint8_t* rgb, int width, int height) {
  if (width <= 0 || height <= 0) throw Error(MPTC_EINVAL, "png " + path + ": bad size");
  const size_t row = (size_t)width * 3;
  std::vector<uint8_t> raw((row + 1) * height);
  for (int y = 0; y < height; ++y) {
    raw[y * (row + 1)] = 0;
    std::memcpy(&raw[y * (row + 1) + 1], rgb + y * row, row);
  }
  uLongf zlen = compressBound(raw.size());
  std::vector<uint8_t> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), raw.size(), 6) != Z_OK)
    throw Error(MPTC_EHIP, "png " + path + ": deflate failed");
  z.resize(zlen);
  std::vector<uint8_t> out(kSignature, kSignature + 8);
  auto chunk = [&](const char* type, const uint8_t* d, size_t n) {
    put_be32(out, (uint32_t)n);
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), d, d + n);
    put_be32(out, (uint32_t)crc32(0, &out[start], (uInt)(n + 4)));
  };
  std::vector<uint8_t> ihdr;
  put_be32(ihdr, (uint32_t)width);
  put_be32(ihdr, (uint32_t)height);
  const uint8_t rest[5] = {8, 2, 0, 0, 0};
  ihdr.insert(ihdr.end(), rest, rest + 5);
  chunk("IHDR", ihdr.data(), ihdr.size());
  chunk("IDAT", z.data(), z.size());
  chunk("IEND", nullptr, 0);
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw Error(MPTC_EINVAL, "png " + path + ": cannot create");
  const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
  if (std::fclose(f) != 0 || !ok) throw Error(MPTC_EINVAL, "png " + path + ": write failed");
}

}  // namespace host
}  // namespace mptc

// C entry points of the PNG helpers (used by the tests through ctypes).
extern "C" {

// Loads a PNG as RGB8 into rgb_out (capacity bytes); *w and *h receive the
// size.  Returns 0, MPTC_EINVAL on a bad file or a buffer too small.
int32_t mptc_png_load(const char* path, uint8_t* rgb_out, int64_t capacity, int32_t* w, int32_t* h) {
  try {
    int fw = 0, fh = 0;
    const std::vector<uint8_t> px = mptc::host::LoadPNG(path, &fw, &fh);
    *w = fw;
    *h = fh;
    if ((int64_t)px.size() > capacity) return MPTC_EINVAL;
    std::memcpy(rgb_out, px.data(), px.size());
    return MPTC_OK;
  } catch (const mptc::host::Error& e) {
    return e.code();
  }
}

int32_t mptc_png_save(const char* path, const uint8_t* rgb, int32_t w, int32_t h) {
  try {
    mptc::host::SavePNG(path, rgb, w, h);
    return MPTC_OK;
  } catch (const mptc::host::Error& e) {
    return e.code();
  }
}

}  // extern "C"
