// libmptc_codec.so: the CPU entropy stage and file format of MPTC around the
// encode path (include/mptc_codec.h).
//
// What must match the reference byte for byte, and why each part looks the
// way it does:
//   * The arithmetic coder is Said and Pearlman's fast adaptive coder (32-bit
//     interval, 15-bit model scale, periodic model updates, table-assisted
//     decoding; A. Said, "Introduction to Arithmetic Coding -- Theory and
//     Practice", HP Labs HPL-2004-76), which the reference vendors as
//     entropy/arithmetic_codec.cpp.  Its code bytes depend on every constant
//     and rounding step of the model update, so those follow the published
//     method exactly.
//   * The endpoint planes are RGB565 -> YCoCg 6-6-7 (codec/image_processing.cpp
//     :10-27), a 5/3 integer wavelet over 64x64 blocks (codec/wavelet.cpp
//     :30-157, codec/image_processing.h:268-334), coefficients narrowed to
//     int8 and offset by 128 (codec/pixel_traits.h:261-274,
//     codec/image_utils.h:217-240).
//   * The stream layout is CompressMultiUnique's (codec/codec.cpp:1307-1532).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mptc.h"
#include "../../include/mptc_codec.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// ---------------------------------------------------------------------------
// Adaptive arithmetic coding over a fixed alphabet.
// ---------------------------------------------------------------------------
constexpr uint32_t kMinLength = 0x01000000u;  // renormalise below 2^24
constexpr uint32_t kMaxLength = 0xFFFFFFFFu;
constexpr uint32_t kLengthShift = 15;         // model probabilities are 15-bit
constexpr uint32_t kMaxCount = 1u << kLengthShift;

// Symbol counts, their cumulative distribution (scaled to 2^15) and, for the
// decoder, a table of where each 1/table_size of the range starts.
class Model {
 public:
  explicit Model(uint32_t symbols) : n_(symbols), last_(symbols - 1), count_(symbols), dist_(symbols) {
    if (n_ > 16) {
      uint32_t bits = 3;
      while (n_ > (1u << (bits + 2))) ++bits;
      table_size_ = 1u << bits;
      table_shift_ = kLengthShift - bits;
      table_.assign(table_size_ + 2, 0);
    }
    reset();
  }

  void reset() {
    total_ = 0;
    cycle_ = n_;
    std::fill(count_.begin(), count_.end(), 1u);
    update(false);
    until_update_ = cycle_ = (n_ + 6) >> 1;
  }

  // Counts are halved when their total would pass 2^15; the distribution is
  // recomputed every `cycle_` symbols, the cycle growing by 5/4 up to
  // 8 (n + 6).
  void update(bool encoder) {
    if ((total_ += cycle_) > kMaxCount) {
      total_ = 0;
      for (uint32_t k = 0; k < n_; ++k) total_ += (count_[k] = (count_[k] + 1) >> 1);
    }
    const uint32_t scale = 0x80000000u / total_;
    uint32_t sum = 0, s = 0;
    if (encoder || table_size_ == 0) {
      for (uint32_t k = 0; k < n_; ++k) {
        dist_[k] = (scale * sum) >> (31 - kLengthShift);
        sum += count_[k];
      }
    } else {
      for (uint32_t k = 0; k < n_; ++k) {
        dist_[k] = (scale * sum) >> (31 - kLengthShift);
        sum += count_[k];
        const uint32_t w = dist_[k] >> table_shift_;
        while (s < w) table_[++s] = k - 1;
      }
      table_[0] = 0;
      while (s <= table_size_) table_[++s] = n_ - 1;
    }
    cycle_ = (5 * cycle_) >> 2;
    const uint32_t max_cycle = (n_ + 6) << 3;
    if (cycle_ > max_cycle) cycle_ = max_cycle;
    until_update_ = cycle_;
  }

  void seen(uint32_t sym, bool encoder) {
    ++count_[sym];
    if (--until_update_ == 0) update(encoder);
  }

  uint32_t n_, last_;
  uint32_t total_ = 0, cycle_ = 0, until_update_ = 0;
  uint32_t table_size_ = 0, table_shift_ = 0;
  std::vector<uint32_t> count_, dist_, table_;
};

class Encoder {
 public:
  void encode(uint32_t sym, Model& m) {
    uint32_t x;
    const uint32_t init_base = base_;
    if (sym == m.last_) {
      x = m.dist_[sym] * (length_ >> kLengthShift);
      base_ += x;
      length_ -= x;
    } else {
      x = m.dist_[sym] * (length_ >>= kLengthShift);
      base_ += x;
      length_ = m.dist_[sym + 1] * length_ - x;
    }
    if (init_base > base_) carry();
    if (length_ < kMinLength) renorm();
    m.seen(sym, true);
  }

  // Flushes the interval: one more byte when it is still wide, else two.
  std::vector<uint8_t>& finish() {
    const uint32_t init_base = base_;
    if (length_ > 2 * kMinLength) {
      base_ += kMinLength;
      length_ = kMinLength >> 1;
    } else {
      base_ += kMinLength >> 1;
      length_ = kMinLength >> 9;
    }
    if (init_base > base_) carry();
    renorm();
    return out_;
  }

 private:
  void carry() {
    size_t p = out_.size();
    while (p > 0 && out_[p - 1] == 0xFFu) out_[--p] = 0;
    if (p > 0) ++out_[p - 1];
  }
  void renorm() {
    do {
      out_.push_back((uint8_t)(base_ >> 24));
      base_ <<= 8;
    } while ((length_ <<= 8) < kMinLength);
  }

  uint32_t base_ = 0, length_ = kMaxLength;
  std::vector<uint8_t> out_;
};

class Decoder {
 public:
  Decoder(const uint8_t* code, size_t bytes) : code_(code, code + bytes) {
    code_.resize(bytes + 8, 0);  // the decoder reads up to 4 bytes past the end
    value_ = ((uint32_t)code_[0] << 24) | ((uint32_t)code_[1] << 16) | ((uint32_t)code_[2] << 8) | code_[3];
    pos_ = 3;
  }

  uint32_t decode(Model& m) {
    uint32_t x, y = length_, s;
    const uint32_t dv = value_ / (length_ >>= kLengthShift);
    const uint32_t t = dv >> m.table_shift_;
    s = m.table_[t];
    uint32_t n = m.table_[t + 1] + 1;
    while (n > s + 1) {
      const uint32_t mid = (s + n) >> 1;
      if (m.dist_[mid] > dv) n = mid; else s = mid;
    }
    x = m.dist_[s] * length_;
    if (s != m.last_) y = m.dist_[s + 1] * length_;
    value_ -= x;
    length_ = y - x;
    if (length_ < kMinLength) {
      do {
        value_ = (value_ << 8) | code_[++pos_];
      } while ((length_ <<= 8) < kMinLength);
    }
    m.seen(s, false);
    return s;
  }

 private:
  std::vector<uint8_t> code_;
  size_t pos_ = 0;
  uint32_t value_ = 0, length_ = kMaxLength;
};

// Bytes as symbols of one adaptive model over 257 symbols (the reference's
// EntropyEncode of a byte vector, codec/codec.cpp:156-213).
std::vector<uint8_t> code_bytes_of(const uint8_t* sym, size_t n) {
  Model m(257);
  Encoder e;
  for (size_t i = 0; i < n; ++i) e.encode(sym[i], m);
  return e.finish();
}

void decode_bytes(const uint8_t* code, size_t bytes, size_t n, uint8_t* out) {
  Model m(257);
  Decoder d(code, bytes);
  for (size_t i = 0; i < n; ++i) out[i] = (uint8_t)d.decode(m);
}

// ---------------------------------------------------------------------------
// Endpoint planes.
// ---------------------------------------------------------------------------
constexpr int kWaveBlock = 64;

// Mirror an index into [0, len) (..., 2, 1, 0, 1, 2, ..., len-2, len-1, len-2, ...).
inline int mirror(int i, int len) {
  if (i < 0) return -i;
  if (i >= len) return 2 * len - 2 - i;
  return i;
}

// One level of the 5/3 lifting transform of v[0..len) (stride st): low-pass
// to the first ceil(len/2) entries, high-pass after them.  Integer division
// truncates toward zero, as in the reference's C arithmetic.
void fwd53(int16_t* v, int len, int st, int16_t* tmp) {
  if (len < 2) return;
  const int mid = len - len / 2;
  for (int i = 1; i < len; i += 2) {
    const int a = v[mirror(i - 1, len) * st], b = v[mirror(i + 1, len) * st];
    tmp[mid + i / 2] = (int16_t)(v[i * st] - (a + b) / 2);
  }
  for (int i = 0; i < len; i += 2) {
    const int a = tmp[mid + mirror(i - 1, len) / 2], b = tmp[mid + mirror(i + 1, len) / 2];
    tmp[i / 2] = (int16_t)(v[i * st] + (a + b + 2) / 4);
  }
  for (int i = 0; i < len; ++i) v[i * st] = tmp[i];
}

void inv53(int16_t* v, int len, int st, int16_t* tmp) {
  if (len < 2) return;
  const int mid = len - len / 2;
  for (int i = 0; i < len; i += 2) {
    const int a = v[(mid + mirror(i - 1, len) / 2) * st], b = v[(mid + mirror(i + 1, len) / 2) * st];
    tmp[i] = (int16_t)(v[(i / 2) * st] - (a + b + 2) / 4);
  }
  for (int i = 1; i < len; i += 2) {
    const int a = tmp[mirror(i - 1, len)], b = tmp[mirror(i + 1, len)];
    tmp[i] = (int16_t)(v[(mid + i / 2) * st] + (a + b) / 2);
  }
  for (int i = 0; i < len; ++i) v[i * st] = tmp[i];
}

// Forward transform of one 64x64 block (row stride 64): levels of side 64,
// 32, ..., 2 on the top-left corner, each transforming the columns and then
// the rows.
void fwd_block(int16_t* b) {
  int16_t tmp[kWaveBlock];
  for (int dim = kWaveBlock; dim > 1; dim /= 2) {
    for (int x = 0; x < dim; ++x) fwd53(b + x, dim, kWaveBlock, tmp);
    for (int y = 0; y < dim; ++y) fwd53(b + y * kWaveBlock, dim, 1, tmp);
  }
}

void inv_block(int16_t* b) {
  int16_t tmp[kWaveBlock];
  for (int dim = 2; dim <= kWaveBlock; dim *= 2) {
    for (int y = 0; y < dim; ++y) inv53(b + y * kWaveBlock, dim, 1, tmp);
    for (int x = 0; x < dim; ++x) inv53(b + x, dim, kWaveBlock, tmp);
  }
}

// RGB565 -> YCoCg 6-6-7 (all integer, C division).
inline void to_ycocg(uint32_t v, int& y, int& co, int& cg) {
  const int r = (int)(v >> 11), g = (int)((v >> 5) & 63), b = (int)(v & 31);
  co = r - b;
  const int t = r + b + (b >> 4);
  cg = g - t;
  y = t + cg / 2;
}

inline uint32_t from_ycocg(int y, int co, int cg) {
  const int t = y - cg / 2;
  const int g = cg + t;
  const int b = (t - co) / 2;
  const int r = b + co;
  return ((uint32_t)(r & 31) << 11) | ((uint32_t)(g & 63) << 5) | (uint32_t)(b & 31);
}

// The byte planes of one endpoint (`shift` 0: ep1, 16: ep2) of a frame:
// Y (bw x bh bytes) and Co then Cg (2 bw x bh bytes), each the plane's
// 64x64-block wavelet coefficients narrowed to int8, + 128, raster order.
void endpoint_planes(const uint64_t* phys, int bw, int bh, int shift, std::vector<uint8_t>& ys,
                     std::vector<uint8_t>& cs) {
  const size_t nb = (size_t)bw * bh;
  std::vector<int16_t> plane[3];
  for (auto& p : plane) p.resize(nb);
  for (size_t k = 0; k < nb; ++k) {
    int y, co, cg;
    to_ycocg((uint32_t)((phys[k] >> shift) & 0xFFFFu), y, co, cg);
    plane[0][k] = (int16_t)y;
    plane[1][k] = (int16_t)co;
    plane[2][k] = (int16_t)cg;
  }
  int16_t blk[kWaveBlock * kWaveBlock];
  ys.resize(nb);
  cs.resize(2 * nb);
  for (int c = 0; c < 3; ++c) {
    uint8_t* out = c == 0 ? ys.data() : cs.data() + (size_t)(c - 1) * nb;
    for (int by = 0; by < bh; by += kWaveBlock)
      for (int bx = 0; bx < bw; bx += kWaveBlock) {
        for (int y = 0; y < kWaveBlock; ++y)
          std::memcpy(blk + y * kWaveBlock, &plane[c][(size_t)(by + y) * bw + bx], sizeof(int16_t) * kWaveBlock);
        fwd_block(blk);
        for (int y = 0; y < kWaveBlock; ++y)
          for (int x = 0; x < kWaveBlock; ++x)
            out[(size_t)(by + y) * bw + bx + x] = (uint8_t)((int8_t)blk[y * kWaveBlock + x] + 128);
      }
  }
}

// Inverse of endpoint_planes: the endpoint values (exact when no coefficient
// left the int8 range).
void endpoints_from_planes(const uint8_t* ys, const uint8_t* cs, int bw, int bh, int shift, uint64_t* phys) {
  const size_t nb = (size_t)bw * bh;
  std::vector<int16_t> plane[3];
  for (auto& p : plane) p.resize(nb);
  int16_t blk[kWaveBlock * kWaveBlock];
  for (int c = 0; c < 3; ++c) {
    const uint8_t* in = c == 0 ? ys : cs + (size_t)(c - 1) * nb;
    for (int by = 0; by < bh; by += kWaveBlock)
      for (int bx = 0; bx < bw; bx += kWaveBlock) {
        for (int y = 0; y < kWaveBlock; ++y)
          for (int x = 0; x < kWaveBlock; ++x)
            blk[y * kWaveBlock + x] = (int16_t)((int)in[(size_t)(by + y) * bw + bx + x] - 128);
        inv_block(blk);
        for (int y = 0; y < kWaveBlock; ++y)
          std::memcpy(&plane[c][(size_t)(by + y) * bw + bx], blk + y * kWaveBlock, sizeof(int16_t) * kWaveBlock);
      }
  }
  const uint64_t keep = ~((uint64_t)0xFFFFu << shift);
  for (size_t k = 0; k < nb; ++k)
    phys[k] = (phys[k] & keep) | ((uint64_t)from_ycocg(plane[0][k], plane[1][k], plane[2][k]) << shift);
}

void put_u32(std::vector<uint8_t>& v, uint32_t x) {
  for (int i = 0; i < 4; ++i) v.push_back((uint8_t)(x >> (8 * i)));
}

void put_code(std::vector<uint8_t>& v, const std::vector<uint8_t>& code) {
  put_u32(v, (uint32_t)code.size());
  v.insert(v.end(), code.begin(), code.end());
}

uint32_t get_u32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

int reconstruct(const uint8_t* motion, int bw, int bh, int S, const uint32_t* unique, uint32_t n_unique,
                const uint64_t* prev, uint64_t* cur) {
  const int nb = bw * bh;
  uint32_t next_unique = 0;
  for (int k = 0; k < nb; ++k) {
    const int x = motion[2 * k], y = motion[2 * k + 1];
    const int bx = k % bw, by = k / bw;
    uint32_t word;
    if (x == 255 && y == 255) {
      if (next_unique >= n_unique) return fail(MPTC_EINVAL, "more unique blocks than palette words");
      word = unique[next_unique++];
    } else if ((x & 0x80) && (y & 0x80)) {
      if (!prev) return fail(MPTC_EINVAL, "inter block without a previous frame");
      const int i = bx + (x & 0x7F) - S, j = by + (y & 0x7F) - S;
      if (i < 0 || i >= bw || j < 0 || j >= bh) return fail(MPTC_EINVAL, "inter offset outside the frame");
      word = (uint32_t)(prev[j * bw + i] >> 32);
    } else {
      const int i = bx + x - S, j = by + y - 2 * S + 1;
      if (i < 0 || i >= bw || j < 0 || j >= bh || j * bw + i >= k)
        return fail(MPTC_EINVAL, "intra offset not an earlier block of the frame");
      word = (uint32_t)(cur[j * bw + i] >> 32);
    }
    cur[k] = (cur[k] & 0xFFFFFFFFull) | ((uint64_t)word << 32);
  }
  return MPTC_OK;
}

}  // namespace

struct mptc_writer {
  int w = 0, h = 0, bw = 0, bh = 0, S = 0, interval = 0, total = 0, added = 0;
  std::vector<uint8_t> out;       // the stream
  std::vector<uint8_t> frames;    // the current group's frame records
  std::vector<uint8_t> palette;   // the current group's palette bytes
  int in_group = 0;
  uint32_t max_pal_bytes = 0, max_pal_code = 0, max_motion = 0, max_y = 0, max_c = 0;
};

struct mptc_reader {
  std::vector<uint8_t> data;
  mptc_stream_info info{};
  size_t pos = 0;
  int bw = 0, bh = 0;
  int group = 0, frame_in_group = 0;
  std::vector<uint8_t> palette;   // the current group's palette bytes
  size_t pal_pos = 0;             // next unused palette byte
  std::vector<uint64_t> prev;     // the previous frame's blocks
  bool have_prev = false;
};

extern "C" {

const char* mptc_codec_last_error(void) { return g_err.c_str(); }

int32_t mptc_writer_create(int32_t width, int32_t height, int32_t search_area, int32_t unique_interval,
                           int32_t total_frames, mptc_writer** out) {
  if (!out) return fail(MPTC_EINVAL, "null argument");
  *out = nullptr;
  if (width <= 0 || height <= 0 || (width & 3) || (height & 3))
    return fail(MPTC_EINVAL, "width and height must be positive multiples of 4");
  if ((width / 4) % kWaveBlock || (height / 4) % kWaveBlock)
    return fail(MPTC_EINVAL, "the block grid must be a multiple of 64 blocks in each direction");
  if (search_area < 1 || search_area > 63) return fail(MPTC_EINVAL, "search_area must be in 1..63");
  if (unique_interval < 1 || unique_interval > 255) return fail(MPTC_EINVAL, "unique_interval must be in 1..255");
  if (total_frames < 0) return fail(MPTC_EINVAL, "negative frame count");
  mptc_writer* w = new mptc_writer();
  w->w = width;
  w->h = height;
  w->bw = width / 4;
  w->bh = height / 4;
  w->S = search_area;
  w->interval = unique_interval;
  w->total = total_frames;
  // header: the maxima are patched by mptc_writer_finish
  put_u32(w->out, (uint32_t)height);
  put_u32(w->out, (uint32_t)width);
  w->out.push_back((uint8_t)unique_interval);
  w->out.push_back((uint8_t)search_area);
  put_u32(w->out, (uint32_t)(total_frames / unique_interval));
  for (int i = 0; i < 5; ++i) put_u32(w->out, 0u);
  *out = w;
  return MPTC_OK;
}

int32_t mptc_writer_add_frame(mptc_writer* w, const uint64_t* phys, const uint8_t* motion, const uint32_t* unique,
                              uint32_t n_unique) {
  if (!w || !phys || !motion || (n_unique && !unique)) return fail(MPTC_EINVAL, "null argument");
  const size_t nb = (size_t)w->bw * w->bh;
  if (n_unique > nb) return fail(MPTC_EINVAL, "more unique words than blocks");
  // frame record (codec/codec.cpp:1115-1158, 804-899)
  const std::vector<uint8_t> mcode = code_bytes_of(motion, 2 * nb);
  std::vector<uint8_t>& f = w->frames;
  put_u32(f, n_unique);
  put_code(f, mcode);
  w->max_motion = std::max(w->max_motion, (uint32_t)mcode.size());
  for (int shift : {0, 16}) {
    std::vector<uint8_t> ys, cs;
    endpoint_planes(phys, w->bw, w->bh, shift, ys, cs);
    const std::vector<uint8_t> ycode = code_bytes_of(ys.data(), ys.size());
    const std::vector<uint8_t> ccode = code_bytes_of(cs.data(), cs.size());
    put_code(f, ycode);
    put_code(f, ccode);
    w->max_y = std::max(w->max_y, (uint32_t)ycode.size());
    w->max_c = std::max(w->max_c, (uint32_t)ccode.size());
  }
  const uint8_t* pb = reinterpret_cast<const uint8_t*>(unique);
  w->palette.insert(w->palette.end(), pb, pb + 4 * (size_t)n_unique);
  ++w->added;
  if (++w->in_group == w->interval) {
    // group record (codec/codec.cpp:1471-1506)
    const std::vector<uint8_t> pcode = code_bytes_of(w->palette.data(), w->palette.size());
    put_code(w->out, pcode);
    put_u32(w->out, (uint32_t)w->palette.size());
    w->out.insert(w->out.end(), f.begin(), f.end());
    w->max_pal_code = std::max(w->max_pal_code, (uint32_t)pcode.size());
    w->max_pal_bytes = std::max(w->max_pal_bytes, (uint32_t)w->palette.size());
    f.clear();
    w->palette.clear();
    w->in_group = 0;
  }
  return MPTC_OK;
}

int32_t mptc_writer_finish(mptc_writer* w, const uint8_t** data, uint64_t* size) {
  if (!w || !data || !size) return fail(MPTC_EINVAL, "null argument");
  if (w->added != w->total) return fail(MPTC_ESTATE, "fewer or more frames added than announced");
  const uint32_t maxima[5] = {w->max_pal_bytes, w->max_pal_code, w->max_motion, w->max_y, w->max_c};
  for (int i = 0; i < 5; ++i)
    for (int b = 0; b < 4; ++b) w->out[14 + 4 * i + b] = (uint8_t)(maxima[i] >> (8 * b));
  *data = w->out.data();
  *size = w->out.size();
  return MPTC_OK;
}

void mptc_writer_destroy(mptc_writer* w) { delete w; }

int32_t mptc_reader_open(const uint8_t* data, uint64_t size, mptc_reader** out, mptc_stream_info* info) {
  if (!data || !out) return fail(MPTC_EINVAL, "null argument");
  *out = nullptr;
  if (size < 34) return fail(MPTC_EINVAL, "stream shorter than its header");
  mptc_reader* r = new mptc_reader();
  r->data.assign(data, data + size);
  mptc_stream_info& in = r->info;
  in.height = (int32_t)get_u32(data);
  in.width = (int32_t)get_u32(data + 4);
  in.unique_interval = data[8];
  in.search_area = data[9];
  in.groups = (int32_t)get_u32(data + 10);
  in.max_palette_bytes = get_u32(data + 14);
  in.max_palette_code = get_u32(data + 18);
  in.max_motion_code = get_u32(data + 22);
  in.max_y_code = get_u32(data + 26);
  in.max_c_code = get_u32(data + 30);
  if (in.width <= 0 || in.height <= 0 || (in.width & 3) || (in.height & 3) || in.unique_interval < 1 ||
      in.search_area < 1) {
    delete r;
    return fail(MPTC_EINVAL, "bad stream header");
  }
  r->bw = in.width / 4;
  r->bh = in.height / 4;
  r->pos = 34;
  if (info) *info = in;
  *out = r;
  return MPTC_OK;
}

int32_t mptc_reader_next_frame(mptc_reader* r, uint64_t* phys_out, uint8_t* motion_out, uint32_t* unique_out,
                               uint32_t* n_unique_out) {
  if (!r || !phys_out || !motion_out || !unique_out || !n_unique_out) return fail(MPTC_EINVAL, "null argument");
  const size_t nb = (size_t)r->bw * r->bh;
  const size_t end = r->data.size();
  auto need = [&](size_t n) { return r->pos + n <= end; };
  auto read_code = [&](const uint8_t*& code, uint32_t& bytes) -> bool {
    if (!need(4)) return false;
    bytes = get_u32(&r->data[r->pos]);
    r->pos += 4;
    if (!need(bytes)) return false;
    code = &r->data[r->pos];
    r->pos += bytes;
    return true;
  };
  if (r->frame_in_group == 0) {
    if (r->group >= r->info.groups) return 1;
    const uint8_t* pcode;
    uint32_t pbytes;
    if (!read_code(pcode, pbytes) || !need(4)) return fail(MPTC_EINVAL, "truncated palette record");
    const uint32_t n_pal = get_u32(&r->data[r->pos]);
    r->pos += 4;
    r->palette.resize(n_pal);
    decode_bytes(pcode, pbytes, n_pal, r->palette.data());
    r->pal_pos = 0;
  }
  if (!need(4)) return fail(MPTC_EINVAL, "truncated frame record");
  const uint32_t n_unique = get_u32(&r->data[r->pos]);
  r->pos += 4;
  if (n_unique > nb || r->pal_pos + 4 * (size_t)n_unique > r->palette.size())
    return fail(MPTC_EINVAL, "frame uses more palette words than its group holds");
  const uint8_t* code;
  uint32_t bytes;
  if (!read_code(code, bytes)) return fail(MPTC_EINVAL, "truncated motion record");
  decode_bytes(code, bytes, 2 * nb, motion_out);
  std::memset(phys_out, 0, sizeof(uint64_t) * nb);
  for (int shift : {0, 16}) {
    const uint8_t *yc, *cc;
    uint32_t yb, cb;
    if (!read_code(yc, yb) || !read_code(cc, cb)) return fail(MPTC_EINVAL, "truncated endpoint record");
    std::vector<uint8_t> ys(nb), cs(2 * nb);
    decode_bytes(yc, yb, nb, ys.data());
    decode_bytes(cc, cb, 2 * nb, cs.data());
    endpoints_from_planes(ys.data(), cs.data(), r->bw, r->bh, shift, phys_out);
  }
  std::memcpy(unique_out, &r->palette[r->pal_pos], 4 * (size_t)n_unique);
  r->pal_pos += 4 * (size_t)n_unique;
  *n_unique_out = n_unique;
  const int rc = reconstruct(motion_out, r->bw, r->bh, r->info.search_area, unique_out, n_unique,
                             r->have_prev ? r->prev.data() : nullptr, phys_out);
  if (rc) return rc;
  r->prev.assign(phys_out, phys_out + nb);
  r->have_prev = true;
  if (++r->frame_in_group == r->info.unique_interval) {
    r->frame_in_group = 0;
    ++r->group;
  }
  return MPTC_OK;
}

void mptc_reader_destroy(mptc_reader* r) { delete r; }

int32_t mptc_entropy_encode(const uint8_t* symbols, uint64_t n, uint8_t* code_out, uint64_t capacity,
                            uint64_t* code_bytes) {
  if ((n && !symbols) || !code_out || !code_bytes) return fail(MPTC_EINVAL, "null argument");
  const std::vector<uint8_t> code = code_bytes_of(symbols, (size_t)n);
  *code_bytes = code.size();
  if (code.size() > capacity) return fail(MPTC_EINVAL, "code does not fit the output buffer");
  std::memcpy(code_out, code.data(), code.size());
  return MPTC_OK;
}

int32_t mptc_entropy_decode(const uint8_t* code, uint64_t code_bytes, uint64_t n, uint8_t* symbols_out) {
  if (!code || (n && !symbols_out)) return fail(MPTC_EINVAL, "null argument");
  if (code_bytes < 1) return fail(MPTC_EINVAL, "empty code");
  decode_bytes(code, (size_t)code_bytes, (size_t)n, symbols_out);
  return MPTC_OK;
}

int32_t mptc_reconstruct_indices(const uint8_t* motion, int32_t blocks_x, int32_t blocks_y, int32_t search_area,
                                 const uint32_t* unique, uint32_t n_unique, const uint64_t* prev_phys,
                                 uint64_t* phys_inout) {
  if (!motion || !phys_inout || (n_unique && !unique)) return fail(MPTC_EINVAL, "null argument");
  if (blocks_x <= 0 || blocks_y <= 0) return fail(MPTC_EINVAL, "bad block grid");
  if (search_area < 1 || search_area > 63) return fail(MPTC_EINVAL, "search_area must be in 1..63");
  return reconstruct(motion, blocks_x, blocks_y, search_area, unique, n_unique, prev_phys, phys_inout);
}

}  // extern "C"
