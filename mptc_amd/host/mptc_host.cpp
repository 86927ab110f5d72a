// C++ host side of the MPTC encode path (see mptc_host.h).
#include "mptc_host.h"

#include <dirent.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "../../include/mptc_codec.h"

namespace mptc {
namespace host {

namespace {
void check(int rc) {
  if (rc != MPTC_OK) throw Error(rc, std::string("mptc: ") + mptc_last_error());
}
}  // namespace

Encoder::Encoder(const Options& opt) : opt_(opt) {
  mptc_config cfg{};
  cfg.device = opt.device;
  cfg.max_width = opt.max_width;
  cfg.max_height = opt.max_height;
  cfg.search_area = opt.search_area;
  cfg.err_threshold = opt.err_threshold;
  cfg.max_batch_frames = 1;
  cfg.max_batch_blocks = 0;
  check(mptc_create(&cfg, &ctx_));
}

Encoder::~Encoder() {
  if (ctx_) mptc_destroy(ctx_);
}

void Encoder::SetParams(int search_area, int err_threshold) {
  check(mptc_set_params(ctx_, search_area, err_threshold));
  opt_.search_area = search_area;
  opt_.err_threshold = err_threshold;
}

std::vector<uint64_t> Encoder::Fit(const uint8_t* rgb, int width, int height) {
  if (width <= 0 || height <= 0) throw Error(MPTC_EINVAL, "mptc: bad frame size");
  std::vector<uint64_t> out((size_t)(width / 4) * (height / 4));
  check(mptc_fit(ctx_, rgb, width, height, out.data()));
  return out;
}

EncodedFrame Encoder::Encode(const uint8_t* rgb, int width, int height, bool is_intra,
                             const EncodedFrame* reference) {
  EncodedFrame f;
  f.width = width;
  f.height = height;
  f.blocks_x = width / 4;
  f.blocks_y = height / 4;
  f.is_intra = is_intra;
  const size_t nb = (size_t)f.blocks_x * f.blocks_y;
  if (!is_intra) {
    if (!reference) throw Error(MPTC_EINVAL, "mptc: inter frame needs the previous frame");
    if (reference->width != width || reference->height != height)
      throw Error(MPTC_EINVAL, "mptc: reference frame has another size");
  }
  f.physical_blocks = Fit(rgb, width, height);
  f.motion_indices.resize(2 * nb);
  f.unique_palette.resize(nb);
  f.logical_endpoints.resize(6 * nb);
  f.kinds.resize(nb);
  uint32_t n_unique = 0;
  check(mptc_reencode(ctx_, rgb, width, height, is_intra ? 1 : 0,
                      is_intra ? nullptr : reference->physical_blocks.data(), f.physical_blocks.data(),
                      f.motion_indices.data(), f.unique_palette.data(), &n_unique, f.logical_endpoints.data(),
                      f.kinds.data()));
  f.unique_palette.resize(n_unique);
  return f;
}

const EncodedFrame& SequenceEncoder::Add(const uint8_t* rgb, int width, int height) {
  EncodedFrame f = enc_.Encode(rgb, width, height, set_intra_, set_intra_ ? nullptr : &prev_);
  prev_ = std::move(f);
  ++count_;
  // intra schedule (codec/codec.cpp:1462-1469)
  const int interval = enc_.options().intra_interval;
  if (num_ <= interval) {
    set_intra_ = false;
    if (num_ == interval) {
      set_intra_ = true;
      num_ = 1;
    } else {
      ++num_;
    }
  }
  return prev_;
}

DirectoryResult CompressDirectory(const std::string& dir, const std::string& out_file, const Options& opt_in,
                                  int unique_interval) {
  // every entry but . and .., sorted by name (codec/codec.cpp:1318-1328)
  DIR* d = opendir(dir.c_str());
  if (!d) throw Error(MPTC_EINVAL, "mptc: cannot open directory " + dir);
  std::vector<std::string> names;
  while (dirent* e = readdir(d)) {
    if (std::strcmp(e->d_name, ".") == 0 || std::strcmp(e->d_name, "..") == 0) continue;
    names.push_back(e->d_name);
  }
  closedir(d);
  std::sort(names.begin(), names.end());
  if (names.empty()) throw Error(MPTC_EINVAL, "mptc: no frames in " + dir);
  int w = 0, h = 0;
  std::vector<uint8_t> rgb = LoadPNG(dir + "/" + names[0], &w, &h);
  Options opt = opt_in;
  opt.max_width = w;
  opt.max_height = h;
  SequenceEncoder seq(opt);
  mptc_writer* wr = nullptr;
  if (mptc_writer_create(w, h, opt.search_area, unique_interval, (int)names.size(), &wr) != MPTC_OK)
    throw Error(MPTC_EINVAL, std::string("mptc: ") + mptc_codec_last_error());
  DirectoryResult res;
  try {
    for (size_t i = 0; i < names.size(); ++i) {
      int fw = w, fh = h;
      if (i > 0) rgb = LoadPNG(dir + "/" + names[i], &fw, &fh);
      if (fw != w || fh != h) throw Error(MPTC_EINVAL, "mptc: " + names[i] + " differs in size from the first frame");
      const EncodedFrame& f = seq.Add(rgb.data(), w, h);
      if (mptc_writer_add_frame(wr, f.physical_blocks.data(), f.motion_indices.data(), f.unique_palette.data(),
                                (uint32_t)f.unique_palette.size()) != MPTC_OK)
        throw Error(MPTC_EINVAL, std::string("mptc: ") + mptc_codec_last_error());
      ++res.frames;
    }
    const uint8_t* data = nullptr;
    uint64_t size = 0;
    if (mptc_writer_finish(wr, &data, &size) != MPTC_OK)
      throw Error(MPTC_ESTATE, std::string("mptc: ") + mptc_codec_last_error());
    FILE* f = std::fopen(out_file.c_str(), "wb");
    if (!f) throw Error(MPTC_EINVAL, "mptc: cannot create " + out_file);
    const bool ok = std::fwrite(data, 1, size, f) == size;
    if (std::fclose(f) != 0 || !ok) throw Error(MPTC_EINVAL, "mptc: cannot write " + out_file);
    res.bytes = size;
    res.frames_written = (res.frames / unique_interval) * unique_interval;
  } catch (...) {
    mptc_writer_destroy(wr);
    throw;
  }
  mptc_writer_destroy(wr);
  return res;
}

}  // namespace host
}  // namespace mptc
