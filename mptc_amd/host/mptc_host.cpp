// C++ host side of the MPTC encode path (see mptc_host.h).
#include "mptc_host.h"

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

}  // namespace host
}  // namespace mptc
