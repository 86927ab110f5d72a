// Harness around the reference encoder, compiled from the reference's own
// sources where they lie (codec/dxt_image.cpp is included into this unit so
// that its file-local helpers -- CompressRGB, ToFiveBits, CompressedBlock --
// can be driven one at a time).  TEST INFRASTRUCTURE ONLY: it produces golden
// vectors and the reference CPU timing; the product never links it.
//
// Build: oracle/build_ref.sh (outputs only into oracle/_ref/).

#include "dxt_image.cpp"  // codec/dxt_image.cpp (with stb_dxt, stb_image, stb_image_write)

#include <chrono>
#include <cstdio>
#include <memory>
#include <string>

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

// CompressRGB (codec/dxt_image.cpp:123-149) on one block laid out as a 4x4 image.
uint64_t ref_fit_block(const uint8_t px[48]) { return CompressRGB(px, 4); }

uint8_t ref_to_five_bits(uint8_t x) { return ToFiveBits(x); }
uint8_t ref_to_six_bits(uint8_t x) { return ToSixBits(x); }

// PhysicalToLogical (codec/dxt_image.cpp:198-227): 40-byte LogicalDXTBlock.
void ref_p2l(uint64_t phys, uint8_t out[40]) {
  PhysicalDXTBlock b;
  b.dxt_block = phys;
  LogicalDXTBlock l = MPTC::PhysicalToLogical(b);
  std::memcpy(out, &l, sizeof(l));
}

// LogicalToPhysical (codec/dxt_image.cpp:165-196).
uint64_t ref_l2p(const uint8_t in[40]) {
  LogicalDXTBlock l;
  std::memcpy(&l, in, sizeof(l));
  return MPTC::LogicalToPhysical(l).dxt_block;
}

// CompressedBlock::Error (codec/dxt_image.cpp:258-276).
int ref_error(const uint8_t logical[40], const uint8_t px[48]) {
  MPTC::CompressedBlock blk;
  std::memcpy(&blk._logical, logical, sizeof(blk._logical));
  blk._uncompressed.assign(px, px + 48);
  return static_cast<int>(blk.Error());
}

// CompressedBlock::RecalculateEndpoints (codec/dxt_image.cpp:290-351) in place.
void ref_refit(uint8_t logical[40], const uint8_t px[48]) {
  MPTC::CompressedBlock blk;
  std::memcpy(&blk._logical, logical, sizeof(blk._logical));
  blk._uncompressed.assign(px, px + 48);
  blk.RecalculateEndpoints();
  std::memcpy(logical, &blk._logical, sizeof(blk._logical));
}

// One candidate of IntraSearch / InterBlockSearch, using the reference's
// CompressedBlock exactly as the search bodies do (codec/dxt_image.cpp:677-698).
int ref_eval_candidate(uint64_t own_phys, const uint8_t px[48], uint32_t cand, int* err_diff,
                       int* re_assigned, uint8_t eps[6]) {
  PhysicalDXTBlock own;
  own.dxt_block = own_phys;
  MPTC::CompressedBlock blk;
  blk._logical = MPTC::PhysicalToLogical(own);
  blk._uncompressed.assign(px, px + 48);
  const int orig_err = static_cast<int>(blk.Error());
  MPTC::CompressedBlock blk2 = blk;
  blk2.AssignIndices(cand);
  bool reset = false;
  if (!(blk == blk2)) {
    blk2.RecalculateEndpoints();
    reset = true;
    PhysicalDXTBlock maybe = MPTC::LogicalToPhysical(blk2._logical);
    bool ok = maybe.interp == cand;
    ok = ok && blk2._logical.palette[3][3] == 0xFF;
    if (!ok) return 0;
  }
  *err_diff = static_cast<int>(blk2.Error()) - orig_err;
  *re_assigned = reset ? 1 : 0;
  std::memcpy(eps, blk2._logical.ep1, 3);
  std::memcpy(eps + 3, blk2._logical.ep2, 3);
  return 1;
}

// Runs the reference frame loop (codec/codec.cpp:1379-1508 minus the entropy
// stage) over n_frames RGB8 frames: each frame is written as a PNG into
// tmpdir, loaded by the DXTImage constructor (codec/dxt_image.cpp:385) and
// re-encoded against the previous frame (codec/dxt_image.cpp:868).
// Outputs per frame f (nb = blocks per frame):
//   phys[f*nb..], motion[f*2nb..], unique[f*nb..] (n_unique[f] valid),
//   eps[f*6nb..] (logical ep1/ep2 rgb after re-encode), t_fit[f], t_reencode[f],
//   psnr_enc[f]: DXTImage::PSNR (codec/dxt_image.cpp:363-383) right after
//     Reencode, over the encoder's logical blocks (re-fit endpoints as the
//     encoder keeps them),
//   psnr_dec[f]: the same function after SetLogicalBlocks
//     (codec/dxt_image.cpp:452-455), i.e. over the blocks a decoder rebuilds
//     from the physical words (both optional: NULL skips them).
// t_fit is the constructor time minus a separate PNG load of the same file.
int ref_encode_sequence(const uint8_t* frames, int w, int h, int n_frames, int search_area,
                        int threshold, int intra_interval, const char* tmpdir, uint64_t* phys,
                        uint8_t* motion, uint32_t* unique, int* n_unique, uint8_t* eps,
                        double* t_fit, double* t_reencode, double* psnr_enc, double* psnr_dec) {
  const int bw = (w + 3) >> 2, bh = (h + 3) >> 2, nb = bw * bh;
  const std::string path = std::string(tmpdir) + "/ref_frame.png";
  std::unique_ptr<MPTC::DXTImage> prev_frame(nullptr), curr_frame(nullptr);
  bool set_intra = true;
  uint32_t curr_intra_num = 1;
  for (int f = 0; f < n_frames; ++f) {
    const uint8_t* rgb = frames + static_cast<size_t>(f) * w * h * 3;
    if (!stbi_write_png(path.c_str(), w, h, 3, rgb, w * 3)) return -1;
    double t0 = now_s();
    int lw, lh, lc;
    stbi_uc* probe = stbi_load(path.c_str(), &lw, &lh, &lc, 3);
    if (!probe) return -2;
    stbi_image_free(probe);
    double t1 = now_s();
    curr_frame.reset(new MPTC::DXTImage(path, set_intra, search_area, threshold));
    double t2 = now_s();
    if (curr_frame->_num_blocks != nb) return -3;
    curr_frame->Reencode(prev_frame, -1);
    double t3 = now_s();
    if (t_fit) t_fit[f] = (t2 - t1) - (t1 - t0);
    if (t_reencode) t_reencode[f] = t3 - t2;
    for (int k = 0; k < nb; ++k) {
      phys[static_cast<size_t>(f) * nb + k] = curr_frame->_physical_blocks[k].dxt_block;
      motion[(static_cast<size_t>(f) * nb + k) * 2 + 0] = std::get<0>(curr_frame->_motion_indices[k]);
      motion[(static_cast<size_t>(f) * nb + k) * 2 + 1] = std::get<1>(curr_frame->_motion_indices[k]);
      if (eps) {
        const LogicalDXTBlock& lb = curr_frame->_logical_blocks[k];
        std::memcpy(eps + (static_cast<size_t>(f) * nb + k) * 6, lb.ep1, 3);
        std::memcpy(eps + (static_cast<size_t>(f) * nb + k) * 6 + 3, lb.ep2, 3);
      }
    }
    n_unique[f] = static_cast<int>(curr_frame->_unique_palette.size());
    for (int u = 0; u < n_unique[f]; ++u)
      unique[static_cast<size_t>(f) * nb + u] = curr_frame->_unique_palette[u];
    if (psnr_enc) psnr_enc[f] = curr_frame->PSNR();
    if (psnr_dec) {
      // the next frame's inter search reads only _physical_blocks of this one
      curr_frame->SetLogicalBlocks();
      psnr_dec[f] = curr_frame->PSNR();
    }
    // intra schedule of codec/codec.cpp:1462-1469
    if (curr_intra_num <= static_cast<uint32_t>(intra_interval)) {
      set_intra = false;
      if (curr_intra_num == static_cast<uint32_t>(intra_interval)) {
        set_intra = true;
        curr_intra_num = 1;
      } else {
        curr_intra_num++;
      }
    }
    prev_frame = std::move(curr_frame);
  }
  std::remove(path.c_str());
  return 0;
}

// Times the reference fit loop (CompressRGB over every block, the constructor
// body at codec/dxt_image.cpp:419-431) on an in-memory frame.
double ref_time_fit(const uint8_t* rgb, int w, int h, uint64_t* phys_out) {
  const int bw = (w + 3) >> 2, bh = (h + 3) >> 2;
  double t0 = now_s();
  for (int j = 0; j < bh; ++j)
    for (int i = 0; i < bw; ++i)
      phys_out[j * bw + i] = CompressRGB(rgb + (static_cast<size_t>(j) * 4 * w + i * 4) * 3, w);
  return now_s() - t0;
}

}  // extern "C"
