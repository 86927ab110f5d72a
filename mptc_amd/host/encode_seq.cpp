// Encodes a seeded synthetic sequence with the C++ host layer and writes each
// frame's outputs as raw files (frameNNN.{phys,motion,unique,eps}) into a
// directory.  Used by the tests to check the host layer against the
// reference's fixtures.
//
//   encode_seq OUT_DIR WIDTH HEIGHT SEED RECTS VEL_MODE VEL CUT_PERMILLE FRAMES S THR INTRA_INTERVAL
//   encode_seq --version
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../csrc/framegen.h"
#include "mptc_host.h"

static bool write_file(const std::string& path, const void* data, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = bytes == 0 || std::fwrite(data, 1, bytes, f) == bytes;
  return std::fclose(f) == 0 && ok;
}

int main(int argc, char** argv) {
  if (argc == 2 && std::strcmp(argv[1], "--version") == 0) {
    std::printf("encode_seq (mptc host layer over libmptc_hip)\n");
    return 0;
  }
  if (argc != 13) {
    std::fprintf(stderr,
                 "usage: %s OUT_DIR WIDTH HEIGHT SEED RECTS VEL_MODE VEL CUT_PERMILLE FRAMES S THR INTRA_INTERVAL\n",
                 argv[0]);
    return 2;
  }
  const std::string out = argv[1];
  mptc_synth_cfg cfg{};
  cfg.width = std::atoi(argv[2]);
  cfg.height = std::atoi(argv[3]);
  cfg.seed = std::strtoull(argv[4], nullptr, 10);
  cfg.n_rects = std::atoi(argv[5]);
  cfg.vel_mode = std::atoi(argv[6]);
  cfg.vel = std::atoi(argv[7]);
  cfg.noise_amp = 8;
  cfg.cut_permille = std::atoi(argv[8]);
  const int frames = std::atoi(argv[9]);
  mptc::host::Options opt;
  opt.max_width = cfg.width;
  opt.max_height = cfg.height;
  opt.search_area = std::atoi(argv[10]);
  opt.err_threshold = std::atoi(argv[11]);
  opt.intra_interval = std::atoi(argv[12]);
  try {
    mptc::host::SequenceEncoder seq(opt);
    std::vector<uint8_t> rgb((size_t)cfg.width * cfg.height * 3);
    for (int t = 0; t < frames; ++t) {
      if (mptc_synth_frames(&cfg, t, 1, rgb.data()) != MPTC_OK) throw std::runtime_error(mptc_last_error());
      const mptc::host::EncodedFrame& f = seq.Add(rgb.data(), cfg.width, cfg.height);
      char name[64];
      std::snprintf(name, sizeof name, "/frame%03d", t);
      const std::string base = out + name;
      if (!write_file(base + ".phys", f.physical_blocks.data(), f.physical_blocks.size() * 8) ||
          !write_file(base + ".motion", f.motion_indices.data(), f.motion_indices.size()) ||
          !write_file(base + ".unique", f.unique_palette.data(), f.unique_palette.size() * 4) ||
          !write_file(base + ".eps", f.logical_endpoints.data(), f.logical_endpoints.size())) {
        std::fprintf(stderr, "cannot write %s.*\n", base.c_str());
        return 1;
      }
      std::printf("frame %d %s unique %zu\n", t, f.is_intra ? "intra" : "inter", f.unique_palette.size());
      std::fflush(stdout);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
