// The encoder command line over the host layer: the reference's mptc_encoder
// (codec/encoder.cpp:461-506) with the same positional arguments, the encode
// path on the GPU and the entropy stage on the CPU.
//
//   mptc_encode search_area intra_interval unique_interval threshold thread_count dir out_file [out_dir [ep_dir]]
//
// thread_count, out_dir and ep_dir are accepted and not used (the reference
// ignores them too on this path).
#include <cstdio>
#include <cstdlib>
#include <string>

#include "mptc_host.h"

int main(int argc, char** argv) {
  if (argc < 8) {
    std::fprintf(stderr,
                 "usage: %s search_area intra_interval unique_interval threshold thread_count dir out_file "
                 "[out_dir [ep_dir]]\n",
                 argv[0]);
    return 2;
  }
  mptc::host::Options opt;
  opt.search_area = std::atoi(argv[1]);
  opt.intra_interval = std::atoi(argv[2]);
  const int unique_interval = std::atoi(argv[3]);
  opt.err_threshold = std::atoi(argv[4]);
  try {
    const mptc::host::DirectoryResult r = mptc::host::CompressDirectory(argv[6], argv[7], opt, unique_interval);
    std::printf("%d frames encoded, %d written in whole groups, %llu bytes -> %s\n", r.frames, r.frames_written,
                (unsigned long long)r.bytes, argv[7]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
