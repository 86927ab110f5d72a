#!/bin/sh
# Builds the reference harnesses from the reference sources where they lie,
# with the flags of the reference's CMake Release build (-std=gnu++11 -O3
# -DNDEBUG -fms-extensions).  No -march: the float code stays on baseline
# x86-64 SSE, with no fused multiply-add, as the reference runs.
#   oracle/_ref/libmptc_ref.so        the encode path: codec/dxt_image.cpp,
#                                     codec/image.cpp, the vendored stb headers
#   oracle/_ref/libmptc_ref_codec.so  the whole file writer: codec/codec.cpp,
#                                     dxt_image.cpp, image.cpp, wavelet.cpp,
#                                     image_processing.cpp and
#                                     entropy/arithmetic_codec.cpp (the source
#                                     list of the reference's mptc_codec and
#                                     arith_codec libraries)
# Output goes only to oracle/_ref/.
set -e
REF="${MPTC_REFERENCE:-/root/reference}"
HERE="$(cd "$(dirname "$0")" && pwd)"
if [ ! -f "$REF/codec/dxt_image.cpp" ]; then
  echo "reference sources not found at $REF; skipping oracle/_ref" >&2
  exit 0
fi
mkdir -p "$HERE/_ref"
g++ -std=gnu++11 -O3 -DNDEBUG -fms-extensions -fPIC -shared -w \
  -I"$REF/codec" -I"$REF/Include" \
  "$HERE/ref_harness.cpp" "$REF/codec/image.cpp" \
  -o "$HERE/_ref/libmptc_ref.so.tmp$$"
mv -f "$HERE/_ref/libmptc_ref.so.tmp$$" "$HERE/_ref/libmptc_ref.so"
echo "built $HERE/_ref/libmptc_ref.so"
g++ -std=gnu++11 -O3 -DNDEBUG -fms-extensions -fPIC -shared -w \
  -I"$REF/codec" -I"$REF/Include" -I"$REF/entropy" \
  "$HERE/ref_codec_harness.cpp" "$REF/codec/codec.cpp" "$REF/codec/dxt_image.cpp" "$REF/codec/image.cpp" \
  "$REF/codec/wavelet.cpp" "$REF/codec/image_processing.cpp" "$REF/entropy/arithmetic_codec.cpp" \
  -lpthread -o "$HERE/_ref/libmptc_ref_codec.so.tmp$$"
mv -f "$HERE/_ref/libmptc_ref_codec.so.tmp$$" "$HERE/_ref/libmptc_ref_codec.so"
echo "built $HERE/_ref/libmptc_ref_codec.so"
