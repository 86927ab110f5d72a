#!/bin/sh
# Builds the reference harness, oracle/_ref/libmptc_ref.so, from the reference
# sources where they lie (codec/dxt_image.cpp, codec/image.cpp and the vendored
# stb headers), with the flags of the reference's CMake Release build
# (-std=gnu++11 -O3 -DNDEBUG -fms-extensions).  No -march: the float code stays
# on baseline x86-64 SSE, with no fused multiply-add, as the reference runs.
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
