#!/usr/bin/env bash
# Test infrastructure only.  Builds oracle/_ref/libpfref.so: the reference's
# hybrid-mode CPU tiler compiled from its own unchanged sources where they lie
# under /root/reference, linked with oracle/ref_harness.cpp (our driver).
# The reference's build system is not used; each translation unit is compiled
# directly with the flags its CMakeLists.txt:51-67 selects on Linux/x86
# (-msse4.1, PATHFINDER_ENABLE_SIMD).  Outputs go to oracle/_ref/ only.
set -euo pipefail
REF=${PF_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/pathfinder" ]; then
  echo "reference tree not present at $REF; skipping oracle/_ref build" >&2
  exit 0
fi
mkdir -p "$OUT/obj"
P="$REF/pathfinder"
SRCS=(
  common/color.cpp common/math/transform2.cpp common/math/unit_vector.cpp common/timestamp.cpp
  core/data/built_path.cpp core/data/contour.cpp core/data/line_segment.cpp core/data/path.cpp
  core/data/segment.cpp
  core/d3d9/tiler.cpp core/d3d9/object_builder.cpp core/d3d9/scene_builder.cpp core/d3d9/data/alpha_tile_id.cpp
  core/scene.cpp core/scene_builder.cpp core/stroke.cpp core/dash.cpp core/path2d.cpp
  core/paint/palette.cpp core/paint/paint.cpp core/paint/gradient.cpp core/paint/pattern.cpp
  core/paint/texture_allocator.cpp core/renderer.cpp gpu_mem/allocator.cpp common/io.cpp
  gpu/device.cpp gpu/command_encoder.cpp gpu/shader.cpp
)
CXXFLAGS=(-std=c++17 -O2 -fPIC -msse4.1 -ffp-contract=off -DPATHFINDER_ENABLE_SIMD
          -I"$REF" -I"$P" -I"$REF/third_party" -w)
objs=()
for s in "${SRCS[@]}"; do
  o="$OUT/obj/$(echo "$s" | tr '/' '_' | sed 's/\.cpp$/.o/')"
  if [ ! -f "$o" ] || [ "$P/$s" -nt "$o" ]; then
    g++ "${CXXFLAGS[@]}" -c "$P/$s" -o "$o"
  fi
  objs+=("$o")
done
g++ "${CXXFLAGS[@]}" -I"$P" -c "$HERE/ref_harness.cpp" -o "$OUT/obj/ref_harness.o"
g++ -shared -o "$OUT/libpfref.so" "$OUT/obj/ref_harness.o" "${objs[@]}" -lpthread
echo "built $OUT/libpfref.so"
