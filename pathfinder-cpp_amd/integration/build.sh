#!/usr/bin/env bash
# Builds integration/_build/check: the reference-side binding (hip_gpu.cpp, renderer_hip.cpp) and
# its driver (check_main.cpp) compiled against the reference's headers, linked with the reference's
# own scene / path / stroke / paint / renderer / gpu sources (compiled here from /root/reference,
# unchanged, with the flags its CMakeLists.txt:51-67 selects on Linux/x86) and with libpfh.so.
# Outputs go to integration/_build/ only.  Needs the reference tree: skipped without it.
set -euo pipefail
REF=${PF_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
PKG=$(cd "$HERE/.." && pwd)
OUT="$HERE/_build"
if [ ! -d "$REF/pathfinder" ]; then
  echo "reference tree not present at $REF; skipping the integration build" >&2
  exit 0
fi
mkdir -p "$OUT/obj"
P="$REF/pathfinder"
SRCS=(
  common/color.cpp common/math/transform2.cpp common/math/unit_vector.cpp common/timestamp.cpp common/io.cpp
  core/data/built_path.cpp core/data/contour.cpp core/data/line_segment.cpp core/data/path.cpp core/data/segment.cpp
  core/scene.cpp core/scene_builder.cpp core/stroke.cpp core/path2d.cpp
  core/paint/palette.cpp core/paint/paint.cpp core/paint/gradient.cpp core/paint/pattern.cpp
  core/paint/texture_allocator.cpp core/renderer.cpp gpu_mem/allocator.cpp
  gpu/device.cpp gpu/command_encoder.cpp gpu/shader.cpp
)
CXXFLAGS=(-std=c++17 -O2 -fPIC -msse4.1 -DPATHFINDER_ENABLE_SIMD -I"$REF" -I"$P" -I"$REF/third_party" -w)
objs=()
for s in "${SRCS[@]}"; do
  o="$OUT/obj/$(echo "$s" | tr '/' '_' | sed 's/\.cpp$/.o/')"
  if [ ! -f "$o" ] || [ "$P/$s" -nt "$o" ]; then
    g++ "${CXXFLAGS[@]}" -c "$P/$s" -o "$o"
  fi
  objs+=("$o")
done
HIPFLAGS=(-D__HIP_PLATFORM_AMD__ -I/opt/rocm/include -I"$PKG/../include")
for s in hip_gpu renderer_hip check_main; do
  g++ "${CXXFLAGS[@]}" -Wall -Wno-unused-function "${HIPFLAGS[@]}" -I"$HERE" -c "$HERE/$s.cpp" -o "$OUT/obj/$s.o"
done
g++ -o "$OUT/check" "$OUT/obj/check_main.o" "$OUT/obj/renderer_hip.o" "$OUT/obj/hip_gpu.o" "${objs[@]}" \
  -L"$PKG" -lpfh -L/opt/rocm/lib -lamdhip64 -Wl,-rpath,'$ORIGIN/../..' -Wl,-rpath,/opt/rocm/lib -lpthread
echo "built $OUT/check"
