#!/usr/bin/env bash
# Builds petakit5d_amd/libpkdecon.so (HIP, gfx950) from petakit5d_amd/csrc.
# Translation units compile in parallel (at most $PK_JOBS at a time, default 8).
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
SRC="$HERE/csrc"
OBJ="$HERE/csrc/obj"
mkdir -p "$OBJ"
FLAGS=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC -Wall -Wno-unused-function
       -Xclang -target-feature -Xclang -packed-fp32-ops)
JOBS="${PK_JOBS:-8}"
units=()
for f in "$SRC"/*.hip; do units+=("$(basename "$f" .hip)"); done
printf '%s\n' "${units[@]}" | xargs -P "$JOBS" -I{} bash -c \
  'hipcc "$@" -c "'"$SRC"'/{}.hip" -o "'"$OBJ"'/{}.o" 2> >(grep -v "packed-fp32-ops" >&2)' _ "${FLAGS[@]}"
objs=()
for u in "${units[@]}"; do objs+=("$OBJ/$u.o"); done
hipcc --offload-arch=gfx950 -shared -fPIC -o "$HERE/libpkdecon.so.tmp" "${objs[@]}"
mv -f "$HERE/libpkdecon.so.tmp" "$HERE/libpkdecon.so"
