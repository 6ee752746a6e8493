#!/usr/bin/env bash
# Builds petakit5d_amd/libpkdecon.so (HIP, gfx950) from petakit5d_amd/csrc.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
SRC="$HERE/csrc"
OBJ="$HERE/csrc/obj"
mkdir -p "$OBJ"
FLAGS=(--offload-arch=gfx950 -O3 -std=c++17 -fPIC -Wall -Wno-unused-function
       -Xclang -target-feature -Xclang -packed-fp32-ops)
pids=()
for f in pk_kern_x pk_kern_col pk_misc pk_engine pk_capi; do
  hipcc "${FLAGS[@]}" -c "$SRC/$f.hip" -o "$OBJ/$f.o" 2> >(grep -v "packed-fp32-ops" >&2) &
  pids+=($!)
done
rc=0
for p in "${pids[@]}"; do wait "$p" || rc=1; done
[ $rc -eq 0 ] || { echo "compile failed" >&2; exit 1; }
hipcc --offload-arch=gfx950 -shared -fPIC -o "$HERE/libpkdecon.so.tmp" "$OBJ"/pk_kern_x.o "$OBJ"/pk_kern_col.o \
  "$OBJ"/pk_misc.o "$OBJ"/pk_engine.o "$OBJ"/pk_capi.o
mv -f "$HERE/libpkdecon.so.tmp" "$HERE/libpkdecon.so"
