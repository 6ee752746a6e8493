#!/usr/bin/env bash
# AddressSanitizer + UndefinedBehaviorSanitizer build of libpkdecon's host
# code: every petakit5d_amd/csrc/*.hip source is compiled for the host only
# (each translation unit embeds an empty offload bundle in place of its device
# code objects, so it registers no kernels and runs only where nothing
# launches one: a host without a GPU), then linked into
#   tests/sanitize/_build/pk_host_san   (tests/sanitize/pk_host_san.hip)
#   tests/sanitize/_build/mexdrv_san    (matlab/pk_decon_mex.c on tests/mex_harness)
# Objects are rebuilt only when their source or a header is newer.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
SRC="$ROOT/petakit5d_amd/csrc"
OUT="$HERE/_build"
mkdir -p "$OUT/obj"
SAN=(-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined
     -Xarch_host -fno-sanitize-recover=undefined -fno-omit-frame-pointer)
# an offload bundle holding only an empty host entry
BUNDLER=/opt/rocm/lib/llvm/bin/clang-offload-bundler
: > "$OUT/empty.bin"
"$BUNDLER" -type=o -targets=host-x86_64-unknown-linux-gnu -input="$OUT/empty.bin" -output="$OUT/empty.hipfb"
FLAGS=(--cuda-host-only -O1 -g -std=c++17 -fPIC -Wno-unused-function "${SAN[@]}"
       -Xclang -fcuda-include-gpubinary -Xclang "$OUT/empty.hipfb")
newest_hdr=$(ls -t "$SRC"/*.hpp "$ROOT/include/pkdecon.h" | head -1)
units=()
for f in "$SRC"/*.hip "$HERE/pk_host_san.hip"; do
  o="$OUT/obj/$(basename "$f" .hip).o"
  if [[ ! -f "$o" || "$f" -nt "$o" || "$newest_hdr" -nt "$o" ]]; then units+=("$f"); fi
done
if ((${#units[@]})); then
  printf '%s\n' "${units[@]}" | xargs -P "${PK_JOBS:-8}" -I{} bash -c \
    'f="{}"; hipcc "$@" -c "$f" -o "'"$OUT"'/obj/$(basename "$f" .hip).o" 2> >(grep -v "packed-fp32-ops" >&2)' _ "${FLAGS[@]}"
fi
lib_objs=()
for f in "$SRC"/*.hip; do lib_objs+=("$OUT/obj/$(basename "$f" .hip).o"); done
hipcc --cuda-host-only "${SAN[@]}" -o "$OUT/pk_host_san" "$OUT/obj/pk_host_san.o" "${lib_objs[@]}" -lpthread
# the MEX gateway and its stand-in mx API are C: built with the same sanitizers
for c in "$ROOT/matlab/pk_decon_mex.c" "$ROOT/tests/mex_harness/mx_fake.c"; do
  hipcc -x c --cuda-host-only -O1 -g "${SAN[@]}" -I "$ROOT/tests/mex_harness" -I "$ROOT/include" \
    -c "$c" -o "$OUT/obj/$(basename "$c" .c).c.o"
done
hipcc --cuda-host-only "${SAN[@]}" -o "$OUT/mexdrv_san" "$OUT/obj/pk_decon_mex.c.o" "$OUT/obj/mx_fake.c.o" \
  "${lib_objs[@]}" -lpthread
echo "built $OUT/pk_host_san $OUT/mexdrv_san"
