#!/bin/bash
# Test infrastructure: compiles the library's .hip sources for the host with the CPU stand-in
# of hip_runtime.h in this folder, under AddressSanitizer, into libtd3_emu.so.  One object
# per source in _obj/, compiled in parallel and only when the source, a header or this
# script is newer than it.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
OUT="$HERE/libtd3_emu.so"
OBJ="$HERE/_obj"
mkdir -p "$OBJ"
export EMU_FLAGS="-std=c++20 -O1 -g -fPIC -pthread -fsanitize=address -fno-omit-frame-pointer \
-DTD3_TRAVERSE_BLOCKS=4 -I$HERE -I$ROOT/tree-d-3d_amd/csrc -I$ROOT/include -Wno-unused-result"
HEADERS=("$HERE"/hip/*.h "$ROOT"/tree-d-3d_amd/csrc/*.h "$ROOT/include/td3.h" "$HERE/build.sh")
stale=()
objects=()
for src in "$ROOT"/tree-d-3d_amd/csrc/*.hip; do
    obj="$OBJ/$(basename "$src" .hip).o"
    objects+=("$obj")
    if [ ! -e "$obj" ] || [ "$src" -nt "$obj" ]; then
        stale+=("$src")
        continue
    fi
    for h in "${HEADERS[@]}"; do
        if [ "$h" -nt "$obj" ]; then
            stale+=("$src")
            break
        fi
    done
done
if [ ${#stale[@]} -gt 0 ]; then
    printf '%s\n' "${stale[@]}" | xargs -P 8 -I{} sh -c \
        'g++ $EMU_FLAGS -x c++ -c "$1" -o "$2/$(basename "$1" .hip).o"' _ {} "$OBJ"
fi
relink=0
[ -e "$OUT" ] || relink=1
for obj in "${objects[@]}"; do
    [ "$obj" -nt "$OUT" ] && relink=1
done
if [ $relink -eq 1 ]; then
    g++ -shared -fsanitize=address -pthread "${objects[@]}" -o "$OUT"
fi
echo "$OUT"
