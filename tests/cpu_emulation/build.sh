#!/bin/bash
# Test infrastructure: compiles the library's .hip sources for the host with the CPU stand-in
# of hip_runtime.h in this folder, under AddressSanitizer, into libtd3_emu.so.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
OUT="$HERE/libtd3_emu.so"
SRCS=$(ls "$ROOT"/tree-d-3d_amd/csrc/*.hip)
g++ -std=c++20 -O1 -g -fPIC -shared -pthread -fsanitize=address -fno-omit-frame-pointer \
    -DTD3_TRAVERSE_BLOCKS=4 -I"$HERE" -I"$ROOT/tree-d-3d_amd/csrc" -I"$ROOT/include" \
    -Wno-unused-result -x c++ $SRCS -o "$OUT"
echo "$OUT"
