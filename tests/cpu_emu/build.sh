#!/bin/sh
# Builds libhsba_cpu.so: the library's own sources (kernels included) compiled
# for the host against the CPU HIP emulation, with AddressSanitizer on.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
SRC=$HERE/../../hs_sfm_amd/csrc
OUT=$HERE/_build
mkdir -p "$OUT"
FLAGS="-std=c++20 -O1 -g -fPIC -fsanitize=address -fno-omit-frame-pointer -I$HERE/include -Wno-unused-result"
for f in hsba_kernels.hip hsba_cholesky.hip hsba_layout.cpp hsba_engine.cpp hsba_server.cpp hsba_api.cpp; do
  [ "$SRC/$f" -nt "$OUT/$f.o" ] && g++ $FLAGS -x c++ -c "$SRC/$f" -o "$OUT/$f.o" &
done
g++ $FLAGS -c "$HERE/hipemu.cpp" -o "$OUT/hipemu.o" &
wait
g++ -shared -fsanitize=address -o "$OUT/libhsba_cpu.so" "$OUT"/*.o -lpthread
echo "$OUT/libhsba_cpu.so"
