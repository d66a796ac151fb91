#!/bin/bash
# Builds libhsba_cpu.so: the library's own sources (kernels included) compiled
# for the host against the CPU HIP emulation, with AddressSanitizer on.
# Fails if any compile fails.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
SRC=$HERE/../../hs_sfm_amd/csrc
OUT=$HERE/_build
mkdir -p "$OUT"
CXX=/opt/rocm/llvm/bin/clang++
FLAGS="-std=c++20 -O1 -g -fPIC -DHSBA_CPU_EMULATION -fsanitize=address -fno-omit-frame-pointer -I$HERE/include -Wno-unused-result -Wno-psabi -Wno-unknown-attributes"
pids=()
for f in hsba_kernels.hip hsba_cholesky.hip hsba_scratch.hip hsba_dense.hip hsba_layout.cpp hsba_engine.cpp hsba_server.cpp hsba_naive.cpp hsba_io.cpp hsba_api.cpp hsba_tune.cpp; do
  if [ "$SRC/$f" -nt "$OUT/$f.o" ] || [ "$HERE/include/hip/hip_runtime.h" -nt "$OUT/$f.o" ] || [ "$SRC/hsba_device.h" -nt "$OUT/$f.o" ] || [ "$SRC/hsba_kernels.h" -nt "$OUT/$f.o" ] || [ "$SRC/hsba_host.h" -nt "$OUT/$f.o" ]; then
    $CXX $FLAGS -x c++ -c "$SRC/$f" -o "$OUT/$f.o" &
    pids+=($!)
  fi
done
if [ "$HERE/hipemu.cpp" -nt "$OUT/hipemu.o" ] || [ "$HERE/include/hip/hip_runtime.h" -nt "$OUT/hipemu.o" ]; then
  $CXX $FLAGS -c "$HERE/hipemu.cpp" -o "$OUT/hipemu.o" &
  pids+=($!)
fi
status=0
for p in "${pids[@]}"; do wait "$p" || status=1; done
if [ $status -ne 0 ]; then echo "cpu emulation build failed" >&2; exit 1; fi
$CXX -shared -fsanitize=address -shared-libasan -o "$OUT/libhsba_cpu.so" "$OUT"/*.o -lpthread
echo "preload: $($CXX -print-file-name=libclang_rt.asan-x86_64.so)" >&2
echo "$OUT/libhsba_cpu.so"
