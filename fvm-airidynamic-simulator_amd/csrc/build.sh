#!/bin/bash
# Builds libfvm_amd.so (gfx950) in-tree.  Usage: build.sh [outdir]
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="${1:-$HERE/../lib}"
mkdir -p "$OUT" "$HERE/../build"
FLAGS="--offload-arch=gfx950 -O3 -std=c++17 -fPIC -Wall -Xclang -target-feature -Xclang -packed-fp32-ops"
pids=()
for f in fvm_kernels fvm_engine fvm_api fvm_noalloc; do
  hipcc $FLAGS -c "$HERE/$f.hip" -o "$HERE/../build/$f.o" 2> "$HERE/../build/$f.log" &
  pids+=($!)
done
rc=0
for i in "${!pids[@]}"; do
  if ! wait "${pids[$i]}"; then rc=1; fi
done
for f in fvm_kernels fvm_engine fvm_api fvm_noalloc; do grep -v "packed-fp32-ops" "$HERE/../build/$f.log" | grep -v "^$" || true; done
[ $rc -eq 0 ] || { echo "compile failed"; exit 1; }
hipcc --offload-arch=gfx950 -shared -fPIC -o "$OUT/libfvm_amd.so.tmp" "$HERE/../build/fvm_kernels.o" "$HERE/../build/fvm_engine.o" "$HERE/../build/fvm_api.o" "$HERE/../build/fvm_noalloc.o"
mv "$OUT/libfvm_amd.so.tmp" "$OUT/libfvm_amd.so"
echo "built $OUT/libfvm_amd.so"
# KernelBenchmark counterpart (host C++ over the C ABI)
g++ -std=c++17 -O2 -I "$HERE/../../include" "$HERE/../apps/kernel_benchmark.cpp" -o "$OUT/kernel_benchmark" \
    -L "$OUT" -lfvm_amd -Wl,-rpath,'$ORIGIN' -Wl,-rpath,/opt/rocm/lib
echo "built $OUT/kernel_benchmark"
# the reference front end's call order through the C++ drop-in (drop-in GPU test)
g++ -std=c++17 -O2 -D__HIP_PLATFORM_AMD__ -I "$HERE/../../include" -I /opt/rocm/include \
    "$HERE/../apps/front_end_sequence.cpp" -o "$OUT/front_end_sequence" \
    -L "$OUT" -lfvm_amd -L /opt/rocm/lib -lamdhip64 -Wl,-rpath,'$ORIGIN' -Wl,-rpath,/opt/rocm/lib
echo "built $OUT/front_end_sequence"
