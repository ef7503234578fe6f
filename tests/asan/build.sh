#!/bin/bash
# Host-side AddressSanitizer build of libfvm_amd (device code as usual) and of
# driver.c, a C program that walks the C ABI's host paths: the plain solver
# with graph replay, snapshots, obstacle update, tile override, timing, resize
# and reset; the batch with admission, retirement and case save / load.
# Output goes to tests/asan/_build (not committed).  Run on a GPU host:
#   bash tests/asan/build.sh
#   ASAN_OPTIONS=detect_leaks=0:verify_asan_link_order=0 tests/asan/_build/driver
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
R="$(cd "$HERE/../.." && pwd)"
C=$R/fvm-airidynamic-simulator_amd/csrc
O=$HERE/_build
mkdir -p $O/obj $O/lib
FLAGS="--offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -Xclang -target-feature -Xclang -packed-fp32-ops -Xarch_host -fsanitize=address -Xarch_host -fno-omit-frame-pointer"
for f in fvm_kernels fvm_engine fvm_api; do
  hipcc $FLAGS -c $C/$f.hip -o $O/obj/$f.o 2>&1 | grep -v "packed-fp32-ops" | grep -v "^$" || true
done
hipcc --offload-arch=gfx950 -shared -fPIC -Xarch_host -fsanitize=address -o $O/lib/libfvm_amd.so \
    $O/obj/fvm_kernels.o $O/obj/fvm_engine.o $O/obj/fvm_api.o
RT="$(dirname "$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)")"
/opt/rocm/llvm/bin/clang -std=c99 -g -fsanitize=address -shared-libasan -fno-omit-frame-pointer -I $R/include \
    $HERE/driver.c -o $O/driver -L $O/lib -lfvm_amd -Wl,-rpath,$O/lib -Wl,-rpath,/opt/rocm/lib -Wl,-rpath,$RT
echo "built $O/driver"
