#!/bin/bash
# Host-side AddressSanitizer build of libthp's sources linked into the
# serving-stream driver (device code is not instrumented): stream_asan in
# this folder.  Run on a GPU host as
#   ASAN_OPTIONS=detect_leaks=0:halt_on_error=1 tests/asan/stream_asan [systems] [requests] [rate]
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
S=$HERE/../../tinker-hp_amd/csrc
B=$(mktemp -d)
F="--offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -Wno-unused-function -Xclang -target-feature -Xclang -packed-fp32-ops -munsafe-fp-atomics -fno-omit-frame-pointer"
/opt/rocm/bin/hipcc $F -Xarch_host -fsanitize=address -c -o $B/thp_api.o $S/thp_api.hip &
/opt/rocm/bin/hipcc $F -Xarch_host -fsanitize=address -c -o $B/thp_serve.o $S/thp_serve.hip &
/opt/rocm/llvm/bin/clang++ -O1 -g -std=c++17 -fPIC -fsanitize=address -fno-omit-frame-pointer -c -o $B/thp_systems.o $S/thp_systems.cpp &
/opt/rocm/llvm/bin/clang++ -O1 -g -std=c++17 -fsanitize=address -fno-omit-frame-pointer -c -o $B/stream_driver.o $HERE/stream_driver.cpp &
wait
/opt/rocm/llvm/bin/clang++ -fsanitize=address -o $HERE/stream_asan $B/stream_driver.o $B/thp_api.o $B/thp_serve.o $B/thp_systems.o -L/opt/rocm/lib -lamdhip64 -Wl,-rpath,/opt/rocm/lib -lpthread
rm -rf $B
