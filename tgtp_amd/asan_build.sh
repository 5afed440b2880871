#!/bin/bash
# Host-side AddressSanitizer build of libtgtp_hip and of the tgtp / benchmark
# binaries, into tgtp_amd/_asan/: the host code of every HIP source (plan
# building, argument checks, launches) and the host-only sources are
# instrumented; device code is not (no GPU sanitizer is used).  The binaries
# take the same arguments as tgtp_amd/bin/tgtp and tgtp_amd/bin/benchmark.
#
# usage: bash tgtp_amd/asan_build.sh
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
C=$ROOT/tgtp_amd/csrc
O=$ROOT/tgtp_amd/_asan
mkdir -p "$O"
HIPFLAGS=(-std=c++20 -O1 -g --offload-arch=gfx950 -fPIC -ffp-contract=off
          -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-command-line-argument
          "-I$ROOT/include" "-I$C"
          -Xarch_host -fsanitize=address -Xarch_host -fno-omit-frame-pointer)
pids=()
for s in capi saba_walk mt_walk tgtp tgtp_fast fast_block taus_fast tgtp_batch; do
  /opt/rocm/bin/hipcc "${HIPFLAGS[@]}" -c "$C/$s.hip" -o "$O/$s.o" & pids+=($!)
done
for s in graph_io host_tgtp; do
  /opt/rocm/llvm/bin/clang++ -std=c++20 -O1 -g -fPIC -fopenmp -ffp-contract=off \
    "-I$ROOT/include" "-I$C" -fsanitize=address -fno-omit-frame-pointer \
    -c "$C/$s.cpp" -o "$O/$s.o" & pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
objs=()
for s in capi saba_walk mt_walk tgtp tgtp_fast fast_block taus_fast tgtp_batch graph_io host_tgtp; do
  objs+=("$O/$s.o")
done
/opt/rocm/bin/hipcc -shared -fPIC --offload-arch=gfx950 -o "$O/libtgtp_hip_asan.so" "${objs[@]}" -fopenmp
for c in tgtp benchmark; do
  /opt/rocm/llvm/bin/clang++ -std=c++20 -O1 -g -fsanitize=address "-I$ROOT/include" \
    "$C/cli_$c.cpp" -o "$O/$c" -L"$O" -ltgtp_hip_asan -Wl,-rpath,'$ORIGIN' \
    -L/opt/rocm/lib -lamdhip64 -Wl,-rpath,/opt/rocm/lib -fopenmp
done
rm -f "${objs[@]}"
echo "built $O"
