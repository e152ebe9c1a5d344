#!/bin/bash
# Build libdml_sim.so: libdml's own kernel and host sources compiled for the
# CPU against the SIMT stand-in for HIP (include/hip/hip_runtime.h, simt.cpp),
# with AddressSanitizer unless SIMT_ASAN=0.  Test infrastructure only.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
src="$here/../../din-mol-li_amd/csrc"
out="${SIMT_OUT:-$here/_build/libdml_sim.so}"
mkdir -p "$(dirname "$out")"
san="-fsanitize=address -fno-omit-frame-pointer"
[[ "${SIMT_ASAN:-1}" == "0" ]] && san=""
flags="-std=c++17 -O2 -g -fPIC -I$here/include -Wno-attributes -Wno-unused-result $san"
g++ $flags -x c++ -c "$src/dml_kernels.hip" -o "$out.k.o"
g++ $flags -x c++ -c "$src/dml_api.hip" -o "$out.a.o"
g++ $flags -c "$here/simt.cpp" -o "$out.s.o"
g++ -shared $san "$out.k.o" "$out.a.o" "$out.s.o" -o "$out"
rm -f "$out.k.o" "$out.a.o" "$out.s.o"
echo "built $out"
