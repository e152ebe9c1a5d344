#!/bin/bash
# Build libdml.so (HIP, gfx950) in-tree.  Used by __graft_entry__.build().
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
src="$here/csrc"
out="$here/dml/libdml.so"
hipcc=${HIPCC:-/opt/rocm/bin/hipcc}
"$hipcc" -O3 --offload-arch=gfx950 -std=c++17 -fPIC -shared \
  -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Wall -Wno-unused-function -Wno-unused-value -Wno-unused-result \
  "$src/dml_kernels.hip" "$src/dml_api.hip" -o "$out.tmp" 2>&1 \
  | grep -v "packed-fp32-ops" || true
[[ -f "$out.tmp" ]] || { echo "build failed" >&2; exit 1; }
mv "$out.tmp" "$out"
echo "built $out"

# host-side known-answer check of the Philox generator the kernels use
"$hipcc" -O2 -std=c++17 "$src/philox_kat.cpp" -o "$here/dml/philox_kat" 2>&1 \
  | grep -v "packed-fp32-ops" || true
[[ -x "$here/dml/philox_kat" ]] || { echo "philox_kat build failed" >&2; exit 1; }

# the Fortran side: ISO_C_BINDING interfaces and the dana driver on libdml
fc=${FC_AMD:-/opt/rocm/bin/amdflang}
if [[ -x "$fc" ]]; then
  mod="$here/build/fmod"
  mkdir -p "$mod"
  "$fc" -O2 -c "$here/fortran/dml_f.F90" -module-dir "$mod" -o "$mod/dml_f.o"
  "$fc" -O2 "$here/fortran/dana_dml.F90" "$mod/dml_f.o" -module-dir "$mod" -I "$mod" \
    -L "$here/dml" -ldml -Wl,-rpath,'$ORIGIN' -o "$here/dml/dana_dml"
  echo "built $here/dml/dana_dml"
fi
