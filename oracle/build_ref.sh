#!/bin/bash
# Build the reference-side checkers into oracle/_ref/ (TEST INFRASTRUCTURE).
#
# The reference's cell/neighbour modules are compiled unmodified from where
# they lie under /root/reference/src with amdflang (its build system needs
# autotools, which this image lacks, so the few files are compiled directly,
# in the order of src/Makefile.am's module dependencies).  Only our own
# harness (oracle/ref_harness/nbdump.F90) is linked against them.  Nothing is
# copied from the reference; outputs go to oracle/_ref/ only.
#
# Usage: oracle/build_ref.sh      (no-op with a message when /root/reference
#                                  or amdflang is missing)
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
ref=/root/reference
out="$here/_ref"
fc=${FC:-/opt/rocm/bin/amdflang}
if [[ ! -d "$ref/src" ]]; then
  echo "build_ref: $ref not present; keeping prebuilt $out" >&2
  exit 0
fi
if [[ ! -x "$fc" ]]; then
  echo "build_ref: $fc not found; skipping" >&2
  exit 0
fi
mkdir -p "$out/obj"
flags=(-cpp -O0 -DDIM3 "-I$ref/lib/fpt/include" -J "$out/obj")
objs=()
for f in Constants.F90 Strings.f90 Errors.f90 Elements.F90 Program_Types.F90 \
         Groups.F90 Cells.F90 Neighbor.F90; do
  o="$out/obj/${f%.*}.o"
  "$fc" "${flags[@]}" -c "$ref/src/$f" -o "$o" 2>/dev/null
  objs+=("$o")
done
"$fc" "${flags[@]}" -c "$here/ref_harness/nbdump.F90" -o "$out/obj/nbdump.o"
"$fc" -o "$out/nbdump" "$out/obj/nbdump.o" "${objs[@]}"
echo "build_ref: built $out/nbdump"
