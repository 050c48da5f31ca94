#!/bin/bash
# Writes fortran_edits.txt: the E15.7 / G15.7 edits of a few values and the E9.2 frequency tag as
# printed by a Fortran compiler (AMD flang by default), the expected strings of
# tests/test_obsgrid.py::test_fortran_edits_match_compiler.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
FC=${FC:-/opt/rocm/llvm/bin/flang}
TMP=$(mktemp -d)
"$FC" -O0 "$HERE/fortran_edits.f90" -o "$TMP/fortran_edits"
"$TMP/fortran_edits" > "$HERE/fortran_edits.txt"
rm -rf "$TMP"
