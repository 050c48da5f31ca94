#!/usr/bin/env bash
# Builds the reference's own hot-path Fortran (read from /root/reference, never copied)
# plus oracle/ref_harness.f90 into oracle/_ref/ref_harness with AMD flang. Test
# infrastructure only. Built with -ffp-contract=off: the reference Makefile targets -xAVX, which
# has no fused multiply-add, so every product is rounded on its own. forward_bxyz_mpi.f90 also holds forward_bxyz, whose solvePARDISO
# call needs Intel MKL (not in this image): that one symbol is left unresolved at link
# time; the harness never calls forward_bxyz.
set -euo pipefail
REF=${REF:-/root/reference/src}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-/opt/rocm/llvm/bin/flang}
if [ ! -d "$REF" ]; then
  echo "reference sources not found at $REF; skipping oracle/_ref" >&2
  exit 0
fi
mkdir -p "$OUT/mod" "$OUT/obj"
SRCS=(
  solver/m_constants.f90
  common/m_outerinnerproduct.f90
  common/m_matrix.f90
  solver/m_fem_util.f90
  solver/m_fem_edge_util.f90
  common/m_line_type.f90
  common/m_param.f90
  common/m_mesh_type.f90
  solver/m_iccg_var_takuto.f90
  src_tide/m_readfvxyz.f90
  solver/COCG13_single.f90
  solver/sort_index.f90
  solver_mpi/forward_bxyz_mpi.f90
)
OBJS=()
for s in "${SRCS[@]}"; do
  o="$OUT/obj/$(basename "$s" .f90).o"
  "$FC" -O2 -ffp-contract=off -J"$OUT/mod" -c "$REF/$s" -o "$o" 2>"$OUT/obj/$(basename "$s" .f90).log"
  OBJS+=("$o")
done
"$FC" -O2 -J"$OUT/mod" -I"$OUT/mod" -c "$HERE/ref_harness.f90" -o "$OUT/obj/ref_harness.o"
"$FC" -O2 "$OUT/obj/ref_harness.o" "${OBJS[@]}" -o "$OUT/ref_harness" -Wl,--unresolved-symbols=ignore-in-object-files
echo "built $OUT/ref_harness"
