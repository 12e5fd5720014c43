#!/bin/bash
# Builds oracle/_ref/libref_gs_sort.so from the reference's own Fortran:
# the integer-only subroutine gs_qsort_dofmap of
# src/gs/gather_scatter.f90 (a procedure contained in gs_init_mapping),
# compiled with the image's flang inside a one-procedure wrapper module
# that gives it a C entry point.  The reference text is read where it lies
# and assembled in a temporary directory outside the repository; only the
# shared library is written to oracle/_ref/ (kept out of git).
#
# Test infrastructure only: the built library is the checker of the
# oracle's C restatement of the sort (tie order inside gs blocks).
set -euo pipefail
REF=${NEKO_REFERENCE:-/root/reference}
SRC="$REF/src/gs/gather_scatter.f90"
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/_ref"
FC=${FC:-/opt/rocm/llvm/bin/flang}
[ -f "$SRC" ] || { echo "build_ref: $SRC not found (reference absent)"; exit 3; }
[ -x "$FC" ] || { echo "build_ref: no Fortran compiler at $FC"; exit 3; }
TMP="$(mktemp -d)"
trap 'rm -rf "$TMP"' EXIT
{
  echo "module ref_gs_sort"
  echo "  use iso_c_binding"
  echo "  implicit none"
  echo "contains"
  awk '/recursive subroutine gs_qsort_dofmap/{f=1} f{print} /end subroutine gs_qsort_dofmap/{f=0}' "$SRC"
  cat <<'F90'
  subroutine ref_gs_qsort_dofmap(dg, gd, n) bind(c, name='ref_gs_qsort_dofmap')
    integer(c_int), value :: n
    integer(c_int), intent(inout) :: dg(n), gd(n)
    integer :: nn
    nn = n
    if (nn .gt. 0) call gs_qsort_dofmap(dg, gd, nn, 1, nn)
  end subroutine ref_gs_qsort_dofmap
end module ref_gs_sort
F90
} > "$TMP/ref_gs_sort.f90"
grep -q "recursive subroutine gs_qsort_dofmap" "$TMP/ref_gs_sort.f90"
mkdir -p "$OUT"
(cd "$TMP" && "$FC" -O2 -fPIC -shared -o "$OUT/libref_gs_sort.so.tmp" ref_gs_sort.f90)
mv "$OUT/libref_gs_sort.so.tmp" "$OUT/libref_gs_sort.so"
echo "built $OUT/libref_gs_sort.so"
