#!/bin/sh
# Compiles the reference's hot-path CPython extension modules (dip_cfuns.c,
# sof_cfuns.c, sof3d_cfuns.c, soint2d_cfuns.c, soint3d_cfuns.c) and the
# band-pass module next to the path (bp_cfuns.c)
# from their own sources, where they lie, into oracle/_ref/.  The outputs are
# test/baseline infrastructure only: they check the C restatement in oracle/
# and serve as the CPU baseline; they are never shipped in a commit.
# Usage: sh oracle/build_ref.sh [reference_root]
set -e
REF="${1:-${PYSEISTR_REF:-/root/reference}}/pyseistr/src"
OUT="$(cd "$(dirname "$0")" && pwd)/_ref"
if [ ! -f "$REF/dip_cfuns.c" ]; then
  echo "build_ref: reference sources not found under $REF; skipping" >&2
  exit 0
fi
PY=${PYTHON:-python3}
PYINC=$($PY -c "import sysconfig; print(sysconfig.get_paths()['include'])")
NPINC=$($PY -c "import numpy; print(numpy.get_include())")
SUF=$($PY -c "import sysconfig; print(sysconfig.get_config_var('EXT_SUFFIX'))")
mkdir -p "$OUT"
for m in dip:dipcfun sof:sofcfun sof3d:sof3dcfun soint2d:soint2dcfun soint3d:soint3dcfun bp:bpcfun; do
  src=${m%%:*}; mod=${m##*:}
  gcc -O2 -fPIC -shared -ffp-contract=off -w -I"$PYINC" -I"$NPINC" \
      "$REF/${src}_cfuns.c" -o "$OUT/$mod$SUF" -lm
done
echo "build_ref: built $(ls "$OUT" | wc -l) modules into $OUT"
