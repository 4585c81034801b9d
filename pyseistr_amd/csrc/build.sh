#!/bin/sh
# Builds libseistr_hip.so for gfx950 (in-tree, next to this script's package).
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/../libseistr_hip.so"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
JOBS=${JOBS:-8}
SRCS="common dip"
[ -f "$HERE/spray.hip" ] && SRCS="$SRCS spray"
[ -f "$HERE/interp.hip" ] && SRCS="$SRCS interp"
[ -f "$HERE/cnn.hip" ] && SRCS="$SRCS cnn"
[ -f "$HERE/batch.hip" ] && SRCS="$SRCS batch"
[ -f "$HERE/bandpass.hip" ] && SRCS="$SRCS bandpass"
[ -f "$HERE/patch.hip" ] && SRCS="$SRCS patch"
[ -f "$HERE/fk.hip" ] && SRCS="$SRCS fk"
[ -f "$HERE/boundary.hip" ] && SRCS="$SRCS boundary"
OBJDIR="$HERE/_build"
mkdir -p "$OBJDIR"
FLAGS="--offload-arch=gfx950 -O3 -fPIC -std=c++17 -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-result"
pids=""
for s in $SRCS; do
  $HIPCC $FLAGS -c "$HERE/$s.hip" -o "$OBJDIR/$s.o" &
  pids="$pids $!"
done
for p in $pids; do wait $p; done
$HIPCC --offload-arch=gfx950 -shared -fPIC $(for s in $SRCS; do echo "$OBJDIR/$s.o"; done) -L/opt/rocm/lib -lhipfft \
  -Wl,-rpath,/opt/rocm/lib -o "$OUT.tmp"
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
