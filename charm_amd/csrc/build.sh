#!/bin/sh
# Builds libcharm_amd.so for gfx950 from the sources in this folder.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=${1:-$HERE/../libcharm_amd.so}
ROCM=${ROCM_PATH:-/opt/rocm}
mkdir -p "$HERE/_build"
HIPFLAGS="--offload-arch=gfx950 -O3 -std=c++17 -fPIC -I$HERE/../../include -I$HERE \
 -Xclang -target-feature -Xclang -packed-fp32-ops -Wall -Wno-unused-function"
pids=""
for f in kernels fft multi point_api server gfm noalloc; do
  rm -f "$HERE/_build/$f.o"
  "$ROCM/bin/hipcc" $HIPFLAGS -c "$HERE/$f.hip" -o "$HERE/_build/$f.o" &
  pids="$pids $!"
done
for f in host_api shc_io shc_ops; do
  rm -f "$HERE/_build/$f.o"
  g++ -O2 -std=c++17 -fPIC -Wall -I$HERE/../../include -I$HERE -c "$HERE/$f.cpp" \
    -o "$HERE/_build/$f.o" &
  pids="$pids $!"
done
# every compile must succeed (a bare "wait" would hide a failed one)
for p in $pids; do wait "$p"; done
"$ROCM/bin/hipcc" --offload-arch=gfx950 -shared -fPIC -o "$OUT.tmp" "$HERE/_build/kernels.o" \
  "$HERE/_build/fft.o" "$HERE/_build/multi.o" "$HERE/_build/point_api.o" "$HERE/_build/server.o" "$HERE/_build/gfm.o" "$HERE/_build/noalloc.o" "$HERE/_build/host_api.o" "$HERE/_build/shc_io.o" "$HERE/_build/shc_ops.o" -L"$ROCM/lib" -lhipfft -lamdhip64 \
  -Wl,-rpath,"$ROCM/lib" -lpthread -Wl,--no-undefined
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
