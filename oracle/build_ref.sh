#!/bin/bash
# Builds oracle/_ref/libref_cpukernels.so from the reference's own CPU Stokes
# kernels (src/CPUKernels.cc, which compiles standalone) plus ref_shim.cc.
# Only the sources under /root/reference are read; nothing is copied.
# The library is a test-time checker: the product never loads it.
# Loading it needs VES3D_DIR set (include/ves3d_common.h:79 reads it at load)
# and lazy binding (RTLD_LAZY): the SSE kernels' error paths name Logger's
# stream manipulator `alert`, defined in src/Logger.cc, which does not compile
# with g++ 11; those paths are only reached for src_stride % 4 != 0, which the
# tests never pass to the SSE kernels.
set -e
REF=${VES3D_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
if [ ! -f "$REF/src/CPUKernels.cc" ]; then
  echo "reference sources not present; skipping oracle/_ref build"; exit 0
fi
mkdir -p "$HERE/_ref"
g++ -O2 -fopenmp -msse3 -fPIC -shared -DDOUBLE_PREC -DNDEBUG \
    -I"$REF/include" -I"$REF/src" \
    "$REF/src/CPUKernels.cc" "$HERE/ref_shim.cc" \
    -Wl,-z,lazy -o "$HERE/_ref/libref_cpukernels.so"
echo "built $HERE/_ref/libref_cpukernels.so"
