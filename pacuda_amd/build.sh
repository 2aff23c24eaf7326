#!/bin/sh
# SPDX-License-Identifier: 0BSD
# Builds librecurrence.so for gfx950 with runner/build.sh's arguments, so that callers of that
# script (pacuda.py's retarget step) can call this one instead.
# usage: [PACUDA_KERNELS=kernels.hpp] pacuda_amd/build.sh [OUT_DIR] [TARGET...]
#   TARGET: gfx950 (the default and the only target); CUDA SM numbers (90, 120, ...) name no
#   AMD target and are ignored. PACUDA_KERNELS: the symjit kernels to build in, a header written
#   by `pacuda-gen hip` (else the kernels of pacuda_amd/kernels); a CUDA header of the reference's
#   generator is refused (CUDA device code does not build for gfx950).
set -e
here=$(cd "$(dirname "$0")" && pwd)
out=${1:-$here/lib}
[ $# -gt 0 ] && shift
for t in "$@"; do
    case "$t" in
    gfx950) ;;
    gfx*) echo "pacuda_amd/build.sh: target $t is not supported (gfx950 only)" >&2; exit 1 ;;
    *) echo "pacuda_amd/build.sh: ignoring $t (not an AMD target; building for gfx950)" >&2 ;;
    esac
done
mkdir -p "$out"
out=$(cd "$out" && pwd)
kernels=""
if [ -n "$PACUDA_KERNELS" ]; then
    kernels=$(cd "$(dirname "$PACUDA_KERNELS")" && pwd)/$(basename "$PACUDA_KERNELS")
fi
(cd "$here/.." && PACUDA_KERNELS_HEADER="$kernels" PACUDA_OUT="$out/librecurrence.so" ${PYTHON:-python3} -c "
import os
from pacuda_amd import build
if not os.environ['PACUDA_KERNELS_HEADER']:
    build.build_kernels()
build.build_library(kernels_header=os.environ['PACUDA_KERNELS_HEADER'] or None, out=os.environ['PACUDA_OUT'])")
echo "$out/librecurrence.so (gfx950)"
