#!/bin/sh
# SPDX-License-Identifier: 0BSD
# Builds librecurrence.so for gfx950 with runner/build.sh's arguments, so that callers of that
# script (pacuda.py's retarget step) can call this one instead.
# usage: [PACUDA_KERNELS=kernels.cuh] pacuda_amd/build.sh [OUT_DIR] [TARGET...]
#   TARGET: gfx950 (the default and the only target); CUDA SM numbers (90, 120, ...) name no
#   AMD target and are ignored. PACUDA_KERNELS: generated symjit kernels, which this runner
#   cannot take; it then reports none (capability bit 2 clear) and refuses plans with symjit rows.
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
if [ -n "$PACUDA_KERNELS" ]; then
    echo "pacuda_amd/build.sh: $PACUDA_KERNELS not built in (no symjit kernels in this runner)" >&2
fi
mkdir -p "$out"
lib=$(cd "$here/.." && ${PYTHON:-python3} -c "from pacuda_amd import build; print(build.build_library())")
[ "$(cd "$(dirname "$lib")" && pwd)" = "$(cd "$out" && pwd)" ] || cp "$lib" "$out/librecurrence.so"
echo "$out/librecurrence.so (gfx950)"
