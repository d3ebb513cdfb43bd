#!/usr/bin/env bash
# Builds oracle/_ref/liblq_ref.so: the reference tree's OpenSteer locality-query
# database (Samples/OpensteerDemo/src/lq.c, compiled from where it lies under
# the reference checkout) linked with oracle/lq_driver.c, for the CPU
# baseline of the neighbour search.  Needs the reference checkout (default
# /root/reference, or $BRT_REFERENCE); without it there is nothing to build.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
ref="${BRT_REFERENCE:-/root/reference}"
src="$ref/Samples/OpensteerDemo/src/lq.c"
inc="$ref/Samples/OpensteerDemo/include"
if [ ! -f "$src" ]; then
    echo "build_ref.sh: $src not found; oracle/_ref not built" >&2
    exit 0
fi
mkdir -p "$here/_ref"
gcc -O2 -fPIC -shared -fopenmp -I "$inc" -o "$here/_ref/liblq_ref.so" "$here/lq_driver.c" "$src"
