#!/bin/bash
# Builds the reference's own host program (main.c, allocDataHost.c,
# openFiles.c from /root/reference, unmodified, compiled where they lie)
# against this repository's kernel.h and libdifxfx.so.  kernel.c is not
# compiled: it needs Intel IPP, which is not available.  -I- stops the
# quoted includes from resolving next to main.c (the reference's IPP
# kernel.h) so they resolve through the -I list instead.
# Output: oracle/_ref/ref_main (git-ignored).
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
REPO="$(dirname "$HERE")"
REF=${REF:-/root/reference}
[ -f "$REF/main.c" ] || { echo "reference not present at $REF; nothing to build"; exit 0; }
mkdir -p "$HERE/_ref"
gcc -O2 -w -I- -I"$REPO/include" -I"$REF" \
    "$REF/main.c" "$REF/allocDataHost.c" "$REF/openFiles.c" \
    -o "$HERE/_ref/ref_main" \
    -L"$REPO/difx_tests_amd/lib" -ldifxfx -Wl,-rpath,'$ORIGIN/../../difx_tests_amd/lib' -Wl,-rpath,/opt/rocm/lib -lm 2>&1 \
    | grep -v "obsolete option '-I-'" || true
test -x "$HERE/_ref/ref_main"
echo "built $HERE/_ref/ref_main"
