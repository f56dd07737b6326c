#!/usr/bin/env bash
# Build the InflationEasy reference program from its own sources, where they lie
# under /root/reference/src, for one compile-time configuration.
#
# The reference selects N and its feature toggles at compile time in
# src/parameters.h (src/parameters.h:22-44); its own regression harness changes N
# by regex-patching that header (tests/regression_main_n16.py:71-76).  This recipe
# does the same edit on a throw-away header in a temporary directory that holds
# symbolic links to the other reference source files, so every .cpp file and header
# is read from /root/reference/src unchanged (quote includes resolve next to the
# link, which picks up the edited configuration header) and only the binary is
# written, to oracle/_ref/.  The temporary directory is removed afterwards.  The flags follow Makefile:2,28-31 (-std=c++17 -O3 -DNDEBUG, plus
# -fopenmp when available).
#
# usage: oracle/build_ref.sh N NUMERICAL(0|1) SIGW(0|1) DELTAN(0|1) [OMP(0|1)] [POST(0|1)]
# output: oracle/_ref/ie_ref_n${N}_p${NUMERICAL}_g${SIGW}_d${DELTAN}_o${OMP}[_post]
set -euo pipefail

REF=${IE_REFERENCE_DIR:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"

N=${1:?N}
NUM=${2:?numerical_potential}
SIGW=${3:?calculate_SIGW}
DN=${4:?perform_deltaN}
OMP=${5:-1}
POST=${6:-0}

if [ ! -f "$REF/src/parameters.h" ]; then
    echo "reference sources not found under $REF" >&2
    exit 3
fi

mkdir -p "$OUT"
TAG="n${N}_p${NUM}_g${SIGW}_d${DN}_o${OMP}"
[ "$POST" = "1" ] && TAG="${TAG}_post"
BIN="$OUT/ie_ref_${TAG}"

TMP=$(mktemp -d /tmp/ie_ref_cfg.XXXXXX)
trap 'rm -rf "$TMP"' EXIT

# Compile-time edits only: N, the three feature toggles, the post-inflation
# toggle (it requires calculate_SIGW = 1, src/parameters.h:33-35) and the OpenMP
# toggle.
sed -E \
    -e "s/^const int N = [0-9]+;/const int N = ${N};/" \
    -e "s/^#define numerical_potential [01]/#define numerical_potential ${NUM}/" \
    -e "s/^#define perform_deltaN [01]/#define perform_deltaN ${DN}/" \
    -e "s/^#define calculate_SIGW [01]/#define calculate_SIGW ${SIGW}/" \
    -e "s/^#define post_inflation [01]/#define post_inflation ${POST}/" \
    -e "s/^#define parallel_calculation [01]/#define parallel_calculation ${OMP}/" \
    "$REF/src/parameters.h" > "$TMP/parameters.h"

grep -q "^const int N = ${N};" "$TMP/parameters.h" || { echo "failed to set N" >&2; exit 4; }

OMPFLAG=""
if [ "$OMP" = "1" ] && echo | ${CXX:-g++} -fopenmp -dM -E - >/dev/null 2>&1; then
    OMPFLAG="-fopenmp"
fi

for src in "$REF"/src/*; do
    base=$(basename "$src")
    [ "$base" = "parameters.h" ] && continue
    ln -s "$src" "$TMP/$base"
done

${CXX:-g++} -std=c++17 -O3 -DNDEBUG -w $OMPFLAG "$TMP"/*.cpp -o "$BIN" -lm
echo "$BIN"
