#!/usr/bin/env bash
# Builds the reference's own CPU MBE code (baselines/MBE.zip: MbeaAdvFinder,
# ImbeaAdvFinder, PmbeAdvFinder; pure C++11) into oracle/_ref/mbe_ref.
# TEST INFRASTRUCTURE ONLY: the binary is used to pin the oracle's counts.
# The zip's sources are unpacked into a temporary directory outside the
# repository and compiled there; only the binary lands in oracle/_ref/.
set -euo pipefail
REF=${GMBE_REFERENCE:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
ZIP="$REF/baselines/MBE.zip"
if [ ! -f "$ZIP" ]; then
  echo "build_ref: $ZIP not found; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
TMP="$(mktemp -d /tmp/gmbe_ref.XXXXXX)"
trap 'rm -rf "$TMP"' EXIT
unzip -q -o "$ZIP" 'MBE/*.cpp' 'MBE/*.h' -d "$TMP"
g++ -std=c++11 -O3 -w -I"$TMP/MBE" -o "$OUT/mbe_ref" \
  "$TMP/MBE/main.cpp" "$TMP/MBE/BicliqueFinder.cpp" "$TMP/MBE/BiGraph.cpp" \
  "$TMP/MBE/Utility.cpp" "$TMP/MBE/PMBE.cpp"
echo "build_ref: built $OUT/mbe_ref"
