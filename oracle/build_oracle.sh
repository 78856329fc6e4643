#!/usr/bin/env bash
# Builds the CPU oracle (test infrastructure) into oracle/liboracle.so.
# -ffp-contract=off: no FMA contraction, so each float expression is evaluated
# as the sequence of IEEE operations written in the restated reference code.
set -euo pipefail
here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
gcc -O2 -std=c11 -fPIC -shared -ffp-contract=off -fno-fast-math -Wall -Wextra \
    -o "${here}/liboracle.so.tmp" "${here}/wrs_oracle.c" -lm
mv "${here}/liboracle.so.tmp" "${here}/liboracle.so"
