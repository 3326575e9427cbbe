#!/bin/sh
# Builds the oracle's scalar C psi loop (test infrastructure; CPU only).
set -e
cd "$(dirname "$0")"
gcc -O2 -fPIC -shared -o libpsi_oracle.so csrc/psi_loop.c -lm
