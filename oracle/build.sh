#!/bin/sh
# Builds the CPU oracle (test infrastructure) into oracle/build/.
set -e
cd "$(dirname "$0")"
mkdir -p build
g++ -O3 -mavx2 -ffp-contract=off -fPIC -shared -std=c++17 -pthread \
    hh_oracle.cpp -o build/libhh_oracle.so
