#!/bin/bash
# The reference's size sweeps (test/1d_script.sh, 2d_script.sh,
# 3d_script.sh) through the drop-in harnesses; run from the repository root
# on a machine with an MI355X after building the library.  One line per
# size, as the harnesses print it; stops at the first failing size.
set -e
cd "$(dirname "$0")/../.."
B=$(mktemp -d)
for h in harness1d harness2d harness3d harness_real; do
    gcc -std=gnu99 -O2 -Wall -I include tope-fft_amd/examples/$h.c -L tope-fft_amd/lib -ltopefft -lm \
        -Wl,-rpath,$PWD/tope-fft_amd/lib -o "$B/$h"
done
echo "1D: lengths 3 .. 27 (test/1d_script.sh), then powers of two 2^3 .. 2^24 and a few longer non-powers"
for n in $(seq 3 27); do "$B/harness1d" $n; done
for j in $(seq 3 24); do "$B/harness1d" $((2 ** j)); done
for n in 100 1000 4097 65537; do "$B/harness1d" $n; done
echo "2D horizontal, vertical (4 .. 2^22, test/2d_script.sh), planar (4 .. 4096)"
for j in $(seq 2 22); do "$B/harness2d" $((2 ** j)) 4; done
for j in $(seq 2 22); do "$B/harness2d" 4 $((2 ** j)); done
for j in $(seq 2 12); do "$B/harness2d" $((2 ** j)) $((2 ** j)); done
echo "3D horizontal, vertical (4 .. 2^20), planar (4 .. 2048), cubic (4 .. 64) (test/3d_script.sh)"
for j in $(seq 2 20); do "$B/harness3d" $((2 ** j)) 4 4; done
for j in $(seq 2 20); do "$B/harness3d" 4 $((2 ** j)) 4; done
for j in $(seq 2 11); do "$B/harness3d" $((2 ** j)) $((2 ** j)) 4; done
for j in $(seq 2 6); do "$B/harness3d" $((2 ** j)) $((2 ** j)) $((2 ** j)); done
echo "Real input (R2C then C2R)"
for d in "16 1 1" "4096 1 1" "256 256 1" "4096 4096 1" "64 64 64" "512 512 512"; do "$B/harness_real" $d; done
rm -rf "$B"
