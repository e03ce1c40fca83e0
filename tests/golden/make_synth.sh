#!/bin/bash
# Golden summary lines of the synthetic stress batch (seed 0x5EED): the
# oracle's restatement of run_full() over tables 0 .. n-1 with each table's cap
# as every stage deadline.  Run from the repository root after
# `make -C oracle oracle`.
set -euo pipefail
out=tests/golden/synth/synth_5eed.txt
: > "$out"
for n in 65536 4194304; do
  ./oracle/_build/e13oracle synthsum 0x5EED 0 "$n" >> "$out"
done
cat "$out"
