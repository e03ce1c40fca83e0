#!/usr/bin/env bash
# Golden outputs of the reference's set_sym() + setup() (e13search.c:315-355,
# 422-426) over a sweep of problems on erdos_14: deleted state X = -1..13,
# edit budgets, symmetric ranges LO..HI and free-cell sets.  Each line is
#   <X> <FMAX> <LO> <HI> [CELL...] => <what the reference printed>
# where the right side is `setup budget=.. sym_lo=.. sym_hi=.. sym_n0=..
# tab=..` or `exit1 <its message>` when setup() refuses the problem.
# Needs `make -C oracle ref` (oracle/_ref/e13ref_dump, this container only).
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
BIN=$ROOT/oracle/_ref/e13ref_dump
M=$HERE/machines/erdos_14.txt
OUT=$HERE/ref_dump/setup_sweep.txt
: > "$OUT"
one() {
  local res
  if res=$("$BIN" setup "$M" "$@" 2>/tmp/e13_setup_err); then
    echo "$* => $res" >> "$OUT"
  else
    echo "$* => exit1 $(head -1 /tmp/e13_setup_err)" >> "$OUT"
  fi
}
XS="-1 0 1 2 3 4 5 6 7 8 9 10 11 12 13"
# delete mode, no symmetry
for X in $XS; do for F in 0 1 2 4; do one "$X" "$F" -1 -1; done; done
# delete mode with a symmetric range (the estimate mode's argv allows it)
for X in $XS; do
  for LH in "0 12" "1 3" "10 13" "11 13" "12 13" "13 13" "5 4" "1 0"; do one "$X" 2 $LH; done
done
# free mode, with and without symmetry
for X in -1 0 5 8 10 12 13; do
  for CELLS in "M0 M1" "D0 G0 K0 K1 L1 M1" "D0 G0 K0 K1 L0 L1 M0 M1" "K0 K1 L0 L1" "N0 N1" "A0" \
               "E0 E1 F0 F1 G0 G1 H0 H1 I0 I1"; do
    for LH in "-1 -1" "0 12" "10 13" "11 13" "12 13" "9 13" "13 13" "5 9" "4 9"; do one "$X" 0 $LH $CELLS; done
  done
done
rm -f /tmp/e13_setup_err
wc -l "$OUT"

# Free-mode searches with no deleted state (X = -1), where the shifted LO and
# the 2*(NQ-1)-cell scan decide the symmetry: the reference program's sorted
# SURVIVOR lines, its summary line (wall time dropped) and, per survivor, the
# counter hashes its run() leaves (e13ref_dump hash).
SRCH=$ROOT/oracle/_ref/e13search
OUT2=$HERE/ref_dump/freesym_xm1.txt
: > "$OUT2"
for args in "-1 0 12 M0 M1" "-1 13 13 D0 M0 M1" "-1 12 13 D0 L0 L1 M0 M1"; do
  echo "case freesym $args" >> "$OUT2"
  OMP_NUM_THREADS=8 "$SRCH" freesym "$M" $args > /tmp/e13_fs_out 2> /tmp/e13_fs_err
  sort /tmp/e13_fs_out > /tmp/e13_fs_sorted
  cat /tmp/e13_fs_sorted >> "$OUT2"
  grep 'leaves=' /tmp/e13_fs_err | sed -E 's/ wall=[0-9.]+s//' >> "$OUT2"
  "$BIN" hash < /tmp/e13_fs_sorted >> "$OUT2"
done
rm -f /tmp/e13_fs_out /tmp/e13_fs_err /tmp/e13_fs_sorted
wc -l "$OUT2"
