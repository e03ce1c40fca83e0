#!/usr/bin/env bash
# Regenerates the golden fixtures under tests/golden/ from the reference tree.
#
#   ref_results/  data files copied from the reference's results/ and
#                 machines/ trees (logs, survivor lines, validation outputs):
#                 inputs and expected outputs only.
#   ref_dump/     outputs of the reference program compiled by oracle/Makefile
#                 (oracle/_ref/): per-X summaries with step totals for the
#                 local-edit searches, the Zobrist KEY/TGT tables, and the
#                 counter hashes ch[0..2] its run() leaves on every fixture
#                 survivor table, and the per-probe sums of its probe()
#                 estimator (probe_*.txt) on six estimate problems.
#
# Needs /root/reference (this container only) and `make -C oracle ref`.
# FMAX=4 takes over an hour on 8 cores (62 min here): pass --fmax4 to include it.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
BIN=$ROOT/oracle/_ref
mkdir -p "$HERE/ref_results" "$HERE/ref_dump" "$HERE/machines"

cp "$REF/machines/erdos_14.txt" "$REF/machines/split_15.txt" "$REF/machines/c0_0LD.txt" "$HERE/machines/"
for f in 1_local_edits/log_f2.txt 1_local_edits/log_f3.txt 1_local_edits/log_f4.txt \
         1_local_edits/surv_f2.txt 1_local_edits/surv_f3.txt 1_local_edits/surv_f4.txt \
         2_counter_3/log_mod.txt 2_counter_3/log_modsym.txt 2_counter_3/g_modsym.log \
         2_counter_3/surv_modsym.txt 2_counter_3/surv_mod.txt \
         3_doubling_2/dbl_X0.log 3_doubling_2/dbl_X1.log 3_doubling_2/dbl_X2.log \
         4_division_6/v4/log_v4.txt 4_division_6/v4/g_v4.log \
         4_division_6/v3/log_v3.txt 4_division_6/v3/g_v3.log \
         validation/base_erdos14.txt validation/mutants.txt \
         validation/split_test.txt validation/split_test_free.txt \
         4_division_6/v1/v1c0_est.txt validation/deep_base14.txt validation/check_erdos14.txt \
         4_division_6/v2/surv_v2.txt 4_division_6/v2/deep_v2.txt 4_division_6/v2/check_v2.txt \
         4_division_6/v1/classify_1RD.txt 4_division_6/v1/offtape3.txt 4_division_6/v1/runaways_deep_big.txt; do
  # fixtures keep a .txt suffix whatever the source suffix
  cp "$REF/results/$f" "$HERE/ref_results/$(echo "$f" | tr '/' '_' | sed 's/\.log$/.txt/')"
done

# deep stage check (src/deep.c): the first survivors of V1's 1RD search and
# three that leave the tape at stage 12, with the reference's results
sed -n '1,4p;104,106p' "$REF/results/4_division_6/v1/surv_1RD.txt" > "$HERE/ref_results/4_division_6_v1_surv_1RD_sample.txt"
sed -n '1,4p;104,106p' "$REF/results/4_division_6/v1/deep15_1RD.txt" > "$HERE/ref_results/4_division_6_v1_deep15_1RD_sample.txt"
# all of V1's 1RD survivors and their deep check, compressed (the classification's input)
gzip -c -n "$REF/results/4_division_6/v1/surv_1RD.txt" > "$HERE/ref_results/4_division_6_v1_surv_1RD.txt.gz"
gzip -c -n "$REF/results/4_division_6/v1/deep15_1RD.txt" > "$HERE/ref_results/4_division_6_v1_deep15_1RD.txt.gz"
# the counter re-check (src/recheck.c 15, validation/recheck_1RD.txt, run in chunks of ~359 lines):
# survivor lines 1, 2, 53, 58 (chunk 0) and 2840-2842 (chunk 7, ids 330-332)
sed -n '1p;2p;53p;58p;2840p;2841p;2842p' "$REF/results/4_division_6/v1/surv_1RD.txt" \
    > "$HERE/ref_results/4_division_6_v1_surv_1RD_recheck_sample.txt"
sed -n '4p;5p;56p;61p;2850p;2851p;2852p' "$REF/results/validation/recheck_1RD.txt" \
    > "$HERE/ref_results/validation_recheck_1RD_sample.txt"

M=$REF/machines/erdos_14.txt
XS="0 1 2 3 4 5 6 7 8 9 10 11 12 13"
export OMP_NUM_THREADS=${OMP_NUM_THREADS:-8}
"$BIN/e13search" delete "$M" 2 $XS > /dev/null 2> "$HERE/ref_dump/delete_f2.err"
"$BIN/e13search" delete "$M" 3 $XS > /dev/null 2> "$HERE/ref_dump/delete_f3.err"
if [[ "${1:-}" == "--fmax4" ]]; then
  "$BIN/e13search" delete "$M" 4 $XS > /dev/null 2> "$HERE/ref_dump/delete_f4.err"
fi
# keep only the summary lines (progress lines and wall times are schedule noise)
for f in "$HERE"/ref_dump/delete_f*.err; do
  grep 'leaves=' "$f" | sed -E 's/ wall=[0-9.]+s//' > "$f.tmp" && mv "$f.tmp" "${f%.err}.txt" && rm -f "$f"
done
# a bounded sample of the 3-state counter search's family (the bench's CPU leg
# beside it): states 11..13 symmetric, free cells D0 G0 L1 M1; summary (wall
# time removed) and the sorted survivor lines
"$BIN/e13search" freesym "$M" 13 11 13 D0 G0 L1 M1 > "$HERE/ref_dump/_cs.out" 2> "$HERE/ref_dump/_cs.err"
{ grep 'leaves=' "$HERE/ref_dump/_cs.err" | sed -E 's/ wall=[0-9.]+s//'; grep '^SURVIVOR' "$HERE/ref_dump/_cs.out" | sort; } \
    > "$HERE/ref_dump/freesym_x13_s11_D0G0L1M1.txt"
rm -f "$HERE/ref_dump/_cs.out" "$HERE/ref_dump/_cs.err"
"$BIN/e13ref_dump" keys > "$HERE/ref_dump/keys_tgt.txt"
cat "$REF"/results/1_local_edits/surv_f{2,3,4}.txt "$REF"/results/2_counter_3/surv_modsym.txt \
    "$REF"/results/2_counter_3/surv_mod.txt > "$HERE/ref_dump/_surv_in.txt"
grep SURVIVOR "$REF/results/validation/split_test.txt" >> "$HERE/ref_dump/_surv_in.txt"
"$BIN/e13ref_dump" hash < "$HERE/ref_dump/_surv_in.txt" > "$HERE/ref_dump/survivor_hashes.txt"
rm -f "$HERE/ref_dump/_surv_in.txt"
# per-probe sums of the reference's probe() (estimate mode), one stream per
# probe; the first line records the arguments: machine X FMAX LO HI SEED N CELL...
probe_dump() {
  local name=$1; shift
  { echo "args $*"; "$BIN/e13ref_dump" probe "$HERE/machines/$1" "${@:2}"; } > "$HERE/ref_dump/probe_$name.txt"
}
probe_dump f2_x8 erdos_14.txt 8 2 -1 -1 0 600
probe_dump f3_x0 erdos_14.txt 0 3 -1 -1 1 400
probe_dump f1_nodel erdos_14.txt -1 1 -1 -1 2 400
probe_dump free_m1n0 erdos_14.txt -1 0 -1 -1 3 600 M1 N0
probe_dump counter3 erdos_14.txt 13 0 10 13 4 600 D0 G0 K0 K1 L1 M1
probe_dump v1_0ld c0_0LD.txt 9 0 3 9 5 600 D0 D1 E0 E1 F0 F1 G0 G1 H0 H1 I0 I1
echo "golden fixtures written under $HERE"
