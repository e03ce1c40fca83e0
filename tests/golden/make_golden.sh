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
#                 survivor table.
#
# Needs /root/reference (this container only) and `make -C oracle ref`.
# FMAX=4 takes over an hour on 8 cores (62 min here): pass --fmax4 to include it.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
BIN=$ROOT/oracle/_ref
mkdir -p "$HERE/ref_results" "$HERE/ref_dump" "$HERE/machines"

cp "$REF/machines/erdos_14.txt" "$REF/machines/split_15.txt" "$HERE/machines/"
for f in 1_local_edits/log_f2.txt 1_local_edits/log_f3.txt 1_local_edits/log_f4.txt \
         1_local_edits/surv_f2.txt 1_local_edits/surv_f3.txt 1_local_edits/surv_f4.txt \
         2_counter_3/log_mod.txt 2_counter_3/log_modsym.txt 2_counter_3/g_modsym.log \
         2_counter_3/surv_modsym.txt 2_counter_3/surv_mod.txt \
         3_doubling_2/dbl_X0.log 3_doubling_2/dbl_X1.log 3_doubling_2/dbl_X2.log \
         4_division_6/v4/log_v4.txt 4_division_6/v4/g_v4.log \
         4_division_6/v3/log_v3.txt 4_division_6/v3/g_v3.log \
         validation/base_erdos14.txt validation/mutants.txt \
         validation/split_test.txt validation/split_test_free.txt; do
  # fixtures keep a .txt suffix whatever the source suffix
  cp "$REF/results/$f" "$HERE/ref_results/$(echo "$f" | tr '/' '_' | sed 's/\.log$/.txt/')"
done

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
"$BIN/e13ref_dump" keys > "$HERE/ref_dump/keys_tgt.txt"
cat "$REF"/results/1_local_edits/surv_f{2,3,4}.txt "$REF"/results/2_counter_3/surv_modsym.txt \
    "$REF"/results/2_counter_3/surv_mod.txt > "$HERE/ref_dump/_surv_in.txt"
grep SURVIVOR "$REF/results/validation/split_test.txt" >> "$HERE/ref_dump/_surv_in.txt"
"$BIN/e13ref_dump" hash < "$HERE/ref_dump/_surv_in.txt" > "$HERE/ref_dump/survivor_hashes.txt"
rm -f "$HERE/ref_dump/_surv_in.txt"
echo "golden fixtures written under $HERE"
