"""The command line e13search_gpu against the reference's own command lines
and logs: the C program's argv modes with its summary line
(e13search.c:413-416), and the GPU driver's argv with its summary line
(e13gpu.py:456-459).  Each test prints its case before touching the GPU.
"""
import os
import re
import subprocess

import pytest

from conftest import PKG_DIR, REF_DUMP, REF_RESULTS, machine_path, read_text

pytestmark = pytest.mark.gpu

CLI = os.path.join(PKG_DIR, "lib", "e13search_gpu")


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def test_cli_gpu_driver_argv_prints_the_drivers_summary():
    """`e13search_gpu gpu machine X LO HI CELL...` takes the GPU driver's argv
    (src/e13gpu.py) and prints its summary line: equal, up to the wall time, to
    the driver's own log of the counter search (results/2_counter_3/g_modsym.log)."""
    say("e13search_gpu gpu erdos_14 13 10 13 D0 G0 K0 K1 L1 M1")
    r = subprocess.run([CLI, "gpu", machine_path(), "13", "10", "13"] + "D0 G0 K0 K1 L1 M1".split(),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    strip = lambda l: re.sub(r" wall=[0-9.]+s", "", l)
    got = [strip(l) for l in r.stderr.splitlines() if " gpu leaves=" in l]
    want = [strip(l) for l in read_text(REF_RESULTS, "2_counter_3_g_modsym.txt").splitlines() if " gpu leaves=" in l]
    assert got == want
    assert len([l for l in r.stdout.splitlines() if l.startswith("SURVIVOR")]) == 61


# fields of the C program's summary that depend on the schedule (split, front,
# wall; e13search.c:413-416)
def c_fields(line):
    line = re.sub(r" split=\d+", "", re.sub(r" front=\d+", "", line))
    return re.sub(r" wall=[0-9.]+s", "", line).strip()


@pytest.mark.parametrize("args,fixture", [
    (["delete", "3"] + [str(x) for x in range(14)], ("ref_dump", "delete_f3.txt")),
    (["freesym", "13", "10", "13"] + "D0 G0 K0 K1 L1 M1".split(), ("ref_results", "2_counter_3_log_modsym.txt")),
    (["free", "14", "K0", "K1", "L1"], ("ref_results", "validation_split_test_free.txt")),
])
def test_cli_c_modes_print_the_c_programs_summary(args, fixture):
    """delete / free / freesym print the C program's summary line: equal to the
    reference's own output up to the schedule-dependent split, front and wall."""
    machine = machine_path("split_15.txt" if args[0] == "free" else "erdos_14.txt")
    say(f"e13search_gpu {args[0]} {os.path.basename(machine)} {' '.join(args[1:])}")
    r = subprocess.run([CLI, args[0], machine] + args[1:], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    got = [c_fields(l) for l in r.stderr.splitlines() if " leaves=" in l]
    assert all(" split=" in l for l in r.stderr.splitlines() if " leaves=" in l)
    assert not any(" gpu leaves=" in l for l in r.stderr.splitlines())
    base = REF_DUMP if fixture[0] == "ref_dump" else REF_RESULTS
    want = [c_fields(l) for l in read_text(base, fixture[1]).splitlines() if " leaves=" in l]
    assert got == want
