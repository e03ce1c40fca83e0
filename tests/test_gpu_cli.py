"""The command line e13search_gpu against the reference's own command lines
and logs (the C program's argv modes are covered by the parity tests through
the library; here the GPU driver's argv and summary line).  Each test prints
its case before touching the GPU.
"""
import os
import re
import subprocess

import pytest

from conftest import PKG_DIR, REF_RESULTS, machine_path, read_text

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
