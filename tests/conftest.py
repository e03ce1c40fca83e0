"""Shared helpers for the test suite.

`gpu` tests run the HIP library on a real MI355X; everything else runs on the
CPU: the oracle against the golden fixtures, the host-side C code of the
library, and the library's exported symbols.
"""
import ctypes
import importlib.util
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
MACHINES = os.path.join(GOLDEN, "machines")
REF_RESULTS = os.path.join(GOLDEN, "ref_results")
REF_DUMP = os.path.join(GOLDEN, "ref_dump")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_BIN = os.path.join(ORACLE_DIR, "_build", "e13oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "libe13oracle.so")
PKG_DIR = os.path.join(ROOT, "erdos-machine-13_amd")
XS = list(range(14))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs the HIP library)")


def load_pkg():
    """The ctypes binding in erdos-machine-13_amd/ (a directory name that is not
    an identifier, so it is loaded by path)."""
    if "e13amd" in sys.modules:
        return sys.modules["e13amd"]
    spec = importlib.util.spec_from_file_location("e13amd", os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["e13amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def ensure_oracle():
    if not (os.path.exists(ORACLE_BIN) and os.path.exists(ORACLE_LIB)):
        subprocess.run(["make", "-C", ORACLE_DIR, "oracle"], check=True, capture_output=True)


def oracle_lib():
    ensure_oracle()
    L = ctypes.CDLL(ORACLE_LIB)
    L.eo_init_keys()
    return L


def run_oracle(*args, threads=None):
    """Run the oracle command line; returns (stdout, stderr)."""
    ensure_oracle()
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    out = subprocess.run([ORACLE_BIN] + [str(a) for a in args], capture_output=True, text=True, env=env,
                         check=True)
    return out.stdout, out.stderr


def machine_path(name="erdos_14.txt"):
    return os.path.join(MACHINES, name)


def read_machine(name="erdos_14.txt"):
    with open(machine_path(name)) as f:
        return f.read()


SUMMARY_RE = re.compile(r"X=(-?\d+) .*?leaves=(\d+) survivors=(\d+)(?: steps=(\d+))?.*dead-at-stage\(0\.\.9\):((?: \d+)+)")


def parse_summaries(text):
    """{X: {leaves, survivors, steps (or None), hist}} from summary lines."""
    out = {}
    for line in text.splitlines():
        m = SUMMARY_RE.search(line)
        if not m:
            continue
        out[int(m.group(1))] = {
            "leaves": int(m.group(2)), "survivors": int(m.group(3)),
            "steps": int(m.group(4)) if m.group(4) else None,
            "hist": [int(x) for x in m.group(5).split()],
        }
    return out


def read_text(*parts):
    with open(os.path.join(*parts)) as f:
        return f.read()


def survivor_lines(text):
    return sorted(l.strip() for l in text.splitlines() if l.startswith("SURVIVOR"))


def hash_lines_by_survivor(text):
    """Pairs each SURVIVOR line with the HASH line that follows it."""
    lines = [l.strip() for l in text.splitlines() if l.startswith(("SURVIVOR", "HASH"))]
    pairs = []
    for i, l in enumerate(lines):
        if l.startswith("SURVIVOR"):
            h = lines[i + 1] if i + 1 < len(lines) and lines[i + 1].startswith("HASH") else None
            pairs.append((l, h))
    return sorted(pairs)


def ref_hashes():
    """table string -> (ch0, ch1, ch2) from the compiled-reference dump."""
    surv = []
    for f in ("1_local_edits_surv_f2.txt", "1_local_edits_surv_f3.txt", "1_local_edits_surv_f4.txt",
              "2_counter_3_surv_modsym.txt", "2_counter_3_surv_mod.txt"):
        surv += [l.strip() for l in read_text(REF_RESULTS, f).splitlines() if l.startswith("SURVIVOR")]
    surv += [l.strip() for l in read_text(REF_RESULTS, "validation_split_test.txt").splitlines()
             if l.startswith("SURVIVOR")]
    hs = [l.split() for l in read_text(REF_DUMP, "survivor_hashes.txt").splitlines() if l.startswith("HASH")]
    assert len(hs) == len(surv)
    out = {}
    for s, h in zip(surv, hs):
        tab = s.split("table=")[1].replace("?", "")
        out[(int(s.split("X=")[1].split()[0]), tab)] = tuple(int(x.split("=")[1], 16) for x in h[2:5])
    return out


@pytest.fixture(scope="session")
def e13():
    return load_pkg()


@pytest.fixture(scope="session")
def gpu_ctx(e13):
    """One device context for all GPU tests (one process on the card)."""
    ctx = e13.Context(0)
    yield ctx
    ctx.close()


class OracleSim(ctypes.Structure):
    """eo_sim of oracle/e13oracle.h."""
    _fields_ = [("head", ctypes.c_int), ("st", ctypes.c_int), ("nt", ctypes.c_int), ("stage", ctypes.c_int),
                ("st0", ctypes.c_int), ("head0", ctypes.c_int), ("abs_ok", ctypes.c_int),
                ("rel_ok", ctypes.c_int), ("cs", ctypes.c_int * 3), ("t", ctypes.c_longlong),
                ("deadline", ctypes.c_longlong), ("hb", ctypes.c_uint64), ("hc", ctypes.c_uint64),
                ("ch", ctypes.c_uint64 * 3), ("tape", ctypes.c_uint8 * 1364)]


def oracle_full(L, tab, dl):
    """The oracle's run_full(): (result, stage, t, st0, head0, abs_ok, rel_ok, ch)."""
    s = OracleSim()
    t32 = (ctypes.c_int * 32)(*(list(tab) + [-1] * (32 - len(tab))))
    r = L.eo_run_full(t32, ctypes.byref(s), (ctypes.c_int32 * 10)(*dl) if dl else None)
    return (r, s.stage, s.t, s.st0, s.head0, s.abs_ok, s.rel_ok, tuple(s.ch))


def synth_golden():
    """{n: summary line} of the synthetic stress batch (seed 0x5EED) from the oracle."""
    with open(os.path.join(GOLDEN, "synth", "synth_5eed.txt")) as f:
        return {int(re.search(r" n=(\d+)", l).group(1)): l.strip() for l in f if l.startswith("synth ")}

