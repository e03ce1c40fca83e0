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


class EoProblem(ctypes.Structure):
    _fields_ = [("nq_base", ctypes.c_int), ("base", ctypes.c_int * 32), ("X", ctypes.c_int), ("fmax", ctypes.c_int),
                ("freemode", ctypes.c_int), ("freebase", ctypes.c_uint8 * 32), ("sym_lo", ctypes.c_int),
                ("sym_hi", ctypes.c_int), ("nq", ctypes.c_int), ("budget", ctypes.c_int), ("sym_n0", ctypes.c_int),
                ("tab", ctypes.c_int * 32)]


def oracle_setup(base, X, fmax, free=None, sym=None):
    L = oracle_lib()
    p = EoProblem()
    p.nq_base = len(base) // 2
    for i, v in enumerate(base):
        p.base[i] = v
    p.X, p.fmax = X, fmax
    p.sym_lo = p.sym_hi = -1
    if free is not None:
        p.freemode = 1
        for c in free:
            p.freebase[2 * (ord(c[0]) - 65) + int(c[1])] = 1
    if sym:
        L.eo_set_sym(ctypes.byref(p), sym[0], sym[1])
    rc = L.eo_setup(ctypes.byref(p))
    return rc, p


def probe_fixture(name):
    """A per-probe dump of the reference's probe(): (args, rows) with args the
    estimate problem (machine, X, FMAX, LO, HI, SEED, N, cells) and rows the
    four sums of each probe as the bit patterns of the doubles."""
    lines = read_text(REF_DUMP, f"probe_{name}.txt").splitlines()
    a = lines[0].split()[1:]
    args = dict(machine=a[0], X=int(a[1]), fmax=int(a[2]), lo=int(a[3]), hi=int(a[4]), seed=int(a[5]), n=int(a[6]),
                cells=a[7:])
    rows = [tuple(int(v, 16) for v in l.split()[1:5]) for l in lines[2:]]
    assert len(rows) == args["n"]
    return args, rows


PROBE_CASES = ["f2_x8", "f3_x0", "f1_nodel", "free_m1n0", "counter3", "v1_0ld"]


def oracle_probe_problem(args):
    """eo_problem of an estimate problem (free mode when cells are given)."""
    base = []
    for cell in read_text(MACHINES, args["machine"]).split():
        base.append(-1 if cell.startswith("-") else (int(cell[2:]) << 2) | ((cell[1] == "R") << 1) | int(cell[0]))
    sym = (args["lo"], args["hi"]) if args["lo"] >= 0 else None
    return oracle_setup(base, args["X"], args["fmax"], args["cells"] or None, sym)


def oracle_probes(L, p, seed, n):
    """Per-probe sums of eo_probe on streams eo_probe_stream(seed, 0..n-1), as bit patterns."""
    L.eo_probe_stream.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.eo_probe_stream.restype = ctypes.c_uint64
    out = []
    for i in range(n):
        rng = ctypes.c_uint64(L.eo_probe_stream(seed, i))
        o = (ctypes.c_double * 4)()
        L.eo_probe(ctypes.byref(p), ctypes.byref(rng), o)
        out.append(tuple(ctypes.c_uint64.from_buffer_copy(ctypes.c_double(v)).value for v in o))
    return out


def v1c0_rows():
    """Every value row of 4_division_6/v1/v1c0_est.txt: (value, machine path,
    leaves, leaves standard error, steps, steps standard error), the machine
    being c0_0LD.txt with C0 set to the value (the file's own definition)."""
    import tempfile
    base = read_text(MACHINES, "c0_0LD.txt").split("\n")
    d = os.path.join(tempfile.gettempdir(), "e13_c0_machines")
    os.makedirs(d, exist_ok=True)
    rows = []
    for line in read_text(REF_RESULTS, "4_division_6_v1_v1c0_est.txt").splitlines():
        f = line.split()
        if not line.startswith("c0_") or len(f) < 7:
            continue
        v = f[0][3:]
        rows_m = list(base)
        rows_m[2] = f"{v[0]}{v[1]}{ord(v[2]) - 65} " + rows_m[2].split()[1]
        path = os.path.join(d, f"c0_{v}.txt")
        with open(path, "w") as fh:
            fh.write("\n".join(rows_m))
        rows.append((v, path, float(f[2]), float(f[3]), float(f[5]), float(f[6])))
    return rows


def v1c0_exact_rows():
    """Rows of 4_division_6/v1/v1c0_est.txt whose standard errors are 0 (the
    value of C0 leaves a one-leaf tree, so the estimate is exact): (value,
    machine path, leaves, steps), the machine being c0_0LD.txt with C0 set to
    the value (the file's own definition of its rows)."""
    import tempfile
    base = read_text(MACHINES, "c0_0LD.txt").split("\n")
    d = os.path.join(tempfile.gettempdir(), "e13_c0_machines")
    os.makedirs(d, exist_ok=True)
    rows = []
    for line in read_text(REF_RESULTS, "4_division_6_v1_v1c0_est.txt").splitlines():
        f = line.split()
        if not line.startswith("c0_") or len(f) < 7 or f[3] != "0.0e+00" or f[6] != "0.0e+00":
            continue
        v = f[0][3:]
        cell = f"{v[0]}{v[1]}{ord(v[2]) - 65}"
        rows_m = list(base)
        rows_m[2] = cell + " " + rows_m[2].split()[1]
        path = os.path.join(d, f"c0_{v}.txt")
        with open(path, "w") as fh:
            fh.write("\n".join(rows_m))
        rows.append((v, path, float(f[2]), float(f[5])))
    return rows
