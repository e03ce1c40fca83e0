"""ctypes binding of libe13.so, the gfx950 TNF-enumeration + simulate-and-filter
library (C ABI in include/e13.h).

The reference keeps this path inside src/e13search.c (run, explore,
run_delete) and src/e13gpu.py (expand kernel and its host loop).  This module
mirrors that interface for Python callers: `setup()` is setup()/set_sym(),
`Context.search()` is run_delete(), `Context.simulate_batch()` is run_full()
over many tables, and the survivor/summary formatting follows report() and
the summary line.  Every call goes through the HIP library; there is no CPU
fallback, and a missing library raises.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libe13.so")
CLI_PATH = os.path.join(LIB_DIR, "e13search_gpu")
CSRC = os.path.join(HERE, "csrc")

NS = 9
MAXCELLS = 32
DEAD, SURV = -1, -2
HALT, FORCED = -1, -2
T14S = (0, 2, 22, 90, 268, 944, 3434, 13450, 52068, 208864)  # e13search.c:70

ERRORS = {
    0: "ok", -1: "invalid argument", -2: "HIP runtime error", -3: "out of memory",
    -4: "forced cells exceed fmax", -5: "symmetric state is named by a fixed cell or has one",
    -6: "start state in the symmetric states", -7: "node stack overflow", -8: "no HIP device",
    -9: "stopped cleanly; the checkpoint is saved", -10: "checkpoint file unreadable, damaged or for another search",
    -11: "a mailbox delivery never arrived complete", -12: "a wave waited longer than the idle watchdog",
    -13: "dynamic LDS does not start at LDS address 0",
    -14: "more probes than the static scratch holds",
}
STATIC_PROBES = 1 << 16  # e13_estimate_static's limit (E13_STATIC_PROBES)


class E13Error(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{what}: {ERRORS.get(code, 'error')} ({code})" if what else ERRORS.get(code, str(code)))


class Problem(ctypes.Structure):
    _fields_ = [("nq", ctypes.c_int32), ("X", ctypes.c_int32), ("budget", ctypes.c_int32),
                ("sym_lo", ctypes.c_int32), ("sym_hi", ctypes.c_int32), ("sym_n0", ctypes.c_int32),
                ("nforced", ctypes.c_int32), ("tab", ctypes.c_int8 * MAXCELLS)]


class Counts(ctypes.Structure):
    _fields_ = [("leaves", ctypes.c_uint64), ("steps", ctypes.c_uint64), ("survivors", ctypes.c_uint64),
                ("dead_at_stage", ctypes.c_uint64 * (NS + 1))]

    def as_dict(self):
        return {"leaves": self.leaves, "steps": self.steps, "survivors": self.survivors,
                "hist": list(self.dead_at_stage)}


class Survivor(ctypes.Structure):
    _fields_ = [("tab", ctypes.c_int8 * MAXCELLS), ("read_mask", ctypes.c_uint32),
                ("st0", ctypes.c_int32), ("head0", ctypes.c_int32), ("abs_ok", ctypes.c_int32),
                ("rel_ok", ctypes.c_int32), ("steps", ctypes.c_int64), ("ch", ctypes.c_uint64 * 3),
                ("recheck_ok", ctypes.c_int32), ("prob", ctypes.c_int32)]


class Leaf(ctypes.Structure):
    _fields_ = [("result", ctypes.c_int32), ("stage", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("st0", ctypes.c_int32), ("head0", ctypes.c_int32), ("abs_ok", ctypes.c_int32),
                ("rel_ok", ctypes.c_int32), ("event", ctypes.c_int32), ("ch", ctypes.c_uint64 * 3)]


class BatchSum(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("digest", ctypes.c_uint64), ("steps", ctypes.c_uint64),
                ("survivors", ctypes.c_uint64), ("dead_at_stage", ctypes.c_uint64 * 10),
                ("kernel_ms", ctypes.c_double)]

    def line(self, seed0, k0):
        """The summary line of the synthetic batch (the e13search_gpu synth line)."""
        return (f"synth seed0=0x{seed0:x} k0={k0} n={self.n} digest={self.digest:016x} steps={self.steps} "
                f"survivors={self.survivors} dead-at-stage(0..9): " + " ".join(str(x) for x in self.dead_at_stage))


NFEAT = 32
MLP_PARAMS = 6402


class CostRec(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float * NFEAT), ("leaves", ctypes.c_uint64), ("steps", ctypes.c_uint64),
                ("prob", ctypes.c_int32), ("depth", ctypes.c_int32), ("est_steps", ctypes.c_double)]


class TuneReport(ctypes.Structure):
    _fields_ = [("trials", ctypes.c_int32), ("pad", ctypes.c_int32), ("first_ms", ctypes.c_double),
                ("best_ms", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [("form", ctypes.c_int32), ("slice", ctypes.c_int32), ("wg_threads", ctypes.c_int32),
                ("wgs_per_cu", ctypes.c_int32), ("cap", ctypes.c_int32), ("batch", ctypes.c_int32),
                ("gate", ctypes.c_int32), ("refill_tries", ctypes.c_int32), ("stack_nodes", ctypes.c_int64),
                ("seed_per_wave", ctypes.c_int32), ("sched_every", ctypes.c_int32), ("donate_max", ctypes.c_int32),
                ("idle_naps", ctypes.c_int32), ("mail_every", ctypes.c_int32), ("trace_events", ctypes.c_int32),
                ("graph_chain", ctypes.c_int32), ("cost_log", ctypes.c_int32),
                ("learned", ctypes.c_int32), ("solo_steps", ctypes.c_int32), ("evict_after", ctypes.c_int32),
                ("probe_slice", ctypes.c_int32), ("cut_live", ctypes.c_int32),
                ("phase_cycles", ctypes.c_int32), ("nt1_skip", ctypes.c_int32), ("tape_words", ctypes.c_int32),
                ("wide_nodes", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("launches", ctypes.c_int64), ("rounds", ctypes.c_int64),
                ("lane_slots", ctypes.c_int64), ("peak_stack", ctypes.c_int64), ("slice", ctypes.c_int32),
                ("wg_threads", ctypes.c_int32), ("grid", ctypes.c_int32), ("gate", ctypes.c_int32),
                ("form", ctypes.c_int32), ("profile", ctypes.c_int32), ("donations", ctypes.c_int64),
                ("pool_takes", ctypes.c_int64), ("cycles", ctypes.c_int64 * 4), ("graphs", ctypes.c_int64),
                ("search_ms", ctypes.c_double), ("search_launches", ctypes.c_int64), ("search_steps", ctypes.c_int64),
                ("frontier", ctypes.c_int64), ("tape_words", ctypes.c_int32), ("pad_s", ctypes.c_int32),
                ("wide_nodes", ctypes.c_int64), ("wide_launches", ctypes.c_int64), ("wide_ms", ctypes.c_double),
                ("wide_steps", ctypes.c_int64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "cycles"}
        d["cycles"] = list(self.cycles)
        return d



class EstimateSum(ctypes.Structure):
    """e13_estimate_sum: per-quantity sums, squares, means and standard errors
    (leaves, nodes, steps, survivors), plus the device's own figures."""
    _fields_ = [("probes", ctypes.c_int64), ("sum", ctypes.c_double * 4), ("sq", ctypes.c_double * 4),
                ("mean", ctypes.c_double * 4), ("se", ctypes.c_double * 4), ("sim_steps", ctypes.c_uint64),
                ("rerun_steps", ctypes.c_uint64),
                ("kernel_ms", ctypes.c_double), ("launches", ctypes.c_int64), ("rounds", ctypes.c_int64),
                ("lane_slots", ctypes.c_int64)]

    NAMES = ("leaves", "nodes", "steps", "survivors")

    def as_dict(self):
        d = {n: (self.mean[j], self.se[j]) for j, n in enumerate(self.NAMES)}
        d.update(probes=self.probes, sim_steps=self.sim_steps, rerun_steps=self.rerun_steps,
                 kernel_ms=self.kernel_ms, launches=self.launches,
                 rounds=self.rounds, lane_slots=self.lane_slots)
        return d


EXPORTS = ("e13_open", "e13_close", "e13_strerror", "e13_set_tuning", "e13_get_tuning", "e13_get_stats",
           "e13_keys", "e13_parse_machine", "e13_setup", "e13_search", "e13_search_many", "e13_simulate_batch",
           "e13_simulate_batch_device", "e13_sync", "e13_recheck", "e13_format_survivor", "e13_format_hash",
           "e13_synth_table", "e13_get_trace", "e13_batch_digest", "e13_synth_run", "e13_get_cost_log",
           "e13_cost_targets", "e13_mlp_init", "e13_mlp_forward", "e13_mlp_grad", "e13_mlp_train",
           "e13_set_cost_model", "e13_autotune", "e13_save_tuning", "e13_load_tuning", "e13_tuning_file",
           "e13_tuning_file_small", "e13_search_profile", "e13_estimate", "e13_estimate_static", "e13_set_checkpoint",
           "e13_deep_check_line", "e13_deep_check", "e13_recheck_counter_line", "e13_recheck_counter",
           "e13_check_survivor", "e13_classify_deep", "e13_deep_check_gpu", "e13_recheck_counter_gpu")

_lib = None


def build(quiet=True):
    """Compile libe13.so and the e13search_gpu command line for gfx950 (hipcc)."""
    out = subprocess.run(["make", "-C", CSRC, "-j8"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("building libe13.so failed:\n" + out.stdout[-4000:] + out.stderr[-4000:])
    if not quiet:
        print(out.stdout)


def lib():
    """The loaded libe13.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (make -C {CSRC})")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    L.e13_open.argtypes = [ctypes.c_int, P(Tuning), P(ctypes.c_void_p)]
    L.e13_close.argtypes = [ctypes.c_void_p]
    L.e13_close.restype = None
    L.e13_strerror.restype = ctypes.c_char_p
    L.e13_set_tuning.argtypes = [ctypes.c_void_p, P(Tuning)]
    L.e13_get_tuning.argtypes = [ctypes.c_void_p, P(Tuning)]
    L.e13_get_stats.argtypes = [ctypes.c_void_p, P(Stats)]
    L.e13_keys.argtypes = [P(ctypes.c_uint64), P(ctypes.c_uint64)]
    L.e13_parse_machine.argtypes = [ctypes.c_char_p, P(ctypes.c_int32), ctypes.c_int]
    L.e13_setup.argtypes = [P(ctypes.c_int32), ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_uint8),
                            ctypes.c_int, ctypes.c_int, P(Problem)]
    L.e13_search.argtypes = [ctypes.c_void_p, P(Problem), P(Counts), P(Survivor), ctypes.c_uint64,
                             P(ctypes.c_uint64)]
    L.e13_search_many.argtypes = [ctypes.c_void_p, P(Problem), ctypes.c_int, P(Counts), P(Survivor),
                                  ctypes.c_uint64, P(ctypes.c_uint64)]
    L.e13_simulate_batch.argtypes = [ctypes.c_void_p, P(ctypes.c_int8), ctypes.c_int, ctypes.c_int,
                                     P(ctypes.c_int32), P(Leaf)]
    L.e13_simulate_batch_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_void_p]
    L.e13_sync.argtypes = [ctypes.c_void_p]
    L.e13_recheck.argtypes = [P(ctypes.c_int8), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int64, P(ctypes.c_uint32)]
    L.e13_format_survivor.argtypes = [P(Problem), P(Survivor), ctypes.c_char_p, ctypes.c_int]
    L.e13_format_hash.argtypes = [P(Problem), P(Survivor), ctypes.c_char_p, ctypes.c_int]
    L.e13_synth_table.argtypes = [ctypes.c_uint64, ctypes.c_uint64, P(ctypes.c_int8), P(ctypes.c_int32)]
    L.e13_synth_table.restype = None
    L.e13_batch_digest.argtypes = [P(Leaf), ctypes.c_int64, ctypes.c_uint64, P(BatchSum)]
    L.e13_synth_run.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64, P(BatchSum)]
    FP = P(ctypes.c_float)
    L.e13_get_cost_log.argtypes = [ctypes.c_void_p, P(CostRec), ctypes.c_int64, P(ctypes.c_int64)]
    L.e13_cost_targets.argtypes = [ctypes.c_uint64, ctypes.c_uint64, FP]
    L.e13_cost_targets.restype = None
    L.e13_mlp_init.argtypes = [ctypes.c_uint64, FP]
    L.e13_mlp_init.restype = None
    L.e13_mlp_forward.argtypes = [ctypes.c_void_p, FP, FP, ctypes.c_int64, FP]
    L.e13_mlp_grad.argtypes = [ctypes.c_void_p, FP, FP, FP, ctypes.c_int64, FP, P(ctypes.c_double)]
    L.e13_mlp_train.argtypes = [ctypes.c_void_p, FP, FP, FP, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                ctypes.c_float, ctypes.c_uint64, FP]
    L.e13_set_cost_model.argtypes = [ctypes.c_void_p, FP]
    L.e13_autotune.argtypes = [ctypes.c_void_p, P(Problem), ctypes.c_int, ctypes.c_int, P(Tuning), P(TuneReport)]
    L.e13_save_tuning.argtypes = [P(Tuning), ctypes.c_char_p]
    L.e13_load_tuning.argtypes = [ctypes.c_char_p, P(Tuning)]
    L.e13_tuning_file.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
    L.e13_tuning_file_small.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
    L.e13_search_profile.argtypes = [ctypes.c_void_p, P(Problem), ctypes.c_int, P(Tuning)]
    L.e13_deep_check_line.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                      ctypes.c_int]
    L.e13_recheck_counter_line.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    L.e13_recheck_counter.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int64]
    L.e13_classify_deep.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64]
    L.e13_check_survivor.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    L.e13_deep_check.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                 ctypes.c_int64]
    L.e13_deep_check_gpu.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int64]
    L.e13_recheck_counter_gpu.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                          ctypes.c_int64]
    L.e13_set_checkpoint.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_int64]
    L.e13_estimate.argtypes = [ctypes.c_void_p, P(Problem), ctypes.c_uint64, ctypes.c_int64, P(ctypes.c_double),
                               P(EstimateSum)]
    L.e13_estimate_static.argtypes = L.e13_estimate.argtypes
    L.e13_get_trace.argtypes = [ctypes.c_void_p, P(ctypes.c_uint64), ctypes.c_uint64, P(ctypes.c_int32),
                                P(ctypes.c_int32)]
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        raise E13Error(rc, what)


def keys():
    """KEY[0..1363] and TGT[0..9] (e13search.c:437-444)."""
    k = (ctypes.c_uint64 * 1364)()
    t = (ctypes.c_uint64 * (NS + 1))()
    lib().e13_keys(k, t)
    return list(k), list(t)


def parse_machine(text):
    """Machine text -> list of codes (load/parse_cell, e13search.c:105-121)."""
    base = (ctypes.c_int32 * 32)()
    nq = lib().e13_parse_machine(text.encode(), base, 16)
    if nq < 0:
        raise E13Error(nq, "parse_machine")
    return list(base[:2 * nq])


def cell_index(name):
    """'D0' -> 6 (cell = 2 * state + symbol, states A = 0 ...)."""
    return 2 * (ord(name[0]) - ord("A")) + int(name[1])


def setup(base, X, fmax=0, free_cells=None, sym=None):
    """setup() + set_sym() (e13search.c:317-355, 423-426).  free_cells: names or
    indices of base cells (free mode); sym: (LO, HI) base states or None."""
    nq_base = len(base) // 2
    b = (ctypes.c_int32 * 32)(*base)
    fb = None
    if free_cells is not None:
        fb = (ctypes.c_uint8 * 32)()
        for c in free_cells:
            fb[cell_index(c) if isinstance(c, str) else int(c)] = 1
    lo, hi = sym if sym else (-1, -1)
    p = Problem()
    rc = lib().e13_setup(b, nq_base, X, fmax, fb, lo, hi, ctypes.byref(p))
    if rc != 0:
        raise E13Error(rc, f"setup X={X}")
    return p


def format_survivor(p, s):
    buf = ctypes.create_string_buffer(1024)
    lib().e13_format_survivor(ctypes.byref(p), ctypes.byref(s), buf, 1024)
    return buf.value.decode()


def format_hash(p, s):
    buf = ctypes.create_string_buffer(256)
    lib().e13_format_hash(ctypes.byref(p), ctypes.byref(s), buf, 256)
    return buf.value.decode()


def summary_line(p, counts):
    """The parity fields of the reference summary (e13search.c:413-416)."""
    return (f"X={p.X} budget={p.budget} leaves={counts.leaves} survivors={counts.survivors} "
            f"steps={counts.steps} dead-at-stage(0..9): " + " ".join(str(x) for x in counts.dead_at_stage))


def synth_table(seed0, k):
    t = (ctypes.c_int8 * 28)()
    cap = ctypes.c_int32()
    lib().e13_synth_table(seed0, k, t, ctypes.byref(cap))
    return list(t), cap.value


def batch_digest(leaves, k0=0):
    """Totals and order-independent digest of complete-table results (e13_batch_digest)."""
    n = len(leaves)
    arr = leaves if isinstance(leaves, ctypes.Array) else (Leaf * n)(*leaves)
    out = BatchSum()
    _check(lib().e13_batch_digest(arr, n, k0, ctypes.byref(out)), "e13_batch_digest")
    return out


def _floats(a, n):
    """A ctypes float array of n values from a sequence or a numpy array."""
    try:
        import numpy as np
        arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))
        if arr.size != n:
            raise ValueError(f"expected {n} floats, got {arr.size}")
        return (ctypes.c_float * n).from_buffer_copy(arr.tobytes())
    except ImportError:
        return (ctypes.c_float * n)(*a)


def mlp_init(seed):
    """Deterministic initial parameters (6402 floats)."""
    p = (ctypes.c_float * MLP_PARAMS)()
    lib().e13_mlp_init(seed, p)
    return list(p)


def cost_targets(leaves, steps):
    t = (ctypes.c_float * 2)()
    lib().e13_cost_targets(leaves, steps, t)
    return list(t)


def save_tuning(tuning, path):
    t = Tuning(**{k: v for k, v in tuning.items() if k in dict(Tuning._fields_)})
    _check(lib().e13_save_tuning(ctypes.byref(t), path.encode()), "e13_save_tuning")


def load_tuning(path):
    t = Tuning()
    _check(lib().e13_load_tuning(path.encode(), ctypes.byref(t)), "e13_load_tuning")
    return t.as_dict()


def deep_check_line(line, survivor_id, nmax, big=False):
    """The deep stage check (src/deep.c; big: src/deep_big.c) of one SURVIVOR line: its result line."""
    buf = ctypes.create_string_buffer(4096)
    _check(min(lib().e13_deep_check_line(line.encode(), survivor_id, nmax, int(big), buf, 4096), 0),
           "e13_deep_check_line")
    return buf.value.decode()


def recheck_counter_line(line, survivor_id, nmax):
    """The counter re-check (src/recheck.c) of one SURVIVOR line: its verdict line."""
    buf = ctypes.create_string_buffer(65536)
    _check(min(lib().e13_recheck_counter_line(line.encode(), survivor_id, nmax, buf, 65536), 0),
           "e13_recheck_counter_line")
    return buf.value.decode()


def recheck_counter(text, nmax, threads=8):
    """The counter re-check of every line of text with table=, numbered from 1."""
    cap = 20480 * (len(text) // 64 + 16)
    buf = ctypes.create_string_buffer(cap)
    _check(min(lib().e13_recheck_counter(text.encode(), nmax, threads, buf, cap), 0), "e13_recheck_counter")
    return buf.value.decode()


def classify_deep(survivors, deep):
    """classify_deep.py's lines for survivor lines and their deep-check lines."""
    cap = 1 << 20
    buf = ctypes.create_string_buffer(cap)
    _check(min(lib().e13_classify_deep(survivors.encode(), deep.encode(), buf, cap), 0), "e13_classify_deep")
    return buf.value.decode()


def check_survivor(line):
    """The crafted-start check (src/check.py) of one SURVIVOR line: its lines."""
    buf = ctypes.create_string_buffer(1 << 16)
    _check(min(lib().e13_check_survivor(line.encode(), buf, 1 << 16), 0), "e13_check_survivor")
    return buf.value.decode()


def deep_check(text, nmax, threads=8, big=False):
    """The deep stage check of every SURVIVOR line of text, numbered from 1."""
    cap = 2048 * (len(text) // 64 + 16)
    buf = ctypes.create_string_buffer(cap)
    _check(min(lib().e13_deep_check(text.encode(), nmax, int(big), threads, buf, cap), 0), "e13_deep_check")
    return buf.value.decode()


def recheck(tab, nq, st0, head0, abs_ok, rel_ok, steps):
    t = (ctypes.c_int8 * 32)(*([int(x) for x in tab] + [HALT] * (32 - len(tab))))
    m = ctypes.c_uint32()
    ok = lib().e13_recheck(t, nq, st0, head0, abs_ok, rel_ok, steps, ctypes.byref(m))
    return ok, m.value


class Context:
    """One HIP device context (e13_open/e13_close)."""

    def __init__(self, device=0, **tuning):
        self._h = ctypes.c_void_p()
        t = Tuning(**tuning)
        _check(lib().e13_open(device, ctypes.byref(t), ctypes.byref(self._h)), "e13_open")

    def close(self):
        if self._h:
            lib().e13_close(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_tuning(self, **kw):
        _check(lib().e13_set_tuning(self._h, ctypes.byref(Tuning(**kw))), "e13_set_tuning")

    def reset_tuning(self):
        """Back to the settings the context opened with (stored profiles, automatic choice)."""
        _check(lib().e13_set_tuning(self._h, None), "e13_set_tuning")

    # fields where 0 reads back as "off" but a set of 0 means "keep"
    _OFF_FIELDS = ("seed_per_wave", "trace_events", "graph_chain", "cost_log", "learned", "solo_steps", "evict_after",
                   "cut_live", "phase_cycles", "tape_words")

    def restore_tuning(self, t):
        """Set every field of a tuning dict read back by tuning() (off fields included)."""
        kw = {k: v for k, v in t.items() if k in dict(Tuning._fields_)}
        for k in self._OFF_FIELDS:
            if kw.get(k, 1) == 0:
                kw[k] = -1
        self.set_tuning(**kw)

    def tuning(self):
        t = Tuning()
        _check(lib().e13_get_tuning(self._h, ctypes.byref(t)), "e13_get_tuning")
        return t.as_dict()

    def stats(self):
        s = Stats()
        _check(lib().e13_get_stats(self._h, ctypes.byref(s)), "e13_get_stats")
        return s.as_dict()

    def deep_check(self, text, nmax):
        """src/deep.c's stage check of every SURVIVOR line of text on this GPU
        (numbered from 1): its result lines."""
        cap = 2048 * (len(text) // 64 + 16)
        buf = ctypes.create_string_buffer(cap)
        _check(min(lib().e13_deep_check_gpu(self._h, text.encode(), nmax, buf, cap), 0), "e13_deep_check_gpu")
        return buf.value.decode()

    def recheck_counter(self, text, nmax):
        """src/recheck.c's counter re-check of every line with table= on this GPU:
        its verdict lines."""
        cap = 20480 * (len(text) // 64 + 16)
        buf = ctypes.create_string_buffer(cap)
        _check(min(lib().e13_recheck_counter_gpu(self._h, text.encode(), nmax, buf, cap), 0),
               "e13_recheck_counter_gpu")
        return buf.value.decode()

    def search(self, p, surv_cap=200000):
        """run_delete() for one problem: (Counts, [Survivor], n_survivors)."""
        c = Counts()
        sv = (Survivor * surv_cap)()
        ns = ctypes.c_uint64()
        _check(lib().e13_search(self._h, ctypes.byref(p), ctypes.byref(c), sv, surv_cap, ctypes.byref(ns)),
               f"e13_search X={p.X}")
        return c, [sv[i] for i in range(min(ns.value, surv_cap))], ns.value

    def search_many(self, problems, surv_cap=200000):
        """Several problems in one persistent launch: ([Counts], [Survivor], n)."""
        n = len(problems)
        ps = (Problem * n)(*problems)
        cs = (Counts * n)()
        sv = (Survivor * surv_cap)()
        ns = ctypes.c_uint64()
        _check(lib().e13_search_many(self._h, ps, n, cs, sv, surv_cap, ctypes.byref(ns)), "e13_search_many")
        return [cs[i] for i in range(n)], [sv[i] for i in range(min(ns.value, surv_cap))], ns.value

    def simulate_batch(self, tabs, caps=None):
        """run_full() on each table (list of equal-length code lists)."""
        n = len(tabs)
        ncells = len(tabs[0])
        flat = (ctypes.c_int8 * (n * ncells))(*[int(v) for t in tabs for v in t])
        cp = (ctypes.c_int32 * n)(*caps) if caps is not None else None
        out = (Leaf * n)()
        _check(lib().e13_simulate_batch(self._h, flat, ncells, n, cp, out), "e13_simulate_batch")
        return out

    def simulate_batch_device(self, tabs_ptr, nq, n, caps_ptr, out_ptr):
        _check(lib().e13_simulate_batch_device(self._h, ctypes.c_void_p(tabs_ptr), nq, n,
                                               ctypes.c_void_p(caps_ptr) if caps_ptr else None,
                                               ctypes.c_void_p(out_ptr)), "e13_simulate_batch_device")

    def synth_run(self, seed0, k0, n):
        """The synthetic stress batch (tables k0..k0+n-1): a BatchSum."""
        out = BatchSum()
        _check(lib().e13_synth_run(self._h, seed0, k0, n, ctypes.byref(out)), "e13_synth_run")
        return out

    def cost_log(self):
        """[CostRec] of the last persistent search run with cost_log=1."""
        n = ctypes.c_int64()
        _check(lib().e13_get_cost_log(self._h, None, 0, ctypes.byref(n)), "e13_get_cost_log")
        out = (CostRec * max(n.value, 1))()
        _check(lib().e13_get_cost_log(self._h, out, n.value, ctypes.byref(n)), "e13_get_cost_log")
        return [out[i] for i in range(n.value)]

    def mlp_forward(self, params, x, n):
        y = (ctypes.c_float * (2 * n))()
        _check(lib().e13_mlp_forward(self._h, _floats(params, MLP_PARAMS), _floats(x, n * NFEAT), n, y),
               "e13_mlp_forward")
        return list(y)

    def mlp_grad(self, params, x, t, n):
        g = (ctypes.c_float * MLP_PARAMS)()
        loss = ctypes.c_double()
        _check(lib().e13_mlp_grad(self._h, _floats(params, MLP_PARAMS), _floats(x, n * NFEAT), _floats(t, 2 * n), n,
                                  g, ctypes.byref(loss)), "e13_mlp_grad")
        return list(g), loss.value

    def mlp_train(self, params, x, t, n, epochs=200, batch=256, lr=2e-3, seed=1):
        p = _floats(params, MLP_PARAMS)
        losses = (ctypes.c_float * epochs)()
        _check(lib().e13_mlp_train(self._h, p, _floats(x, n * NFEAT), _floats(t, 2 * n), n, epochs, batch, lr, seed,
                                   losses), "e13_mlp_train")
        return list(p), list(losses)

    def set_cost_model(self, params):
        _check(lib().e13_set_cost_model(self._h, _floats(params, MLP_PARAMS) if params is not None else None),
               "e13_set_cost_model")

    def autotune(self, problems, reps=1):
        """Measure settings on a persistent search of `problems`; installs and returns
        (tuning dict, report)."""
        arr = (Problem * len(problems))(*problems)
        t = Tuning()
        r = TuneReport()
        _check(lib().e13_autotune(self._h, arr, len(problems), reps, ctypes.byref(t), ctypes.byref(r)),
               "e13_autotune")
        return t.as_dict(), {"trials": r.trials, "first_ms": r.first_ms, "best_ms": r.best_ms}

    def tuning_file(self, small=False):
        """Path of this device's stored settings (small=True: the small-search profile)."""
        buf = ctypes.create_string_buffer(1024)
        fn = lib().e13_tuning_file_small if small else lib().e13_tuning_file
        _check(fn(self._h, buf, 1024), "e13_tuning_file")
        return buf.value.decode()

    def search_profile(self, problems):
        """(profile, settings) a persistent search of these problems runs with:
        profile 0 = the main settings, 1 = the small-search profile."""
        n = len(problems)
        ps = (Problem * n)(*problems)
        t = Tuning()
        prof = lib().e13_search_profile(self._h, ps, n, ctypes.byref(t))
        _check(min(prof, 0), "e13_search_profile")
        return prof, t.as_dict()

    def set_checkpoint(self, path, every_sec=0.0, max_launches=0):
        """Checkpoints of the host-loop search (path None: off); a stop raises
        E13Error with code -9 (E13_ESTOPPED) after the save."""
        _check(lib().e13_set_checkpoint(self._h, path.encode() if path else None, every_sec, max_launches),
               "e13_set_checkpoint")

    def estimate(self, p, probes, seed=0, per_probe=False):
        """The probe estimator (the reference's estimate mode): EstimateSum, and
        with per_probe=True also the list of each probe's four sums as doubles."""
        es = EstimateSum()
        buf = (ctypes.c_double * (4 * probes))() if per_probe else None
        _check(lib().e13_estimate(self._h, ctypes.byref(p), seed, probes, buf, ctypes.byref(es)), "e13_estimate")
        if per_probe:
            return es, [tuple(buf[4 * i:4 * i + 4]) for i in range(probes)]
        return es

    def estimate_static(self, p, probes, seed=0):
        """The probe estimator over the device's static scratch (no allocation
        in the library call; probes <= STATIC_PROBES, else E13Error -14):
        (EstimateSum, list of each probe's four sums).  Calls that overlap on
        one GPU race on that scratch."""
        es = EstimateSum()
        buf = (ctypes.c_double * (4 * max(probes, 1)))()
        _check(lib().e13_estimate_static(self._h, ctypes.byref(p), seed, probes, buf, ctypes.byref(es)),
               "e13_estimate_static")
        return es, [tuple(buf[4 * i:4 * i + 4]) for i in range(probes)]

    def sync(self):
        _check(lib().e13_sync(self._h), "e13_sync")
