"""Throughput of the TNF enumeration + simulate-and-filter path on one MI355X.

Workload (one bench step): the local-edit search around erdos_14 with at most
3 changed cells, every deleted state X = 0..13 -- 36,383,467 leaves and
4,006,276,021 simulated steps.  Each step's leaf counts, step totals and
dead-at-stage histograms are compared with the reference's (the golden
fixture tests/golden/ref_dump/delete_f3.txt); a mismatch aborts the run.

value = simulated TM steps per second over the K timed steps (the whole
search: kernels, the host driver loop and the survivor re-check).  Leaves/s
and the CPU baseline are reported beside it.

roofline: the dominant kernel is the expansion kernel.  Its algorithmic work
is 13 integer lane-ops per simulated step; achieved = steps x 13 / kernel
time, with the kernel time taken from HIP events on the library's own stream
around every launch of the timed region.  peak = the gfx950 VALU int32 issue
rate, 256 CU x 64 lanes/clk x 2.4 GHz = 3.93e13 op/s.

cpu_baseline: the reference program compiled from its own source
(oracle/_ref/e13search, kind "reference"), or the oracle's restatement when
that binary is absent (kind "port"), on the same search, timed on this host.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
MACHINE = os.path.join(ROOT, "tests", "golden", "machines", "erdos_14.txt")
GOLDEN = os.path.join(ROOT, "tests", "golden", "ref_dump", "delete_f3.txt")
PROFILES = os.path.join(ROOT, "profiles")
FMAX = 3
XS = list(range(14))
OPS_PER_STEP = 13
VALU_PEAK = 256 * 64 * 2.4e9  # int32 lane-ops/s


def golden():
    want = {}
    with open(GOLDEN) as f:
        for line in f:
            m = re.search(r"X=(\d+) .*leaves=(\d+) survivors=(\d+) steps=(\d+) dead-at-stage\(0\.\.9\):((?: \d+)+)",
                          line)
            if m:
                want[int(m.group(1))] = (int(m.group(2)), int(m.group(3)), int(m.group(4)),
                                         [int(x) for x in m.group(5).split()])
    return want


def one_pass(e13, ctx, problems, want):
    leaves = steps = 0
    for X, p in problems:
        c, sv, ns = ctx.search(p, surv_cap=1024)
        got = (c.leaves, c.survivors, c.steps, list(c.dead_at_stage))
        if got != want[X]:
            raise SystemExit(f"bench: parity failure at X={X}: {got} != {want[X]}")
        if any(s.recheck_ok != 1 for s in sv):
            raise SystemExit(f"bench: survivor re-check failed at X={X}")
        leaves += c.leaves
        steps += c.steps
    return leaves, steps


def cpu_baseline(max_s=120):
    """The reference CPU search (or the oracle port) on the same workload."""
    ref = os.path.join(ROOT, "oracle", "_ref", "e13search")
    port = os.path.join(ROOT, "oracle", "_build", "e13oracle")
    exe, kind = (ref, "reference") if os.path.exists(ref) else (port, "port")
    if not os.path.exists(exe):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    t0 = time.time()
    try:
        out = subprocess.run([exe, "delete", MACHINE, str(FMAX)] + [str(x) for x in XS], capture_output=True,
                             text=True, env=env, timeout=max_s)
    except subprocess.TimeoutExpired:
        return None
    wall = time.time() - t0
    summaries = [l for l in out.stderr.splitlines() if "leaves=" in l]  # not the progress lines
    steps = sum(int(re.search(r"steps=(\d+)", l).group(1)) for l in summaries)
    leaves = sum(int(re.search(r"leaves=(\d+)", l).group(1)) for l in summaries)
    return {"value": steps / wall, "unit": "steps/s", "cores": threads, "kind": kind,
            "sample": f"the full workload (delete FMAX={FMAX}, X=0..13): {leaves} leaves, {steps} steps in "
                      f"{wall:.1f} s wall, OpenMP schedule(dynamic,1), E13_SPLIT=2",
            "leaves_per_s": leaves / wall}


def traffic_from_profiles():
    """HBM bytes per expansion launch from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    path = os.path.join(PROFILES, "expand_fmax3_counters.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = re.search(r"^traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read(), re.M)
    return float(m.group(1)) if m else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the search shards by subtree across GPUs, but no multi-GPU "
              "run is part of it (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import importlib.util
    spec = importlib.util.spec_from_file_location("e13amd", os.path.join(ROOT, "erdos-machine-13_amd",
                                                                         "__init__.py"))
    e13 = importlib.util.module_from_spec(spec)
    sys.modules["e13amd"] = e13
    spec.loader.exec_module(e13)
    e13.lib()  # raises when the HIP library has not been built

    with open(MACHINE) as f:
        base = e13.parse_machine(f.read())
    want = golden()
    problems = [(X, e13.setup(base, X, FMAX)) for X in XS]
    ctx = e13.Context(0)
    for _ in range(args.warmup):
        one_pass(e13, ctx, problems, want)
    ctx.sync()
    s0 = ctx.stats()
    t0 = time.perf_counter()
    leaves = steps = 0
    for _ in range(args.steps):
        l, s = one_pass(e13, ctx, problems, want)
        leaves += l
        steps += s
    ctx.sync()
    dt = time.perf_counter() - t0
    s1 = ctx.stats()
    kernel_s = (s1["kernel_ms"] - s0["kernel_ms"]) / 1e3
    launches = s1["launches"] - s0["launches"]
    achieved = steps * OPS_PER_STEP / kernel_s
    line = {
        "metric": "simulated TM steps/s, local-edit search FMAX=3 (X=0..13) on erdos_14, bit-identical to the "
                  "reference",
        "value": steps / dt,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "the reference's own search input (erdos_14, deleted states 0..13, <=3 changed cells); no "
                "synthetic data",
        "config": {"workload": "delete FMAX=3 X=0..13 (36,383,467 leaves, 4,006,276,021 steps per pass)",
                   "form": s1["form"], "slice": s1["slice"], "wg_threads": s1["wg_threads"], "grid": s1["grid"],
                   "gate": "ones-count" if s1["gate"] == 1 else "zobrist-hash"},
        "leaves_per_s": leaves / dt,
        "bit_identical": True,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK / 1e12, "unit": "TOP/s",
                     "frac": achieved / VALU_PEAK, "traffic": traffic_from_profiles(),
                     "kernel": "k_expand", "launches": launches,
                     "kernel_ms_per_launch": kernel_s * 1e3 / max(launches, 1),
                     "steps_per_s_kernel": steps / kernel_s},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(),
    }
    ctx.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
