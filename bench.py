"""Throughput of the TNF enumeration + simulate-and-filter path on one MI355X.

Workload (one bench step): the local-edit search around erdos_14 with at most
4 changed cells, every deleted state X = 0..13 -- 5,986,617,022 leaves and
586,158,995,069 simulated steps, the largest search of the reference's
configurations that one GPU runs.  Each step's leaf counts, step totals and
dead-at-stage histograms are compared with the reference's (the golden fixture
tests/golden/ref_dump/delete_f4.txt, the compiled reference's own run); a
mismatch aborts the run.

value = simulated TM steps per second over the K timed steps (the whole
search: every X in one persistent launch after a breadth-first frontier, plus
the survivor re-check), with the device's stored autotuned settings
(erdos-machine-13_amd/tuning/).  Leaves/s and the CPU baseline are reported
beside it.  Once each, with counts checked: the symmetric 3-state counter
search (3,078,656,272 leaves; with the reference program and the GPU side by
side on a bounded sample of its family, freesym X=13 11..13 D0 G0 L1 M1), the FMAX=3 search (36,383,467 leaves), the
2^22-table synthetic stress batch (refilling vs plain form, digest checked),
the counter search in the plain per-leaf-launch form vs the persistent form,
FMAX=4 in the plain form and with the static, learned, random and
probe-estimated frontier orders (interleaved, median of 3), and the probe
estimator (the reference's estimate mode) on
the counter search, its first probes matched bit for bit against the dump of
the reference's probe(), and the deep stage check (the reference's deep.c) of
all 2868 V1 survivors to stage 15 on the GPU, its lines equal to the
reference's file, with the reference program on a sample of it on the host.

roofline: the dominant kernel is the persistent search kernel (k_search) on
its short lane tape (k_search<0, true>, three workgroups per CU); the
machines that go past that tape finish in a second launch of k_search on the
whole tape (k_search<0, false>), reported beside it.  Its algorithmic work is
13 integer lane-ops and 8 LDS bytes per simulated step; achieved = steps
simulated inside it x 13 / its time, the time from HIP events on the
library's own stream around every launch of the timed region.
peak = the gfx950 VALU int32 issue rate, 256 CU x 64 lanes/clk x 2.4 GHz =
3.93e13 op/s (LDS: 256 CU x 128 B/clk x 2.4 GHz).  traffic = HBM bytes per
launch from the committed rocprofv3 counter summary of the same workload.

cpu_baseline: the reference program compiled from its own source
(oracle/_ref/e13search, kind "reference"), or the oracle's restatement when
that binary is absent (kind "port"), on a bounded sample of the same workload:
the FMAX=4 search of deleted state X=4 alone (52,214,301 leaves,
5,601,350,435 steps), on all host cores, median of 3 wall times; the GPU is
timed on the same sample beside it.
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
REF_DUMP = os.path.join(ROOT, "tests", "golden", "ref_dump")
PROFILES = os.path.join(ROOT, "profiles")
XS = list(range(14))
CPU_SAMPLE_X = 4  # the bounded CPU sample: FMAX=4 for this deleted state alone
OPS_PER_STEP = 13
VALU_PEAK = 256 * 64 * 2.4e9  # int32 lane-ops/s
LDS_BYTES_PER_STEP = 8        # one tape word read and written back
LDS_PEAK = 256 * 128 * 2.4e9  # B/s


def golden(fmax):
    want = {}
    with open(os.path.join(REF_DUMP, f"delete_f{fmax}.txt")) as f:
        for line in f:
            m = re.search(r"X=(\d+) .*leaves=(\d+) survivors=(\d+) steps=(\d+) dead-at-stage\(0\.\.9\):((?: \d+)+)",
                          line)
            if m:
                want[int(m.group(1))] = (int(m.group(2)), int(m.group(3)), int(m.group(4)),
                                         [int(x) for x in m.group(5).split()])
    return want


def one_pass(e13, ctx, problems, want):
    """The whole workload: every X in one search (one persistent launch)."""
    cs, sv, ns = ctx.search_many([p for _, p in problems], surv_cap=1024)
    leaves = steps = 0
    for (X, _), c in zip(problems, cs):
        got = (c.leaves, c.survivors, c.steps, list(c.dead_at_stage))
        if got != want[X]:
            raise SystemExit(f"bench: parity failure at X={X}: {got} != {want[X]}")
        leaves += c.leaves
        steps += c.steps
    if ns != len(sv) or any(s.recheck_ok != 1 for s in sv):
        raise SystemExit("bench: survivor re-check failed")
    return leaves, steps


def extra_workload(e13, ctx, name, problems, want_leaves, want_steps, want_surv):
    """One timed run of a further north-star search (counts checked)."""
    ctx.sync()
    s0 = ctx.stats()
    t0 = time.perf_counter()
    cs, sv, ns = ctx.search_many(problems, surv_cap=4096)
    dt = time.perf_counter() - t0
    s1 = ctx.stats()
    leaves = sum(c.leaves for c in cs)
    steps = sum(c.steps for c in cs)
    ok = leaves == want_leaves and ns == want_surv and (want_steps is None or steps == want_steps)
    ok = ok and all(s.recheck_ok == 1 for s in sv)
    ws = (s1["wide_ms"] - s0["wide_ms"]) / 1e3
    ks = (s1["search_ms"] - s0["search_ms"]) / 1e3 - ws
    ksteps = s1["search_steps"] - s0["search_steps"] - (s1["wide_steps"] - s0["wide_steps"])
    return {"workload": name, "leaves": leaves, "steps": steps, "survivors": ns, "bit_identical_counts": ok,
            "seconds": dt, "steps_per_s": steps / dt, "leaves_per_s": leaves / dt,
            "kernel_ms": s1["kernel_ms"] - s0["kernel_ms"],
            "settings_profile": "small-search" if s1["profile"] == 1 else "main",
            "k_search": {"lane_tape_words": s1["tape_words"], "ms": ks * 1e3, "steps": ksteps, "steps_per_s": ksteps / ks,
                         "whole_tape_rest_ms": ws * 1e3,
                         "valu_frac": ksteps * OPS_PER_STEP / ks / VALU_PEAK,
                         "lds_frac": ksteps * LDS_BYTES_PER_STEP / ks / LDS_PEAK,
                         "waves_per_simd": s1["grid"] * s1["wg_threads"] / 64 / (256 * 4),
                         "active_lanes_per_round": (s1["lane_slots"] - s0["lane_slots"]) / max(s1["rounds"] - s0["rounds"], 1),
                         "steps_per_lane_round": ksteps / max(s1["lane_slots"] - s0["lane_slots"], 1)}}


def synth_workload(e13, ctx, n=1 << 22):
    """The synthetic stress batch: 2^22 seeded random tables with log-uniform
    caps (10..2e5 steps), in the refilling form and the plain one-table-per-lane
    form; totals and digest checked against the oracle's golden line."""
    with open(os.path.join(ROOT, "tests", "golden", "synth", "synth_5eed.txt")) as f:
        want = [l.strip() for l in f if f" n={n} " in l][0]
    out = {"workload": f"synthetic stress batch, {n} tables, seed 0x5EED"}
    for form, name in ((2, "refill"), (1, "plain")):
        ctx.set_tuning(form=form)
        sm = ctx.synth_run(0x5EED, 0, n)
        out[name] = {"kernel_ms": sm.kernel_ms, "steps_per_s": sm.steps / (sm.kernel_ms * 1e-3),
                     "tables_per_s": n / (sm.kernel_ms * 1e-3), "bit_identical_digest": sm.line(0x5EED, 0) == want}
    ctx.reset_tuning()
    out["steps"] = sm.steps
    out["refill_over_plain"] = out["plain"]["kernel_ms"] / out["refill"]["kernel_ms"]
    return out


def plain_vs_batched(e13, ctx, problem):
    """The 3-state counter search in the plain per-leaf-launch form (form 1:
    one machine per lane, one launch per batch, host loop) against the
    persistent batched form (form 3), counts checked both ways."""
    res = {"workload": "3-state counter search: plain per-leaf-launch form vs persistent form"}
    for form in (1, 3):
        ctx.set_tuning(form=form)
        ctx.sync()
        s0 = ctx.stats()
        t0 = time.perf_counter()
        cs, sv, ns = ctx.search_many([problem], surv_cap=4096) if form == 3 else (lambda r: ([r[0]], r[1], r[2]))(
            ctx.search(problem, surv_cap=4096))
        dt = time.perf_counter() - t0
        s1 = ctx.stats()
        ok = cs[0].leaves == 3078656272 and cs[0].steps == 319112357176 and ns == 61
        res[f"form{form}"] = {"seconds": dt, "kernel_ms": s1["kernel_ms"] - s0["kernel_ms"],
                              "launches": s1["launches"] - s0["launches"], "steps_per_s": cs[0].steps / dt,
                              "bit_identical_counts": ok}
    ctx.reset_tuning()
    res["persistent_over_plain"] = res["form1"]["seconds"] / res["form3"]["seconds"]
    return res


def learned_vs_static(e13, ctx, base):
    """The FMAX=4 search (5,986,617,022 leaves) three ways: the plain per-leaf-launch
    form (form 1, host loop), the persistent form with the static split of a
    breadth-first frontier, and the same with the frontier split and ordered by
    the cost model trained on the GPU (hand-written kernels) from the FMAX=2 and
    FMAX=3 searches' own cost logs.  Counts checked every time."""
    import numpy as np
    want = (5986617022, 586158995069, 57)

    def run(problems, **tune):
        ctx.set_tuning(**tune)
        ctx.sync()
        s0 = ctx.stats()
        t0 = time.perf_counter()
        if tune.get("form", 3) == 3:
            cs, sv, ns = ctx.search_many(problems, surv_cap=4096)
        else:
            cs, ns = [], 0
            for p in problems:
                c, sv, n1 = ctx.search(p, surv_cap=4096)
                cs.append(c)
                ns += n1
        dt = time.perf_counter() - t0
        s1 = ctx.stats()
        ok = (sum(c.leaves for c in cs), sum(c.steps for c in cs), ns) == want
        return {"seconds": dt, "kernel_ms": s1["kernel_ms"] - s0["kernel_ms"], "launches": s1["launches"] - s0["launches"],
                "steps_per_s": want[1] / dt, "bit_identical_counts": ok}

    def log(fmax):
        ctx.set_tuning(seed_per_wave=4, cost_log=1, learned=-1)
        ctx.search_many([e13.setup(base, X, fmax) for X in XS], surv_cap=4096)
        ctx.set_tuning(cost_log=-1)
        recs = ctx.cost_log()
        return (np.array([list(r.x) for r in recs], dtype=np.float32),
                np.array([e13.cost_targets(r.leaves, r.steps) for r in recs], dtype=np.float32))

    f4 = [e13.setup(base, X, 4) for X in XS]
    out = {"workload": "delete FMAX=4 X=0..13: plain per-leaf-launch form vs persistent static split vs learned "
                       "split vs random order vs probe-estimated split (16 probes per frontier node; the four "
                       "persistent runs interleaved, median of 3)"}
    try:
        x2, t2 = log(2)
        x3, t3 = log(3)
        x, t = np.concatenate([x2, x3]), np.concatenate([t2, t3])
        t0 = time.perf_counter()
        params, losses = ctx.mlp_train(e13.mlp_init(1), x, t, len(x), epochs=150, batch=256, lr=2e-3, seed=1)
        out["training"] = {"samples": len(x), "epochs": 150, "seconds": time.perf_counter() - t0,
                           "loss_first_epoch": losses[0], "loss_last_epoch": losses[-1]}
        ctx.set_cost_model(params)
        # static, learned and random frontier orders, interleaved, median of 3 each
        runs = {"static_split": [], "learned_split": [], "random_order": [], "probe_order": []}
        for _ in range(3):
            for name, learned in (("static_split", -1), ("learned_split", 1), ("random_order", 2), ("probe_order", 3)):
                ctx.reset_tuning()
                runs[name].append(run(f4, seed_per_wave=4, learned=learned))
        for name, rs in runs.items():
            med = sorted(rs, key=lambda r: r["seconds"])[1]
            out[name] = dict(med, runs_s=[round(r["seconds"], 4) for r in rs],
                             bit_identical_counts=all(r["bit_identical_counts"] for r in rs))
        ctx.reset_tuning()
        out["plain_form1"] = run(f4, form=1, seed_per_wave=-1, learned=-1)
    finally:
        ctx.reset_tuning()
        ctx.set_cost_model(None)
    out["learned_over_plain"] = out["plain_form1"]["seconds"] / out["learned_split"]["seconds"]
    out["learned_over_static"] = out["static_split"]["seconds"] / out["learned_split"]["seconds"]
    out["random_over_static"] = out["static_split"]["seconds"] / out["random_order"]["seconds"]
    out["probe_over_static"] = out["static_split"]["seconds"] / out["probe_order"]["seconds"]
    out["learned_over_probe"] = out["probe_order"]["seconds"] / out["learned_split"]["seconds"]
    return out


def estimator_workload(e13, ctx, base, probes=1 << 20):
    """The probe estimator (estimate mode) on the 3-state counter search: the
    committed dump of the reference's own probe() is matched bit for bit
    first, then 2^20 probes are timed (kernel time from HIP events)."""
    p = e13.setup(base, 13, free_cells="D0 G0 K0 K1 L1 M1".split(), sym=(10, 13))
    with open(os.path.join(ROOT, "tests", "golden", "ref_dump", "probe_counter3.txt")) as f:
        lines = f.read().splitlines()
    seed, n = int(lines[0].split()[6]), int(lines[0].split()[7])
    rows = [tuple(int(v, 16) for v in l.split()[1:5]) for l in lines[2:2 + n]]
    import struct
    _, per = ctx.estimate(p, n, seed=seed, per_probe=True)
    got = [tuple(struct.unpack("<Q", struct.pack("<d", v))[0] for v in t) for t in per]
    if got != rows:
        raise SystemExit("estimator: per-probe sums differ from the reference's probe()")
    ctx.estimate(p, 4096, seed=1)  # warm
    t0 = time.perf_counter()
    es = ctx.estimate(p, probes, seed=0)
    wall = time.perf_counter() - t0
    ks = es.kernel_ms / 1e3
    return {"workload": f"estimate (probe estimator) of freesym X=13 10..13 D0 G0 K0 K1 L1 M1, {probes} probes",
            "kernel": "k_probe", "bit_identical_to_reference_probe": True, "probes": probes,
            "seconds": wall, "kernel_ms": es.kernel_ms, "probes_per_s": probes / ks,
            "steps_per_s": es.sim_steps / ks, "steps": es.sim_steps,
            "valu_frac": es.sim_steps * OPS_PER_STEP / ks / VALU_PEAK,
            "active_lanes_per_round": es.lane_slots / max(es.rounds, 1),
            "estimate": {"leaves": [es.mean[0], es.se[0]], "steps": [es.mean[2], es.se[2]],
                         "exact_leaves": 3078656272, "exact_steps": 319112357176}}


def deep_workload(e13, ctx, nmax=15, sample=64, with_cpu=True):
    """The deep stage check (the reference's src/deep.c) of all 2868 survivors
    of V1's 1RD search to stage 15 on the GPU, its lines compared with the
    reference's own output (results/4_division_6/v1/deep15_1RD.txt, whose
    survivor numbers restart per chunk, so compared after the number); and on
    the first `sample` survivors the GPU beside the reference program compiled
    from its source (oracle/_ref/deep, one process per host core, each on its
    share of the sample), or the library's host threads without it."""
    import gzip
    rr = os.path.join(ROOT, "tests", "golden", "ref_results")
    surv = gzip.open(os.path.join(rr, "4_division_6_v1_surv_1RD.txt.gz"), "rt").read().splitlines()
    want = gzip.open(os.path.join(rr, "4_division_6_v1_deep15_1RD.txt.gz"), "rt").read().splitlines()
    s0 = ctx.stats()
    t0 = time.perf_counter()
    got = ctx.deep_check("\n".join(surv) + "\n", nmax).splitlines()
    dt = time.perf_counter() - t0
    s1 = ctx.stats()
    ok = len(got) == len(want) and all(g.split(" ", 2)[2] == w.split(" ", 2)[2] for g, w in zip(got, want))
    out = {"workload": f"deep stage check (deep.c) of the {len(surv)} V1 1RD survivors to stage {nmax}",
           "kernel": "k_deep", "seconds": dt, "kernel_ms": s1["kernel_ms"] - s0["kernel_ms"],
           "launches": s1["launches"] - s0["launches"], "survivors_per_s": len(surv) / dt,
           "lines_equal_reference": ok}
    text = "\n".join(surv[:sample]) + "\n"
    t0 = time.perf_counter()
    ctx.deep_check(text, nmax)
    gdt = time.perf_counter() - t0
    out["sample"] = {"survivors": sample, "gpu_seconds": gdt}
    if with_cpu:
        model, ncpu = host_cpu()
        ref = os.path.join(ROOT, "oracle", "_ref", "deep")
        t0 = time.perf_counter()
        if os.path.exists(ref):
            parts = [surv[i:sample:ncpu] for i in range(min(ncpu, sample))]
            procs = [subprocess.Popen([ref, str(nmax)], stdin=subprocess.PIPE, stdout=subprocess.DEVNULL, text=True)
                     for _ in parts]
            for pr, part in zip(procs, parts):
                pr.stdin.write("\n".join(part) + "\n")
                pr.stdin.close()
            for pr in procs:
                pr.wait()
            kind, cores = "reference", len(parts)
        else:
            e13.deep_check(text, nmax, threads=ncpu)
            kind, cores = "port", ncpu
        cdt = time.perf_counter() - t0
        out["sample"]["cpu"] = {"seconds": cdt, "kind": kind, "cores": cores, "cpu_model": model,
                                "survivors_per_s": sample / cdt}
        out["sample"]["gpu_over_cpu"] = cdt / gdt
    return out


def cpu_estimate_baseline(probes=8000000, max_s=120):
    """The reference program's own estimate mode on the same problem, all host cores."""
    ref = os.path.join(ROOT, "oracle", "_ref", "e13search")
    port = os.path.join(ROOT, "oracle", "_build", "e13oracle")
    exe, kind = (ref, "reference") if os.path.exists(ref) else (port, "port")
    if not os.path.exists(exe):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    t0 = time.time()
    try:
        subprocess.run([exe, "estimate", MACHINE, str(probes), "13", "0", "10", "13"] + "D0 G0 K0 K1 L1 M1".split(),
                       capture_output=True, text=True, env=env, timeout=max_s, check=True)
    except (subprocess.TimeoutExpired, subprocess.CalledProcessError):
        return None
    wall = time.time() - t0
    return {"value": probes / wall, "unit": "probes/s", "cores": threads, "kind": kind,
            "sample": f"{probes} probes of the same estimate problem in {wall:.1f} s wall"}


def host_cpu():
    """The host's CPU model and core count (lscpu / nproc)."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        m = re.search(r"^Model name:\s*(.+)$", out, re.M)
        model = m.group(1).strip() if m else None
    except (OSError, subprocess.TimeoutExpired):
        pass
    return model, os.cpu_count()


def cpu_baseline(runs=3, max_s=150):
    """The reference CPU search (or the oracle port) on the bounded sample:
    delete FMAX=4 for X=CPU_SAMPLE_X alone, all host cores, median of `runs`
    wall times (process start to exit); counts checked against the fixture."""
    ref = os.path.join(ROOT, "oracle", "_ref", "e13search")
    port = os.path.join(ROOT, "oracle", "_build", "e13oracle")
    exe, kind = (ref, "reference") if os.path.exists(ref) else (port, "port")
    if not os.path.exists(exe):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    want = golden(4)[CPU_SAMPLE_X]
    walls = []
    for _ in range(runs):
        t0 = time.time()
        try:
            out = subprocess.run([exe, "delete", MACHINE, "4", str(CPU_SAMPLE_X)], capture_output=True, text=True,
                                 env=env, timeout=max_s)
        except subprocess.TimeoutExpired:
            return None
        walls.append(time.time() - t0)
        summ = [l for l in out.stderr.splitlines() if "leaves=" in l]  # not the progress lines
        leaves = int(re.search(r"leaves=(\d+)", summ[0]).group(1))
        steps = int(re.search(r"steps=(\d+)", summ[0]).group(1))
        if (leaves, steps) != (want[0], want[2]):
            raise SystemExit(f"cpu baseline: counts differ from the reference fixture: {leaves} {steps}")
        print(f"bench: cpu sample run {len(walls)}: {walls[-1]:.1f} s", file=sys.stderr, flush=True)
    wall = sorted(walls)[len(walls) // 2]
    model, ncpu = host_cpu()
    return {"value": steps / wall, "unit": "steps/s", "cores": threads, "kind": kind,
            "sample": f"delete FMAX=4 X={CPU_SAMPLE_X} alone ({leaves} leaves, {steps} steps), median of "
                      f"{runs} wall times {[round(w, 2) for w in walls]} s, OpenMP schedule(dynamic,1), E13_SPLIT=2",
            "leaves_per_s": leaves / wall, "cpu_model": model, "nproc": ncpu}


def counter_cpu_sample(e13, ctx, base, runs=3, max_s=150):
    """The counter search's family on a bounded sample both ways: the reference
    program (oracle/_ref/e13search, or the oracle port) on freesym X=13 with
    states 11..13 symmetric and free cells D0 G0 L1 M1 (96,010,408 leaves,
    10,716,181,112 steps), all host cores, median of `runs` wall times; the
    GPU on the same problem beside it, its counts equal to the CPU's."""
    ref = os.path.join(ROOT, "oracle", "_ref", "e13search")
    port = os.path.join(ROOT, "oracle", "_build", "e13oracle")
    exe, kind = (ref, "reference") if os.path.exists(ref) else (port, "port")
    if not os.path.exists(exe):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    cells = "D0 G0 L1 M1".split()
    walls, summ = [], None
    for _ in range(runs):
        t0 = time.time()
        try:
            out = subprocess.run([exe, "freesym", MACHINE, "13", "11", "13"] + cells, capture_output=True, text=True,
                                 env=env, timeout=max_s)
        except subprocess.TimeoutExpired:
            return None
        walls.append(time.time() - t0)
        summ = [l for l in out.stderr.splitlines() if "leaves=" in l][0]
        print(f"bench: counter cpu sample run {len(walls)}: {walls[-1]:.1f} s", file=sys.stderr, flush=True)
    m = re.search(r"leaves=(\d+) survivors=(\d+) steps=(\d+) .*dead-at-stage\(0\.\.9\):((?: \d+)+)", summ)
    cpu = (int(m.group(1)), int(m.group(2)), int(m.group(3)), [int(x) for x in m.group(4).split()])
    wall = sorted(walls)[len(walls) // 2]
    p = e13.setup(base, 13, free_cells=cells, sym=(11, 13))
    gw = []
    for _ in range(runs):
        ctx.sync()
        t0 = time.perf_counter()
        cs, sv, ns = ctx.search_many([p], surv_cap=1024)
        gw.append(time.perf_counter() - t0)
    c = cs[0]
    same = (c.leaves, c.survivors, c.steps, list(c.dead_at_stage)) == cpu
    if not same:
        raise SystemExit(f"bench: counter sample parity failure: {c.leaves} {c.steps} != the CPU's {cpu}")
    gwall = sorted(gw)[len(gw) // 2]
    model, ncpu = host_cpu()
    return {"sample": f"freesym X=13 11..13 D0 G0 L1 M1 ({cpu[0]} leaves, {cpu[2]} steps)",
            "cpu": {"value": cpu[2] / wall, "unit": "steps/s", "cores": threads, "kind": kind,
                    "walls_s": [round(w, 2) for w in walls], "cpu_model": model, "nproc": ncpu},
            "gpu": {"value": cpu[2] / gwall, "unit": "steps/s", "walls_s": [round(w, 4) for w in gw]},
            "counts_equal": same, "gpu_over_cpu": wall / gwall}


def gpu_on_cpu_sample(e13, ctx, base, runs=3):
    """The GPU on the CPU baseline's sample (FMAX=4, X=CPU_SAMPLE_X alone):
    median of `runs` search calls, counts checked."""
    want = golden(4)[CPU_SAMPLE_X]
    p = e13.setup(base, CPU_SAMPLE_X, 4)
    walls = []
    for _ in range(runs):
        ctx.sync()
        t0 = time.perf_counter()
        cs, sv, ns = ctx.search_many([p], surv_cap=1024)
        walls.append(time.perf_counter() - t0)
        if (cs[0].leaves, cs[0].survivors, cs[0].steps, list(cs[0].dead_at_stage)) != want:
            raise SystemExit("bench: parity failure on the CPU sample")
    wall = sorted(walls)[len(walls) // 2]
    return {"steps_per_s": want[2] / wall, "leaves_per_s": want[0] / wall, "seconds": wall,
            "runs_s": [round(w, 4) for w in walls]}


def traffic_from_profiles(workload):
    """HBM bytes per k_search launch from the committed rocprofv3 counter
    summary of the same workload (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration)."""
    path = os.path.join(PROFILES, f"k_search_{workload}_counters.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = re.search(r"^traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read(), re.M)
    return float(m.group(1)) if m else None


WORKLOADS = {
    "fmax4": "delete FMAX=4 X=0..13 (5,986,617,022 leaves, 586,158,995,069 steps per pass)",
    "fmax3": "delete FMAX=3 X=0..13 (36,383,467 leaves, 4,006,276,021 steps per pass)",
    "counter3": "freesym X=13 10..13 D0 G0 K0 K1 L1 M1 (3,078,656,272 leaves, 319,112,357,176 steps per pass)",
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the further north-star workloads")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="fmax4",
                    help="timed workload (fmax4 is the bench line; the others for profiling the kernel)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (rocprof FETCH/WRITE_SIZE calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 18  # float32 elements
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 3}))
        return
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
    counter = e13.setup(base, 13, free_cells="D0 G0 K0 K1 L1 M1".split(), sym=(10, 13))
    if args.workload == "counter3":
        problems = [(13, counter)]
        want = {13: (3078656272, 61, 319112357176, [0, 2957671647, 119843176, 1115937, 22970, 2236, 1, 0, 244, 0])}
    else:
        fmax = 4 if args.workload == "fmax4" else 3
        problems = [(X, e13.setup(base, X, fmax)) for X in XS]
        want = golden(fmax)
    ctx = e13.Context(0)
    # the stored settings profile this search runs with (small searches have their own)
    profile, tuning = ctx.search_profile([p for _, p in problems])
    tuning_file = os.path.relpath(ctx.tuning_file(small=profile == 1), ROOT)
    tuning_file = tuning_file if os.path.exists(os.path.join(ROOT, tuning_file)) else None
    for _ in range(args.warmup):
        one_pass(e13, ctx, problems, want)
        print("bench: warmup pass done", file=sys.stderr, flush=True)
    ctx.sync()
    s0 = ctx.stats()
    t0 = time.perf_counter()
    leaves = steps = 0
    for k in range(args.steps):
        l, st = one_pass(e13, ctx, problems, want)
        leaves += l
        steps += st
        if k % 5 == 4:
            print(f"bench: {k + 1} timed passes", file=sys.stderr, flush=True)
    ctx.sync()
    dt = time.perf_counter() - t0
    s1 = ctx.stats()
    # one more pass, untimed, with the per-phase wave cycle counts on (they cost about 2%)
    ctx.restore_tuning(dict(tuning, phase_cycles=1))
    c0 = ctx.stats()["cycles"]
    one_pass(e13, ctx, problems, want)
    cyc = [b - a for a, b in zip(c0, ctx.stats()["cycles"])]
    ctx.reset_tuning()
    # the roofline is the persistent kernel's own: its HIP-event time, launches and
    # the steps simulated inside it (the frontier launches before it are apart).
    # With a short lane tape the search is two launches of k_search per pass: the
    # dominant short-tape one, and one on the whole tape for the machines that
    # went past it (reported beside it)
    wide_s = (s1["wide_ms"] - s0["wide_ms"]) / 1e3
    wide_launches = s1["wide_launches"] - s0["wide_launches"]
    wide_steps = s1["wide_steps"] - s0["wide_steps"]
    kernel_s = (s1["search_ms"] - s0["search_ms"]) / 1e3 - wide_s
    launches = s1["search_launches"] - s0["search_launches"] - wide_launches
    ksteps = s1["search_steps"] - s0["search_steps"] - wide_steps
    frontier_ms = (s1["kernel_ms"] - s0["kernel_ms"]) - (kernel_s + wide_s) * 1e3
    achieved = ksteps * OPS_PER_STEP / kernel_s
    short_tape = s1["tape_words"] < 43
    kname = "k_search<0, true>" if short_tape else "k_search<0, false>"
    extra = []
    if not args.no_extra:
        extra.append(extra_workload(e13, ctx, "3-state counter, freesym X=13 10..13 D0 G0 K0 K1 L1 M1", [counter],
                                    3078656272, 319112357176, 61))
        if not args.no_cpu:
            extra[-1]["cpu_same_sample"] = counter_cpu_sample(e13, ctx, base)
        print("bench: counter search done", file=sys.stderr, flush=True)
        f3 = [e13.setup(base, X, 3) for X in XS]
        extra.append(extra_workload(e13, ctx, "delete FMAX=3 X=0..13", f3, 36383467, 4006276021, 11))
        extra.append(synth_workload(e13, ctx))
        print("bench: synthetic batch done", file=sys.stderr, flush=True)
        extra.append(plain_vs_batched(e13, ctx, counter))
        print("bench: plain vs batched done", file=sys.stderr, flush=True)
        extra.append(learned_vs_static(e13, ctx, base))
        print("bench: frontier orders done", file=sys.stderr, flush=True)
        est = estimator_workload(e13, ctx, base)
        if not args.no_cpu:
            est["cpu"] = cpu_estimate_baseline()
        extra.append(est)
        print("bench: estimator done", file=sys.stderr, flush=True)
        extra.append(deep_workload(e13, ctx, with_cpu=not args.no_cpu))
        print("bench: deep check done", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu and args.workload == "fmax4":
        cpu = cpu_baseline()
        if cpu is not None:
            cpu["gpu_same_sample"] = gpu_on_cpu_sample(e13, ctx, base)
            cpu["gpu_over_cpu_same_sample"] = cpu["gpu_same_sample"]["steps_per_s"] / cpu["value"]
    line = {
        "metric": "simulated TM steps/s (leaves/s beside it) on 1 MI355X, bit-identical survivors, vs CPU",
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
        "data": "the reference's own search input (erdos_14 and its edits); no synthetic data",
        "config": {"workload": WORKLOADS[args.workload],
                   "form": s1["form"], "slice": s1["slice"], "wg_threads": s1["wg_threads"], "grid": s1["grid"],
                   "gate": "ones-count" if s1["gate"] == 1 else "zobrist-hash",
                   "settings_profile": "small-search" if s1["profile"] == 1 else "main",
                   "tuning": {k: v for k, v in tuning.items() if k in ("wgs_per_cu", "slice", "refill_tries", "solo_steps",
                                                                       "seed_per_wave", "donate_max", "sched_every",
                                                                       "idle_naps", "evict_after")},
                   "tuning_file": tuning_file},
        "leaves_per_s": leaves / dt,
        "bit_identical": True,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK / 1e12, "unit": "TOP/s",
                     "frac": achieved / VALU_PEAK, "traffic": traffic_from_profiles(args.workload),
                     "lds": {"achieved_GBps": ksteps * LDS_BYTES_PER_STEP / kernel_s / 1e9,
                             "peak_GBps": LDS_PEAK / 1e9, "frac": ksteps * LDS_BYTES_PER_STEP / kernel_s / LDS_PEAK},
                     "kernel": kname, "lane_tape_words": s1["tape_words"], "launches": launches,
                     "kernel_ms_per_launch": kernel_s * 1e3 / max(launches, 1),
                     "steps_per_launch": ksteps / max(launches, 1),
                     "steps_per_s_kernel": ksteps / kernel_s,
                     "frontier_kernels_ms_per_pass": frontier_ms / args.steps,
                     "whole_tape_rest": {"kernel": "k_search<0, false>", "launches": wide_launches,
                                         "ms_per_pass": wide_s * 1e3 / args.steps,
                                         "steps_per_pass": wide_steps / args.steps,
                                         "nodes_per_pass": (s1["wide_nodes"] - s0["wide_nodes"]) / args.steps},
                     # occupancy and divergence: resident waves per SIMD (256 CUs x 4 SIMDs),
                     # running lanes per wave round (of 64), steps per lane per round
                     "waves_per_simd": s1["grid"] * s1["wg_threads"] / 64 / (256 * 4),
                     "active_lanes_per_round": (s1["lane_slots"] - s0["lane_slots"]) / max(s1["rounds"] - s0["rounds"], 1),
                     "steps_per_lane_round": ksteps / max(s1["lane_slots"] - s0["lane_slots"], 1),
                     "wave_cycle_split": {"refill": cyc[0] / max(sum(cyc), 1), "steps": cyc[1] / max(sum(cyc), 1),
                                          "finish": cyc[2] / max(sum(cyc), 1), "pool_idle": cyc[3] / max(sum(cyc), 1)}},
        "north_star_workloads": extra,
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
