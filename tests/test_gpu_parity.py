"""GPU parity: the HIP library against the oracle and the golden fixtures.

Integer work, so every comparison is exact: leaf counts, step totals,
dead-at-stage histograms, sorted survivor lines and counter hashes.  Each test
prints the case it runs before touching the GPU.
"""
import ctypes

import pytest

from conftest import (REF_DUMP, REF_RESULTS, XS, oracle_full, oracle_lib, parse_summaries, read_machine, read_text,
                      ref_hashes, survivor_lines, synth_golden)

pytestmark = pytest.mark.gpu

T14S = (0, 2, 22, 90, 268, 944, 3434, 13450, 52068, 208864)


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def search_all(e13, ctx, base, Xs, fmax=0, free=None, sym=None):
    """Summaries {X: dict}, sorted SURVIVOR lines, and (line, hash) pairs."""
    sums, lines, hashes = {}, [], []
    for X in Xs:
        try:
            p = e13.setup(base, X, fmax, free, sym)
        except e13.E13Error as ex:
            assert ex.code == -4
            continue
        c, sv, ns = ctx.search(p)
        d = c.as_dict()
        sums[X] = {"leaves": d["leaves"], "survivors": d["survivors"], "steps": d["steps"], "hist": d["hist"]}
        assert ns == len(sv) == c.survivors
        for s in sv:
            assert s.recheck_ok == 1
            line = e13.format_survivor(p, s)
            lines.append(line)
            hashes.append((X, line.split("table=")[1].replace("?", ""), tuple(s.ch)))
    return sums, sorted(lines), hashes


@pytest.fixture(scope="module")
def base(e13):
    return e13.parse_machine(read_machine())


def test_base_machine(e13, gpu_ctx, base):
    say("simulate_batch: erdos_14 base")
    lv = gpu_ctx.simulate_batch([base])
    want = read_text(REF_RESULTS, "validation_base_erdos14.txt").splitlines()[1]
    got = (f"base states=14 result={'SURV' if lv[0].result == -2 else 'DEAD'} stage={lv[0].stage} "
           f"st0={lv[0].st0} head0={lv[0].head0} abs={lv[0].abs_ok} rel={lv[0].rel_ok} steps={lv[0].steps}")
    assert got == want.strip()


def test_mutants(e13, gpu_ctx, base):
    say("simulate_batch: 1568 single-cell mutants")
    tabs, keys = [], []
    for i in range(28):
        for v in range(-1, 56):
            if v == base[i]:
                continue
            t = list(base)
            t[i] = v
            tabs.append(t)
            keys.append((i, v))
    lv = gpu_ctx.simulate_batch(tabs)
    got = [f"{i} {v} {lv[k].stage} {int(lv[k].result == -2)}" for k, (i, v) in enumerate(keys)]
    want = [l for l in read_text(REF_RESULTS, "validation_mutants.txt").splitlines() if not l.startswith("#")]
    assert got == want


@pytest.mark.parametrize("gate", [1, 2])
def test_synthetic_batch_matches_oracle(e13, gpu_ctx, gate):
    n = 1 << 13
    say(f"simulate_batch: {n} synthetic stress tables, gate={gate}")
    tabs, caps = [], []
    for k in range(n):
        t, c = e13.synth_table(0x5EED, k)
        tabs.append(t)
        caps.append(c)
    gpu_ctx.set_tuning(gate=gate)
    try:
        lv = gpu_ctx.simulate_batch(tabs, caps)
    finally:
        gpu_ctx.reset_tuning()
    L = oracle_lib()
    for k in range(n):
        want = oracle_full(L, tabs[k], [caps[k]] * 10)
        got = (lv[k].result, lv[k].stage, lv[k].steps, lv[k].st0, lv[k].head0, lv[k].abs_ok, lv[k].rel_ok,
               tuple(lv[k].ch))
        assert got == want, (k, got, want)


@pytest.mark.parametrize("n,form,gate", [(1 << 16, 2, 2), (1 << 16, 1, 1), (1 << 22, 2, 1), (1 << 22, 1, 1)])
def test_synthetic_stress_batch_matches_golden(e13, gpu_ctx, n, form, gate):
    """The 2^22-table stress batch (uneven lifetimes, 10..2e5 steps): totals and
    the order-independent digest of every table's result equal the oracle's."""
    say(f"synth_run: {n} synthetic tables form={form} gate={gate}")
    gpu_ctx.set_tuning(form=form, gate=gate)
    try:
        sm = gpu_ctx.synth_run(0x5EED, 0, n)
    finally:
        gpu_ctx.reset_tuning()
    assert sm.line(0x5EED, 0) == synth_golden()[n]


def check_against(sums, golden, keys=("leaves", "survivors", "steps", "hist")):
    for X, g in golden.items():
        for k in keys:
            if g.get(k) is None:
                continue
            assert sums[X][k] == g[k], (X, k, sums[X][k], g[k])


@pytest.mark.parametrize("form,gate", [(3, 1), (3, 2), (2, 1), (1, 1)])
def test_local_edits_fmax2(e13, gpu_ctx, base, form, gate):
    say(f"search: delete FMAX=2 X=0..13 form={form} gate={gate}")
    gpu_ctx.set_tuning(form=form, gate=gate)
    try:
        sums, lines, hashes = search_all(e13, gpu_ctx, base, XS, fmax=2)
    finally:
        gpu_ctx.reset_tuning()
    check_against(sums, parse_summaries(read_text(REF_DUMP, "delete_f2.txt")))
    check_against(sums, parse_summaries(read_text(REF_RESULTS, "1_local_edits_log_f2.txt")))
    assert lines == survivor_lines(read_text(REF_RESULTS, "1_local_edits_surv_f2.txt"))
    ref = ref_hashes()
    for X, tab, ch in hashes:
        assert ref[(X, tab)] == ch


def search_many(e13, ctx, base, Xs, fmax):
    probs = [e13.setup(base, X, fmax) for X in Xs]
    cs, sv, ns = ctx.search_many(probs)
    assert ns == len(sv) == sum(c.survivors for c in cs)
    sums = {X: {"leaves": c.leaves, "survivors": c.survivors, "steps": c.steps, "hist": list(c.dead_at_stage)}
            for X, c in zip(Xs, cs)}
    lines, hashes = [], []
    for s in sv:
        assert s.recheck_ok == 1
        p = probs[s.prob]
        line = e13.format_survivor(p, s)
        lines.append(line)
        hashes.append((p.X, line.split("table=")[1].replace("?", ""), tuple(s.ch)))
    return sums, sorted(lines), hashes


def test_local_edits_fmax2_one_launch(e13, gpu_ctx, base):
    say("search_many: delete FMAX=2 X=0..13 in one persistent launch")
    sums, lines, hashes = search_many(e13, gpu_ctx, base, XS, 2)
    check_against(sums, parse_summaries(read_text(REF_DUMP, "delete_f2.txt")))
    assert lines == survivor_lines(read_text(REF_RESULTS, "1_local_edits_surv_f2.txt"))


@pytest.mark.parametrize("many", [True, False])
def test_local_edits_fmax3(e13, gpu_ctx, base, many):
    say(f"search: delete FMAX=3 X=0..13 one_launch={many}")
    if many:
        sums, lines, hashes = search_many(e13, gpu_ctx, base, XS, 3)
    else:
        sums, lines, hashes = search_all(e13, gpu_ctx, base, XS, fmax=3)
    check_against(sums, parse_summaries(read_text(REF_DUMP, "delete_f3.txt")))
    assert lines == survivor_lines(read_text(REF_RESULTS, "1_local_edits_surv_f3.txt"))
    ref = ref_hashes()
    for X, tab, ch in hashes:
        assert ref[(X, tab)] == ch


def test_local_edits_fmax4(e13, gpu_ctx, base):
    """5,986,617,022 leaves and 586,158,995,069 steps: per-X counts and
    histograms against the compiled reference's run (ref_dump/delete_f4.txt)
    and the reference's own log and survivor file."""
    say("search_many: delete FMAX=4 X=0..13 in one persistent launch")
    sums, lines, hashes = search_many(e13, gpu_ctx, base, XS, 4)
    check_against(sums, parse_summaries(read_text(REF_DUMP, "delete_f4.txt")))
    check_against(sums, parse_summaries(read_text(REF_RESULTS, "1_local_edits_log_f4.txt")))
    assert sum(v["steps"] for v in sums.values()) == 586158995069
    assert lines == survivor_lines(read_text(REF_RESULTS, "1_local_edits_surv_f4.txt"))
    ref = ref_hashes()
    for X, tab, ch in hashes:
        assert ref[(X, tab)] == ch


def test_split_tests(e13, gpu_ctx):
    split = e13.parse_machine(read_text(REF_RESULTS, "..", "machines", "split_15.txt"))
    say("search: split_15 delete X=14 FMAX=1")
    sums, lines, _ = search_all(e13, gpu_ctx, split, [14], fmax=1)
    assert sums[14] == {"leaves": 56, "survivors": 3, "steps": 632803, "hist": [0, 52, 1, 0, 0, 0, 0, 0, 0, 0]}
    assert lines == survivor_lines(read_text(REF_RESULTS, "validation_split_test.txt"))
    say("search: split_15 free X=14 K0 K1 L1")
    sums, lines, _ = search_all(e13, gpu_ctx, split, [14], free=["K0", "K1", "L1"])
    check_against(sums, parse_summaries(read_text(REF_RESULTS, "validation_split_test_free.txt")))
    assert lines == survivor_lines(read_text(REF_RESULTS, "validation_split_test_free.txt"))


@pytest.mark.parametrize("X,cells", [(0, "B0 B1 C0 C1"), (1, "A0 A1 C0 C1"), (2, "A0 A1 B0 B1")])
def test_doubling_free(e13, gpu_ctx, base, X, cells):
    say(f"search: doubling free X={X} {cells}")
    sums, lines, _ = search_all(e13, gpu_ctx, base, [X], free=cells.split())
    check_against(sums, parse_summaries(read_text(REF_RESULTS, f"3_doubling_2_dbl_X{X}.txt")))
    assert lines == []


def test_division_v4(e13, gpu_ctx, base):
    say("search: freesym X=9 3..9 F0..I1 (V4)")
    sums, lines, _ = search_all(e13, gpu_ctx, base, [9], free="F0 F1 G0 G1 H0 H1 I0 I1".split(), sym=(3, 9))
    check_against(sums, parse_summaries(read_text(REF_RESULTS, "4_division_6_v4_log_v4.txt")))
    check_against(sums, parse_summaries(read_text(REF_RESULTS, "4_division_6_v4_g_v4.txt")))
    assert lines == []


def test_counter_3state_symmetric(e13, gpu_ctx, base):
    say("search: freesym X=13 10..13 D0 G0 K0 K1 L1 M1 (3-state counter)")
    sums, lines, hashes = search_all(e13, gpu_ctx, base, [13], free="D0 G0 K0 K1 L1 M1".split(), sym=(10, 13))
    check_against(sums, parse_summaries(read_text(REF_RESULTS, "2_counter_3_log_modsym.txt")))
    check_against(sums, parse_summaries(read_text(REF_RESULTS, "2_counter_3_g_modsym.txt")))
    assert lines == survivor_lines(read_text(REF_RESULTS, "2_counter_3_surv_modsym.txt"))
    ref = ref_hashes()
    for X, tab, ch in hashes:
        assert ref[(X, tab)] == ch


@pytest.mark.parametrize("tune", [dict(form=2, slice=1, cap=7), dict(form=2, slice=5, cap=97, batch=3),
                                  dict(form=2, slice=200, cap=100000, batch=1 << 18),
                                  dict(form=2, wg_threads=64, wgs_per_cu=1, refill_tries=1),
                                  dict(form=1, cap=33, batch=17),
                                  dict(form=1, cap=33, batch=17, graph_chain=8),
                                  dict(form=2, slice=5, cap=97, batch=3, graph_chain=16),
                                  dict(form=2, graph_chain=4),
                                  dict(form=3, slice=1), dict(form=3, slice=7, refill_tries=1),
                                  dict(form=3, solo_steps=1), dict(form=3, solo_steps=-1),
                                  dict(form=3, evict_after=1, solo_steps=7), dict(form=3, evict_after=-1),
                                  dict(form=2, solo_steps=3),
                                  dict(form=3, slice=300, wg_threads=128, wgs_per_cu=3),
                                  # more workgroups per CU than can be resident, with a frontier
                                  # dealt to every wave: the grid is clamped to the resident ones
                                  dict(form=3, wgs_per_cu=4, seed_per_wave=8),
                                  dict(form=3, wg_threads=64, wgs_per_cu=1, stack_nodes=1 << 16),
                                  # rounds cut short once few lanes still run (the rest refilled)
                                  dict(form=3, cut_live=32), dict(form=3, cut_live=64, slice=256),
                                  dict(form=3, cut_live=2, slice=7, solo_steps=-1),
                                  # the stage-1 gate at every step, or only after a change of the ones count
                                  dict(form=3, nt1_skip=2), dict(form=3, nt1_skip=1, cut_live=-1, slice=3)])
def test_results_invariant_under_tuning(e13, gpu_ctx, base, tune):
    say(f"search: delete FMAX=2 X=7,8,12 tuning={tune}")
    gpu_ctx.set_tuning(**tune)
    try:
        sums, lines, hashes = search_all(e13, gpu_ctx, base, [7, 8, 12], fmax=2)
    finally:
        gpu_ctx.reset_tuning()
    golden = parse_summaries(read_text(REF_DUMP, "delete_f2.txt"))
    check_against(sums, {X: golden[X] for X in (7, 8, 12)})
    assert lines == survivor_lines(read_text(REF_RESULTS, "1_local_edits_surv_f2.txt"))


def test_autotune_keeps_results(e13, gpu_ctx, base):
    """The autotuner's trials all give the same counts (it checks each against
    the first) and the settings it installs still give the reference's."""
    say("autotune: delete FMAX=2 X=0..13")
    ps = [e13.setup(base, X, 2) for X in XS]
    try:
        best, rep = gpu_ctx.autotune(ps, reps=1)
        assert rep["trials"] >= 10 and rep["best_ms"] <= rep["first_ms"] * 1.0001
        cs, sv, ns = gpu_ctx.search_many(ps, surv_cap=4096)
    finally:
        gpu_ctx.reset_tuning()
    golden = parse_summaries(read_text(REF_DUMP, "delete_f2.txt"))
    for X, c in zip(XS, cs):
        d = c.as_dict()
        assert (d["leaves"], d["steps"], d["hist"]) == (golden[X]["leaves"], golden[X]["steps"], golden[X]["hist"])



def test_search_profiles(e13, base):
    """A context picks the stored small-search profile for searches of edit
    depth <= 3 and the main settings for deeper ones; settings installed with
    set_tuning then hold for every search.  Counts are the reference's under
    either profile."""
    import os
    say("profiles: delete FMAX=3 X=0,1 (small) and freesym X=13 (main) on a fresh context")
    small = [e13.setup(base, X, 3) for X in (0, 1)]
    large = [e13.setup(base, 13, 0, "D0 G0 K0 K1 L1 M1".split(), (10, 13))]
    with e13.Context(0) as ctx:
        path = ctx.tuning_file(small=True)
        assert path.endswith("_small.txt") and path.replace("_small.txt", ".txt") == ctx.tuning_file()
        has_small = os.path.exists(path)
        prof, t = ctx.search_profile(small)
        assert prof == (1 if has_small else 0)
        if has_small:
            stored = e13.load_tuning(path)
            for k, v in stored.items():
                if v > 0:
                    assert t[k] == v, k
        assert ctx.search_profile(large)[0] == 0
        cs, _, _ = ctx.search_many(small)
        assert ctx.stats()["profile"] == prof
        golden = parse_summaries(read_text(REF_DUMP, "delete_f3.txt"))
        for X, c in zip((0, 1), cs):
            d = c.as_dict()
            assert (d["leaves"], d["steps"], d["hist"]) == (golden[X]["leaves"], golden[X]["steps"], golden[X]["hist"])
        ctx.set_tuning(slice=ctx.tuning()["slice"])
        assert ctx.search_profile(small)[0] == 0
        ctx.search_many(small)
        assert ctx.stats()["profile"] == 0
