"""GPU searches against reference fixtures that the parity suite's shared
context does not cover: fresh contexts with the stored settings profiles
(including searches whose breadth-first frontier finishes the whole tree),
free-mode searches with no deleted state (X = -1), the raw 3-state counter
search without symmetry, the division search V3, and FMAX=3 in the plain
per-leaf-launch form, and the persistent search on a short lane tape.  Exact
comparisons throughout."""
import pytest

from conftest import REF_DUMP, REF_RESULTS, XS, parse_summaries, read_machine, read_text, ref_hashes, survivor_lines

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def base(e13):
    return e13.parse_machine(read_machine())


def run_one(e13, ctx, p):
    c, sv, ns = ctx.search(p)
    assert ns == len(sv) == c.survivors
    lines, hashes = [], []
    for s in sv:
        assert s.recheck_ok == 1
        line = e13.format_survivor(p, s)
        lines.append(line)
        hashes.append((p.X, line.split("table=")[1].replace("?", ""), tuple(s.ch)))
    d = c.as_dict()
    return {"leaves": d["leaves"], "survivors": d["survivors"], "steps": d["steps"], "hist": d["hist"]}, \
        sorted(lines), hashes


def same(got, want):
    for k in ("leaves", "survivors", "steps", "hist"):
        if want.get(k) is not None:
            assert got[k] == want[k], (k, got[k], want[k])


@pytest.mark.parametrize("X", [2, 13, 8])
def test_frontier_finishes_the_tree(e13, base, X):
    """FMAX=2 searches of one deleted state each on a fresh context, in the
    persistent form with a breadth-first frontier of 2 nodes per wave: for X=2
    and X=13 the frontier expansion ends every machine of the tree, so no
    persistent launch is needed (one with no work would never end); X=8 has
    work left for the launch."""
    golden = parse_summaries(read_text(REF_DUMP, "delete_f2.txt"))
    with e13.Context(0) as ctx:
        say(f"search: delete FMAX=2 X={X} alone, form 3, seed_per_wave=2, fresh context")
        ctx.set_tuning(form=3, seed_per_wave=2)
        got, lines, _ = run_one(e13, ctx, e13.setup(base, X, 2))
        same(got, golden[X])
        want = [l for l in survivor_lines(read_text(REF_RESULTS, "1_local_edits_surv_f2.txt")) if f"X={X} " in l]
        assert lines == want


def test_stored_profiles_fresh_context(e13, base):
    """The settings a context opens with (the stored main and small-search
    profiles, chosen per search): every X of FMAX=2 and FMAX=3 one at a time,
    then the symmetric counter search."""
    f2 = parse_summaries(read_text(REF_DUMP, "delete_f2.txt"))
    f3 = parse_summaries(read_text(REF_DUMP, "delete_f3.txt"))
    ref = ref_hashes()
    with e13.Context(0) as ctx:
        for fmax, golden, surv in ((2, f2, "1_local_edits_surv_f2.txt"), (3, f3, "1_local_edits_surv_f3.txt")):
            lines = []
            for X in XS:
                say(f"search: delete FMAX={fmax} X={X}, stored profiles, fresh context")
                got, ls, hs = run_one(e13, ctx, e13.setup(base, X, fmax))
                same(got, golden[X])
                lines += ls
                for x, tab, ch in hs:
                    assert ref[(x, tab)] == ch
            assert sorted(lines) == survivor_lines(read_text(REF_RESULTS, surv))
        say("search: freesym X=13 10..13 D0 G0 K0 K1 L1 M1, stored profiles, fresh context")
        got, lines, _ = run_one(e13, ctx, e13.setup(base, 13, 0, "D0 G0 K0 K1 L1 M1".split(), (10, 13)))
        same(got, parse_summaries(read_text(REF_RESULTS, "2_counter_3_log_modsym.txt"))[13])
        assert lines == survivor_lines(read_text(REF_RESULTS, "2_counter_3_surv_modsym.txt"))


def freesym_xm1_cases():
    cases, cur = [], None
    for line in read_text(REF_DUMP, "freesym_xm1.txt").splitlines():
        if line.startswith("case "):
            cur = {"args": line.split()[2:], "surv": [], "hash": []}
            cases.append(cur)
        elif line.startswith("SURVIVOR"):
            cur["surv"].append(line.strip())
        elif line.startswith("HASH"):
            cur["hash"].append(tuple(int(f.split("=")[1], 16) for f in line.split()[2:5]))
        elif "leaves=" in line:
            cur["sum"] = parse_summaries(line)[-1]
    return cases


@pytest.mark.parametrize("case", freesym_xm1_cases(), ids=lambda c: "_".join(c["args"]))
def test_freesym_no_deleted_state(e13, gpu_ctx, base, case):
    """freesym with X = -1 (no deleted state): symmetry from the shifted LO..HI,
    checked over 2*(NQ-1) cells, or off when LO lands below 0 -- the compiled
    reference's counts, survivors and counter hashes."""
    a = case["args"]
    say(f"search: freesym {' '.join(a)}")
    p = e13.setup(base, int(a[0]), 0, a[3:], (int(a[1]), int(a[2])))
    got, lines, hashes = run_one(e13, gpu_ctx, p)
    same(got, case["sum"])
    assert lines == sorted(case["surv"])
    # the fixture's HASH lines follow its SURVIVOR lines one for one
    want = dict(zip(case["surv"], case["hash"]))
    got_h = {tab: ch for _, tab, ch in hashes}
    for line, ch in want.items():
        assert got_h[line.split("table=")[1].replace("?", "")] == ch


def test_counter_3state_raw(e13, gpu_ctx, base):
    """The 3-state counter search without symmetry: 18,471,584,227 leaves and
    366 = 61 * 3! survivors (results/2_counter_3/log_mod.txt, surv_mod.txt),
    the symmetric search's leaves times the relabellings it removes."""
    say("search: free X=13 D0 G0 K0 K1 L1 M1 (raw 3-state counter)")
    p = e13.setup(base, 13, 0, "D0 G0 K0 K1 L1 M1".split())
    got, lines, hashes = run_one(e13, gpu_ctx, p)
    same(got, parse_summaries(read_text(REF_RESULTS, "2_counter_3_log_mod.txt"))[13])
    assert lines == survivor_lines(read_text(REF_RESULTS, "2_counter_3_surv_mod.txt"))
    ref = ref_hashes()
    for X, tab, ch in hashes:
        assert ref[(X, tab)] == ch
    say(f"raw counter: {got['steps']} steps")


def test_division_v3(e13, gpu_ctx, base):
    """Division search V3: freesym X=9 3..9 E0..I1, 53,282,072,641 leaves and
    5,803,777,946,442 steps (results/4_division_6/v3/log_v3.txt, g_v3.log)."""
    say("search: freesym X=9 3..9 E0 E1 F0 F1 G0 G1 H0 H1 I0 I1 (V3)")
    p = e13.setup(base, 9, 0, "E0 E1 F0 F1 G0 G1 H0 H1 I0 I1".split(), (3, 9))
    got, lines, _ = run_one(e13, gpu_ctx, p)
    same(got, parse_summaries(read_text(REF_RESULTS, "4_division_6_v3_log_v3.txt"))[9])
    same(got, parse_summaries(read_text(REF_RESULTS, "4_division_6_v3_g_v3.txt"))[9])
    assert lines == []


def test_local_edits_fmax3_plain_form(e13, gpu_ctx, base):
    """FMAX=3 in the plain form (one machine per lane, one launch per batch of
    nodes, host loop)."""
    f3 = parse_summaries(read_text(REF_DUMP, "delete_f3.txt"))
    gpu_ctx.set_tuning(form=1)
    lines = []
    try:
        for X in XS:
            say(f"search: delete FMAX=3 X={X} form=1")
            got, ls, _ = run_one(e13, gpu_ctx, e13.setup(base, X, 3))
            same(got, f3[X])
            lines += ls
    finally:
        gpu_ctx.reset_tuning()
    assert sorted(lines) == survivor_lines(read_text(REF_RESULTS, "1_local_edits_surv_f3.txt"))


def check_many(e13, ctx, problems, golden, surv_file):
    cs, sv, ns = ctx.search_many([p for _, p in problems], surv_cap=4096)
    assert ns == len(sv)
    for (X, _), c in zip(problems, cs):
        d = c.as_dict()
        same({"leaves": d["leaves"], "survivors": d["survivors"], "steps": d["steps"], "hist": d["hist"]}, golden[X])
    lines = []
    for s in sv:
        assert s.recheck_ok == 1
        lines.append(e13.format_survivor(problems[s.prob][1], s))
    assert sorted(lines) == survivor_lines(read_text(REF_RESULTS, surv_file))


@pytest.mark.parametrize("tune", [dict(wgs_per_cu=2, tape_words=20), dict(wgs_per_cu=3, tape_words=18),
                                  dict(wgs_per_cu=3, tape_words=-1), dict(wgs_per_cu=4, tape_words=-1),
                                  dict(wgs_per_cu=1, tape_words=24, slice=64)],
                         ids=lambda t: "-".join(f"{k}{v}" for k, v in t.items()))
def test_short_lane_tape(e13, base, tune):
    """The persistent search on a short lane tape in LDS (more workgroups per
    CU): machines whose head goes past it are paused and finish in a second
    launch with the whole tape; FMAX=3 over X=0..13 gives the reference's
    per-X counts and survivors, and the second launch did run."""
    f3 = parse_summaries(read_text(REF_DUMP, "delete_f3.txt"))
    with e13.Context(0) as ctx:
        say(f"search: delete FMAX=3 X=0..13, one persistent search, {tune}, fresh context")
        ctx.set_tuning(form=3, **tune)
        check_many(e13, ctx, [(X, e13.setup(base, X, 3)) for X in XS], f3, "1_local_edits_surv_f3.txt")
        st = ctx.stats()
        assert 18 <= st["tape_words"] < 43
        assert st["wide_nodes"] > 0 and st["wide_launches"] >= 1


@pytest.mark.parametrize("cap", [1, 5])
def test_short_lane_tape_full_list(e13, base, cap):
    """A list of machines past the short tape that fills at once: the launch
    stops cleanly, and the rest of the search runs on the whole tape; FMAX=3
    over X=0..13 still gives the reference's counts and survivors."""
    f3 = parse_summaries(read_text(REF_DUMP, "delete_f3.txt"))
    with e13.Context(0) as ctx:
        say(f"search: delete FMAX=3 X=0..13, tape_words=20, list of {cap} wide nodes, fresh context")
        ctx.set_tuning(form=3, wgs_per_cu=3, tape_words=20, wide_nodes=cap)
        check_many(e13, ctx, [(X, e13.setup(base, X, 3)) for X in XS], f3, "1_local_edits_surv_f3.txt")
        st = ctx.stats()
        assert st["tape_words"] == 20 and st["wide_launches"] >= 1


def test_short_lane_tape_checkpoints(e13, base, tmp_path):
    """Clean stops and resumes on a short lane tape: the counter search with a
    checkpoint every 50 ms on a 20-word lane tape (machines past it wait in the
    saved nodes and finish on the whole tape) gives the reference's counts and
    61 survivors."""
    want = parse_summaries(read_text(REF_RESULTS, "2_counter_3_log_modsym.txt"))[13]
    p = e13.setup(base, 13, 0, "D0 G0 K0 K1 L1 M1".split(), (10, 13))
    with e13.Context(0) as ctx:
        say("search: freesym X=13 10..13 D0 G0 K0 K1 L1 M1, tape_words=20, checkpoint every 50 ms")
        ctx.set_tuning(form=3, wgs_per_cu=3, tape_words=20)
        ctx.set_checkpoint(str(tmp_path / "short"), every_sec=0.05)
        got, lines, _ = run_one(e13, ctx, p)
        ctx.set_checkpoint(None)
        same(got, want)
        assert lines == survivor_lines(read_text(REF_RESULTS, "2_counter_3_surv_modsym.txt"))
        st = ctx.stats()
        assert st["search_launches"] >= 3 and st["wide_nodes"] > 0


def test_phase_clocks(e13, base):
    """The short-tape search built with the per-phase wave clock reads (the
    kernel the bench's untimed pass runs) gives the reference's counts and
    survivors for the counter search on the stored main profile, and counts
    wave cycles in every phase."""
    want = parse_summaries(read_text(REF_RESULTS, "2_counter_3_log_modsym.txt"))[13]
    p = e13.setup(base, 13, 0, "D0 G0 K0 K1 L1 M1".split(), (10, 13))
    with e13.Context(0) as ctx:
        _, tuning = ctx.search_profile([p])
        ctx.restore_tuning(dict(tuning, phase_cycles=1, trace_events=4096))
        say("search: freesym X=13 10..13 D0 G0 K0 K1 L1 M1, main profile, per-phase clocks and event trace")
        got, lines, _ = run_one(e13, ctx, p)
        same(got, want)
        assert lines == survivor_lines(read_text(REF_RESULTS, "2_counter_3_surv_modsym.txt"))
        st = ctx.stats()
        assert st["tape_words"] < 43
        assert all(c > 0 for c in st["cycles"]), st["cycles"]
        # the event trace of the short-tape launch: every wave logged events and,
        # where its log did not fill, ended with its exit event (5)
        import ctypes
        L = e13.lib()
        waves, per = ctypes.c_int32(), ctypes.c_int32()
        assert L.e13_get_trace(ctx._h, None, 0, ctypes.byref(waves), ctypes.byref(per)) == 0
        assert waves.value > 0 and per.value == 4097
        buf = (ctypes.c_uint64 * (waves.value * per.value))()
        assert L.e13_get_trace(ctx._h, buf, len(buf), ctypes.byref(waves), ctypes.byref(per)) == 0
        for w in range(waves.value):
            n = buf[w * per.value]
            assert n >= 1, w
            if n < per.value - 1:
                assert buf[w * per.value + n] & 0xFF == 5, (w, n)


def test_counter_family_sample(e13, gpu_ctx, base):
    """The counter search's family on the bench's CPU sample: freesym X=13
    with states 11..13 symmetric and free cells D0 G0 L1 M1 -- the compiled
    reference's counts (96,010,408 leaves, 10,716,181,112 steps) and its 19
    survivor lines (tests/golden/ref_dump/freesym_x13_s11_D0G0L1M1.txt)."""
    text = read_text(REF_DUMP, "freesym_x13_s11_D0G0L1M1.txt")
    say("search: freesym X=13 11..13 D0 G0 L1 M1")
    p = e13.setup(base, 13, 0, "D0 G0 L1 M1".split(), (11, 13))
    got, lines, _ = run_one(e13, gpu_ctx, p)
    same(got, parse_summaries(text)[13])
    assert lines == sorted(l for l in text.splitlines() if l.startswith("SURVIVOR"))
