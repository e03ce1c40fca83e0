"""The CPU oracle (oracle/e13oracle*.c) against what pins it.

Pins, by oracle function:
  eo_init_keys / splitmix   KEY/TGT dump of the compiled reference
                            (ref_dump/keys_tgt.txt) and the closed form
                            TGT[n] = XOR of KEY over cells -1..2^n-2.
  eo_run / stage rule       validation/base_erdos14.txt (208,864 steps),
                            validation/mutants.txt (1568 known answers), and a
                            brute-force set simulator below (exact block and
                            counter sets, no hashing) on random tables.
  eo_setup / eo_search      per-X leaves, histograms and survivors of
                            1_local_edits/log_f2.txt, log_f3.txt; step totals
                            from the compiled reference (ref_dump/delete_f*.txt);
                            split tests; the doubling logs; survivor lines.
  counter hashes ch[]       ref_dump/survivor_hashes.txt (the reference's own
                            run() on every fixture survivor table).
  eo_recheck                every fixture survivor line (502, including the raw
                            counter search's 366) re-checks with the '?' marks of
                            its line equal to the unread cells; claims one step,
                            state, head or start kind off are refused; and on
                            random edits of erdos_14 it accepts exactly the
                            tables eo_run_full calls survivors, with their
                            stage-1 state, head and step count.
  eo_synth_table            the brute-force simulator on the same tables.
  eo_probe / eo_estimate    per-probe sums of the reference's own probe()
                            (ref_dump/probe_*.txt, one stream per probe), the
                            exact rows of 4_division_6/v1/v1c0_est.txt (values
                            of C0 whose tree is one leaf), and the exact search
                            of a small free-mode tree, which the mean of many
                            probes must estimate without bias.
"""
import ctypes
import random

import pytest

from conftest import (MACHINES, PROBE_CASES, REF_DUMP, REF_RESULTS, XS, hash_lines_by_survivor, machine_path,
                      oracle_lib, oracle_probe_problem, oracle_probes, parse_summaries, probe_fixture, read_text,
                      ref_hashes, run_oracle, survivor_lines, v1c0_exact_rows)

T14S = (0, 2, 22, 90, 268, 944, 3434, 13450, 52068, 208864)


def ref_keys():
    vals = [int(l, 16) for l in read_text(REF_DUMP, "keys_tgt.txt").split()]
    return vals[:1364], [0] + vals[1364:]


def test_keys_match_reference_dump():
    L = oracle_lib()
    key = list((ctypes.c_uint64 * 1364).in_dll(L, "eo_key"))
    tgt = list((ctypes.c_uint64 * 10).in_dll(L, "eo_tgt"))
    rk, rt = ref_keys()
    assert key == rk
    assert tgt[1:] == rt[1:]


def test_targets_closed_form():
    rk, rt = ref_keys()
    for n in range(1, 10):
        h = 0
        for cell in range(-1, 2 ** n - 1):
            h ^= rk[cell + 64]
        assert h == rt[n]


def test_splitmix_first_outputs():
    # splitmix64 from seed 20261028, written out from the published algorithm
    M = (1 << 64) - 1
    s = 20261028
    out = []
    for _ in range(4):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        out.append(z ^ (z >> 31))
    assert out == ref_keys()[0][:4]


def test_base_machine():
    out, _ = run_oracle("base", machine_path())
    assert out.strip() == read_text(REF_RESULTS, "validation_base_erdos14.txt").splitlines()[1].strip()


def test_mutants():
    out, _ = run_oracle("mutant", machine_path())
    want = [l for l in read_text(REF_RESULTS, "validation_mutants.txt").splitlines() if not l.startswith("#")]
    assert out.splitlines() == want


def test_split_delete():
    out, err = run_oracle("delete", machine_path("split_15.txt"), 1, 14)
    s = parse_summaries(err)[14]
    assert (s["leaves"], s["survivors"], s["steps"]) == (56, 3, 632803)
    assert s["hist"] == [0, 52, 1, 0, 0, 0, 0, 0, 0, 0]
    assert survivor_lines(out) == survivor_lines(read_text(REF_RESULTS, "validation_split_test.txt"))


def test_split_free():
    out, err = run_oracle("free", machine_path("split_15.txt"), 14, "K0", "K1", "L1")
    s = parse_summaries(err)[14]
    want = parse_summaries(read_text(REF_RESULTS, "validation_split_test_free.txt"))[14]
    assert s == want
    assert survivor_lines(out) == survivor_lines(read_text(REF_RESULTS, "validation_split_test_free.txt"))


@pytest.mark.parametrize("X,cells", [(0, "B0 B1 C0 C1"), (1, "A0 A1 C0 C1"), (2, "A0 A1 B0 B1")])
def test_doubling_free(X, cells):
    _, err = run_oracle("free", machine_path(), X, *cells.split())
    want = parse_summaries(read_text(REF_RESULTS, f"3_doubling_2_dbl_X{X}.txt"))[X]
    assert parse_summaries(err)[X] == want


def test_local_edits_fmax2():
    out, err = run_oracle("delete", machine_path(), 2, *XS, threads=4)
    got = parse_summaries(err)
    old = parse_summaries(read_text(REF_RESULTS, "1_local_edits_log_f2.txt"))
    dump = parse_summaries(read_text(REF_DUMP, "delete_f2.txt"))
    for X in XS:
        assert got[X]["leaves"] == old[X]["leaves"] and got[X]["hist"] == old[X]["hist"]
        assert got[X] == dump[X]
    assert survivor_lines(out) == survivor_lines(read_text(REF_RESULTS, "1_local_edits_surv_f2.txt"))


@pytest.mark.parametrize("X", [2, 3, 4, 9, 10, 12, 13])
def test_local_edits_fmax3_subset(X):
    """FMAX=3 for the deleted states whose trees are small (the others are
    covered on the GPU against the same fixtures)."""
    out, err = run_oracle("delete", machine_path(), 3, X, threads=4)
    got = parse_summaries(err)[X]
    assert got == parse_summaries(read_text(REF_DUMP, "delete_f3.txt"))[X]
    want = [l for l in survivor_lines(read_text(REF_RESULTS, "1_local_edits_surv_f3.txt")) if f"X={X} " in l]
    assert survivor_lines(out) == want


def test_survivor_hashes_match_reference():
    out, _ = run_oracle("delete", machine_path(), 2, 8)
    out3, _ = run_oracle("delete", machine_path(), 3, 12)
    ref = ref_hashes()
    n = 0
    for line, h in hash_lines_by_survivor(out + out3):
        X = int(line.split("X=")[1].split()[0])
        tab = line.split("table=")[1].replace("?", "")
        chs = tuple(int(f.split("=")[1], 16) for f in h.split()[2:5])
        assert ref[(X, tab)] == chs
        n += 1
    assert n == 7


# ---------------------------------------------------------------------------
# Brute-force simulator: tape as a set of one-cells, block and counter compared
# as exact sets (no hashing), stage rule as stated in the reference's header
# comment (e13search.c:13-23) and README "Filter".
# ---------------------------------------------------------------------------
def erdos14():
    text = read_text(MACHINES, "erdos_14.txt").split()
    return [-1 if c[0] == "-" else (int(c[2:]) << 2) | ((c[1] == "R") << 1) | int(c[0]) for c in text]


def brute_run(tab, deadlines, mistake=None):
    ones = set()
    head = st = t = 0
    nt, stage, st0, head0 = 1, 0, -1, 0
    abs_ok = rel_ok = 1
    counters = {}
    nblock = 0
    while True:
        if nblock == 2 ** nt and all(c in ones for c in range(-1, 2 ** nt - 1)):
            ok, a, r = True, abs_ok, rel_ok
            if nt > 1:
                a = a and head == head0
                r = r and head - 2 ** nt == head0 - 2
                ok = st == st0 and (a or r)
            if ok:
                cls = 0 if nt <= 2 else (1 if nt <= 8 else 2)
                cnt = frozenset(c for c in ones if c <= -3)
                if cls in counters and counters[cls] != cnt:
                    return -1, stage, t
                if mistake != "distinct" and any(j != cls and v == cnt for j, v in counters.items()):
                    return -1, stage, t
                counters[cls] = cnt
                if nt == 1:
                    st0, head0 = st, head
                abs_ok, rel_ok, stage = int(a), int(r), nt
                if nt == 9:
                    return -2, stage, t
                nt += 1
        if (t >= deadlines[nt]) if mistake == "deadline" else (t > deadlines[nt]):
            return -1, stage, t
        sym = 1 if head in ones else 0
        code = tab[2 * st + sym]
        if code < 0:
            return -1, stage, t
        w = code & 1
        if w != sym:
            if w:
                ones.add(head)
            else:
                ones.discard(head)
            if head >= -2:
                nblock += 1 if w else -1
        head += 1 if code & 2 else -1
        if head <= (-63 if mistake == "window" else -64) or head >= 1299:
            return -1, stage, t
        st = code >> 2
        t += 1


def oracle_run(L, tab, deadlines):
    class Sim(ctypes.Structure):
        _fields_ = [("head", ctypes.c_int), ("st", ctypes.c_int), ("nt", ctypes.c_int), ("stage", ctypes.c_int),
                    ("st0", ctypes.c_int), ("head0", ctypes.c_int), ("abs_ok", ctypes.c_int),
                    ("rel_ok", ctypes.c_int), ("cs", ctypes.c_int * 3), ("t", ctypes.c_longlong),
                    ("deadline", ctypes.c_longlong), ("hb", ctypes.c_uint64), ("hc", ctypes.c_uint64),
                    ("ch", ctypes.c_uint64 * 3), ("tape", ctypes.c_uint8 * 1364)]
    s = Sim()
    t32 = (ctypes.c_int * 32)(*(list(tab) + [-1] * (32 - len(tab))))
    dl = (ctypes.c_int32 * 10)(*deadlines)
    r = L.eo_run_full(t32, ctypes.byref(s), dl)
    return r, s.stage, s.t


def test_run_matches_brute_force_on_random_edits():
    """Random 1..3-cell edits of erdos_14 with the reference deadlines, cut to
    tables whose runs stay short enough for the set simulator."""
    L = oracle_lib()
    conv = erdos14()
    dl = [0] + [4 * T14S[n] + 64 for n in range(1, 10)]
    rng = random.Random(7)
    checked = 0
    for _ in range(400):
        tab = list(conv)
        for _ in range(rng.randint(1, 3)):
            tab[rng.randrange(28)] = rng.randrange(-1, 56)
        r = oracle_run(L, tab, dl)
        if r[2] > 20000:
            continue
        assert brute_run(tab, dl) == r, tab
        checked += 1
    assert checked > 300


def test_run_matches_brute_force_on_synthetic_tables():
    """The synthetic stress tables: cap replaces every stage deadline."""
    L = oracle_lib()
    L.eo_synth_table.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_int32)]
    checked = 0
    for k in range(600):
        tab = (ctypes.c_int * 32)()
        cap = ctypes.c_int32()
        L.eo_synth_table(0x5EED, k, tab, ctypes.byref(cap))
        if cap.value > 20000:
            continue
        dl = [cap.value] * 10
        tb = list(tab[:28])
        assert brute_run(tb, dl) == oracle_run(L, tb, dl), k
        checked += 1
    assert checked > 400


@pytest.mark.parametrize("mistake", ["deadline", "distinct", "window"])
def test_planted_mistake_is_caught(mistake):
    """A mistake planted in a copy of the rule (deadline off by one, counters of
    different classes allowed to be equal, window one cell narrower) makes the
    same comparison against the oracle fail on the same random tables: the
    check above has the power to see such slips."""
    L = oracle_lib()
    conv = erdos14()
    dl = [0] + [4 * T14S[n] + 64 for n in range(1, 10)]
    rng = random.Random(7)
    differ = 0
    for _ in range(400):
        tab = list(conv)
        for _ in range(rng.randint(1, 3)):
            tab[rng.randrange(28)] = rng.randrange(-1, 56)
        r = oracle_run(L, tab, dl)
        if r[2] > 20000:
            continue
        differ += brute_run(tab, dl, mistake) != r
    assert differ > 0


@pytest.mark.parametrize("name", PROBE_CASES)
def test_probe_matches_reference_probe(name):
    """eo_probe gives the same four doubles, bit for bit, as the reference's own
    probe() on the same stream, probe by probe."""
    args, rows = probe_fixture(name)
    rc, p = oracle_probe_problem(args)
    assert rc == 0
    got = oracle_probes(oracle_lib(), p, args["seed"], args["n"])
    assert got == rows
    assert len(set(rows)) > 5  # the probes take different paths


def test_estimate_exact_rows_of_v1c0():
    """Values of C0 whose whole tree is one leaf: every probe sees the same leaf,
    so the estimate is exact and equals the reference's printed row."""
    rows = v1c0_exact_rows()
    assert len(rows) >= 20
    for v, path, leaves, steps in rows[:8]:
        out, _ = run_oracle("estimate", path, "200", "9", "0", "3", "9",
                            *"D0 D1 E0 E1 F0 F1 G0 G1 H0 H1 I0 I1".split())
        got = {l.split()[0]: float(l.split()[1]) for l in out.splitlines()[1:]}
        assert got["leaves"] == leaves and got["steps"] == steps, v


def test_estimate_mean_is_unbiased_on_a_small_tree():
    """The small free-mode tree (erdos_14, cells M1 N0 free): the mean over many
    probes lands within 5 standard errors of the exact search's leaves, steps
    and survivors."""
    _, err = run_oracle("free", machine_path(), "-1", "M1", "N0")
    exact = parse_summaries(err)[-1]
    out, _ = run_oracle("estimate", machine_path(), "200000", "-1", "0", "-1", "-1", "M1", "N0")
    est = {l.split()[0]: (float(l.split()[1]), float(l.split()[3])) for l in out.splitlines()[1:]}
    for key, ex in (("leaves", exact["leaves"]), ("steps", exact["steps"]), ("survivors", exact["survivors"])):
        m, se = est[key]
        assert se > 0 and abs(m - ex) < 5 * se, (key, m, se, ex)


# ---------------------------------------------------------------------------
# eo_recheck: the plain unhashed re-check of a survivor claim
# ---------------------------------------------------------------------------
def recheck_fixture_lines():
    out = []
    for f in ("1_local_edits_surv_f2.txt", "1_local_edits_surv_f3.txt", "1_local_edits_surv_f4.txt",
              "2_counter_3_surv_modsym.txt", "2_counter_3_surv_mod.txt", "validation_split_test.txt",
              "validation_split_test_free.txt"):
        out += [l.strip() for l in read_text(REF_RESULTS, f).splitlines() if l.startswith("SURVIVOR")]
    return out


def parse_claim(line):
    f = dict(kv.split("=", 1) for kv in line.split(" table=")[0].split()[1:])
    cells = []
    for row in line.split("table=")[1].split("|"):
        cells += row.split()
    tab, unread = [], 0
    for i, c in enumerate(cells):
        if c.startswith("?"):
            unread |= 1 << i
            c = c[1:]
        tab.append(-1 if c == "---" else (int(c[2:]) << 2) | ((c[1] == "R") << 1) | int(c[0]))
    claim = (int(f["st0"]), int(f["head0"]), int(f["abs"]), int(f["rel"]), int(f["steps"]))
    return tab, unread, claim


def eo_recheck(L, tab, claim):
    L.eo_recheck.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             ctypes.c_int, ctypes.c_longlong, ctypes.POINTER(ctypes.c_uint32)]
    t32 = (ctypes.c_int * 32)(*(list(tab) + [-1] * (32 - len(tab))))
    m = ctypes.c_uint32()
    ok = L.eo_recheck(t32, len(tab) // 2, *claim, ctypes.byref(m))
    return ok, m.value


def test_recheck_accepts_every_fixture_survivor():
    L = oracle_lib()
    lines = recheck_fixture_lines()
    assert len(lines) >= 500
    for line in lines:
        tab, unread, claim = parse_claim(line)
        ok, mask = eo_recheck(L, tab, claim)
        assert ok == 1, line
        assert mask == ((1 << len(tab)) - 1) & ~unread, line


def test_recheck_refuses_wrong_claims():
    """A claim one step long or short, with another stage-1 state or head, or
    with the other start kind (abs <-> rel) where only one holds, is refused."""
    L = oracle_lib()
    refused = 0
    for line in recheck_fixture_lines()[::7]:
        tab, _, (st0, head0, a, r, steps) = parse_claim(line)
        nq = len(tab) // 2
        wrong = [(st0, head0, a, r, steps + 1), (st0, head0, a, r, steps - 1), ((st0 + 1) % nq, head0, a, r, steps),
                 (st0, head0 + 1, a, r, steps), (st0, head0 - 1, a, r, steps)]
        if a != r:
            wrong.append((st0, head0, r, a, steps))
        for w in wrong:
            assert eo_recheck(L, tab, w)[0] == 0, (line, w)
            refused += 1
    assert refused > 350


def test_recheck_agrees_with_run_on_random_edits():
    """Random 1..3-cell edits of erdos_14 plus the cells that make fixture
    survivors: eo_recheck accepts the claim that eo_run_full's final state
    makes exactly when the run ends as a survivor."""
    from conftest import oracle_full
    L = oracle_lib()
    conv = erdos14()
    rng = random.Random(11)
    tabs = []
    for _ in range(300):
        tab = list(conv)
        for _ in range(rng.randint(1, 3)):
            tab[rng.randrange(28)] = rng.randrange(-1, 56)
        tabs.append(tab)
    tabs += [parse_claim(l)[0] for l in recheck_fixture_lines()]
    surv = dead = 0
    for tab in tabs:
        r, stage, t, st0, head0, a, rl, _ = oracle_full(L, tab, None)
        if stage < 1:
            continue
        ok, _ = eo_recheck(L, tab, (st0, head0, a, rl, t))
        assert ok == (1 if r == -2 else 0), (tab, r, stage)
        surv += r == -2
        dead += r != -2
    assert surv >= 500 and dead >= 50
