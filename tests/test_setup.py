"""setup() + set_sym() (e13search.c:315-355, 422-426) in the oracle
(eo_setup) and in the library (e13_setup) against the reference's own
functions: ref_dump/setup_sweep.txt holds what the compiled reference's
set_sym() + setup() leave (budget, SYM_LO, SYM_HI, SYM_N0, table codes) or
the message it exits with, over X = -1..13, edit budgets, symmetric ranges
and free-cell sets on erdos_14 (tests/golden/make_setup.sh).  The sweep
includes the X = -1 cases where the reference scans only 2*(NQ-1) cells for
fixed cells naming a symmetric state and where a LO shifted below 0 turns
symmetry off."""
import ctypes

import pytest

from conftest import REF_DUMP, machine_path, oracle_setup, parse_summaries, read_machine, read_text, run_oracle,\
    survivor_lines

E13_EFORCED, E13_ESYM, E13_ESYMSTART = -4, -5, -6


def sweep():
    cases = []
    for line in read_text(REF_DUMP, "setup_sweep.txt").splitlines():
        args, res = line.split(" => ")
        a = args.split()
        X, fmax, lo, hi = (int(v) for v in a[:4])
        cells = a[4:]
        if res.startswith("exit1"):
            want = ("sym_start" if "start state" in res else "sym",)
        else:
            f = dict(kv.split("=", 1) for kv in res.split()[1:])
            want = (int(f["budget"]), int(f["sym_lo"]), int(f["sym_hi"]), int(f["sym_n0"]),
                    [int(v) for v in f["tab"].split(",")])
        cases.append((X, fmax, lo, hi, cells, want))
    return cases


CASES = sweep()


def base_codes(e13):
    return e13.parse_machine(read_machine())


def test_sweep_covers_the_edges():
    kinds = {c[5][0] if isinstance(c[5][0], str) else ("off" if c[5][1] < 0 else "on") for c in CASES}
    assert kinds == {"sym", "sym_start", "off", "on"}
    # X = -1 runs with symmetry where a scan of the whole table would refuse it
    assert any(c[0] == -1 and c[4] and len(c[5]) == 5 and c[5][1] >= 0 for c in CASES)
    # X = -1 with LO = 0: the reference runs it with symmetry off
    assert any(c[0] == -1 and c[2] == 0 and len(c[5]) == 5 and c[5][1] == -1 for c in CASES)
    assert len(CASES) > 600


def _oracle(base, X, fmax, lo, hi, cells):
    rc, p = oracle_setup(base, X, fmax, cells or None, (lo, hi) if lo >= 0 else None)
    if rc == -2:
        return ("sym",)
    if rc == -3:
        return ("sym_start",)
    return (p.budget, p.sym_lo, p.sym_hi if p.sym_lo >= 0 else None, p.sym_n0, list(p.tab[:2 * p.nq]))


def _library(e13, base, X, fmax, lo, hi, cells):
    b = (ctypes.c_int32 * 32)(*base)
    fb = None
    if cells:
        fb = (ctypes.c_uint8 * 32)()
        for c in cells:
            fb[e13.cell_index(c)] = 1
    p = e13.Problem()
    rc = e13.lib().e13_setup(b, len(base) // 2, X, fmax, fb, lo, hi, ctypes.byref(p))
    if rc == E13_ESYM:
        return ("sym",)
    if rc == E13_ESYMSTART:
        return ("sym_start",)
    assert rc in (0, E13_EFORCED), rc
    budget = -1 if rc == E13_EFORCED else p.budget
    return (budget, p.sym_lo, p.sym_hi if p.sym_lo >= 0 else None, p.sym_n0, list(p.tab[:2 * p.nq]))


def _want(w):
    if len(w) == 1:
        return w
    budget, lo, hi, n0, tab = w
    return (budget, lo, hi if lo >= 0 else None, n0, tab)


def test_oracle_setup_matches_reference(e13):
    base = base_codes(e13)
    bad = [c[:5] for c in CASES if _oracle(base, *c[:5]) != _want(c[5])]
    assert not bad, bad[:5]


def test_library_setup_matches_reference(e13):
    base = base_codes(e13)
    bad = [c[:5] for c in CASES if _library(e13, base, *c[:5]) != _want(c[5])]
    assert not bad, bad[:5]


@pytest.mark.parametrize("mistake", ["whole_table_scan", "lo_negative_refused"])
def test_planted_setup_mistakes_are_caught(e13, mistake):
    """The two ways the restatements once departed from the reference, planted
    in a Python copy of the rule, make the same sweep fail: the fixture has the
    cases that tell them apart."""
    base = base_codes(e13)
    nq_base = len(base) // 2
    differ = 0
    for X, fmax, lo, hi, cells, want in CASES:
        if len(want) == 1 or not cells and lo < 0:
            continue
        slo = lo - (lo > X) if lo >= 0 else -1
        shi = hi - (hi >= X)
        tab = want[4]
        if mistake == "lo_negative_refused" and lo >= 0 and slo < 0:
            differ += 1  # the mistaken rule refuses a problem the reference runs
            continue
        if slo < 0:
            continue
        ncell = len(tab) if mistake == "whole_table_scan" else 2 * (nq_base - 1)
        named = [0] * 16
        for i in range(ncell):
            c = tab[i]
            if c >= 0 and slo <= c >> 2 <= shi:
                named[(c >> 2) - slo] = 1
        n0 = 0
        while n0 <= shi - slo and named[n0]:
            n0 += 1
        ok = all(not named[q - slo] and tab[2 * q] == -2 and tab[2 * q + 1] == -2 for q in range(slo + n0, shi + 1))
        differ += (not ok) or n0 != want[3]
    assert differ > 0


def test_oracle_freesym_without_deletion():
    """The oracle's free-mode searches with X = -1 against the compiled
    reference's output (ref_dump/freesym_xm1.txt): counts, histograms and
    survivor lines."""
    cases, cur = [], None
    for line in read_text(REF_DUMP, "freesym_xm1.txt").splitlines():
        if line.startswith("case "):
            cur = {"args": line.split()[2:], "surv": []}
            cases.append(cur)
        elif line.startswith("SURVIVOR"):
            cur["surv"].append(line.strip())
        elif "leaves=" in line:
            cur["sum"] = parse_summaries(line)[-1]
    assert len(cases) == 3
    for c in cases:
        out, err = run_oracle("freesym", machine_path(), *c["args"], threads=8)
        assert parse_summaries(err)[-1] == c["sum"], c["args"]
        assert survivor_lines(out) == sorted(c["surv"]), c["args"]
