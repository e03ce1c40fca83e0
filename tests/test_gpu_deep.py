"""The deep stage check (src/deep.c) and the counter re-check (src/recheck.c)
of survivors on the GPU (k_deep: one wave per survivor, its tape in LDS),
against the reference's own output files; each case prints itself before
touching the GPU."""
import gzip
import os

import pytest

from conftest import REF_RESULTS, read_text

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def dctx(e13):
    ctx = e13.Context(0)
    yield ctx
    ctx.close()


def test_deep_check_erdos14(e13, dctx):
    """erdos_14 to stage 14 (results/validation/deep_base14.txt)."""
    line = [l for l in read_text(REF_RESULTS, "validation_check_erdos14.txt").splitlines() if l.startswith("SURVIVOR")][0]
    say("deep check of erdos_14 to stage 14")
    assert dctx.deep_check(line, 14).splitlines() == read_text(REF_RESULTS, "validation_deep_base14.txt").splitlines()[:1]


def test_deep_check_v2(e13, dctx):
    """All 12 survivors of V2 to stage 15 (results/4_division_6/v2/deep_v2.txt)."""
    say("deep check of the 12 V2 survivors to stage 15")
    text = read_text(REF_RESULTS, "4_division_6_v2_surv_v2.txt")
    assert dctx.deep_check(text, 15) == read_text(REF_RESULTS, "4_division_6_v2_deep_v2.txt")


def test_deep_check_all_v1_survivors(e13, dctx):
    """All 2868 survivors of V1's 1RD search to stage 15: the reference's whole
    results/4_division_6/v1/deep15_1RD.txt, line for line (2865 reach stage 15,
    three leave the tape at stage 12).  The reference ran it in chunks whose
    survivor numbers restart, so the lines are compared after the number; ours
    count from 1 in input order."""
    surv = gzip.open(os.path.join(REF_RESULTS, "4_division_6_v1_surv_1RD.txt.gz"), "rt").read()
    want = gzip.open(os.path.join(REF_RESULTS, "4_division_6_v1_deep15_1RD.txt.gz"), "rt").read().splitlines()
    say("deep check of the 2868 V1 1RD survivors to stage 15")
    got = dctx.deep_check(surv, 15).splitlines()
    assert len(got) == len(want) == 2868
    assert [g.split(" ", 2)[1] for g in got] == [str(i) for i in range(1, 2869)]
    bad = [i for i, (g, w) in enumerate(zip(got, want)) if g.split(" ", 2)[2] != w.split(" ", 2)[2]]
    assert not bad, (len(bad), got[bad[0]], want[bad[0]])


def test_counter_recheck(e13, dctx):
    """src/recheck.c 15 on the V1 survivor lines of validation/recheck_1RD.txt
    (counter changes at stage 15 and 13, stage 12 not reached), numbered from 1
    in input order: the reference's verdicts."""
    say("counter re-check of 7 V1 survivors to stage 15")
    text = read_text(REF_RESULTS, "4_division_6_v1_surv_1RD_recheck_sample.txt")
    want = [w.split(": ", 1)[1] for w in read_text(REF_RESULTS, "validation_recheck_1RD_sample.txt").splitlines()]
    got = dctx.recheck_counter(text, 15).splitlines()
    assert [g.split(": ", 1)[0] for g in got] == [str(i) for i in range(1, 8)]
    assert [g.split(": ", 1)[1] for g in got] == want


def test_gpu_equals_host_on_v1_sample(e13, dctx):
    """The GPU forms equal the host forms (e13_deep_check, e13_recheck_counter)
    on the V1 sample at stages 13 and 15."""
    text = read_text(REF_RESULTS, "4_division_6_v1_surv_1RD_recheck_sample.txt")
    for nmax in (13, 15):
        say(f"deep check and re-check of the V1 sample to stage {nmax}, GPU against host")
        assert dctx.deep_check(text, nmax) == e13.deep_check(text, nmax, threads=8)
        assert dctx.recheck_counter(text, nmax) == e13.recheck_counter(text, nmax, threads=8)


def test_command_line_gpu_modes():
    """`e13search_gpu deep_gpu NMAX < survivors` and `recheck_gpu NMAX`: the
    reference's lines for V2's survivors (deep_v2.txt) and recheck.c's verdicts
    and summary line for the V1 sample."""
    import subprocess
    from conftest import PKG_DIR
    cli = os.path.join(PKG_DIR, "lib", "e13search_gpu")
    say("e13search_gpu deep_gpu 15 < V2 survivors")
    r = subprocess.run([cli, "deep_gpu", "15"], input=read_text(REF_RESULTS, "4_division_6_v2_surv_v2.txt"),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout == read_text(REF_RESULTS, "4_division_6_v2_deep_v2.txt")
    say("e13search_gpu recheck_gpu 15 < V1 sample")
    r = subprocess.run([cli, "recheck_gpu", "15"], input=read_text(REF_RESULTS, "4_division_6_v1_surv_1RD_recheck_sample.txt"),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    want = [w.split(": ", 1)[1] for w in read_text(REF_RESULTS, "validation_recheck_1RD_sample.txt").splitlines()]
    assert [l.split(": ", 1)[1] for l in r.stdout.splitlines()] == want
    assert "machines 7 consistent 0 failed 7" in r.stderr
