"""Checkpoint, clean stop and resume of the host-loop search (the GPU driver's
E13_CKPT handling, src/e13gpu.py:19-23, 349-409).

A search stopped between launches and resumed, any number of times, must give
exactly the counts, histogram and survivors of an uninterrupted run (and of the
reference's log for the same problem).  Each test prints its case before
touching the GPU.
"""
import os
import re
import subprocess

import pytest

from conftest import PKG_DIR, REF_DUMP, REF_RESULTS, machine_path, parse_summaries, read_machine, read_text

pytestmark = pytest.mark.gpu

CLI = os.path.join(PKG_DIR, "lib", "e13search_gpu")


def say(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def cctx(e13):
    """A context of its own, on the batched host loop with small batches (many launches)."""
    ctx = e13.Context(0)
    ctx.set_tuning(form=2, batch=256)
    yield ctx
    ctx.close()


def result(e13, p, c, sv):
    d = c.as_dict()
    lines = sorted(e13.format_survivor(p, s) + " " + e13.format_hash(p, s) for s in sv)
    assert all(s.recheck_ok == 1 for s in sv)
    return (d["leaves"], d["steps"], d["survivors"], tuple(d["hist"])), lines


def test_stop_and_resume_equals_uninterrupted(e13, cctx, tmp_path):
    base = e13.parse_machine(read_machine())
    p = e13.setup(base, 12, 3)
    say("search FMAX=3 X=12, host loop, uninterrupted")
    cctx.set_checkpoint(None)
    c0, sv0, _ = cctx.search(p)
    want = result(e13, p, c0, sv0)
    ref = parse_summaries(read_text(REF_DUMP, "delete_f3.txt"))[12]
    assert want[0] == (ref["leaves"], ref["steps"], ref["survivors"], tuple(ref["hist"]))
    path = str(tmp_path / "ck")
    cctx.set_checkpoint(path, every_sec=0, max_launches=5)
    stops = 0
    say("search FMAX=3 X=12 with a clean stop every 5 launches")
    while True:
        try:
            c, sv, _ = cctx.search(p)
            break
        except e13.E13Error as ex:
            assert ex.code == -9
            stops += 1
            assert stops < 2000
            with open(path) as fh:  # the pointer names a complete slot
                slot = fh.read().strip()
            assert slot in ("a", "b") and os.path.exists(f"{path}.{slot}")
    assert stops >= 3
    assert result(e13, p, c, sv) == want
    # the files stay after the end (as the driver's do): the same search again
    # resumes from the last checkpoint and ends with the same result
    say("search FMAX=3 X=12 again from the last checkpoint")
    cctx.set_checkpoint(path, every_sec=3600)
    c, sv, _ = cctx.search(p)
    assert result(e13, p, c, sv) == want
    # a checkpoint belongs to its problem, and a damaged one is refused
    say("checkpoint of another problem; damaged checkpoint")
    with pytest.raises(e13.E13Error) as ex:
        cctx.search(e13.setup(base, 13, 3))
    assert ex.value.code == -10
    with open(path) as fh:
        slot = fh.read().strip()
    with open(f"{path}.{slot}", "r+b") as fh:
        fh.truncate(os.path.getsize(f"{path}.{slot}") - 16)
    with pytest.raises(e13.E13Error) as ex:
        cctx.search(p)
    assert ex.value.code == -10
    cctx.set_checkpoint(None)


def test_leftover_stop_file_is_removed(e13, cctx, tmp_path):
    """A stop file present at the start is the one of the last clean stop: it is
    removed and the search runs to the end (e13gpu.py:377-378)."""
    base = e13.parse_machine(read_machine())
    p = e13.setup(base, 4, 3)
    path = str(tmp_path / "ck2")
    open(path + ".stop", "w").close()
    say("search FMAX=3 X=4 with a leftover stop file")
    cctx.set_checkpoint(path, every_sec=3600)
    try:
        c, _, _ = cctx.search(p)
    finally:
        cctx.set_checkpoint(None)
    assert not os.path.exists(path + ".stop")
    ref = parse_summaries(read_text(REF_DUMP, "delete_f3.txt"))[4]
    assert (c.leaves, c.steps) == (ref["leaves"], ref["steps"])


def test_cli_exit_code_3_and_resume(tmp_path):
    """The command line: E13_CKPT with a launch budget stops with exit code 3
    and the 'run the same command to resume' line; re-running finishes with the
    reference's summary line."""
    ck = str(tmp_path / "cli_ck")
    env = dict(os.environ, E13_CKPT=ck, E13_CKPT_LAUNCHES="6", E13_FORM="2", E13_BATCH="256")
    args = [CLI, "delete", machine_path(), "3", "12"]
    codes = []
    for _ in range(400):
        say(f"e13search_gpu delete FMAX=3 X=12 with E13_CKPT, run {len(codes) + 1}")
        r = subprocess.run(args, env=env, capture_output=True, text=True, timeout=120)
        codes.append(r.returncode)
        if r.returncode != 3:
            break
        assert "STOPPED cleanly" in r.stderr
    assert codes[-1] == 0 and codes.count(3) >= 2, codes
    got = parse_summaries(r.stderr)[12]
    ref = parse_summaries(read_text(REF_DUMP, "delete_f3.txt"))[12]
    assert (got["leaves"], got["steps"], got["survivors"], got["hist"]) == (ref["leaves"], ref["steps"],
                                                                             ref["survivors"], ref["hist"])


COUNTER = dict(leaves=3078656272, steps=319112357176, survivors=61,
               hist=[0, 2957671647, 119843176, 1115937, 22970, 2236, 1, 0, 244, 0])


@pytest.fixture(scope="module")
def pctx(e13):
    """A context of its own on the persistent form (the default settings)."""
    ctx = e13.Context(0)
    yield ctx
    ctx.close()


def counter_problem(e13):
    base = e13.parse_machine(read_machine())
    return e13.setup(base, 13, free_cells="D0 G0 K0 K1 L1 M1".split(), sym=(10, 13))


def test_persistent_stops_and_relaunches(e13, pctx, tmp_path):
    """The persistent search with a checkpoint every 50 ms: each launch is
    stopped cleanly (running machines paused as nodes on the waves' stacks),
    saved, and relaunched from the gathered nodes; the counter search still
    gives the reference's counts and 61 survivors."""
    p = counter_problem(e13)
    path = str(tmp_path / "p3")
    pctx.set_checkpoint(path, every_sec=0.05)
    say("counter search, persistent form, checkpoint every 50 ms")
    try:
        c, sv, n = pctx.search(p)
    finally:
        pctx.set_checkpoint(None)
    d = c.as_dict()
    assert (d["leaves"], d["steps"], d["survivors"], d["hist"]) == (COUNTER["leaves"], COUNTER["steps"],
                                                                     COUNTER["survivors"], COUNTER["hist"])
    assert n == 61 and all(s.recheck_ok == 1 for s in sv)
    assert os.path.exists(path)  # it did stop and save at least once
    assert pctx.stats()["search_launches"] >= 3


def test_persistent_clean_stop_and_resume(e13, pctx, tmp_path):
    """E13_ESTOPPED after two stopped launches, resumed call after call until
    the search ends: FMAX=4 over X=0..13 in one persistent search per call gives
    the reference's per-X counts; a checkpoint of the host loop or of another
    search is refused."""
    base = e13.parse_machine(read_machine())
    ps = [e13.setup(base, X, 4) for X in range(14)]
    path = str(tmp_path / "p4")
    pctx.set_checkpoint(path, every_sec=0.25, max_launches=2)
    stops = 0
    say("delete FMAX=4 X=0..13, persistent form, stop after every 2 launches of 250 ms")
    try:
        while True:
            try:
                cs, sv, n = pctx.search_many(ps)
                break
            except e13.E13Error as ex:
                assert ex.code == -9
                stops += 1
                assert stops < 200
        with pytest.raises(e13.E13Error) as ex:  # the same file for other problems
            pctx.search_many(ps[:13])
        assert ex.value.code == -10
    finally:
        pctx.set_checkpoint(None)
    assert stops >= 1
    ref = parse_summaries(read_text(REF_DUMP, "delete_f4.txt"))
    for X in range(14):
        d = cs[X].as_dict()
        assert (d["leaves"], d["steps"], d["hist"]) == (ref[X]["leaves"], ref[X]["steps"], ref[X]["hist"]), X
    assert n == sum(ref[X]["survivors"] for X in range(14)) == 57
    assert all(s.recheck_ok == 1 for s in sv)


def test_persistent_stop_file(e13, pctx, tmp_path):
    """The stop file (<path>.stop) stops a running persistent search within
    about a second; the same call again resumes and finishes it."""
    import threading
    import time
    p = counter_problem(e13)
    base = e13.parse_machine(read_machine())
    ps = [e13.setup(base, X, 4) for X in range(14)]
    path = str(tmp_path / "p5")
    pctx.set_checkpoint(path, every_sec=3600)
    timer = threading.Timer(0.4, lambda: open(path + ".stop", "w").close())
    say("delete FMAX=4 X=0..13, persistent form, stop file after 0.4 s")
    try:
        timer.start()
        t0 = time.time()
        with pytest.raises(e13.E13Error) as ex:
            pctx.search_many(ps)
        assert ex.value.code == -9 and time.time() - t0 < 2.5
        assert not os.path.exists(path + ".stop")
        say("resume")
        cs, sv, n = pctx.search_many(ps)
    finally:
        timer.cancel()
        pctx.set_checkpoint(None)
    ref = parse_summaries(read_text(REF_DUMP, "delete_f4.txt"))
    assert sum(cs[X].leaves for X in range(14)) == sum(ref[X]["leaves"] for X in range(14))
    assert sum(cs[X].steps for X in range(14)) == sum(ref[X]["steps"] for X in range(14))
    assert n == 57


def test_cli_persistent_multi_x_checkpoint(tmp_path):
    """The command line on the persistent form: delete FMAX=3 over X=0..13 in
    one search with one checkpoint file, stopped after every launch of 20 ms
    (exit code 3) and re-run until it ends with the reference's summaries."""
    ck = str(tmp_path / "cli_p3")
    env = dict(os.environ, E13_CKPT=ck, E13_CKPT_SEC="0.02", E13_CKPT_LAUNCHES="1")
    args = [CLI, "delete", machine_path(), "3"] + [str(x) for x in range(14)]
    codes = []
    for _ in range(200):
        say(f"e13search_gpu delete FMAX=3 X=0..13 with E13_CKPT, run {len(codes) + 1}")
        r = subprocess.run(args, env=env, capture_output=True, text=True, timeout=120)
        codes.append(r.returncode)
        if r.returncode != 3:
            break
        assert "STOPPED cleanly" in r.stderr
    assert codes[-1] == 0 and codes.count(3) >= 1, codes
    got = parse_summaries(r.stderr)
    ref = parse_summaries(read_text(REF_DUMP, "delete_f3.txt"))
    for X in range(14):
        assert (got[X]["leaves"], got[X]["steps"], got[X]["survivors"], got[X]["hist"]) == (
            ref[X]["leaves"], ref[X]["steps"], ref[X]["survivors"], ref[X]["hist"]), X



def test_cli_gpu_driver_v3_checkpointed(tmp_path):
    """The division search V3 (freesym X=9 3..9 E0..I1: 53,282,072,641 leaves,
    5.8e12 steps) through the GPU driver's argv with E13_CKPT, the way the
    reference runs V1/V2 (src/e13gpu.py:293-300, 349-409): stopped after every
    two launches of one second (exit code 3, clean stop, nodes saved) and re-run
    until it ends, each run a new process resuming from the last slot.  The
    final summary line equals the driver's own log (g_v3.log) up to the wall
    time."""
    ck = str(tmp_path / "v3")
    env = dict(os.environ, E13_CKPT=ck, E13_CKPT_SEC="1", E13_CKPT_LAUNCHES="2")
    args = [CLI, "gpu", machine_path(), "9", "3", "9"] + "E0 E1 F0 F1 G0 G1 H0 H1 I0 I1".split()
    codes = []
    for _ in range(60):
        say(f"e13search_gpu gpu erdos_14 9 3 9 E0..I1 with E13_CKPT, run {len(codes) + 1}")
        r = subprocess.run(args, env=env, capture_output=True, text=True, timeout=120)
        codes.append(r.returncode)
        if r.returncode != 3:
            break
        assert "STOPPED cleanly" in r.stderr
    assert codes[-1] == 0 and codes.count(3) >= 2, (codes, r.stderr[-2000:])
    strip = lambda l: re.sub(r" wall=[0-9.]+s", "", l).strip()
    got = [strip(l) for l in r.stderr.splitlines() if " gpu leaves=" in l]
    want = [strip(l) for l in read_text(REF_RESULTS, "4_division_6_v3_g_v3.txt").splitlines() if " gpu leaves=" in l]
    assert got == want
    assert not [l for l in r.stdout.splitlines() if l.startswith("SURVIVOR")]
