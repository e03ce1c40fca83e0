"""Checkpoint, clean stop and resume of the host-loop search (the GPU driver's
E13_CKPT handling, src/e13gpu.py:19-23, 349-409).

A search stopped between launches and resumed, any number of times, must give
exactly the counts, histogram and survivors of an uninterrupted run (and of the
reference's log for the same problem).  Each test prints its case before
touching the GPU.
"""
import os
import subprocess

import pytest

from conftest import PKG_DIR, REF_DUMP, machine_path, parse_summaries, read_machine, read_text

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
