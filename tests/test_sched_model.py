"""CPU model of the persistent search kernel's work-sharing protocol
(tests/sched_model.c): host threads stand in for waves, with random pauses
between shared-memory operations, deliveries whose elements are written in a
random order and often after the flag, hungry maps of narrow words (several
map words with few waves) and delivery tags of a small width (so the tag
sequence wraps many times in one run).

The protocol as k_search runs it must end every run through the end mark with
the exact leaf, machine and step totals of a sequential walk of the same tree.
Each planted mistake must make some run fail:
  - a launch with no seeds and no early return: the idle watchdog (the hang
    of a search whose breadth-first seeding finished the whole tree);
  - tags made of the donor's id and its give count modulo 2^bits, or a single
    tag value: a receiver sees its last tag again and skips a delivery the
    donor counted as busy, so the search ends only at the watchdog;
  - a delivery not counted as busy.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("sched") / "sched_model")
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-pthread", "-o", exe, os.path.join(ROOT, "tests", "sched_model.c")],
                   check=True)
    return exe


def run(model, *args, timeout=60):
    r = subprocess.run([model, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout.strip()


# (waves, seeds, map bits, tag bits, pool_min_idle, evict_after, flag-first probability)
CASES = [(1, 1, 64, 32, 4, 30, 0.2), (2, 1, 64, 2, 1, 30, 0.5), (3, 1, 64, 2, 8, 0, 0.5), (8, 64, 64, 3, 4, 30, 0.2),
         (16, 3, 4, 2, 2, 30, 0.8), (29, 7, 4, 2, 8, 0, 0.5), (64, 64, 64, 2, 4, 30, 0.5), (65, 2, 8, 3, 1, 30, 0.3),
         (130, 9, 8, 2, 4, 30, 0.5), (200, 200, 64, 2, 8, 10, 0.5)]


@pytest.mark.parametrize("case", CASES, ids=[f"w{c[0]}_s{c[1]}_m{c[2]}_t{c[3]}" for c in CASES])
def test_protocol_ends_with_exact_totals(model, case):
    w, s, mb, tb, pmi, ev, ff = case
    for tree in (1, 2, 3, 4, 5):
        rc, out = run(model, "--waves", w, "--seeds", s, "--depth", 15, "--slots", 6, "--map-bits", mb, "--tag-bits",
                      tb, "--pool-min-idle", pmi, "--evict", ev, "--flag-first", ff, "--tree", tree,
                      "--idle-limit", 10)
        assert rc == 0 and out.startswith("ok "), (case, tree, out)


def test_protocol_on_two_cpus(model):
    """Many waves on two CPUs: the threads are preempted in the middle of
    deliveries, takes and withdrawals."""
    for tree in range(10, 30):
        r = subprocess.run(["taskset", "-c", "0,1", model, "--waves", str(12 + tree % 5), "--seeds",
                            str(tree % 4 + 1), "--depth", "14", "--slots", "6", "--map-bits", "4", "--tag-bits", "2",
                            "--flag-first", "0.6", "--tree", str(tree), "--idle-limit", "10"],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and r.stdout.startswith("ok "), (tree, r.stdout)


def test_no_seeds_leaves_at_once(model):
    rc, out = run(model, "--waves", 16, "--seeds", 0, "--idle-limit", 10, timeout=5)
    assert rc == 0 and out == "ok 0 0 0 0 0"


def test_stack_overflow_is_an_error_not_a_hang(model):
    rc, out = run(model, "--waves", 1, "--seeds", 64, "--depth", 14, "--slots", 6, "--stack", 64)
    assert rc == 1 and out.startswith("error 1 "), out


def test_planted_empty_launch_hits_the_watchdog(model):
    rc, out = run(model, "--waves", 16, "--seeds", 0, "--plant-empty", "--idle-limit", 0.2, timeout=10)
    assert rc == 1 and out.startswith("error 3 "), out


def failures(model, *args, trees=range(1, 13)):
    bad = []
    for tree in trees:
        rc, out = run(model, *args, "--tree", tree, "--idle-limit", 0.3)
        if rc != 0:
            bad.append(out)
    return bad


def test_planted_donor_count_tag_loses_a_delivery(model):
    bad = failures(model, "--waves", 3, "--seeds", 1, "--depth", 18, "--slots", 6, "--tag-bits", 1,
                   "--plant-donor-tag")
    assert bad and all(re.match(r"error 3 ", o) for o in bad), bad


def test_single_tag_value_loses_a_delivery(model):
    bad = failures(model, "--waves", 8, "--seeds", 8, "--depth", 16, "--slots", 6, "--tag-bits", 1)
    assert bad and all(o.startswith("error 3 ") for o in bad), bad


def test_planted_uncounted_delivery_is_caught(model):
    bad = failures(model, "--waves", 8, "--seeds", 8, "--depth", 16, "--slots", 6, "--plant-nobusy")
    assert bad, "an uncounted delivery went unnoticed"
