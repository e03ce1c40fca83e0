"""The estimator entry point that allocates nothing (e13_estimate_static, over
each GPU's static device scratch) against e13_estimate, through the C ABI.

Two calls in a row on one context, the second with fewer probes (and another
problem) than the first, each must give the same per-probe sums and totals as
e13_estimate bit for bit: the second call shares the scratch the first one
left behind, so stale counters or sums would show.  A request above the
scratch's size is rejected with E13_ESCRATCH before any device work.
Each test prints its case before touching the GPU.
"""
import struct

import pytest

from conftest import MACHINES, probe_fixture, read_text

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def bits(t):
    return tuple(struct.unpack("<Q", struct.pack("<d", v))[0] for v in t)


def gpu_problem(e13, args):
    base = e13.parse_machine(read_text(MACHINES, args["machine"]))
    sym = (args["lo"], args["hi"]) if args["lo"] >= 0 else None
    return e13.setup(base, args["X"], args["fmax"], args["cells"] or None, sym)


@pytest.fixture(scope="module")
def nctx(e13):
    ctx = e13.Context(0)
    ctx.reset_tuning()
    yield ctx
    ctx.close()


def same(e13, ctx, name, seed, n):
    args, _ = probe_fixture(name)
    p = gpu_problem(e13, args)
    say(f"estimate_static {name}: {n} probes, seed {seed}")
    ses, sper = ctx.estimate_static(p, n, seed=seed)
    say(f"estimate {name}: {n} probes, seed {seed}")
    es, per = ctx.estimate(p, n, seed=seed, per_probe=True)
    bad = [i for i in range(n) if bits(sper[i]) != bits(per[i])]
    assert not bad, (name, n, bad[:5])
    assert ses.probes == es.probes == n
    assert ses.sim_steps == es.sim_steps and ses.launches == es.launches == 1
    for j in range(4):
        assert bits([ses.sum[j], ses.sq[j], ses.mean[j], ses.se[j]]) == bits([es.sum[j], es.sq[j], es.mean[j],
                                                                             es.se[j]]), (name, j)


def test_static_matches_estimate_then_smaller(e13, nctx):
    same(e13, nctx, "counter3", 21, e13.STATIC_PROBES)  # the scratch's full size
    same(e13, nctx, "v1_0ld", 22, 3000)                  # smaller, on what the first call left
    same(e13, nctx, "f2_x8", 23, 257)


def test_static_rejects_more_than_scratch(e13, nctx):
    args, _ = probe_fixture("f2_x8")
    p = gpu_problem(e13, args)
    say(f"estimate_static f2_x8: {e13.STATIC_PROBES + 1} probes (rejected on the host)")
    with pytest.raises(e13.E13Error) as ei:
        nctx.estimate_static(p, e13.STATIC_PROBES + 1)
    assert ei.value.code == -14
    same(e13, nctx, "f1_nodel", 24, 500)  # the scratch still serves after a rejection
