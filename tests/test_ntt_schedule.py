"""The NTT's two radix-exchange schedules (the exchange between the register
phases of each 32-point half through LDS, or through DPP quad permutes in
registers) give the same transforms: the standalone IncompleteNTT under each
against the oracle, and a batch's step under each schedule, eager and
replayed from its graph, against the plain per-call path; the tuner reports
its choice and both times."""
import sys

import numpy as np
import pytest

from conftest import Q

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sa():
    import salsaa_amd
    return salsaa_amd


@pytest.mark.parametrize("xch", ["lds", "dpp"])
@pytest.mark.parametrize("count", [1, 127, 1000, 1 << 16])
def test_standalone_ntt_schedules_vs_oracle(sa, orc, monkeypatch, xch, count):
    say("case: standalone IncompleteNTT xch=%s count=%d" % (xch, count))
    monkeypatch.setenv("SALSAA_NTT_XCH", xch)
    ctx = sa.Context(Q, 64)
    try:
        x = orc.sample_uniform(31, 3, (count, 64), Q)
        d = ctx.to_device(x)
        ctx.convert(d, sa.REP_COEFF, sa.REP_INCOMPLETE)
        got = d.download()
        idx = np.unique(np.linspace(0, count - 1, min(count, 512)).astype(int))
        assert np.array_equal(got[idx], orc.to_incomplete(x[idx], Q))
        ctx.convert(d, sa.REP_INCOMPLETE, sa.REP_COEFF)
        assert np.array_equal(d.download(), x)
    finally:
        ctx.close()


def test_batch_step_under_both_schedules(sa, orc):
    say("case: batch step under both NTT schedules")
    ctx = sa.Context(Q, 64)
    rng = np.random.default_rng(17)
    n = 24
    m = [int(1 << rng.integers(0, 11)) for _ in range(n)]
    kappa = [int(rng.choice([8, 16, 32])) for _ in range(n)]
    A = orc.sample_uniform(5, sa.STREAM_A, (32, 1024, 64), Q)
    key = ctx.key(ctx.to_device(A))
    b = sa.Batch(ctx, key, m, kappa, reps=2, fq2=True, tune=True)
    tn = b.tuning()
    say("  tuning", tn)
    assert tn["ntt_schedule"] in ("lds", "dpp")
    assert tn["ntt_pass_ms"]["lds"] > 0 and tn["ntt_pass_ms"]["dpp"] > 0
    for i in range(n):
        s = b.slot(i)
        s["witness"].upload(orc.sample_bounded(200 + i, sa.STREAM_W, (m[i], 64), 16, Q))
        nr = s["nrounds"]
        s["challenges"].upload(np.stack([orc.sample_uniform(300 + i, 0x5C000 + r, (2 * nr,), Q)
                                         .reshape(nr, 2) for r in range(2)]))
    keys = ("claims", "msgs", "finals", "evals", "rhs", "ok")

    def snap(plain):
        out = []
        for i in range(n):
            s = b.slot(i, plain)
            d = {k: s[k].download() for k in keys}
            d["commitment"] = s["commitment"].download()[: kappa[i]]
            d["witness"] = s["witness"].download()
            out.append(d)
        return out

    b.step_plain()
    ctx.sync()
    want = snap(True)
    for xch in ("lds", "dpp", "lds"):
        b.set_ntt_schedule(xch)
        for graph in (False, True, True):
            b.step(graph=graph)
            ctx.sync()
            got = snap(False)
            for i in range(n):
                for k in keys + ("commitment",):
                    assert np.array_equal(got[i][k], want[i][k]), (xch, graph, i, k)
                assert np.array_equal(got[i]["witness"], want[i]["witness"]), (xch, i)
    b.close()
    key.close()
    ctx.close()
