"""The continuous-batching server: every job's result is the same served in
a mix of concurrent jobs of different sizes, served alone, and by the oracle."""
import sys
import time

import numpy as np
import pytest

from conftest import NR, Q

pytestmark = pytest.mark.gpu

KAPPA, MMAX = 16, 1024
FIELDS = ("commitment", "claims", "msgs", "finals", "evals", "rhs", "ok")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sa():
    import salsaa_amd
    return salsaa_amd


@pytest.fixture(scope="module")
def ctx(sa):
    c = sa.Context(Q, 64)
    yield c
    c.close()


@pytest.fixture(scope="module")
def keyA(sa, ctx, orc):
    A_h = orc.sample_uniform(11, sa.STREAM_A, (KAPPA, MMAX, 64), Q)
    key = ctx.key(ctx.to_device(A_h))
    yield key, A_h
    key.close()


def jobs_list(orc, n=48, seed=4):
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n):
        m = int(1 << rng.integers(0, 11))
        host = orc.sample_bounded(900 + k, 0xB, (m, 64), 16, Q) if k % 7 == 3 else None
        jobs.append((m, 100 + k, host))
    return jobs


def serve(sa, ctx, key, jobs, batch_jobs, inflight, stagger):
    srv = sa.Server(ctx, key, KAPPA, arena_elems=1 << 13, max_batch_jobs=batch_jobs,
                    max_batch_elems=1 << 12, inflight=inflight, reps=2, fq2=True)
    ids = []
    for i, (m, seed, host) in enumerate(jobs):
        ids.append(srv.submit(m, seed, 16, host))
        if stagger and i % 8 == 7:
            time.sleep(0.002)  # later jobs arrive while earlier batches run
    srv.drain()
    res = []
    for jid, (m, _, _) in zip(ids, jobs):
        r = srv.poll(jid, m)
        assert r is not None
        res.append(r)
    srv.close()
    return res


def test_server_mixed_equals_alone_and_oracle(sa, ctx, orc, keyA):
    key, A_h = keyA
    jobs = jobs_list(orc)
    say("case: server mixed (batched, 2 in flight)")
    mixed = serve(sa, ctx, key, jobs, 64, 2, True)
    say("  batches:", sorted(set(r["batch"] for r in mixed))[-1],
        "max jobs/batch:", max(r["batch_jobs"] for r in mixed))
    assert max(r["batch_jobs"] for r in mixed) > 1
    say("case: server alone (one job per batch)")
    alone = serve(sa, ctx, key, jobs, 1, 1, False)
    assert all(r["batch_jobs"] == 1 for r in alone)
    for i, (x, y) in enumerate(zip(mixed, alone)):
        for f in FIELDS:
            assert np.array_equal(x[f], y[f]), (i, f)
        assert x["ok"].all()
    # oracle on a few jobs, sampled or host witnesses
    for i in (0, 3, 10, len(jobs) - 1):
        m, seed, host = jobs[i]
        W = host if host is not None else orc.sample_bounded(seed, 0xB, (m, 64), 16, Q)
        want_c = orc.commit(np.ascontiguousarray(A_h[:, :m]), W.reshape(1, m, 64), Q, threads=8)
        assert np.array_equal(mixed[i]["commitment"], want_c), i
        n = (m * 64).bit_length() - 1
        cs = np.stack([orc.sample_uniform(seed, 0x5C000 + r, (n * 2,), Q).reshape(n, 2)
                       for r in range(2)])
        want = orc.sumcheck(W.reshape(1, m * 64), cs, Q, fq2=True, nr=NR, threads=8)
        for f in ("claims", "msgs", "finals", "evals"):
            assert np.array_equal(mixed[i][f], want[f]), (i, f)


def test_server_rejects_oversized_jobs(sa, ctx, keyA):
    key, _ = keyA
    srv = sa.Server(ctx, key, KAPPA, arena_elems=256, max_batch_jobs=4, max_batch_elems=256)
    with pytest.raises(sa.SalsaaError):
        srv.submit(512, 1)     # larger than the arena
    with pytest.raises(sa.SalsaaError):
        srv.submit(48, 1)      # not a power of two
    srv.close()


def test_server_with_generated_crs_key(sa, ctx, orc):
    # the server over a CRS key kept as its module only (the key rows made in
    # the batched commit kernel): mixed batches = one job at a time = oracle
    say("case: server, implicit CRS key")
    module = orc.sample_uniform(12, sa.STREAM_A, (KAPPA, 64), Q)
    key = ctx.crs_key(ctx.to_device(module), MMAX, implicit=True)
    jobs = jobs_list(orc, n=24, seed=5)
    mixed = serve(sa, ctx, key, jobs, 64, 2, True)
    alone = serve(sa, ctx, key, jobs, 1, 1, False)
    for i, (x, y) in enumerate(zip(mixed, alone)):
        for f in FIELDS:
            assert np.array_equal(x[f], y[f]), (i, f)
        assert x["ok"].all()
    A = orc.compute_commitment_keys(module, MMAX, Q)
    for i in (0, 3, len(jobs) - 1):
        m, seed, host = jobs[i]
        W = host if host is not None else orc.sample_bounded(seed, 0xB, (m, 64), 16, Q)
        want_c = orc.commit(np.ascontiguousarray(A[:, :m]), W.reshape(1, m, 64), Q, threads=8)
        assert np.array_equal(mixed[i]["commitment"], want_c), i
    key.close()
