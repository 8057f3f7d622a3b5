"""Whole prover steps of the workload configurations on the HIP path, checked
bit-exact against the oracle on the same seeded inputs."""
import sys

import numpy as np
import pytest

from conftest import NR, Q

pytestmark = pytest.mark.gpu


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


def _cs(orc, reps, nrounds, fq2, seed):
    cs = np.stack([orc.sample_uniform(seed, 0x5C000 + rep, (nrounds * 2,), Q).reshape(nrounds, 2)
                   for rep in range(reps)])
    if not fq2:
        cs[:, :, 1] = 0
    return cs


def test_config1_prover_step(sa, ctx, orc):
    """Small plain step: 4096 uniform ring elements (seed 0) through forward
    and inverse NTT; commit with A 16 x 256 uniform and W 16 rows x 256 in
    [-8, 8]; norm check over all 16 x 2^14 coefficients of W with F_{q^2}
    challenges, 2 repetitions, verifier replay and the norm_check_3 values."""
    say("case: config 1 prover step")
    seed = 0
    x = ctx.uniform((4096, 64), seed, sa.STREAM_NTT)
    x_h = orc.sample_uniform(seed, sa.STREAM_NTT, (4096, 64), Q)
    assert np.array_equal(x.download(), x_h)
    ctx.convert(x, sa.REP_COEFF, sa.REP_INCOMPLETE)
    assert np.array_equal(x.download(), orc.to_incomplete(x_h, Q))
    ctx.convert(x, sa.REP_INCOMPLETE, sa.REP_COEFF)
    assert np.array_equal(x.download(), x_h)

    kappa, m, r = 16, 256, 16
    A = ctx.uniform((kappa, m, 64), seed, sa.STREAM_A)
    W = ctx.bounded((r, m, 64), 8, seed, sa.STREAM_W)
    A_h = orc.sample_uniform(seed, sa.STREAM_A, (kappa, m, 64), Q)
    W_h = orc.sample_bounded(seed, sa.STREAM_W, (r, m, 64), 8, Q)
    assert np.array_equal(W.download(), W_h)
    key = ctx.key(A)
    com = ctx.parallel_dot_series_matrix(key, W).download()
    assert np.array_equal(com, orc.commit(A_h, W_h, Q, threads=8))
    key.close()

    nrounds = 14
    cs = ctx.challenges(2, nrounds, 1, seed)
    assert np.array_equal(cs.download(), _cs(orc, 2, nrounds, 1, seed))
    outs = ctx.norm_check(W, cs, 1)
    want = orc.sumcheck(W_h.reshape(r, m * 64), cs.download(), Q, fq2=True, nr=NR, threads=8)
    for k in ("claims", "msgs", "finals", "evals"):
        assert np.array_equal(outs[k].download(), want[k]), k
    assert outs["ok"].download().all() and want["ok"].all()
    rhs = outs["rhs"].download()
    for rep in range(2):
        assert np.array_equal(rhs[rep], orc.norm_check_3_fq2(W_h, cs.download()[rep], Q, NR))
