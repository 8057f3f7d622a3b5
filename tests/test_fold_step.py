"""The fused folding step (decompose + fold in one pass, then commit and norm
check) against the same steps as separate calls and against the oracle."""
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


# C: "ternary" (the reference's fold challenges: the matrix-core fold when the
# radix is at most 128), "uniform" (the NTT-domain fold), "mixed" (instance 0
# uniform, the others ternary: both kernels in one step).  parts = 1 with
# 2^12-bounded witnesses gives radix 2^13, which also takes the NTT domain.
@pytest.mark.parametrize("n,r,m,parts,r_out,bound,kappa,cmode", [
    (3, 8, 128, 4, 8, 1 << 12, 16, "ternary"),
    (2, 3, 256, 5, 5, 1000, 9, "ternary"),
    (2, 3, 256, 5, 5, 1000, 9, "uniform"),
    (3, 4, 128, 2, 3, 1 << 12, 8, "mixed"),
    (2, 2, 128, 1, 4, 1 << 12, 8, "ternary"),
    (2, 8, 256, 8, 8, (1 << 20) - 1, 8, "ternary"),
])
def test_fold_step_vs_separate_calls_and_oracle(sa, ctx, orc, n, r, m, parts, r_out, bound,
                                                kappa, cmode):
    say("case: fold step n=%d r=%d m=%d parts=%d r_out=%d C %s" % (n, r, m, parts, r_out, cmode))
    reps, nrounds = 2, (m * 64).bit_length() - 1
    W_h = orc.sample_bounded(21, sa.STREAM_W, (n * r, m, 64), bound, Q)
    if cmode == "uniform":
        C_h = orc.sample_uniform(22, sa.STREAM_FOLD, (n, r_out, r * parts, 64), Q)
    else:
        C_h = orc.sample_bounded(22, sa.STREAM_FOLD, (n, r_out, r * parts, 64), 1, Q)
    if cmode == "mixed":
        C_h[0] = orc.sample_uniform(24, sa.STREAM_FOLD, C_h[0].shape, Q)
    A_h = orc.sample_uniform(23, sa.STREAM_A, (kappa, m, 64), Q)
    cs_h = np.stack([np.stack([orc.sample_uniform(30 + i, 0x5C000 + rep, (nrounds * 2,), Q)
                               .reshape(nrounds, 2) for rep in range(reps)]) for i in range(n)])
    key = ctx.key(ctx.to_device(A_h))
    W, C, cs = ctx.to_device(W_h), ctx.to_device(C_h), ctx.to_device(cs_h)
    got = ctx.fold_step(key, W, n, parts, C, cs)
    g = {k: v.download() for k, v in got.items()}
    assert g["ok"].all()
    rows = n * r_out
    # separate calls: decompose, fold, Coefficient form, commit, norm checks
    coeff = np.empty((rows, m, 64), np.uint64)
    for i in range(n):
        Wi = ctx.to_device(W_h[i * r:(i + 1) * r])
        limbs, d_radix = ctx.decompose(Wi, parts)
        assert int(d_radix.download()[0]) == int(g["radix"][i])
        fo = ctx.fold(limbs, ctx.to_device(C_h[i]))
        sl = slice(i * r_out, (i + 1) * r_out)
        assert np.array_equal(fo.download(), g["folded"][sl]), i
        ctx.convert(fo, sa.REP_INCOMPLETE, sa.REP_COEFF)
        coeff[sl] = fo.download()
        nc = ctx.norm_check(fo, ctx.to_device(cs_h[i]), 1)
        assert np.array_equal(nc["claims"].download(), g["claims"][sl])
        assert np.array_equal(nc["msgs"].download(), g["msgs"][:, :, sl])
        assert np.array_equal(nc["finals"].download(), g["finals"][:, sl])
        assert np.array_equal(nc["evals"].download(), g["evals"][:, sl])
        assert np.array_equal(nc["rhs"].download(), g["nc_rhs"][:, sl])
    assert np.array_equal(coeff, g["folded_coeff"])
    com = ctx.parallel_dot_series_matrix(key, ctx.to_device(coeff)).download()
    assert np.array_equal(com, g["rhs"])
    # oracle: decomposition + fold of the first and last instance, commitment
    # of their rows
    for i in sorted({0, n - 1}):
        limbs, radix = orc.decompose(W_h[i * r:(i + 1) * r], parts, Q)
        assert radix == int(g["radix"][i])
        want_fold = orc.fold(limbs, C_h[i], Q, threads=8)
        sl = slice(i * r_out, (i + 1) * r_out)
        assert np.array_equal(want_fold, g["folded"][sl]), i
        want_com = orc.commit(A_h, coeff[sl], Q, threads=8)
        assert np.array_equal(want_com, g["rhs"][:, sl]), i
    key.close()


def test_fold_step_with_generated_crs_key(sa, ctx, orc):
    # the folding step under a CRS key kept as its module only (rows made in
    # the commit kernel) equals the same step under the stored CRS key
    say("case: fold step, implicit CRS key")
    n, r, m, parts, r_out, kappa = 2, 4, 256, 3, 4, 12
    reps, nrounds = 2, (m * 64).bit_length() - 1
    W = ctx.to_device(orc.sample_bounded(41, sa.STREAM_W, (n * r, m, 64), 1 << 10, Q))
    C = ctx.to_device(orc.sample_bounded(42, sa.STREAM_FOLD, (n, r_out, r * parts, 64), 1, Q))
    cs = ctx.to_device(np.stack([np.stack([orc.sample_uniform(50 + i, 0x5C000 + rep,
                                                              (nrounds * 2,), Q)
                                           .reshape(nrounds, 2) for rep in range(reps)])
                                 for i in range(n)]))
    module = ctx.to_device(orc.sample_uniform(43, sa.STREAM_A, (kappa, 64), Q))
    outs = []
    for implicit in (True, False):
        key = ctx.crs_key(module, m, implicit=implicit)
        got = ctx.fold_step(key, W, n, parts, C, cs)
        outs.append({k: v.download() for k, v in got.items()})
        key.close()
    assert outs[0]["ok"].all()
    for k in outs[1]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
