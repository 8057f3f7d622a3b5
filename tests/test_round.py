"""One SNARK round's prover operators (protocol.rs:164-452) in the reference's
order on device-resident data under a CRS key generated in the commit kernel,
with each of the verifier's relations checked on the device's outputs:
verify_decomp, the norm check's verifier replay, verifier_split, the batching
assertion, verifier_fold and the commitment part of verifier_join."""
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


@pytest.fixture(scope="module")
def ctx(sa):
    c = sa.Context(Q, 64)
    yield c
    c.close()


def ring_combine(sa, ctx, C, X):
    """out[j][l] = sum_k C[j][k] X[l][k] for C [r_out][r_in][d] and X
    [rows][r_in][d] (both Coefficient); out [rows][r_out][d] IncompleteNTT."""
    rows, r_in, d = X.shape
    r_out = C.shape[0]
    a = np.ascontiguousarray(np.broadcast_to(C[None], (rows, r_out, r_in, d)))
    b = np.ascontiguousarray(np.broadcast_to(X[:, None], (rows, r_out, r_in, d)))
    prod = ctx.ring_mul(ctx.to_device(a), ctx.to_device(b)).download()
    out = np.zeros((rows, r_out, d), np.uint64)
    for k in range(r_in):
        out = (out + prod[:, :, k]) % np.uint64(Q)
    return out


def coeff(sa, ctx, x):
    dx = ctx.to_device(x)
    ctx.convert(dx, sa.REP_INCOMPLETE, sa.REP_COEFF)
    return dx.download()


def test_snark_round_relations(sa, ctx, orc):
    say("case: one SNARK round, implicit CRS key")
    kappa, r, m, parts = 4, 2, 4096, 2
    module = ctx.to_device(orc.sample_uniform(60, sa.STREAM_A, (kappa, 64), Q))
    key = ctx.crs_key(module, m, implicit=True)
    W = ctx.bounded((r, m, 64), 1, 61, sa.STREAM_W)            # sample_random_short_mat
    com = ctx.parallel_dot_series_matrix(key, W)              # statement commitment
    # decomp_ell + verify_decomp (decomp.rs:32-112)
    W2, rhs_dec, radix = ctx.decomp_ell(key, W, parts)
    assert int(ctx.verify_decomp(rhs_dec, com, parts, radix).download()[0]) == 0
    # norm checks, two repetitions with F_q challenges (protocol.rs:222-259)
    nrounds = (m * 64).bit_length() - 1
    nc = ctx.norm_check(W2, ctx.challenges(2, nrounds, 0, 62), 0)
    assert nc["ok"].download().all()
    # split + verifier_split (split.rs:45-76, 171-221): L + m^(m/2) R = rhs
    lr = ctx.split(key, W2).download()
    rw = r * parts
    R = coeff(sa, ctx, np.ascontiguousarray(lr[:, rw:]))
    mult = orc.compute_commitment_keys(module.download(), m // 2, Q)[:, m // 2 - 1]
    prod = ctx.ring_mul(ctx.to_device(np.repeat(mult[:, None], rw, axis=1)),
                        ctx.to_device(R)).download()
    assert np.array_equal((lr[:, :rw] + prod) % np.uint64(Q), rhs_dec.download())
    W2h = W2.download()
    W3 = ctx.to_device(np.concatenate([W2h[:, : m // 2], W2h[:, m // 2:]]))   # [2 rw][m / 2]
    r3, m3 = 2 * rw, m // 2
    # projection + its commitment (project.rs:52-98)
    rng = np.random.default_rng(63)
    z = ctx.to_device(rng.integers(0, 2, (256, r3 * 256), dtype=np.uint8), np.uint8)
    sg = ctx.to_device(rng.integers(0, 2, (256, r3 * 256), dtype=np.uint8), np.uint8)
    P = ctx.project(W3, z, sg)
    assert P.shape == (m3, 64)
    P1 = sa.DeviceArray.view(ctx, P.ptr, (1, m3, 64))
    pcom = ctx.parallel_dot_series_matrix(key, P1)           # [kappa][1]
    # batch_projections + verify_batching (project.rs:131-217)
    one = np.zeros(64, np.uint64)
    one[0] = 1
    plhs, wlhs, prhs, wrhs, ok = ctx.batch_projections(P, W3, z, sg, ctx.to_device(one))
    assert ok.download()[0] == 1
    # fold + verifier_fold (fold.rs:22-66): commit(fold(W)) = C . rhs
    r_out = 2
    Ch = ctx.bounded((r_out, r3, 64), 1, 64, sa.STREAM_FOLD).download()  # challenge_for_fold
    W4 = ctx.fold(W3, ctx.to_device(Ch))
    ctx.convert(W4, sa.REP_INCOMPLETE, sa.REP_COEFF)
    rhs3 = coeff(sa, ctx, lr)                                 # split output, [kappa][r3]
    vfold = ring_combine(sa, ctx, Ch, rhs3)                   # [kappa][r_out]
    assert np.array_equal(ctx.parallel_dot_series_matrix(key, W4).download(), vfold)
    # cross terms and the join (project.rs:229-275): the witness gains the
    # projection as a row, and the statement rows' commitments are the folded
    # ones followed by the projection's commitment
    l0, l1 = ctx.send_cross_terms_before_join(plhs, wlhs, W4, P)
    assert np.array_equal(l0.download(), ctx.parallel_dot_single_series_matrix(
        plhs, sa.DeviceArray.view(ctx, W4.ptr, (1, m3, 64))).download()[0])
    joined = ctx.to_device(np.concatenate([W4.download(), P.download()[None]]))
    jcom = ctx.parallel_dot_series_matrix(key, joined).download()
    assert np.array_equal(jcom, np.concatenate([vfold, pcom.download()], axis=1))
    key.close()
