"""The protocol run on the device (protocol.rs:125-483) through
salsaa_protocol: every verifier relation of every round holds (status 0), the
shapes follow the reference's round structure, and the final verifier rhs is
checked independently of the driver: its CRS rows against the plain commit of
the final witness, and its last row (projected_lhs with the batching
challenge 1, an all-ones layer) against the sum of the witness rows' ring
elements computed on the host.  The oracle's replay of protocol.rs (same
seeded challenges, the reference's own route for every statement row) gives
the same final witness and every row of the verifier rhs, and each verifier
check, given one prover value off by one, reports its own code on the device
and in the replay."""
import os
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


def setup(sa, ctx, orc, kappa, r, m, seed=60):
    module = ctx.to_device(orc.sample_uniform(seed, sa.STREAM_A, (kappa, 64), Q))
    W = ctx.bounded((r, m, 64), 1, seed + 1, sa.STREAM_W)   # sample_random_short_mat
    return module, W


def ones_row(sa, ctx, W):
    """<all-ones layer, W[k]> for every row k, IncompleteNTT."""
    s = (W.download().sum(axis=1, dtype=object) % Q).astype(np.uint64)
    ds = ctx.to_device(s)
    ctx.convert(ds, sa.REP_COEFF, sa.REP_INCOMPLETE)
    return ds.download()


@pytest.mark.parametrize("kappa,m,rounds,parts", [(4, 1 << 13, 2, (2, 2)), (6, 1 << 12, 1, (4,)),
                                                  (3, 1 << 14, 3, (2, 1, 2))])
def test_snark_rounds(sa, ctx, orc, kappa, m, rounds, parts):
    say("case: snark kappa=%d m=%d rounds=%d parts=%s" % (kappa, m, rounds, parts))
    module, W = setup(sa, ctx, orc, kappa, 1, m)
    out = ctx.protocol(module, W, rounds, parts, seed=7)
    assert out["status"] == 0
    Wf, rhs = out["W"], out["rhs"].download()
    rf, mf, _ = Wf.shape
    assert rf == 2 and mf == m >> rounds
    assert rhs.shape[0] == kappa + 4 * rounds
    key = ctx.crs_key(module, m, implicit=True)
    assert np.array_equal(ctx.parallel_dot_series_matrix(key, Wf).download(), rhs[:kappa])
    key.close()
    assert np.array_equal(ones_row(sa, ctx, Wf), rhs[-1])
    assert all(t >= 0 for t in out["times"].values())


def test_folding_scheme(sa, ctx, orc):
    say("case: folding scheme kappa=4 m=4096 r=4")
    module, W = setup(sa, ctx, orc, 4, 4, 1 << 12, seed=70)
    out = ctx.protocol(module, W, 1, folding=True, seed=9)
    assert out["status"] == 0
    Wf, rhs = out["W"], out["rhs"]
    assert Wf.shape == (4, 1 << 12, 64) and rhs.shape == (8, 4, 64)
    # after the closing decomp_ell(.., 2) (verify_decomp held: status 0) the
    # verifier rhs's CRS rows are the plain commitment of the limbs
    key = ctx.crs_key(module, 1 << 12, implicit=True)
    com = ctx.parallel_dot_series_matrix(key, Wf).download()
    assert np.array_equal(com, rhs.download()[:4])
    key.close()


def test_seeded_and_deterministic(sa, ctx, orc):
    say("case: protocol determinism kappa=4 m=8192")
    module, W = setup(sa, ctx, orc, 4, 1, 1 << 13, seed=80)
    a = ctx.protocol(module, W, 1, (2,), seed=11)
    b = ctx.protocol(module, W, 1, (2,), seed=11)
    c = ctx.protocol(module, W, 1, (2,), seed=12)
    assert a["status"] == b["status"] == c["status"] == 0
    assert np.array_equal(a["W"].download(), b["W"].download())
    assert np.array_equal(a["rhs"].download(), b["rhs"].download())
    assert not np.array_equal(a["W"].download(), c["W"].download())


def test_projection_too_wide_is_an_error(sa, ctx, orc):
    say("case: protocol with 256 r > m")
    module, W = setup(sa, ctx, orc, 4, 1, 1 << 9, seed=90)
    with pytest.raises(sa.SalsaaError, match="256 r <= m"):
        ctx.protocol(module, W, 1, (2,), seed=1)


THREADS = max(1, min(16, os.cpu_count() or 1))


def replay(orc, module_h, W_h, rounds, parts, folding, seed, fault=None):
    from oracle import protocol as op
    return op.protocol(module_h, W_h, rounds, parts, folding, seed, Q, fault=fault,
                       threads=THREADS)


def host_setup(sa, ctx, orc, kappa, r, m, seed):
    module_h = orc.sample_uniform(seed, sa.STREAM_A, (kappa, 64), Q)
    W_h = orc.sample_bounded(seed + 1, sa.STREAM_W, (r, m, 64), 1, Q)
    return module_h, W_h, ctx.to_device(module_h), ctx.to_device(W_h)


@pytest.mark.parametrize("kappa,r,m,rounds,parts,folding,seed",
                         [(3, 1, 1 << 11, 1, (2,), False, 7),
                          (4, 1, 1 << 13, 2, (1, 2), False, 11),
                          (3, 4, 1 << 10, 1, None, True, 9)])
def test_protocol_matches_oracle_replay(sa, ctx, orc, kappa, r, m, rounds, parts, folding, seed):
    say("case: protocol vs oracle replay kappa=%d r=%d m=%d rounds=%d parts=%s folding=%d"
        % (kappa, r, m, rounds, parts, folding))
    module_h, W_h, module, W = host_setup(sa, ctx, orc, kappa, r, m, 100 + seed)
    out = ctx.protocol(module, W, rounds, parts, folding=folding, seed=seed)
    want = replay(orc, module_h, W_h, rounds, parts, folding, seed)
    assert out["status"] == 0 and want["status"] == 0
    assert np.array_equal(out["W"].download(), want["W"])
    got_rhs = out["rhs"].download()
    assert got_rhs.shape == want["rhs"].shape
    for e in range(got_rhs.shape[0]):
        assert np.array_equal(got_rhs[e], want["rhs"][e]), e


@pytest.mark.parametrize("check", [1, 2, 3, 4, 5, 6])
def test_fault_codes_match_oracle(sa, ctx, orc, check):
    # one prover value off by one before check `check` reads it: the device
    # and the replay both report 100 * round + check
    folding = check == 5
    kappa, r, m, parts = (3, 4, 1 << 10, None) if folding else (3, 1, 1 << 11, (2,))
    say("case: protocol fault check=%d" % check)
    module_h, W_h, module, W = host_setup(sa, ctx, orc, kappa, r, m, 300)
    out = ctx.protocol(module, W, 1, parts, folding=folding, seed=5, fault=(0, check))
    want = replay(orc, module_h, W_h, 1, parts, folding, 5, fault=(0, check))
    expect = 100 + 6 if check == 6 else check
    assert out["status"] == want["status"] == expect, (out["status"], want["status"])


@pytest.mark.parametrize("fault,expect", [((1, 1), 101), ((1, 2), 102), ((1, 3), 103),
                                          ((1, 4), 104), ((0, 3), 3), ((0, 6), 206)])
def test_fault_codes_second_round(sa, ctx, orc, fault, expect):
    say("case: protocol fault %s in two rounds" % (fault,))
    module, W = setup(sa, ctx, orc, 4, 1, 1 << 13, seed=310)
    out = ctx.protocol(module, W, 2, (1, 2), seed=3, fault=fault)
    assert out["status"] == expect


def test_many_statement_rows(sa, ctx, orc):
    # kappa = 36 and 8 rounds: 68 statement rows (two tables of split
    # multipliers) and 16 norm-check rows (two passes of the row evaluation)
    say("case: protocol kappa=36 m=2^20 8 rounds")
    kappa, m, rounds = 36, 1 << 20, 8
    module, W = setup(sa, ctx, orc, kappa, 1, m, seed=320)
    out = ctx.protocol(module, W, rounds, (1,) + (2,) * (rounds - 1), seed=13)
    assert out["status"] == 0
    Wf, rhs = out["W"], out["rhs"].download()
    assert rhs.shape[0] == kappa + 4 * rounds == 68
    key = ctx.crs_key(module, m, implicit=True)
    assert np.array_equal(ctx.parallel_dot_series_matrix(key, Wf).download(), rhs[:kappa])
    key.close()
    assert np.array_equal(ones_row(sa, ctx, Wf), rhs[-1])


def test_outputs_reused_and_memory_trimmed(sa, ctx, orc):
    # a run into the outputs of the one before it (the joined witness is
    # written in place), after the context's cached memory was given back,
    # equals the first run
    say("case: protocol outputs reused, context trimmed")
    for folding, r, m, parts in ((False, 1, 1 << 12, (2,)), (True, 4, 1 << 11, None)):
        module, W = setup(sa, ctx, orc, 5, r, m, seed=330)
        a = ctx.protocol(module, W, 1, parts, folding=folding, seed=4)
        Wa, Va = a["W"].download(), a["rhs"].download()
        ctx.trim()
        b = ctx.protocol(module, W, 1, parts, folding=folding, seed=4, outs=a)
        assert b["W"] is a["W"] and b["status"] == a["status"] == 0
        assert np.array_equal(b["W"].download(), Wa)
        assert np.array_equal(b["rhs"].download(), Va)
