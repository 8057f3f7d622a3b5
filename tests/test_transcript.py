"""The transcript on the device against the oracle: the Fiat-Shamir norm check
(salsaa_norm_check_fs: challenges from the SHA-256 chain over the messages)
gives the oracle's challenges and every output, over F_{q^2} and F_q, at
lengths down to one ring element; the transcript bytes of a norm check (with
and without Fiat-Shamir), of a batch instance's prover step and of a
protocol run equal the oracle's byte for byte, and a reader that sees only
the bytes accepts them."""
import os
import sys

import numpy as np
import pytest

from conftest import NR, Q

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


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
def tr():
    from oracle import transcript
    return transcript


@pytest.mark.parametrize("rows,m,reps,fq2,bound", [(3, 64, 2, True, 16), (1, 1, 1, True, 16),
                                                   (2, 2, 2, False, None), (8, 256, 2, True, 1 << 20),
                                                   (1, 4096, 2, True, 16)])
def test_norm_check_fs_vs_oracle(sa, ctx, orc, tr, rows, m, reps, fq2, bound):
    say("case: norm_check_fs rows=%d m=%d reps=%d fq2=%d" % (rows, m, reps, fq2))
    L = m * 64
    w = (orc.sample_uniform(7, 0xB, (rows, L), Q) if bound is None
         else orc.sample_bounded(7, 0xB, (rows, L), bound, Q))
    seed = bytes((5 * i + rows) % 256 for i in range(32))
    got = ctx.norm_check_fs(ctx.to_device(w), reps, fq2, NR, seed)
    g = {k: v.download() for k, v in got.items()}
    want = tr.sumcheck_fs(w, reps, Q, fq2=fq2, nr=NR, seed=seed)
    for k in ("cs", "claims", "msgs", "finals", "evals", "ok"):
        assert np.array_equal(g[k], want[k]), k
    assert g["ok"].all()
    for rep in range(reps):
        want_rhs = orc.norm_check_3_fq2(w.reshape(rows, m, 64), want["cs"][rep], Q, NR)
        assert np.array_equal(g["rhs"][rep], want_rhs), rep


def test_norm_check_transcripts_bytes(sa, ctx, orc, tr):
    say("case: norm check transcript bytes, Fiat-Shamir and seeded")
    rows, m, reps = 3, 32, 2
    L = m * 64
    w = orc.sample_bounded(8, 0xB, (rows, L), 16, Q)
    dw = ctx.to_device(w)
    seed = bytes(range(32, 64))
    # Fiat-Shamir
    o = ctx.norm_check_fs(dw, reps, True, NR, seed)
    b = ctx.transcript_norm_check(o, L, o["cs"], True, NR, fs=True, seed=seed)
    want = tr.sumcheck_fs(w, reps, Q, seed=seed)
    rhs = np.stack([orc.norm_check_3_fq2(w.reshape(rows, m, 64), want["cs"][r], Q, NR)
                    for r in range(reps)])
    assert b == tr.norm_check_bytes(want, Q, L, True, NR, True, seed, rhs)
    assert tr.verify_norm_check_bytes(b)
    # seeded challenges (salsaa_norm_check)
    cs = ctx.challenges(reps, L.bit_length() - 1, 1, 9)
    o2 = ctx.norm_check(dw, cs, 1)
    b2 = ctx.transcript_norm_check(o2, L, cs, True, NR)
    want2 = orc.sumcheck(w, cs.download(), Q, fq2=True, nr=NR, threads=THREADS)
    want2["cs"] = cs.download()
    rhs2 = o2["rhs"].download()
    assert np.array_equal(rhs2, np.stack([orc.norm_check_3_fq2(w.reshape(rows, m, 64),
                                                               want2["cs"][r], Q, NR)
                                          for r in range(reps)]))
    assert b2 == tr.norm_check_bytes(want2, Q, L, True, NR, False, None, rhs2)
    assert tr.verify_norm_check_bytes(b2)


def test_prover_step_transcript_of_a_batch_instance(sa, ctx, orc, tr):
    say("case: prover-step transcript of batch instances")
    ms, ks = [64, 16, 128], [16, 8, 32]
    A = orc.sample_uniform(11, sa.STREAM_A, (32, 128, 64), Q)
    key = ctx.key(ctx.to_device(A))
    b = sa.Batch(ctx, key, ms, ks, reps=2, fq2=True, tune=False)
    Ws = []
    for i, m in enumerate(ms):
        sl = b.slot(i)
        Ws.append(orc.sample_bounded(60 + i, sa.STREAM_W, (m, 64), 16, Q))
        sl["witness"].upload(Ws[i])
        n = sl["nrounds"]
        sl["challenges"].upload(np.stack([orc.sample_uniform(70 + i, 0x5C000 + r, (2 * n,), Q)
                                          .reshape(n, 2) for r in range(2)]))
    b.step(graph=False)
    ctx.sync()
    for i, m in enumerate(ms):
        sl = b.slot(i)
        com = sl["commitment"]
        got = ctx.transcript_prover_step(com, ks[i], sl, m * 64, sl["challenges"])
        cs = sl["challenges"].download()
        o = orc.sumcheck(Ws[i].reshape(1, -1), cs, Q, fq2=True, nr=NR, threads=THREADS)
        o["cs"] = cs
        rhs = np.stack([orc.norm_check_3_fq2(Ws[i].reshape(1, m, 64), cs[r], Q, NR)
                        for r in range(2)])
        want_com = orc.commit(np.ascontiguousarray(A[: ks[i], :m]), Ws[i].reshape(1, m, 64), Q,
                              threads=THREADS)
        assert got == tr.prover_step_bytes(want_com, o, Q, m * 64, True, NR, False, None, rhs), i
    b.close()
    key.close()


@pytest.mark.parametrize("folding", [False, True])
def test_protocol_transcript(sa, ctx, orc, tr, folding):
    say("case: protocol transcript folding=%d" % folding)
    from oracle import protocol as op
    kappa, r, m, parts = (3, 4, 1 << 10, None) if folding else (3, 1, 1 << 11, (2,))
    module_h = orc.sample_uniform(400, sa.STREAM_A, (kappa, 64), Q)
    W_h = orc.sample_bounded(401, sa.STREAM_W, (r, m, 64), 1, Q)
    out = ctx.protocol(ctx.to_device(module_h), ctx.to_device(W_h), 1, parts, folding=folding,
                       seed=21)
    got = ctx.transcript_protocol(out, kappa, 1, folding, 21)
    want = op.protocol(module_h, W_h, 1, parts, folding, 21, Q, threads=THREADS)
    assert out["status"] == want["status"] == 0
    assert got == tr.protocol_bytes(Q, kappa, 1, folding, 21, 0, want["rhs"], want["W"])
    h = tr.parse_header(got)
    assert (h["kind"], h["rows"], h["nrounds"], h["kappa"]) == (3, want["W"].shape[0],
                                                               want["rhs"].shape[0], kappa)
