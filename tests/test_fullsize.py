"""Oracle parity at the configurations' real sizes.

Config 2: the commit of W (8 x 16384, |w| <= 16) under A (32 x 16384) in full,
decomp_ell base 2^8 into 6 limbs in full, limb commitments and the device's
verify_decomp.  Config 3: a batch whose instances reach m = 2^18 (2^24
coefficients, 24 rounds) with kappa 8, 16 and 32 under a 32 x 2^18 key:
commitment rows and every norm-check output of long, medium and short
instances.  Config 4: the fused folding step at 8 instances of 8 x 4096 with
|w| <= 2^12: one instance's decomposition and fold, one folded row's whole
norm check, and the commitments of that instance's folded rows.

The oracle runs on host threads beside each other (ctypes releases the GIL),
so the whole file stays within about a minute of CPU time on 16 threads."""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

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


def test_cfg2_commit_and_decomp_ell_full_size(sa, ctx, orc):
    say("case: config 2 commit 32 x 16384 . 8 x 16384, decomp_ell 2^8 x 6")
    kappa, r, m, parts, radix = 32, 8, 16384, 6, 256
    A_h = orc.sample_uniform(0, sa.STREAM_A, (kappa, m, 64), Q)
    W_h = orc.sample_bounded(0, sa.STREAM_W, (r, m, 64), 16, Q)
    key = ctx.key(ctx.to_device(A_h))
    W = ctx.to_device(W_h)
    com = ctx.parallel_dot_series_matrix(key, W)
    new_W, rhs_dec, d_radix = ctx.decomp_ell(key, W, parts, radix)
    mism = ctx.verify_decomp(rhs_dec, com, parts, d_radix)
    got_com, got_limbs, got_dec = com.download(), new_W.download(), rhs_dec.download()
    assert int(d_radix.download()[0]) == radix
    assert int(mism.download()[0]) == 0
    say("  oracle: commit in full (%d ring products)" % (kappa * r * m))
    assert np.array_equal(got_com, orc.commit(A_h, W_h, Q, threads=THREADS))
    limbs, used = orc.decompose(W_h, parts, Q, radix)
    assert used == radix
    assert np.array_equal(got_limbs, limbs)
    # limb commitments: the top and bottom limbs of several rows, all key rows
    pick = [0, 5, 23, 47]
    want = orc.commit(A_h, np.ascontiguousarray(limbs[pick]), Q, threads=THREADS)
    assert np.array_equal(got_dec[:, pick], want)
    key.close()


def _instance_oracle(orc, A_rows, W, cs, nr):
    """commitment rows, and every norm-check output of one instance"""
    m = W.shape[0]
    com = orc.commit(A_rows, W.reshape(1, m, 64), Q, threads=1)
    nc = orc.sumcheck(W.reshape(1, m * 64), cs, Q, fq2=True, nr=nr, threads=1)
    rhs = np.stack([orc.norm_check_3_fq2(W.reshape(1, m, 64), cs[rep], Q, nr)
                    for rep in range(cs.shape[0])])
    return com, nc, rhs


def test_cfg3_batch_instances_at_full_lengths(sa, ctx, orc):
    say("case: config 3 batch, instances of 2^13, 2^16 and 2^18 ring elements")
    M, K = 1 << 18, 32
    named = [(1 << 18, 32), (1 << 16, 8), (1 << 13, 16), (1 << 18, 8), (1 << 16, 32),
             (1 << 13, 32), (1 << 16, 16)]
    rng = np.random.default_rng(3)
    fill = [(int(1 << rng.integers(10, 19)), int(rng.choice([8, 16, 32]))) for _ in range(17)]
    inst = named + fill
    ms, ks = [x[0] for x in inst], [x[1] for x in inst]
    A = ctx.uniform((K, M, 64), 3, sa.STREAM_A)
    key = ctx.key(A)
    A.free()
    b = sa.Batch(ctx, key, ms, ks, reps=2, fq2=True, tune=True)
    say("  tuning", b.tuning())
    for i in range(b.n):
        s = b.slot(i)
        w = s["witness"]
        sa._check(sa.lib.salsaa_sample_bounded(ctx.h, w.ptr, w.shape[0] * 64, 16, 1000 + i,
                                               sa.STREAM_W))
        c = s["challenges"]
        sa._check(sa.lib.salsaa_sample_challenges(ctx.h, c.ptr, 2, s["nrounds"], 1, 5000 + i))
    b.step(graph=True)
    b.step_plain()
    ctx.sync()
    picks = [0, 1, 2, 4, 5]
    jobs, got = {}, {}
    with ThreadPoolExecutor(max_workers=len(picks)) as ex:
        for i in picks:
            s, p = b.slot(i), b.slot(i, plain=True)
            g = {k: s[k].download() for k in ("claims", "msgs", "finals", "evals", "rhs", "ok")}
            g["commitment"] = s["commitment"].download()[: ks[i]]
            for k in g:
                want_plain = (p[k].download()[: ks[i]] if k == "commitment"
                              else p[k].download())
                assert np.array_equal(g[k], want_plain), (i, k)
            assert g["ok"].all()
            got[i] = g
            rows = [0, ks[i] - 1]
            A_rows = np.stack([orc.sample_uniform_at(3, sa.STREAM_A, j * M * 64, (ms[i], 64), Q)
                               for j in rows])
            say("  oracle: instance %d (m = %d, kappa = %d)" % (i, ms[i], ks[i]))
            jobs[i] = ex.submit(_instance_oracle, orc, A_rows, s["witness"].download(),
                                s["challenges"].download(), NR)
        for i in picks:
            com, nc, rhs = jobs[i].result()
            g = got[i]
            assert np.array_equal(g["commitment"][[0, ks[i] - 1]], com), i
            for k in ("claims", "msgs", "finals", "evals"):
                assert np.array_equal(g[k], nc[k]), (i, k)
            assert np.array_equal(g["rhs"], rhs), i
    b.close()
    key.close()


def test_cfg4_fold_step_full_shape(sa, ctx, orc):
    say("case: config 4 folding step, 8 instances of 8 x 4096, |w| <= 2^12")
    n, r, m, parts, r_out, kappa, reps = 8, 8, 4096, 4, 8, 16, 2
    nrounds = (m * 64).bit_length() - 1
    W_h = orc.sample_bounded(21, sa.STREAM_W, (n * r, m, 64), 1 << 12, Q)
    C_h = orc.sample_bounded(22, sa.STREAM_FOLD, (n, r_out, r * parts, 64), 1, Q)
    A_h = orc.sample_uniform(23, sa.STREAM_A, (kappa, m, 64), Q)
    cs_h = np.stack([np.stack([orc.sample_uniform(30 + i, 0x5C000 + rep, (nrounds * 2,), Q)
                               .reshape(nrounds, 2) for rep in range(reps)]) for i in range(n)])
    key = ctx.key(ctx.to_device(A_h))
    got = ctx.fold_step(key, ctx.to_device(W_h), n, parts, ctx.to_device(C_h),
                        ctx.to_device(cs_h))
    g = {k: v.download() for k, v in got.items()}
    assert g["ok"].all()
    i, row = 5, 2  # instance 5, its folded row 2
    sl = slice(i * r_out, (i + 1) * r_out)
    fr = i * r_out + row
    limbs, radix = orc.decompose(W_h[i * r:(i + 1) * r], parts, Q)
    assert radix == int(g["radix"][i]) == 16
    with ThreadPoolExecutor(max_workers=3) as ex:
        f_fold = ex.submit(orc.fold, limbs, C_h[i], Q, max(1, THREADS - 2))
        wrow = g["folded_coeff"][fr]
        f_nc = ex.submit(orc.sumcheck, wrow.reshape(1, m * 64), cs_h[i], Q, True, NR, 1)
        f_com = ex.submit(orc.commit, A_h, g["folded_coeff"][sl], Q, 1)
        want_fold = f_fold.result()
        assert np.array_equal(g["folded"][sl], want_fold)
        assert np.array_equal(g["folded_coeff"][sl], orc.from_incomplete(want_fold, Q))
        nc = f_nc.result()
        assert np.array_equal(g["claims"][fr:fr + 1], nc["claims"])
        assert np.array_equal(g["msgs"][:, :, fr:fr + 1], nc["msgs"])
        assert np.array_equal(g["finals"][:, fr:fr + 1], nc["finals"])
        assert np.array_equal(g["evals"][:, fr:fr + 1], nc["evals"])
        for rep in range(reps):
            want_rhs = orc.norm_check_3_fq2(wrow.reshape(1, m, 64), cs_h[i][rep], Q, NR)
            assert np.array_equal(g["nc_rhs"][rep, fr:fr + 1], want_rhs)
        assert np.array_equal(g["rhs"][:, sl], f_com.result())
    key.close()
