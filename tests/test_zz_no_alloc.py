"""salsaa_norm_check_noalloc (the norm check whose scratch is the library's
fixed device array) gives exactly salsaa_norm_check's outputs: two calls in a
row, the second smaller (so stale scratch from the first must not leak), over
F_{q^2} and F_q, the multi-pass and the single-kernel structure, and a call
too large for the fixed array is rejected before any device work."""
import numpy as np
import pytest

import salsaa_amd as sa

pytestmark = pytest.mark.gpu

OUTS = ("claims", "msgs", "finals", "evals", "rhs", "ok")


def _pair(ctx, rows, m, reps, fq2, seed):
    print("no-alloc norm check: rows=%d m=%d reps=%d fq2=%d seed=%d" % (rows, m, reps, fq2, seed),
          flush=True)
    W = ctx.bounded((rows, m, sa.D), 16, seed, sa.STREAM_W)
    nrounds = (m * sa.D).bit_length() - 1
    cs = ctx.challenges(reps, nrounds, fq2, seed + 1)
    got = ctx.norm_check_noalloc(W, cs, fq2)
    want = ctx.norm_check(W, cs, fq2)
    ctx.sync()
    for k in OUTS:
        a, b = got[k].download(), want[k].download()
        assert np.array_equal(a, b), (k, rows, m, reps, fq2)
    assert got["ok"].download().all()


def test_noalloc_matches_norm_check_two_calls():
    ctx = sa.Context()
    try:
        # multi-pass (Gram pass, 4-round passes, final kernel), then smaller
        _pair(ctx, 4, 1 << 12, 2, 1, 11)
        _pair(ctx, 2, 1 << 8, 2, 1, 12)
        # F_q, then the single-kernel structure (L <= the final kernel's size)
        _pair(ctx, 3, 1 << 10, 3, 0, 13)
        _pair(ctx, 1, 16, 1, 1, 14)
    finally:
        ctx.close()


def test_noalloc_rejects_oversize():
    ctx = sa.Context()
    try:
        W = ctx.empty((1, sa.D))
        cs = ctx.empty((2, 22, 2))
        o = {k: ctx.empty((1,)) for k in ("claims", "msgs", "finals", "evals")}
        o["rhs"] = None
        o["ok"] = ctx.empty((2, 16), np.int32)
        # 2 reps x 16 rows x 2^22 coefficients = 128 MiB of folded witnesses
        st = sa.lib.salsaa_norm_check_noalloc(ctx.h, W.ptr, 16, 1 << 22, cs.ptr, 2, 1, sa.NR,
                                              o["claims"].ptr, o["msgs"].ptr, o["finals"].ptr,
                                              o["evals"].ptr, None, o["ok"].ptr)
        assert st == -1
        assert b"64 MiB" in sa.lib.salsaa_last_error()
    finally:
        ctx.close()
