"""Batched prover steps (grouped launches, hipGraph replay) against the plain
per-call path and the oracle: every output of every instance bit-exact."""
import ctypes
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


OUTS = ("claims", "msgs", "finals", "evals", "rhs", "ok")


def fill(sa, ctx, b, seed0):
    for i in range(b.n):
        s = b.slot(i)
        w = s["witness"]
        sa._check(sa.lib.salsaa_sample_bounded(ctx.h, w.ptr, w.shape[0] * 64, 16, seed0 + i,
                                               sa.STREAM_W))
        cs = s["challenges"]
        sa._check(sa.lib.salsaa_sample_challenges(ctx.h, cs.ptr, b.reps, s["nrounds"],
                                                  int(b.fq2), seed0 + 7919 * (i + 1)))


def snapshot(b, plain):
    res = []
    for i in range(b.n):
        s = b.slot(i, plain)
        d = {k: s[k].download() for k in OUTS}
        d["commitment"] = s["commitment"].download()[: s["kappa"]]
        res.append(d)
    return res


def same(a, b):
    for i, (x, y) in enumerate(zip(a, b)):
        for k in x:
            assert np.array_equal(x[k], y[k]), (i, k)


@pytest.mark.parametrize("fq2,reps", [(1, 2), (0, 1)])
def test_batch_matches_plain_and_oracle(sa, ctx, orc, fq2, reps):
    say("case: batch fq2=%d reps=%d" % (fq2, reps))
    rng = np.random.default_rng(5 + fq2)
    n = 37
    m = [int(1 << rng.integers(0, 10)) for _ in range(n)]
    kappa = [int(rng.choice([8, 16, 32])) for _ in range(n)]
    kmax, mmax = 32, 512
    A_h = orc.sample_uniform(3, sa.STREAM_A, (kmax, mmax, 64), Q)
    key = ctx.key(ctx.to_device(A_h))
    b = sa.Batch(ctx, key, m, kappa, reps=reps, fq2=fq2, tune=True)
    t = b.tuning()
    say("  tuning", t)
    # the tuner's choices are among its timed candidates, and its synthetic
    # witnesses are gone (the commitments below also need the arena's padding
    # zero again)
    assert str(t["ntt_blocks"]) in t["ntt_blocks_ms"]
    if t["step_ms_by_second_pass_blocks"]:
        assert str(t["step_second_pass_blocks"]) in t["step_ms_by_second_pass_blocks"]
    assert not any(b.slot(i)["witness"].download().any() for i in range(n))
    fill(sa, ctx, b, 100)
    W0 = [b.slot(i)["witness"].download() for i in range(n)]
    b.step_plain()
    plain = snapshot(b, True)
    b.step(graph=False)
    eager = snapshot(b, False)
    same(eager, plain)
    for i in range(n):
        assert np.array_equal(b.slot(i)["witness"].download(), W0[i])
        assert plain[i]["ok"].all()
    # oracle on a few instances of different lengths
    picks = sorted(set([int(np.argmax(m)), int(np.argmin(m)), 0, n - 1]))
    for i in picks:
        s = b.slot(i)
        Wi = W0[i].reshape(1, m[i], 64)
        want_c = orc.commit(np.ascontiguousarray(A_h[: kappa[i], : m[i]]), Wi, Q, threads=8)
        assert np.array_equal(eager[i]["commitment"], want_c), i
        cs = s["challenges"].download()
        want = orc.sumcheck(W0[i].reshape(1, m[i] * 64), cs, Q, fq2=bool(fq2), nr=NR, threads=8)
        for k in ("claims", "msgs", "finals", "evals"):
            assert np.array_equal(eager[i][k], want[k]), (i, k)
    # graph replays, with eager work and synchronisation between them
    for rep in range(3):
        b.step(graph=True)
        ctx.sync()
        same(snapshot(b, False), plain)
        b.step_plain()
        ctx.sync()
    b.close()
    key.close()


def test_batch_rejects_bad_plans(sa, ctx, orc):
    say("case: batch argument checks")
    A = ctx.uniform((8, 64, 64), 1, sa.STREAM_A)
    key = ctx.key(A)
    with pytest.raises(sa.SalsaaError):
        sa.Batch(ctx, key, [3], [8], tune=False)      # not a power of two
    with pytest.raises(sa.SalsaaError):
        sa.Batch(ctx, key, [128], [8], tune=False)    # longer than the key
    with pytest.raises(sa.SalsaaError):
        sa.Batch(ctx, key, [64], [9], tune=False)     # more rows than the key
    key.close()


def test_batch_unfused_phases_agree(sa, ctx, orc):
    # The unfused alternatives of the step's NTT pass (forward alone, forward
    # with planes, inverse, planes from the Coefficient witnesses), run through
    # the phase API, give the fused pass's commitments and restore the witnesses.
    say("case: batch unfused phases")
    rng = np.random.default_rng(11)
    n = 21
    m = [int(1 << rng.integers(0, 10)) for _ in range(n)]
    kappa = [int(rng.choice([8, 16, 32])) for _ in range(n)]
    A_h = orc.sample_uniform(4, sa.STREAM_A, (32, 512, 64), Q)
    key = ctx.key(ctx.to_device(A_h))
    b = sa.Batch(ctx, key, m, kappa, reps=1, fq2=1, tune=False)
    fill(sa, ctx, b, 300)
    W0 = [b.slot(i)["witness"].download() for i in range(n)]

    def commitments():
        return [b.slot(i)["commitment"].download()[: kappa[i]] for i in range(n)]

    def witnesses_restored():
        for i in range(n):
            assert np.array_equal(b.slot(i)["witness"].download(), W0[i]), i

    b.step(graph=False)
    want = commitments()
    for seq in (("ntt_fwd_planes", "ntt_inv", "commit"),
                ("ntt_fwd", "ntt_inv", "planes", "commit"),
                ("planes", "commit")):
        say("  phases", seq)
        for i in range(n):  # clear the commitments between runs
            c = b.slot(i)["commitment"]
            c.upload(np.zeros(c.shape, np.uint64))
        for name in seq:
            b.phase(name)
        ctx.sync()
        witnesses_restored()
        got = commitments()
        for i in range(n):
            assert np.array_equal(got[i], want[i]), (seq, i)
    for i in (0, n - 1):
        Wi = W0[i].reshape(1, m[i], 64)
        assert np.array_equal(want[i], orc.commit(np.ascontiguousarray(A_h[: kappa[i], : m[i]]),
                                                  Wi, Q, threads=8)), i
    b.close()
    key.close()


def test_batch_short_copy_mixed(sa, ctx, orc):
    # The step's norm checks read the round trip's two-byte copy of the
    # witnesses wherever a tile (128 ring elements) holds only residues below
    # 2^14 in balanced form.  Here one instance is uniform, two have a single
    # residue just past the bound (2^14 and -2^14 - 1), one holds only the
    # extreme short values, and the rest are short: first-pass blocks of one
    # row take either path.  The step (eager, graph, and phase by phase) must
    # equal the plain per-call path and the oracle.
    say("case: batch short copy, mixed magnitudes")
    m = [4096, 4096, 4096, 512, 256, 2048, 1024, 64]
    kappa = [8, 16, 32, 8, 16, 8, 32, 8]
    n = len(m)
    A_h = orc.sample_uniform(6, sa.STREAM_A, (32, 4096, 64), Q)
    key = ctx.key(ctx.to_device(A_h))
    b = sa.Batch(ctx, key, m, kappa, reps=2, fq2=1, tune=False)
    fill(sa, ctx, b, 500)
    W0 = [b.slot(i)["witness"].download() for i in range(n)]

    def put(i, flat_idx, value):
        W0[i].reshape(-1)[flat_idx] = value

    put(0, 3 * (4 * m[0]) + 100, 1 << 14)            # row segment 3, column 100
    put(3, 7 * (4 * m[3]) + 4 * m[3] - 1, Q - (1 << 14) - 1)
    W0[1] = orc.sample_uniform(501, sa.STREAM_W, W0[1].shape, Q)
    pick = np.random.default_rng(9).random(W0[2].shape) < 0.5
    W0[2] = np.where(pick, (1 << 14) - 1, Q - (1 << 14)).astype(np.uint64)
    for i in range(n):
        b.slot(i)["witness"].upload(W0[i])
    b.step_plain()
    plain = snapshot(b, True)
    b.step(graph=False)
    same(snapshot(b, False), plain)
    for i in range(n):
        s = b.slot(i)
        assert np.array_equal(s["witness"].download(), W0[i]), i
        Wi = W0[i].reshape(1, m[i], 64)
        want_c = orc.commit(np.ascontiguousarray(A_h[: kappa[i], : m[i]]), Wi, Q, threads=8)
        assert np.array_equal(plain[i]["commitment"], want_c), i
        want = orc.sumcheck(W0[i].reshape(1, m[i] * 64), s["challenges"].download(), Q,
                            fq2=True, nr=NR, threads=8)
        for k in ("claims", "msgs", "finals", "evals"):
            assert np.array_equal(plain[i][k], want[k]), (i, k)
        assert plain[i]["ok"].all()
    for rep in range(2):
        b.step(graph=True)
        ctx.sync()
        same(snapshot(b, False), plain)
        b.step_plain()
        ctx.sync()

    def clear():
        for i in range(n):
            for k in ("msgs", "claims"):
                t = b.slot(i)[k]
                t.upload(np.zeros(t.shape, t.download().dtype))

    for seq in (("ntt_roundtrip_planes", "commit", "norm_check"), ("norm_check_witness",)):
        say("  phases", seq)
        clear()
        for name in seq:
            b.phase(name)
        ctx.sync()
        same(snapshot(b, False), plain)
    b.close()
    key.close()


def test_batch_short_copy_widths(sa, ctx, orc):
    # The round trip keeps each slice of 512 coefficients (8 ring elements)
    # one byte wide when its balanced residues lie in [-128, 127], two bytes
    # wide when they lie in [-2^14, 2^14), and the norm check's first pass
    # reads a sub-block from the copy only when its slices share one width.
    # Instances: the byte edges -128 and 127 (64-segment pass); byte slices
    # with every other slice widened by one 128 or -129; a byte half and a
    # two-byte half; random slices of either width (16-segment pass); and
    # only the two-byte edges -129 and 128.  Eager and graph steps must equal
    # the plain per-call path, and the plain path the oracle.
    say("case: batch short copy, byte and two-byte slices")
    m = [4096, 2048, 1024, 512, 256]
    kappa = [8, 16, 8, 32, 8]
    n = len(m)
    A_h = orc.sample_uniform(16, sa.STREAM_A, (32, 4096, 64), Q)
    key = ctx.key(ctx.to_device(A_h))
    b = sa.Batch(ctx, key, m, kappa, reps=2, fq2=1, tune=False)
    fill(sa, ctx, b, 1600)
    rng = np.random.default_rng(16)

    def res(v):
        return np.where(v < 0, v + Q, v).astype(np.uint64)

    def byte_vals(k):
        return rng.integers(-128, 128, k)

    W0 = []
    v = rng.choice(np.array([-128, 127, 0]), (m[0], 64))
    W0.append(res(v))
    v = byte_vals(m[1] * 64)
    for sl in range(0, m[1] * 64 // 512, 2):
        v[sl * 512 + rng.integers(0, 512)] = 128 if sl % 4 == 0 else -129
    W0.append(res(v).reshape(m[1], 64))
    v = byte_vals(m[2] * 64)
    h = m[2] * 32
    v[h:] = rng.integers(-(1 << 14), 1 << 14, h)
    W0.append(res(v).reshape(m[2], 64))
    v = byte_vals(m[3] * 64).reshape(-1, 512)
    wide = rng.random(v.shape[0]) < 0.5
    v[wide] = rng.integers(-(1 << 14), 1 << 14, (int(wide.sum()), 512))
    W0.append(res(v).reshape(m[3], 64))
    v = rng.choice(np.array([-129, 128, -(1 << 14), (1 << 14) - 1]), (m[4], 64))
    W0.append(res(v))
    for i in range(n):
        b.slot(i)["witness"].upload(W0[i])
    b.step_plain()
    plain = snapshot(b, True)
    for i in range(n):
        s = b.slot(i)
        want = orc.sumcheck(W0[i].reshape(1, m[i] * 64), s["challenges"].download(), Q,
                            fq2=True, nr=NR, threads=8)
        for k in ("claims", "msgs", "finals", "evals"):
            assert np.array_equal(plain[i][k], want[k]), (i, k)
        assert plain[i]["ok"].all()
    b.step(graph=False)
    same(snapshot(b, False), plain)
    for i in range(n):
        assert np.array_equal(b.slot(i)["witness"].download(), W0[i]), i
    for rep in range(2):
        b.step(graph=True)
        ctx.sync()
        same(snapshot(b, False), plain)
        b.step_plain()
        ctx.sync()
    b.close()
    key.close()


def test_batch_wide_rows_long_walk(sa, ctx, orc):
    # A class of many long rows whose residues are all wide (no tile short):
    # the first pass's blocks that take the rows from the witnesses walk many
    # sub-blocks, and digit rows of 127 (residue 0x7F7F7F7F7F7F) make the
    # int32 digit sums of the Gram matrix grow fastest (6 x 127^2 per column
    # for E_5).  The step must equal the plain per-call path.
    say("case: batch, 130 wide rows of m = 2^15")
    n, m = 130, 1 << 15
    A_h = orc.sample_uniform(12, sa.STREAM_A, (8, m, 64), Q)
    key = ctx.key(ctx.to_device(A_h))
    b = sa.Batch(ctx, key, [m] * n, [8] * n, reps=2, fq2=1, tune=False)
    fill(sa, ctx, b, 1300)
    wide = np.full((m, 64), 0x7F7F7F7F7F7F, np.uint64)
    for i in range(n):
        b.slot(i)["witness"].upload(wide)
    b.step_plain()
    plain = snapshot(b, True)
    b.step(graph=False)
    ctx.sync()
    same(snapshot(b, False), plain)
    assert all(p["ok"].all() for p in plain)
    b.close()
    key.close()


def test_batch_split_pass_equals_single_pass(sa, ctx, orc):
    # The step's NTT pass runs in two launches (SALSAA_STEP_SPLIT=1: the
    # shorter classes first and, up to 80 % of the arena, rows of the next
    # class, whose norm check then runs as two groups (here the 1024 class,
    # one row in each part); 2: the longer classes first; the first part's
    # norm checks overlap the rest of the pass) or one (0): every order gives
    # the plain path's outputs.
    import os
    say("case: batch split pass vs single pass")
    m = [1024, 1024, 512, 256, 256, 128, 64, 32, 8, 1]
    kappa = [8, 32, 16, 8, 16, 32, 8, 16, 8, 32]
    n = len(m)
    A_h = orc.sample_uniform(8, sa.STREAM_A, (32, 1024, 64), Q)
    key = ctx.key(ctx.to_device(A_h))
    outs = []
    for env in ("1", "2", "0"):
        os.environ["SALSAA_STEP_SPLIT"] = env
        try:
            b = sa.Batch(ctx, key, m, kappa, reps=2, fq2=1, tune=False)
        finally:
            del os.environ["SALSAA_STEP_SPLIT"]
        fill(sa, ctx, b, 900)
        b.step_plain()
        plain = snapshot(b, True)
        for graph in (False, True, True):
            b.step(graph=graph)
            ctx.sync()
            same(snapshot(b, False), plain)
        outs.append(plain)
        b.close()
    same(outs[0], outs[1])
    same(outs[0], outs[2])
    key.close()


def test_batch_commit_schedules(sa, ctx, orc):
    # The commit runs as one launch per class of key row tiles or as one
    # launch over every class (blocks of all classes that read one key chunk
    # side by side); each class here holds several jobs of 16 instances, the
    # last one partial, and kappa values inside each class.  Both schedules
    # give the plain path's outputs, eagerly and from graph replays, and the
    # oracle's commitments.
    say("case: batch commit schedules")
    rng = np.random.default_rng(17)
    n = 75
    m = [int(1 << rng.integers(0, 10)) for _ in range(n)]
    kappa = [int(rng.choice([3, 8, 9, 12, 16, 18, 20, 32])) for _ in range(n)]
    A_h = orc.sample_uniform(15, sa.STREAM_A, (32, 512, 64), Q)
    key = ctx.key(ctx.to_device(A_h))
    b = sa.Batch(ctx, key, m, kappa, reps=1, fq2=1, tune=False)
    fill(sa, ctx, b, 2100)
    W0 = [b.slot(i)["witness"].download() for i in range(n)]
    b.step_plain()
    plain = snapshot(b, True)
    for sched in ("mixed", "per_class", "mixed_small", "mixed"):
        say("  schedule", sched)
        b.set_commit_schedule(sched)
        for graph in (False, True, True):
            for i in range(n):
                c = b.slot(i)["commitment"]
                c.upload(np.zeros(c.shape, np.uint64))
            b.step(graph=graph)
            ctx.sync()
            same(snapshot(b, False), plain)
    for i in sorted(set([int(np.argmax(m)), 0, 1, 2, n - 1])):
        Wi = W0[i].reshape(1, m[i], 64)
        want = orc.commit(np.ascontiguousarray(A_h[: kappa[i], : m[i]]), Wi, Q, threads=8)
        assert np.array_equal(plain[i]["commitment"], want), i
    b.close()
    key.close()


def test_batch_ntt_grid(sa, ctx, orc):
    # The step's NTT pass runs on any number of persistent blocks (the tuner
    # picks the fastest of one per CU, 7/8 and 3/4 of that): every grid gives
    # the plain path's outputs.
    say("case: batch NTT pass grids")
    m = [2048, 1024, 1024, 512, 256, 256, 128, 64, 16, 2]
    kappa = [32, 8, 16, 32, 8, 16, 32, 8, 16, 8]
    A_h = orc.sample_uniform(13, sa.STREAM_A, (32, 2048, 64), Q)
    key = ctx.key(ctx.to_device(A_h))
    b = sa.Batch(ctx, key, m, kappa, reps=2, fq2=1, tune=False)
    fill(sa, ctx, b, 1700)
    b.step_plain()
    plain = snapshot(b, True)
    for blocks in (0, 200, 7, 1):
        b.set_ntt_grid(blocks)
        for graph in (False, True, True):
            b.step(graph=graph)
            ctx.sync()
            same(snapshot(b, False), plain)
    b.close()
    key.close()


def test_batch_with_generated_crs_key(sa, ctx, orc):
    # the batched step over a CRS key kept as its module only (the key rows
    # made in the batched commit kernel): eager step and graph replays = the
    # plain path (per-instance views, the single commit kernel) = the oracle
    say("case: batch, implicit CRS key")
    m = [2048, 1024, 1024, 512, 256, 128, 64, 8]
    kappa = [8, 32, 16, 8, 32, 16, 8, 32]
    n = len(m)
    module = orc.sample_uniform(13, sa.STREAM_A, (32, 64), Q)
    key = ctx.crs_key(ctx.to_device(module), 2048, implicit=True)
    b = sa.Batch(ctx, key, m, kappa, reps=2, fq2=1, tune=True)
    with pytest.raises(sa.SalsaaError):
        b.set_commit_schedule("mixed")  # the mixed launch reads a stored key
    fill(sa, ctx, b, 1300)
    W0 = [b.slot(i)["witness"].download() for i in range(n)]
    b.step_plain()
    plain = snapshot(b, True)
    for graph in (False, True, True):
        b.step(graph=graph)
        ctx.sync()
        same(snapshot(b, False), plain)
    A = orc.compute_commitment_keys(module, 2048, Q)
    for i in (0, 1, 4, n - 1):
        want = orc.commit(np.ascontiguousarray(A[: kappa[i], : m[i]]), W0[i].reshape(1, m[i], 64),
                          Q, threads=8)
        assert np.array_equal(plain[i]["commitment"], want), i
        assert plain[i]["ok"].all()
    b.close()
    key.close()
