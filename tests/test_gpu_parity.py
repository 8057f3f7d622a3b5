"""Parity of the HIP path (through the C ABI) with the CPU oracle: bit-exact
on seeded inputs at sizes the oracle finishes in seconds, and size-independent
properties at the configs' full sizes."""
import sys

import numpy as np
import pytest

from conftest import NR, Q

pytestmark = pytest.mark.gpu

SEED = 0


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


def rand_ring(orc, shape, seed, stream, bound=None):
    if bound is None:
        return orc.sample_uniform(seed, stream, shape, Q)
    return orc.sample_bounded(seed, stream, shape, bound, Q)


# ------------------------------------------------------------------ sampling
def test_samplers_match_oracle(ctx, orc):
    say("case: samplers")
    u = ctx.uniform((300, 64), 7, 11).download()
    assert np.array_equal(u, orc.sample_uniform(7, 11, (300, 64), Q))
    b = ctx.bounded((5000,), 16, 3, 12).download()
    assert np.array_equal(b, orc.sample_bounded(3, 12, (5000,), 16, Q))
    for fq2 in (0, 1):
        cs = ctx.challenges(3, 9, fq2, 42).download()
        for rep in range(3):
            want = orc.sample_uniform(42, 0x5C000 + rep, (18,), Q).reshape(9, 2)
            if not fq2:
                want[:, 1] = 0
            assert np.array_equal(cs[rep], want)


# ------------------------------------------------------------------ NTT
@pytest.mark.parametrize("count", [1, 127, 128, 1000])
def test_ntt_incomplete_and_full_vs_oracle(sa, ctx, orc, count):
    say("case: ntt count=%d" % count)
    x = rand_ring(orc, (count, 64), SEED, 3)
    dx = ctx.to_device(x)
    ctx.convert(dx, sa.REP_COEFF, sa.REP_INCOMPLETE)
    inc = dx.download()
    assert np.array_equal(inc, orc.to_incomplete(x, Q))
    ctx.convert(dx, sa.REP_INCOMPLETE, sa.REP_FULL)  # via Coefficient
    full = dx.download()
    assert np.array_equal(full, orc.to_full(x, Q))
    ctx.convert(dx, sa.REP_FULL, sa.REP_COEFF)
    assert np.array_equal(dx.download(), x)


def test_ntt_edge_values(sa, ctx, orc):
    say("case: ntt edge values")
    x = np.zeros((8, 64), np.uint64)
    x[1] = Q - 1
    x[2, 0] = 1
    x[3, 63] = Q - 1
    x[4] = np.arange(64, dtype=np.uint64)
    x[5] = Q - 1 - np.arange(64, dtype=np.uint64)
    x[6, ::2] = Q - 1
    x[7, 1::2] = (Q - 1) // 2
    dx = ctx.to_device(x)
    ctx.convert(dx, sa.REP_COEFF, sa.REP_INCOMPLETE)
    assert np.array_equal(dx.download(), orc.to_incomplete(x, Q))
    ctx.convert(dx, sa.REP_INCOMPLETE, sa.REP_COEFF)
    assert np.array_equal(dx.download(), x)


def test_ring_mul_vs_oracle(ctx, orc):
    say("case: ring_mul")
    a = rand_ring(orc, (200, 64), 1, 1)
    b = rand_ring(orc, (200, 64), 1, 2)
    out = ctx.ring_mul(ctx.to_device(a), ctx.to_device(b)).download()
    for i in range(0, 200, 13):
        assert np.array_equal(out[i], orc.ring_mul_incomplete(a[i], b[i], Q))


# ------------------------------------------------------------------ commit
@pytest.mark.parametrize("kappa,r,m,bound", [(4, 3, 40, None), (16, 16, 256, 8), (8, 1, 1024, 16),
                                             (32, 8, 64, 16), (5, 9, 300, 4)])
def test_commit_vs_oracle(ctx, orc, kappa, r, m, bound):
    say("case: commit kappa=%d r=%d m=%d" % (kappa, r, m))
    A = rand_ring(orc, (kappa, m, 64), SEED, 0xA)
    W = rand_ring(orc, (r, m, 64), SEED, 0xB, bound)
    key = ctx.key(ctx.to_device(A))
    got = ctx.parallel_dot_series_matrix(key, ctx.to_device(W)).download()
    want = orc.commit(A, W, Q, threads=8)
    assert np.array_equal(got, want)
    # a shorter witness uses the key's prefix layer
    m2 = m // 2
    got2 = ctx.parallel_dot_series_matrix(key, ctx.to_device(np.ascontiguousarray(W[:, :m2]))).download()
    want2 = orc.commit(np.ascontiguousarray(A[:, :m2]), np.ascontiguousarray(W[:, :m2]), Q, 8)
    assert np.array_equal(got2, want2)
    key.close()


def test_split_vs_oracle(ctx, orc):
    say("case: split")
    kappa, r, m = 6, 3, 128
    A = rand_ring(orc, (kappa, m, 64), 5, 0xA)
    W = rand_ring(orc, (r, m, 64), 5, 0xB, 16)
    key = ctx.key(ctx.to_device(A))
    got = ctx.split(key, ctx.to_device(W)).download()
    halves = np.concatenate([W[:, : m // 2], W[:, m // 2:]], axis=0)
    want = orc.commit(np.ascontiguousarray(A[:, : m // 2]), halves, Q, 8)
    assert np.array_equal(got, want)
    key.close()


# ------------------------------------------------------------------ power series
@pytest.mark.parametrize("m0,n", [(1, 1), (5, 2), (16, 4)])
def test_complete_power_series_vs_oracle(ctx, orc, m0, n):
    say("case: complete_power_series m0=%d n=%d" % (m0, n))
    base = rand_ring(orc, (m0, 64), 41 + n, 0xC)
    tensors = rand_ring(orc, (n, 2, 64), 42 + n, 0xD)
    got = ctx.complete_power_series(ctx.to_device(base), ctx.to_device(tensors)).download()
    assert np.array_equal(got, orc.complete_power_series(base, tensors, Q))


@pytest.mark.parametrize("m,r", [(1, 1), (48, 3), (200, 5), (3000, 16), (700, 17)])
def test_dot_single_series_vs_oracle(ctx, orc, m, r):
    # r <= 16: one pass in the NTT domain; r = 17: the commit with the layer as its key
    say("case: parallel_dot_single_series_matrix m=%d r=%d" % (m, r))
    layer = rand_ring(orc, (m, 64), 43, 0xC)
    W = rand_ring(orc, (r, m, 64), 44, 0xC, 16)
    got = ctx.parallel_dot_single_series_matrix(ctx.to_device(layer), ctx.to_device(W))
    assert np.array_equal(got.download(), orc.dot_single_series(layer, W, Q))


def test_power_series_then_dot_vs_oracle(ctx, orc):
    # the batching step of the projection (project.rs:174-193): a layer
    # expanded from a base and tensors, then its inner product with the rows
    say("case: expanded layer, then inner product")
    base = rand_ring(orc, (8, 64), 45, 0xC)
    tensors = rand_ring(orc, (3, 2, 64), 46, 0xD)
    W = rand_ring(orc, (2, 64, 64), 47, 0xC)
    lay = ctx.complete_power_series(ctx.to_device(base), ctx.to_device(tensors))
    got = ctx.parallel_dot_single_series_matrix(lay, ctx.to_device(W)).download()
    want = orc.dot_single_series(orc.complete_power_series(base, tensors, Q), W, Q)
    assert np.array_equal(got, want)


# ------------------------------------------------------------------ CRS
@pytest.mark.parametrize("kappa,length", [(2, 256), (3, 300), (1, 1)])
def test_compute_commitment_keys_vs_oracle(ctx, orc, kappa, length):
    say("case: compute_commitment_keys kappa=%d len=%d" % (kappa, length))
    module = rand_ring(orc, (kappa, 64), 61 + kappa, 0xC)
    got = ctx.compute_commitment_keys(ctx.to_device(module), length).download()
    assert np.array_equal(got, orc.compute_commitment_keys(module, length, Q))


@pytest.mark.parametrize("kappa,m,r", [(4, 200, 3), (16, 1024, 8), (40, 64, 2)])
def test_crs_key_commit_vs_oracle(ctx, orc, kappa, m, r):
    # the key generated on the device (digit planes for kappa <= 36, slot-major
    # doubles above) commits as the key made from the oracle's CRS rows, and as
    # the oracle; prefix views too
    say("case: CRS key kappa=%d m=%d r=%d" % (kappa, m, r))
    module = rand_ring(orc, (kappa, 64), 70 + kappa, 0xC)
    A = orc.compute_commitment_keys(module, m, Q)
    W = rand_ring(orc, (r, m, 64), 71, 0xB, 16)
    dW = ctx.to_device(W)
    key = ctx.crs_key(ctx.to_device(module), m)
    got = ctx.parallel_dot_series_matrix(key, dW).download()
    assert np.array_equal(got, orc.commit(A, W, Q, threads=8))
    ref_key = ctx.key(ctx.to_device(A))
    assert np.array_equal(got, ctx.parallel_dot_series_matrix(ref_key, dW).download())
    k2, m2 = max(1, kappa // 2), m // 2
    v = key.view(k2, m2)
    W2 = np.ascontiguousarray(W[:, :m2])
    got2 = ctx.parallel_dot_series_matrix(v, ctx.to_device(W2)).download()
    assert np.array_equal(got2, orc.commit(np.ascontiguousarray(A[:k2, :m2]), W2, Q, threads=8))
    v.close()
    ref_key.close()
    key.close()


@pytest.mark.parametrize("kappa,m,r", [(4, 200, 3), (13, 4096, 1), (16, 1024, 8), (36, 256, 20)])
def test_crs_key_generated_in_commit(sa, ctx, orc, kappa, m, r):
    # the CRS key kept as its module only, the rows made inside the commit
    # kernel: commit, split, decomp_ell and a prefix view equal the stored CRS
    # key's and the oracle's
    say("case: implicit CRS key kappa=%d m=%d r=%d" % (kappa, m, r))
    module = rand_ring(orc, (kappa, 64), 90 + kappa, 0xC)
    W = rand_ring(orc, (r, m, 64), 91, 0xB, 16)
    dW, dmod = ctx.to_device(W), ctx.to_device(module)
    imp = ctx.crs_key(dmod, m, implicit=True)
    got = ctx.parallel_dot_series_matrix(imp, dW).download()
    A = orc.compute_commitment_keys(module, m, Q)
    assert np.array_equal(got, orc.commit(A, W, Q, threads=8))
    stored = ctx.crs_key(dmod, m)
    assert np.array_equal(ctx.split(imp, dW).download(), ctx.split(stored, dW).download())
    k2, m2 = max(1, kappa // 3), max(1, m // 3)
    v = imp.view(k2, m2)
    W2 = np.ascontiguousarray(W[:, :m2])
    got2 = ctx.parallel_dot_series_matrix(v, ctx.to_device(W2)).download()
    assert np.array_equal(got2, orc.commit(np.ascontiguousarray(A[:k2, :m2]), W2, Q, threads=8))
    v.close()
    nw_i, rhs_i, _ = ctx.decomp_ell(imp, dW, 2, 0)
    nw_s, rhs_s, _ = ctx.decomp_ell(stored, dW, 2, 0)
    assert np.array_equal(nw_i.download(), nw_s.download())
    assert np.array_equal(rhs_i.download(), rhs_s.download())
    stored.close()
    imp.close()


def test_verifier_split_identity_on_crs_key(sa, ctx, orc):
    # verifier_split (split.rs:187-211): under a power-series key, the split
    # halves' commitments L, R and the layer's tensor [1, m^c] (c = m / 2)
    # recombine to the commitment of the whole witness, L + m^c R, row by row
    say("case: verifier_split identity")
    kappa, r, m = 3, 2, 256
    c = m // 2
    module = rand_ring(orc, (kappa, 64), 81, 0xC)
    W = rand_ring(orc, (r, m, 64), 82, 0xB, 16)
    dW = ctx.to_device(W)
    key = ctx.crs_key(ctx.to_device(module), m)
    full = ctx.parallel_dot_series_matrix(key, dW).download()
    lr = ctx.split(key, dW)                       # [kappa][2r][d]: L rows, then R rows
    R = ctx.to_device(np.ascontiguousarray(lr.download()[:, r:]))
    ctx.convert(R, sa.REP_INCOMPLETE, sa.REP_COEFF)
    mult = orc.compute_commitment_keys(module, c, Q)[:, c - 1]   # m_j^c = row_j[c - 1]
    mrep = ctx.to_device(np.repeat(mult[:, None], r, axis=1))
    prod = ctx.ring_mul(mrep, R).download()
    L = lr.download()[:, :r]
    assert np.array_equal((L + prod) % np.uint64(Q), full)
    # the identity fails for the other tensor entry (a plausible slip)
    assert not np.array_equal((L + R.download()) % np.uint64(Q), full)
    key.close()


def test_fold_homomorphism_on_device(sa, ctx, orc):
    # test_fold (fold.rs:36-55) on the device: with a ternary challenge
    # (challenge_for_fold, fold.rs:68-73), committing the folded witness gives
    # the challenge-combined commitment (verifier_fold, fold.rs:57-66)
    say("case: fold homomorphism")
    kappa, r_in, r_out, m = 4, 3, 2, 128
    A = rand_ring(orc, (kappa, m, 64), 83, 0xA)
    W = rand_ring(orc, (r_in, m, 64), 84, 0xB)
    key = ctx.key(ctx.to_device(A))
    dW = ctx.to_device(W)
    com = ctx.parallel_dot_series_matrix(key, dW)             # [kappa][r_in][d]
    C = ctx.bounded((r_out, r_in, 64), 1, 85, 0xF)            # {-1, 0, 1} coefficients
    folded = ctx.fold(dW, C)                                  # [r_out][m][d] IncompleteNTT
    ctx.convert(folded, sa.REP_INCOMPLETE, sa.REP_COEFF)
    lhs = ctx.parallel_dot_series_matrix(key, folded).download()
    ctx.convert(com, sa.REP_INCOMPLETE, sa.REP_COEFF)
    comh, Ch = com.download(), C.download()
    a = np.ascontiguousarray(np.broadcast_to(Ch[None], (kappa, r_out, r_in, 64)))
    b = np.ascontiguousarray(np.broadcast_to(comh[:, None], (kappa, r_out, r_in, 64)))
    prod = ctx.ring_mul(ctx.to_device(a), ctx.to_device(b)).download()
    rhs = np.zeros((kappa, r_out, 64), np.uint64)
    for k in range(r_in):
        rhs = (rhs + prod[:, :, k]) % np.uint64(Q)
    assert np.array_equal(lhs, rhs)
    key.close()


# ------------------------------------------------------------------ projection
@pytest.mark.parametrize("r,m", [(1, 256), (2, 1536), (3, 1536), (4, 2048), (5, 2560)])
def test_project_vs_oracle(ctx, orc, r, m):
    # m = 1536 at r = 2: the last 512 elements of each row are outside every chunk
    say("case: project r=%d m=%d" % (r, m))
    rng = np.random.default_rng(7 * r + m)
    W = rand_ring(orc, (r, m, 64), 48, 0xE)
    z = rng.integers(0, 2, (256, r * 256), dtype=np.uint8)
    sg = rng.integers(0, 2, (256, r * 256), dtype=np.uint8)
    got = ctx.project(ctx.to_device(W), ctx.to_device(z, np.uint8), ctx.to_device(sg, np.uint8))
    assert np.array_equal(got.download(), orc.project(W, z, sg, Q))


def test_project_extremes_and_checks(sa, ctx, orc):
    # the largest part sums (every residue q - 1, every sign +1: cw (q - 1) in
    # u64), all zero, all -1; and the reference's chunk-width precondition
    say("case: project extremes")
    r, m = 4, 1024
    W = np.full((r, m, 64), Q - 1, np.uint64)
    dW = ctx.to_device(W)
    for zv, sv in ((0, 1), (1, 1), (0, 0)):
        z = np.full((256, r * 256), zv, np.uint8)
        sg = np.full((256, r * 256), sv, np.uint8)
        got = ctx.project(dW, ctx.to_device(z, np.uint8), ctx.to_device(sg, np.uint8)).download()
        assert np.array_equal(got, orc.project(W, z, sg, Q)), (zv, sv)
    # the balanced extremes (q - 1) / 2 and (q + 1) / 2 (the largest digits),
    # 0 and 1, with every sign +1 and with random signs
    W2 = np.resize(np.array([(Q - 1) // 2, (Q + 1) // 2, 0, 1, Q - 1], np.uint64), W.shape)
    rng = np.random.default_rng(5)
    for z, sg in ((np.zeros((256, r * 256), np.uint8), np.ones((256, r * 256), np.uint8)),
                  (rng.integers(0, 2, (256, r * 256), dtype=np.uint8),
                   rng.integers(0, 2, (256, r * 256), dtype=np.uint8))):
        got = ctx.project(ctx.to_device(W2), ctx.to_device(z, np.uint8),
                          ctx.to_device(sg, np.uint8)).download()
        assert np.array_equal(got, orc.project(W2, z, sg, Q))
    z = np.zeros((256, 2 * 256), np.uint8)
    with pytest.raises(sa.SalsaaError):
        ctx.project(ctx.to_device(np.zeros((2, 256, 64), np.uint64)),
                    ctx.to_device(z, np.uint8), ctx.to_device(z, np.uint8))


@pytest.mark.parametrize("r,m,unit", [(1, 256, False), (2, 2048, True), (4, 4096, False)])
def test_batch_projections_vs_oracle(ctx, orc, r, m, unit):
    # batch_projections (project.rs:131-203) and the cross terms before the
    # join (:229-245) on the projection of W; c0 = 1 is the reference's own
    # batching challenge, a random c0 the general case
    say("case: batch_projections r=%d m=%d unit=%d" % (r, m, unit))
    rng = np.random.default_rng(100 + r)
    W = rng.integers(0, Q, (r, m, 64), dtype=np.uint64)
    z = rng.integers(0, 2, (256, r * 256), dtype=np.uint8)
    sg = rng.integers(0, 2, (256, r * 256), dtype=np.uint8)
    c0 = np.zeros(64, np.uint64)
    if unit:
        c0[0] = 1
    else:
        c0 = rng.integers(0, Q, 64, dtype=np.uint64)
    dW, dz, dsg = ctx.to_device(W), ctx.to_device(z, np.uint8), ctx.to_device(sg, np.uint8)
    dP = ctx.project(dW, dz, dsg)
    P = dP.download()
    assert np.array_equal(P, orc.project(W, z, sg, Q))
    plhs, wlhs, prhs, wrhs, ok = ctx.batch_projections(dP, dW, dz, dsg, ctx.to_device(c0))
    w_plhs, w_wlhs, w_prhs, w_wrhs, w_ok = orc.batch_projections(P, W, z, sg, c0, Q)
    assert w_ok and ok.download()[0] == 1
    assert np.array_equal(plhs.download(), w_plhs)
    assert np.array_equal(wlhs.download(), w_wlhs)
    assert np.array_equal(prhs.download(), w_prhs)
    assert np.array_equal(wrhs.download(), w_wrhs)
    # the right-hand sides alone (layers not written) are the same
    _, _, prhs2, wrhs2, ok2 = ctx.batch_projections(dP, dW, dz, dsg, ctx.to_device(c0),
                                                    layers=False)
    assert np.array_equal(prhs2.download(), w_prhs) and np.array_equal(wrhs2.download(), w_wrhs)
    assert ok2.download()[0] == 1
    # cross terms against a second witness of the same shape (the folded one)
    Wf = rng.integers(0, Q, (2, m, 64), dtype=np.uint64)
    l0, l1 = ctx.send_cross_terms_before_join(plhs, wlhs, ctx.to_device(Wf), dP)
    w_l0, w_l1 = orc.send_cross_terms_before_join(w_plhs, w_wlhs, Wf, P, Q)
    assert np.array_equal(l0.download(), w_l0) and np.array_equal(l1.download(), w_l1)


def test_batch_projections_mismatch_and_checks(sa, ctx, orc):
    # a projection that is not the witness's fails the assertion (ok = 0, the
    # outputs still the oracle's); shapes with no matching layer are refused
    say("case: batch_projections mismatch")
    r, m = 2, 1024
    rng = np.random.default_rng(7)
    W = rng.integers(0, Q, (r, m, 64), dtype=np.uint64)
    z = rng.integers(0, 2, (256, r * 256), dtype=np.uint8)
    sg = rng.integers(0, 2, (256, r * 256), dtype=np.uint8)
    c0 = rng.integers(0, Q, 64, dtype=np.uint64)
    P = orc.project(W, z, sg, Q)
    P[17, 3] = (P[17, 3] + 1) % Q
    dz, dsg = ctx.to_device(z, np.uint8), ctx.to_device(sg, np.uint8)
    _, _, prhs, wrhs, ok = ctx.batch_projections(ctx.to_device(P), ctx.to_device(W), dz, dsg,
                                                 ctx.to_device(c0), layers=False)
    w = orc.batch_projections(P, W, z, sg, c0, Q)
    assert not w[4] and ok.download()[0] == 0
    assert np.array_equal(prhs.download(), w[2]) and np.array_equal(wrhs.download(), w[3])
    for rr, mm in ((3, 768 * 2), (2, 512 * 3)):
        zz = np.zeros((256, rr * 256), np.uint8)
        with pytest.raises(sa.SalsaaError):
            ctx.batch_projections(ctx.to_device(np.zeros((mm, 64), np.uint64)),
                                  ctx.to_device(np.zeros((rr, mm, 64), np.uint64)),
                                  ctx.to_device(zz, np.uint8), ctx.to_device(zz, np.uint8),
                                  ctx.to_device(c0), layers=False)


# ------------------------------------------------------------------ norm check
def _oracle_cs(reps, nrounds, fq2, seed):
    from oracle import oracle as orc
    cs = np.stack([orc.sample_uniform(seed, 0x5C000 + rep, (nrounds * 2,), Q).reshape(nrounds, 2)
                   for rep in range(reps)])
    if not fq2:
        cs[:, :, 1] = 0
    return cs


@pytest.mark.parametrize("fq2", [0, 1])
@pytest.mark.parametrize("rows,L,reps", [(2, 64, 1), (3, 1024, 2), (2, 2048, 2), (3, 1 << 14, 3),
                                         (1, 1 << 16, 2), (2, 1 << 13, 1)])
def test_norm_check_vs_oracle(ctx, orc, fq2, rows, L, reps):
    say("case: norm_check fq2=%d rows=%d L=%d reps=%d" % (fq2, rows, L, reps))
    n = L.bit_length() - 1
    w = rand_ring(orc, (rows, L), 9, 0xB, 16)
    cs = ctx.challenges(reps, n, fq2, 77)
    assert np.array_equal(cs.download(), _oracle_cs(reps, n, fq2, 77))
    outs = ctx.norm_check(ctx.to_device(w), cs, fq2)
    want = orc.sumcheck(w, cs.download(), Q, fq2=bool(fq2), nr=NR, threads=8)
    assert np.array_equal(outs["claims"].download(), want["claims"])
    assert np.array_equal(outs["msgs"].download(), want["msgs"])
    assert np.array_equal(outs["finals"].download(), want["finals"])
    assert np.array_equal(outs["evals"].download(), want["evals"])
    assert outs["ok"].download().all() and want["ok"].all()
    rhs = outs["rhs"].download()
    m = L // 64
    W3 = w.reshape(rows, m, 64)
    csd = cs.download()
    for rep in range(reps):
        if fq2:
            want_rhs = orc.norm_check_3_fq2(W3, csd[rep], Q, NR)
            assert np.array_equal(rhs[rep], want_rhs)
        else:
            want_rhs = orc.norm_check_3(W3, csd[rep, :, 0], Q)
            assert np.array_equal(rhs[rep, :, 0], want_rhs)
            assert not rhs[rep, :, 1].any()


@pytest.mark.parametrize("reps,nr", [(6, NR), (2, 20)])
def test_norm_check_reps_and_nonresidue(ctx, orc, reps, nr):
    # more repetitions than one launch of the first pass folds (the rest run
    # fold-only), and a non-residue of 8 or more (the general F_{q^2} product)
    assert pow(nr, (Q - 1) // 2, Q) == Q - 1
    rows, L = 2, 1 << 13
    say("case: norm_check reps=%d nr=%d rows=%d L=%d" % (reps, nr, rows, L))
    n = L.bit_length() - 1
    w = rand_ring(orc, (rows, L), 21, 0xB)
    cs = ctx.challenges(reps, n, 1, 78)
    outs = ctx.norm_check(ctx.to_device(w), cs, 1, nr=nr)
    want = orc.sumcheck(w, cs.download(), Q, fq2=True, nr=nr, threads=8)
    for k in ("claims", "msgs", "finals", "evals"):
        assert np.array_equal(outs[k].download(), want[k]), k
    assert outs["ok"].download().all() and want["ok"].all()
    rhs, csd = outs["rhs"].download(), cs.download()
    W3 = w.reshape(rows, L // 64, 64)
    for rep in range(reps):
        assert np.array_equal(rhs[rep], orc.norm_check_3_fq2(W3, csd[rep], Q, nr))


@pytest.mark.parametrize("fq2,L", [(1, 1 << 16), (0, 1 << 16), (1, 2048)])
def test_norm_check_mixed_magnitudes(ctx, orc, fq2, L):
    # The first pass takes a short path for a wave (64 columns) whose residues
    # are all below 2^14 in balanced form.  Rows here mix waves of short
    # values, of the extreme short values 2^14 - 1 and -2^14, of short values
    # with one residue just past the bound (2^14 or -2^14 - 1), and of uniform
    # residues; at L = 2048 half of the block's lanes are past the row's end.
    rows, reps = 2, 2
    say("case: norm_check mixed magnitudes fq2=%d L=%d" % (fq2, L))
    n = L.bit_length() - 1
    S = L // 16
    w = rand_ring(orc, (rows, L), 31, 0xB, 16)
    uni = rand_ring(orc, (rows, L), 32, 0xB)
    W, U = w.reshape(rows, 16, S), uni.reshape(rows, 16, S)
    pick = np.random.default_rng(5).random((rows, 16, S)) < 0.5
    for k in range(S // 64):
        c = slice(64 * k, 64 * k + 64)
        kind = k % 5 if S > 128 else (0, 4)[k]
        if kind == 1:
            W[:, :, c] = np.where(pick[:, :, c], (1 << 14) - 1, Q - (1 << 14))
        elif kind == 2:
            W[:, 3, 64 * k + 17] = 1 << 14
        elif kind == 3:
            W[:, 15, 64 * k + 63] = Q - (1 << 14) - 1
        elif kind == 4:
            W[:, :, c] = U[:, :, c]
    cs = ctx.challenges(reps, n, fq2, 79)
    outs = ctx.norm_check(ctx.to_device(w), cs, fq2)
    want = orc.sumcheck(w, cs.download(), Q, fq2=bool(fq2), nr=NR, threads=8)
    for k in ("claims", "msgs", "finals", "evals"):
        assert np.array_equal(outs[k].download(), want[k]), k
    assert outs["ok"].download().all() and want["ok"].all()
    rhs, csd = outs["rhs"].download(), cs.download()
    W3 = w.reshape(rows, L // 64, 64)
    for rep in range(reps):
        if fq2:
            assert np.array_equal(rhs[rep], orc.norm_check_3_fq2(W3, csd[rep], Q, NR))
        else:
            assert np.array_equal(rhs[rep, :, 0], orc.norm_check_3(W3, csd[rep, :, 0], Q))


def test_norm_check_row_of_2_30(sa, ctx):
    # one row of 2^30 coefficients: more first-pass blocks than one u64 Gram
    # sum can take at the default columns per thread.  A uniform row passes
    # the verifier's checks; a constant row c has claim L c^2, evaluation c
    # (the tensor weights sum to 1) and final claim c^2, for F_q and F_{q^2}
    L = 1 << 30
    say("case: norm_check one row of L=2^30")
    w = ctx.uniform((1, L), 41, 0xB)
    for fq2 in (1, 0):
        cs = ctx.challenges(1, 30, fq2, 91)
        outs = ctx.norm_check(w, cs, fq2)
        assert outs["ok"].download().all(), fq2
    c = 123456789
    chunk = ctx.to_device(np.full(1 << 26, c, np.uint64))
    for k in range(L >> 26):
        part = sa.DeviceArray.view(ctx, w.ptr + (k << 26) * 8, (1 << 26,))
        ctx.copy(part, chunk)
    for fq2 in (1, 0):
        cs = ctx.challenges(2, 30, fq2, 92)
        outs = ctx.norm_check(w, cs, fq2)
        assert outs["ok"].download().all(), fq2
        assert int(outs["claims"].download()[0]) == L * c * c % Q
        ev, fin = outs["evals"].download(), outs["finals"].download()
        for rep in range(2):
            assert (int(ev[rep, 0, 0]), int(ev[rep, 0, 1])) == (c, 0)
            assert (int(fin[rep, 0, 0]), int(fin[rep, 0, 1])) == (c * c % Q, 0)
    w.free()
    chunk.free()


def test_norm_check_small_lengths(ctx, orc):
    for L in (2, 4, 32):
        say("case: norm_check L=%d (no rhs)" % L)
        n = L.bit_length() - 1
        w = rand_ring(orc, (2, L), 1, 2)
        cs = ctx.challenges(2, n, 1, 5)
        outs = ctx.norm_check(ctx.to_device(w), cs, 1, want_rhs=False)
        want = orc.sumcheck(w, cs.download(), Q, fq2=True, nr=NR)
        assert np.array_equal(outs["msgs"].download(), want["msgs"])
        assert np.array_equal(outs["finals"].download(), want["finals"])
        assert outs["ok"].download().all()


def test_norm_check_rejects_wrong_claim_path(ctx, orc):
    # a witness row and its negation have equal claims and messages: the
    # verifier replay must accept both (sign symmetry of the square)
    say("case: norm_check sign symmetry")
    L = 4096
    w = rand_ring(orc, (1, L), 4, 4)
    wn = np.where(w == 0, w, np.uint64(Q) - w)
    cs = ctx.challenges(1, 12, 0, 3)
    a = ctx.norm_check(ctx.to_device(w), cs, 0)
    b = ctx.norm_check(ctx.to_device(wn), cs, 0)
    assert np.array_equal(a["msgs"].download(), b["msgs"].download())
    assert a["ok"].download().all() and b["ok"].download().all()


# ------------------------------------------------------------------ decomposition
@pytest.mark.parametrize("r,m,parts,bound,radix", [(8, 32, 4, 4096, 0), (8, 16, 6, 16, 256),
                                                   (2, 8, 2, 3, 0), (3, 5, 3, 1 << 20, 0)])
def test_decomp_ell_vs_oracle(ctx, orc, r, m, parts, bound, radix):
    say("case: decomp r=%d m=%d parts=%d radix=%d" % (r, m, parts, radix))
    kappa = 4
    A = rand_ring(orc, (kappa, m, 64), 2, 0xA)
    W = rand_ring(orc, (r, m, 64), 2, 0xB, bound)
    key = ctx.key(ctx.to_device(A))
    dW = ctx.to_device(W)
    new_W, rhs_dec, d_radix = ctx.decomp_ell(key, dW, parts, radix)
    want_limbs, want_radix = orc.decompose(W, parts, Q, radix)
    assert int(d_radix.download()[0]) == want_radix
    assert np.array_equal(new_W.download(), want_limbs)
    assert np.array_equal(rhs_dec.download(), orc.commit(A, want_limbs, Q, 8))
    rhs = ctx.parallel_dot_series_matrix(key, dW)
    mism = ctx.verify_decomp(rhs_dec, rhs, parts, d_radix)
    assert int(mism.download()[0]) == 0
    assert np.array_equal(ctx.compose(rhs_dec, parts, d_radix).download(), rhs.download())
    # witness recomposition (rows mode)
    mism2 = ctx.verify_decomp(new_W, dW, parts, d_radix, rows_mode=True)
    assert int(mism2.download()[0]) == 0
    # a corrupted limb commitment is counted
    bad = rhs_dec.download()
    bad[0, 0, 5] = (int(bad[0, 0, 5]) + 1) % Q
    mism3 = ctx.verify_decomp(ctx.to_device(bad), rhs, parts, d_radix)
    assert int(mism3.download()[0]) == 1
    key.close()


# ------------------------------------------------------------------ fold
@pytest.mark.parametrize("r_in,r_out,m", [(4, 2, 16), (32, 8, 64), (3, 5, 7)])
def test_fold_vs_oracle(ctx, orc, r_in, r_out, m):
    say("case: fold r_in=%d r_out=%d m=%d" % (r_in, r_out, m))
    W = rand_ring(orc, (r_in, m, 64), 6, 0xB, 4096)
    C = rand_ring(orc, (r_out, r_in, 64), 6, 0xF, 1)
    got = ctx.fold(ctx.to_device(W), ctx.to_device(C)).download()
    assert np.array_equal(got, orc.fold(W, C, Q, threads=8))


# ------------------------------------------------------------------ full-size properties
def test_cfg2_ntt_full_size_roundtrip(sa, ctx, orc):
    say("case: 2^20-element NTT round trip")
    n = 1 << 20
    dx = ctx.uniform((n, 64), 0, 0xC)
    ctx.convert(dx, sa.REP_COEFF, sa.REP_INCOMPLETE)
    inc = dx.download()
    x = orc.sample_uniform(0, 0xC, (n, 64), Q)
    idx = np.random.default_rng(0).choice(n, 64, replace=False)
    assert np.array_equal(inc[idx], orc.to_incomplete(x[idx], Q))
    ctx.convert(dx, sa.REP_INCOMPLETE, sa.REP_COEFF)
    assert np.array_equal(dx.download(), x)


def test_cfg2_norm_check_full_size(ctx, orc):
    say("case: 8 x 2^20 norm check, F_q^2, 2 repetitions")
    rows, L = 8, 1 << 20
    w = ctx.bounded((rows, L), 16, 0, 0xB)
    cs = ctx.challenges(2, 20, 1, 1)
    outs = ctx.norm_check(w, cs, 1)
    assert outs["ok"].download().all()
    # rows 0 and 5 in full against the oracle
    wh = w.download()
    want = orc.sumcheck(wh[[0, 5]], cs.download(), Q, fq2=True, nr=NR, threads=2)
    got_msgs = outs["msgs"].download()[:, :, [0, 5]]
    assert np.array_equal(got_msgs, want["msgs"])
    assert np.array_equal(outs["evals"].download()[:, [0, 5]], want["evals"])


def test_commit_valu_path_matches_oracle(sa, orc, monkeypatch):
    # the fp64 VALU commit (selected per context) against the oracle, and the
    # default int8-MFMA path against it on the same inputs
    say("case: commit VALU path")
    kappa, r, m = 8, 5, 300
    A = rand_ring(orc, (kappa, m, 64), 3, 0xA)
    W = rand_ring(orc, (r, m, 64), 3, 0xB, 16)
    want = orc.commit(A, W, Q, threads=8)
    monkeypatch.setenv("SALSAA_COMMIT", "valu")
    cv = sa.Context(Q, 64)
    kv = cv.key(cv.to_device(A))
    got_v = cv.parallel_dot_series_matrix(kv, cv.to_device(W)).download()
    limbs_v = cv.decomp_ell(kv, cv.to_device(W), 3, 0)[1].download()
    kv.close()
    cv.close()
    monkeypatch.delenv("SALSAA_COMMIT")
    cm = sa.Context(Q, 64)
    km = cm.key(cm.to_device(A))
    got_m = cm.parallel_dot_series_matrix(km, cm.to_device(W)).download()
    limbs_m = cm.decomp_ell(km, cm.to_device(W), 3, 0)[1].download()
    km.close()
    cm.close()
    assert np.array_equal(got_v, want)
    assert np.array_equal(got_m, want)
    assert np.array_equal(limbs_v, limbs_m)


@pytest.mark.parametrize("kappa,r,m", [(40, 3, 200), (32, 48, 130)])
def test_commit_large_tiles_vs_oracle(ctx, orc, kappa, r, m):
    # more than 36 key rows (several M blocks) and more than 8 witness rows
    say("case: commit tiles kappa=%d r=%d m=%d" % (kappa, r, m))
    A = rand_ring(orc, (kappa, m, 64), 8, 0xA)
    W = rand_ring(orc, (r, m, 64), 8, 0xB)
    key = ctx.key(ctx.to_device(A))
    got = ctx.parallel_dot_series_matrix(key, ctx.to_device(W)).download()
    assert np.array_equal(got, orc.commit(A, W, Q, threads=8))
    key.close()
