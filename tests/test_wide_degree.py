"""The ring at the reference's own degrees N = 128, 256, 512 with their moduli
(protocol.rs:28-47) through the C ABI, against the CPU oracle: representation
changes, the ring product, the commit (int8 matrix cores over the
IncompleteNTT's slot pairs), the norm check with norm_check_3's rhs,
decomposition with decomp_ell, and the fold.  Entry points whose kernels exist
at d = 64 only refuse such a context."""
import sys

import numpy as np
import pytest

from conftest import REF_RINGS

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sa():
    import salsaa_amd
    return salsaa_amd


@pytest.fixture(scope="module", params=REF_RINGS, ids=["d128", "d256", "d512"])
def ring(request, sa):
    d, q, nr = request.param
    c = sa.Context(q, d)
    yield c, d, q, nr
    c.close()


def uniform(orc, shape, q, seed, stream):
    return orc.sample_uniform(seed, stream, shape, q)


def bounded(orc, shape, q, seed, stream, bound):
    return orc.sample_bounded(seed, stream, shape, bound, q)


@pytest.mark.parametrize("count", [1, 33, 300])
def test_wide_ntt_vs_oracle(sa, ring, orc, count):
    ctx, d, q, _ = ring
    say("case: wide ntt d=%d count=%d" % (d, count))
    x = uniform(orc, (count, d), q, 1, 3)
    dx = ctx.to_device(x)
    ctx.convert(dx, sa.REP_COEFF, sa.REP_INCOMPLETE)
    assert np.array_equal(dx.download(), orc.to_incomplete(x, q))
    ctx.convert(dx, sa.REP_INCOMPLETE, sa.REP_COEFF)
    assert np.array_equal(dx.download(), x)
    if (q - 1) % (2 * d) == 0:
        ctx.convert(dx, sa.REP_COEFF, sa.REP_FULL)
        assert np.array_equal(dx.download(), orc.to_full(x, q))
        ctx.convert(dx, sa.REP_FULL, sa.REP_INCOMPLETE)  # via Coefficient
        assert np.array_equal(dx.download(), orc.to_incomplete(x, q))
        ctx.convert(dx, sa.REP_INCOMPLETE, sa.REP_COEFF)
        assert np.array_equal(dx.download(), x)
    else:
        # no primitive 2d-th root of unity mod q: the Full representation does not exist
        with pytest.raises(RuntimeError, match="2d"):
            ctx.convert(dx, sa.REP_COEFF, sa.REP_FULL)


def test_wide_ntt_edge_values(sa, ring, orc):
    ctx, d, q, _ = ring
    say("case: wide ntt edge values d=%d" % d)
    x = np.zeros((6, d), np.uint64)
    x[1] = q - 1
    x[2, 0] = 1
    x[3, d - 1] = q - 1
    x[4] = np.arange(d, dtype=np.uint64)
    x[5, 1::2] = (q - 1) // 2
    dx = ctx.to_device(x)
    ctx.convert(dx, sa.REP_COEFF, sa.REP_INCOMPLETE)
    assert np.array_equal(dx.download(), orc.to_incomplete(x, q))
    ctx.convert(dx, sa.REP_INCOMPLETE, sa.REP_COEFF)
    assert np.array_equal(dx.download(), x)


def test_wide_ntt_round_trip_large(sa, ring, orc):
    # 2^15 elements: the round trip restores the input; sampled elements
    # against the oracle
    ctx, d, q, _ = ring
    n = 1 << 15
    say("case: wide ntt round trip d=%d n=%d" % (d, n))
    dx = ctx.uniform((n, d), 5, 7)
    x = dx.download()
    ctx.convert(dx, sa.REP_COEFF, sa.REP_INCOMPLETE)
    inc = dx.download()
    for i in (0, 1, n // 3, n - 1):
        assert np.array_equal(inc[i], orc.to_incomplete(x[i:i + 1], q)[0])
    ctx.convert(dx, sa.REP_INCOMPLETE, sa.REP_COEFF)
    assert np.array_equal(dx.download(), x)


def test_wide_ring_mul_vs_oracle(ring, orc):
    ctx, d, q, _ = ring
    say("case: wide ring_mul d=%d" % d)
    a = uniform(orc, (40, d), q, 1, 1)
    b = uniform(orc, (40, d), q, 1, 2)
    out = ctx.ring_mul(ctx.to_device(a), ctx.to_device(b)).download()
    for i in range(0, 40, 7):
        assert np.array_equal(out[i], orc.ring_mul_incomplete(a[i], b[i], q))


@pytest.mark.parametrize("kappa,r,m,bound", [(4, 3, 40, None), (12, 8, 256, 16),
                                             (13, 1, 200, 8), (7, 16, 64, None),
                                             (18, 5, 1, 16)])
def test_wide_commit_vs_oracle(ring, orc, kappa, r, m, bound):
    ctx, d, q, _ = ring
    say("case: wide commit d=%d kappa=%d r=%d m=%d" % (d, kappa, r, m))
    A = uniform(orc, (kappa, m, d), q, 2, 0xA)
    W = uniform(orc, (r, m, d), q, 2, 0xB) if bound is None else bounded(orc, (r, m, d), q, 2,
                                                                           0xB, bound)
    key = ctx.key(ctx.to_device(A))
    got = ctx.parallel_dot_series_matrix(key, ctx.to_device(W)).download()
    assert np.array_equal(got, orc.commit(A, W, q, threads=8))
    if m >= 2 and kappa >= 2:
        # a view: the key's first kappa - 1 rows and half its length
        k2, m2 = kappa - 1, m // 2
        v = key.view(k2, m2)
        W2 = np.ascontiguousarray(W[:, :m2])
        got2 = ctx.parallel_dot_series_matrix(v, ctx.to_device(W2)).download()
        want2 = orc.commit(np.ascontiguousarray(A[:k2, :m2]), W2, q, threads=8)
        assert np.array_equal(got2, want2)
        v.close()
    key.close()


def test_wide_commit_long_witness(ring, orc):
    # 2^17 ring elements, the commit's K range split into chunks: a witness
    # whose nonzero elements are the first and the last 128 commits like the
    # oracle's commit of those 256 elements under the matching key columns
    ctx, d, q, _ = ring
    kappa = {128: 12, 256: 7, 512: 4}[d]
    m, e = 1 << 17, 128
    say("case: wide commit long witness d=%d kappa=%d m=%d" % (d, kappa, m))
    dA = ctx.uniform((kappa, m, d), 3, 0xA)
    key = ctx.key(dA)
    dA.free()
    W = np.zeros((1, m, d), np.uint64)
    W[0, :e] = bounded(orc, (e, d), q, 3, 0xB, 2)
    W[0, m - e:] = bounded(orc, (e, d), q, 4, 0xB, 2)
    got = ctx.parallel_dot_series_matrix(key, ctx.to_device(W)).download()
    A_sel = np.zeros((kappa, 2 * e, d), np.uint64)
    for j in range(kappa):
        A_sel[j, :e] = orc.sample_uniform_at(3, 0xA, j * m * d, (e, d), q)
        A_sel[j, e:] = orc.sample_uniform_at(3, 0xA, (j * m + m - e) * d, (e, d), q)
    W_sel = np.concatenate([W[:, :e], W[:, m - e:]], axis=1)
    assert np.array_equal(got, orc.commit(A_sel, W_sel, q, threads=8))
    key.close()


def test_wide_commit_kappa_limit(ring, orc):
    ctx, d, q, _ = ring
    A = ctx.uniform((19, 8, d), 1, 1)
    with pytest.raises(RuntimeError, match="kappa <= 18"):
        ctx.key(A)


@pytest.mark.parametrize("fq2,m", [(1, 1), (0, 64), (1, 64), (1, 4096)])
def test_wide_norm_check_vs_oracle(ring, orc, fq2, m):
    ctx, d, q, nr = ring
    rows, reps = 2, 2
    L = m * d
    n = L.bit_length() - 1
    say("case: wide norm_check d=%d fq2=%d m=%d L=%d" % (d, fq2, m, L))
    w = bounded(orc, (rows, L), q, 9, 0xB, 16)
    cs = ctx.challenges(reps, n, fq2, 77)
    csd = cs.download()
    outs = ctx.norm_check(ctx.to_device(w), cs, fq2, nr=nr)
    want = orc.sumcheck(w, csd, q, fq2=bool(fq2), nr=nr, threads=8)
    for k in ("claims", "msgs", "finals", "evals"):
        assert np.array_equal(outs[k].download(), want[k]), k
    assert outs["ok"].download().all() and want["ok"].all()
    rhs = outs["rhs"].download()
    W3 = w.reshape(rows, m, d)
    for rep in range(reps):
        if fq2:
            assert np.array_equal(rhs[rep], orc.norm_check_3_fq2(W3, csd[rep], q, nr))
        else:
            assert np.array_equal(rhs[rep, :, 0], orc.norm_check_3(W3, csd[rep, :, 0], q))
            assert not rhs[rep, :, 1].any()


def test_wide_decomp_ell_vs_oracle(ring, orc):
    ctx, d, q, _ = ring
    kappa, r, m, parts = 6, 3, 96, 4
    say("case: wide decomp_ell d=%d" % d)
    A = uniform(orc, (kappa, m, d), q, 4, 0xA)
    W = bounded(orc, (r, m, d), q, 4, 0xB, 1 << 20)
    key = ctx.key(ctx.to_device(A))
    dW = ctx.to_device(W)
    new_W, rhs, d_radix = ctx.decomp_ell(key, dW, parts)
    want_W, radix = orc.decompose(W, parts, q)
    assert int(d_radix.download()[0]) == radix
    assert np.array_equal(new_W.download(), want_W)
    assert np.array_equal(rhs.download(), orc.commit(A, want_W, q, threads=8))
    # the limb commitments compose to W's commitment, the limbs to W
    com = ctx.parallel_dot_series_matrix(key, dW)
    assert int(ctx.verify_decomp(rhs, com, parts, d_radix).download()[0]) == 0
    assert np.array_equal(ctx.compose(rhs, parts, d_radix).download(), com.download())
    assert int(ctx.verify_decomp(new_W, dW, parts, d_radix, rows_mode=True).download()[0]) == 0
    bad = rhs.download()
    bad[0, 0, d - 1] = (int(bad[0, 0, d - 1]) + 1) % q
    assert int(ctx.verify_decomp(ctx.to_device(bad), com, parts, d_radix).download()[0]) == 1
    # decompose alone, explicit radix
    out, dr = ctx.decompose(dW, parts, radix=1 << 6)
    want2, _ = orc.decompose(W, parts, q, radix=1 << 6)
    assert np.array_equal(out.download(), want2)
    key.close()


def test_wide_fold_vs_oracle(ring, orc):
    ctx, d, q, _ = ring
    r_in, r_out, m = 3, 2, 50
    say("case: wide fold d=%d" % d)
    W = bounded(orc, (r_in, m, d), q, 6, 0xB, 1 << 12)
    C = bounded(orc, (r_out, r_in, d), q, 6, 0xF, 1)
    got = ctx.fold(ctx.to_device(W), ctx.to_device(C)).download()
    assert np.array_equal(got, orc.fold(W, C, q, threads=8))


def test_wide_refuses_d64_only_entry_points(ring):
    ctx, d, q, _ = ring
    W = ctx.uniform((1, 256, d), 1, 1)
    isz = ctx.to_device(np.zeros((256, 256), np.uint8), np.uint8)
    with pytest.raises(RuntimeError, match="d = 64 only"):
        ctx.project(W, isz, isz)
    with pytest.raises(RuntimeError, match="d = 64 only"):
        ctx.crs_key(ctx.uniform((2, d), 1, 2), 8, implicit=True)
