"""The benchmark's CPU baseline (oracle/cpu_baseline.cpp) computes what the
oracle computes, bit for bit, in its AVX-512 IFMA lanes and in its scalar
variant: IncompleteNTT forms, the NTT round trip, the commit (key prefixes,
lengths that are not a multiple of the lane count), the F_{q^2} norm check
with norm_check_3 (chunk boundaries off the lane grid) and the
decomposition with verify_decomp."""
import numpy as np
import pytest

from conftest import NR, Q


@pytest.fixture(scope="module")
def cb():
    from oracle import cpu_baseline
    return cpu_baseline


@pytest.fixture(params=[True, False], ids=["ifma", "scalar"])
def variant(cb, request):
    cb.init(4, request.param)
    yield request.param
    cb.init(0, True)


def test_incomplete_and_roundtrip(cb, orc, variant):
    x = orc.sample_uniform(1, 2, (40, 64), Q)
    assert np.array_equal(cb.to_incomplete(x, Q), orc.to_incomplete(x, Q))
    y = x.copy()
    cb.ntt_roundtrip(y, Q)
    assert np.array_equal(y, x)


@pytest.mark.parametrize("kappa,r,m,use_kappa,use_m", [(5, 3, 64, 5, 64), (4, 2, 24, 3, 16),
                                                       (3, 2, 20, 3, 20)])
def test_commit(cb, orc, variant, kappa, r, m, use_kappa, use_m):
    A = orc.sample_uniform(3, 0xA, (kappa, m, 64), Q)
    W = orc.sample_bounded(3, 0xB, (r, use_m, 64), 16, Q)
    key = cb.Key(A, Q)
    got = cb.commit(key, W, use_kappa)
    want = orc.commit(np.ascontiguousarray(A[:use_kappa, :use_m]), W, Q, threads=4)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("rows,L,reps,bound", [(2, 2048, 2, 16), (1, 1 << 14, 1, None),
                                               (3, 1 << 12, 2, 1 << 12)])
def test_norm_check(cb, orc, variant, rows, L, reps, bound):
    w = (orc.sample_uniform(4, 0xB, (rows, L), Q) if bound is None
         else orc.sample_bounded(4, 0xB, (rows, L), bound, Q))
    n = L.bit_length() - 1
    cs = np.stack([orc.sample_uniform(5, 0x5C000 + rep, (2 * n,), Q).reshape(n, 2)
                   for rep in range(reps)])
    got = cb.norm_check(w, cs, Q, NR)
    want = orc.sumcheck(w, cs, Q, fq2=True, nr=NR, threads=4)
    for k in ("claims", "msgs", "finals", "evals", "ok"):
        assert np.array_equal(got[k], want[k]), k
    assert got["ok"].all()
    for rep in range(reps):
        want_rhs = orc.norm_check_3_fq2(w.reshape(rows, L // 64, 64), cs[rep], Q, NR)
        assert np.array_equal(got["rhs"][rep], want_rhs)


@pytest.mark.parametrize("parts,radix,bound", [(6, 256, 16), (4, 0, 1 << 12), (3, 0, 5)])
def test_decompose_and_verify(cb, orc, variant, parts, radix, bound):
    W = orc.sample_bounded(6, 0xB, (3, 16, 64), bound, Q)
    got, gr = cb.decompose(W, parts, Q, radix)
    want, wr = orc.decompose(W, parts, Q, radix)
    assert gr == wr and np.array_equal(got, want)
    A = orc.sample_uniform(6, 0xA, (2, 16, 64), Q)
    key = cb.Key(A, Q)
    rhs, rhs_dec = cb.commit(key, W), cb.commit(key, got)
    assert cb.verify_decomp(rhs_dec, rhs, parts, gr, Q) == 0
    rhs_dec[1, 2, 3] ^= 1
    assert cb.verify_decomp(rhs_dec, rhs, parts, gr, Q) == 1
