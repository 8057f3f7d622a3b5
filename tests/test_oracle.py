"""Pins the CPU oracle against the reference's known-answer tests, closed forms,
invariants and brute force on small inputs.  Each check is chosen so that a
plausible mistake (wrong root, wrong bit order, dropped term, wrong sign,
transposed operand) fails it."""
import random

import numpy as np
import pytest

from conftest import D, NR, Q, REF_RINGS
import pyref

Q62 = 4546383823830515713  # reference test modulus (cyclotomic_ring.rs:612)


def rnd(rng, n, q):
    return [rng.randrange(q) for _ in range(n)]


# ---------------------------------------------------------------- scalar KATs
def test_multiply_mod_kat(orc):
    # hexl/bindings.rs:99-106
    assert orc.mulmod(12345, 67890, 1000000007) == 838102050


def test_eltwise_mult_mod_kat(orc):
    # hexl/bindings.rs:108-118
    got = [orc.mulmod(a, b, 100) for a, b in zip([1, 2, 3, 4, 5], [6, 7, 8, 9, 10])]
    assert got == [6, 14, 24, 36, 50]


def test_mulmod_matches_int128(orc):
    rng = random.Random(1)
    for q in (Q, Q62, 65537, 17, 1000000007, (1 << 61) - 1):
        for _ in range(2000):
            a, b = rng.randrange(q), rng.randrange(q)
            assert orc.mulmod(a, b, q) == a * b % q
    # extremes
    for q in (Q, Q62):
        for a in (0, 1, q - 1, q - 2):
            for b in (0, 1, q - 1):
                assert orc.mulmod(a, b, q) == a * b % q


def test_sum_sq_kat(orc):
    # fast_norm_check.rs:264-269
    assert orc.sum_sq_chunked_raw(np.arange(1, 9, dtype=np.uint64), 71) == 62


def test_tensor_kat(orc):
    # fast_norm_check.rs:41-57
    assert orc.tensor([3, 5], 17).tolist() == [8, 7, 5, 15]


def test_tensor_matches_eq_weights(orc):
    rng = random.Random(2)
    cs = rnd(rng, 6, Q)
    assert orc.tensor(cs, Q).tolist() == pyref.eq_weights(cs, Q)


def test_quadratic_extension_kat(orc):
    # quadratic_extension.rs:89-111 (shift 1)
    assert orc.fq2_mul([3, 5], [7, 9], 1, 17) == [15, 11]


# ---------------------------------------------------------------- NTT
def test_minimal_roots_cfg_modulus(orc):
    # Independent check: every primitive 64th/128th root is psi0^odd for a
    # psi0 built from a primitive root found by sympy; HEXL takes the minimum.
    sympy = pytest.importorskip("sympy")
    g = int(sympy.primitive_root(Q))
    for order in (64, 128):
        psi0 = pow(g, (Q - 1) // order, Q)
        want = min(pow(psi0, k, Q) for k in range(1, order, 2))
        assert orc.lib.or_min_primitive_root(order, Q) == want
    assert orc.lib.or_min_primitive_root(64, Q) == 20344578268491
    assert orc.lib.or_min_primitive_root(128, Q) == 66971972259682


def test_roots_table_bitreversed(orc):
    for n in (4, 8, 32, 64):
        psi = orc.lib.or_min_primitive_root(2 * n, Q)
        roots = orc.get_roots(n, Q).tolist()
        bits = n.bit_length() - 1
        for i in range(n):
            j = int(format(i, "0%db" % bits)[::-1], 2)
            assert roots[j] == pow(psi, i, Q)


@pytest.mark.parametrize("n,q", [(2, 17), (4, 17), (8, 65537), (32, Q), (64, Q), (16, Q62),
                                 (64, REF_RINGS[0][1]), (128, REF_RINGS[0][1]),
                                 (128, REF_RINGS[1][1]), (256, REF_RINGS[2][1])])
def test_ntt_matches_closed_form(orc, n, q):
    rng = random.Random(n)
    x = rnd(rng, n, q)
    fwd = orc.ntt_forward(x, q).tolist()
    psi = orc.lib.or_min_primitive_root(2 * n, q)
    bits = max(n.bit_length() - 1, 0)
    # evaluation at psi^(2 brv(j) + 1), in Python ints
    for j in range(n):
        b = int(format(j, "0%db" % bits)[::-1], 2) if bits else 0
        assert fwd[j] == pyref.eval_poly(x, pow(psi, 2 * b + 1, q), q)
    assert orc.ntt_forward_direct(x, q).tolist() == fwd
    assert orc.ntt_inverse(fwd, q).tolist() == x


def test_multiply_poly_kat(orc):
    # hexl/bindings.rs:119-148: NTT product of half-zero size-8 inputs equals
    # the schoolbook product.
    q, n = 65537, 8
    a = [1, 2, 3, 1, 0, 0, 0, 0]
    b = [8, 7, 6, 1, 0, 0, 0, 0]
    fa, fb = orc.ntt_forward(a, q), orc.ntt_forward(b, q)
    prod = orc.ntt_inverse([int(x) * int(y) % q for x, y in zip(fa, fb)], q).tolist()
    want = [0] * n
    for i in range(4):
        for j in range(4):
            want[i + j] = (want[i + j] + a[i] * b[j]) % q
    assert prod == want


def test_full_ntt_product_theorem(orc):
    rng = random.Random(3)
    for _ in range(5):
        a, b = rnd(rng, D, Q), rnd(rng, D, Q)
        fa, fb = orc.to_full(a, Q), orc.to_full(b, Q)
        c = orc.from_full(np.array([int(x) * int(y) % Q for x, y in zip(fa, fb)],
                                   dtype=np.uint64), Q).tolist()
        assert c == pyref.negacyclic_mul(a, b, Q)


# ---------------------------------------------------------------- ring
def test_ring_add_sub_kat():
    # cyclotomic_ring.rs:201-227 are plain eltwise add/sub mod 17
    assert [(a + b) % 17 for a, b in zip([1, 2, 3, 4], [4, 3, 2, 1])] == [5, 5, 5, 5]
    assert [(a - b) % 17 for a, b in zip([5, 6, 7, 8], [4, 3, 2, 1])] == [1, 3, 5, 7]


def test_multiplication_same_representation_kat(orc):
    # cyclotomic_ring.rs:229-242
    c = orc.ring_mul_incomplete([1, 2, 1, 0], [1, 1, 1, 0], 17)
    assert orc.from_incomplete(c, 17).tolist() == [0, 3, 4, 3]


@pytest.mark.parametrize("d,q", [(4, 17), (8, Q62), (64, Q), (32, Q62)] +
                         [(d, q) for d, q, _ in REF_RINGS])
def test_incomplete_product_is_negacyclic(orc, d, q):
    rng = random.Random(d)
    for _ in range(4):
        a, b = rnd(rng, d, q), rnd(rng, d, q)
        c = orc.ring_mul_incomplete(a, b, q)
        assert orc.from_incomplete(c, q).tolist() == pyref.negacyclic_mul(a, b, q)
        assert orc.ring_mul_naive(a, b, q).tolist() == pyref.negacyclic_mul(a, b, q)


def test_incomplete_representation_closed_form(orc):
    # slot j of IncompleteNTT holds (e(omega_j), o(omega_j)) with omega_j the
    # slot's root of Y^{d/2}+1 and omega = NTT_{d/2}(Y)
    rng = random.Random(4)
    a = rnd(rng, D, Q)
    inc = orc.to_incomplete(a, Q).tolist()
    om = orc.shift_factors(D, Q).tolist()
    h = D // 2
    for j in range(h):
        assert pow(om[j], h, Q) == Q - 1
        assert inc[j] == pyref.eval_poly(a[0::2], om[j], Q)
        assert inc[h + j] == pyref.eval_poly(a[1::2], om[j], Q)
    assert orc.from_incomplete(inc, Q).tolist() == a


def test_conjugate_inner_product(orc):
    rng = random.Random(5)
    a, b = rnd(rng, D, Q), rnd(rng, D, Q)
    prod = pyref.negacyclic_mul(orc.conjugate(a, Q).tolist(), b, Q)
    assert prod[0] == sum(x * y for x, y in zip(a, b)) % Q


# ---------------------------------------------------------------- commit
def const_ring(vals, d, q):
    out = np.zeros((len(vals), d), np.uint64)
    out[:, 0] = [v % q for v in vals]
    return out


def test_split_commit_kat(orc):
    # split.rs:109-143 (N = 8, q = Q62): <[1,2,..,128],[1..8]> = 1793;
    # halves commit to 49 and 109 under the layer [1,2,4,8].
    d = 8
    A = const_ring([1, 2, 4, 8, 16, 32, 64, 128], d, Q62)[None]
    W = const_ring(list(range(1, 9)), d, Q62)[None]
    rhs = orc.commit(A, W, Q62)
    want = orc.to_incomplete(const_ring([1793], d, Q62)[0], Q62)
    assert rhs[0, 0].tolist() == want.tolist()
    assert want.tolist() == [1793] * 4 + [0] * 4
    A2 = const_ring([1, 2, 4, 8], d, Q62)[None]
    Ws = np.stack([W[0, :4], W[0, 4:]])
    rhs2 = orc.commit(A2, Ws, Q62)
    assert orc.from_incomplete(rhs2[0, 0], Q62).tolist()[0] == 49
    assert orc.from_incomplete(rhs2[0, 1], Q62).tolist()[0] == 109


@pytest.mark.parametrize("kappa,r,m,d,q", [(2, 3, 4, 16, Q), (2, 2, 2, 128, REF_RINGS[0][1]),
                                           (1, 2, 2, 256, REF_RINGS[1][1])])
def test_commit_matches_schoolbook(orc, kappa, r, m, d, q):
    rng = random.Random(6)
    A = np.array(rnd(rng, kappa * m * d, q), np.uint64).reshape(kappa, m, d)
    W = np.array([rng.randrange(-8, 9) % q for _ in range(r * m * d)],
                 np.uint64).reshape(r, m, d)
    out = orc.commit(A, W, q)
    for j in range(kappa):
        for k in range(r):
            acc = [0] * d
            for i in range(m):
                p = pyref.negacyclic_mul(A[j, i].tolist(), W[k, i].tolist(), q)
                acc = [(x + y) % q for x, y in zip(acc, p)]
            assert orc.from_incomplete(out[j, k], q).tolist() == acc


# ---------------------------------------------------------------- sumcheck
def test_sum_check_invariant_small_modulus(orc):
    # fast_norm_check.rs:237-262: q = 65537, len 128, final claim = <w, tensor(cs)>^2
    rng = random.Random(7)
    q = 65537
    w = np.array(rnd(rng, 128, q), np.uint64)[None]
    cs = np.zeros((1, 7, 2), np.uint64)
    cs[0, :, 0] = rnd(rng, 7, q)
    res = orc.sumcheck(w, cs, q)
    assert res["ok"].all()
    ip = sum(int(a) * int(b) for a, b in zip(w[0], orc.tensor(cs[0, :, 0], q))) % q
    assert int(res["finals"][0, 0, 0]) == ip * ip % q
    assert int(res["evals"][0, 0, 0]) == ip


def test_sumcheck_round_polys_brute_force(orc):
    rng = random.Random(8)
    L = 32
    w = np.array(rnd(rng, L, Q), np.uint64)[None]
    cs = np.zeros((1, 5, 2), np.uint64)
    cs[0, :, 0] = rnd(rng, 5, Q)
    res = orc.sumcheck(w, cs, Q)
    assert int(res["claims"][0]) == sum(int(x) ** 2 for x in w[0]) % Q
    for t in range(5):
        f0, f1, f2 = (int(res["msgs"][0, t, 0, i, 0]) for i in range(3))
        g = pyref.round_poly_brute(w[0].tolist(), cs[0, :t, 0].tolist(), Q)
        # g(X) = f0 + 2 f1 X + f2 X^2
        assert g == [f0 % Q, (f0 + 2 * f1 + f2) % Q, (f0 + 4 * f1 + 4 * f2) % Q]


def test_sumcheck_raw_chunk_sums_congruent(orc):
    # The reference sends sum_fast_par's unreduced sums of chunk residues; they
    # are congruent to the canonical messages.
    rng = random.Random(9)
    L = 1 << 13
    w = np.array(rnd(rng, L, Q), np.uint64)
    h = L // 2
    w0, w1 = w[:h], w[h:]
    raw_f0 = orc.sum_chunked_raw(np.array([int(a) * int(a) % Q for a in w0], np.uint64), Q)
    cs = np.zeros((1, 13, 2), np.uint64)
    res = orc.sumcheck(w[None], cs, Q)
    assert raw_f0 % Q == int(res["msgs"][0, 0, 0, 0, 0])
    assert orc.sum_sq_chunked_raw(w, Q) % Q == int(res["claims"][0])
    assert raw_f0 != int(res["msgs"][0, 0, 0, 0, 0])  # raw really is unreduced


def test_sumcheck_fq2_reduces_to_fq(orc):
    rng = random.Random(10)
    L = 64
    w = np.array(rnd(rng, 2 * L, Q), np.uint64).reshape(2, L)
    cs = np.zeros((2, 6, 2), np.uint64)
    cs[:, :, 0] = np.array(rnd(rng, 12, Q), np.uint64).reshape(2, 6)
    a = orc.sumcheck(w, cs, Q, fq2=False)
    b = orc.sumcheck(w, cs, Q, fq2=True, nr=NR)
    for k in ("claims", "msgs", "finals", "evals", "ok"):
        assert np.array_equal(a[k], b[k])
    assert not b["msgs"][..., 1].any()


def test_sumcheck_fq2_invariants(orc):
    rng = random.Random(11)
    L = 16
    w = np.array(rnd(rng, L, Q), np.uint64)[None]
    cs = np.array(rnd(rng, 8, Q), np.uint64).reshape(1, 4, 2)
    res = orc.sumcheck(w, cs, Q, fq2=True, nr=NR)
    assert res["ok"].all()
    f2 = pyref.Fq2(Q, NR)
    cl = [tuple(int(v) for v in c) for c in cs[0]]
    ev = pyref.mle_eval(w[0].tolist(), cl, Q, f2)
    assert tuple(int(v) for v in res["evals"][0, 0]) == ev
    assert tuple(int(v) for v in res["finals"][0, 0]) == f2.mul(ev, ev)


def test_fq2_is_a_field_for_nr5():
    # 5 is a quadratic non-residue mod Q, so Y^2 - 5 is irreducible
    assert pow(NR, (Q - 1) // 2, Q) == Q - 1


# ---------------------------------------------------------------- norm_check_3
@pytest.mark.parametrize("d,q", [(16, Q), (128, REF_RINGS[0][1])])
def test_norm_check_3_constant_term_is_evaluation(orc, d, q):
    rng = random.Random(12)
    rows, m = 2, 8
    nc = d.bit_length() - 1
    W = np.array([rng.randrange(-16, 17) % q for _ in range(rows * m * d)],
                 np.uint64).reshape(rows, m, d)
    cs = rnd(rng, 3 + nc, q)
    rhs = orc.norm_check_3(W, cs, q)
    tens = pyref.eq_weights(cs, q)
    for j in range(rows):
        coeff = orc.from_incomplete(rhs[j], q).tolist()
        ip = sum(int(x) * t for x, t in zip(W[j].reshape(-1), tens)) % q
        assert coeff[0] == ip
        # full ring element: base * sum_i T_ring[i] W_j[i]
        tr = pyref.eq_weights(cs[:3], q)
        tc = pyref.eq_weights(cs[3:], q)
        base = orc.conjugate(tc, q).tolist()
        v = [sum(tr[i] * int(W[j, i, c]) for i in range(m)) % q for c in range(d)]
        assert coeff == pyref.negacyclic_mul(base, v, q)


def test_norm_check_3_fq2_matches_evaluation(orc):
    rng = random.Random(13)
    rows, m, d = 1, 4, 8
    W = np.array(rnd(rng, rows * m * d, Q), np.uint64).reshape(rows, m, d)
    cs2 = np.array(rnd(rng, 2 * 5, Q), np.uint64).reshape(5, 2)
    rhs = orc.norm_check_3_fq2(W, cs2, Q, NR)
    f2 = pyref.Fq2(Q, NR)
    ev = pyref.mle_eval(W[0].reshape(-1).tolist(), [tuple(int(v) for v in c) for c in cs2],
                        Q, f2)
    c0 = orc.from_incomplete(rhs[0, 0], Q).tolist()[0]
    c1 = orc.from_incomplete(rhs[0, 1], Q).tolist()[0]
    assert (c0, c1) == ev
    # F_q challenges: component 0 equals the F_q norm_check_3
    cs2[:, 1] = 0
    rhs2 = orc.norm_check_3_fq2(W, cs2, Q, NR)
    assert np.array_equal(rhs2[:, 0], orc.norm_check_3(W, cs2[:, 0], Q))
    assert not rhs2[:, 1].any()


# ---------------------------------------------------------------- decomposition
def digits_signed(x, q, radix, parts):
    v = x if 2 * x <= q else x - q
    s = -1 if v < 0 else 1
    a = abs(v)
    out = []
    for _ in range(parts):
        a, dgt = divmod(a, radix)
        out.append((s * dgt) % q)
    return out


def test_decompose_digit_brute_force(orc):
    rng = random.Random(14)
    r, m, d, parts = 2, 3, 8, 4
    W = np.array([rng.randrange(-(1 << 12), (1 << 12) + 1) % Q for _ in range(r * m * d)],
                 np.uint64).reshape(r, m, d)
    W[0, 0, 0] = Q - (1 << 12)  # the extreme -2^12 sets the bit width to 13
    limbs, radix = orc.decompose(W, parts, Q)
    assert radix == 16  # |w| <= 2^12: bit width 13, ceil(13/4) = 4
    for rho in range(r):
        for i in range(m):
            for c in range(d):
                want = digits_signed(int(W[rho, i, c]), Q, radix, parts)
                got = [int(limbs[rho * parts + t, i, c]) for t in range(parts)]
                assert got == want
    assert np.array_equal(orc.compose_rows(limbs, parts, radix, Q), W)


def test_decompose_explicit_radix_cfg2(orc):
    rng = random.Random(15)
    W = np.array([rng.randrange(-16, 17) % Q for _ in range(2 * 4 * 8)],
                 np.uint64).reshape(2, 4, 8)
    limbs, radix = orc.decompose(W, 6, Q, radix=256)
    assert radix == 256 and limbs.shape == (12, 4, 8)
    assert not limbs[1::6].any()  # |w| <= 16 < 256: only limb 0 is nonzero
    assert np.array_equal(orc.compose_rows(limbs, 6, 256, Q), W)


def test_decomp_invariant(orc):
    # decomp.rs:78-92: parts = 1, compose(decomp(W)) = W and compose(rhs_dec) = rhs
    rng = random.Random(16)
    d = 8
    A = np.array(rnd(rng, 8 * 2 * d, Q62), np.uint64).reshape(8, 2, d)
    W = np.array([rng.randrange(-2, 3) % Q62 for _ in range(8 * 2 * d)],
                 np.uint64).reshape(8, 2, d)
    limbs, radix = orc.decompose(W, 1, Q62)
    assert np.array_equal(orc.compose(limbs, 1, radix, Q62), W)
    rhs = orc.commit(A, W, Q62)
    rhs_dec = orc.commit(A, limbs, Q62)
    assert np.array_equal(orc.compose(rhs_dec, 1, radix, Q62), rhs)
    # and with 2 parts: compose over the limb commitments gives the commitment
    limbs2, radix2 = orc.decompose(W, 2, Q62)
    rhs_dec2 = orc.commit(A, limbs2, Q62)
    assert np.array_equal(orc.compose(rhs_dec2, 2, radix2, Q62), rhs)


def test_decompose_all_zero(orc):
    W = np.zeros((1, 2, 4), np.uint64)
    limbs, radix = orc.decompose(W, 3, Q)
    assert radix == 1 and not limbs.any()


# ---------------------------------------------------------------- fold
def test_fold_homomorphism(orc):
    # fold.rs:36-55: commit(fold(W, C)) = C . commit(W)
    rng = random.Random(17)
    d, kappa, r_in, r_out, m = 8, 3, 4, 2, 4
    A = np.array(rnd(rng, kappa * m * d, Q), np.uint64).reshape(kappa, m, d)
    W = np.array(rnd(rng, r_in * m * d, Q), np.uint64).reshape(r_in, m, d)
    C = np.array([rng.choice([Q - 1, 0, 1]) for _ in range(r_out * r_in * d)],
                 np.uint64).reshape(r_out, r_in, d)
    Wf = orc.fold(W, C, Q)
    lhs = orc.commit(A, orc.from_incomplete(Wf, Q), Q)
    cm = orc.commit(A, W, Q)
    for j in range(kappa):
        for jo in range(r_out):
            acc = [0] * d
            for k in range(r_in):
                p = pyref.negacyclic_mul(orc.from_incomplete(cm[j, k], Q).tolist(),
                                         C[jo, k].tolist(), Q)
                acc = [(x + y) % Q for x, y in zip(acc, p)]
            assert orc.from_incomplete(lhs[j, jo], Q).tolist() == acc


# ---------------------------------------------------------------- power series
def _const(v, d=D):
    a = np.zeros(d, np.uint64)
    a[0] = v
    return a


def test_kronecker_order_kat(orc):
    # arithmetic.rs:866-872: kronecker_product([1, 2, 3], [4, 5]) =
    # [4, 5, 8, 10, 12, 15].  With the 2-vectors [1, 2] and [2, 3] as the one
    # tensor over the base [4, 5], the layers are its first and last four entries.
    base = np.stack([_const(4), _const(5)])
    for t, want in (([1, 2], [4, 5, 8, 10]), ([2, 3], [8, 10, 12, 15])):
        lay = orc.complete_power_series(base, np.stack([[_const(t[0]), _const(t[1])]]), Q)
        assert lay[:, 0].tolist() == want and not lay[:, 1:].any()


def test_complete_power_series_brute_force(orc):
    # the index formula (the first tensor the most significant bit) with
    # random ring elements against schoolbook products; all-ones tensors (the
    # projection's layers, project.rs:177-186) repeat the base
    rng = random.Random(23)
    d, m0, n = 8, 3, 3
    base = np.array(rnd(rng, m0 * d, Q), np.uint64).reshape(m0, d)
    tensors = np.array(rnd(rng, n * 2 * d, Q), np.uint64).reshape(n, 2, d)
    lay = orc.complete_power_series(base, tensors, Q)
    for idx in range(m0 << n):
        bits, j = idx // m0, idx % m0
        acc = base[j].tolist()
        for k in range(n):
            acc = pyref.negacyclic_mul(acc, tensors[k, (bits >> (n - 1 - k)) & 1].tolist(), Q)
        assert lay[idx].tolist() == acc, idx
    ones = np.stack([np.stack([_const(1, d), _const(1, d)])] * n)
    assert np.array_equal(orc.complete_power_series(base, ones, Q), np.tile(base, (1 << n, 1)))


def test_dot_single_series_is_commit_kappa1(orc):
    # arithmetic.rs:507-536 is the commit (arithmetic.rs:459-505) with the layer
    # as the only key row; the two oracle routes share no product code
    rng = random.Random(29)
    d, r, m = 8, 3, 5
    layer = np.array(rnd(rng, m * d, Q), np.uint64).reshape(m, d)
    W = np.array(rnd(rng, r * m * d, Q), np.uint64).reshape(r, m, d)
    assert np.array_equal(orc.dot_single_series(layer, W, Q), orc.commit(layer[None], W, Q)[0])


# ---------------------------------------------------------------- projection
def _proj_inputs(seed, r, m, d):
    rng = np.random.default_rng(seed)
    W = rng.integers(0, Q, (r, m, d), dtype=np.uint64)
    z = rng.integers(0, 2, (256, r * 256), dtype=np.uint8)
    sg = rng.integers(0, 2, (256, r * 256), dtype=np.uint8)
    return W, z, sg


def test_project_signed_product(orc):
    # project.rs:52-98: chunk j of row i gives 256 elements, the signed sums
    # over the chunk with C[k][l] = 0 (is_zero), +1 or -1 (sign): a plain
    # signed matrix product in int64 (|sum| < 2^58) reduced mod q
    d, r, m = 8, 2, 1536  # cw = 512: three chunks per row, the last 512 columns dropped
    W, z, sg = _proj_inputs(3, r, m, d)
    got = orc.project(W, z, sg, Q)
    C = np.where(z == 1, 0, np.where(sg == 1, 1, -1)).astype(np.int64)
    cw, nch = r * 256, m // (r * 256)
    assert got.shape == (r * nch * 256, d)
    for i in range(r):
        for j in range(nch):
            chunk = W[i, j * cw:(j + 1) * cw].astype(np.int64)
            want = np.mod(C @ chunk, Q).astype(np.uint64)
            assert np.array_equal(got[(i * nch + j) * 256:(i * nch + j + 1) * 256], want), (i, j)


def test_project_linear(orc):
    d, r, m = 8, 1, 512
    W1, z, sg = _proj_inputs(4, r, m, d)
    W2, _, _ = _proj_inputs(5, r, m, d)
    W12 = ((W1.astype(object) + W2.astype(object)) % Q).astype(np.uint64)
    lhs = orc.project(W12, z, sg, Q)
    rhs = (orc.project(W1, z, sg, Q).astype(object) + orc.project(W2, z, sg, Q).astype(object)) % Q
    assert np.array_equal(lhs, rhs.astype(np.uint64))


def test_project_batching_identity(orc):
    # project.rs:154-190 (batch_projections): with c_j = sum_i C[i][j] c0^i,
    # <(1, c0, .., c0^255), projected chunk> = <(c_j), chunk> in R_q; the
    # reference asserts it ("Inner product mismatch") with c0 = 1 and random
    # batching challenges; here c0 is a random ring element
    d, r, m = 8, 1, 256
    W, z, sg = _proj_inputs(6, r, m, d)
    P = orc.project(W, z, sg, Q)
    rng = random.Random(31)
    c0 = rnd(rng, d, Q)
    pw = [[1] + [0] * (d - 1)]
    for _ in range(255):
        pw.append(pyref.negacyclic_mul(pw[-1], c0, Q))
    C = np.where(z == 1, 0, np.where(sg == 1, 1, -1))
    lhs = [0] * d
    for k in range(256):
        lhs = [(a + b) % Q for a, b in zip(lhs, pyref.negacyclic_mul(pw[k], P[k].tolist(), Q))]
    rhs = [0] * d
    for l in range(256):
        cj = [0] * d
        for k in range(256):
            if C[k, l]:
                cj = [(a + int(C[k, l]) * b) % Q for a, b in zip(cj, pw[k])]
        rhs = [(a + b) % Q for a, b in zip(rhs, pyref.negacyclic_mul(cj, W[0, l].tolist(), Q))]
    assert lhs == rhs


def test_consecutive_powers_monomial(orc):
    # project.rs:118-129 with c = X: X^k = (-1)^(k // d) X^(k mod d) in R_q
    d = 8
    X = np.zeros(d, np.uint64)
    X[1] = 1
    pw = orc.consecutive_powers(X, 256, Q)
    for k in range(256):
        want = np.zeros(d, np.uint64)
        want[k % d] = 1 if (k // d) % 2 == 0 else Q - 1
        assert np.array_equal(pw[k], want), k


def test_batch_projections_constant_challenge(orc):
    # batching_challenge_for_project (project.rs:104-116) draws c0 = 1: every
    # power is 1, so c_l is the constant column sum of C and both layers are
    # constants; the assertion holds
    d, r, m = 8, 2, 1024
    W, z, sg = _proj_inputs(7, r, m, d)
    P = orc.project(W, z, sg, Q)
    one = np.zeros(d, np.uint64)
    one[0] = 1
    plhs, wlhs, prhs, wrhs, ok = orc.batch_projections(P, W, z, sg, one, Q)
    assert ok
    C = np.where(z == 1, 0, np.where(sg == 1, 1, -1)).astype(np.int64)
    col = np.mod(C.sum(axis=0), Q)
    assert plhs.shape == (m, d) and wlhs.shape == (m, d)
    assert (plhs[:, 0] == 1).all() and (plhs[:, 1:] == 0).all()
    assert np.array_equal(wlhs[:, 0], np.tile(col, m // (r * 256)).astype(np.uint64))
    assert (wlhs[:, 1:] == 0).all()


def test_batch_projections_identity_and_plants(orc):
    # random c0: the assertion of project.rs:198 holds; the projection rhs is
    # the direct sum_k c0^(k mod 256) P[k]; a changed projection residue and
    # a transposed challenge (both planted here, in the oracle's inputs) fail
    d, r, m = 8, 1, 512
    W, z, sg = _proj_inputs(8, r, m, d)
    P = orc.project(W, z, sg, Q)
    rng = random.Random(41)
    c0 = np.array(rnd(rng, d, Q), np.uint64)
    plhs, wlhs, prhs, wrhs, ok = orc.batch_projections(P, W, z, sg, c0, Q)
    assert ok and wrhs.shape == (r, d)
    pw = [[1] + [0] * (d - 1)]
    for _ in range(255):
        pw.append(pyref.negacyclic_mul(pw[-1], c0.tolist(), Q))
    direct = [0] * d
    for x in range(m):
        direct = [(a + b) % Q for a, b in zip(direct, pyref.negacyclic_mul(pw[x % 256], P[x].tolist(), Q))]
    assert np.array_equal(prhs, orc.to_incomplete(np.array(direct, np.uint64), Q))
    P2 = P.copy()
    P2[3, 5] = (P2[3, 5] + 1) % Q
    assert not orc.batch_projections(P2, W, z, sg, c0, Q)[4]
    assert not orc.batch_projections(P, W, z.T.copy(), sg.T.copy(), c0, Q)[4]


def test_cross_terms_are_row0_products(orc):
    # send_cross_terms_before_join (project.rs:229-245) keeps entry 0 of each
    # inner product: rows past the first change nothing
    d, r, m = 8, 2, 512
    W, z, sg = _proj_inputs(9, r, m, d)
    P = orc.project(W, z, sg, Q)
    rng = random.Random(43)
    c0 = np.array(rnd(rng, d, Q), np.uint64)
    plhs, wlhs, _, _, _ = orc.batch_projections(P, W, z, sg, c0, Q)
    l0, l1 = orc.send_cross_terms_before_join(plhs, wlhs, W, P, Q)
    l0b, l1b = orc.send_cross_terms_before_join(plhs, wlhs, W[:1], P, Q)
    assert np.array_equal(l0, l0b) and np.array_equal(l1, l1b)
    assert np.array_equal(l0, orc.dot_single_series(plhs, W[:1], Q)[0])


def test_crs_monomial(orc):
    # compute_commitment_keys (crs.rs:43-72) of m = X: row[i] = X^(i+1) in R_q
    d = 8
    X = np.zeros((1, d), np.uint64)
    X[0, 1] = 1
    row = orc.compute_commitment_keys(X, 40, Q)[0]
    for i in range(40):
        k = i + 1
        want = np.zeros(d, np.uint64)
        want[k % d] = 1 if (k // d) % 2 == 0 else Q - 1
        assert np.array_equal(row[i], want), i


def test_crs_layers_form_a_power_series(orc):
    # crs.rs:58-69: the layers are the row's prefixes and the tensors
    # [1, m^c] rebuild each layer from the one below (complete_power_series,
    # tensors in the order the reference pushes them = most significant first)
    d = 8
    rng = random.Random(51)
    module = np.array([rnd(rng, d, Q), rnd(rng, d, Q)], np.uint64)
    rows = orc.compute_commitment_keys(module, 24, Q)
    for row in rows:
        layers, tensors = orc.crs_power_series(row)
        assert [len(x) for x in layers] == [24, 12, 6, 3] and len(tensors) == 3
        full = orc.complete_power_series(layers[-1], np.stack(tensors), Q)
        assert np.array_equal(full, row)
        # a reversed tensor order (a plausible slip) does not rebuild it
        assert not np.array_equal(orc.complete_power_series(layers[-1], np.stack(tensors[::-1]), Q), row)


# ---------------------------------------------------------------- sampling
def test_samplers_support(orc):
    u = orc.sample_uniform(0, 1, (4096,), Q)
    assert (u < Q).all() and len(set(u.tolist())) > 4000
    b = orc.sample_bounded(0, 2, (20000,), 8, Q)
    signed = np.where(b > Q // 2, b.astype(object) - Q, b.astype(object))
    assert set(signed.tolist()) == set(range(-8, 9))


M64 = (1 << 64) - 1


def _mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_prng_definition(orc):
    # The seeded stream is this library's definition (the reference rng is
    # unseeded): SplitMix64 keyed by (seed, stream), counter = element index.
    for seed, stream in ((0, 1), (7, 3), (123456789, 42)):
        key = _mix64(seed ^ ((stream * 0xD1B54A32D192ED03) & M64))
        for idx in (0, 1, 2, 1000, 1 << 40):
            want = _mix64((key + (idx + 1) * 0x9E3779B97F4A7C15) & M64)
            assert orc.prng_draw(seed, stream, idx) == want
    u = orc.sample_uniform(5, 9, (10,), Q).tolist()
    assert u == [(orc.prng_draw(5, 9, i) * Q) >> 64 for i in range(10)]


def test_full_slots_from_incomplete_halves(orc):
    """a(zeta_s) = E[j] + zeta_s O[j] with omega_j = zeta_s^2: the relation the
    fused fold uses to get full-NTT slots (the commit's domain) from the
    IncompleteNTT form, checked on random elements."""
    # roots straight from the forward transforms of X and Y
    x = np.zeros((1, 64), np.uint64)
    x[0, 1] = 1
    zeta = orc.to_full(x, Q)[0]              # zeta_s = NTT_64(X)[s]
    y = np.zeros((1, 64), np.uint64)
    y[0, 2] = 1                                # X^2 = Y in the even half
    omega = orc.to_incomplete(y, Q)[0][:32]    # omega_j = NTT_32(Y)[j]
    zj = []
    for s in range(64):
        z2 = int(zeta[s]) * int(zeta[s]) % Q
        js = [j for j in range(32) if int(omega[j]) == z2]
        assert len(js) == 1
        zj.append(js[0])
    a = orc.sample_uniform(3, 1, (20, 64), Q)
    full, inc = orc.to_full(a, Q), orc.to_incomplete(a, Q)
    for e in range(20):
        for s in range(64):
            j = zj[s]
            want = (int(inc[e, j]) + int(zeta[s]) * int(inc[e, 32 + j])) % Q
            assert int(full[e, s]) == want


# ---------------------------------------------------------------- the reference's rings
@pytest.mark.parametrize("d,q,nr", REF_RINGS)
def test_reference_moduli(d, q, nr):
    # protocol.rs:28-47: each MOD_Q is a prime below 2^50 with d | q - 1 (the
    # IncompleteNTT's two length-d/2 negacyclic transforms need a primitive
    # d-th root of unity); nr is its smallest quadratic non-residue
    assert q < 1 << 50 and pow(2, q - 1, q) == 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x, e = q - 1, 0
        while x % 2 == 0:
            x, e = x // 2, e + 1
        y = pow(a, x, q)
        if y in (1, q - 1):
            continue
        for _ in range(e - 1):
            y = y * y % q
            if y == q - 1:
                break
        assert y == q - 1, "composite witness %d" % a
    assert (q - 1) % d == 0
    assert all(pow(a, (q - 1) // 2, q) == 1 for a in range(2, nr))
    assert pow(nr, (q - 1) // 2, q) == q - 1
