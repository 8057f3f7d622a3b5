"""CPU replay of the SALSAA protocol (protocol.rs:125-483) -- TEST
INFRASTRUCTURE ONLY.

This is the checker for the device's salsaa_protocol: it runs the same
rounds with the same seeded challenges and returns the final witness, the
verifier's rhs and the status code of the first verifier relation that
failed.  Only tests/ load it; the product library never does.

It follows the reference's own route, not the device's shortcuts: every
statement row's layer of the needed length is expanded (the CRS rows'
powers, complete_power_series for norm_check_3's series, prefixes of the
repeated batching layers) and committed against the witness with the
oracle's per-product commit (arithmetic.rs:459-505), and every verifier
relation is checked as the reference asserts it.

Challenge sources (the reference's are unseeded; the device draws these):
round i's seed is the key of stream 0x7000 + i under `seed`; its norm-check
challenges are sample_uniform(round seed, 0x5C000 + rep) (coordinate 0,
F_q), the projection's is_zero / sign bits the low bits of streams 1 / 2,
and the fold challenge C [1][r][d] sample_bounded(round seed, 3, bound 1).
The batching challenge c0 is the constant 1 (project.rs:100-111).

Fault injection (salsaa_protocol_opts): check 1..6 in round fault_round adds
1 mod q to the prover value that check reads, as the device does.
"""
import numpy as np

from . import oracle as orc

HEIGHT = 256  # project.rs:17


def _one(d):
    o = np.zeros(d, np.uint64)
    o[0] = 1
    return o


def _add1(x, q):
    x = x.copy()
    flat = x.reshape(-1)
    flat[0] = (int(flat[0]) + 1) % q
    return x


def _lg2(x):
    return int(x).bit_length() - 1


class _CrsRow:
    """A CRS row (crs.rs:43-72): layer l = the first l powers of m_j; the
    split multiplier at length l is [1, m_j^(l/2)]."""

    def __init__(self, powers):
        self.powers = powers

    def layer(self, l, q):
        return self.powers[:l]

    def mult(self, l, q):
        d = self.powers.shape[1]
        return np.stack([_one(d), self.powers[l // 2 - 1]])


class _SeriesRow:
    """norm_check_3's series (fast_norm_check.rs:401-441): base [1][d],
    constant tensors [n][2][d]; the layer of length 2^k is the base times
    the last k tensors (complete_power_series, arithmetic.rs:158-176), and its
    split multiplier is tensor n - k (split.rs:187-211)."""

    def __init__(self, base, tens):
        self.base, self.tens = base, tens

    def layer(self, l, q):
        k = _lg2(l)
        return orc.complete_power_series(self.base, self.tens[len(self.tens) - k:], q)

    def mult(self, l, q):
        return self.tens[len(self.tens) - _lg2(l)]


class _RepeatRow:
    """witness_lhs / projected_lhs (project.rs:168-181): all-ones tensors, so
    each layer is a prefix of the longest and every multiplier is [1, 1]."""

    def __init__(self, longest):
        self.longest = longest

    def layer(self, l, q):
        return self.longest[:l]

    def mult(self, l, q):
        d = self.longest.shape[1]
        return np.stack([_one(d), _one(d)])


def _commit(stmt, W, q, threads):
    """parallel_dot_series_matrix(statement, W) (arithmetic.rs:459-505)"""
    l = W.shape[1]
    A = np.stack([row.layer(l, q) for row in stmt])
    return orc.commit(A, W, q, threads=threads)


def _nc_series(cs, q, d):
    """norm_check_3's base conj(tensor(last log2 d challenges)) and tensors
    [1 - c, c] of the others, Coefficient (fast_norm_check.rs:406-430)"""
    split = _lg2(d)
    ring, coef = cs[:-split], cs[-split:]
    base = orc.conjugate(orc.tensor(coef, q), q).reshape(1, d)
    tens = np.zeros((len(ring), 2, d), np.uint64)
    for t, c in enumerate(ring):
        c = int(c) % q
        tens[t, 0, 0] = (1 - c) % q
        tens[t, 1, 0] = c
    return base, tens


def _norm_check_ok(msgs, claims, finals_sq, rep, q, cs, fault_claim):
    """verify_norm_check_2 round by round (fast_norm_check.rs:352-399), then
    verify_norm_check_3 (:443-462): ct(rhs)^2 equals the last claim."""
    rows = claims.shape[0]
    for j in range(rows):
        claim = (int(claims[j]) + (1 if fault_claim and j == 0 else 0)) % q
        for t in range(msgs.shape[1]):
            f0, f1, f2 = (int(msgs[rep, t, j, k, 0]) for k in range(3))
            if (2 * f0 + 2 * f1 + f2) % q != claim:
                return False
            c = int(cs[t])
            claim = (f0 + 2 * f1 * c + f2 * c * c) % q
        if finals_sq[j] != claim:
            return False
    return True


def protocol(module, W, nrounds, parts, folding, seed, q, fault=None, threads=8):
    """Returns dict(status, W, rhs): the final witness [r][m][d]
    (Coefficient), the verifier rhs [R][r][d] (IncompleteNTT) and 0 or
    100 * round + check of the first failed relation."""
    module = np.ascontiguousarray(module, np.uint64)
    W = np.ascontiguousarray(W, np.uint64)
    kappa, d = module.shape
    r, m, _ = W.shape
    parts = list(parts) if parts is not None else [1] * nrounds
    status = [0]

    def fail(code):
        if status[0] == 0:
            status[0] = code

    def fault_at(rnd, check):
        return fault is not None and fault[1] == check and (check == 6 or fault[0] == rnd)

    # the CRS and the commitment (protocol.rs:132-145)
    crs = orc.compute_commitment_keys(module, m, q)
    stmt = [_CrsRow(crs[j]) for j in range(kappa)]
    V = _commit(stmt, W, q, threads)

    def decomp(W, V, p, code, corrupt):
        # decomp_ell + verify_decomp (decomp.rs:32-112)
        W2, radix = orc.decompose(W, p, q)
        Vd = _commit(stmt, W2, q, threads)
        if corrupt:
            Vd = _add1(Vd, q)
        if not np.array_equal(orc.compose(Vd, p, radix, q), V):
            fail(code)
        return W2, Vd

    rounds_done = 0
    for i in range(nrounds):
        rseed = orc.stream_key(seed, 0x7000 + i)
        p = parts[i] if i < len(parts) else 1
        if not folding and p > 1:
            W, V = decomp(W, V, p, 100 * i + 1, fault_at(i, 1))
            r = W.shape[0]
        # norm_check_1/2/3 with their verifiers, 2 repetitions (protocol.rs:198-259)
        L = m * d
        nr = _lg2(L)
        cs = np.stack([orc.sample_uniform(rseed, 0x5C000 + rep, (2 * nr,), q).reshape(nr, 2)
                       for rep in range(2)])
        cs[:, :, 1] = 0
        nc = orc.sumcheck(W.reshape(r, L), cs, q, fq2=False, threads=threads)
        new_rows = []
        for rep in range(2):
            rhs = orc.norm_check_3(W, cs[rep, :, 0], q)
            ct = orc.from_incomplete(rhs, q)[:, 0]
            sq = [int(x) * int(x) % q for x in ct]
            if not _norm_check_ok(nc["msgs"], nc["claims"], sq, rep, q, cs[rep, :, 0],
                                  fault_at(i, 2)):
                fail(100 * i + 2)
            base, tens = _nc_series(cs[rep, :, 0], q, d)
            stmt.append(_SeriesRow(base, tens))
            new_rows.append(rhs)
        V = np.concatenate([V, np.stack(new_rows)])
        # split + verifier_split (protocol.rs:261-291, split.rs:45-76, 171-221)
        if not folding:
            h = m // 2
            W3 = np.concatenate([W[:, :h], W[:, h:]])
            S = _commit(stmt, W3, q, threads)
            if fault_at(i, 3):
                S = _add1(S, q)
            Sc = orc.from_incomplete(S, q)
            for e, row in enumerate(stmt):
                t0, t1 = row.mult(m, q)
                for k in range(r):
                    a = orc.ring_mul_naive(Sc[e, k], t0, q)
                    b = orc.ring_mul_naive(Sc[e, r + k], t1, q)
                    if not np.array_equal(orc.to_incomplete((a + b) % np.uint64(q), q), V[e, k]):
                        fail(100 * i + 3)
            W, V, r, m = W3, S, 2 * r, h
        # challenge_for_project, project, batch_projections, verify_batching (:293-348)
        width = HEIGHT * r
        assert width <= m and m % width == 0
        z = orc.sample_bits(rseed, 1, (HEIGHT, width))
        sg = orc.sample_bits(rseed, 2, (HEIGHT, width))
        P = orc.project(W, z, sg, q)
        pcom = _commit(stmt, P[None], q, threads)
        if fault_at(i, 4):
            P = _add1(P, q)
        plhs, wlhs, prhs, wrhs, ok = orc.batch_projections(P, W, z, sg, _one(d), q)
        if not ok:
            fail(100 * i + 4)
        stmt.append(_RepeatRow(wlhs))
        V = np.concatenate([V, wrhs[None]])
        # challenge_for_fold, fold, verifier_fold with WIT_REP = 1 (:349-386)
        C = orc.sample_bounded(rseed, 3, (1, r, d), 1, q)
        W4 = orc.from_incomplete(orc.fold(W, C, q, threads=threads), q)
        Vc = orc.from_incomplete(V, q)
        Vf = orc.fold(np.ascontiguousarray(Vc.transpose(1, 0, 2)), C, q, threads=threads)[0]
        # send_cross_terms_before_join, verifier_join, witness.append (:387-410)
        l0, l1 = orc.send_cross_terms_before_join(plhs, wlhs, W4, P, q)
        R = len(stmt)
        Vj = np.zeros((R + 1, 2, d), np.uint64)
        Vj[:R, 0] = Vf
        Vj[:R - 1, 1] = pcom[:, 0]
        Vj[R - 1, 1] = l1
        Vj[R, 0], Vj[R, 1] = l0, prhs
        stmt.append(_RepeatRow(plhs))
        V = Vj
        W = np.concatenate([W4, P[None]])
        r = 2
        rounds_done = i + 1
        if folding:
            # the folding scheme closes with decomp_ell(.., 2) + verify_decomp
            W, V = decomp(W, V, 2, 100 * i + 5, fault_at(i, 5))
            r = W.shape[0]
            break
    if not folding:
        # the final check (protocol.rs:464-472)
        Wf = _add1(W, q) if fault_at(rounds_done, 6) else W
        if not np.array_equal(_commit(stmt, Wf, q, threads), V):
            fail(100 * rounds_done + 6)
    return dict(status=status[0], W=W, rhs=V)
