"""The transcript's checker on the CPU: the oracle's round-synchronous
sumcheck (the rows sharing each round's challenge) equals the pinned
or_sumcheck when its callback hands out fixed challenges; the Fiat-Shamir
challenges depend on the messages and the seed; the byte layout parses back
and a reader that sees only the bytes accepts an honest transcript and
rejects a changed byte."""
import ctypes

import numpy as np
import pytest

from conftest import NR, Q


@pytest.fixture(scope="module")
def tr():
    from oracle import transcript
    return transcript


def _fixed(cs):
    def chal(_u, rep, t, _msgs, out):
        out[0] = int(cs[rep, t, 0])
        out[1] = int(cs[rep, t, 1])
    return chal


@pytest.mark.parametrize("rows,L,reps,fq2", [(3, 1024, 2, True), (2, 256, 1, False)])
def test_round_synchronous_sumcheck_equals_or_sumcheck(orc, tr, rows, L, reps, fq2):
    w = orc.sample_bounded(1, 0xB, (rows, L), 16, Q)
    n = L.bit_length() - 1
    cs = np.stack([orc.sample_uniform(2, 0x5C000 + r, (2 * n,), Q).reshape(n, 2)
                   for r in range(reps)])
    if not fq2:
        cs[:, :, 1] = 0
    want = orc.sumcheck(w, cs, Q, fq2=fq2, nr=NR, threads=1)
    o = dict(claims=np.zeros(rows, np.uint64), msgs=np.zeros((reps, n, rows, 3, 2), np.uint64),
             cs=np.zeros((reps, n, 2), np.uint64), finals=np.zeros((reps, rows, 2), np.uint64),
             evals=np.zeros((reps, rows, 2), np.uint64), ok=np.zeros((reps, rows), np.int32))
    cb = tr._CHAL(_fixed(cs))
    p = orc._p
    orc.lib.or_sumcheck_cb(p(w), rows, L, reps, int(fq2), NR, Q, p(o["claims"]), p(o["msgs"]),
                           p(o["cs"]), p(o["finals"]), p(o["evals"]),
                           o["ok"].ctypes.data_as(ctypes.POINTER(ctypes.c_int)), cb, None)
    for k in ("claims", "msgs", "finals", "evals", "ok"):
        assert np.array_equal(o[k], want[k]), k
    assert np.array_equal(o["cs"], cs)


def test_fiat_shamir_challenges_follow_the_messages(orc, tr):
    rows, L = 2, 512
    w = orc.sample_bounded(3, 0xB, (rows, L), 16, Q)
    a = tr.sumcheck_fs(w, 2, Q, seed=bytes(32))
    assert a["ok"].all()
    # the same challenges, handed in, reproduce the messages
    b = orc.sumcheck(w, a["cs"], Q, fq2=True, nr=NR, threads=1)
    for k in ("claims", "msgs", "finals", "evals"):
        assert np.array_equal(a[k], b[k]), k
    # repetitions draw different chains; another seed changes every challenge
    assert not np.array_equal(a["cs"][0], a["cs"][1])
    c = tr.sumcheck_fs(w, 2, Q, seed=bytes([1]) + bytes(31))
    assert not np.any(c["cs"] == a["cs"])
    # the same claims with other messages (w1 negated keeps sum w^2) change
    # the challenges from round 0 on
    w2 = w.copy()
    w2[0, L // 2:] = np.where(w[0, L // 2:] == 0, 0, np.uint64(Q) - w[0, L // 2:])
    d = tr.sumcheck_fs(w2, 2, Q)
    assert np.array_equal(d["claims"], a["claims"])
    assert not np.any(d["cs"][:, 0] == a["cs"][:, 0])


def test_byte_layout_and_reader(orc, tr):
    rows, L, reps = 3, 1024, 2
    w = orc.sample_bounded(4, 0xB, (rows, L), 16, Q)
    seed = bytes(range(32))
    o = tr.sumcheck_fs(w, reps, Q, seed=seed)
    rhs = np.stack([orc.norm_check_3_fq2(w.reshape(rows, L // 64, 64), o["cs"][r], Q, NR)
                    for r in range(reps)])
    b = tr.norm_check_bytes(o, Q, L, True, NR, True, seed, rhs)
    h = tr.parse_header(b)
    assert (h["kind"], h["flags"], h["q"], h["nr"], h["L"], h["rows"], h["reps"],
            h["nrounds"], h["d"]) == (1, 7, Q, NR, L, rows, reps, 10, 64)
    assert h["seed"] == seed
    n = 10
    assert len(b) == 96 + 8 * (rows + reps * n * rows * 6 + reps * n * 2 + 4 * reps * rows
                               + reps * rows * 128)
    assert tr.verify_norm_check_bytes(b)
    # a message byte, a slot of an rhs that coefficient 0 depends on (the even
    # half: verify_norm_check_3 reads ct only), the seed
    rhs_start = len(b) - 8 * reps * rows * 128
    for pos in (96 + 8 * rows + 5, rhs_start + 8 * 129 + 2, 100):
        bad = bytearray(b)
        bad[pos] ^= 1
        assert not tr.verify_norm_check_bytes(bytes(bad)), pos
