"""The proof transcript's byte format and the Fiat-Shamir norm check, on the
CPU -- TEST INFRASTRUCTURE ONLY (the checker of salsaa_norm_check_fs and the
salsaa_transcript_* entry points; the product library never loads it).

The reference has no transcript: it logs values (protocol.rs:474-482) and
its verifier draws each norm-check challenge at random, independent of the
round's message (fast_norm_check.rs:375-385).  The Fiat-Shamir variant takes
round t's challenge, shared by every row as the reference's is
(fast_norm_check.rs:352-399), from a SHA-256 chain (hashlib, FIPS 180-4):
    h_0     = SHA-256(b"SALSAA-NC-FS-v1" || seed || q || L || rows || rep || claims)
    h_{t+1} = SHA-256(h_t || round t's messages of every row)
    c_t     = (mulhi(u64(h_{t+1}[0:8]), q), mulhi(u64(h_{t+1}[8:16]), q) or 0 over F_q)
integers little-endian (q, L, claims, messages u64; rows, rep u32).

Byte layout (a 96-byte header, then canonical u64 sections):
    0 b"SLSATRN1"  8 u32 kind  12 u32 flags  16 u64 q  24 u64 nr  32 u64 L
   40 u32 rows  44 u32 reps  48 u32 nrounds  52 u32 d  56 u32 kappa  60 u32 0
   64 32-byte seed
  kind 1 (norm check): claims, msgs [reps][nrounds][rows][3][2], challenges
    [reps][nrounds][2], finals [reps][rows][2], evals [reps][rows][2], rhs
    [reps][rows][2][d] (flag bit 2)
  kind 2 (prover step): commitment [kappa][rows][d], then kind 1's sections
  kind 3 (protocol; rows = final witness rows, reps = rounds, nrounds =
    statement rows R, L = final witness length): the verifier rhs [R][rows][d],
    the final witness [rows][L][d]; flags bit 3 folding, status in bits 16..31
flags: bit 0 Fiat-Shamir, bit 1 F_{q^2}, bit 2 rhs present.
"""
import ctypes
import hashlib
import struct

import numpy as np

from . import oracle as orc

D = 64
DOMAIN = b"SALSAA-NC-FS-v1"

_CHAL = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                         ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64))
_P = ctypes.POINTER(ctypes.c_uint64)
orc.lib.or_sumcheck_cb.restype = None
orc.lib.or_sumcheck_cb.argtypes = [_P, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                   ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, _P, _P, _P, _P,
                                   _P, ctypes.POINTER(ctypes.c_int), _CHAL, ctypes.c_void_p]


def _mulhi(x, q):
    return (x * q) >> 64


def challenge_chain_start(seed, q, L, rows, rep, claims):
    return hashlib.sha256(DOMAIN + bytes(seed) + struct.pack("<QQII", q, L, rows, rep)
                          + np.ascontiguousarray(claims, "<u8").tobytes()).digest()


def sumcheck_fs(w, reps, q, fq2=True, nr=5, seed=bytes(32)):
    """The Fiat-Shamir norm check of the rows w [rows][L]: dict of claims,
    msgs, cs (the derived challenges), finals, evals, ok."""
    w = np.ascontiguousarray(w, np.uint64)
    rows, L = w.shape
    n = L.bit_length() - 1
    seed = bytes(seed)
    assert len(seed) == 32
    o = dict(claims=np.zeros(rows, np.uint64), msgs=np.zeros((reps, n, rows, 3, 2), np.uint64),
             cs=np.zeros((reps, n, 2), np.uint64), finals=np.zeros((reps, rows, 2), np.uint64),
             evals=np.zeros((reps, rows, 2), np.uint64), ok=np.zeros((reps, rows), np.int32))
    state = {}

    def chal(_user, rep, t, msgs, c_out):
        if t == 0:
            state[rep] = challenge_chain_start(seed, q, L, rows, rep, o["claims"])
        raw = np.ctypeslib.as_array(msgs, shape=(rows * 6,)).astype("<u8").tobytes()
        h = hashlib.sha256(state[rep] + raw).digest()
        state[rep] = h
        c_out[0] = _mulhi(int.from_bytes(h[0:8], "little"), q)
        c_out[1] = _mulhi(int.from_bytes(h[8:16], "little"), q) if fq2 else 0

    cb = _CHAL(chal)
    p = orc._p
    orc.lib.or_sumcheck_cb(p(w), rows, L, reps, int(fq2), nr, q, p(o["claims"]), p(o["msgs"]),
                           p(o["cs"]), p(o["finals"]), p(o["evals"]),
                           o["ok"].ctypes.data_as(ctypes.POINTER(ctypes.c_int)), cb, None)
    return o


def _header(kind, flags, q, nr, L, rows, reps, nrounds, kappa, seed):
    seed = bytes(seed) if seed is not None else bytes(32)
    return (b"SLSATRN1" + struct.pack("<IIQQQIIIIII", kind, flags, q, nr, L, rows, reps, nrounds,
                                      D, kappa, 0) + seed)


def _u64(a):
    return np.ascontiguousarray(a, np.uint64).astype("<u8").tobytes()


def _nc_body(o, rhs):
    body = _u64(o["claims"]) + _u64(o["msgs"]) + _u64(o["cs"]) + _u64(o["finals"]) + \
        _u64(o["evals"])
    return body + (_u64(rhs) if rhs is not None else b"")


def norm_check_bytes(o, q, L, fq2, nr, fs, seed=None, rhs=None):
    """kind 1 from a norm check's outputs (o: claims, msgs, cs, finals, evals)"""
    reps, n, rows = o["msgs"].shape[:3]
    flags = (1 if fs else 0) | (2 if fq2 else 0) | (4 if rhs is not None else 0)
    return _header(1, flags, q, nr if fq2 else 0, L, rows, reps, n, 0,
                   seed if fs else None) + _nc_body(o, rhs)


def prover_step_bytes(commitment, o, q, L, fq2, nr, fs, seed=None, rhs=None):
    """kind 2: the commitment [kappa][rows][d], then the norm check"""
    reps, n, rows = o["msgs"].shape[:3]
    kappa = commitment.shape[0]
    flags = (1 if fs else 0) | (2 if fq2 else 0) | (4 if rhs is not None else 0)
    return (_header(2, flags, q, nr if fq2 else 0, L, rows, reps, n, kappa, seed if fs else None)
            + _u64(commitment) + _nc_body(o, rhs))


def protocol_bytes(q, kappa, nrounds, folding, seed, status, rhs, W):
    """kind 3: the verifier rhs [R][r][d] and the final witness [r][m][d]"""
    R, r = rhs.shape[:2]
    m = W.shape[1]
    flags = (8 if folding else 0) | ((status & 0xFFFF) << 16)
    sd = struct.pack("<Q", seed) + bytes(24)
    return _header(3, flags, q, 0, m, r, nrounds, R, kappa, sd) + _u64(rhs) + _u64(W)


def parse_header(b):
    """The header's fields as a dict (for tests and readers)."""
    assert b[:8] == b"SLSATRN1"
    kind, flags, q, nr, L, rows, reps, nrounds, d, kappa, _ = struct.unpack("<IIQQQIIIIII",
                                                                            b[8:64])
    return dict(kind=kind, flags=flags, q=q, nr=nr, L=L, rows=rows, reps=reps, nrounds=nrounds,
                d=d, kappa=kappa, seed=b[64:96])


def verify_norm_check_bytes(b):
    """A verifier reading only a kind-1 transcript: with Fiat-Shamir, every
    challenge must be the hash chain's; every round identity 2 f0 + 2 f1 + f2
    = claim must hold with the claim update (fast_norm_check.rs:352-399); the
    last claim must be eval^2 and, with the rhs, ct(rhs)^2 (:443-462)."""
    h = parse_header(b)
    assert h["kind"] == 1
    q, rows, reps, n, nr = h["q"], h["rows"], h["reps"], h["nrounds"], h["nr"]
    fs, fq2, has_rhs = h["flags"] & 1, h["flags"] & 2, h["flags"] & 4
    a = np.frombuffer(b[96:], "<u8").astype(np.uint64)
    off = 0

    def take(*shape):
        nonlocal off
        k = int(np.prod(shape))
        v = a[off:off + k].reshape(shape)
        off += k
        return v

    claims, msgs, cs = take(rows), take(reps, n, rows, 3, 2), take(reps, n, 2)
    finals, evals = take(reps, rows, 2), take(reps, rows, 2)
    rhs = take(reps, rows, 2, D) if has_rhs else None
    if off != a.size:
        return False

    def mul(x, y):
        return ((x[0] * y[0] + nr * x[1] * y[1]) % q, (x[0] * y[1] + x[1] * y[0]) % q)

    for rep in range(reps):
        if fs:
            st = challenge_chain_start(h["seed"], q, h["L"], rows, rep, claims)
            for t in range(n):
                st = hashlib.sha256(st + _u64(msgs[rep, t])).digest()
                want = (_mulhi(int.from_bytes(st[0:8], "little"), q),
                        _mulhi(int.from_bytes(st[8:16], "little"), q) if fq2 else 0)
                if (int(cs[rep, t, 0]), int(cs[rep, t, 1])) != want:
                    return False
        for row in range(rows):
            cl = (int(claims[row]), 0)
            for t in range(n):
                f = [(int(msgs[rep, t, row, k, 0]), int(msgs[rep, t, row, k, 1])) for k in range(3)]
                lhs = tuple((2 * f[0][i] + 2 * f[1][i] + f[2][i]) % q for i in range(2))
                if lhs != cl:
                    return False
                c = (int(cs[rep, t, 0]), int(cs[rep, t, 1]))
                f1c = mul((2 * f[1][0], 2 * f[1][1]), c)
                f2cc = mul(f[2], mul(c, c))
                cl = tuple((f[0][i] + f1c[i] + f2cc[i]) % q for i in range(2))
            if (int(finals[rep, row, 0]), int(finals[rep, row, 1])) != cl:
                return False
            ev = (int(evals[rep, row, 0]), int(evals[rep, row, 1]))
            if mul(ev, ev) != cl:
                return False
            if rhs is not None:
                ct = orc.from_incomplete(rhs[rep, row], q)[:, 0]
                ctv = (int(ct[0]), int(ct[1]))
                if mul(ctv, ctv) != cl:
                    return False
    return True
