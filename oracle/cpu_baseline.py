"""ctypes binding of libcpu_baseline.so, the benchmark's CPU baseline (the
reference's CPU prover path in C++ on all host cores; see cpu_baseline.cpp).
BENCHMARK AND TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg and tests/
load it; the product library never does."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcpu_baseline.so")


def _load():
    if not os.path.exists(_LIB_PATH):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", _HERE, "libcpu_baseline.so"])
    lib = ctypes.CDLL(_LIB_PATH)
    P, sz, u64, vp, i = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_void_p,
                         ctypes.c_int)
    sig = {
        "cb_variant": (i, []),
        "cb_init": (i, [i, i]),
        "cb_to_incomplete": (None, [P, sz, u64]),
        "cb_ntt_roundtrip": (None, [P, sz, u64]),
        "cb_key_create": (vp, [P, sz, sz, u64]),
        "cb_key_destroy": (None, [vp]),
        "cb_commit": (i, [vp, sz, P, sz, sz, P]),
        "cb_norm_check": (None, [P, sz, sz, P, sz, u64, u64, P, P, P, P, P, P]),
        "cb_decompose": (u64, [P, sz, sz, sz, u64, u64, P]),
        "cb_verify_decomp": (sz, [P, P, sz, sz, sz, u64, u64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def init(threads=0, ifma=True):
    """Threads of the pool (0: all hardware threads); returns the count."""
    return lib.cb_init(threads, int(ifma))


def variant():
    return "avx512-ifma" if lib.cb_variant() else "scalar"


def to_incomplete(a, q):
    a = np.ascontiguousarray(a, np.uint64).copy()
    lib.cb_to_incomplete(_p(a), a.size // 64, q)
    return a


def ntt_roundtrip(W, q):
    """in place"""
    lib.cb_ntt_roundtrip(_p(W), W.size // 64, q)


class Key:
    def __init__(self, A, q):
        A = np.ascontiguousarray(A, np.uint64)
        self.kappa, self.m = A.shape[0], A.shape[1]
        self.h = lib.cb_key_create(_p(A), self.kappa, self.m, q)

    def close(self):
        if self.h:
            lib.cb_key_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def commit(key, W, kappa=None):
    W = np.ascontiguousarray(W, np.uint64)
    r, m, _ = W.shape
    kappa = key.kappa if kappa is None else kappa
    out = np.zeros((kappa, r, 64), np.uint64)
    assert lib.cb_commit(key.h, kappa, _p(W), r, m, _p(out)) == 0
    return out


def norm_check(w, cs, q, nr=5, want_rhs=True):
    """w [rows][L]; cs [reps][log2 L][2] (F_{q^2}); outputs as the oracle's
    sumcheck, plus rhs [reps][rows][2][64]."""
    w = np.ascontiguousarray(w, np.uint64)
    cs = np.ascontiguousarray(cs, np.uint64)
    rows, L = w.shape
    reps, n, _ = cs.shape
    o = dict(claims=np.zeros(rows, np.uint64), msgs=np.zeros((reps, n, rows, 3, 2), np.uint64),
             finals=np.zeros((reps, rows, 2), np.uint64), evals=np.zeros((reps, rows, 2), np.uint64),
             ok=np.zeros((reps, rows), np.int32),
             rhs=np.zeros((reps, rows, 2, 64), np.uint64) if want_rhs else None)
    lib.cb_norm_check(_p(w), rows, L, _p(cs), reps, nr, q, _p(o["claims"]), _p(o["msgs"]),
                      _p(o["finals"]), _p(o["evals"]), _p(o["ok"]),
                      _p(o["rhs"]) if want_rhs else None)
    return o


def decompose(W, parts, q, radix=0):
    W = np.ascontiguousarray(W, np.uint64)
    r, m, _ = W.shape
    out = np.zeros((r * parts, m, 64), np.uint64)
    radix = lib.cb_decompose(_p(W), r, m, parts, radix, q, _p(out))
    return out, radix


def verify_decomp(rhs_dec, rhs, parts, radix, q):
    kappa, r, _ = rhs.shape
    return lib.cb_verify_decomp(_p(np.ascontiguousarray(rhs_dec, np.uint64)),
                                _p(np.ascontiguousarray(rhs, np.uint64)), kappa, r, parts,
                                radix, q)
