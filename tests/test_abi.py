"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, and its legacy hexl_wrapper entry points (host code) agree with the
oracle and the reference's known answers.  No device call is made here."""
import ctypes
import random

import numpy as np
import pytest

from conftest import Q

Q62 = 4546383823830515713


@pytest.fixture(scope="module")
def sa():
    import salsaa_amd
    return salsaa_amd


def ptr(a):
    return a.ctypes.data


def test_library_exports_every_header_symbol(sa):
    names = sa.header_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(sa.lib, n)]
    assert not missing, missing
    # the 21 legacy hexl_wrapper symbols bound by src/hexl/bindings.rs:3-73
    legacy = ["multiply_mod", "power_mod", "add_mod", "sub_mod", "eltwise_mult_mod",
              "get_roots", "inv_mod", "get_inv_roots", "sum_sq", "sum_sq_fast", "sum",
              "sum_fast", "eltwise_fma_mod", "eltwise_add_mod", "eltwise_sub_mod",
              "multiply_poly", "eltwise_reduce_mod", "eltwise_reduce_mod_naive",
              "polynomial_multiply_cyclotomic_mod", "ntt_forward_in_place",
              "ntt_inverse_in_place"]
    assert len(legacy) == 21 and set(legacy) <= set(names)


def test_legacy_scalar_kats(sa):
    assert sa.lib.multiply_mod(12345, 67890, 1000000007) == 838102050
    assert sa.lib.add_mod(10, 9, 17) == 2 and sa.lib.sub_mod(3, 5, 17) == 15
    assert sa.lib.power_mod(3, Q - 1, Q) == 1
    assert sa.lib.multiply_mod(sa.lib.inv_mod(123456789, Q), 123456789, Q) == 1


def test_legacy_eltwise_kats(sa):
    a = np.array([1, 2, 3, 4, 5], np.uint64)
    b = np.array([6, 7, 8, 9, 10], np.uint64)
    r = np.zeros(5, np.uint64)
    sa.lib.eltwise_mult_mod(ptr(r), ptr(a), ptr(b), 5, 100)
    assert r.tolist() == [6, 14, 24, 36, 50]
    x = np.arange(1, 9, dtype=np.uint64)
    assert sa.lib.sum_sq(ptr(x), 8, 71) == 62
    assert sa.lib.sum_sq_fast(ptr(x), 8, 71) == 62
    assert sa.lib.sum(ptr(x), 8, 7) == 36 % 7 == sa.lib.sum_fast(ptr(x), 8, 7)
    sa.lib.eltwise_fma_mod(ptr(r), ptr(a), 3, ptr(b), 5, 17)
    assert r.tolist() == [(3 * i + j) % 17 for i, j in zip(a.tolist(), b.tolist())]
    big = np.array([2 ** 64 - 1, Q, 5], np.uint64)
    out = np.zeros(3, np.uint64)
    sa.lib.eltwise_reduce_mod(ptr(out), ptr(big), 3, Q)
    assert out.tolist() == [(2 ** 64 - 1) % Q, 0, 5]


@pytest.mark.parametrize("n,q", [(4, 17), (8, 65537), (32, Q), (64, Q), (16, Q62)])
def test_legacy_ntt_matches_oracle(sa, orc, n, q):
    rng = random.Random(n)
    x = np.array([rng.randrange(q) for _ in range(n)], np.uint64)
    y = x.copy()
    sa.lib.ntt_forward_in_place(ptr(y), n, q)
    assert y.tolist() == orc.ntt_forward(x, q).tolist()
    sa.lib.ntt_inverse_in_place(ptr(y), n, q)
    assert y.tolist() == x.tolist()
    roots = np.ctypeslib.as_array(sa.lib.get_roots(n, q), shape=(n,)).copy()
    assert roots.tolist() == orc.get_roots(n, q).tolist()
    inv = np.ctypeslib.as_array(sa.lib.get_inv_roots(n, q), shape=(n,)).copy()
    assert inv.tolist() == orc.get_inv_roots(n, q).tolist()


def test_legacy_multiply_poly_kat(sa):
    # hexl/bindings.rs:119-148
    a = np.array([1, 2, 3, 1, 0, 0, 0, 0], np.uint64)
    b = np.array([8, 7, 6, 1, 0, 0, 0, 0], np.uint64)
    r = np.zeros(8, np.uint64)
    sa.lib.multiply_poly(ptr(r), ptr(a), ptr(b), 8, 65537)
    want = [0] * 8
    for i in range(4):
        for j in range(4):
            want[i + j] += int(a[i]) * int(b[j])
    assert r.tolist() == want
    r2 = np.zeros(8, np.uint64)
    sa.lib.polynomial_multiply_cyclotomic_mod(ptr(r2), ptr(a), ptr(b), 8, 65537)
    assert r2.tolist() == want


def test_legacy_negacyclic_product(sa, orc):
    rng = random.Random(3)
    a = np.array([rng.randrange(Q) for _ in range(64)], np.uint64)
    b = np.array([rng.randrange(Q) for _ in range(64)], np.uint64)
    r = np.zeros(64, np.uint64)
    sa.lib.multiply_poly(ptr(r), ptr(a), ptr(b), 64, Q)
    assert r.tolist() == orc.ring_mul_naive(a, b, Q).tolist()


def test_legacy_threads_share_tables(sa, orc):
    # the table cache is built under a lock; concurrent first use is safe
    import threading
    q = 998244353 * 0 + 7340033  # NTT-friendly prime, q = 1 mod 2^20
    errs = []

    def work():
        x = np.arange(1, 1025, dtype=np.uint64) % q
        y = x.copy()
        sa.lib.ntt_forward_in_place(ptr(y), 1024, q)
        sa.lib.ntt_inverse_in_place(ptr(y), 1024, q)
        if y.tolist() != x.tolist():
            errs.append(1)

    ts = [threading.Thread(target=work) for _ in range(8)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs
