"""ctypes view of the CPU oracle (test infrastructure only).

Only the tests, the smoke check and the benchmark's CPU-baseline leg import
this module: it is the checker, never the product.  All arrays are numpy
uint64 in the layouts documented in salsaa_oracle.h.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsalsaa_oracle.so")

u64 = ctypes.c_uint64
sz = ctypes.c_size_t
P = ctypes.POINTER(ctypes.c_uint64)
IP = ctypes.POINTER(ctypes.c_int)


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    sig = {
        "or_mulmod": (u64, [u64, u64, u64]),
        "or_mulmod_slow": (u64, [u64, u64, u64]),
        "or_addmod": (u64, [u64, u64, u64]),
        "or_submod": (u64, [u64, u64, u64]),
        "or_powmod": (u64, [u64, u64, u64]),
        "or_invmod": (u64, [u64, u64]),
        "or_min_primitive_root": (u64, [u64, u64]),
        "or_get_roots": (None, [sz, u64, P]),
        "or_get_inv_roots": (None, [sz, u64, P]),
        "or_complete_power_series": (None, [P, sz, P, sz, sz, u64, P]),
        "or_dot_single_series": (None, [P, P, P, sz, sz, sz, u64]),
        "or_project": (None, [P, sz, sz, sz, ctypes.c_void_p, ctypes.c_void_p, u64, P]),
        "or_ntt_forward": (None, [P, sz, u64]),
        "or_ntt_inverse": (None, [P, sz, u64]),
        "or_ntt_forward_direct": (None, [P, P, sz, u64]),
        "or_to_incomplete": (None, [P, sz, u64]),
        "or_from_incomplete": (None, [P, sz, u64]),
        "or_to_full": (None, [P, sz, u64]),
        "or_from_full": (None, [P, sz, u64]),
        "or_shift_factors": (None, [sz, u64, P]),
        "or_ring_mul_incomplete": (None, [P, P, P, sz, u64]),
        "or_ring_mul_naive": (None, [P, P, P, sz, u64]),
        "or_conjugate": (None, [P, P, sz, u64]),
        "or_commit": (None, [P, P, P, sz, sz, sz, sz, u64, ctypes.c_int]),
        "or_tensor": (None, [P, sz, u64, P]),
        "or_sum_chunked_raw": (u64, [P, sz, u64]),
        "or_sum_sq_chunked_raw": (u64, [P, sz, u64]),
        "or_sumcheck": (None, [P, sz, sz, P, sz, ctypes.c_int, u64, u64, P, P, P, P,
                               IP, ctypes.c_int]),
        "or_norm_check_3": (None, [P, sz, sz, sz, P, u64, P]),
        "or_norm_check_3_fq2": (None, [P, sz, sz, sz, P, u64, u64, P]),
        "or_decompose": (u64, [P, sz, sz, sz, sz, u64, u64, P]),
        "or_compose": (None, [P, sz, sz, sz, sz, u64, u64, P]),
        "or_compose_rows": (None, [P, sz, sz, sz, sz, u64, u64, P]),
        "or_fold": (None, [P, P, P, sz, sz, sz, sz, u64, ctypes.c_int]),
        "or_fq2_mul": (None, [u64, u64, u64, u64, u64, u64, P]),
        "or_prng_draw": (u64, [u64, u64, u64]),
        "or_sample_uniform": (None, [u64, u64, sz, u64, P]),
        "or_sample_uniform_at": (None, [u64, u64, sz, sz, u64, P]),
        "or_sample_bounded": (None, [u64, u64, sz, u64, u64, P]),
        "or_sample_bits": (None, [u64, u64, sz, ctypes.c_void_p]),
        "or_stream_key": (u64, [u64, u64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _p(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P)


def _u(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


# ---------------------------------------------------------------- scalars
def mulmod(a, b, q):
    return lib.or_mulmod(a, b, q)


def ntt_forward(x, q):
    x = _u(x).copy()
    lib.or_ntt_forward(_p(x), x.size, q)
    return x


def ntt_inverse(x, q):
    x = _u(x).copy()
    lib.or_ntt_inverse(_p(x), x.size, q)
    return x


def ntt_forward_direct(x, q):
    x = _u(x)
    out = np.zeros_like(x)
    lib.or_ntt_forward_direct(_p(x), _p(out), x.size, q)
    return out


def get_roots(n, q):
    out = np.zeros(n, np.uint64)
    lib.or_get_roots(n, q, _p(out))
    return out


def get_inv_roots(n, q):
    out = np.zeros(n, np.uint64)
    lib.or_get_inv_roots(n, q, _p(out))
    return out


def _per_elem(fn, a, q):
    a = _u(a).copy()
    d = a.shape[-1]
    flat = a.reshape(-1, d)
    for row in flat:
        fn(_p(row), d, q)
    return flat.reshape(a.shape)


def to_incomplete(a, q):
    return _per_elem(lib.or_to_incomplete, a, q)


def from_incomplete(a, q):
    return _per_elem(lib.or_from_incomplete, a, q)


def to_full(a, q):
    return _per_elem(lib.or_to_full, a, q)


def from_full(a, q):
    return _per_elem(lib.or_from_full, a, q)


def shift_factors(d, q):
    out = np.zeros(d // 2, np.uint64)
    lib.or_shift_factors(d, q, _p(out))
    return out


def ring_mul_incomplete(a, b, q):
    a, b = _u(a), _u(b)
    out = np.zeros_like(a)
    lib.or_ring_mul_incomplete(_p(a), _p(b), _p(out), a.size, q)
    return out


def ring_mul_naive(a, b, q):
    a, b = _u(a), _u(b)
    out = np.zeros_like(a)
    lib.or_ring_mul_naive(_p(a), _p(b), _p(out), a.size, q)
    return out


def conjugate(a, q):
    a = _u(a)
    out = np.zeros_like(a)
    lib.or_conjugate(_p(a), _p(out), a.size, q)
    return out


def commit(A, W, q, threads=0):
    """A [kappa][m][d], W [r][m][d] Coefficient -> [kappa][r][d] IncompleteNTT."""
    A, W = _u(A), _u(W)
    kappa, m, d = A.shape
    r = W.shape[0]
    assert W.shape == (r, m, d)
    out = np.zeros((kappa, r, d), np.uint64)
    lib.or_commit(_p(A), _p(W), _p(out), kappa, r, m, d, q, threads)
    return out


def complete_power_series(base, tensors, q):
    """complete_power_series (arithmetic.rs:158-176): base [m0][d], tensors
    [n][2][d] (IncompletePowerSeries order), Coefficient -> the largest
    expanded layer [m0 << n][d], Coefficient."""
    base, tensors = _u(base), _u(tensors)
    m0, d = base.shape
    n = tensors.shape[0]
    assert tensors.shape == (n, 2, d)
    out = np.zeros((m0 << n, d), np.uint64)
    lib.or_complete_power_series(_p(base), m0, _p(tensors), n, d, q, _p(out))
    return out


def dot_single_series(layer, W, q):
    """parallel_dot_single_series_matrix (arithmetic.rs:507-536): layer [m][d],
    W [r][m][d] Coefficient -> [r][d] IncompleteNTT."""
    layer, W = _u(layer), _u(W)
    m, d = layer.shape
    r = W.shape[0]
    assert W.shape == (r, m, d)
    out = np.zeros((r, d), np.uint64)
    lib.or_dot_single_series(_p(layer), _p(W), _p(out), r, m, d, q)
    return out


def project(W, is_zero, sign, q):
    """project (project.rs:52-98) before its commit: W [r][m][d], is_zero and
    sign [256][r*256] (0/1) -> [r*(m // (r*256))*256][d]."""
    W = _u(W)
    r, m, d = W.shape
    cw = r * 256
    z = np.ascontiguousarray(is_zero, np.uint8)
    sg = np.ascontiguousarray(sign, np.uint8)
    assert z.shape == (256, cw) and sg.shape == (256, cw)
    out = np.zeros((r * (m // cw) * 256, d), np.uint64)
    lib.or_project(_p(W), r, m, d, z.ctypes.data, sg.ctypes.data, q, _p(out))
    return out


def consecutive_powers(c, width, q):
    """consecutive_powers (project.rs:118-129): [1, c, c^2, ...] [width][d]
    Coefficient, each power the schoolbook product of the last with c."""
    c = _u(c)
    out = np.zeros((width, c.size), np.uint64)
    out[0, 0] = 1
    for k in range(1, width):
        out[k] = ring_mul_naive(out[k - 1], c, q)
    return out


def compute_commitment_keys(module, length, q):
    """compute_commitment_keys (crs.rs:43-72): row j = m_j, m_j^2, ...,
    m_j^length, each power the schoolbook product of the last with m_j.
    module [kappa][d] Coefficient -> [kappa][length][d]."""
    module = _u(module)
    kappa, d = module.shape
    out = np.zeros((kappa, length, d), np.uint64)
    for j in range(kappa):
        out[j, 0] = module[j]
        for i in range(1, length):
            out[j, i] = ring_mul_naive(out[j, i - 1], module[j], q)
    return out


def crs_power_series(row):
    """The PowerSeries of one CRS row (crs.rs:58-69): layers row[:len],
    row[:len/2], ... while the length is even, and below the layer of length
    2c the tensor [1, row[c - 1]] (1 as the ring's one).  Returns (layers,
    tensors), tensors in the order the reference pushes them."""
    d = row.shape[1]
    one = np.zeros(d, np.uint64)
    one[0] = 1
    layers, tensors = [], []
    cur = row.shape[0]
    while cur % 2 == 0:
        layers.append(row[:cur])
        cur //= 2
        tensors.append(np.stack([one, row[cur - 1]]))
    layers.append(row[:cur])
    return layers, tensors


def _signed_col_sums(is_zero, sign, pw, q):
    """c_l = sum_k C[k][l] pw[k] mod q with C from (is_zero, sign) [H][w]:
    the positive and negative parts summed apart in u64 (H q < 2^64)."""
    keep = np.asarray(is_zero) == 0
    pos = (keep & (np.asarray(sign) != 0)).astype(np.uint64).T
    neg = (keep & (np.asarray(sign) == 0)).astype(np.uint64).T
    a, b = (pos @ pw) % np.uint64(q), (neg @ pw) % np.uint64(q)
    return np.where(a >= b, a - b, a + np.uint64(q) - b).astype(np.uint64)


def batch_projections(P, W, is_zero, sign, c0, q):
    """batch_projections (project.rs:131-203), step by step as the reference
    takes them: c_l = sum_k C[k][l] c0^k (:154-166), both layers by
    complete_power_series with [1, 1] tensors (:168-181), both right-hand
    sides by dot_single_series (:183-188) and the assertion (:190-198).
    P [m][d], W [r][m][d], c0 [d] Coefficient.  Returns (projected_lhs,
    witness_lhs, projection_rhs [d], witness_rhs [r][d], ok)."""
    P, W = _u(P), _u(W)
    r, m, d = W.shape
    cw = r * 256
    pw = consecutive_powers(c0, 256, q)
    cj = _signed_col_sums(is_zero, sign, pw, q)
    one = np.zeros(d, np.uint64)
    one[0] = 1
    n_w = int(m // cw).bit_length() - 1
    n_p = n_w + r.bit_length() - 1
    wlhs = complete_power_series(cj, np.tile(one, (n_w, 2, 1)), q)
    plhs = complete_power_series(pw, np.tile(one, (n_p, 2, 1)), q)
    wrhs = dot_single_series(wlhs, W, q)
    prhs = dot_single_series(plhs, P[None], q)[0]
    total = np.zeros(d, np.uint64)
    for x in wrhs:
        total = (total + x) % np.uint64(q)
    return plhs, wlhs, prhs, wrhs, bool(np.array_equal(total, prhs))


def send_cross_terms_before_join(projected_lhs, witness_lhs, W, P, q):
    """send_cross_terms_before_join (project.rs:229-245): entry 0 of
    dot_single_series(projected_lhs, W) and of dot_single_series(witness_lhs,
    [P]), each [d] IncompleteNTT."""
    l0 = dot_single_series(projected_lhs, W, q)[0]
    l1 = dot_single_series(witness_lhs, _u(P)[None], q)[0]
    return l0, l1


def tensor(cs, q):
    cs = _u(cs)
    out = np.zeros(1 << cs.size, np.uint64)
    lib.or_tensor(_p(cs), cs.size, q, _p(out))
    return out


def sum_chunked_raw(x, q):
    x = _u(x)
    return lib.or_sum_chunked_raw(_p(x), x.size, q)


def sum_sq_chunked_raw(x, q):
    x = _u(x)
    return lib.or_sum_sq_chunked_raw(_p(x), x.size, q)


def sumcheck(w, cs, q, fq2=False, nr=5, threads=0):
    """w [rows][L]; cs [reps][nrounds][2].  Returns dict of numpy arrays."""
    w, cs = _u(w), _u(cs)
    rows, L = w.shape
    reps, nrounds, _ = cs.shape
    assert (1 << nrounds) == L
    claims = np.zeros(rows, np.uint64)
    msgs = np.zeros((reps, nrounds, rows, 3, 2), np.uint64)
    finals = np.zeros((reps, rows, 2), np.uint64)
    evals = np.zeros((reps, rows, 2), np.uint64)
    ok = np.zeros((reps, rows), np.int32)
    lib.or_sumcheck(_p(w), rows, L, _p(cs), reps, int(fq2), nr, q, _p(claims), _p(msgs),
                    _p(finals), _p(evals), ok.ctypes.data_as(IP), threads)
    return dict(claims=claims, msgs=msgs, finals=finals, evals=evals, ok=ok)


def norm_check_3(W, cs, q):
    W, cs = _u(W), _u(cs)
    rows, m, d = W.shape
    out = np.zeros((rows, d), np.uint64)
    lib.or_norm_check_3(_p(W), rows, m, d, _p(cs), q, _p(out))
    return out


def norm_check_3_fq2(W, cs2, q, nr=5):
    W, cs2 = _u(W), _u(cs2)
    rows, m, d = W.shape
    out = np.zeros((rows, 2, d), np.uint64)
    lib.or_norm_check_3_fq2(_p(W), rows, m, d, _p(cs2), nr, q, _p(out))
    return out


def decompose(W, parts, q, radix=0):
    W = _u(W)
    r, m, d = W.shape
    out = np.zeros((r * parts, m, d), np.uint64)
    radix = lib.or_decompose(_p(W), r, m, d, parts, radix, q, _p(out))
    return out, radix


def compose(x, parts, radix, q):
    """compose_with_radix_mod: [rows][n*parts][d] -> [rows][n][d]."""
    x = _u(x)
    rows, np_, d = x.shape
    n = np_ // parts
    out = np.zeros((rows, n, d), np.uint64)
    lib.or_compose(_p(x), rows, n, d, parts, radix, q, _p(out))
    return out


def compose_rows(limbs, parts, radix, q):
    """gadget recomposition of limb rows: [r*parts][m][d] -> [r][m][d]."""
    limbs = _u(limbs)
    rp, m, d = limbs.shape
    r = rp // parts
    out = np.zeros((r, m, d), np.uint64)
    lib.or_compose_rows(_p(limbs), r, m, d, parts, radix, q, _p(out))
    return out


def fold(W, C, q, threads=0):
    """W [r_in][m][d], C [r_out][r_in][d] -> [r_out][m][d] IncompleteNTT."""
    W, C = _u(W), _u(C)
    r_in, m, d = W.shape
    r_out = C.shape[0]
    out = np.zeros((r_out, m, d), np.uint64)
    lib.or_fold(_p(W), _p(C), _p(out), r_in, r_out, m, d, q, threads)
    return out


def fq2_mul(a, b, nr, q):
    out = np.zeros(2, np.uint64)
    lib.or_fq2_mul(a[0], a[1], b[0], b[1], nr, q, _p(out))
    return [int(out[0]), int(out[1])]


def prng_draw(seed, stream, idx):
    return lib.or_prng_draw(seed, stream, idx)


def sample_uniform(seed, stream, shape, q):
    n = int(np.prod(shape))
    out = np.zeros(n, np.uint64)
    lib.or_sample_uniform(seed, stream, n, q, _p(out))
    return out.reshape(shape)


def sample_uniform_at(seed, stream, start, shape, q):
    """Elements start, start + 1, .. of sample_uniform's stream."""
    n = int(np.prod(shape))
    out = np.zeros(n, np.uint64)
    lib.or_sample_uniform_at(seed, stream, start, n, q, _p(out))
    return out.reshape(shape)


def sample_bits(seed, stream, shape):
    """uint8 0/1, one bit per element of the stream"""
    n = int(np.prod(shape))
    out = np.zeros(n, np.uint8)
    lib.or_sample_bits(seed, stream, n, out.ctypes.data)
    return out.reshape(shape)


def stream_key(seed, stream):
    return int(lib.or_stream_key(seed, stream))


def sample_bounded(seed, stream, shape, bound, q):
    n = int(np.prod(shape))
    out = np.zeros(n, np.uint64)
    lib.or_sample_bounded(seed, stream, n, bound, q, _p(out))
    return out.reshape(shape)
