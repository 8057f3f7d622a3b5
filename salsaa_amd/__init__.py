"""salsaa_amd -- MI355X (gfx950) implementation of SALSAA's prover hot path.

The compute lives in libsalsaa_amd.so (hand-written HIP kernels behind the C
ABI declared in include/salsaa_amd.h).  This module is the host-side mirror
of the reference's operator API over that ABI: it loads the library (and
fails loudly if it has not been built), manages device buffers, and exposes
the reference's function names (parallel_dot_series_matrix, norm_check,
decomp_ell, parallel_dot_matrix_matrix, split) on HBM-resident matrices.
"""
import ctypes
import weakref
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsalsaa_amd.so")

Q = 1125899906840833  # largest prime < 2^50 with q = 1 mod 128
D = 64
NR = 5  # smallest quadratic non-residue mod Q

REP_COEFF, REP_FULL, REP_INCOMPLETE = 0, 1, 2

# sampler streams used by the configs (any distinct values work)
STREAM_A, STREAM_W, STREAM_NTT, STREAM_FOLD = 0xA, 0xB, 0xC, 0xF

_u64 = ctypes.c_uint64
_sz = ctypes.c_size_t
_vp = ctypes.c_void_p
_int = ctypes.c_int
_PP = ctypes.POINTER(ctypes.c_void_p)
_PSZ = ctypes.POINTER(ctypes.c_size_t)


class FoldOut(ctypes.Structure):
    """salsaa_fold_outputs."""
    _fields_ = [(n, ctypes.c_void_p) for n in ("folded", "folded_coeff", "rhs", "radix", "claims",
                                                 "msgs", "finals", "evals", "nc_rhs", "ok")]


class ServerOpts(ctypes.Structure):
    """salsaa_server_opts."""
    _fields_ = [("arena_elems", ctypes.c_size_t), ("max_batch_jobs", ctypes.c_size_t),
                ("max_batch_elems", ctypes.c_size_t), ("inflight", ctypes.c_int),
                ("reps", ctypes.c_size_t), ("fq2", ctypes.c_int), ("nr", ctypes.c_uint64),
                ("kchunk", ctypes.c_size_t)]


class ProtocolOpts(ctypes.Structure):
    """salsaa_protocol_opts."""
    _fields_ = [("fault_round", ctypes.c_size_t), ("fault_check", ctypes.c_int)]


class SlotT(ctypes.Structure):
    """salsaa_slot_t: device slots of one batch instance."""
    _fields_ = [(n, ctypes.c_void_p) for n in ("witness", "challenges", "commitment", "claims",
                                                 "msgs", "finals", "evals", "rhs", "ok")] + \
               [(n, ctypes.c_size_t) for n in ("m", "kappa", "nrounds")]


class SalsaaError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libsalsaa_amd.so is not built (expected at %s); run "
            "`python -c 'import __graft_entry__ as g; g.build()'` or "
            "`make -C salsaa_amd/csrc`" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        # legacy hexl_wrapper ABI
        "multiply_mod": (_u64, [_u64, _u64, _u64]),
        "power_mod": (_u64, [_u64, _u64, _u64]),
        "add_mod": (_u64, [_u64, _u64, _u64]),
        "sub_mod": (_u64, [_u64, _u64, _u64]),
        "inv_mod": (_u64, [_u64, _u64]),
        "eltwise_mult_mod": (None, [_vp, _vp, _vp, _u64, _u64]),
        "eltwise_fma_mod": (None, [_vp, _vp, _u64, _vp, _u64, _u64]),
        "eltwise_add_mod": (None, [_vp, _vp, _vp, _u64, _u64]),
        "eltwise_sub_mod": (None, [_vp, _vp, _vp, _u64, _u64]),
        "eltwise_reduce_mod": (None, [_vp, _vp, _u64, _u64]),
        "eltwise_reduce_mod_naive": (None, [_vp, _vp, _u64, _u64]),
        "get_roots": (ctypes.POINTER(_u64), [_u64, _u64]),
        "get_inv_roots": (ctypes.POINTER(_u64), [_u64, _u64]),
        "sum": (_u64, [_vp, _u64, _u64]),
        "sum_fast": (_u64, [_vp, _u64, _u64]),
        "sum_sq": (_u64, [_vp, _u64, _u64]),
        "sum_sq_fast": (_u64, [_vp, _u64, _u64]),
        "multiply_poly": (None, [_vp, _vp, _vp, _u64, _u64]),
        "polynomial_multiply_cyclotomic_mod": (None, [_vp, _vp, _vp, _u64, _u64]),
        "ntt_forward_in_place": (None, [_vp, _sz, _u64]),
        "ntt_inverse_in_place": (None, [_vp, _sz, _u64]),
        # device API
        "salsaa_last_error": (ctypes.c_char_p, []),
        "salsaa_version": (ctypes.c_char_p, []),
        "salsaa_ctx_create": (_int, [_u64, ctypes.c_uint32, _PP]),
        "salsaa_ctx_destroy": (_int, [_vp]),
        "salsaa_ctx_sync": (_int, [_vp]),
        "salsaa_ctx_trim": (_int, [_vp]),
        "salsaa_ctx_stream": (_vp, [_vp]),
        "salsaa_dev_alloc": (_int, [_vp, _sz, _PP]),
        "salsaa_dev_free": (_int, [_vp, _vp]),
        "salsaa_memcpy_h2d": (_int, [_vp, _vp, _vp, _sz]),
        "salsaa_memcpy_d2h": (_int, [_vp, _vp, _vp, _sz]),
        "salsaa_memcpy_d2d": (_int, [_vp, _vp, _vp, _sz]),
        "salsaa_memstream": (_int, [_vp, _vp, _sz, _int, _int, _int, _int]),
        "salsaa_event_create": (_int, [_PP]),
        "salsaa_event_destroy": (_int, [_vp]),
        "salsaa_event_record": (_int, [_vp, _vp]),
        "salsaa_event_elapsed_ms": (_int, [_vp, _vp, ctypes.POINTER(ctypes.c_float)]),
        "salsaa_sample_uniform": (_int, [_vp, _vp, _sz, _u64, _u64]),
        "salsaa_sample_bounded": (_int, [_vp, _vp, _sz, _u64, _u64, _u64]),
        "salsaa_sample_challenges": (_int, [_vp, _vp, _sz, _sz, _int, _u64]),
        "salsaa_convert": (_int, [_vp, _vp, _sz, _int, _int]),
        "salsaa_ring_mul": (_int, [_vp, _vp, _vp, _vp, _sz]),
        "salsaa_complete_power_series": (_int, [_vp, _vp, _sz, _vp, _sz, _vp]),
        "salsaa_parallel_dot_single_series_matrix": (_int, [_vp, _vp, _sz, _vp, _sz, _vp]),
        "salsaa_project": (_int, [_vp, _vp, _sz, _sz, _vp, _vp, _vp]),
        "salsaa_batch_projections": (_int, [_vp, _vp, _vp, _sz, _sz, _vp, _vp, _vp, _vp, _vp,
                                            _vp, _vp, _vp]),
        "salsaa_send_cross_terms_before_join": (_int, [_vp, _vp, _vp, _sz, _vp, _vp, _vp, _vp]),
        "salsaa_key_create": (_int, [_vp, _vp, _sz, _sz, _PP]),
        "salsaa_key_create_crs": (_int, [_vp, _vp, _sz, _sz, _PP]),
        "salsaa_key_crs_implicit": (_int, [_vp, _vp, _sz, _sz, _PP]),
        "salsaa_compute_commitment_keys": (_int, [_vp, _vp, _sz, _sz, _vp]),
        "salsaa_key_destroy": (_int, [_vp]),
        "salsaa_parallel_dot_series_matrix": (_int, [_vp, _vp, _vp, _sz, _sz, _vp]),
        "salsaa_split": (_int, [_vp, _vp, _vp, _sz, _sz, _vp]),
        "salsaa_norm_check": (_int, [_vp, _vp, _sz, _sz, _vp, _sz, _int, _u64, _vp, _vp, _vp,
                                     _vp, _vp, _vp]),
        "salsaa_norm_check_noalloc": (_int, [_vp, _vp, _sz, _sz, _vp, _sz, _int, _u64, _vp, _vp,
                                             _vp, _vp, _vp, _vp]),
        "salsaa_decompose": (_int, [_vp, _vp, _sz, _sz, _sz, _u64, _vp, _vp]),
        "salsaa_decomp_ell": (_int, [_vp, _vp, _vp, _sz, _sz, _sz, _u64, _vp, _vp, _vp]),
        "salsaa_compose": (_int, [_vp, _vp, _sz, _sz, _sz, _vp, _vp]),
        "salsaa_verify_decomp": (_int, [_vp, _vp, _vp, _sz, _sz, _sz, _vp, _int, _vp]),
        "salsaa_parallel_dot_matrix_matrix": (_int, [_vp, _vp, _sz, _sz, _vp, _sz, _vp]),
        "salsaa_key_view": (_int, [_vp, _sz, _sz, _PP]),
        "salsaa_batch_create": (_int, [_vp, _vp, _sz, _PSZ, _PSZ, _sz, _int, _u64, _int, _PP]),
        "salsaa_batch_destroy": (_int, [_vp]),
        "salsaa_batch_slots": (_int, [_vp, _sz, _int, ctypes.POINTER(SlotT)]),
        "salsaa_batch_step": (_int, [_vp, _int]),
        "salsaa_batch_step_plain": (_int, [_vp]),
        "salsaa_batch_phase": (_int, [_vp, _int]),
        "salsaa_fold_step": (_int, [_vp, _vp, _sz, _vp, _sz, _sz, _sz, _u64, _vp, _sz, _vp, _sz,
                                     _int, _u64, ctypes.POINTER(FoldOut)]),
        "salsaa_protocol": (_int, [_vp, _vp, _sz, _sz, _vp, _sz, _sz, _PSZ, _int, _u64, _vp, _sz,
                                    _vp, _sz, _PSZ, ctypes.POINTER(ctypes.c_int32),
                                    ctypes.POINTER(ctypes.c_double)]),
        "salsaa_protocol_ex": (_int, [_vp, _vp, _sz, _sz, _vp, _sz, _sz, _PSZ, _int, _u64, _vp,
                                       _sz, _vp, _sz, _PSZ, ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ProtocolOpts)]),
        "salsaa_norm_check_fs": (_int, [_vp, _vp, _sz, _sz, _sz, _int, _u64, ctypes.c_char_p, _vp,
                                        _vp, _vp, _vp, _vp, _vp, _vp]),
        "salsaa_transcript_norm_check": (_int, [_vp, _sz, _sz, _sz, _int, _u64, _int,
                                                ctypes.c_char_p, _vp, _vp, _vp, _vp, _vp, _vp,
                                                _vp, _sz, _PSZ]),
        "salsaa_transcript_prover_step": (_int, [_vp, _sz, _vp, _sz, _sz, _sz, _int, _u64, _int,
                                                 ctypes.c_char_p, _vp, _vp, _vp, _vp, _vp, _vp,
                                                 _vp, _sz, _PSZ]),
        "salsaa_transcript_protocol": (_int, [_vp, _sz, _sz, _int, _u64, ctypes.c_int32, _PSZ,
                                              _vp, _vp, _vp, _sz, _PSZ]),
        "salsaa_server_create": (_int, [_vp, _vp, _sz, ctypes.POINTER(ServerOpts), _PP]),
        "salsaa_server_submit": (_int, [_vp, _sz, _vp, _u64, _u64, ctypes.POINTER(ctypes.c_uint64)]),
        "salsaa_server_poll": (_int, [_vp, _u64, _vp, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_uint64)]),
        "salsaa_server_drain": (_int, [_vp]),
        "salsaa_server_layout": (_int, [_vp, _PSZ]),
        "salsaa_server_destroy": (_int, [_vp]),
        "salsaa_batch_tuning": (_int, [_vp, _PSZ, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_float),
                                       ctypes.POINTER(ctypes.c_float)]),
        "salsaa_batch_tuning_ntt": (_int, [_vp, ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_float)]),
        "salsaa_batch_set_ntt_schedule": (_int, [_vp, _int]),
        "salsaa_batch_tuning_grid": (_int, [_vp, ctypes.POINTER(ctypes.c_uint),
                                            ctypes.POINTER(ctypes.c_uint),
                                            ctypes.POINTER(ctypes.c_float)]),
        "salsaa_batch_set_ntt_grid": (_int, [_vp, ctypes.c_uint]),
        "salsaa_batch_tuning_step_grid": (_int, [_vp, ctypes.POINTER(ctypes.c_uint),
                                                 ctypes.POINTER(ctypes.c_uint),
                                                 ctypes.POINTER(ctypes.c_float)]),
        "salsaa_batch_tuning_commit": (_int, [_vp, ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_float)]),
        "salsaa_batch_set_commit_schedule": (_int, [_vp, _int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()

#: every C-ABI symbol declared in include/salsaa_amd.h
EXPORTED = None


def _check(status):
    if status != 0:
        msg = lib.salsaa_last_error()
        raise SalsaaError("salsaa call failed (%d): %s" % (status, msg.decode() if msg else ""))


class DeviceArray:
    """A device buffer of uint64 (or int32/uint32) owned by a Context, or a
    non-owning view of device memory owned elsewhere (DeviceArray.view)."""

    def __init__(self, ctx, shape, dtype=np.uint64):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.owner = True
        p = ctypes.c_void_p()
        _check(lib.salsaa_dev_alloc(ctx.h, max(self.nbytes, 1), ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def view(cls, ctx, ptr, shape, dtype=np.uint64):
        a = cls.__new__(cls)
        a.ctx, a.ptr, a.owner = ctx, ptr, False
        a.shape = tuple(int(s) for s in shape)
        a.dtype = np.dtype(dtype)
        a.nbytes = int(np.prod(a.shape)) * a.dtype.itemsize
        return a

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes, (arr.shape, self.shape)
        _check(lib.salsaa_memcpy_h2d(self.ctx.h, self.ptr, arr.ctypes.data, self.nbytes))
        return self

    def download(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib.salsaa_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.owner and self.ctx.h:
            lib.salsaa_dev_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Key:
    """Resident commitment key (A in the slot-major full-NTT layout)."""

    _close_order = 3

    def __init__(self, ctx, A_dev, kappa, m):
        self.ctx, self.kappa, self.m = ctx, kappa, m
        p = ctypes.c_void_p()
        _check(lib.salsaa_key_create(ctx.h, A_dev.ptr, kappa, m, ctypes.byref(p)))
        self.h = p.value
        ctx._own(self)

    @classmethod
    def crs_implicit(cls, ctx, module, m):
        """The same CRS key stored as its module only: commits generate the
        key rows in the kernel."""
        k = cls.__new__(cls)
        k.ctx, k.kappa, k.m = ctx, module.shape[0], m
        p = ctypes.c_void_p()
        _check(lib.salsaa_key_crs_implicit(ctx.h, module.ptr, k.kappa, m, ctypes.byref(p)))
        k.h = p.value
        ctx._own(k)
        return k

    @classmethod
    def crs(cls, ctx, module, m):
        """The commitment key of the power-series CRS of `module` [kappa][d]
        (Coefficient) with rows of length m, generated on the device
        (compute_commitment_keys, crs.rs:43-72)."""
        k = cls.__new__(cls)
        k.ctx, k.kappa, k.m = ctx, module.shape[0], m
        p = ctypes.c_void_p()
        _check(lib.salsaa_key_create_crs(ctx.h, module.ptr, k.kappa, m, ctypes.byref(p)))
        k.h = p.value
        ctx._own(k)
        return k

    def view(self, kappa, m):
        """The key's first kappa rows and m ring elements (shares storage;
        close the view before the key)."""
        v = Key.__new__(Key)
        v.ctx, v.kappa, v.m = self.ctx, kappa, m
        p = ctypes.c_void_p()
        _check(lib.salsaa_key_view(self.h, kappa, m, ctypes.byref(p)))
        v.h = p.value
        v._close_order = 2  # views before keys
        self.ctx._own(v)
        return v

    def close(self):
        if self.h:
            lib.salsaa_key_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


#: phases of salsaa_protocol's timers, in the order of its times[] output
PROTOCOL_PHASES = ("commit", "decomp", "norm_check", "split", "projection", "fold", "join",
                   "final")


class Context:
    """A modulus/degree context on the current GPU (one stream)."""

    def __init__(self, q=Q, d=D):
        p = ctypes.c_void_p()
        _check(lib.salsaa_ctx_create(q, d, ctypes.byref(p)))
        self.h, self.q, self.d = p.value, q, d
        # keys, batches and servers made on this context: closed before it
        self._deps = weakref.WeakSet()

    def _own(self, obj):
        self._deps.add(obj)

    def close(self):
        if self.h:
            # servers and batches first, then key views, then keys (a native
            # object must not outlive its context)
            for obj in sorted(list(self._deps), key=lambda o: o._close_order):
                obj.close()
            lib.salsaa_ctx_destroy(self.h)
            self.h = None

    def sync(self):
        _check(lib.salsaa_ctx_sync(self.h))

    def trim(self):
        """Free the device memory the context keeps for reuse (salsaa_ctx_trim)."""
        _check(lib.salsaa_ctx_trim(self.h))

    # ---- buffers
    def empty(self, shape, dtype=np.uint64):
        return DeviceArray(self, shape, dtype)

    def to_device(self, arr, dtype=np.uint64):
        arr = np.ascontiguousarray(arr, dtype=dtype)
        return DeviceArray(self, arr.shape, dtype).upload(arr)

    def copy(self, dst, src):
        assert dst.nbytes == src.nbytes
        _check(lib.salsaa_memcpy_d2d(self.h, dst.ptr, src.ptr, src.nbytes))
        return dst

    def memstream(self, buf, n16, rn, wn, inplace=False, blocks_per_cu=8):
        """Bandwidth probe (salsaa_memstream): rn read and wn write streams of
        n16 16-byte vectors in buf."""
        need = (rn + wn - (1 if inplace else 0)) * n16 * 16
        assert buf.nbytes >= need, "stream buffer too small"
        _check(lib.salsaa_memstream(self.h, buf.ptr, n16, rn, wn, int(inplace), blocks_per_cu))

    def uniform(self, shape, seed, stream):
        a = self.empty(shape)
        _check(lib.salsaa_sample_uniform(self.h, a.ptr, int(np.prod(shape)), seed, stream))
        return a

    def bounded(self, shape, bound, seed, stream):
        a = self.empty(shape)
        _check(lib.salsaa_sample_bounded(self.h, a.ptr, int(np.prod(shape)), bound, seed, stream))
        return a

    def challenges(self, reps, nrounds, fq2, seed):
        cs = self.empty((reps, nrounds, 2))
        _check(lib.salsaa_sample_challenges(self.h, cs.ptr, reps, nrounds, int(fq2), seed))
        return cs

    # ---- events
    def event(self):
        p = ctypes.c_void_p()
        _check(lib.salsaa_event_create(ctypes.byref(p)))
        return p.value

    def record(self, ev):
        _check(lib.salsaa_event_record(self.h, ev))

    @staticmethod
    def elapsed_ms(a, b):
        ms = ctypes.c_float()
        _check(lib.salsaa_event_elapsed_ms(a, b, ctypes.byref(ms)))
        return ms.value

    # ---- operators (reference names)
    def convert(self, arr, from_rep, to_rep):
        """to_{incomplete_ntt,ntt,coeff}_representation on every element, in place."""
        count = int(np.prod(arr.shape[:-1]))
        _check(lib.salsaa_convert(self.h, arr.ptr, count, from_rep, to_rep))
        return arr

    def ring_mul(self, a, b):
        out = self.empty(a.shape)
        _check(lib.salsaa_ring_mul(self.h, a.ptr, b.ptr, out.ptr, int(np.prod(a.shape[:-1]))))
        return out

    def complete_power_series(self, base, tensors):
        """The largest layer of the power series with base [m0][d] and length-2
        tensors [n][2][d] (complete_power_series, arithmetic.rs:158-176),
        Coefficient form: [m0 << n][d]."""
        m0, d = base.shape
        n = tensors.shape[0]
        assert d == self.d and tensors.shape == (n, 2, d)
        out = self.empty((m0 << n, d))
        _check(lib.salsaa_complete_power_series(self.h, base.ptr, m0, tensors.ptr, n, out.ptr))
        return out

    def parallel_dot_single_series_matrix(self, layer, W):
        """out[j] = sum_i layer[i] W[j][i] (arithmetic.rs:507-536): layer [m][d],
        W [r][m][d] Coefficient -> [r][d] IncompleteNTT."""
        m, d = layer.shape
        r = W.shape[0]
        assert d == self.d and W.shape == (r, m, d)
        out = self.empty((r, d))
        _check(lib.salsaa_parallel_dot_single_series_matrix(self.h, layer.ptr, m, W.ptr, r,
                                                            out.ptr))
        return out

    def project(self, W, is_zero, sign):
        """project (project.rs:52-98) before its commit: W [r][m][d], is_zero and
        sign device bytes [256][r*256] -> [r*(m // (r*256))*256][d]."""
        r, m, d = W.shape
        assert d == self.d and is_zero.shape == (256, r * 256) and sign.shape == is_zero.shape
        out = self.empty((r * (m // (r * 256)) * 256, d))
        _check(lib.salsaa_project(self.h, W.ptr, r, m, is_zero.ptr, sign.ptr, out.ptr))
        return out

    def batch_projections(self, P, W, is_zero, sign, c0, layers=True):
        """batch_projections (project.rs:131-203): P [m][d] = project(W) for
        W [r][m][d], the challenge bytes [256][r*256], c0 [d] (Coefficient) ->
        (projected_lhs [m][d] or None, witness_lhs [m][d] or None,
        projection_rhs [d], witness_rhs [r][d], ok [1] int32)."""
        r, m, d = W.shape
        assert d == self.d and P.shape == (m, d) and c0.shape == (d,)
        assert is_zero.shape == (256, r * 256) and sign.shape == is_zero.shape
        plhs = self.empty((m, d)) if layers else None
        wlhs = self.empty((m, d)) if layers else None
        prhs, wrhs = self.empty((d,)), self.empty((r, d))
        ok = self.empty((1,), np.int32)
        _check(lib.salsaa_batch_projections(self.h, P.ptr, W.ptr, r, m, is_zero.ptr, sign.ptr,
                                            c0.ptr, plhs.ptr if layers else None,
                                            wlhs.ptr if layers else None, prhs.ptr, wrhs.ptr,
                                            ok.ptr))
        return plhs, wlhs, prhs, wrhs, ok

    def send_cross_terms_before_join(self, projected_lhs, witness_lhs, W, P):
        """send_cross_terms_before_join (project.rs:229-245): (l0, l1), each [d]
        IncompleteNTT, from the layers [m][d], the folded witness W [r][m][d]
        (row 0 used) and the projection P [m][d]."""
        m, d = projected_lhs.shape
        assert witness_lhs.shape == (m, d) and W.shape[1:] == (m, d) and P.shape == (m, d)
        l0, l1 = self.empty((d,)), self.empty((d,))
        _check(lib.salsaa_send_cross_terms_before_join(self.h, projected_lhs.ptr, witness_lhs.ptr,
                                                       m, W.ptr, P.ptr, l0.ptr, l1.ptr))
        return l0, l1

    def key(self, A_dev):
        kappa, m, d = A_dev.shape
        assert d == self.d
        return Key(self, A_dev, kappa, m)

    def crs_key(self, module, m, implicit=False):
        """Key.crs: the power-series CRS key of module [kappa][d], length m
        (implicit: generated inside each commit, nothing stored)."""
        assert module.shape[1:] == (self.d,)
        return Key.crs_implicit(self, module, m) if implicit else Key.crs(self, module, m)

    def compute_commitment_keys(self, module, length):
        """compute_commitment_keys (crs.rs:43-72): [kappa][length][d]
        Coefficient, row j = m_j, m_j^2, ..., m_j^length."""
        kappa, d = module.shape
        assert d == self.d
        out = self.empty((kappa, length, d))
        _check(lib.salsaa_compute_commitment_keys(self.h, module.ptr, kappa, length, out.ptr))
        return out

    def parallel_dot_series_matrix(self, key, W, out=None):
        r, m, d = W.shape
        out = out or self.empty((key.kappa, r, d))
        _check(lib.salsaa_parallel_dot_series_matrix(self.h, key.h, W.ptr, r, m, out.ptr))
        return out

    def split(self, key, W, out=None):
        r, m, d = W.shape
        out = out or self.empty((key.kappa, 2 * r, d))
        _check(lib.salsaa_split(self.h, key.h, W.ptr, r, m, out.ptr))
        return out

    def norm_check(self, w, cs, fq2, nr=NR, want_rhs=True, outs=None):
        """norm_check_1 + norm_check_2 (+ verifier) + norm_check_3 for every
        row of w [rows][L] (any shape whose leading axis is the row)."""
        rows = w.shape[0]
        L = int(np.prod(w.shape[1:]))
        reps, nrounds, _ = cs.shape
        if outs is None:
            outs = dict(
                claims=self.empty((rows,)),
                msgs=self.empty((reps, nrounds, rows, 3, 2)),
                finals=self.empty((reps, rows, 2)),
                evals=self.empty((reps, rows, 2)),
                rhs=self.empty((reps, rows, 2, self.d)) if want_rhs else None,
                ok=self.empty((reps, rows), np.int32),
            )
        rhs = outs["rhs"].ptr if outs.get("rhs") is not None else None
        _check(lib.salsaa_norm_check(self.h, w.ptr, rows, L, cs.ptr, reps, int(fq2), nr,
                                     outs["claims"].ptr, outs["msgs"].ptr, outs["finals"].ptr,
                                     outs["evals"].ptr, rhs, outs["ok"].ptr))
        return outs

    def norm_check_noalloc(self, w, cs, fq2, nr=NR, want_rhs=True, outs=None):
        """norm_check through salsaa_norm_check_noalloc: the same outputs, with
        the library's fixed device scratch instead of allocated scratch (the
        output buffers here are still allocated by this mirror)."""
        rows = w.shape[0]
        L = int(np.prod(w.shape[1:]))
        reps, nrounds, _ = cs.shape
        if outs is None:
            outs = dict(
                claims=self.empty((rows,)),
                msgs=self.empty((reps, nrounds, rows, 3, 2)),
                finals=self.empty((reps, rows, 2)),
                evals=self.empty((reps, rows, 2)),
                rhs=self.empty((reps, rows, 2, self.d)) if want_rhs else None,
                ok=self.empty((reps, rows), np.int32),
            )
        rhs = outs["rhs"].ptr if outs.get("rhs") is not None else None
        _check(lib.salsaa_norm_check_noalloc(self.h, w.ptr, rows, L, cs.ptr, reps, int(fq2), nr,
                                             outs["claims"].ptr, outs["msgs"].ptr,
                                             outs["finals"].ptr, outs["evals"].ptr, rhs,
                                             outs["ok"].ptr))
        return outs

    def norm_check_fs(self, w, reps, fq2=True, nr=NR, seed=bytes(32), want_rhs=True):
        """The norm check with Fiat-Shamir challenges (salsaa_norm_check_fs):
        norm_check's outputs plus cs [reps][nrounds][2], the derived
        challenges."""
        rows = w.shape[0]
        L = int(np.prod(w.shape[1:]))
        n = L.bit_length() - 1
        seed = bytes(seed)
        assert len(seed) == 32
        o = dict(cs=self.empty((reps, n, 2)), claims=self.empty((rows,)),
                 msgs=self.empty((reps, n, rows, 3, 2)), finals=self.empty((reps, rows, 2)),
                 evals=self.empty((reps, rows, 2)),
                 rhs=self.empty((reps, rows, 2, self.d)) if want_rhs else None,
                 ok=self.empty((reps, rows), np.int32))
        _check(lib.salsaa_norm_check_fs(self.h, w.ptr, rows, L, reps, int(fq2), nr, seed,
                                        o["cs"].ptr, o["claims"].ptr, o["msgs"].ptr,
                                        o["finals"].ptr, o["evals"].ptr,
                                        o["rhs"].ptr if want_rhs else None, o["ok"].ptr))
        return o

    def _transcript(self, call, *args):
        n = ctypes.c_size_t()
        _check(call(self.h, *args, None, 0, ctypes.byref(n)))
        buf = self.empty((max(1, n.value),), np.uint8)
        _check(call(self.h, *args, buf.ptr, n.value, ctypes.byref(n)))
        out = buf.download()[: n.value].tobytes()
        buf.free()
        return out

    def transcript_norm_check(self, o, L, cs, fq2=True, nr=NR, fs=False, seed=None):
        """The kind-1 transcript bytes of a norm check's outputs o (claims,
        msgs, finals, evals, rhs or None) with its challenges cs."""
        reps, n, rows = o["msgs"].shape[:3]
        rhs = o.get("rhs")
        return self._transcript(lib.salsaa_transcript_norm_check, rows, L, reps, int(fq2), nr,
                                int(fs), bytes(seed) if seed else None, o["claims"].ptr,
                                o["msgs"].ptr, cs.ptr, o["finals"].ptr, o["evals"].ptr,
                                rhs.ptr if rhs is not None else None)

    def transcript_prover_step(self, commitment, kappa, o, L, cs, fq2=True, nr=NR, fs=False,
                               seed=None):
        """kind 2: commitment (its first kappa rows) then the norm check."""
        reps, n, rows = o["msgs"].shape[:3]
        rhs = o.get("rhs")
        return self._transcript(lib.salsaa_transcript_prover_step, kappa, commitment.ptr, rows, L,
                                reps, int(fq2), nr, int(fs), bytes(seed) if seed else None,
                                o["claims"].ptr, o["msgs"].ptr, cs.ptr, o["finals"].ptr,
                                o["evals"].ptr, rhs.ptr if rhs is not None else None)

    def transcript_protocol(self, out, kappa, nrounds, folding, seed):
        """kind 3 from protocol()'s result."""
        r, m, _ = out["W"].shape
        R = out["rhs"].shape[0]
        shape = (ctypes.c_size_t * 4)(r, m, R, r)
        return self._transcript(lib.salsaa_transcript_protocol, kappa, nrounds, int(folding), seed,
                                out["status"], shape, out["rhs"].ptr, out["W"].ptr)

    def decompose(self, W, parts, radix=0, out=None, d_radix=None):
        r, m, d = W.shape
        out = out or self.empty((r * parts, m, d))
        d_radix = d_radix or self.empty((1,))
        _check(lib.salsaa_decompose(self.h, W.ptr, r, m, parts, radix, out.ptr, d_radix.ptr))
        return out, d_radix

    def decomp_ell(self, key, W, parts, radix=0, new_W=None, rhs=None, d_radix=None):
        r, m, d = W.shape
        new_W = new_W or self.empty((r * parts, m, d))
        rhs = rhs or self.empty((key.kappa, r * parts, d))
        d_radix = d_radix or self.empty((1,))
        _check(lib.salsaa_decomp_ell(self.h, key.h, W.ptr, r, m, parts, radix, new_W.ptr, rhs.ptr,
                                     d_radix.ptr))
        return new_W, rhs, d_radix

    def compose(self, x, parts, d_radix):
        rows, np_, d = x.shape
        out = self.empty((rows, np_ // parts, d))
        _check(lib.salsaa_compose(self.h, x.ptr, rows, np_ // parts, parts, d_radix.ptr, out.ptr))
        return out

    def verify_decomp(self, rhs_dec, rhs, parts, d_radix, rows_mode=False, mismatch=None):
        """Counts mismatches of compose(rhs_dec) against rhs on the device."""
        if rows_mode:
            kappa, r = rhs.shape[0], rhs.shape[1]
        else:
            kappa, r = rhs.shape[0], rhs.shape[1]
        if mismatch is None:
            mismatch = self.to_device(np.zeros(1, np.uint32), np.uint32)
        _check(lib.salsaa_verify_decomp(self.h, rhs_dec.ptr, rhs.ptr, kappa, r, parts,
                                        d_radix.ptr, int(rows_mode), mismatch.ptr))
        return mismatch

    def parallel_dot_matrix_matrix(self, W, C, out=None):
        """fold: out[j][i] = sum_k W[k][i] * C[j][k] (IncompleteNTT)."""
        r_in, m, d = W.shape
        r_out = C.shape[0]
        out = out or self.empty((r_out, m, d))
        _check(lib.salsaa_parallel_dot_matrix_matrix(self.h, W.ptr, r_in, m, C.ptr, r_out,
                                                     out.ptr))
        return out

    fold = parallel_dot_matrix_matrix

    def fold_step(self, key, W, n, parts, C, cs, fq2=True, radix=0, nr=NR, outs=None):
        """Folding step of n instances: W [n*r][m][d], C [n][r_out][r*parts][d]
        (ternary challenges), cs [n][reps][log2(m d)][2].  Returns the fold
        output (IncompleteNTT), its Coefficient form, the commitment of the
        stacked folded witnesses and the norm check outputs of every row."""
        rows_in, m, d = W.shape
        r = rows_in // n
        _, r_out, rk, _ = C.shape
        assert rk == r * parts and C.shape[0] == n
        reps, nrounds = cs.shape[1], cs.shape[2]
        rows = n * r_out
        o = outs or dict(folded=self.empty((rows, m, d)), folded_coeff=self.empty((rows, m, d)),
                 rhs=self.empty((key.kappa, rows, d)), radix=self.empty((n,)),
                 claims=self.empty((rows,)), msgs=self.empty((reps, nrounds, rows, 3, 2)),
                 finals=self.empty((reps, rows, 2)), evals=self.empty((reps, rows, 2)),
                 nc_rhs=self.empty((reps, rows, 2, d)), ok=self.empty((reps, rows), np.int32))
        fo = FoldOut(*[o[k].ptr for k in ("folded", "folded_coeff", "rhs", "radix", "claims",
                                           "msgs", "finals", "evals", "nc_rhs", "ok")])
        _check(lib.salsaa_fold_step(self.h, key.h, n, W.ptr, r, m, parts, radix, C.ptr, r_out,
                                    cs.ptr, reps, int(fq2), nr, ctypes.byref(fo)))
        return o


    @staticmethod
    def protocol_shape(kappa, r, m, nrounds, parts=None, folding=False):
        """(rows, length) of the protocol's final witness and the statement
        rows R of its verifier rhs [R][rows][d]."""
        parts = list(parts) if parts is not None else [1] * nrounds
        rf, mf, R = r, m, kappa
        for i in range(nrounds):
            if not folding and parts[i] > 1:
                rf *= parts[i]
            R += 2
            if not folding:
                rf, mf = rf * 2, mf // 2
            R += 2
            rf = 2
            if folding:
                rf *= 2
                break
        return rf, mf, R

    def protocol(self, module, W, nrounds, parts=None, folding=False, seed=0, fault=None,
                 outs=None):
        """protocol (protocol.rs:125-483) on the device under the CRS of module
        [kappa][d], from the witness W [r][m][d] (Coefficient).  Returns a dict:
        status (0 when every verifier relation held, else 100 * round +
        check), W (final witness), rhs (verifier rhs [R][r][d]) and times (ms
        per phase, PROTOCOL_PHASES).  fault = (round, check) injects a fault
        for the verifier check `check` (salsaa_protocol_opts); outs = the W
        and rhs of an earlier call of the same shape, written again instead
        of allocating new outputs."""
        kappa = module.shape[0]
        r, m, d = W.shape
        parts = list(parts) if parts is not None else [1] * nrounds
        rf, mf, R = self.protocol_shape(kappa, r, m, nrounds, parts, folding)
        if outs is not None and outs["W"].shape == (rf, mf, d) and outs["rhs"].shape == (R, rf, d):
            Wout, rhs = outs["W"], outs["rhs"]
        else:
            Wout, rhs = self.empty((rf, mf, d)), self.empty((R, rf, d))
        pa = (ctypes.c_size_t * max(1, len(parts)))(*parts)
        shape = (ctypes.c_size_t * 4)()
        st = ctypes.c_int32(-1)
        times = (ctypes.c_double * 8)()
        opts = ProtocolOpts(fault[0], fault[1]) if fault else None
        _check(lib.salsaa_protocol_ex(self.h, module.ptr, kappa, m, W.ptr, r, nrounds, pa,
                                      int(folding), seed, Wout.ptr, Wout.nbytes // 8, rhs.ptr,
                                      rhs.nbytes // 8, shape, ctypes.byref(st), times,
                                      ctypes.byref(opts) if opts else None))
        assert tuple(shape) == (rf, mf, R, rf), tuple(shape)
        return dict(status=st.value, W=Wout, rhs=rhs,
                    times=dict(zip(PROTOCOL_PHASES, list(times))))


#: the batched commit's schedules, in salsaa_batch_set_commit_schedule's order
COMMIT_SCHEDULES = ("per_class", "mixed", "mixed_small")


class Batch:
    """Many independent prover steps (NTT forward + inverse, commit, norm
    check) over one resident key, in grouped launches (salsaa_batch_*)."""

    _close_order = 1

    def __init__(self, ctx, key, m, kappa, reps=2, fq2=True, nr=NR, tune=True):
        self.ctx, self.key = ctx, key
        self.n = len(m)
        self.m, self.kappa = [int(x) for x in m], [int(x) for x in kappa]
        self.reps, self.fq2 = reps, bool(fq2)
        self.krows = max(self.kappa)
        M = (ctypes.c_size_t * self.n)(*self.m)
        K = (ctypes.c_size_t * self.n)(*self.kappa)
        p = ctypes.c_void_p()
        _check(lib.salsaa_batch_create(ctx.h, key.h, self.n, M, K, reps, int(fq2), nr, int(tune),
                                       ctypes.byref(p)))
        self.h = p.value
        ctx._own(self)

    def slot(self, i, plain=False):
        """Device views of instance i's witness, challenges and outputs
        (output shapes as norm_check's with one row)."""
        st = SlotT()
        _check(lib.salsaa_batch_slots(self.h, i, int(plain), ctypes.byref(st)))
        c, m, n, r = self.ctx, st.m, st.nrounds, self.reps
        V = DeviceArray.view
        return dict(
            witness=V(c, st.witness, (m, D)),
            challenges=V(c, st.challenges, (r, n, 2)),
            commitment=V(c, st.commitment, (self.krows, 1, D)),
            claims=V(c, st.claims, (1,)),
            msgs=V(c, st.msgs, (r, n, 1, 3, 2)),
            finals=V(c, st.finals, (r, 1, 2)),
            evals=V(c, st.evals, (r, 1, 2)),
            rhs=V(c, st.rhs, (r, 1, 2, D)),
            ok=V(c, st.ok, (r, 1), np.int32),
            m=m, kappa=st.kappa, nrounds=n)

    def step(self, graph=True):
        _check(lib.salsaa_batch_step(self.h, int(graph)))

    def step_plain(self):
        _check(lib.salsaa_batch_step_plain(self.h))

    # the batched step's phases in order (forward NTT, the commit's digit
    # planes, inverse NTT and the witnesses' short copy in one pass; commit;
    # norm checks from the short copy, valid after the first phase over the
    # current witnesses), and the alternatives, which are not part of the
    # step: the norm checks from the witnesses alone and the unfused passes
    # of the first phase
    PHASES = ("ntt_roundtrip_planes", "commit", "norm_check")
    UNFUSED = ("ntt_fwd", "ntt_fwd_planes", "ntt_inv", "planes", "norm_check_witness")
    _PHASE_CODE = {"ntt_roundtrip_planes": 0, "ntt_inv": 1, "planes": 2, "commit": 3,
                   "norm_check_witness": 4, "ntt_fwd": 5, "ntt_fwd_planes": 6, "norm_check": 7}

    def phase(self, name):
        """Run one phase eagerly by name (PHASES or UNFUSED)."""
        _check(lib.salsaa_batch_phase(self.h, self._PHASE_CODE[name]))

    def arena(self):
        """The witness arena: every instance's witness back to back (sorted
        by length), each padded with zeros to whole tiles of ring
        elements, as one [total][d] view (tiles of 128, the fused forward
        NTT + planes kernel's)."""
        p = min((self.slot(i)["witness"] for i in range(self.n)), key=lambda a: a.ptr)
        total = sum((m + 127) // 128 * 128 for m in self.m)
        return DeviceArray.view(self.ctx, p.ptr, (total, D))

    def tuning(self):
        kc, cpt = ctypes.c_size_t(), ctypes.c_int()
        tc, tn = ctypes.c_float(), ctypes.c_float()
        _check(lib.salsaa_batch_tuning(self.h, ctypes.byref(kc), ctypes.byref(cpt),
                                       ctypes.byref(tc), ctypes.byref(tn)))
        x, tl, td = ctypes.c_int(), ctypes.c_float(), ctypes.c_float()
        _check(lib.salsaa_batch_tuning_ntt(self.h, ctypes.byref(x), ctypes.byref(tl),
                                           ctypes.byref(td)))
        g, gc, gt = ctypes.c_uint(), (ctypes.c_uint * 3)(), (ctypes.c_float * 3)()
        _check(lib.salsaa_batch_tuning_grid(self.h, ctypes.byref(g), gc, gt))
        cm, cms = ctypes.c_int(), (ctypes.c_float * 3)()
        _check(lib.salsaa_batch_tuning_commit(self.h, ctypes.byref(cm), cms))
        g2, g2c, g2t = ctypes.c_uint(), (ctypes.c_uint * 3)(), (ctypes.c_float * 3)()
        _check(lib.salsaa_batch_tuning_step_grid(self.h, ctypes.byref(g2), g2c, g2t))
        return dict(kchunk=kc.value, cpt_first=cpt.value, commit_ms=tc.value,
                    norm_check_ms=tn.value, ntt_schedule=("lds", "dpp")[x.value],
                    ntt_pass_ms={"lds": tl.value, "dpp": td.value}, ntt_blocks=g.value,
                    ntt_blocks_ms={str(gc[i]): gt[i] for i in range(3) if gc[i]},
                    step_second_pass_blocks=g2.value,
                    step_ms_by_second_pass_blocks={str(g2c[i]): g2t[i] for i in range(3)
                                                   if g2c[i]},
                    commit_schedule=COMMIT_SCHEDULES[cm.value],
                    commit_schedule_ms={n: cms[i] for i, n in enumerate(COMMIT_SCHEDULES)})

    def set_ntt_schedule(self, name):
        """The NTT pass's exchange: "lds" or "dpp" (quad permutes)."""
        _check(lib.salsaa_batch_set_ntt_schedule(self.h, {"lds": 0, "dpp": 1}[name]))

    def set_ntt_grid(self, blocks):
        """The NTT pass's persistent blocks (0: one per CU)."""
        _check(lib.salsaa_batch_set_ntt_grid(self.h, int(blocks)))

    def set_commit_schedule(self, name):
        """The commit's launches (COMMIT_SCHEDULES): "per_class", "mixed" (one
        launch over every class) or "mixed_small" (one launch over the classes
        of kappa <= 18 beside the others'); the mixed ones need a stored key."""
        _check(lib.salsaa_batch_set_commit_schedule(self.h, COMMIT_SCHEDULES.index(name)))

    def close(self):
        if self.h:
            lib.salsaa_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching commit + norm-check service (salsaa_server_*).
    max_batch_jobs=1, inflight=1 serves each job alone (the plain path)."""

    _close_order = 0

    def __init__(self, ctx, key, kappa, arena_elems, max_batch_jobs, max_batch_elems,
                 inflight=2, reps=2, fq2=True, nr=NR, kchunk=64):
        self.ctx, self.key, self.kappa, self.reps = ctx, key, kappa, reps
        o = ServerOpts(arena_elems, max_batch_jobs, max_batch_elems, inflight, reps, int(fq2), nr,
                       kchunk)
        p = ctypes.c_void_p()
        _check(lib.salsaa_server_create(ctx.h, key.h, kappa, ctypes.byref(o), ctypes.byref(p)))
        self.h = p.value
        ctx._own(self)
        offs = (ctypes.c_size_t * 8)()
        _check(lib.salsaa_server_layout(self.h, offs))
        self.off = list(offs)
        self._keep = {}

    def submit(self, m, seed, bound=16, witness=None):
        """Queue a job; witness (host [m][d] uint64) or None (sampled from seed)."""
        jid = ctypes.c_uint64()
        ptr = None
        if witness is not None:
            witness = np.ascontiguousarray(witness, dtype=np.uint64)
            assert witness.shape == (m, D)
            ptr = witness.ctypes.data
        _check(lib.salsaa_server_submit(self.h, m, ptr, seed, bound, ctypes.byref(jid)))
        if witness is not None:
            self._keep[jid.value] = witness
        return jid.value

    def poll(self, jid, m):
        """None while pending; else the job's outputs, times and batch."""
        rec = np.empty(self.off[7], np.uint64)
        t = (ctypes.c_double * 4)()
        bi = (ctypes.c_uint64 * 2)()
        st = lib.salsaa_server_poll(self.h, jid, rec.ctypes.data, t, bi)
        if st < 0:
            _check(st)
        if st == 0:
            return None
        self._keep.pop(jid, None)
        n = (m * D).bit_length() - 1
        r, o = self.reps, self.off
        return dict(
            commitment=rec[o[0]: o[0] + self.kappa * D].reshape(self.kappa, 1, D),
            claims=rec[o[1]: o[1] + 1],
            msgs=rec[o[2]: o[2] + r * n * 6].reshape(r, n, 1, 3, 2),
            finals=rec[o[3]: o[3] + 2 * r].reshape(r, 1, 2),
            evals=rec[o[4]: o[4] + 2 * r].reshape(r, 1, 2),
            rhs=rec[o[5]: o[5] + 2 * D * r].reshape(r, 1, 2, D),
            ok=rec[o[6]:].view(np.int32)[:r].reshape(r, 1),
            times=list(t), batch=int(bi[0]), batch_jobs=int(bi[1]))

    def drain(self):
        _check(lib.salsaa_server_drain(self.h))

    def close(self):
        if self.h:
            lib.salsaa_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def header_symbols():
    """Names of every function declared in include/salsaa_amd.h."""
    import re
    root = os.path.dirname(_HERE)
    with open(os.path.join(root, "include", "salsaa_amd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)) -
                  {"if", "defined"})
