// api.hip -- the batched device C ABI: contexts, keys, memory helpers and the
// operator entry points that replace the reference's arithmetic.rs and
// subroutines/*.rs functions on HBM-resident matrices.
#include <cstdlib>
#include <cstring>
#include <new>

#include "handles.hpp"

// Scratch of salsaa_norm_check_noalloc (the message accumulators and the
// folded witnesses between passes of one norm check): a fixed array that
// every device the library runs on holds until the process ends.  Host code
// never passes `scratch` itself to a kernel: it uses the device address that
// hipGetSymbolAddress gives for each device (noalloc_scratch below).
static constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace salsaa {

namespace {
thread_local std::string tl_error;

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kMaxDevices] = {};

// Resolves the device address of `scratch` on device `dev`, which must be
// current; once per device, at context creation.
int noalloc_scratch_init(int dev) {
    if (dev < 0 || dev >= kMaxDevices) {
        tl_error = "device index out of range";
        return SALSAA_ERR_ARG;
    }
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            tl_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return SALSAA_ERR_HIP;
        }
        g_noalloc_ptr[dev] = p;
    }
    return SALSAA_OK;
}

void *noalloc_scratch(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    return g_noalloc_ptr[dev];
}

uint64_t h_min_root(uint64_t order, uint64_t q) {
    uint64_t w = 0;
    for (uint64_t g = 2; g < q && !w; g++) {
        uint64_t c = h_powmod(g, (q - 1) / order, q);
        if (h_powmod(c, order / 2, q) == q - 1) w = c;
    }
    if (!w) return 0;
    uint64_t w2 = h_mulmod(w, w, q), cur = w, best = w;
    for (uint64_t i = 0; i < order / 2; i++) {
        if (cur < best) best = cur;
        cur = h_mulmod(cur, w2, q);
    }
    return best;
}

uint64_t h_brv(uint64_t x, int bits) {
    uint64_t r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1ULL) << (bits - 1 - i);
    return r;
}

bool is_prime_u64(uint64_t n) {
    if (n < 2) return false;
    for (uint64_t p : {2ULL, 3ULL, 5ULL, 7ULL, 11ULL, 13ULL, 17ULL, 19ULL, 23ULL, 29ULL, 31ULL, 37ULL}) {
        if (n % p == 0) return n == p;
    }
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        s++;
    }
    for (uint64_t a : {2ULL, 3ULL, 5ULL, 7ULL, 11ULL, 13ULL, 17ULL, 19ULL, 23ULL, 29ULL, 31ULL, 37ULL}) {
        uint64_t x = h_powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int r = 1; r < s; r++) {
            x = h_mulmod(x, x, n);
            if (x == n - 1) {
                comp = false;
                break;
            }
        }
        if (comp) return false;
    }
    return true;
}

// HEXL-convention tables for n = 32 and n = 64 plus the slot moduli omega.
// The length-n negacyclic NTT's twiddles (HEXL convention: psi the minimal
// primitive 2n-th root, tab[fwd_off + i] = psi^brv(i)), their inverses with
// the last inverse stage's twiddle pre-scaled by n^{-1}, and n^{-1}.
static std::vector<uint64_t> fill_roots(uint64_t q, int n, std::vector<double> &tab, int fwd_off,
                                        int inv_off, int ninv_off) {
    int bits = 0;
    while ((1 << bits) < n) bits++;
    uint64_t psi = h_min_root(2 * (uint64_t)n, q);
    std::vector<uint64_t> roots(n);
    uint64_t p = 1;
    for (int i = 0; i < n; i++) {
        roots[h_brv(i, bits)] = p;
        p = h_mulmod(p, psi, q);
    }
    uint64_t ninv = h_invmod((uint64_t)n, q);
    for (int i = 0; i < n; i++) {
        tab[fwd_off + i] = (double)roots[i];
        tab[inv_off + i] = (double)h_invmod(roots[i], q);
    }
    // last inverse stage uses inv(roots[1]) * n^{-1}
    tab[inv_off + 0] = (double)h_mulmod(h_invmod(roots[1], q), ninv, q);
    tab[ninv_off] = (double)ninv;
    return roots;
}

// Tables of a degree-d context, d > 64 (WideTab layout); *full_ok says
// whether the length-d Full transform exists (2d | q - 1).
void build_wide_tables(uint64_t q, int d, std::vector<double> &tab, bool *full_ok) {
    const WideTab W(d);
    tab.assign(W.size, 0.0);
    fill_roots(q, W.n, tab, W.fwd_h, W.inv_h, W.ninv_h);
    *full_ok = (q - 1) % (2 * (uint64_t)d) == 0;
    if (*full_ok) fill_roots(q, d, tab, W.fwd_f, W.inv_f, W.ninv_f);
    // slot moduli: NTT_n of Y evaluates Y at the slot roots psi_d^(2 brv(j) + 1)
    int bits = 0;
    while ((1 << bits) < W.n) bits++;
    const uint64_t psi = h_min_root((uint64_t)d, q);
    for (int j = 0; j < W.n; j++)
        tab[W.omega + j] = (double)h_powmod(psi, 2 * h_brv(j, bits) + 1, q);
}

void build_tables(uint64_t q, std::vector<double> &tab) {
    tab.assign(TAB_SIZE, 0.0);
    auto fill = [&](int n, int fwd_off, int inv_off, int ninv_off) {
        return fill_roots(q, n, tab, fwd_off, inv_off, ninv_off);
    };
    std::vector<uint64_t> r32 = fill(32, TAB_FWD32, TAB_INV32, TAB_NINV32);
    fill(64, TAB_FWD64, TAB_INV64, TAB_NINV64);
    // omega = NTT_32(Y): the forward transform of (0, 1, 0, ...) is
    // omega_j = psi^(2 brv(j) + 1) (evaluation at the slot roots)
    uint64_t psi32 = h_min_root(64, q);
    for (int j = 0; j < 32; j++) tab[TAB_OMEGA + j] = (double)h_powmod(psi32, 2 * h_brv(j, 5) + 1, q);
    (void)r32;
    // full slots from the IncompleteNTT halves: slot s evaluates at
    // zeta_s = psi_128^(2 brv(s) + 1); a(zeta) = e(zeta^2) + zeta o(zeta^2)
    // and zeta^2 is one of the 32-point slot roots omega_j
    {
        const uint64_t psi128 = h_min_root(128, q);
        for (int s = 0; s < 64; s++) {
            const uint64_t z = h_powmod(psi128, 2 * h_brv(s, 6) + 1, q), z2 = h_mulmod(z, z, q);
            int zj = -1;
            for (int j = 0; j < 32; j++)
                if ((uint64_t)tab[TAB_OMEGA + j] == z2) zj = j;
            tab[TAB_ZETA + s] = (double)z;
            tab[TAB_ZJ + s] = (double)zj;  // always found: zeta^2 is a root of Y^32 + 1
        }
    }
    // commit output map: a vector of 64 full-NTT slot values to the
    // IncompleteNTT form of the ring element it represents (linear in the slots)
    for (int t = 0; t <= 12; t++) tab[TAB_W256 + t] = (double)h_powmod(2, 8 * (uint64_t)t, q);
    for (int s = 0; s < 64; s++) {
        uint64_t e[64] = {0}, ev[32], od[32];
        e[s] = 1;
        ntt_inverse_in_place(e, 64, q);
        for (int i = 0; i < 32; i++) {
            ev[i] = e[2 * i];
            od[i] = e[2 * i + 1];
        }
        ntt_forward_in_place(ev, 32, q);
        ntt_forward_in_place(od, 32, q);
        for (int c = 0; c < 32; c++) {
            tab[TAB_FIN + s * 64 + c] = (double)ev[c];
            tab[TAB_FIN + s * 64 + 32 + c] = (double)od[c];
        }
    }
}

}  // namespace

void set_error(const std::string &msg) { tl_error = msg; }

int scratch_get(Context *ctx, int slot, size_t bytes, void **out) {
    ScratchBuf &b = ctx->scratch[slot];
    if (b.bytes < bytes) {
        if (b.ptr) SALSAA_HIP(hipFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
        size_t want = bytes + bytes / 8 + 256;
        hipError_t e = hipMalloc(&b.ptr, want);
        if (e != hipSuccess) {
            set_error(std::string("scratch allocation failed: ") + hipGetErrorString(e));
            return SALSAA_ERR_NOMEM;
        }
        b.bytes = want;
    }
    *out = b.ptr;
    return SALSAA_OK;
}

__global__ void k_fill_zero(uint64_t *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// rn read streams and wn write streams of n 16-byte vectors, back to back in
// buf (reads first); with inplace the first write stream is read stream 0.
// Vector i of every write stream gets the xor of vector i of the read streams
// plus the stream index, so no load or store can be dropped.
__global__ __launch_bounds__(256) void k_memstream(ulonglong2 *buf, size_t n, int rn, int wn,
                                                   int inplace) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    // write stream w >= 1 follows the read streams (w >= 0 without inplace)
    ulonglong2 *wb = buf + ((size_t)rn - (inplace ? 1 : 0)) * n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        unsigned long long a = 0, b = 0;
        for (int r = 0; r < rn; ++r) {
            const ulonglong2 v = buf[(size_t)r * n + i];
            a ^= v.x;
            b ^= v.y;
        }
        for (int w = 0; w < wn; ++w) {
            ulonglong2 *p = (inplace && w == 0) ? buf + i : wb + (size_t)w * n + i;
            *p = make_ulonglong2(a + (unsigned)w, b);
        }
    }
}

int zero_row(Context *ctx, size_t bytes, const int8_t **out) {
    ScratchBuf &b = ctx->zeros;
    if (b.bytes < bytes) {
        if (b.ptr) SALSAA_HIP(hipFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
        size_t want = (bytes + 4095) / 4096 * 4096 * 2;
        hipError_t e = hipMalloc(&b.ptr, want);
        if (e != hipSuccess) {
            set_error(std::string("zero buffer allocation failed: ") + hipGetErrorString(e));
            return SALSAA_ERR_NOMEM;
        }
        hipLaunchKernelGGL(k_fill_zero, dim3(ceil_div(want / 8, 256)), dim3(256), 0, ctx->stream,
                           (uint64_t *)b.ptr, want / 8);
        SALSAA_HIP(hipGetLastError());
        b.bytes = want;
    }
    *out = (const int8_t *)b.ptr;
    return SALSAA_OK;
}

}  // namespace salsaa

using namespace salsaa;

extern "C" {

const char *salsaa_last_error(void) { return tl_error.c_str(); }

const char *salsaa_version(void) { return "salsaa_amd 0.1 gfx950"; }

int salsaa_ctx_create(uint64_t q, uint32_t d, salsaa_ctx **out) {
    SALSAA_CHECK(out != nullptr, "null output pointer");
    SALSAA_CHECK(d == 64 || d == 128 || d == 256 || d == 512,
                 "device kernels support d = 64, 128, 256 and 512");
    SALSAA_CHECK(q > 2 && q < (1ULL << 50), "q must be below 2^50");
    SALSAA_CHECK((q - 1) % 128 == 0, "q must be 1 mod 128 (128-th roots of unity)");
    SALSAA_CHECK((q - 1) % d == 0, "q must be 1 mod d (the IncompleteNTT's d-th roots of unity)");
    SALSAA_CHECK(is_prime_u64(q), "q must be prime");
    salsaa_ctx *c = new (std::nothrow) salsaa_ctx();
    if (!c) return SALSAA_ERR_NOMEM;
    Context &x = c->c;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error(std::string("no HIP device: ") + hipGetErrorString(e));
        delete c;
        return SALSAA_ERR_HIP;
    }
    x.device = dev;
    if (hipDeviceGetAttribute(&x.ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        x.ncu < 1)
        x.ncu = 256;
    x.q = q;
    x.d = (int)d;
    x.qd = (double)q;
    x.qinv = 1.0 / (double)q;
    const char *path = getenv("SALSAA_COMMIT");
    x.mfma_commit = !(path && std::string(path) == "valu");
    const char *ppath = getenv("SALSAA_PROJECT");
    x.mfma_project = !(ppath && std::string(ppath) == "valu");
    const char *xpath = getenv("SALSAA_NTT_XCH");
    x.ntt_xch = (xpath && std::string(xpath) == "dpp") ? 1 : 0;
    build_tables(q, x.h_tables);
    std::vector<double> wtab;
    if (d > (uint32_t)kD) build_wide_tables(q, (int)d, wtab, &x.full_ok);
    if (hipStreamCreateWithFlags(&x.stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&x.d_tables, TAB_SIZE * sizeof(double)) != hipSuccess ||
        hipMemcpy(x.d_tables, x.h_tables.data(), TAB_SIZE * sizeof(double),
                  hipMemcpyHostToDevice) != hipSuccess ||
        (!wtab.empty() &&
         (hipMalloc(&x.d_wtab, wtab.size() * sizeof(double)) != hipSuccess ||
          hipMemcpy(x.d_wtab, wtab.data(), wtab.size() * sizeof(double), hipMemcpyHostToDevice) !=
              hipSuccess))) {
        set_error("context setup failed: " + std::string(hipGetErrorString(hipGetLastError())));
        if (x.stream) (void)hipStreamDestroy(x.stream);
        if (x.d_tables) (void)hipFree(x.d_tables);
        if (x.d_wtab) (void)hipFree(x.d_wtab);
        delete c;
        return SALSAA_ERR_HIP;
    }
    const int st = noalloc_scratch_init(dev);
    if (st != SALSAA_OK) {
        (void)hipStreamDestroy(x.stream);
        (void)hipFree(x.d_tables);
        if (x.d_wtab) (void)hipFree(x.d_wtab);
        delete c;
        return st;
    }
    *out = c;
    return SALSAA_OK;
}

int salsaa_ctx_destroy(salsaa_ctx *ctx) {
    if (!ctx) return SALSAA_OK;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    for (auto &b : ctx->c.scratch)
        if (b.ptr) (void)hipFree(b.ptr);
    if (ctx->c.zeros.ptr) (void)hipFree(ctx->c.zeros.ptr);
    for (auto &b : ctx->c.slabs) (void)hipFree(b.ptr);
    (void)hipFree(ctx->c.d_tables);
    if (ctx->c.d_wtab) (void)hipFree(ctx->c.d_wtab);
    (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
    return SALSAA_OK;
}

int salsaa_ctx_trim(salsaa_ctx *ctx) {
    CTX_OK(ctx);
    SALSAA_HIP(hipStreamSynchronize(ctx->c.stream));
    for (auto &b : ctx->c.scratch) {
        if (b.ptr) SALSAA_HIP(hipFree(b.ptr));
        b = ScratchBuf();
    }
    if (ctx->c.zeros.ptr) SALSAA_HIP(hipFree(ctx->c.zeros.ptr));
    ctx->c.zeros = ScratchBuf();
    for (auto &b : ctx->c.slabs) SALSAA_HIP(hipFree(b.ptr));
    ctx->c.slabs.clear();
    return SALSAA_OK;
}

int salsaa_ctx_sync(salsaa_ctx *ctx) {
    CTX_OK(ctx);
    SALSAA_HIP(hipStreamSynchronize(ctx->c.stream));
    return SALSAA_OK;
}

void *salsaa_ctx_stream(salsaa_ctx *ctx) { return ctx ? (void *)ctx->c.stream : nullptr; }

int salsaa_dev_alloc(salsaa_ctx *ctx, size_t bytes, void **d_ptr) {
    CTX_OK(ctx);
    SALSAA_CHECK(d_ptr != nullptr, "null output pointer");
    hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return SALSAA_ERR_NOMEM;
    }
    return SALSAA_OK;
}

int salsaa_dev_free(salsaa_ctx *ctx, void *d_ptr) {
    CTX_OK(ctx);
    if (d_ptr) SALSAA_HIP(hipFree(d_ptr));
    return SALSAA_OK;
}

int salsaa_memcpy_h2d(salsaa_ctx *ctx, void *d_dst, const void *h_src, size_t bytes) {
    CTX_OK(ctx);
    SALSAA_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
    SALSAA_HIP(hipStreamSynchronize(ctx->c.stream));
    return SALSAA_OK;
}

int salsaa_memcpy_d2h(salsaa_ctx *ctx, void *h_dst, const void *d_src, size_t bytes) {
    CTX_OK(ctx);
    SALSAA_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
    SALSAA_HIP(hipStreamSynchronize(ctx->c.stream));
    return SALSAA_OK;
}

int salsaa_memcpy_d2d(salsaa_ctx *ctx, void *d_dst, const void *d_src, size_t bytes) {
    CTX_OK(ctx);
    SALSAA_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->c.stream));
    return SALSAA_OK;
}

int salsaa_memstream(salsaa_ctx *ctx, void *d_buf, size_t n16, int rn, int wn, int inplace,
                     int blocks_per_cu) {
    CTX_OK(ctx);
    SALSAA_CHECK(d_buf != nullptr && n16 > 0, "empty stream buffer");
    SALSAA_CHECK(rn >= 0 && rn <= 16 && wn >= 0 && wn <= 16 && rn + wn > 0,
                 "stream counts must be 0..16, not both zero");
    SALSAA_CHECK(!inplace || (rn > 0 && wn > 0), "an in-place write stream needs a read stream");
    SALSAA_CHECK(blocks_per_cu >= 1 && blocks_per_cu <= 64, "blocks per CU must be 1..64");
    const unsigned g = (unsigned)(ctx->c.ncu * blocks_per_cu);
    hipLaunchKernelGGL(k_memstream, dim3(g), dim3(256), 0, ctx->c.stream,
                       static_cast<ulonglong2 *>(d_buf), n16, rn, wn, inplace);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int salsaa_event_create(void **ev) {
    SALSAA_CHECK(ev != nullptr, "null output pointer");
    hipEvent_t e;
    SALSAA_HIP(hipEventCreate(&e));
    *ev = (void *)e;
    return SALSAA_OK;
}

int salsaa_event_destroy(void *ev) {
    if (ev) SALSAA_HIP(hipEventDestroy((hipEvent_t)ev));
    return SALSAA_OK;
}

int salsaa_event_record(salsaa_ctx *ctx, void *ev) {
    CTX_OK(ctx);
    SALSAA_HIP(hipEventRecord((hipEvent_t)ev, ctx->c.stream));
    return SALSAA_OK;
}

int salsaa_event_elapsed_ms(void *start, void *stop, float *ms) {
    SALSAA_HIP(hipEventSynchronize((hipEvent_t)stop));
    SALSAA_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return SALSAA_OK;
}

int salsaa_sample_uniform(salsaa_ctx *ctx, uint64_t *d_out, size_t n, uint64_t seed,
                          uint64_t stream) {
    CTX_OK(ctx);
    return launch_sample_uniform(&ctx->c, d_out, n, seed, stream);
}

int salsaa_sample_bounded(salsaa_ctx *ctx, uint64_t *d_out, size_t n, uint64_t bound,
                          uint64_t seed, uint64_t stream) {
    CTX_OK(ctx);
    return launch_sample_bounded(&ctx->c, d_out, n, bound, seed, stream);
}

int salsaa_sample_challenges(salsaa_ctx *ctx, uint64_t *d_cs, size_t reps, size_t nrounds,
                             int fq2, uint64_t seed) {
    CTX_OK(ctx);
    return launch_challenges(&ctx->c, d_cs, reps, nrounds, fq2, seed);
}

int salsaa_convert(salsaa_ctx *ctx, uint64_t *d_data, size_t count, int from_rep, int to_rep) {
    CTX_OK(ctx);
    SALSAA_CHECK(from_rep >= 0 && from_rep <= 2 && to_rep >= 0 && to_rep <= 2,
                 "unknown representation");
    if (from_rep == to_rep) return SALSAA_OK;
    // NTT <-> IncompleteNTT goes through Coefficient, as the reference does
    // (cyclotomic_ring.rs:488-490, 519-521)
    if (from_rep != SALSAA_REP_COEFF)
        SALSAA_TRY(launch_ntt_any(&ctx->c, d_data, count, from_rep, 1));
    if (to_rep != SALSAA_REP_COEFF) SALSAA_TRY(launch_ntt_any(&ctx->c, d_data, count, to_rep, 0));
    return SALSAA_OK;
}

int salsaa_ring_mul(salsaa_ctx *ctx, const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_out,
                    size_t count) {
    CTX_OK(ctx);
    return launch_ring_mul(&ctx->c, d_a, d_b, d_out, count);
}

}  // extern "C"

// A key of kappa rows and length m whose device layout `fill` writes: the
// MFMA commit's digit planes (kappa <= kMfmaMaxKappa) or slot-major doubles.
template <typename Fill>
static int key_make(salsaa_ctx *ctx, size_t kappa, size_t m, salsaa_key **out, Fill fill) {
    CTX_OK(ctx);
    SALSAA_CHECK(out != nullptr, "null output pointer");
    SALSAA_CHECK(kappa >= 1 && m >= 1, "key needs kappa, m >= 1");
    salsaa_key *k = new (std::nothrow) salsaa_key();
    if (!k) return SALSAA_ERR_NOMEM;
    k->k.ctx = &ctx->c;
    k->k.kappa = kappa;
    k->k.m = m;
    k->k.a_rows = kappa;
    k->k.a_m = m;
    int s = SALSAA_OK;
    if (ctx->c.d != kD) {
        // degree d > 64: digit planes of the doubled rows over the d/2 slot pairs
        if (2 * kappa > kMfmaMaxKappa) {
            delete k;
            set_error("a key at degree d > 64 needs kappa <= 18");
            return SALSAA_ERR_ARG;
        }
        k->k.a_kt = (2 * kappa * 7 + 15) / 16;
        k->k.a_ktiles = plane_ktiles(m);
        hipError_t e = hipMalloc(&k->k.d_AB, wide_plane_bytes(ctx->c.d, kappa, m));
        if (e != hipSuccess) {
            delete k;
            set_error(std::string("key allocation: ") + hipGetErrorString(e));
            return SALSAA_ERR_NOMEM;
        }
        s = fill(true, k->k);
    } else if (ctx->c.mfma_commit && kappa <= kMfmaMaxKappa) {
        k->k.a_kt = (kappa * 7 + 15) / 16;
        k->k.a_ktiles = plane_ktiles(m);
        hipError_t e = hipMalloc(&k->k.d_AB, plane_bytes(kappa, m));
        if (e != hipSuccess) {
            delete k;
            set_error(std::string("key allocation: ") + hipGetErrorString(e));
            return SALSAA_ERR_NOMEM;
        }
        s = fill(true, k->k);
    } else {
        hipError_t e = hipMalloc(&k->k.d_A, (size_t)kD * kappa * m * sizeof(double));
        if (e != hipSuccess) {
            delete k;
            set_error(std::string("key allocation: ") + hipGetErrorString(e));
            return SALSAA_ERR_NOMEM;
        }
        s = fill(false, k->k);
    }
    if (s != SALSAA_OK) {
        (void)hipFree(k->k.d_A);
        (void)hipFree(k->k.d_AB);
        delete k;
        return s;
    }
    *out = k;
    return SALSAA_OK;
}

extern "C" {

int salsaa_key_create(salsaa_ctx *ctx, const uint64_t *d_A, size_t kappa, size_t m,
                      salsaa_key **out) {
    return key_make(ctx, kappa, m, out, [&](bool planes, Key &k) {
        if (ctx->c.d != kD) return launch_wide_planes(&ctx->c, d_A, kappa, m, k.d_AB, k.a_ktiles);
        return planes ? launch_ntt_full_to_planes(&ctx->c, d_A, kappa, m, kappa, m * kD, 0, k.d_AB,
                                                  k.a_ktiles)
                      : launch_ntt_full_to_slotmajor(&ctx->c, d_A, kappa, m, kappa, m * kD, 0,
                                                     k.d_A);
    });
}

int salsaa_key_create_crs(salsaa_ctx *ctx, const uint64_t *d_module, size_t kappa, size_t m,
                          salsaa_key **out) {
    CTX_D64(ctx, "salsaa_key_create_crs");
    return key_make(ctx, kappa, m, out, [&](bool planes, Key &k) {
        return planes ? launch_crs(&ctx->c, d_module, kappa, m, 2, k.d_AB, k.a_ktiles)
                      : launch_crs(&ctx->c, d_module, kappa, m, 1, k.d_A, 0);
    });
}

int salsaa_key_crs_implicit(salsaa_ctx *ctx, const uint64_t *d_module, size_t kappa, size_t m,
                            salsaa_key **out) {
    CTX_D64(ctx, "salsaa_key_crs_implicit");
    SALSAA_CHECK(out != nullptr, "null output pointer");
    SALSAA_CHECK(kappa >= 1 && kappa <= kMfmaMaxKappa && m >= 1,
                 "a generated CRS key needs 1 <= kappa <= 36 and m >= 1");
    salsaa_key *k = new (std::nothrow) salsaa_key();
    if (!k) return SALSAA_ERR_NOMEM;
    k->k.ctx = &ctx->c;
    k->k.kappa = kappa;
    k->k.m = m;
    k->k.a_rows = kappa;
    k->k.a_m = m;
    k->k.a_ktiles = plane_ktiles(m);  // the K extent the key covers (nothing stored)
    hipError_t e = hipMalloc(&k->k.d_mhat, kappa * kD * sizeof(uint64_t));
    if (e != hipSuccess) {
        delete k;
        set_error(std::string("key allocation: ") + hipGetErrorString(e));
        return SALSAA_ERR_NOMEM;
    }
    // the module in NTT form: each slot of a power is the slot's scalar power
    int st = SALSAA_OK;
    if (hipMemcpyAsync(k->k.d_mhat, d_module, kappa * kD * sizeof(uint64_t),
                       hipMemcpyDeviceToDevice, ctx->c.stream) != hipSuccess)
        st = SALSAA_ERR_HIP;
    if (st == SALSAA_OK) st = launch_ntt(&ctx->c, k->k.d_mhat, kappa, SALSAA_REP_FULL, 0);
    if (st != SALSAA_OK) {
        (void)hipFree(k->k.d_mhat);
        delete k;
        return st;
    }
    *out = k;
    return SALSAA_OK;
}

int salsaa_compute_commitment_keys(salsaa_ctx *ctx, const uint64_t *d_module, size_t kappa,
                                   size_t len, uint64_t *d_out) {
    CTX_D64(ctx, "salsaa_compute_commitment_keys");
    SALSAA_TRY(launch_crs(&ctx->c, d_module, kappa, len, 0, d_out, 0));
    return launch_ntt(&ctx->c, d_out, kappa * len, SALSAA_REP_FULL, 1);
}

int salsaa_key_view(const salsaa_key *key, size_t kappa, size_t m, salsaa_key **out) {
    SALSAA_CHECK(key != nullptr && out != nullptr, "null key or output pointer");
    SALSAA_CHECK(kappa >= 1 && kappa <= key->k.kappa, "view rows must be 1..kappa of the key");
    SALSAA_CHECK(m >= 1 && m <= key->k.m, "view length must be 1..m of the key");
    salsaa_key *v = new (std::nothrow) salsaa_key();
    if (!v) return SALSAA_ERR_NOMEM;
    v->k = key->k;
    v->k.kappa = kappa;
    v->k.m = m;
    v->k.owner = false;
    *out = v;
    return SALSAA_OK;
}

int salsaa_key_destroy(salsaa_key *key) {
    if (!key) return SALSAA_OK;
    if (key->k.owner) {
        (void)hipStreamSynchronize(key->k.ctx->stream);
        if (key->k.d_A) (void)hipFree(key->k.d_A);
        if (key->k.d_AB) (void)hipFree(key->k.d_AB);
        if (key->k.d_mhat) (void)hipFree(key->k.d_mhat);
    }
    delete key;
    return SALSAA_OK;
}

static int commit_rows(salsaa_ctx *ctx, const salsaa_key *key, const uint64_t *d_W, size_t rows,
                       size_t m, size_t row_lo, size_t stride_lo, size_t stride_hi,
                       uint64_t *d_out) {
    SALSAA_CHECK(key != nullptr, "null key");
    SALSAA_CHECK(key->k.ctx == &ctx->c, "key belongs to another context");
    SALSAA_CHECK(m >= 1 && m <= key->k.m, "witness longer than the key");
    SALSAA_CHECK(rows >= 1, "commit needs at least one witness row");
    if (ctx->c.d != kD) {
        // degree d > 64: the slot-pair commit over plain [rows][m][d] rows
        SALSAA_CHECK(key->k.d_AB != nullptr, "key has no digit planes");
        SALSAA_CHECK(row_lo == rows && stride_lo == m * (size_t)ctx->c.d && stride_hi == 0,
                     "only whole matrices commit at degree d > 64");
        const size_t ps = plane_ktiles(m);
        int8_t *wb = nullptr;
        SALSAA_TRY(scratch_get(&ctx->c, 6, wide_plane_bytes(ctx->c.d, rows, m), (void **)&wb));
        SALSAA_TRY(launch_wide_planes(&ctx->c, d_W, rows, m, wb, ps));
        return launch_commit_mfma_pairs(&ctx->c, key->k.d_AB, 2 * key->k.kappa, key->k.a_kt,
                                        key->k.a_ktiles, wb, 2 * rows, m, ps, d_out);
    }
    if (key->k.d_mhat) {
        // a generated CRS key: its rows are made inside the commit kernel
        const size_t ps = plane_ktiles(m);
        int8_t *wb = nullptr;
        SALSAA_TRY(scratch_get(&ctx->c, 6, plane_bytes(rows, m), (void **)&wb));
        SALSAA_TRY(launch_ntt_full_to_planes(&ctx->c, d_W, rows, m, row_lo, stride_lo, stride_hi,
                                             wb, ps));
        return launch_commit_crs(&ctx->c, key->k.d_mhat, key->k.kappa, wb, rows, m, ps, d_out);
    }
    if (key->k.d_AB) {
        const size_t ps = plane_ktiles(m);
        int8_t *wb = nullptr;
        SALSAA_TRY(scratch_get(&ctx->c, 6, plane_bytes(rows, m), (void **)&wb));
        SALSAA_TRY(launch_ntt_full_to_planes(&ctx->c, d_W, rows, m, row_lo, stride_lo, stride_hi,
                                             wb, ps));
        return launch_commit_mfma(&ctx->c, key->k.d_AB, key->k.kappa, key->k.a_kt,
                                  key->k.a_ktiles, wb, rows, m, ps, d_out);
    }
    double *w_sm = nullptr;
    SALSAA_TRY(scratch_get(&ctx->c, 6, (size_t)kD * rows * m * sizeof(double), (void **)&w_sm));
    SALSAA_TRY(launch_ntt_full_to_slotmajor(&ctx->c, d_W, rows, m, row_lo, stride_lo, stride_hi,
                                            w_sm));
    return launch_commit_mac(&ctx->c, key->k.d_A, key->k.kappa, key->k.a_rows, key->k.a_m, w_sm,
                             rows, m, d_out);
}

int salsaa_parallel_dot_series_matrix(salsaa_ctx *ctx, const salsaa_key *key,
                                      const uint64_t *d_W, size_t r, size_t m, uint64_t *d_out) {
    CTX_OK(ctx);
    return commit_rows(ctx, key, d_W, r, m, r, m * (size_t)ctx->c.d, 0, d_out);
}

int salsaa_complete_power_series(salsaa_ctx *ctx, const uint64_t *d_base, size_t m0,
                                 const uint64_t *d_tensors, size_t n, uint64_t *d_out) {
    CTX_D64(ctx, "salsaa_complete_power_series");
    SALSAA_CHECK(m0 >= 1 && n <= 40 && ((m0 << n) >> n) == m0, "power series size out of range");
    // copies of the base and the tensors in NTT form, where each product is pointwise
    uint64_t *tmp = nullptr;
    const size_t nin = m0 + 2 * n;
    SALSAA_TRY(scratch_get(&ctx->c, 9, nin * kD * sizeof(uint64_t), (void **)&tmp));
    SALSAA_HIP(hipMemcpyAsync(tmp, d_base, m0 * kD * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                              ctx->c.stream));
    if (n)
        SALSAA_HIP(hipMemcpyAsync(tmp + m0 * kD, d_tensors, 2 * n * kD * sizeof(uint64_t),
                                  hipMemcpyDeviceToDevice, ctx->c.stream));
    SALSAA_TRY(launch_ntt(&ctx->c, tmp, nin, SALSAA_REP_FULL, 0));
    SALSAA_TRY(launch_series_expand(&ctx->c, tmp, m0, tmp + m0 * kD, n, d_out));
    return launch_ntt(&ctx->c, d_out, m0 << n, SALSAA_REP_FULL, 1);
}

int salsaa_parallel_dot_single_series_matrix(salsaa_ctx *ctx, const uint64_t *d_layer, size_t m,
                                             const uint64_t *d_W, size_t r, uint64_t *d_out) {
    CTX_D64(ctx, "salsaa_parallel_dot_single_series_matrix");
    SALSAA_CHECK(m >= 1 && r >= 1, "inner product needs m, r >= 1");
    // up to 16 rows: one pass over the layer and the rows in the NTT domain
    if (r <= 16) return launch_dot_series(&ctx->c, d_layer, m, d_W, r, d_out);
    // more rows: the commit with the layer as the only key row
    salsaa_key *key = nullptr;
    SALSAA_TRY(salsaa_key_create(ctx, d_layer, 1, m, &key));
    const int st = salsaa_parallel_dot_series_matrix(ctx, key, d_W, r, m, d_out);
    const int st2 = salsaa_key_destroy(key);
    return st != SALSAA_OK ? st : st2;
}

int salsaa_project(salsaa_ctx *ctx, const uint64_t *d_W, size_t r, size_t m,
                   const uint8_t *d_is_zero, const uint8_t *d_sign, uint64_t *d_out) {
    CTX_D64(ctx, "salsaa_project");
    return launch_project(&ctx->c, d_W, r, m, d_is_zero, d_sign, d_out);
}

static bool is_pow2(size_t x) { return x && !(x & (x - 1)); }

int salsaa_batch_projections(salsaa_ctx *ctx, const uint64_t *d_P, const uint64_t *d_W, size_t r,
                             size_t m, const uint8_t *d_is_zero, const uint8_t *d_sign,
                             const uint64_t *d_c0, uint64_t *d_projected_lhs,
                             uint64_t *d_witness_lhs, uint64_t *d_projection_rhs,
                             uint64_t *d_witness_rhs, int32_t *d_ok) {
    CTX_D64(ctx, "salsaa_batch_projections");
    SALSAA_CHECK(d_projection_rhs && d_witness_rhs, "null output pointer");
    const size_t cw = r * 256;
    // the reference's series have layers of lengths cw 2^t and 256 2^t only,
    // so its inner products find no layer for other shapes
    SALSAA_CHECK(is_pow2(r) && cw <= m && m % cw == 0 && is_pow2(m / cw) && cw < 65536,
                 "batch_projections needs r and m / (256 r) powers of two (project.rs:168-186)");
    Context *c = &ctx->c;
    // scratch: c0 [1], powers [256], c_l [cw], column sums of W [r][cw] and of
    // P [256], and the split sums' accumulator [max(r cw, 256)]
    uint64_t *buf = nullptr;
    const size_t nacc = std::max(r * cw, (size_t)256);
    const size_t nbuf = 1 + 256 + cw + r * cw + 256 + nacc;
    SALSAA_TRY(scratch_get(c, 9, nbuf * kD * sizeof(uint64_t), (void **)&buf));
    uint64_t *c0 = buf, *pw = c0 + kD, *cj = pw + 256 * kD, *wsum = cj + cw * kD,
             *psum = wsum + r * cw * kD;
    unsigned long long *acc = reinterpret_cast<unsigned long long *>(psum + 256 * kD);
    // consecutive powers of c0 (project.rs:118-129), slot by slot in NTT form
    SALSAA_HIP(hipMemcpyAsync(c0, d_c0, kD * sizeof(uint64_t), hipMemcpyDeviceToDevice, c->stream));
    SALSAA_TRY(launch_ntt(c, c0, 1, SALSAA_REP_FULL, 0));
    SALSAA_TRY(launch_powers(c, c0, pw));
    SALSAA_TRY(launch_ntt(c, pw, 256, SALSAA_REP_FULL, 1));
    // c_l = sum_k C[k][l] c0^k (project.rs:154-166)
    SALSAA_TRY(launch_project_transposed(c, pw, cw, d_is_zero, d_sign, cj));
    // Both layers repeat their base (all tensors are [1, 1]), so each inner
    // product is the base's against the sum of the matrix's chunks.
    SALSAA_TRY(launch_chunk_sums(c, d_W, r, m, cw, wsum, acc));
    SALSAA_TRY(launch_chunk_sums(c, d_P, 1, m, 256, psum, acc));
    SALSAA_TRY(salsaa_parallel_dot_single_series_matrix(ctx, cj, cw, wsum, r, d_witness_rhs));
    SALSAA_TRY(salsaa_parallel_dot_single_series_matrix(ctx, pw, 256, psum, 1, d_projection_rhs));
    if (d_witness_lhs) SALSAA_TRY(launch_repeat(c, cj, cw, m, d_witness_lhs));
    if (d_projected_lhs) SALSAA_TRY(launch_repeat(c, pw, 256, m, d_projected_lhs));
    if (d_ok) SALSAA_TRY(launch_sum_equals(c, d_witness_rhs, r, d_projection_rhs, d_ok));
    return SALSAA_OK;
}

int salsaa_send_cross_terms_before_join(salsaa_ctx *ctx, const uint64_t *d_projected_lhs,
                                        const uint64_t *d_witness_lhs, size_t m,
                                        const uint64_t *d_W, const uint64_t *d_P, uint64_t *d_l0,
                                        uint64_t *d_l1) {
    CTX_D64(ctx, "salsaa_send_cross_terms_before_join");
    SALSAA_CHECK(d_l0 && d_l1, "null output pointer");
    // the reference takes entry 0 of each inner product: row 0 of each matrix
    SALSAA_TRY(salsaa_parallel_dot_single_series_matrix(ctx, d_projected_lhs, m, d_W, 1, d_l0));
    return salsaa_parallel_dot_single_series_matrix(ctx, d_witness_lhs, m, d_P, 1, d_l1);
}

int salsaa_split(salsaa_ctx *ctx, const salsaa_key *key, const uint64_t *d_W, size_t r, size_t m,
                 uint64_t *d_out) {
    CTX_D64(ctx, "salsaa_split");
    SALSAA_CHECK(m >= 2 && m % 2 == 0, "split needs an even witness length");
    // row k < r is L_k = W[k][0 .. m/2), row r + k is R_k = W[k][m/2 .. m)
    return commit_rows(ctx, key, d_W, 2 * r, m / 2, r, m * kD, (m / 2) * kD, d_out);
}

int salsaa_norm_check(salsaa_ctx *ctx, const uint64_t *d_w, size_t rows, size_t L,
                      const uint64_t *d_cs, size_t reps, int fq2, uint64_t nr, uint64_t *d_claims,
                      uint64_t *d_msgs, uint64_t *d_finals, uint64_t *d_evals, uint64_t *d_rhs,
                      int32_t *d_ok) {
    CTX_OK(ctx);
    return launch_norm_check(&ctx->c, d_w, rows, L, d_cs, reps, fq2, nr, d_claims, d_msgs,
                             d_finals, d_evals, d_rhs, d_ok);
}

int salsaa_norm_check_noalloc(salsaa_ctx *ctx, const uint64_t *d_w, size_t rows, size_t L,
                              const uint64_t *d_cs, size_t reps, int fq2, uint64_t nr,
                              uint64_t *d_claims, uint64_t *d_msgs, uint64_t *d_finals,
                              uint64_t *d_evals, uint64_t *d_rhs, int32_t *d_ok) {
    CTX_OK(ctx);
    SALSAA_CHECK(ctx->c.d == kD || d_rhs == nullptr,
                 "norm_check_noalloc: at degree d > 64 only without rhs (d_rhs = NULL)");
    SALSAA_CHECK(rows > 0 && reps > 0 && rows < 65536 && reps < 65536,
                 "norm_check_noalloc: rows and reps must be in 1..65535");
    SALSAA_CHECK(L >= 2 && (L & (L - 1)) == 0 && __builtin_ctzll(L) <= 48,
                 "norm_check_noalloc: length must be a power of two >= 2");
    // the folded witnesses alone take reps * rows * L bytes
    if ((unsigned __int128)reps * rows * L > kNoAllocScratchBytes) {
        set_error("norm_check_noalloc: reps * rows * L = " +
                  std::to_string((double)reps * rows * L) +
                  " bytes of scratch exceed the fixed 64 MiB (use salsaa_norm_check)");
        return SALSAA_ERR_ARG;
    }
    size_t nm = 0, na = 0, nb = 0;
    nc_scratch_need(rows, L, reps, &nm, &na, &nb);
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    if (up(nm) + up(na) + up(nb) > kNoAllocScratchBytes) {
        set_error("norm_check_noalloc: needs " + std::to_string(up(nm) + up(na) + up(nb)) +
                  " bytes of scratch, more than the fixed 64 MiB (use salsaa_norm_check)");
        return SALSAA_ERR_ARG;
    }
    unsigned char *base = static_cast<unsigned char *>(noalloc_scratch(ctx->c.device));
    SALSAA_CHECK(base != nullptr, "norm_check_noalloc: scratch not set up on this device");
    NcScratch ext;
    ext.macc = reinterpret_cast<unsigned long long *>(base);
    ext.macc_bytes = nm;
    ext.bufA = reinterpret_cast<double *>(base + up(nm));
    ext.bufA_bytes = na;
    ext.bufB = reinterpret_cast<double *>(base + up(nm) + up(na));
    ext.bufB_bytes = nb;
    const NcOut o = nc_out_plain(rows, reps, (size_t)__builtin_ctzll(L), d_claims, d_msgs,
                                 d_finals, d_evals, d_rhs, d_ok);
    return launch_norm_check_ex(&ctx->c, d_w, nullptr, rows, L, d_cs, 0, reps, fq2, nr, o, 0,
                                &ext);
}

int salsaa_decompose(salsaa_ctx *ctx, const uint64_t *d_W, size_t r, size_t m, size_t parts,
                     uint64_t radix, uint64_t *d_out, uint64_t *d_radix) {
    CTX_OK(ctx);
    return launch_decompose(&ctx->c, d_W, r, m, parts, radix, d_out, d_radix);
}

int salsaa_decomp_ell(salsaa_ctx *ctx, const salsaa_key *key, const uint64_t *d_W, size_t r,
                      size_t m, size_t parts, uint64_t radix, uint64_t *d_new_W, uint64_t *d_rhs,
                      uint64_t *d_radix) {
    CTX_OK(ctx);
    SALSAA_CHECK(key != nullptr, "null key");
    SALSAA_CHECK(key->k.ctx == &ctx->c, "key belongs to another context");
    SALSAA_CHECK(m >= 1 && m <= key->k.m, "witness longer than the key");
    SALSAA_CHECK(r >= 1, "decomp_ell needs at least one witness row");
    if (ctx->c.d != kD) {
        // degree d > 64: the limbs, then their slot-pair commit
        SALSAA_TRY(launch_decompose(&ctx->c, d_W, r, m, parts, radix, d_new_W, d_radix));
        return commit_rows(ctx, key, d_new_W, r * parts, m, r * parts, m * (size_t)ctx->c.d, 0,
                           d_rhs);
    }
    SALSAA_TRY(launch_radix(&ctx->c, d_W, r, m, parts, radix, d_radix));
    // one pass over W writes the limb witness and the limbs' slot-major NTT
    const size_t rows = r * parts;
    if (key->k.d_AB || key->k.d_mhat) {
        const size_t ps = plane_ktiles(m);
        int8_t *wb = nullptr;
        SALSAA_TRY(scratch_get(&ctx->c, 6, plane_bytes(rows, m), (void **)&wb));
        SALSAA_TRY(launch_decomp_to_slotmajor(&ctx->c, d_W, r, m, parts, d_radix, d_new_W, wb, 1,
                                              ps));
        if (key->k.d_mhat)
            return launch_commit_crs(&ctx->c, key->k.d_mhat, key->k.kappa, wb, rows, m, ps, d_rhs);
        return launch_commit_mfma(&ctx->c, key->k.d_AB, key->k.kappa, key->k.a_kt,
                                  key->k.a_ktiles, wb, rows, m, ps, d_rhs);
    }
    double *w_sm = nullptr;
    SALSAA_TRY(scratch_get(&ctx->c, 6, (size_t)kD * rows * m * sizeof(double), (void **)&w_sm));
    SALSAA_TRY(launch_decomp_to_slotmajor(&ctx->c, d_W, r, m, parts, d_radix, d_new_W, w_sm, 0, 0));
    return launch_commit_mac(&ctx->c, key->k.d_A, key->k.kappa, key->k.a_rows, key->k.a_m, w_sm,
                             rows, m, d_rhs);
}

int salsaa_compose(salsaa_ctx *ctx, const uint64_t *d_in, size_t rows, size_t n, size_t parts,
                   const uint64_t *d_radix, uint64_t *d_out) {
    CTX_OK(ctx);
    return launch_compose(&ctx->c, d_in, rows, n, parts, d_radix, d_out);
}

int salsaa_verify_decomp(salsaa_ctx *ctx, const uint64_t *d_rhs_dec, const uint64_t *d_rhs,
                         size_t kappa, size_t r, size_t parts, const uint64_t *d_radix,
                         int rows_mode, uint32_t *d_mismatch) {
    CTX_OK(ctx);
    return launch_compose_check(&ctx->c, d_rhs_dec, kappa, r, parts, d_radix, d_rhs, rows_mode,
                                d_mismatch);
}

int salsaa_parallel_dot_matrix_matrix(salsaa_ctx *ctx, const uint64_t *d_W, size_t r_in,
                                      size_t m, const uint64_t *d_C, size_t r_out,
                                      uint64_t *d_out) {
    CTX_OK(ctx);
    return launch_fold(&ctx->c, d_W, r_in, m, d_C, r_out, d_out);
}

}  // extern "C"
