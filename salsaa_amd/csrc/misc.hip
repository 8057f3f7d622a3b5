// misc.hip -- seeded samplers, decomposition (decomp_ell), recomposition
// checks (verify_decomp), batched ring products and the fold kernel.
#include "ntt_device.hpp"

namespace salsaa {

namespace {

__global__ void k_sample_uniform(uint64_t *__restrict__ out, size_t n, uint64_t key, uint64_t q) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = mulhi64(draw(key, i), q);
}

__global__ void k_sample_bounded(uint64_t *__restrict__ out, size_t n, uint64_t key,
                                 uint64_t bound, uint64_t q) {
    const uint64_t span = 2 * bound + 1;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        uint64_t t = mulhi64(draw(key, i), span);
        out[i] = t >= bound ? t - bound : q - (bound - t);
    }
}

// cs[rep][round][2]: coordinate k of round t from stream 0x5C000 + rep, index 2t + k
__global__ void k_challenges(uint64_t *__restrict__ cs, size_t reps, size_t nrounds, int fq2,
                             uint64_t seed, uint64_t q) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= reps * nrounds * 2) return;
    const size_t rep = i / (nrounds * 2), rem = i % (nrounds * 2);
    const size_t t = rem / 2, k = rem % 2;
    uint64_t v = 0;
    if (k == 0 || fq2) v = mulhi64(draw(stream_key(seed, 0x5C000ULL + rep), 2 * t + k), q);
    cs[i] = v;
}

// ---- decomposition (arithmetic.rs:1461-1593)
// OR of |x| (pos / neg split) over W; one atomic per block (a grid of a few
// hundred blocks, so the single accumulator sees little contention)
// instance blockIdx.y: W + y n, OR into acc[y] (of the magnitudes min(x, q - x)).
// Each thread keeps four 16-byte loads in flight per iteration.
__device__ __forceinline__ uint64_t mag(uint64_t x, uint64_t q) { return (2 * x > q) ? q - x : x; }
__global__ __launch_bounds__(256) void k_bitor(const uint64_t *__restrict__ W, size_t n,
                                               uint64_t q, unsigned long long *__restrict__ acc) {
    __shared__ uint64_t part[4];
    uint64_t a = 0;
    W += (size_t)blockIdx.y * n;
    acc += blockIdx.y;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    // whole 16-byte pairs when the instance starts 16-byte aligned
    const bool al = ((reinterpret_cast<uintptr_t>(W) & 15) == 0);
    size_t done = 0;
    if (al) {
        const ulonglong2 *W2 = reinterpret_cast<const ulonglong2 *>(W);
        const size_t n2 = n / 2;
        size_t i = t0;
        for (; i + 3 * stride < n2; i += 4 * stride) {
            const ulonglong2 v0 = W2[i], v1 = W2[i + stride], v2 = W2[i + 2 * stride],
                             v3 = W2[i + 3 * stride];
            a |= mag(v0.x, q) | mag(v0.y, q) | mag(v1.x, q) | mag(v1.y, q) | mag(v2.x, q) |
                 mag(v2.y, q) | mag(v3.x, q) | mag(v3.y, q);
        }
        for (; i < n2; i += stride) {
            const ulonglong2 v = W2[i];
            a |= mag(v.x, q) | mag(v.y, q);
        }
        done = 2 * n2;
    }
    for (size_t i = done + t0; i < n; i += stride) a |= mag(W[i], q);
    for (int off = 32; off >= 1; off >>= 1) a |= __shfl_xor(a, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t b = part[0] | part[1] | part[2] | part[3];
        if (b) atomicOr(acc, (unsigned long long)b);
    }
}

__global__ void k_zero_n(unsigned long long *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ULL;
}

// gadget weight 2^(lg * t) mod q; 0 once the shift leaves 64 bits, as the
// reference's wrapping 2u64.pow does (arithmetic.rs:1739)
__device__ __forceinline__ double gadget_weight(unsigned lg, size_t t, uint64_t q) {
    const size_t s = (size_t)lg * t;
    return s >= 64 ? 0.0 : u2d((1ULL << s) % q);
}

// radix = 2^ceil(bw / parts) from the OR accumulator (or the explicit radix),
// one per instance
__global__ void k_set_radix(const unsigned long long *__restrict__ acc, size_t ninst,
                            uint64_t explicit_radix, size_t parts, uint64_t *__restrict__ d_radix) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    if (explicit_radix) {
        d_radix[i] = explicit_radix;
        return;
    }
    const uint64_t a = acc[i];
    const unsigned bw = a ? 64u - (unsigned)__clzll((long long)a) : 0u;
    const unsigned lg = (unsigned)((bw + parts - 1) / parts);
    d_radix[i] = 1ULL << lg;
}

__global__ void k_decompose(const uint64_t *__restrict__ W, size_t r, size_t md, size_t parts,
                            const uint64_t *__restrict__ d_radix, uint64_t q,
                            uint64_t *__restrict__ out) {
    const uint64_t radix = *d_radix;
    const unsigned lg = (unsigned)__ffsll((long long)radix) - 1u;
    const uint64_t mask = radix - 1;
    const size_t n = r * md;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (size_t)gridDim.x * blockDim.x) {
        const uint64_t x = W[idx];
        const bool neg = 2 * x > q;
        const uint64_t v = neg ? q - x : x;
        const size_t rho = idx / md, rem = idx % md;
        uint64_t *o = out + rho * parts * md + rem;
        for (size_t t = 0; t < parts; t++) {
            const unsigned s = (unsigned)(lg * t);
            const uint64_t dg = (lg == 0) ? (v & mask) : (s >= 64 ? 0 : (v >> s) & mask);
            o[t * md] = neg ? (dg ? q - dg : 0) : dg;
        }
    }
}

// compose_with_radix_mod: in [rows][n*parts][d] -> out [rows][n][d], d = 2^logd
__global__ void k_compose(const uint64_t *__restrict__ in, size_t rows, size_t n, size_t parts,
                          const uint64_t *__restrict__ d_radix, uint64_t *__restrict__ out,
                          uint64_t q, Mod M, int logd) {
    const uint64_t radix = *d_radix;
    const unsigned lg = (unsigned)__ffsll((long long)radix) - 1u;
    const size_t total = rows * n << logd;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const size_t c = idx & ((1u << logd) - 1), ri = idx >> logd;
        const size_t row = ri / n, i = ri % n;
        double acc = 0.0;
        for (size_t t = 0; t < parts; t++) {
            const double x = u2d(in[(((row * n * parts) + i * parts + t) << logd) + c]);
            acc = addmod(acc, mulmod(x, gadget_weight(lg, t, q), M), M);
        }
        out[idx] = d2u(acc);
    }
}

// rows_mode 0: compare compose(in [rows][n*parts]) (columns grouped) with
// expect [rows][n]; rows_mode 1: in = limbs [n_rows*parts][n][d] grouped by
// rows, expect [n_rows][n][d].
__global__ void k_compose_check(const uint64_t *__restrict__ in, size_t rows, size_t n,
                                size_t parts, const uint64_t *__restrict__ d_radix,
                                const uint64_t *__restrict__ expect, int rows_mode,
                                uint32_t *__restrict__ mismatch, uint64_t q, Mod M, int logd) {
    const uint64_t radix = *d_radix;
    const unsigned lg = (unsigned)__ffsll((long long)radix) - 1u;
    const size_t total = rows * n << logd;
    uint32_t bad = 0;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const size_t c = idx & ((1u << logd) - 1), ri = idx >> logd;
        const size_t row = ri / n, i = ri % n;
        double acc = 0.0;
        for (size_t t = 0; t < parts; t++) {
            const size_t src = rows_mode ? ((((row * parts + t) * n + i) << logd) + c)
                                         : ((((row * n * parts) + i * parts + t) << logd) + c);
            acc = addmod(acc, mulmod(u2d(in[src]), gadget_weight(lg, t, q), M), M);
        }
        if (d2u(acc) != expect[idx]) bad++;
    }
    if (bad) atomicAdd(mismatch, bad);
}

// ---- IncompleteNTT slot products
// out[e] = a[e] * b[e], all IncompleteNTT with h = 2^logh slots (d = 2h),
// slot moduli X^2 - omega[j]; thread per (element, slot)
__global__ void k_slot_mul(const uint64_t *__restrict__ a, const uint64_t *__restrict__ b,
                           uint64_t *__restrict__ out, size_t count, int logh,
                           const double *__restrict__ omega, Mod M) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t h = (size_t)1 << logh;
    if (gid >= count * h) return;
    const size_t e = gid >> logh;
    const size_t j = gid & (h - 1);
    const uint64_t *ae = a + 2 * h * e, *be = b + 2 * h * e;
    const double a0 = u2d(ae[j]), a1 = u2d(ae[h + j]);
    const double b0 = u2d(be[j]), b1 = u2d(be[h + j]);
    const double om = omega[j];
    const double re = reduce(mul_lazy(a0, b0, M) + mul_lazy(mulmod(a1, b1, M), om, M), M);
    const double ro = reduce(mul_lazy(a1, b0, M) + mul_lazy(a0, b1, M), M);
    out[2 * h * e + j] = d2u(re);
    out[2 * h * e + h + j] = d2u(ro);
}

// fold MAC in the IncompleteNTT domain (h = 2^logh slots, d = 2h):
// out[j][i] = sum_k Wn[k][i] * Cn[j][k]; thread per (j, i, slot)
__global__ void k_fold_mac(const uint64_t *__restrict__ Wn, size_t r_in, size_t m,
                           const uint64_t *__restrict__ Cn, size_t r_out,
                           uint64_t *__restrict__ out, int logh, const double *__restrict__ omega,
                           Mod M) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t h = (size_t)1 << logh, d = 2 * h;
    if (gid >= r_out * m * h) return;
    const size_t s = gid & (h - 1);
    const size_t ji = gid >> logh;
    const size_t j = ji / m, i = ji % m;
    const double om = omega[s];
    double re = 0.0, ro = 0.0;
    for (size_t k = 0; k < r_in; k++) {
        const uint64_t *we = Wn + (k * m + i) * d;
        const uint64_t *ce = Cn + (j * r_in + k) * d;
        const double w0 = u2d(we[s]), w1 = u2d(we[h + s]);
        const double c0 = u2d(ce[s]), c1 = u2d(ce[h + s]);
        re += mul_lazy(w0, c0, M) + mul_lazy(mulmod(w1, c1, M), om, M);
        ro += mul_lazy(w1, c0, M) + mul_lazy(w0, c1, M);
        re = reduce_lazy(re, M);
        ro = reduce_lazy(ro, M);
    }
    uint64_t *o = out + (j * m + i) * d;
    o[s] = d2u(reduce(re, M));
    o[h + s] = d2u(reduce(ro, M));
}

unsigned grid_for(size_t n, unsigned block) {
    size_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    if (g == 0) g = 1;
    return (unsigned)g;
}

}  // namespace

int launch_sample_uniform(Context *ctx, uint64_t *out, size_t n, uint64_t seed, uint64_t stream) {
    if (!n) return SALSAA_OK;
    hipLaunchKernelGGL(k_sample_uniform, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, out, n,
                       stream_key(seed, stream), ctx->q);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_sample_bounded(Context *ctx, uint64_t *out, size_t n, uint64_t bound, uint64_t seed,
                          uint64_t stream) {
    if (!n) return SALSAA_OK;
    SALSAA_CHECK(bound < ctx->q / 2, "bound must be below q/2");
    hipLaunchKernelGGL(k_sample_bounded, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, out, n,
                       stream_key(seed, stream), bound, ctx->q);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_challenges(Context *ctx, uint64_t *cs, size_t reps, size_t nrounds, int fq2,
                      uint64_t seed) {
    const size_t n = reps * nrounds * 2;
    if (!n) return SALSAA_OK;
    hipLaunchKernelGGL(k_challenges, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, cs, reps,
                       nrounds, fq2, seed, ctx->q);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_radix_n(Context *ctx, const uint64_t *W, size_t ninst, size_t r, size_t m,
                   size_t parts, uint64_t radix, uint64_t *d_radix) {
    SALSAA_CHECK(parts >= 1 && parts <= 64, "parts must be in [1, 64]");
    SALSAA_CHECK(radix == 0 || (radix & (radix - 1)) == 0, "radix must be a power of two");
    SALSAA_CHECK(ninst >= 1 && ninst < 65536, "instances out of range");
    const size_t n = r * m * (size_t)ctx->d;  // per instance
    unsigned long long *acc = nullptr;
    SALSAA_TRY(scratch_get(ctx, 3, ninst * sizeof(unsigned long long), (void **)&acc));
    if (radix == 0) {
        // the OR accumulators start at zero (kernel writes, safe under capture)
        hipLaunchKernelGGL(k_zero_n, dim3(ceil_div(ninst, 256)), dim3(256), 0, ctx->stream, acc,
                           ninst);
        // about four blocks per CU over all instances, each taking at least
        // 4096 residues
        const unsigned g = (unsigned)std::min<size_t>(
            std::max<size_t>(1, (size_t)ctx->ncu * 4 / ninst), std::max<size_t>(1, (n + 4095) / 4096));
        hipLaunchKernelGGL(k_bitor, dim3(g, (unsigned)ninst), dim3(256), 0, ctx->stream, W, n,
                           ctx->q, acc);
    }
    hipLaunchKernelGGL(k_set_radix, dim3(ceil_div(ninst, 256)), dim3(256), 0, ctx->stream, acc,
                       ninst, radix, parts, d_radix);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_radix(Context *ctx, const uint64_t *W, size_t r, size_t m, size_t parts,
                 uint64_t radix, uint64_t *d_radix) {
    return launch_radix_n(ctx, W, 1, r, m, parts, radix, d_radix);
}

int launch_decompose(Context *ctx, const uint64_t *W, size_t r, size_t m, size_t parts,
                     uint64_t radix, uint64_t *out, uint64_t *d_radix) {
    SALSAA_TRY(launch_radix(ctx, W, r, m, parts, radix, d_radix));
    const size_t md = m * (size_t)ctx->d, n = r * md;
    hipLaunchKernelGGL(k_decompose, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, W, r, md,
                       parts, d_radix, ctx->q, out);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_compose(Context *ctx, const uint64_t *in, size_t rows, size_t n, size_t parts,
                   const uint64_t *d_radix, uint64_t *out) {
    Mod M{ctx->qd, ctx->qinv};
    const size_t total = rows * n * (size_t)ctx->d;
    if (!total) return SALSAA_OK;
    hipLaunchKernelGGL(k_compose, dim3(grid_for(total, 256)), dim3(256), 0, ctx->stream, in, rows, n,
                       parts, d_radix, out, ctx->q, M, log2_deg(ctx));
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_compose_check(Context *ctx, const uint64_t *in, size_t rows, size_t n, size_t parts,
                         const uint64_t *d_radix, const uint64_t *expect, int rows_mode,
                         uint32_t *d_mismatch) {
    Mod M{ctx->qd, ctx->qinv};
    const size_t total = rows * n * (size_t)ctx->d;
    if (!total) return SALSAA_OK;
    hipLaunchKernelGGL(k_compose_check, dim3(grid_for(total, 256)), dim3(256), 0, ctx->stream, in,
                       rows, n, parts, d_radix, expect, rows_mode, d_mismatch, ctx->q, M,
                       log2_deg(ctx));
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

// the IncompleteNTT of count Coefficient elements src into dst
static int ntt_copy(Context *ctx, const uint64_t *src, uint64_t *dst, size_t count) {
    if (ctx->d != kD) return launch_wide_ntt(ctx, src, dst, count, SALSAA_REP_INCOMPLETE, 0);
    SALSAA_HIP(hipMemcpyAsync(dst, src, count * kD * 8, hipMemcpyDeviceToDevice, ctx->stream));
    return launch_ntt(ctx, dst, count, SALSAA_REP_INCOMPLETE, 0);
}

int launch_ring_mul(Context *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out,
                    size_t count) {
    if (!count) return SALSAA_OK;
    Mod M{ctx->qd, ctx->qinv};
    const size_t d = (size_t)ctx->d;
    uint64_t *ta = nullptr;
    SALSAA_TRY(scratch_get(ctx, 4, 2 * count * d * sizeof(uint64_t), (void **)&ta));
    uint64_t *tb = ta + count * d;
    SALSAA_TRY(ntt_copy(ctx, a, ta, count));
    SALSAA_TRY(ntt_copy(ctx, b, tb, count));
    hipLaunchKernelGGL(k_slot_mul, dim3(ceil_div(count * d / 2, 256)), dim3(256), 0, ctx->stream,
                       ta, tb, out, count, log2_deg(ctx) - 1, slot_omega(ctx), M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_fold(Context *ctx, const uint64_t *W, size_t r_in, size_t m, const uint64_t *C,
                size_t r_out, uint64_t *out) {
    if (!r_out || !m) return SALSAA_OK;
    SALSAA_CHECK(r_in >= 1, "fold needs r_in >= 1");
    Mod M{ctx->qd, ctx->qinv};
    const size_t d = (size_t)ctx->d;
    uint64_t *wn = nullptr;
    const size_t nw = r_in * m, nc = r_out * r_in;
    SALSAA_TRY(scratch_get(ctx, 5, (nw + nc) * d * sizeof(uint64_t), (void **)&wn));
    uint64_t *cn = wn + nw * d;
    SALSAA_TRY(ntt_copy(ctx, W, wn, nw));
    SALSAA_TRY(ntt_copy(ctx, C, cn, nc));
    const size_t threads = r_out * m * d / 2;
    hipLaunchKernelGGL(k_fold_mac, dim3(ceil_div(threads, 256)), dim3(256), 0, ctx->stream, wn, r_in,
                       m, cn, r_out, out, log2_deg(ctx) - 1, slot_omega(ctx), M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

// ---------------------------------------------------------------- power series
// The largest expanded layer of complete_power_series (arithmetic.rs:158-176)
// in the NTT domain, where every ring product is pointwise:
//   out[((i_0 2 + i_1) 2 + ... + i_{n-1}) m0 + j][s] = T_0[i_0][s] ... T_{n-1}[i_{n-1}][s] B[j][s]
// (kronecker_product, arithmetic.rs:845-864, applied last tensor first, so
// the first tensor is the most significant index bit).  One thread per output
// residue; base [m0][d], tensors [n][2][d], out [m0 << n][d], all NTT form.
namespace {
__global__ void k_series_expand(const uint64_t *__restrict__ base, size_t m0,
                                const uint64_t *__restrict__ tens, int n,
                                uint64_t *__restrict__ out, size_t total, Mod M) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const size_t idx = e / kD, s = e % kD;
        const size_t bits = idx / m0, j = idx % m0;
        double v = u2d(base[j * kD + s]);
        for (int k = 0; k < n; k++) {
            const size_t bit = (bits >> (n - 1 - k)) & 1;
            v = mulmod(v, u2d(tens[((size_t)k * 2 + bit) * kD + s]), M);
        }
        out[e] = d2u(v);
    }
}
}  // namespace

int launch_series_expand(Context *ctx, const uint64_t *base, size_t m0, const uint64_t *tens,
                         size_t n, uint64_t *out) {
    Mod M{ctx->qd, ctx->qinv};
    const size_t total = (m0 << n) * kD;
    const unsigned g = (unsigned)std::min<size_t>(ceil_div(total, (size_t)256), (size_t)ctx->ncu * 32);
    hipLaunchKernelGGL(k_series_expand, dim3(g), dim3(256), 0, ctx->stream, base, m0, tens, (int)n,
                       out, total, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

}  // namespace salsaa
