// project.hip -- the projection of a witness (project, project.rs:52-98,
// before its commit): with cw = r * 256, chunk j < m / cw of row i gives 256
// ring elements
//     out[(i (m / cw) + j) 256 + k] = sum_l C[k][l] W[i][j cw + l]  mod q,
// C[k][l] = 0 where is_zero[k][l], else +1 / -1 by sign[k][l].  As in the
// reference the positive and negative parts are summed in u64 (exact while
// cw q < 2^64) and reduced mod q once.
//
// A block takes one chunk and 32 of its 256 outputs; its four waves own 8
// outputs each, lane = coefficient.  The chunk streams through LDS in tiles of
// 64 ring elements (32 KiB), with the matching 64 x 32 challenge signs stored
// transposed, so a wave reads the 8 signs of one element with one broadcast
// 8-byte read.  Integer work, bound by the sign selects (VALU).
#include "common.hpp"

namespace salsaa {

namespace {

constexpr int PJ_H = 256;       // projections per chunk (HEIGHT, project.rs:17)
constexpr int PJ_KB = 32;       // outputs per block
constexpr int PJ_LT = 64;       // ring elements per LDS tile
constexpr int PJ_THREADS = 256;

// TRANS = false: the projection, block (chunk, 32 of its 256 outputs), summed
// index l < cw, C[k][l] at k cw + l.  TRANS = true: the challenge row of
// batch_projections, c_l = sum_k C[k][l] P[k] over the 256 elements P, one
// chunk, blocks of 32 of the cw outputs l, summed index k < 256, C[k][l] at
// the same place (so the sign tile is read along its rows).
template <bool TRANS>
__global__ __launch_bounds__(PJ_THREADS) void k_project(const uint64_t *__restrict__ W, size_t m,
                                                        size_t cw, size_t nch,
                                                        const uint8_t *__restrict__ zero,
                                                        const uint8_t *__restrict__ sign,
                                                        uint64_t *__restrict__ out, Mod M) {
    __shared__ __align__(16) uint64_t wt[PJ_LT * kD];
    __shared__ __align__(8) int8_t ct[PJ_LT][PJ_KB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t idx = blockIdx.x, i = idx / nch, j = idx % nch;
    const int k0 = blockIdx.y * PJ_KB;
    const size_t nsum = TRANS ? (size_t)PJ_H : cw, nout = TRANS ? cw : (size_t)PJ_H;
    const uint64_t *src = TRANS ? W : W + (i * m + j * cw) * kD;
    uint64_t pos[8], neg[8];
#pragma unroll
    for (int kk = 0; kk < 8; kk++) pos[kk] = neg[kk] = 0;
    for (size_t l0 = 0; l0 < nsum; l0 += PJ_LT) {
        // the tile: 64 elements (4096 u64, 16 per thread, coalesced), and
        // 8 signs per thread: (output k0 + t / 8, elements 8 (t % 8) ..) or,
        // transposed, (element t / 4, outputs k0 + 8 (t % 4) ..)
        const ulonglong2 *s2 = reinterpret_cast<const ulonglong2 *>(src + l0 * kD);
#pragma unroll
        for (int u = 0; u < 8; u++)
            reinterpret_cast<ulonglong2 *>(wt)[tid + PJ_THREADS * u] = s2[tid + PJ_THREADS * u];
        if (!TRANS) {
            const int k = tid >> 3, lb = (tid & 7) * 8;
            const size_t off = (size_t)(k0 + k) * cw + l0 + lb;
#pragma unroll
            for (int e = 0; e < 8; e++)
                ct[lb + e][k] = zero[off + e] ? 0 : (sign[off + e] ? 1 : -1);
        } else {
            const int l = tid >> 2, ob = (tid & 3) * 8;
            const size_t off = (l0 + l) * cw + k0 + ob;
#pragma unroll
            for (int e = 0; e < 8; e++)
                ct[l][ob + e] = zero[off + e] ? 0 : (sign[off + e] ? 1 : -1);
        }
        __syncthreads();
        for (int l = 0; l < PJ_LT; l++) {
            const uint64_t w = wt[l * kD + lane];
            const uint64_t c8 = *reinterpret_cast<const uint64_t *>(&ct[l][8 * wave]);
#pragma unroll
            for (int kk = 0; kk < 8; kk++) {
                const int c = (int)(int8_t)(c8 >> (8 * kk));
                pos[kk] += c > 0 ? w : 0;
                neg[kk] += c < 0 ? w : 0;
            }
        }
        __syncthreads();
    }
    const uint64_t q = (uint64_t)M.q;
#pragma unroll
    for (int kk = 0; kk < 8; kk++) {
        const uint64_t a = pos[kk] % q, b = neg[kk] % q;
        out[(idx * nout + k0 + 8 * wave + kk) * kD + lane] = a >= b ? a - b : a + q - b;
    }
}

// out[i][l] = sum_j src[i][j cw + l] mod q (j < nch) for rows i < rows of
// length m, per coefficient (u64 sums, reduced every 8192 terms).
__global__ void k_chunk_sums(const uint64_t *__restrict__ src, size_t rows, size_t m, size_t cw,
                             size_t nch, uint64_t *__restrict__ out, uint64_t q) {
    const size_t total = rows * cw * kD;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const size_t i = e / (cw * kD), lc = e % (cw * kD);
        const uint64_t *p = src + i * m * kD + lc;
        uint64_t acc = 0;
        for (size_t j = 0; j < nch; j++) {
            acc += p[j * cw * kD];
            if ((j & 8191) == 8191) acc %= q;  // below (8192 + 1) q < 2^64
        }
        out[e] = acc % q;
    }
}

// out[k][s] = c[s]^k for k < 256 and the 64 slots s of one element in NTT form
// (consecutive_powers, project.rs:118-129, slot by slot).
__global__ void k_powers(const uint64_t *__restrict__ c, uint64_t *__restrict__ out, Mod M) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= PJ_H * kD) return;
    const int k = e / kD, s = e % kD;
    const double b = u2d(c[s]);
    double v = 1.0;
    for (int bit = 7; bit >= 0; bit--) {
        v = mulmod(v, v, M);
        if ((k >> bit) & 1) v = mulmod(v, b, M);
    }
    out[e] = d2u(v);
}

// out[x] = base[x mod m0] for x < n (the expanded layer of a power series
// whose tensors are all [1, 1]).
__global__ void k_repeat(const uint64_t *__restrict__ base, size_t m0, size_t n,
                         uint64_t *__restrict__ out) {
    const size_t total = n * kD;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x)
        out[e] = base[((e / kD) % m0) * kD + e % kD];
}

// ok[0] = 1 when sum_i a[i] == b (r elements against one, canonical), else 0:
// the assertion of batch_projections (project.rs:194-198).  One wave.
__global__ void k_sum_equals(const uint64_t *__restrict__ a, size_t r,
                             const uint64_t *__restrict__ b, int32_t *__restrict__ ok, Mod M) {
    const int lane = threadIdx.x;
    double acc = 0.0;
    for (size_t i = 0; i < r; i++) acc = addmod(acc, u2d(a[i * kD + lane]), M);
    const bool eq = d2u(acc) == b[lane];
    const bool all = __all(eq);
    if (lane == 0) ok[0] = all ? 1 : 0;
}

}  // namespace

int launch_project(Context *ctx, const uint64_t *W, size_t r, size_t m, const uint8_t *is_zero,
                   const uint8_t *sign, uint64_t *out) {
    const size_t cw = r * PJ_H;
    SALSAA_CHECK(r >= 1 && cw <= m, "projection needs r * 256 <= m (project.rs:61-64)");
    // the u64 part sums are exact while cw (q - 1) < 2^64 (the reference's own bound)
    SALSAA_CHECK(cw <= ~0ULL / (ctx->q - 1), "projection chunk too wide for exact u64 sums");
    SALSAA_CHECK(cw < 65536, "projection chunk too wide");
    const size_t nch = m / cw;
    SALSAA_CHECK(r * nch < (1u << 31), "projection too large");
    Mod M{ctx->qd, ctx->qinv};
    hipLaunchKernelGGL(k_project<false>, dim3((unsigned)(r * nch), PJ_H / PJ_KB), dim3(PJ_THREADS),
                       0, ctx->stream, W, m, cw, nch, is_zero, sign, out, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_project_transposed(Context *ctx, const uint64_t *P, size_t cw, const uint8_t *is_zero,
                              const uint8_t *sign, uint64_t *out) {
    SALSAA_CHECK(cw >= PJ_H && cw % PJ_H == 0 && cw < 65536, "challenge width out of range");
    Mod M{ctx->qd, ctx->qinv};
    hipLaunchKernelGGL(k_project<true>, dim3(1, (unsigned)(cw / PJ_KB)), dim3(PJ_THREADS), 0,
                       ctx->stream, P, (size_t)PJ_H, cw, (size_t)1, is_zero, sign, out, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_chunk_sums(Context *ctx, const uint64_t *src, size_t rows, size_t m, size_t cw,
                      uint64_t *out) {
    SALSAA_CHECK(cw >= 1 && m % cw == 0, "chunk sums need cw | m");
    const size_t total = rows * cw * kD;
    const unsigned g = (unsigned)std::min<size_t>(ceil_div(total, (size_t)256), (size_t)ctx->ncu * 16);
    hipLaunchKernelGGL(k_chunk_sums, dim3(g), dim3(256), 0, ctx->stream, src, rows, m, cw, m / cw,
                       out, ctx->q);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_powers(Context *ctx, const uint64_t *c_ntt, uint64_t *out) {
    Mod M{ctx->qd, ctx->qinv};
    hipLaunchKernelGGL(k_powers, dim3(PJ_H * kD / 256), dim3(256), 0, ctx->stream, c_ntt, out, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_repeat(Context *ctx, const uint64_t *base, size_t m0, size_t n, uint64_t *out) {
    const size_t total = n * kD;
    const unsigned g = (unsigned)std::min<size_t>(ceil_div(total, (size_t)256), (size_t)ctx->ncu * 32);
    hipLaunchKernelGGL(k_repeat, dim3(g), dim3(256), 0, ctx->stream, base, m0, n, out);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_sum_equals(Context *ctx, const uint64_t *a, size_t r, const uint64_t *b, int32_t *ok) {
    Mod M{ctx->qd, ctx->qinv};
    hipLaunchKernelGGL(k_sum_equals, dim3(1), dim3(kD), 0, ctx->stream, a, r, b, ok, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

}  // namespace salsaa
