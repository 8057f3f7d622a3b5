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
    const uint64_t *src = W + (i * m + j * cw) * kD;
    uint64_t pos[8], neg[8];
#pragma unroll
    for (int kk = 0; kk < 8; kk++) pos[kk] = neg[kk] = 0;
    for (size_t l0 = 0; l0 < cw; l0 += PJ_LT) {
        // the tile: 64 elements (4096 u64, 16 per thread, coalesced), and
        // thread t's 8 signs of output k0 + t / 8, elements 8 (t % 8) ..
        const ulonglong2 *s2 = reinterpret_cast<const ulonglong2 *>(src + l0 * kD);
#pragma unroll
        for (int u = 0; u < 8; u++)
            reinterpret_cast<ulonglong2 *>(wt)[tid + PJ_THREADS * u] = s2[tid + PJ_THREADS * u];
        {
            const int k = tid >> 3, lb = (tid & 7) * 8;
            const size_t off = (size_t)(k0 + k) * cw + l0 + lb;
#pragma unroll
            for (int e = 0; e < 8; e++)
                ct[lb + e][k] = zero[off + e] ? 0 : (sign[off + e] ? 1 : -1);
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
        out[(idx * PJ_H + k0 + 8 * wave + kk) * kD + lane] = a >= b ? a - b : a + q - b;
    }
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
    hipLaunchKernelGGL(k_project, dim3((unsigned)(r * nch), PJ_H / PJ_KB), dim3(PJ_THREADS), 0,
                       ctx->stream, W, m, cw, nch, is_zero, sign, out, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

}  // namespace salsaa
