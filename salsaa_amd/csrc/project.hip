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

// ---- the projection on the int8 matrix cores
//
// Per chunk the projection is one integer matrix product: C (256 x cw, entries
// -1, 0, 1) times the chunk's coefficients (cw x 64).  A coefficient x, shifted
// by h = (q - 1) / 2 (|x - h| <= h < 2^49), is written as 7 signed base-256
// digits: with u = x - h + 0x80..80 (seven 0x80 bytes, so 0 <= u < 2^56), byte
// t of u is d_t + 128, i.e. byte t of u ^ 0x80..80 is d_t as an int8.  Then
//     out[k][c] = sum_t 2^(8t) (C . D_t)[k][c] + h sum_l C[k][l]  mod q,
// every C . D_t is exact in int32 (|sum| <= 128 cw < 2^23), and the last term
// is one constant per output (k_proj_bias).  The digit
// matrices are the B operands of v_mfma_i32_16x16x64_i8: tile t holds digit t
// of 16 coefficients, so the 7 tiles of one output sit in the same lane and
// register of 7 accumulators and are recombined without a transpose.
//
// Block: one chunk, 16 coefficients (blockIdx.y), all 256 outputs; wave w
// owns outputs 64 w .. 64 w + 63 (4 row tiles) x 7 digit tiles = 28
// accumulators.  The chunk goes through in stages of 256 elements (4 K steps
// of 64): thread (ks, g, c) loads coefficient c of elements 16 g .. 16 g + 15
// of K step ks, turns them into digits and writes the 7 digit tiles' LDS slot
// 16 g + c (MFMA lane order: column c, K 16 g ..), double buffered, one block
// barrier per stage.  The signs come pre-tiled (k_proj_cfrag) from L2.
constexpr int PM_THREADS = 256;
constexpr int PM_STAGE = 4 * 7 * 1024;  // LDS bytes of one stage's digit tiles

typedef int v4i_pm __attribute__((ext_vector_type(4)));

// Cf[((ks 16 + rt) 64 + L) 16 + b] = C[16 rt + L % 16][64 ks + 16 (L / 16) + b]
// as an int8 (the A fragments of K step ks, row tile rt, in lane order).
__global__ void k_proj_cfrag(const uint8_t *__restrict__ zero, const uint8_t *__restrict__ sign,
                             size_t cw, uint8_t *__restrict__ Cf) {
    const size_t nslot = (size_t)PJ_H * cw / 16;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nslot;
         e += (size_t)gridDim.x * blockDim.x) {
        const size_t ks = e / 1024;
        const int rt = (int)((e / 64) % 16), L = (int)(e % 64);
        const size_t off = (size_t)(16 * rt + L % 16) * cw + 64 * ks + 16 * (L / 16);
        uint32_t w[4];
#pragma unroll
        for (int h = 0; h < 4; h++) {
            uint32_t v = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const uint8_t z = zero[off + 4 * h + b], sg = sign[off + 4 * h + b];
                v |= (uint32_t)(z ? 0u : (sg ? 1u : 0xFFu)) << (8 * b);
            }
            w[h] = v;
        }
        *reinterpret_cast<uint4 *>(Cf + e * 16) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// bias[k] = h sum_l C[k][l] mod q with h = (q - 1) / 2: the projection of the
// constant h, added back to the projection of the shifted coefficients x - h.
__global__ __launch_bounds__(256) void k_proj_bias(const uint8_t *__restrict__ zero,
                                                   const uint8_t *__restrict__ sign, size_t cw,
                                                   uint64_t *__restrict__ bias, Mod M) {
    __shared__ int part[4];
    const int k = blockIdx.x, tid = threadIdx.x;
    int s = 0;
    for (size_t l = tid; l < cw; l += 256) {
        const size_t o = (size_t)k * cw + l;
        s += zero[o] ? 0 : (sign[o] ? 1 : -1);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((tid & 63) == 0) part[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        const int t = part[0] + part[1] + part[2] + part[3];  // |t| <= cw
        const double h = (double)(((uint64_t)M.q - 1) / 2);
        bias[k] = d2u(mulmod(h, canon((double)t, M), M));
    }
}

__global__ __launch_bounds__(PM_THREADS, 2) void k_project_mfma(
    const uint64_t *__restrict__ W, size_t m, size_t cw, size_t nch,
    const uint8_t *__restrict__ Cf, const uint64_t *__restrict__ bias, uint64_t *__restrict__ out,
    const double *__restrict__ tab, Mod M) {
    __shared__ __align__(16) char lds[2 * PM_STAGE];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const size_t idx = blockIdx.x, i = idx / nch, j = idx % nch;
    const int cg = blockIdx.y;
    // staging role: K step sks, element group sg, coefficient sc
    const int sks = tid >> 6, sg = (tid >> 4) & 3, sc = tid & 15;
    const uint64_t *src = W + (i * m + j * cw + 64 * sks + 16 * sg) * kD + 16 * cg + sc;
    // u = x - h + 0x80..80 (seven 0x80 bytes): 0 <= u < 2^56 for x in [0, q)
    const uint64_t ushift = 0x0080808080808080ULL - (((uint64_t)M.q - 1) >> 1);
    const int nst = (int)(cw / 256);
    uint64_t R[16];
    auto gload = [&](int st) {
        const uint64_t *p = src + (size_t)st * 256 * kD;
#pragma unroll
        for (int e = 0; e < 16; e++) R[e] = p[e * kD];
    };
    auto stage = [&](int buf) {
        uint32_t d[4][7];
#pragma unroll
        for (int h = 0; h < 4; h++) {
            uint32_t lo[4], hi[4];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const uint64_t u = R[4 * h + b] + ushift;
                lo[b] = (uint32_t)u;
                hi[b] = (uint32_t)(u >> 32);
            }
            bytes_transpose4(lo, hi, d[h]);
#pragma unroll
            for (int t = 0; t < 7; t++) d[h][t] ^= 0x80808080u;  // d_t + 128 -> d_t as int8
        }
        char *base = lds + buf * PM_STAGE + sks * 7 * 1024 + (16 * sg + sc) * 16;
#pragma unroll
        for (int t = 0; t < 7; t++)
            *reinterpret_cast<v4i_pm *>(base + t * 1024) = (v4i_pm){(int)d[0][t], (int)d[1][t],
                                                                    (int)d[2][t], (int)d[3][t]};
    };
    v4i_pm acc[4][7];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int t = 0; t < 7; t++) acc[a][t] = (v4i_pm){0, 0, 0, 0};
    const uint8_t *cf = Cf + (size_t)(4 * w) * 1024 + lane * 16;
    const int nk = 4 * nst;
    v4i_pm fa[4], fn[4];
#pragma unroll
    for (int a = 0; a < 4; a++) fa[a] = *reinterpret_cast<const v4i_pm *>(cf + a * 1024);
    gload(0);
    for (int st = 0; st < nst; st++) {
        stage(st & 1);
        if (st + 1 < nst) gload(st + 1);
        __syncthreads();
        const char *bb = lds + (st & 1) * PM_STAGE + lane * 16;
#pragma unroll 2
        for (int ks = 0; ks < 4; ks++) {
            // the next K step's sign fragments (the last one repeats itself)
            const int kn = min(4 * st + ks + 1, nk - 1);
            const uint8_t *ca = cf + (size_t)kn * 16 * 1024;
            // issued here, ahead of this step's products (the barriers keep the
            // compiler from sinking them to the end of the step)
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int a = 0; a < 4; a++) fn[a] = *reinterpret_cast<const v4i_pm *>(ca + a * 1024);
            __builtin_amdgcn_sched_barrier(0);
            asm volatile("" ::: "memory");
#pragma unroll
            for (int t = 0; t < 7; t++) {
                const v4i_pm fb = *reinterpret_cast<const v4i_pm *>(bb + (ks * 7 + t) * 1024);
#pragma unroll
                for (int a = 0; a < 4; a++)
                    acc[a][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[a], fb, acc[a][t], 0, 0, 0);
            }
#pragma unroll
            for (int a = 0; a < 4; a++) fa[a] = fn[a];
        }
    }
    // recombination: lane (g, lr) holds rows 16 (4 w + a) + 4 g + v, column lr.
    // sum_t 2^(8t) acc_t = lo + 2^32 hi with lo = acc_0 + 2^8 acc_1 + 2^16 acc_2
    // + 2^24 acc_3 and hi = acc_4 + 2^8 acc_5 + 2^16 acc_6, both exact in fp64
    // (|acc_t| <= 128 cw < 2^23: |lo| < 2^48, |hi| < 2^40), and one modular
    // product for 2^32 hi.
    const double w32 = tab[TAB_W256 + 4];
    const int lr = lane & 15, g = lane >> 4;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            double d[7];
#pragma unroll
            for (int t = 0; t < 7; t++) d[t] = (double)acc[a][t][v];
            const double lo =
                __fma_rn(d[3], 16777216.0, __fma_rn(d[2], 65536.0, __fma_rn(d[1], 256.0, d[0])));
            const double hi = __fma_rn(d[6], 65536.0, __fma_rn(d[5], 256.0, d[4]));
            const int k = 16 * (4 * w + a) + 4 * g + v;
            const double x = lo + mul_lazy(hi, w32, M) + u2d(bias[k]);  // |x| < 2^48 + 2q
            out[(idx * PJ_H + k) * kD + 16 * cg + lr] = d2u(reduce(x, M));
        }
}

// out[i][l] = sum_j src[i][j cw + l] mod q (j < nch) for rows i < rows of
// length m, per coefficient (u64 sums, reduced every 8192 terms).  With few
// outputs the chunks are split over gridDim.y: each part adds its sum (mod q)
// to acc, and k_chunk_sums_finish reduces acc into out.
__global__ void k_chunk_sums(const uint64_t *__restrict__ src, size_t rows, size_t m, size_t cw,
                             size_t nch, uint64_t *__restrict__ out,
                             unsigned long long *__restrict__ acc, uint64_t q) {
    const size_t total = rows * cw * kD;
    const size_t per = (nch + gridDim.y - 1) / gridDim.y;
    const size_t j0 = (size_t)blockIdx.y * per, j1 = min(nch, j0 + per);
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const size_t i = e / (cw * kD), lc = e % (cw * kD);
        const uint64_t *p = src + i * m * kD + lc;
        uint64_t a = 0;
        for (size_t j = j0; j < j1; j++) {
            a += p[j * cw * kD];
            if (((j - j0) & 8191) == 8191) a %= q;  // below (8192 + 1) q < 2^64
        }
        if (gridDim.y == 1)
            out[e] = a % q;
        else
            atomicAdd(&acc[e], (unsigned long long)(a % q));  // at most gridDim.y q < 2^64
    }
}

__global__ void k_chunk_sums_finish(const unsigned long long *__restrict__ acc, size_t n,
                                    uint64_t *__restrict__ out, uint64_t q) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        out[e] = acc[e] % q;
}

__global__ void k_fill_u64(unsigned long long *__restrict__ p, size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        p[e] = 0;
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
    if (ctx->mfma_project) {
        // scratch: the tiled signs [256 cw] bytes, then the biases [256]
        uint8_t *Cf = nullptr;
        SALSAA_TRY(scratch_get(ctx, 10, (size_t)PJ_H * cw + PJ_H * sizeof(uint64_t), (void **)&Cf));
        uint64_t *bias = reinterpret_cast<uint64_t *>(Cf + (size_t)PJ_H * cw);
        const size_t nslot = (size_t)PJ_H * cw / 16;
        hipLaunchKernelGGL(k_proj_cfrag, dim3(ceil_div(nslot, (size_t)256)), dim3(256), 0,
                           ctx->stream, is_zero, sign, cw, Cf);
        SALSAA_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_proj_bias, dim3(PJ_H), dim3(256), 0, ctx->stream, is_zero, sign, cw,
                           bias, M);
        SALSAA_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_project_mfma, dim3((unsigned)(r * nch), kD / 16), dim3(PM_THREADS), 0,
                           ctx->stream, W, m, cw, nch, Cf, bias, out, ctx->d_tables, M);
    } else {
        hipLaunchKernelGGL(k_project<false>, dim3((unsigned)(r * nch), PJ_H / PJ_KB),
                           dim3(PJ_THREADS), 0, ctx->stream, W, m, cw, nch, is_zero, sign, out, M);
    }
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
                      uint64_t *out, unsigned long long *acc) {
    SALSAA_CHECK(cw >= 1 && m % cw == 0, "chunk sums need cw | m");
    const size_t total = rows * cw * kD, nch = m / cw;
    // about a million summing threads: split the chunks when the outputs are fewer
    size_t split = 1;
    while (acc && split < nch && split < 8192 && total * split < ((size_t)1 << 20)) split *= 2;
    const unsigned g = (unsigned)std::min<size_t>(ceil_div(total, (size_t)256), (size_t)ctx->ncu * 16);
    const unsigned gz = (unsigned)std::min<size_t>(ceil_div(total, (size_t)256), 4096);
    if (split > 1)
        hipLaunchKernelGGL(k_fill_u64, dim3(gz), dim3(256), 0, ctx->stream, acc, total);
    hipLaunchKernelGGL(k_chunk_sums, dim3(g, (unsigned)split), dim3(256), 0, ctx->stream, src, rows,
                       m, cw, nch, out, acc, ctx->q);
    if (split > 1)
        hipLaunchKernelGGL(k_chunk_sums_finish, dim3(gz), dim3(256), 0, ctx->stream, acc, total,
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
