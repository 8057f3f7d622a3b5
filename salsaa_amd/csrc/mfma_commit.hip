// mfma_commit.hip -- the commit's per-slot products on the int8 matrix cores.
//
// In the full NTT domain the commit is, for every slot s, C_s = A_s . W_s over
// F_q with A_s (kappa x m) and W_s (m x r) holding 50-bit residues.  Each
// residue x is written as 7 signed base-256 digits x = sum_u d_u 256^u,
// d_u in [-128, 127], stored as int8 planes [64][rows * 7][stride] (row
// rows*7 + u holds digit u).  Then
//     C_s[j][k] = sum_{u,v} 256^(u+v) * (A_s,u . W_s,v)[j][k],
// and the 49 digit products are ONE integer GEMM per slot with M = 7 kappa,
// N = 7 r, K = m on v_mfma_i32_16x16x64_i8: every digit product is at most
// 2^14 in magnitude, so int32 accumulation over a K chunk of at most 2^16 is
// exact.  A block (4 waves) owns a K chunk of one slot, waves owning strips of
// 9 key rows (63 digit rows) against 8 witness rows (56 digit rows); the int32
// tile goes through LDS and each (j, k) is recombined with the weights
// 2^(8t) mod q in exact fp64 into the same per-slot partial sums the VALU
// kernel produces.
#include "common.hpp"

namespace salsaa {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int MF_WAVES = 4;
constexpr int MJ = 9;        // key rows per wave strip (63 digit rows)
constexpr int NK = 8;        // witness rows per block strip (56 digit rows)
constexpr int KSTEP = 64;    // K per MFMA

__global__ __launch_bounds__(MF_WAVES * 64) void k_commit_mfma(
    const int8_t *__restrict__ AB, int kappa, size_t strideA, const int8_t *__restrict__ WB, int r,
    size_t strideW, size_t kchunk, size_t kend, int nbx, int mbx, double *__restrict__ partial,
    const double *__restrict__ tab, Mod M) {
    __shared__ int Dl[64][65];
    // XCD-aware block order: the 8 XCDs take consecutive runs of logical
    // blocks, so the strips that share one key chunk (nb fastest) share an L2
    const unsigned nblk = gridDim.x * gridDim.y * gridDim.z;
    const unsigned pid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const unsigned lid = (nblk % 8 == 0) ? (pid % 8) * (nblk / 8) + pid / 8 : pid;
    const int nb = (int)(lid % (unsigned)nbx);
    unsigned rest = lid / (unsigned)nbx;
    const int mb = (int)(rest % (unsigned)mbx);
    rest /= (unsigned)mbx;
    const int s = (int)(rest % 64u);
    const int z = (int)(rest / 64u);

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lr = lane & 15, g = lane >> 4;
    const int j0 = mb * (MF_WAVES * MJ) + w * MJ;
    const int k0 = nb * NK;
    const int rowsA = 7 * kappa, rowsB = 7 * r;
    const size_t kb = (size_t)z * kchunk;
    const size_t ke = min(kend, kb + kchunk);

    // this lane's fragment rows (or -1 for zero padding)
    const int8_t *pa[4];
    const int8_t *pb[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int mr = 16 * t + lr, ga = 7 * j0 + mr;
        pa[t] = (mr < 7 * MJ && ga < rowsA) ? AB + ((size_t)s * rowsA + ga) * strideA + 16 * g : nullptr;
        const int nr = 16 * t + lr, gb = 7 * k0 + nr;
        pb[t] = (nr < 7 * NK && gb < rowsB) ? WB + ((size_t)s * rowsB + gb) * strideW + 16 * g : nullptr;
    }
    v4i acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = (v4i){0, 0, 0, 0};
    const v4i zero = {0, 0, 0, 0};
    for (size_t kk = kb; kk < ke; kk += KSTEP) {
        v4i af[4], bf[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            af[t] = pa[t] ? *reinterpret_cast<const v4i *>(pa[t] + kk) : zero;
            bf[t] = pb[t] ? *reinterpret_cast<const v4i *>(pb[t] + kk) : zero;
        }
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++)
                acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    // recombination, one wave at a time through a 64 x 64 int32 tile
    const double *wt = tab + TAB_W256;
    for (int ww = 0; ww < MF_WAVES; ww++) {
        __syncthreads();
        if (w == ww) {
            // D[16a + 4g + v][16b + lr] = acc[a][b][v]
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++)
#pragma unroll
                    for (int v = 0; v < 4; v++) Dl[16 * a + 4 * g + v][16 * b + lr] = acc[a][b][v];
        }
        __syncthreads();
        if (w == ww) {
            for (int o = lane; o < MJ * NK; o += 64) {
                const int jl = o / NK, kl = o % NK;
                const int j = j0 + jl, k = k0 + kl;
                if (j >= kappa || k >= r) continue;
                double x = 0.0;
#pragma unroll
                for (int u = 0; u < 7; u++) {
#pragma unroll
                    for (int v = 0; v < 7; v++)
                        x += mul_lazy((double)Dl[jl * 7 + u][kl * 7 + v], wt[u + v], M);
                    x = reduce_lazy(x, M);
                }
                partial[(((size_t)z * 64 + s) * kappa + j) * r + k] = reduce(x, M);
            }
        }
    }
}

}  // namespace

int launch_commit_mfma(Context *ctx, const int8_t *AB, size_t kappa, size_t strideA,
                       const int8_t *WB, size_t r, size_t m, size_t strideW, uint64_t *out) {
    Mod M{ctx->qd, ctx->qinv};
    const size_t kend = (m + KSTEP - 1) / KSTEP * KSTEP;  // planes are zero past m
    SALSAA_CHECK(kend <= strideW && kend <= strideA, "plane rows shorter than the witness");
    const size_t nbx = (r + NK - 1) / NK, mbx = (kappa + MF_WAVES * MJ - 1) / (MF_WAVES * MJ);
    // K chunks of at most 2^16 (exact int32 accumulation), at least 1024 when
    // that still leaves >= ~1024 blocks
    size_t ksplit = (kend + 65535) / 65536;
    while (nbx * mbx * 64 * ksplit < 1024 && kend / (ksplit * 2) >= 1024) ksplit *= 2;
    size_t kchunk = (kend + ksplit - 1) / ksplit;
    kchunk = (kchunk + KSTEP - 1) / KSTEP * KSTEP;
    ksplit = (kend + kchunk - 1) / kchunk;
    double *partial = nullptr;
    SALSAA_TRY(scratch_get(ctx, 7, ksplit * 64 * kappa * r * sizeof(double), (void **)&partial));
    const size_t nblk = nbx * mbx * 64 * ksplit;
    SALSAA_CHECK(nblk < (1u << 31), "commit grid too large");
    // linear grid in x (decoded in the kernel)
    hipLaunchKernelGGL(k_commit_mfma, dim3((unsigned)nblk), dim3(MF_WAVES * 64), 0, ctx->stream, AB,
                       (int)kappa, strideA, WB, (int)r, strideW, kchunk, kend, (int)nbx, (int)mbx,
                       partial, ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return launch_commit_finalize(ctx, partial, ksplit, kappa, r, out);
}

}  // namespace salsaa
