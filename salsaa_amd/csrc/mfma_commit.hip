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
// exact.  A block owns a K chunk of one slot (layout below); the int32 tiles
// go through LDS and each (j, k) is recombined with the weights 2^(8t) mod q
// in exact fp64 into the same per-slot partial sums the VALU kernel produces.
#include "common.hpp"

namespace salsaa {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int MJ = 9;        // key rows per M strip (63 digit rows, 4 MFMA row tiles)
constexpr int NK = 8;        // witness rows per N strip (56 digit rows)
constexpr int KSTEP = 64;    // K per MFMA

// Block = 4 M-strip waves x NWN N-groups; a wave owns one M strip (4 row
// tiles) and NSPW consecutive N strips (NT = ceil(56 NSPW / 16) column tiles:
// 4 for one strip, 7 for two).  All waves of a block walk the same K range,
// so the key rows and witness rows they share come from L1 and each byte
// plane is streamed from HBM once per block.
template <int NSPW, int NWN>
__global__ __launch_bounds__(4 * NWN * 64) void k_commit_mfma(
    const int8_t *__restrict__ AB, int kappa, size_t strideA, const int8_t *__restrict__ WB, int r,
    size_t strideW, size_t kchunk, size_t kend, int nbx, int mbx, const int8_t *__restrict__ zrow,
    double *__restrict__ partial, const double *__restrict__ tab, Mod M) {
    constexpr int NT = (56 * NSPW + 15) / 16;
    constexpr int NROWS = 56 * NSPW;    // digit columns owned by a wave
    constexpr int NW = 4 * NWN;         // waves per block
    constexpr int LDSW = NROWS + 1;
    __shared__ int Dl[2][64][LDSW];     // 2 waves' tiles at a time
    // XCD-aware block order: the 8 XCDs take consecutive runs of logical
    // blocks, so the blocks sharing one key chunk share an L2
    const unsigned nblk = gridDim.x;
    const unsigned pid = blockIdx.x;
    const unsigned lid = (nblk % 8 == 0) ? (pid % 8) * (nblk / 8) + pid / 8 : pid;
    const int nb = (int)(lid % (unsigned)nbx);
    unsigned rest = lid / (unsigned)nbx;
    const int mb = (int)(rest % (unsigned)mbx);
    rest /= (unsigned)mbx;
    const int s = (int)(rest % 64u);
    const int z = (int)(rest / 64u);

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = w & 3, wn = w >> 2;
    const int lr = lane & 15, g = lane >> 4;
    const int j0 = mb * (4 * MJ) + wm * MJ;
    const int k0 = (nb * NWN + wn) * NSPW * NK;
    const int rowsA = 7 * kappa, rowsB = 7 * r;
    const size_t kb = (size_t)z * kchunk;
    const size_t ke = min(kend, kb + kchunk);

    // fragment rows: out-of-range rows read the zero row `zrow` (at least
    // kend + 64 bytes of zeros), so every load is unconditional and feeds the
    // matrix cores directly
    const int8_t *pa[4];
    const int8_t *pb[NT];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int mr = 16 * t + lr, ga = 7 * j0 + mr;
        const bool ok = mr < 7 * MJ && ga < rowsA;
        pa[t] = (ok ? AB + ((size_t)s * rowsA + ga) * strideA : zrow) + 16 * g;
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int nr = 16 * t + lr, gb = 7 * k0 + nr;
        const bool ok = nr < NROWS && gb < rowsB;
        pb[t] = (ok ? WB + ((size_t)s * rowsB + gb) * strideW : zrow) + 16 * g;
    }
    v4i acc[4][NT];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < NT; b++) acc[a][b] = (v4i){0, 0, 0, 0};
    // Two-stage ping-pong over K (ke - kb is a multiple of 128): buffer 1's
    // loads are in flight while buffer 0 feeds the matrix cores and vice
    // versa; the last refill of buffer 0 re-reads a valid step and is unused.
    v4i a0[4], b0[NT], a1[4], b1[NT];
    auto load = [&](v4i(&af)[4], v4i(&bf)[NT], size_t k) {
#pragma unroll
        for (int t = 0; t < 4; t++) af[t] = *reinterpret_cast<const v4i *>(pa[t] + k);
#pragma unroll
        for (int t = 0; t < NT; t++) bf[t] = *reinterpret_cast<const v4i *>(pb[t] + k);
    };
    auto mma = [&](v4i(&af)[4], v4i(&bf)[NT]) {
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < NT; b++)
                acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[a], bf[b], acc[a][b], 0, 0, 0);
    };
    if (kb < ke) load(a0, b0, kb);
    for (size_t kk = kb; kk < ke; kk += 2 * KSTEP) {
        load(a1, b1, kk + KSTEP);
        mma(a0, b0);
        load(a0, b0, kk + 2 * KSTEP < ke ? kk + 2 * KSTEP : kk);
        mma(a1, b1);
    }
    // recombination through LDS, 2 waves at a time
    const double *wt = tab + TAB_W256;
    for (int grp = 0; grp < NW; grp += 2) {
        __syncthreads();
        const bool mine = w >= grp && w < grp + 2;
        int (*D)[LDSW] = Dl[w & 1];
        if (mine) {
            // D[16a + 4g + v][16b + lr] = acc[a][b][v]
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < NT; b++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int col = 16 * b + lr;
                        if (col < NROWS) D[16 * a + 4 * g + v][col] = acc[a][b][v];
                    }
        }
        __syncthreads();
        if (mine) {
            for (int o = lane; o < MJ * NK * NSPW; o += 64) {
                const int jl = o / (NK * NSPW), kl = o % (NK * NSPW);
                const int j = j0 + jl, k = k0 + kl;
                if (j >= kappa || k >= r) continue;
                // diagonals E_t = sum_{u+v=t} D[u][v]: at most 7 int32 terms,
                // exact in fp64 and below 2^34 < q in magnitude
                double E[13];
#pragma unroll
                for (int t = 0; t < 13; t++) E[t] = 0.0;
#pragma unroll
                for (int u = 0; u < 7; u++)
#pragma unroll
                    for (int v = 0; v < 7; v++) E[u + v] += (double)D[jl * 7 + u][kl * 7 + v];
                double x = 0.0;
#pragma unroll
                for (int t = 0; t < 13; t++) {
                    x += mul_lazy(E[t], wt[t], M);
                    if ((t % 7) == 6) x = reduce_lazy(x, M);
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
    // planes are zero from m up to the next multiple of 128
    const size_t kend = (m + 2 * KSTEP - 1) / (2 * KSTEP) * (2 * KSTEP);
    SALSAA_CHECK(kend <= strideW && kend <= strideA, "plane rows shorter than the witness");
    // N layout: one strip per wave when r <= 8, else two strips per wave and
    // up to 2 N-wave groups per block (8 waves; more would spill the
    // 112 accumulator registers of a two-strip wave)
    const size_t nstrips = (r + NK - 1) / NK;
    const int nspw = nstrips == 1 ? 1 : 2;
    const size_t ngroups = (nstrips + nspw - 1) / nspw;
    const int nwn = ngroups >= 2 ? 2 : 1;
    const size_t nbx = (ngroups + nwn - 1) / nwn;
    const size_t mbx = (kappa + 4 * MJ - 1) / (4 * MJ);
    // K chunks of at most 2^16 (exact int32 accumulation), at least 1024 when
    // that still leaves >= ~1024 blocks
    size_t ksplit = (kend + 65535) / 65536;
    while (nbx * mbx * 64 * ksplit < 1024 && kend / (ksplit * 2) >= 1024) ksplit *= 2;
    size_t kchunk = (kend + ksplit - 1) / ksplit;
    kchunk = (kchunk + 2 * KSTEP - 1) / (2 * KSTEP) * (2 * KSTEP);
    ksplit = (kend + kchunk - 1) / kchunk;
    double *partial = nullptr;
    SALSAA_TRY(scratch_get(ctx, 7, ksplit * 64 * kappa * r * sizeof(double), (void **)&partial));
    const int8_t *zrow = nullptr;
    SALSAA_TRY(zero_row(ctx, kend + 2 * KSTEP, &zrow));
    const size_t nblk = nbx * mbx * 64 * ksplit;
    SALSAA_CHECK(nblk < (1u << 31), "commit grid too large");
#define SALSAA_MFMA_LAUNCH(NSPW, NWN)                                                              \
    hipLaunchKernelGGL((k_commit_mfma<NSPW, NWN>), dim3((unsigned)nblk), dim3(4 * NWN * 64), 0,      \
                       ctx->stream, AB, (int)kappa, strideA, WB, (int)r, strideW, kchunk, kend,    \
                       (int)nbx, (int)mbx, zrow, partial, ctx->d_tables, M)
    if (nspw == 1) SALSAA_MFMA_LAUNCH(1, 1);
    else if (nwn == 1) SALSAA_MFMA_LAUNCH(2, 1);
    else SALSAA_MFMA_LAUNCH(2, 2);
#undef SALSAA_MFMA_LAUNCH
    SALSAA_HIP(hipGetLastError());
    return launch_commit_finalize(ctx, partial, ksplit, kappa, r, out);
}

}  // namespace salsaa
