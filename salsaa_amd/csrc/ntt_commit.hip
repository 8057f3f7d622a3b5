// ntt_commit.hip -- batched negacyclic NTT kernels (d = 64) and the commit
// kernels (ring matrix product A . W in the full 64-slot NTT domain).
//
// NTT: a block stages 128 ring elements (64 KiB) through LDS with coalesced
// 16-byte global accesses; each thread then runs one whole length-32
// (IncompleteNTT half) or length-64 (full) transform in registers.  The LDS
// row stride and the half offset are padded so the per-thread strided LDS
// accesses are bank-conflict free.
//
// Commit: out[j][k] = sum_i A_j[i] * W_k[i] in R_q.  In the full NTT domain
// this is 64 independent (kappa x m) . (m x r) products over F_q, one per
// slot.  A (resident key) and W are stored slot-major [64][rows][m] as exact
// doubles, so lanes walking i read contiguous memory.  Each thread keeps a
// JT x KT register tile of lazy accumulators (one reduction every 8 MACs);
// blocks split K and write per-slot partial sums; a finalize kernel sums the
// partials, runs the inverse 64-point NTT and the forward IncompleteNTT, and
// writes the reference's IncompleteNTT output (arithmetic.rs:459-505,
// cyclotomic_ring.rs:789).
#include "ntt_device.hpp"

namespace salsaa {

namespace {

constexpr int NTT_EL = 128;      // ring elements per block
constexpr int S_INC = 66;        // LDS row stride (u64) for the IncompleteNTT kernel
constexpr int S_FULL = 65;       // LDS row stride (u64) for the full-NTT kernels

// half-layout position of entry k (0..63) in an IncompleteNTT LDS row:
// the odd half starts at 33 instead of 32
__device__ __forceinline__ int half_pos(int k) { return k + (k >> 5); }

template <bool INV>
__global__ __launch_bounds__(256) void k_ntt_incomplete(uint64_t *__restrict__ data,
                                                        size_t count,
                                                        const double *__restrict__ tab,
                                                        Mod M) {
    __shared__ uint64_t lds[NTT_EL * S_INC];
    const size_t e0 = (size_t)blockIdx.x * NTT_EL;
    const int nel = (int)min((size_t)NTT_EL, count - e0);
    const int tid = threadIdx.x;
    const ulonglong2 *src = reinterpret_cast<const ulonglong2 *>(data + e0 * kD);
    // forward input is Coefficient (plain layout); inverse input is
    // IncompleteNTT (half layout)
    for (int c = tid; c < nel * 32; c += 256) {
        int e = c >> 5, k = (c & 31) * 2;
        ulonglong2 v = src[c];
        int p0 = INV ? half_pos(k) : k, p1 = INV ? half_pos(k + 1) : k + 1;
        lds[e * S_INC + p0] = v.x;
        lds[e * S_INC + p1] = v.y;
    }
    __syncthreads();
    const int e = tid >> 1, p = tid & 1;
    double v[32];
    if (e < nel) {
        if (!INV) {
#pragma unroll
            for (int i = 0; i < 32; i++) v[i] = u2d(lds[e * S_INC + 2 * i + p]);
            ntt_fwd_regs<32>(v, tab + TAB_FWD32, M);
        } else {
#pragma unroll
            for (int i = 0; i < 32; i++) v[i] = u2d(lds[e * S_INC + 33 * p + i]);
            ntt_inv_regs<32>(v, tab + TAB_INV32, tab[TAB_NINV32], M);
        }
    }
    __syncthreads();
    if (e < nel) {
        if (!INV) {
#pragma unroll
            for (int i = 0; i < 32; i++) lds[e * S_INC + 33 * p + i] = d2u(v[i]);
        } else {
#pragma unroll
            for (int i = 0; i < 32; i++) lds[e * S_INC + 2 * i + p] = d2u(v[i]);
        }
    }
    __syncthreads();
    ulonglong2 *dst = reinterpret_cast<ulonglong2 *>(data + e0 * kD);
    for (int c = tid; c < nel * 32; c += 256) {
        int e2 = c >> 5, k = (c & 31) * 2;
        int p0 = INV ? k : half_pos(k), p1 = INV ? k + 1 : half_pos(k + 1);
        ulonglong2 v2;
        v2.x = lds[e2 * S_INC + p0];
        v2.y = lds[e2 * S_INC + p1];
        dst[c] = v2;
    }
}

template <bool INV>
__global__ __launch_bounds__(128) void k_ntt_full(uint64_t *__restrict__ data, size_t count,
                                                  const double *__restrict__ tab, Mod M) {
    __shared__ uint64_t lds[NTT_EL * S_FULL];
    const size_t e0 = (size_t)blockIdx.x * NTT_EL;
    const int nel = (int)min((size_t)NTT_EL, count - e0);
    const int tid = threadIdx.x;
    const ulonglong2 *src = reinterpret_cast<const ulonglong2 *>(data + e0 * kD);
    for (int c = tid; c < nel * 32; c += 128) {
        int e = c >> 5, k = (c & 31) * 2;
        ulonglong2 v = src[c];
        lds[e * S_FULL + k] = v.x;
        lds[e * S_FULL + k + 1] = v.y;
    }
    __syncthreads();
    double v[64];
    if (tid < nel) {
#pragma unroll
        for (int i = 0; i < 64; i++) v[i] = u2d(lds[tid * S_FULL + i]);
        if (!INV) ntt_fwd_regs<64>(v, tab + TAB_FWD64, M);
        else ntt_inv_regs<64>(v, tab + TAB_INV64, tab[TAB_NINV64], M);
#pragma unroll
        for (int i = 0; i < 64; i++) lds[tid * S_FULL + i] = d2u(v[i]);
    }
    __syncthreads();
    ulonglong2 *dst = reinterpret_cast<ulonglong2 *>(data + e0 * kD);
    for (int c = tid; c < nel * 32; c += 128) {
        int e = c >> 5, k = (c & 31) * 2;
        ulonglong2 v2;
        v2.x = lds[e * S_FULL + k];
        v2.y = lds[e * S_FULL + k + 1];
        dst[c] = v2;
    }
}

// Coefficient ring elements -> full NTT, written slot-major as exact doubles:
// dst[(s * rows + row) * m + i].  Source row `row` starts at
// (row % row_lo) * stride_lo + (row / row_lo) * stride_hi (u64 units), which
// expresses both plain matrices and the halves of split (split.rs:21-28).
__global__ __launch_bounds__(128) void k_ntt_full_to_slotmajor(
    const uint64_t *__restrict__ src, size_t rows, size_t m, size_t row_lo, size_t stride_lo,
    size_t stride_hi, double *__restrict__ dst, const double *__restrict__ tab, Mod M) {
    __shared__ uint64_t lds[NTT_EL * S_FULL];
    const size_t row = blockIdx.y;
    const size_t i0 = (size_t)blockIdx.x * NTT_EL;
    const int nel = (int)min((size_t)NTT_EL, m - i0);
    const int tid = threadIdx.x;
    const uint64_t *base = src + (row % row_lo) * stride_lo + (row / row_lo) * stride_hi;
    const ulonglong2 *s2 = reinterpret_cast<const ulonglong2 *>(base + i0 * kD);
    for (int c = tid; c < nel * 32; c += 128) {
        int e = c >> 5, k = (c & 31) * 2;
        ulonglong2 v = s2[c];
        lds[e * S_FULL + k] = v.x;
        lds[e * S_FULL + k + 1] = v.y;
    }
    __syncthreads();
    if (tid < nel) {
        double v[64];
#pragma unroll
        for (int i = 0; i < 64; i++) v[i] = u2d(lds[tid * S_FULL + i]);
        ntt_fwd_regs<64>(v, tab + TAB_FWD64, M);
        const size_t i = i0 + tid;
#pragma unroll
        for (int s = 0; s < 64; s++) dst[((size_t)s * rows + row) * m + i] = v[s];
    }
}

// ------------------------------------------------------------------ commit
constexpr int CM_THREADS = 256;

// grid: x = tile (jt * ktiles + kt), y = slot, z = K split.
// partial[z][s][j][k] = sum over this block's i range, in [0, q).
template <int JT, int KT>
__global__ __launch_bounds__(CM_THREADS) void k_commit_mac(
    const double *__restrict__ A_sm, int kappa, size_t mA, const double *__restrict__ W_sm,
    int r, size_t m, size_t chunk, double *__restrict__ partial, Mod M) {
    const int ktiles = (r + KT - 1) / KT;
    const int jt = blockIdx.x / ktiles, kt = blockIdx.x % ktiles;
    const int j0 = jt * JT, k0 = kt * KT;
    const int s = blockIdx.y;
    const size_t ib = (size_t)blockIdx.z * chunk;
    const size_t ie = min(m, ib + chunk);
    const double *Ab = A_sm + ((size_t)s * kappa) * mA;
    const double *Wb = W_sm + ((size_t)s * r) * m;

    double acc[JT][KT];
#pragma unroll
    for (int a = 0; a < JT; a++)
#pragma unroll
        for (int b = 0; b < KT; b++) acc[a][b] = 0.0;

    int cnt = 0;
    for (size_t i = ib + threadIdx.x; i < ie; i += CM_THREADS) {
        double av[JT], wv[KT];
#pragma unroll
        for (int a = 0; a < JT; a++) av[a] = (j0 + a < kappa) ? Ab[(size_t)(j0 + a) * mA + i] : 0.0;
#pragma unroll
        for (int b = 0; b < KT; b++) wv[b] = (k0 + b < r) ? Wb[(size_t)(k0 + b) * m + i] : 0.0;
#pragma unroll
        for (int a = 0; a < JT; a++)
#pragma unroll
            for (int b = 0; b < KT; b++) acc[a][b] += mul_lazy(av[a], wv[b], M);
        if (++cnt == 7) {
            cnt = 0;
#pragma unroll
            for (int a = 0; a < JT; a++)
#pragma unroll
                for (int b = 0; b < KT; b++) acc[a][b] = reduce_lazy(acc[a][b], M);
        }
    }
    // canonical, then block reduction (wave shuffles, then LDS across waves)
    __shared__ double red[CM_THREADS / 64][JT * KT];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < JT; a++)
#pragma unroll
        for (int b = 0; b < KT; b++) {
            double x = reduce(acc[a][b], M);
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) x = addmod(x, __shfl_xor(x, off, 64), M);
            if (lane == 0) red[wave][a * KT + b] = x;
        }
    __syncthreads();
    if (threadIdx.x < JT * KT) {
        double x = red[0][threadIdx.x];
#pragma unroll
        for (int w = 1; w < CM_THREADS / 64; w++) x = addmod(x, red[w][threadIdx.x], M);
        const int a = threadIdx.x / KT, b = threadIdx.x % KT;
        if (j0 + a < kappa && k0 + b < r)
            partial[(((size_t)blockIdx.z * 64 + s) * kappa + (j0 + a)) * r + (k0 + b)] = x;
    }
}

// One thread per output ring element (j, k): sum the K-split partials of all
// 64 slots, inverse full NTT -> coefficients, forward IncompleteNTT.
__global__ __launch_bounds__(64) void k_commit_finalize(const double *__restrict__ partial,
                                                        int ksplit, int kappa, int r,
                                                        uint64_t *__restrict__ out,
                                                        const double *__restrict__ tab, Mod M) {
    const int jk = blockIdx.x * 64 + threadIdx.x;
    if (jk >= kappa * r) return;
    double v[64];
#pragma unroll
    for (int s = 0; s < 64; s++) {
        double x = 0.0;
        for (int z = 0; z < ksplit; z++) x = addmod(x, partial[((size_t)z * 64 + s) * kappa * r + jk], M);
        v[s] = x;
    }
    ntt_inv_regs<64>(v, tab + TAB_INV64, tab[TAB_NINV64], M);
    double e[32], o[32];
#pragma unroll
    for (int i = 0; i < 32; i++) {
        e[i] = v[2 * i];
        o[i] = v[2 * i + 1];
    }
    ntt_fwd_regs<32>(e, tab + TAB_FWD32, M);
    ntt_fwd_regs<32>(o, tab + TAB_FWD32, M);
    uint64_t *dst = out + (size_t)jk * kD;
#pragma unroll
    for (int i = 0; i < 32; i++) {
        dst[i] = d2u(e[i]);
        dst[32 + i] = d2u(o[i]);
    }
}

}  // namespace

int launch_ntt(Context *ctx, uint64_t *data, size_t count, int mode, int inverse) {
    if (count == 0) return SALSAA_OK;
    Mod M{ctx->qd, ctx->qinv};
    unsigned grid = ceil_div(count, NTT_EL);
    if (mode == SALSAA_REP_INCOMPLETE) {
        if (inverse)
            hipLaunchKernelGGL(k_ntt_incomplete<true>, dim3(grid), dim3(256), 0, ctx->stream, data,
                               count, ctx->d_tables, M);
        else
            hipLaunchKernelGGL(k_ntt_incomplete<false>, dim3(grid), dim3(256), 0, ctx->stream, data,
                               count, ctx->d_tables, M);
    } else {
        if (inverse)
            hipLaunchKernelGGL(k_ntt_full<true>, dim3(grid), dim3(128), 0, ctx->stream, data, count,
                               ctx->d_tables, M);
        else
            hipLaunchKernelGGL(k_ntt_full<false>, dim3(grid), dim3(128), 0, ctx->stream, data, count,
                               ctx->d_tables, M);
    }
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_ntt_full_to_slotmajor(Context *ctx, const uint64_t *src, size_t rows, size_t m,
                                 size_t row_lo, size_t stride_lo, size_t stride_hi,
                                 double *dst) {
    if (rows == 0 || m == 0) return SALSAA_OK;
    SALSAA_CHECK(rows < 65536, "too many rows for one launch");
    Mod M{ctx->qd, ctx->qinv};
    dim3 grid(ceil_div(m, NTT_EL), (unsigned)rows);
    hipLaunchKernelGGL(k_ntt_full_to_slotmajor, grid, dim3(128), 0, ctx->stream, src, rows, m,
                       row_lo, stride_lo, stride_hi, dst, ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_commit_mac(Context *ctx, const double *A_sm, size_t kappa, size_t mA,
                      const double *W_sm, size_t r, size_t m, uint64_t *out) {
    Mod M{ctx->qd, ctx->qinv};
    constexpr int JT = 4, KT = 4;
    const size_t tiles = ((kappa + JT - 1) / JT) * ((r + KT - 1) / KT);
    // split K so the launch has at least ~2048 blocks, each >= 1024 i values
    size_t ksplit = 1;
    while (tiles * 64 * ksplit < 2048 && m / (ksplit * 2) >= 1024) ksplit *= 2;
    size_t chunk = (m + ksplit - 1) / ksplit;
    SALSAA_CHECK(tiles < (1u << 31), "commit tile count overflow");
    double *partial = nullptr;
    SALSAA_TRY(scratch_get(ctx, 7, ksplit * 64 * kappa * r * sizeof(double), (void **)&partial));
    dim3 grid((unsigned)tiles, 64, (unsigned)ksplit);
    hipLaunchKernelGGL((k_commit_mac<JT, KT>), grid, dim3(CM_THREADS), 0, ctx->stream, A_sm,
                       (int)kappa, mA, W_sm, (int)r, m, chunk, partial, M);
    SALSAA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_commit_finalize, dim3(ceil_div(kappa * r, 64)), dim3(64), 0, ctx->stream,
                       partial, (int)ksplit, (int)kappa, (int)r, out, ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

}  // namespace salsaa
