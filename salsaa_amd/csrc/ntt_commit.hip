// ntt_commit.hip -- the commit kernels: ring matrix product A . W in the full
// 64-slot NTT domain.
//
// out[j][k] = sum_i A_j[i] * W_k[i] in R_q.  In the full NTT domain this is 64
// independent (kappa x m) . (m x r) products over F_q, one per slot.  A (the
// resident key) and W are stored slot-major [64][rows][m] as exact doubles, so
// lanes walking i read contiguous memory.  Each wave keeps an 8 x 8 register
// tile of lazy accumulators (one reduction every 7 MACs); blocks split K and
// write per-slot partial sums; a finalize kernel sums the partials and maps
// the 64 slots to the IncompleteNTT form (the reference's output
// representation, arithmetic.rs:459-505, cyclotomic_ring.rs:789).
#include "ntt_device.hpp"

namespace salsaa {

namespace {

// ------------------------------------------------------------------ commit
constexpr int CM_THREADS = 256;
constexpr int CM_JT = 8, CM_KT = 8;  // per-wave register tile (kappa x r)

// grid: x = group of 4 tiles, y = slot, z = K split.  Wave w of block b owns
// tile t = 4b + w, enumerated kt-major (t = kt * jtiles + jt) so the 4 waves
// of a block share the same W rows (served from L1) and read distinct A rows.
// Each lane walks i = lane, lane + 64, ... of the block's K chunk keeping an
// 8 x 8 tile of lazy accumulators; the 64 lanes' partial sums are then
// combined by a butterfly reduce-scatter that leaves output (a, b) = lane in
// lane a * 8 + b.  partial[z][s][j][k] in [0, q).
__global__ __launch_bounds__(CM_THREADS) void k_commit_mac(
    const double *__restrict__ A_sm, int kappa, int a_rows, size_t mA, const double *__restrict__ W_sm,
    int r, size_t m, size_t chunk, double *__restrict__ partial, Mod M) {
    const int jtiles = (kappa + CM_JT - 1) / CM_JT, ktiles = (r + CM_KT - 1) / CM_KT;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = blockIdx.x * (CM_THREADS / 64) + wave;
    if (t >= jtiles * ktiles) return;  // whole wave idles; no block barriers below
    const int kt = t / jtiles, jt = t % jtiles;
    const int j0 = jt * CM_JT, k0 = kt * CM_KT;
    const int s = blockIdx.y;
    const size_t ib = (size_t)blockIdx.z * chunk;
    const size_t ie = min(m, ib + chunk);
    const double *Ab = A_sm + ((size_t)s * a_rows) * mA;
    const double *Wb = W_sm + ((size_t)s * r) * m;

    double acc[CM_JT * CM_KT];
#pragma unroll
    for (int x = 0; x < CM_JT * CM_KT; x++) acc[x] = 0.0;

    int cnt = 0;
    for (size_t i = ib + lane; i < ie; i += 64) {
        double av[CM_JT], wv[CM_KT];
#pragma unroll
        for (int a = 0; a < CM_JT; a++) av[a] = (j0 + a < kappa) ? Ab[(size_t)(j0 + a) * mA + i] : 0.0;
#pragma unroll
        for (int b = 0; b < CM_KT; b++) wv[b] = (k0 + b < r) ? Wb[(size_t)(k0 + b) * m + i] : 0.0;
#pragma unroll
        for (int a = 0; a < CM_JT; a++)
#pragma unroll
            for (int b = 0; b < CM_KT; b++) acc[a * CM_KT + b] += mul_lazy(av[a], wv[b], M);
        if (++cnt == 7) {
            cnt = 0;
#pragma unroll
            for (int x = 0; x < CM_JT * CM_KT; x++) acc[x] = reduce_lazy(acc[x], M);
        }
    }
#pragma unroll
    for (int x = 0; x < CM_JT * CM_KT; x++) acc[x] = reduce(acc[x], M);
    // butterfly reduce-scatter over the 64 lanes: at offset o the lane with
    // bit o set keeps the upper half of its live values
#pragma unroll
    for (int step = 0; step < 6; step++) {
        const int o = 32 >> step;
        const int half = o;  // live values before this step: 2 * o
        const bool up = (lane & o) != 0;
#pragma unroll
        for (int x = 0; x < half; x++) {
            const double lo = acc[x], hi = acc[x + half];
            const double keep = up ? hi : lo, send = up ? lo : hi;
            acc[x] = addmod(keep, __shfl_xor(send, o, 64), M);
        }
    }
    const int a = lane / CM_KT, b = lane % CM_KT;
    if (j0 + a < kappa && k0 + b < r)
        partial[(((size_t)blockIdx.z * 64 + s) * kappa + (j0 + a)) * r + (k0 + b)] = acc[0];
}

// One block of 64 threads per output ring element (j, k): thread s sums the
// K-split partials of slot s; thread c then forms IncompleteNTT coefficient c
// as sum_s fin[s][c] * v[s], fin being the linear map IncNTT o INTT_64.
// 64 threads per (j, k): thread s sums slot s's partials over the K chunks,
// then thread c forms coefficient c of the IncompleteNTT result through the
// precomputed slots -> IncompleteNTT map.  partial_jk points at (z = 0, s = 0)
// of this (j, k); consecutive slots are slot_stride apart, chunks
// chunk_stride apart.
__device__ __forceinline__ void finalize_one(const double *__restrict__ partial_jk, int ksplit,
                                             size_t slot_stride, size_t chunk_stride,
                                             uint64_t *__restrict__ out,
                                             const double *__restrict__ tab, Mod M) {
    __shared__ double vs[64];
    const int t = threadIdx.x;
    double x = 0.0;
    for (int z = 0; z < ksplit; z++)
        x = addmod(x, partial_jk[(size_t)z * chunk_stride + (size_t)t * slot_stride], M);
    vs[t] = x;
    __syncthreads();
    const double *fin = tab + TAB_FIN;
    double acc = 0.0;
#pragma unroll 8
    for (int s = 0; s < 64; s++) {
        acc += mul_lazy(fin[s * 64 + t], vs[s], M);
        if ((s & 3) == 3) acc = reduce_lazy(acc, M);
    }
    out[t] = d2u(reduce(acc, M));
}

__global__ __launch_bounds__(64) void k_commit_finalize(const double *__restrict__ partial,
                                                        int ksplit, int kappa, int r,
                                                        uint64_t *__restrict__ out,
                                                        const double *__restrict__ tab, Mod M) {
    const int jk = blockIdx.x;
    const size_t plane = (size_t)kappa * r;
    finalize_one(partial + jk, ksplit, plane, 64 * plane, out + (size_t)jk * kD, tab, M);
}

// batched: job of block b found by binary search over jobs[].blk0
__global__ __launch_bounds__(64) void k_commit_finalize_batch(const FinJob *__restrict__ jobs,
                                                              int njobs, int kappa,
                                                              const double *__restrict__ tab,
                                                              Mod M) {
    const int b = blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].blk0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const FinJob jb = jobs[lo];
    const int jk = b - jb.blk0, j = jk / jb.rows, k = jk % jb.rows;
    const size_t plane = (size_t)kappa * jb.rows;
    finalize_one(jb.partial + jk, jb.ks, plane, 64 * plane, jb.out + j * jb.osj + k * jb.osk,
                 tab, M);
}

}  // namespace

int launch_commit_mac(Context *ctx, const double *A_sm, size_t kappa, size_t a_rows, size_t mA,
                      const double *W_sm, size_t r, size_t m, uint64_t *out) {
    Mod M{ctx->qd, ctx->qinv};
    const size_t tiles = ((kappa + CM_JT - 1) / CM_JT) * ((r + CM_KT - 1) / CM_KT);
    const size_t groups = (tiles + 3) / 4;
    // split K so the launch has >= ~1024 blocks while each wave keeps >= 16
    // iterations (1024 values of i) to amortise its reduce-scatter
    size_t ksplit = 1;
    while (groups * 64 * ksplit < 1024 && m / (ksplit * 2) >= 1024) ksplit *= 2;
    size_t chunk = (m + ksplit - 1) / ksplit;
    SALSAA_CHECK(groups < (1u << 31), "commit tile count overflow");
    double *partial = nullptr;
    SALSAA_TRY(scratch_get(ctx, 7, ksplit * 64 * kappa * r * sizeof(double), (void **)&partial));
    dim3 grid((unsigned)groups, 64, (unsigned)ksplit);
    hipLaunchKernelGGL(k_commit_mac, grid, dim3(CM_THREADS), 0, ctx->stream, A_sm, (int)kappa,
                       (int)a_rows, mA,
                       W_sm, (int)r, m, chunk, partial, M);
    SALSAA_HIP(hipGetLastError());
    return launch_commit_finalize(ctx, partial, ksplit, kappa, r, out);
}

int launch_commit_finalize_batch(Context *ctx, const FinJob *jobs, size_t njobs, size_t nblk,
                                 size_t kappa) {
    SALSAA_CHECK(njobs >= 1 && nblk >= 1 && nblk < (1u << 31), "finalize grid size out of range");
    Mod M{ctx->qd, ctx->qinv};
    hipLaunchKernelGGL(k_commit_finalize_batch, dim3((unsigned)nblk), dim3(64), 0, ctx->stream,
                       jobs, (int)njobs, (int)kappa, ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_commit_finalize(Context *ctx, const double *partial, size_t ksplit, size_t kappa,
                           size_t r, uint64_t *out) {
    Mod M{ctx->qd, ctx->qinv};
    hipLaunchKernelGGL(k_commit_finalize, dim3((unsigned)(kappa * r)), dim3(64), 0, ctx->stream,
                       partial, (int)ksplit, (int)kappa, (int)r, out, ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

}  // namespace salsaa
