// mfma_commit.hip -- the commit's per-slot products on the int8 matrix cores.
//
// In the full NTT domain the commit is, for every slot s, C_s = A_s . W_s over
// F_q with A_s (kappa x m) and W_s (m x r) holding 50-bit residues.  Each
// residue x is written as 7 signed base-256 digits x = sum_u d_u 256^u,
// d_u in [-128, 127]; digit row r*7 + u of a matrix holds digit u of value
// row r.  Then
//     C_s[j][k] = sum_{u,v} 256^(u+v) * (A_s,u . W_s,v)[j][k],
// and the 49 digit products are ONE integer GEMM per slot with M = 7 kappa,
// N = 7 r, K = m on v_mfma_i32_16x16x64_i8: every digit product is at most
// 2^14 in magnitude, so int32 accumulation over a K chunk of at most 2^16 is
// exact.
//
// Operands are fragment-tiled (store_planes in ntt8.hip): the 16 digit rows x
// 64 columns one MFMA reads are one contiguous row-major KiB, so a wave moves
// a whole fragment with one 16-byte load per lane, and the row tiles of one K
// step are consecutive, so a block's loads form one sequential stream per
// operand.  In LDS a fragment is kept in MFMA lane order: slot 16 g + r holds
// row r, columns 16 g .. 16 g + 15, and the staging lane that writes slot l
// is the one that loaded it (the row/column transposition happens in the
// global load's lane-to-address map, within the fragment's 1 KiB), so both
// the staging writes and the operand reads are lane-contiguous, the pattern
// gfx950's LDS serves without bank conflicts.  A block owns one
// slot, one K chunk, all of the key's digit rows (at most 16 row tiles) and
// 16 witness rows (7 column tiles).  Per K step its threads stage the step's
// fragments from HBM through registers into LDS two steps ahead (double
// buffered), and every wave multiplies its own register tile of up to 4 x 4
// fragments out of LDS; HBM sees each fragment once per block.  At the end the
// int32 tiles of each group of 16 key rows go through LDS and each (j, k) is
// recombined with the weights 2^(8t) mod q in exact fp64 into the per-slot
// partial sums that k_commit_finalize reduces over K chunks.
#include "common.hpp"

namespace salsaa {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int FRAG = 1024;  // bytes of one 16 x 64 int8 fragment
constexpr int GT = 7;       // tiles per recombination group (16 rows of values)

// Memory chunk (16 bytes, row-major 4 r + g) that fills LDS slot l = 16 g + r.
__device__ __forceinline__ int slot_chunk(int l) { return 4 * (l & 15) + (l >> 4); }

// TM: row tiles per wave (wave wm owns key tiles wm + 4a, a < TM);
// WN x TN: witness tiles (wave wn owns tiles wn + WN b of the block's 7).
template <int TM, int TN, int WN>
struct CommitShape {
    static constexpr int NW = 4 * WN, NTHR = 64 * NW;
    static constexpr int SA = 4 * TM, SB = WN * TN;          // fragment slots per stage
    static constexpr int STAGE = (SA + SB) * FRAG;
    static constexpr int CPT = (SA + (SB < GT ? SB : GT) + NW - 1) / NW;  // 16-B chunks per thread
    static constexpr int DR = (SA < GT ? SA : GT) * 16;      // recombination rows
    static constexpr int DC = (SB < GT ? SB : GT) * 16 + 1;  // recombination row length
    static constexpr int LDS_BYTES = 2 * STAGE > DR * DC * 4 ? 2 * STAGE : DR * DC * 4;
};

// One block's work: slot s, K tiles [ktb, ktb + nsteps), all key rows
// j < kappa and the nbv witness tiles at WB (rows k < rows of the block).
// Key fragment (s, kt, t) sits at AB + s a_slot + (kt a_kt + t) KiB, witness
// fragment (s, kt, t) at WB + s w_slot + (kt w_kt + t) KiB.  Writes the
// recombined residues to part[j * prow + k].
template <int TM, int TN, int WN>
__device__ __forceinline__ void commit_body(char *lds, const int8_t *AB, size_t a_slot, int a_kt,
                                            int kappa, const int8_t *WB, size_t w_slot, int w_kt,
                                            int nbv, int rows, int s, int ktb, int nsteps,
                                            double *__restrict__ part, int prow,
                                            const double *__restrict__ tab, Mod M) {
    using S = CommitShape<TM, TN, WN>;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = w & 3, wn = w >> 2;
    const int rtA = (7 * kappa + 15) / 16;
    const int nchunk = (rtA + nbv) * 64;

    // staging: lane c = tid + NTHR i moves LDS slot c % 64 of fragment c / 64
    // (key tiles first, then the block's witness tiles); one wave moves one KiB.
    // Chunks past the block's tiles repeat the same-lane chunk of key tile 0
    // (the same bytes to the same LDS address), so loads and stores carry no
    // branches and the wait counts can keep two steps in flight.
    const int8_t *gp[S::CPT];
    int gs[S::CPT], lo[S::CPT];
#pragma unroll
    for (int i = 0; i < S::CPT; i++) {
        const int c0 = tid + S::NTHR * i;
        const int c = c0 < nchunk ? c0 : (c0 & 63);
        const int t = c >> 6, off = slot_chunk(c & 63) * 16;
        const int lslot = (c & 63) * 16;
        if (t < rtA) {
            gp[i] = AB + s * a_slot + ((size_t)ktb * a_kt + t) * FRAG + off;
            gs[i] = a_kt * FRAG;
            lo[i] = t * FRAG + lslot;
        } else {
            const int tb = t - rtA;
            gp[i] = WB + s * w_slot + ((size_t)ktb * w_kt + tb) * FRAG + off;
            gs[i] = w_kt * FRAG;
            lo[i] = (S::SA + tb) * FRAG + lslot;
        }
    }
    v4i R0[S::CPT], R1[S::CPT];
    // the compiler and scheduling barriers keep each step's loads together
    // and in program order, which is what lets the wait before a store skip
    // the newer step's loads (AB and WB are not __restrict__ so that the
    // compiler does not treat these loads as invariant and move them freely)
    auto gload = [&](v4i(&R)[S::CPT], int st) {
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < S::CPT; i++)
            R[i] = *reinterpret_cast<const v4i *>(gp[i] + (size_t)(unsigned)st * (unsigned)gs[i]);
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("" ::: "memory");
    };
    auto sstore = [&](const v4i(&R)[S::CPT], int buf) {
#pragma unroll
        for (int i = 0; i < S::CPT; i++)
            *reinterpret_cast<v4i *>(lds + buf * S::STAGE + lo[i]) = R[i];
    };
    v4i acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; a++)
#pragma unroll
        for (int b = 0; b < TN; b++) acc[a][b] = (v4i){0, 0, 0, 0};
    // fragment slots past the block's tiles hold stale bytes; they only feed
    // rows or columns of the product that are never recombined
    const int lane_slot = lane * 16;
    auto compute = [&](int buf) {
        const char *base = lds + buf * S::STAGE + lane_slot;
        v4i fb[TN];
#pragma unroll
        for (int b = 0; b < TN; b++)
            fb[b] = *reinterpret_cast<const v4i *>(base + (S::SA + wn + WN * b) * FRAG);
#pragma unroll
        for (int a = 0; a < TM; a++) {
            if (wm + 4 * a >= rtA) continue;  // wave-uniform: a key tile past kappa
            const v4i fa = *reinterpret_cast<const v4i *>(base + (wm + 4 * a) * FRAG);
#pragma unroll
            for (int b = 0; b < TN; b++)
                acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa, fb[b], acc[a][b], 0, 0, 0);
        }
    };
    // pipeline: step i is in LDS buffer i & 1 while the registers carry the
    // loads of steps i + 1 and i + 2 (nsteps is even and at least 2; loads
    // past the last step repeat it, and the last store into buffer 0 is
    // never read)
    const int last = nsteps - 1;
    gload(R0, 0);
    gload(R1, 1);
    sstore(R0, 0);
    gload(R0, min(2, last));
    __syncthreads();
    for (int i = 0; i < nsteps; i += 2) {
        compute(0);
        sstore(R1, 1);
        gload(R1, min(i + 3, last));
        __syncthreads();
        compute(1);
        sstore(R0, 0);
        gload(R0, min(i + 4, last));
        __syncthreads();
    }

    // recombination, one group of 7 key tiles (16 key rows) at a time:
    // D[16 tl + 4 g + v][16 b' + lr] = acc[a][b][v] for the group's tiles
    int(*D)[S::DC] = reinterpret_cast<int(*)[S::DC]>(lds);
    const int lr = lane & 15, g = lane >> 4;
    const double *wt = tab + TAB_W256;
    const int ngrp = (rtA + GT - 1) / GT;
    for (int gi = 0; gi < ngrp; gi++) {
        if (gi) __syncthreads();
#pragma unroll
        for (int a = 0; a < TM; a++) {
            const int tl = wm + 4 * a - GT * gi;
            if (tl < 0 || tl >= GT || 16 * tl >= S::DR) continue;
#pragma unroll
            for (int b = 0; b < TN; b++) {
                const int bl = wn + WN * b;
                if (bl >= GT) continue;
#pragma unroll
                for (int v = 0; v < 4; v++) D[16 * tl + 4 * g + v][16 * bl + lr] = acc[a][b][v];
            }
        }
        __syncthreads();
        for (int o = tid; o < 256; o += S::NTHR) {
            const int jl = o >> 4, kl = o & 15;
            const int j = 16 * gi + jl, k = kl;
            if (j >= kappa || k >= rows) continue;
            // diagonals E_t = sum_{u+v=t} D[7jl+u][7kl+v]: at most 7 int32
            // terms, below 2^34 < q in magnitude and exact in fp64
            long long E[13];
#pragma unroll
            for (int t = 0; t < 13; t++) E[t] = 0;
#pragma unroll
            for (int u = 0; u < 7; u++)
#pragma unroll
                for (int v = 0; v < 7; v++) E[u + v] += D[jl * 7 + u][kl * 7 + v];
            double x = 0.0;
#pragma unroll
            for (int t = 0; t < 13; t++) {
                x += mul_lazy((double)E[t], wt[t], M);
                if ((t % 7) == 6) x = reduce_lazy(x, M);
            }
            part[(size_t)j * prow + k] = reduce(x, M);
        }
    }
}

// XCD-aware block order: the 8 XCDs take consecutive runs of logical
// blocks, so blocks that share one slot's key chunk share an L2
__device__ __forceinline__ unsigned logical_block() {
    const unsigned nblk = gridDim.x, pid = blockIdx.x;
    return (nblk % 8 == 0) ? (pid % 8) * (nblk / 8) + pid / 8 : pid;
}

// One commit: W planes with rtB = ceil(7 r / 16) tiles per K step; block
// (nb, s, z) owns witness rows 16 nb .. 16 nb + 15, slot s < nslot and K
// chunk z.
template <int TM, int TN, int WN>
__global__ __launch_bounds__(256 * WN) __attribute__((amdgpu_waves_per_eu(4))) void k_commit_mfma(
    const int8_t *AB, int kappa, int a_kt, int a_ktiles, const int8_t *WB, int r, int ktW,
    int kchunk, int ktend, int nbx, int nslot, double *__restrict__ partial,
    const double *__restrict__ tab, Mod M) {
    __shared__ __align__(16) char lds[CommitShape<TM, TN, WN>::LDS_BYTES];
    const unsigned lid = logical_block();
    const int nb = (int)(lid % (unsigned)nbx);
    const unsigned rest = lid / (unsigned)nbx;
    const int s = (int)(rest % (unsigned)nslot);
    const int z = (int)(rest / (unsigned)nslot);
    const int rtB = (7 * r + 15) / 16;
    const int ktb = z * kchunk;
    commit_body<TM, TN, WN>(lds, AB, (size_t)a_ktiles * a_kt * FRAG, a_kt, kappa,
                            WB + (size_t)GT * nb * FRAG, (size_t)ktW * rtB * FRAG, rtB,
                            min(GT, rtB - GT * nb), min(16, r - 16 * nb), s, ktb,
                            min(ktend, ktb + kchunk) - ktb,
                            partial + ((size_t)z * nslot + s) * kappa * r + 16 * nb, r, tab, M);
}

// A batch of commits against one key: job j is a block of at most 16 witness
// rows with its own planes and K extent; table[b] = (job, z * 64 + s) in
// logical block order.
template <int TM, int TN, int WN>
__global__ __launch_bounds__(256 * WN) __attribute__((amdgpu_waves_per_eu(4))) void k_commit_mfma_batch(
    const int8_t *AB, int kappa, int a_kt, int a_ktiles, const CommitJob *__restrict__ jobs,
    const uint2 *__restrict__ table, int kchunk, const double *__restrict__ tab, Mod M) {
    __shared__ __align__(16) char lds[CommitShape<TM, TN, WN>::LDS_BYTES];
    const uint2 e = table[logical_block()];
    const CommitJob jb = jobs[e.x];
    const int s = (int)(e.y & 63u), z = (int)(e.y >> 6);
    const int ktb = z * kchunk;
    commit_body<TM, TN, WN>(lds, AB, (size_t)a_ktiles * a_kt * FRAG, a_kt, kappa, jb.WB,
                            (size_t)jb.ktiles * jb.rtiles * FRAG, jb.rtiles, jb.ntile, jb.rows, s,
                            ktb, min(jb.ktiles, ktb + kchunk) - ktb,
                            jb.partial + ((size_t)z * 64 + s) * kappa * jb.rows, jb.rows, tab, M);
}

// The same over jobs of several key row classes in one launch: the blocks
// that share one (slot, K chunk) of the key, whatever their jobs' kappa, are
// consecutive in the table and so run side by side on one XCD, which reads
// that key chunk from HBM once for all of them (a launch per class reads it
// once per class).  Every block runs the shape of TM key tiles per wave (the
// largest class's) with its job's kappa: it loads only its job's key tiles
// and skips the products of the others.
template <int TM, int WN>
__global__ __launch_bounds__(256 * WN) __attribute__((amdgpu_waves_per_eu(4))) void k_commit_mfma_mixed(
    const int8_t *AB, int a_kt, int a_ktiles, const CommitJob *__restrict__ jobs,
    const uint2 *__restrict__ table, int kchunk, const double *__restrict__ tab, Mod M) {
    __shared__ __align__(16) char lds[CommitShape<TM, 4, WN>::LDS_BYTES];
    const uint2 e = table[logical_block()];
    const CommitJob jb = jobs[e.x];
    const int s = (int)(e.y & 63u), z = (int)(e.y >> 6);
    const int ktb = z * kchunk;
    commit_body<TM, 4, WN>(lds, AB, (size_t)a_ktiles * a_kt * FRAG, a_kt, jb.kappa, jb.WB,
                           (size_t)jb.ktiles * jb.rtiles * FRAG, jb.rtiles, jb.ntile, jb.rows, s,
                           ktb, min(jb.ktiles, ktb + kchunk) - ktb,
                           jb.partial + ((size_t)z * 64 + s) * jb.kappa * jb.rows, jb.rows, tab, M);
}

template <int TM, int TN, int WN, bool BATCH>
int occupancy() {
    static int occ = 0;
    if (occ == 0) {
        int n = 0;
        const void *fn = BATCH ? reinterpret_cast<const void *>(&k_commit_mfma_batch<TM, TN, WN>)
                               : reinterpret_cast<const void *>(&k_commit_mfma<TM, TN, WN>);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 256 * WN, 0) != hipSuccess || n < 1)
            n = 1;
        occ = n;
    }
    return occ;
}

// nslot slots; wide_d = 0: the 64 full-NTT slots of d = 64 (finalized into
// IncompleteNTT form), else the slot pairs of degree wide_d (kappa and r
// count the doubled rows, see ring_wide.hip)
template <int TM, int TN, int WN>
int run(Context *ctx, const int8_t *AB, size_t kappa, size_t a_kt, size_t a_ktiles,
        const int8_t *WB, size_t r, size_t ktW, size_t ktend, uint64_t *out, size_t nslot,
        int wide_d) {
    Mod M{ctx->qd, ctx->qinv};
    const size_t nbx = (r + 15) / 16;
    // K chunks: at most 1024 tiles (2^16 columns, exact int32 sums), an even
    // number of tiles; pick the split that minimises (block waves x chunk
    // length + a fixed per-block cost) for this variant's occupancy
    const size_t slots = (size_t)ctx->ncu * occupancy<TM, TN, WN, false>();
    size_t best_ks = 0, best_kc = 0;
    double best = 0;
    for (size_t ks = (ktend + 1023) / 1024; ks <= ktend / 2; ks *= 2) {
        size_t kc = (ktend + ks - 1) / ks;
        kc += kc & 1;
        const size_t ksn = (ktend + kc - 1) / kc;
        const size_t nblk = nbx * nslot * ksn;
        const double cost = (double)((nblk + slots - 1) / slots) * (double)(kc + 8);
        if (best_ks == 0 || cost < best) {
            best = cost;
            best_ks = ksn;
            best_kc = kc;
        }
    }
    double *partial = nullptr;
    SALSAA_TRY(scratch_get(ctx, 7, best_ks * nslot * kappa * r * sizeof(double),
                           (void **)&partial));
    const size_t nblk = nbx * nslot * best_ks;
    SALSAA_CHECK(nblk < (1u << 31), "commit grid too large");
    hipLaunchKernelGGL((k_commit_mfma<TM, TN, WN>), dim3((unsigned)nblk), dim3(256 * WN), 0,
                       ctx->stream, AB, (int)kappa, (int)a_kt, (int)a_ktiles, WB, (int)r, (int)ktW,
                       (int)best_kc, (int)ktend, (int)nbx, (int)nslot, partial, ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    if (wide_d) return launch_commit_finalize_pairs(ctx, partial, best_ks, kappa, r, out);
    return launch_commit_finalize(ctx, partial, best_ks, kappa, r, out);
}

// ---- commit against the power-series CRS generated in the kernel
//
// The key row j of the CRS is m_j, m_j^2, ...: in slot s of the full NTT its
// element i is mhat_j[s]^(i + 1).  k_commit_crs is k_commit_mfma with the key's
// fragments made in LDS instead of read: per K step (64 columns), thread
// (j, e4) forms the four values base_j * mhat_j^e (e = 4 e4 .. 4 e4 + 3, the
// powers e < 64 from a per-block table), splits them into digits exactly as
// the stored planes do (byte u of (v + 0x80..80) ^ 0x80..80) and writes digit
// u of the four into row 7 j + u of the fragment tiles; base_j then advances
// by mhat_j^64.  The witness fragments stream from their planes one K step
// ahead in registers; single LDS stage, two block barriers per K step.
__device__ __forceinline__ double crs_powd(double b, unsigned long long e, Mod M) {
    double r = 1.0;
    while (e) {
        if (e & 1) r = mulmod(r, b, M);
        b = mulmod(b, b, M);
        e >>= 1;
    }
    return r;
}

// One block's work: slot s, K tiles [ktb, ktb + nsteps), all key rows j <
// kappa and the nbv witness tiles (rows k < rows) whose fragment (s, kt, t)
// sits at WB + s w_slot + (kt w_kt + t) KiB; the recombined residues go to
// part[j * prow + k].
template <int TM, int TN, int WN>
__device__ __forceinline__ void crs_body(char *lds, double (*tpow)[64], double *step64,
                                         const uint64_t *__restrict__ mhat, int kappa,
                                         const int8_t *WB, size_t w_slot, int w_kt, int nbv,
                                         int rows, int s, int ktb, int nsteps,
                                         double *__restrict__ part, int prow,
                                         const double *__restrict__ tab, Mod M) {
    using S = CommitShape<TM, TN, WN>;
    const int rtA = (7 * kappa + 15) / 16;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = w & 3, wn = w >> 2;
    for (int x = tid; x < kappa * 64; x += S::NTHR)
        tpow[x >> 6][x & 63] = crs_powd(u2d(mhat[(x >> 6) * kD + s]), (unsigned)(x & 63), M);
    for (int x = tid; x < kappa; x += S::NTHR) step64[x] = crs_powd(u2d(mhat[x * kD + s]), 64, M);
    // generator roles (j, e4), at most GPT per thread
    constexpr int GPT = (int)((kMfmaMaxKappa * 16 + S::NTHR - 1) / S::NTHR);
    double base[GPT];
#pragma unroll
    for (int gi = 0; gi < GPT; gi++) {
        const int x = tid + S::NTHR * gi;
        base[gi] = x < kappa * 16
                       ? crs_powd(u2d(mhat[(x >> 4) * kD + s]), 64ull * (unsigned)ktb + 1, M)
                       : 0.0;
    }
    // witness staging: chunk c = tid + NTHR i (i < CW) of the block's nbv
    // tiles; LDS slot c % 64 of tile c / 64 takes memory chunk slot_chunk(.)
    constexpr int CW = (GT * 64 + S::NTHR - 1) / S::NTHR;
    const int8_t *wbase = WB + (size_t)s * w_slot;
    v4i RW[CW];
    auto wload = [&](int st) {
#pragma unroll
        for (int i = 0; i < CW; i++) {
            const int c = tid + S::NTHR * i;
            if (c < nbv * 64)
                RW[i] = *reinterpret_cast<const v4i *>(
                    wbase + ((size_t)(ktb + st) * w_kt + (c >> 6)) * FRAG + slot_chunk(c & 63) * 16);
        }
    };
    auto wstore = [&]() {
#pragma unroll
        for (int i = 0; i < CW; i++) {
            const int c = tid + S::NTHR * i;
            if (c < nbv * 64)
                *reinterpret_cast<v4i *>(lds + (S::SA + (c >> 6)) * FRAG + (c & 63) * 16) = RW[i];
        }
    };
    v4i acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; a++)
#pragma unroll
        for (int b = 0; b < TN; b++) acc[a][b] = (v4i){0, 0, 0, 0};
    __syncthreads();  // the power tables
    wload(0);
    for (int st = 0; st < nsteps; st++) {
        wstore();
        wload(min(st + 1, nsteps - 1));
        // the key's digit rows for this K step
#pragma unroll
        for (int gi = 0; gi < GPT; gi++) {
            const int x = tid + S::NTHR * gi;
            if (x < kappa * 16) {
                const int j = x >> 4, e4 = x & 15;
                uint32_t lo[4], hi[4], dg[7];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint64_t y =
                        (d2u(mulmod(base[gi], tpow[j][4 * e4 + q], M)) + 0x0080808080808080ULL) ^
                        0x8080808080808080ULL;
                    lo[q] = (uint32_t)y;
                    hi[q] = (uint32_t)(y >> 32);
                }
                bytes_transpose4(lo, hi, dg);
#pragma unroll
                for (int u = 0; u < 7; u++) {
                    const int dr = 7 * j + u;
                    // tile dr / 16, LDS slot 16 g + dr % 16 (g = e4 / 4), bytes 4 (e4 % 4) ..
                    *reinterpret_cast<uint32_t *>(lds + (dr >> 4) * FRAG +
                                                  (16 * (e4 >> 2) + (dr & 15)) * 16 + 4 * (e4 & 3)) =
                        dg[u];
                }
                base[gi] = mulmod(base[gi], step64[j], M);
            }
        }
        __syncthreads();
        {
            const char *bb = lds + lane * 16;
            v4i fb[TN];
#pragma unroll
            for (int b = 0; b < TN; b++)
                fb[b] = *reinterpret_cast<const v4i *>(bb + (S::SA + wn + WN * b) * FRAG);
#pragma unroll
            for (int a = 0; a < TM; a++) {
                const v4i fa = *reinterpret_cast<const v4i *>(bb + (wm + 4 * a) * FRAG);
#pragma unroll
                for (int b = 0; b < TN; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa, fb[b], acc[a][b], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // recombination, as in commit_body
    int(*D)[S::DC] = reinterpret_cast<int(*)[S::DC]>(lds);
    const int lr = lane & 15, g = lane >> 4;
    const double *wt = tab + TAB_W256;
    const int ngrp = (rtA + GT - 1) / GT;
    for (int gi = 0; gi < ngrp; gi++) {
        if (gi) __syncthreads();
#pragma unroll
        for (int a = 0; a < TM; a++) {
            const int tl = wm + 4 * a - GT * gi;
            if (tl < 0 || tl >= GT || 16 * tl >= S::DR) continue;
#pragma unroll
            for (int b = 0; b < TN; b++) {
                const int bl = wn + WN * b;
                if (bl >= GT) continue;
#pragma unroll
                for (int v = 0; v < 4; v++) D[16 * tl + 4 * g + v][16 * bl + lr] = acc[a][b][v];
            }
        }
        __syncthreads();
        for (int o = tid; o < 256; o += S::NTHR) {
            const int jl = o >> 4, kl = o & 15;
            const int j = 16 * gi + jl, k = kl;
            if (j >= kappa || k >= rows) continue;
            long long E[13];
#pragma unroll
            for (int t = 0; t < 13; t++) E[t] = 0;
#pragma unroll
            for (int u = 0; u < 7; u++)
#pragma unroll
                for (int v = 0; v < 7; v++) E[u + v] += D[jl * 7 + u][kl * 7 + v];
            double x = 0.0;
#pragma unroll
            for (int t = 0; t < 13; t++) {
                x += mul_lazy((double)E[t], wt[t], M);
                if ((t % 7) == 6) x = reduce_lazy(x, M);
            }
            part[(size_t)j * prow + k] = reduce(x, M);
        }
    }
}

template <int TM, int TN, int WN>
__global__ __launch_bounds__(256 * WN) __attribute__((amdgpu_waves_per_eu(4))) void k_commit_crs(
    const uint64_t *__restrict__ mhat, int kappa, const int8_t *WB, int r, int ktW, int kchunk,
    int ktend, int nbx, double *__restrict__ partial, const double *__restrict__ tab, Mod M) {
    __shared__ __align__(16) char lds[CommitShape<TM, TN, WN>::LDS_BYTES];
    __shared__ double tpow[kMfmaMaxKappa][64];
    __shared__ double step64[kMfmaMaxKappa];
    const unsigned lid = logical_block();
    const int nb = (int)(lid % (unsigned)nbx);
    const unsigned rest = lid / (unsigned)nbx;
    const int s = (int)(rest % 64u);
    const int z = (int)(rest / 64u);
    const int rtB = (7 * r + 15) / 16;
    const int ktb = z * kchunk;
    crs_body<TM, TN, WN>(lds, tpow, step64, mhat, kappa, WB + (size_t)GT * nb * FRAG,
                         (size_t)ktW * rtB * FRAG, rtB, min(GT, rtB - GT * nb),
                         min(16, r - 16 * nb), s, ktb, min(ktend, ktb + kchunk) - ktb,
                         partial + ((size_t)z * 64 + s) * kappa * r + 16 * nb, r, tab, M);
}

// A batch of commits against the generated key: the job tables of
// k_commit_mfma_batch
template <int TM, int TN, int WN>
__global__ __launch_bounds__(256 * WN) __attribute__((amdgpu_waves_per_eu(4))) void k_commit_crs_batch(
    const uint64_t *__restrict__ mhat, int kappa, const CommitJob *__restrict__ jobs,
    const uint2 *__restrict__ table, int kchunk, const double *__restrict__ tab, Mod M) {
    __shared__ __align__(16) char lds[CommitShape<TM, TN, WN>::LDS_BYTES];
    __shared__ double tpow[kMfmaMaxKappa][64];
    __shared__ double step64[kMfmaMaxKappa];
    const uint2 e = table[logical_block()];
    const CommitJob jb = jobs[e.x];
    const int s = (int)(e.y & 63u), z = (int)(e.y >> 6);
    const int ktb = z * kchunk;
    crs_body<TM, TN, WN>(lds, tpow, step64, mhat, kappa, jb.WB,
                         (size_t)jb.ktiles * jb.rtiles * FRAG, jb.rtiles, jb.ntile, jb.rows, s, ktb,
                         min(jb.ktiles, ktb + kchunk) - ktb,
                         jb.partial + ((size_t)z * 64 + s) * kappa * jb.rows, jb.rows, tab, M);
}

template <int TM, int TN, int WN>
int run_crs_batch(Context *ctx, const uint64_t *mhat, size_t kappa, const CommitJob *jobs,
                  const uint2 *table, size_t nblk, size_t kchunk) {
    Mod M{ctx->qd, ctx->qinv};
    hipLaunchKernelGGL((k_commit_crs_batch<TM, TN, WN>), dim3((unsigned)nblk), dim3(256 * WN), 0,
                       ctx->stream, mhat, (int)kappa, jobs, table, (int)kchunk, ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

template <int TM>
int run_crs_batch_n(Context *ctx, const uint64_t *mhat, size_t kappa, const CommitJob *jobs,
                    const uint2 *table, size_t nblk, size_t kchunk, size_t max_tiles) {
    if (max_tiles <= 2) return run_crs_batch<TM, 1, 2>(ctx, mhat, kappa, jobs, table, nblk, kchunk);
    if (max_tiles <= 4) return run_crs_batch<TM, 2, 2>(ctx, mhat, kappa, jobs, table, nblk, kchunk);
    return run_crs_batch<TM, 4, 2>(ctx, mhat, kappa, jobs, table, nblk, kchunk);
}

template <int TM, int TN, int WN>
int run_crs(Context *ctx, const uint64_t *mhat, size_t kappa, const int8_t *WB, size_t r,
            size_t ktW, size_t ktend, uint64_t *out) {
    Mod M{ctx->qd, ctx->qinv};
    const size_t nbx = (r + 15) / 16;
    // K chunks of at most 1024 tiles (exact int32 sums): enough blocks to fill
    // the device several times over
    size_t kc = std::min<size_t>(1024, std::max<size_t>(2, ktend));
    while (kc > 16 && nbx * 64 * ((ktend + kc - 1) / kc) < (size_t)ctx->ncu * 8) kc /= 2;
    const size_t ks = (ktend + kc - 1) / kc;
    double *partial = nullptr;
    SALSAA_TRY(scratch_get(ctx, 7, ks * 64 * kappa * r * sizeof(double), (void **)&partial));
    const size_t nblk = nbx * 64 * ks;
    SALSAA_CHECK(nblk < (1u << 31), "commit grid too large");
    hipLaunchKernelGGL((k_commit_crs<TM, TN, WN>), dim3((unsigned)nblk), dim3(256 * WN), 0,
                       ctx->stream, mhat, (int)kappa, WB, (int)r, (int)ktW, (int)kc, (int)ktend,
                       (int)nbx, partial, ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return launch_commit_finalize(ctx, partial, ks, kappa, r, out);
}

template <int TM>
int run_crs_n(Context *ctx, const uint64_t *mhat, size_t kappa, const int8_t *WB, size_t r,
              size_t ktW, size_t ktend, uint64_t *out) {
    const size_t rtB = (7 * r + 15) / 16;
    if (rtB <= 2) return run_crs<TM, 1, 2>(ctx, mhat, kappa, WB, r, ktW, ktend, out);
    if (rtB <= 4) return run_crs<TM, 2, 2>(ctx, mhat, kappa, WB, r, ktW, ktend, out);
    return run_crs<TM, 4, 2>(ctx, mhat, kappa, WB, r, ktW, ktend, out);
}

template <int TM>
int run_n(Context *ctx, const int8_t *AB, size_t kappa, size_t a_kt, size_t a_ktiles,
          const int8_t *WB, size_t r, size_t ktW, size_t ktend, uint64_t *out, size_t nslot,
          int wide_d) {
    const size_t rtB = (7 * r + 15) / 16;
    if (rtB <= 2)
        return run<TM, 1, 2>(ctx, AB, kappa, a_kt, a_ktiles, WB, r, ktW, ktend, out, nslot,
                             wide_d);
    if (rtB <= 4)
        return run<TM, 2, 2>(ctx, AB, kappa, a_kt, a_ktiles, WB, r, ktW, ktend, out, nslot,
                             wide_d);
    return run<TM, 4, 2>(ctx, AB, kappa, a_kt, a_ktiles, WB, r, ktW, ktend, out, nslot, wide_d);
}

template <int TM, int TN, int WN>
int run_batch(Context *ctx, const int8_t *AB, size_t kappa, size_t a_kt, size_t a_ktiles,
              const CommitJob *jobs, const uint2 *table, size_t nblk, size_t kchunk) {
    Mod M{ctx->qd, ctx->qinv};
    hipLaunchKernelGGL((k_commit_mfma_batch<TM, TN, WN>), dim3((unsigned)nblk), dim3(256 * WN), 0,
                       ctx->stream, AB, (int)kappa, (int)a_kt, (int)a_ktiles, jobs, table,
                       (int)kchunk, ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

template <int TM>
int run_batch_n(Context *ctx, const int8_t *AB, size_t kappa, size_t a_kt, size_t a_ktiles,
                const CommitJob *jobs, const uint2 *table, size_t nblk, size_t kchunk,
                size_t max_tiles) {
    if (max_tiles <= 2)
        return run_batch<TM, 1, 2>(ctx, AB, kappa, a_kt, a_ktiles, jobs, table, nblk, kchunk);
    if (max_tiles <= 4)
        return run_batch<TM, 2, 2>(ctx, AB, kappa, a_kt, a_ktiles, jobs, table, nblk, kchunk);
    return run_batch<TM, 4, 2>(ctx, AB, kappa, a_kt, a_ktiles, jobs, table, nblk, kchunk);
}

}  // namespace

static int commit_mfma_any(Context *ctx, const int8_t *AB, size_t kappa, size_t a_kt,
                           size_t a_ktiles, const int8_t *WB, size_t r, size_t m, size_t ktW,
                           uint64_t *out, size_t nslot, int wide_d) {
    const size_t ktend = plane_ktiles(m);
    SALSAA_CHECK(kappa >= 1 && kappa <= kMfmaMaxKappa, "MFMA commit needs kappa <= 36");
    SALSAA_CHECK(ktend == ktW && ktend <= a_ktiles, "plane K-tiles do not cover the witness");
    SALSAA_CHECK(a_ktiles < (1u << 24) && r < (1u << 20), "commit too large");
    const size_t rtA = (7 * kappa + 15) / 16;
    SALSAA_CHECK(rtA <= a_kt, "key planes hold fewer rows than kappa");
    if (rtA <= 4) return run_n<1>(ctx, AB, kappa, a_kt, a_ktiles, WB, r, ktW, ktend, out, nslot, wide_d);
    if (rtA <= 8) return run_n<2>(ctx, AB, kappa, a_kt, a_ktiles, WB, r, ktW, ktend, out, nslot, wide_d);
    return run_n<4>(ctx, AB, kappa, a_kt, a_ktiles, WB, r, ktW, ktend, out, nslot, wide_d);
}

int launch_commit_mfma(Context *ctx, const int8_t *AB, size_t kappa, size_t a_kt,
                       size_t a_ktiles, const int8_t *WB, size_t r, size_t m, size_t ktW,
                       uint64_t *out) {
    return commit_mfma_any(ctx, AB, kappa, a_kt, a_ktiles, WB, r, m, ktW, out, 64, 0);
}

int launch_commit_mfma_pairs(Context *ctx, const int8_t *AB, size_t kappa2, size_t a_kt,
                             size_t a_ktiles, const int8_t *WB, size_t r2, size_t m, size_t ktW,
                             uint64_t *out) {
    SALSAA_CHECK(ctx->d > kD, "slot-pair commit is for degrees above 64");
    return commit_mfma_any(ctx, AB, kappa2, a_kt, a_ktiles, WB, r2, m, ktW, out,
                           (size_t)ctx->d / 2, ctx->d);
}

int launch_commit_crs(Context *ctx, const uint64_t *mhat, size_t kappa, const int8_t *WB, size_t r,
                      size_t m, size_t ktW, uint64_t *out) {
    const size_t ktend = plane_ktiles(m);
    SALSAA_CHECK(kappa >= 1 && kappa <= kMfmaMaxKappa, "MFMA commit needs kappa <= 36");
    SALSAA_CHECK(ktend == ktW, "plane K-tiles do not cover the witness");
    SALSAA_CHECK(ktend < (1u << 24) && r < (1u << 20), "commit too large");
    const size_t rtA = (7 * kappa + 15) / 16;
    if (rtA <= 4) return run_crs_n<1>(ctx, mhat, kappa, WB, r, ktW, ktend, out);
    if (rtA <= 8) return run_crs_n<2>(ctx, mhat, kappa, WB, r, ktW, ktend, out);
    return run_crs_n<4>(ctx, mhat, kappa, WB, r, ktW, ktend, out);
}

int launch_commit_crs_batch(Context *ctx, const uint64_t *mhat, size_t kappa, const CommitJob *jobs,
                            const uint2 *table, size_t nblk, size_t kchunk, size_t max_tiles) {
    SALSAA_CHECK(kappa >= 1 && kappa <= kMfmaMaxKappa, "MFMA commit needs kappa <= 36");
    SALSAA_CHECK(max_tiles >= 1 && max_tiles <= 7, "a batch job holds at most 16 witness rows");
    SALSAA_CHECK(kchunk >= 2 && kchunk <= 1024 && kchunk % 2 == 0, "K chunk must be even, <= 1024");
    SALSAA_CHECK(nblk >= 1 && nblk < (1u << 31), "commit grid size out of range");
    const size_t rtA = (7 * kappa + 15) / 16;
    if (rtA <= 4) return run_crs_batch_n<1>(ctx, mhat, kappa, jobs, table, nblk, kchunk, max_tiles);
    if (rtA <= 8) return run_crs_batch_n<2>(ctx, mhat, kappa, jobs, table, nblk, kchunk, max_tiles);
    return run_crs_batch_n<4>(ctx, mhat, kappa, jobs, table, nblk, kchunk, max_tiles);
}

int launch_commit_mfma_batch(Context *ctx, const int8_t *AB, size_t kappa, size_t a_kt,
                             size_t a_ktiles, const CommitJob *jobs, const uint2 *table,
                             size_t nblk, size_t kchunk, size_t max_tiles) {
    SALSAA_CHECK(kappa >= 1 && kappa <= kMfmaMaxKappa, "MFMA commit needs kappa <= 36");
    SALSAA_CHECK(max_tiles >= 1 && max_tiles <= 7, "a batch job holds at most 16 witness rows");
    SALSAA_CHECK(kchunk >= 2 && kchunk <= 1024 && kchunk % 2 == 0, "K chunk must be even, <= 1024");
    SALSAA_CHECK(nblk >= 1 && nblk < (1u << 31), "commit grid size out of range");
    const size_t rtA = (7 * kappa + 15) / 16;
    SALSAA_CHECK(rtA <= a_kt, "key planes hold fewer rows than kappa");
    if (rtA <= 4)
        return run_batch_n<1>(ctx, AB, kappa, a_kt, a_ktiles, jobs, table, nblk, kchunk, max_tiles);
    if (rtA <= 8)
        return run_batch_n<2>(ctx, AB, kappa, a_kt, a_ktiles, jobs, table, nblk, kchunk, max_tiles);
    return run_batch_n<4>(ctx, AB, kappa, a_kt, a_ktiles, jobs, table, nblk, kchunk, max_tiles);
}

int launch_commit_mfma_mixed(Context *ctx, const int8_t *AB, size_t a_kt, size_t a_ktiles,
                             const CommitJob *jobs, const uint2 *table, size_t nblk,
                             size_t kchunk, int tm) {
    SALSAA_CHECK(a_kt >= 1 && a_kt <= 16, "key planes hold 1..16 row tiles per K step");
    SALSAA_CHECK(kchunk >= 2 && kchunk <= 1024 && kchunk % 2 == 0, "K chunk must be even, <= 1024");
    SALSAA_CHECK(nblk >= 1 && nblk < (1u << 31), "commit grid size out of range");
    SALSAA_CHECK(tm == 2 || tm == 4, "mixed commit shape must be 2 or 4 key tiles per wave");
    Mod M{ctx->qd, ctx->qinv};
    if (tm == 4)
        hipLaunchKernelGGL((k_commit_mfma_mixed<4, 2>), dim3((unsigned)nblk), dim3(512), 0,
                           ctx->stream, AB, (int)a_kt, (int)a_ktiles, jobs, table, (int)kchunk,
                           ctx->d_tables, M);
    else
        hipLaunchKernelGGL((k_commit_mfma_mixed<2, 2>), dim3((unsigned)nblk), dim3(512), 0,
                           ctx->stream, AB, (int)a_kt, (int)a_ktiles, jobs, table, (int)kchunk,
                           ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int commit_batch_occupancy(size_t kappa, size_t max_tiles) {
    const size_t rtA = (7 * kappa + 15) / 16;
    const int tm = rtA <= 4 ? 1 : rtA <= 8 ? 2 : 4;
    const int tn = max_tiles <= 2 ? 1 : max_tiles <= 4 ? 2 : 4;
#define SALSAA_OCC(TM)                                                                          \
    if (tm == TM) {                                                                              \
        if (tn == 1) return occupancy<TM, 1, 2, true>();                                         \
        if (tn == 2) return occupancy<TM, 2, 2, true>();                                         \
        return occupancy<TM, 4, 2, true>();                                                      \
    }
    SALSAA_OCC(1)
    SALSAA_OCC(2)
    SALSAA_OCC(4)
#undef SALSAA_OCC
    return 1;
}

}  // namespace salsaa
