// fold_step.hip -- the batched folding step: decomposition, fold, re-commit
// and norm check of several accumulated instances.
//
// Instance i has a witness W_i of r rows x m ring elements (Coefficient).
//   decompose   W_i -> r * parts limb rows, limb t of row rho in row
//               rho * parts + t (decompose_matrix_by_chunks,
//               arithmetic.rs:1572-1593; radix per instance, auto or given)
//   fold        W'_i[j] = sum_k C_i[j][k] * limb_k  (fold, fold.rs:22-27 ->
//               parallel_dot_matrix_matrix, arithmetic.rs:429-457), r_out rows,
//               IncompleteNTT
//   commit      the folded witnesses of all instances stacked as one witness
//               (parallel_dot_series_matrix, arithmetic.rs:459-505)
//   norm check  norm_check_1/2/3 of every folded row, each instance with its
//               own challenges (fast_norm_check.rs:282-462)
// With ternary fold challenges (the reference's, cyclotomic_ring.rs:319-339)
// and limbs below 128, k_fold_mfma fuses the first two on the int8 matrix
// cores in the Coefficient domain: the limbs are formed from W in registers
// and never reach HBM, and k_fold_coeff_out forms the outputs below from its
// Coefficient rows.  Otherwise k_decomp_fold fuses them in the NTT
// domain: a block takes 16 columns of all r rows of one instance, forms each
// limb and its IncompleteNTT in LDS, and accumulates the fold for all r_out
// output rows in registers.  The same block then writes the fold output (IncompleteNTT), the
// commit's digit planes (full-NTT slots from the IncompleteNTT halves:
// a(zeta_s) = E[zj_s] + zeta_s O[zj_s]) and the Coefficient form the norm
// check reads.
#include "handles.hpp"
#include "ntt8.hpp"

using namespace salsaa;

namespace {

constexpr int FNT = 1024;    // threads per block
constexpr int FROW = 72;     // LDS row stride (u64), skewed layout
constexpr int FCOLS = 16;    // columns per block

// The column block of a (m / FCOLS) x n grid, remapped so that the four
// blocks whose 16-column plane stores share each 64-byte line run on one XCD
// (blocks are dealt to the 8 XCDs in turn): their partial lines then merge in
// that XCD's L2 before they reach HBM.
__device__ __forceinline__ size_t fold_col_block() {
    const unsigned gx = gridDim.x, bx = blockIdx.x;
    return gx % 8 == 0 ? (size_t)(bx % 8) * (gx / 8) + bx / 8 : bx;
}
constexpr uint64_t DIGIT_BIAS8 = 0x0080808080808080ULL;

__device__ __forceinline__ int sk(int k) { return k + (k >> 3); }

// signed digit t of x (arithmetic.rs:1496-1570)
__device__ __forceinline__ uint64_t limb_digit(uint64_t x, uint64_t q, unsigned lg, uint64_t mask,
                                               int tl) {
    const bool neg = 2 * x > q;
    const uint64_t v = neg ? q - x : x;
    const unsigned s = lg * (unsigned)tl;
    const uint64_t dg = (lg == 0) ? (v & mask) : (s >= 64 ? 0 : (v >> s) & mask);
    return neg ? (dg ? q - dg : 0) : dg;
}

// ---- the fold on the int8 matrix cores (Coefficient domain).  The fold
// challenges of the reference are ternary (CyclotomicRing::random_biased,
// cyclotomic_ring.rs:319-339) and the limbs of a radix <= 128 are int8, so
// per instance and column i the fold is one integer matrix product:
//     W'[j][a](i) = sum_k sum_b T(C[j][k])[a][b] L_k[b](i),
// T(c)[a][b] = c[a - b] for a >= b, -c[a - b + 64] otherwise (the negacyclic
// product as a matrix), |sum| <= 64 rk * 127 < 2^19: exact in int32, and the
// folded rows come out directly in Coefficient form.  An instance whose C is
// not ternary or whose radix exceeds 128 takes k_decomp_fold's NTT-domain
// fold instead (the two kernels select instances by the same flags).
typedef int v4i __attribute__((ext_vector_type(4)));
constexpr int FM_COLS = 128;  // columns per block
constexpr int FM_THREADS = 512;

__device__ __forceinline__ bool fold_mfma_ok(const unsigned *bad, const uint64_t *radix, int inst) {
    return bad[inst] == 0 && radix[inst] <= 128;
}

// T fragments: for instance inst, output row j, limb row k and row tile rt
// (a = 16 rt .. 16 rt + 15), lane (g, r) of the 16 x 16 x 64 fragment holds
// T(C[j][k])[16 rt + r][16 g + x] at byte x (x < 16).  grid (r_out rk, n),
// 256 threads (rt = tid >> 6, lane = tid & 63); bad[inst] is set when C is
// not ternary.
__global__ __launch_bounds__(256) void k_fold_prep(const uint64_t *__restrict__ C, int r_out, int rk,
                                                   uint64_t q, int8_t *__restrict__ tfrag,
                                                   unsigned *__restrict__ bad) {
    __shared__ int8_t c8[kD];
    const int jk = blockIdx.x, inst = blockIdx.y, tid = threadIdx.x;
    const uint64_t *cc = C + (((size_t)inst * r_out) * rk + jk) * kD;
    if (tid < kD) {
        const uint64_t x = cc[tid];
        const bool t = x == 0 || x == 1 || x == q - 1;
        if (!t) atomicOr(bad + inst, 1u);
        c8[tid] = x == 1 ? (int8_t)1 : (x == q - 1 ? (int8_t)-1 : (int8_t)0);
    }
    __syncthreads();
    const int rt = tid >> 6, lane = tid & 63, g = lane >> 4, rr = lane & 15;
    const int a = 16 * rt + rr;
    uint32_t w[4];
#pragma unroll
    for (int x4 = 0; x4 < 4; x4++) {
        uint32_t v = 0;
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const int b = 16 * g + 4 * x4 + y, dlt = a - b;
            const int8_t t = dlt >= 0 ? c8[dlt] : (int8_t)(-c8[dlt + kD]);
            v |= (uint32_t)(uint8_t)t << (8 * y);
        }
        w[x4] = v;
    }
    *reinterpret_cast<uint4 *>(tfrag + (((size_t)inst * r_out * rk + jk) * 4 + rt) * 1024 +
                               lane * 16) = make_uint4(w[0], w[1], w[2], w[3]);
}

// grid (m / 128, n), 512 threads: wave w takes output row j = w (waves past
// r_out only stage limbs) over the block's 128 columns, 4 row tiles x 8
// column tiles of 16 x 16 accumulators.  Per W row rho, each thread holds 16
// consecutive coefficients of one column (i = tid / 4, b = 16 (tid % 4) ..),
// and for each limb tl writes their digits to the LDS tile Ld[i][b] (the
// B fragments: lane (g, c) reads Ld[16 ct + c][16 g .. 16 g + 15]).
__global__ __launch_bounds__(FM_THREADS) void k_fold_mfma(
    const uint64_t *__restrict__ W, int r, size_t m, int parts, const uint64_t *__restrict__ d_radix,
    const unsigned *__restrict__ bad, const int8_t *__restrict__ tfrag, int r_out, uint64_t q,
    uint64_t *__restrict__ folded_coeff) {
    __shared__ __align__(16) int8_t Ld[FM_COLS * kD];
    const int inst = blockIdx.y;
    if (!fold_mfma_ok(bad, d_radix, inst)) return;  // uniform
    const size_t i0 = (size_t)blockIdx.x * FM_COLS;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, cl = lane & 15;
    const uint64_t radix = d_radix[inst];
    const unsigned lg = (unsigned)__ffsll((long long)radix) - 1u;
    const uint64_t mask = radix - 1;
    const int rk = r * parts;
    const bool act = wave < r_out;  // uniform per wave
    const int8_t *tj = tfrag + ((size_t)inst * r_out + (act ? wave : 0)) * rk * 4096;
    v4i acc[4][8];
#pragma unroll
    for (int rt = 0; rt < 4; rt++)
#pragma unroll
        for (int ct = 0; ct < 8; ct++) acc[rt][ct] = (v4i){0, 0, 0, 0};
    const int ci = tid >> 2, bq = tid & 3;  // staging: column ci, coefficients 16 bq ..
    // Software pipeline: the next limb's A fragments load one step ahead, and
    // the next W row's loads go out once the current row's last digits are
    // packed, so both wait behind this step's products.
    auto load_w = [&](uint64_t(&wv)[16], int rho) {
        const uint64_t *src = W + (((size_t)inst * r + rho) * m + i0 + ci) * kD + 16 * bq;
#pragma unroll
        for (int x = 0; x < 16; x += 2) {
            const ulonglong2 v = *reinterpret_cast<const ulonglong2 *>(src + x);
            wv[x] = v.x;
            wv[x + 1] = v.y;
        }
    };
    auto load_a = [&](v4i(&af)[4], int k) {
#pragma unroll
        for (int rt = 0; rt < 4; rt++)
            af[rt] = *reinterpret_cast<const v4i *>(tj + ((size_t)k * 4 + rt) * 1024 + lane * 16);
    };
    uint64_t wv[16];
    v4i af[4];
    load_w(wv, 0);
    if (act) load_a(af, 0);
    for (int rho = 0; rho < r; rho++) {
        for (int tl = 0; tl < parts; tl++) {
            const int k = rho * parts + tl;
            uint32_t pk[4];
#pragma unroll
            for (int x4 = 0; x4 < 4; x4++) {
                uint32_t v = 0;
#pragma unroll
                for (int y = 0; y < 4; y++) {
                    const uint64_t dg = limb_digit(wv[4 * x4 + y], q, lg, mask, tl);
                    const int sv = dg > (q >> 1) ? (int)((int64_t)dg - (int64_t)q) : (int)dg;
                    v |= (uint32_t)(uint8_t)(int8_t)sv << (8 * y);
                }
                pk[x4] = v;
            }
            if (tl == parts - 1 && rho + 1 < r) load_w(wv, rho + 1);
            v4i an[4];
            const bool more = k + 1 < rk;
            if (act && more) load_a(an, k + 1);
            __syncthreads();  // the previous limb's B fragments have been read
            *reinterpret_cast<uint4 *>(Ld + ci * kD + 16 * bq) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
            __syncthreads();
            if (act) {
#pragma unroll
                for (int ct = 0; ct < 8; ct++) {
                    const v4i bf = *reinterpret_cast<const v4i *>(Ld + (16 * ct + cl) * kD + 16 * g);
#pragma unroll
                    for (int rt = 0; rt < 4; rt++)
                        acc[rt][ct] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[rt], bf, acc[rt][ct], 0, 0, 0);
                }
                if (more) {
#pragma unroll
                    for (int rt = 0; rt < 4; rt++) af[rt] = an[rt];
                }
            }
        }
    }
    if (!act) return;
    // lane (g, c) of tile (rt, ct) holds coefficients a = 16 rt + 4 g .. + 3 of
    // column i = 16 ct + c: Coefficient form, canonical
    uint64_t *orow = folded_coeff + ((size_t)inst * r_out + wave) * m * kD;
#pragma unroll
    for (int ct = 0; ct < 8; ct++) {
        uint64_t *col = orow + (i0 + 16 * ct + cl) * kD;
#pragma unroll
        for (int rt = 0; rt < 4; rt++) {
            uint64_t o[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int v = acc[rt][ct][e];
                o[e] = v >= 0 ? (uint64_t)v : q - (uint64_t)(-(int64_t)v);
            }
            ulonglong2 *dst = reinterpret_cast<ulonglong2 *>(col + 16 * rt + 4 * g);
            dst[0] = make_ulonglong2(o[0], o[1]);
            dst[1] = make_ulonglong2(o[2], o[3]);
        }
    }
}

// The block's folded rows (IncompleteNTT, rows e' = jo * 16 + i of the
// skewed LDS layout) out: (a) the fold output, (b) the commit's digit planes
// of the full-NTT slots, and with COEFF_OUT (c) the Coefficient form.
template <bool COEFF_OUT>
__device__ __forceinline__ void fold_outputs(uint64_t *lds, int inst, size_t i0, size_t m, int r_out,
                                             uint64_t *__restrict__ folded,
                                             uint64_t *__restrict__ folded_coeff,
                                             int8_t *__restrict__ planes, size_t ktiles,
                                             size_t rt_n, const double *__restrict__ tab, Mod M,
                                             int tid) {
    const int e = tid >> 3, h = tid & 7, p = h >> 2, t = h & 3;
    const int nout = r_out * FCOLS;
    // (a) fold output, IncompleteNTT
    for (int c = tid; c < nout * 32; c += FNT) {
        const int eo = c >> 5, k2 = (c & 31) * 2, jo = eo >> 4, i = eo & 15;
        ulonglong2 v;
        v.x = d2u(bitsd(lds[eo * FROW + sk(k2)]));
        v.y = d2u(bitsd(lds[eo * FROW + sk(k2 + 1)]));
        *reinterpret_cast<ulonglong2 *>(
            folded + (((size_t)inst * r_out + jo) * m + i0 + i) * kD + k2) = v;
    }
    // (b) digit planes of the full-NTT slots: one 16-byte store per
    // (slot, digit row) covers the block's 16 columns
    for (int c = tid; c < 64 * r_out; c += FNT) {
        const int s = c / r_out, jo = c % r_out;
        const int zj = (int)tab[TAB_ZJ + s];
        const double z = tab[TAB_ZETA + s];
        uint32_t lo[16], hi[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint64_t *row = lds + (jo * FCOLS + i) * FROW;
            const double a = addmod(bitsd(row[sk(zj)]), mulmod(z, bitsd(row[36 + sk(zj)]), M), M);
            const uint64_t y = (d2u(a) + DIGIT_BIAS8) ^ 0x8080808080808080ULL;
            lo[i] = (uint32_t)y;
            hi[i] = (uint32_t)(y >> 32);
        }
        const size_t R = (size_t)inst * r_out + jo;
        int8_t *base = planes + (((size_t)s * ktiles + i0 / 64) * rt_n) * 1024 + R * 7 * 64 + (i0 % 64);
        uint32_t pk[4][7];
#pragma unroll
        for (int w4 = 0; w4 < 4; w4++) {
            const uint32_t l4[4] = {lo[4 * w4], lo[4 * w4 + 1], lo[4 * w4 + 2], lo[4 * w4 + 3]};
            const uint32_t h4[4] = {hi[4 * w4], hi[4 * w4 + 1], hi[4 * w4 + 2], hi[4 * w4 + 3]};
            bytes_transpose4(l4, h4, pk[w4]);
        }
#pragma unroll
        for (int u = 0; u < 7; u++)
            *reinterpret_cast<uint4 *>(base + u * 64) = make_uint4(pk[0][u], pk[1][u], pk[2][u], pk[3][u]);
    }
    if (!COEFF_OUT) return;
    __syncthreads();
    // (c) Coefficient form: inverse IncompleteNTT of the folded rows in place
    const bool olive = e < nout;
    double v[8];
    if (olive) {
#pragma unroll
        for (int x = 0; x < 8; x++) v[x] = bitsd(lds[e * FROW + 36 * p + sk(8 * t + x)]);
        ntt32_inv_b(v, tab + TAB_INV32, t, M);
#pragma unroll
        for (int x = 0; x < 8; x++) lds[e * FROW + 36 * p + sk(8 * t + x)] = dbits(v[x]);
    }
    __syncthreads();
    if (olive) {
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = bitsd(lds[e * FROW + 36 * p + sk(t + 4 * k)]);
        ntt32_inv_a(v, tab + TAB_INV32, tab[TAB_NINV32], M);
    }
    __syncthreads();
    if (olive) {
#pragma unroll
        for (int k = 0; k < 8; k++) lds[e * FROW + 8 * k + 2 * t + p] = d2u(v[k]);
    }
    __syncthreads();
    for (int c = tid; c < nout * 32; c += FNT) {
        const int eo = c >> 5, k2 = (c & 31) * 2, jo = eo >> 4, i = eo & 15;
        *reinterpret_cast<ulonglong2 *>(
            folded_coeff + (((size_t)inst * r_out + jo) * m + i0 + i) * kD + k2) =
            *reinterpret_cast<const ulonglong2 *>(&lds[eo * FROW + k2]);
    }
}


// grid (m / 16, n).  W [n * r][m][d]; Cn [n][r_out][r * parts][d] in
// IncompleteNTT form; outputs for stacked rows R = i * r_out + j.  The
// instances k_fold_mfma takes are skipped.
__global__ __launch_bounds__(FNT) void k_decomp_fold(
    const uint64_t *__restrict__ W, int r, size_t m, int parts,
    const uint64_t *__restrict__ d_radix, const uint64_t *__restrict__ Cn, int r_out,
    uint64_t *__restrict__ folded, uint64_t *__restrict__ folded_coeff,
    int8_t *__restrict__ planes, size_t ktiles, size_t rt_n, uint64_t q,
    const double *__restrict__ tab, Mod M, const unsigned *__restrict__ bad) {
    __shared__ uint64_t lds[128 * FROW];
    __shared__ double cl[8 * 8 * kD];  // this limb's challenges C[jo][rho * parts + tl]
    const int inst = blockIdx.y;
    if (fold_mfma_ok(bad, d_radix, inst)) return;  // uniform
    const size_t i0 = fold_col_block() * FCOLS;
    const int tid = threadIdx.x;
    // NTT role: element e, half p, quarter t (positions 8k + 2t + p)
    const int e = tid >> 3, h = tid & 7, p = h >> 2, t = h & 3;
    const int nel = r * FCOLS;  // element e = rho * 16 + i
    const uint64_t radix = d_radix[inst];
    const unsigned lg = (unsigned)__ffsll((long long)radix) - 1u;
    const uint64_t mask = radix - 1;
    // ---- load the block's r x 16 elements (coalesced 16-byte loads)
    for (int c = tid; c < nel * 32; c += FNT) {
        const int e = c >> 5, k2 = (c & 31) * 2;
        const int rho = e >> 4, i = e & 15;
        const ulonglong2 v = *reinterpret_cast<const ulonglong2 *>(
            W + (((size_t)inst * r + rho) * m + i0 + i) * kD + k2);
        *reinterpret_cast<ulonglong2 *>(&lds[e * FROW + k2]) = v;
    }
    __syncthreads();
    const bool live = e < nel;
    uint64_t raw[8];
    if (live) {
#pragma unroll
        for (int k = 0; k < 8; k++) raw[k] = lds[e * FROW + 8 * k + 2 * t + p];
    }
    __syncthreads();
    // MAC role: slot j, column i, output rows 4 jh .. 4 jh + 3
    const int mj = tid & 31, mi = (tid >> 5) & 15, jh = tid >> 9;
    const double om = tab[TAB_OMEGA + mj];
    double ae[4] = {0.0, 0.0, 0.0, 0.0}, ao[4] = {0.0, 0.0, 0.0, 0.0};
    const int rk = r * parts;
    for (int tl = 0; tl < parts; tl++) {
        // limb tl of every element, IncompleteNTT into the skewed layout
        double v[8];
        if (live) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = u2d(limb_digit(raw[k], q, lg, mask, tl));
            ntt32_fwd_a(v, tab + TAB_FWD32, M);
#pragma unroll
            for (int k = 0; k < 8; k++) lds[e * FROW + 36 * p + sk(t + 4 * k)] = dbits(v[k]);
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int x = 0; x < 8; x++) v[x] = bitsd(lds[e * FROW + 36 * p + sk(8 * t + x)]);
            ntt32_fwd_b(v, tab + TAB_FWD32, t, M);
#pragma unroll
            for (int x = 0; x < 8; x++) lds[e * FROW + 36 * p + sk(8 * t + x)] = dbits(v[x]);
        }
        // this limb's challenge slice into LDS: cl[(jo * 8 + rho) * d + s]
        for (int c = tid; c < r_out * r * kD; c += FNT) {
            const int s = c & (kD - 1), jr = c >> 6, jo = jr / r, rho = jr % r;
            cl[(jo * 8 + rho) * kD + s] =
                u2d(Cn[(((size_t)inst * r_out + jo) * rk + rho * parts + tl) * kD + s]);
        }
        __syncthreads();
        // fold: out[j] += C[j][k] * limb_k in slot mj, k = rho * parts + tl
        for (int rho = 0; rho < r; rho++) {
            const uint64_t *row = lds + (rho * FCOLS + mi) * FROW;
            const double le = bitsd(row[sk(mj)]), lo = bitsd(row[36 + sk(mj)]);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int jo = 4 * jh + u;
                if (jo >= r_out) continue;
                const double *ce = cl + (jo * 8 + rho) * kD;
                const double c0 = ce[mj], c1 = ce[kH + mj];
                ae[u] += mul_lazy(le, c0, M) + mul_lazy(mulmod(lo, c1, M), om, M);
                ao[u] += mul_lazy(lo, c0, M) + mul_lazy(le, c1, M);
                ae[u] = reduce_lazy(ae[u], M);
                ao[u] = reduce_lazy(ao[u], M);
            }
        }
        __syncthreads();  // the next limb rewrites the LDS rows
    }
    // ---- folded rows (IncompleteNTT) into LDS rows e' = jo * 16 + i
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int jo = 4 * jh + u;
        if (jo < r_out) {
            uint64_t *row = lds + (jo * FCOLS + mi) * FROW;
            row[sk(mj)] = dbits(reduce(ae[u], M));
            row[36 + sk(mj)] = dbits(reduce(ao[u], M));
        }
    }
    __syncthreads();
    fold_outputs<true>(lds, inst, i0, m, r_out, folded, folded_coeff, planes, ktiles, rt_n, tab, M,
                       tid);
}

// grid (m / 16, n): for the instances k_fold_mfma takes, its Coefficient rows
// (16 columns of the r_out rows per block) to the IncompleteNTT fold output
// and the commit's digit planes.
__global__ __launch_bounds__(FNT) void k_fold_coeff_out(
    const uint64_t *__restrict__ d_radix, const unsigned *__restrict__ bad, int r_out, size_t m,
    uint64_t *__restrict__ folded, uint64_t *__restrict__ folded_coeff, int8_t *__restrict__ planes,
    size_t ktiles, size_t rt_n, const double *__restrict__ tab, Mod M) {
    __shared__ uint64_t lds[128 * FROW];
    const int inst = blockIdx.y;
    if (!fold_mfma_ok(bad, d_radix, inst)) return;  // uniform
    const size_t i0 = fold_col_block() * FCOLS;
    const int tid = threadIdx.x, nout = r_out * FCOLS;
    const int e = tid >> 3, h = tid & 7, p = h >> 2, t = h & 3;
    for (int c = tid; c < nout * 32; c += FNT) {
        const int eo = c >> 5, k2 = (c & 31) * 2, jo = eo >> 4, i = eo & 15;
        *reinterpret_cast<ulonglong2 *>(&lds[eo * FROW + k2]) = *reinterpret_cast<const ulonglong2 *>(
            folded_coeff + (((size_t)inst * r_out + jo) * m + i0 + i) * kD + k2);
    }
    __syncthreads();
    const bool olive = e < nout;
    double v[8];
    if (olive) {
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = u2d(lds[e * FROW + 8 * k + 2 * t + p]);
    }
    __syncthreads();
    if (olive) {
        ntt32_fwd_a(v, tab + TAB_FWD32, M);
#pragma unroll
        for (int k = 0; k < 8; k++) lds[e * FROW + 36 * p + sk(t + 4 * k)] = dbits(v[k]);
    }
    __syncthreads();
    if (olive) {
#pragma unroll
        for (int x = 0; x < 8; x++) v[x] = bitsd(lds[e * FROW + 36 * p + sk(8 * t + x)]);
        ntt32_fwd_b(v, tab + TAB_FWD32, t, M);
#pragma unroll
        for (int x = 0; x < 8; x++) lds[e * FROW + 36 * p + sk(8 * t + x)] = dbits(v[x]);
    }
    __syncthreads();
    fold_outputs<false>(lds, inst, i0, m, r_out, folded, folded_coeff, planes, ktiles, rt_n, tab, M,
                        tid);
}

__global__ void k_fill_u32(unsigned *p, size_t n, unsigned v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// per-row copies of each instance's challenges: out[R] = cs[R / r_out]
__global__ void k_expand_cs(const uint64_t *__restrict__ cs, size_t per, int r_out, size_t rows,
                            uint64_t *__restrict__ out) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= rows * per) return;
    const size_t R = gid / per, x = gid % per;
    out[gid] = cs[(R / r_out) * per + x];
}

}  // namespace

extern "C" {

int salsaa_fold_step(salsaa_ctx *ctx, const salsaa_key *key, size_t n, const uint64_t *d_W,
                     size_t r, size_t m, size_t parts, uint64_t radix, const uint64_t *d_C,
                     size_t r_out, const uint64_t *d_cs, size_t reps, int fq2, uint64_t nr,
                     const salsaa_fold_outputs *o) {
    CTX_D64(ctx, "salsaa_fold_step");
    Context *c = &ctx->c;
    SALSAA_CHECK(key && o && d_W && d_C && d_cs, "null argument");
    SALSAA_CHECK(key->k.ctx == c && (key->k.d_AB || key->k.d_mhat),
                 "key from another context or without planes");
    SALSAA_CHECK(n >= 1 && n < 65536, "instances out of range");
    SALSAA_CHECK(r >= 1 && r <= 8 && parts >= 1 && r * parts <= 64,
                 "the fused fold takes r <= 8 rows and r*parts <= 64 limb rows");
    SALSAA_CHECK(r_out >= 1 && r_out <= 8, "the fused fold writes r_out <= 8 rows");
    SALSAA_CHECK(m >= 128 && m % 128 == 0 && m <= key->k.m,
                 "witness length must be a multiple of 128 within the key");
    SALSAA_CHECK((m & (m - 1)) == 0, "witness length must be a power of two (norm check)");
    SALSAA_CHECK(o->folded && o->folded_coeff && o->rhs && o->radix && o->ok,
                 "fold outputs missing");
    Mod M{c->qd, c->qinv};
    const size_t rows = n * r_out, rk = r * parts;
    // radix per instance (auto when radix = 0), then the challenge NTTs
    SALSAA_TRY(launch_radix_n(c, d_W, n, r, m, parts, radix, o->radix));
    // Which instances fold on the matrix cores: bad[i] is set for a C that is
    // not ternary (and for all when the fragment table would exceed 256 MiB);
    // the radix bound is checked by the kernels themselves.
    unsigned *bad = nullptr;
    int8_t *tfrag = nullptr;
    const size_t tbytes = n * r_out * rk * 4 * 1024;
    SALSAA_TRY(scratch_get(c, 12, n * sizeof(unsigned), (void **)&bad));
    const bool table = tbytes <= ((size_t)256 << 20);
    hipLaunchKernelGGL(k_fill_u32, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, bad, n,
                       table ? 0u : 1u);
    if (table) {
        SALSAA_TRY(scratch_get(c, 13, tbytes, (void **)&tfrag));
        hipLaunchKernelGGL(k_fold_prep, dim3((unsigned)(r_out * rk), (unsigned)n), dim3(256), 0,
                           c->stream, d_C, (int)r_out, (int)rk, c->q, tfrag, bad);
    }
    SALSAA_HIP(hipGetLastError());
    uint64_t *cn = nullptr;
    SALSAA_TRY(scratch_get(c, 5, n * r_out * rk * kD * sizeof(uint64_t), (void **)&cn));
    SALSAA_HIP(hipMemcpyAsync(cn, d_C, n * r_out * rk * kD * sizeof(uint64_t),
                              hipMemcpyDeviceToDevice, c->stream));
    SALSAA_TRY(launch_ntt(c, cn, n * r_out * rk, SALSAA_REP_INCOMPLETE, 0));
    // fused decomposition + fold -> fold output, planes, Coefficient form
    const size_t ktiles = plane_ktiles(m), rt_n = (7 * rows + 15) / 16;
    int8_t *planes = nullptr;
    SALSAA_TRY(scratch_get(c, 6, plane_bytes(rows, m), (void **)&planes));
    // the matrix-core fold with its outputs, and the NTT-domain fold of the
    // other instances
    if (table) {
        hipLaunchKernelGGL(k_fold_mfma, dim3((unsigned)(m / FM_COLS), (unsigned)n), dim3(FM_THREADS), 0,
                           c->stream, d_W, (int)r, m, (int)parts, o->radix, bad, tfrag, (int)r_out,
                           c->q, o->folded_coeff);
        hipLaunchKernelGGL(k_fold_coeff_out, dim3((unsigned)(m / FCOLS), (unsigned)n), dim3(FNT), 0,
                           c->stream, o->radix, bad, (int)r_out, m, o->folded, o->folded_coeff, planes,
                           ktiles, rt_n, c->d_tables, M);
    }
    hipLaunchKernelGGL(k_decomp_fold, dim3((unsigned)(m / FCOLS), (unsigned)n), dim3(FNT), 0,
                       c->stream, d_W, (int)r, m, (int)parts, o->radix, cn, (int)r_out, o->folded,
                       o->folded_coeff, planes, ktiles, rt_n, c->q, c->d_tables, M, bad);
    SALSAA_HIP(hipGetLastError());
    // one commit of the stacked folded witnesses
    if (key->k.d_mhat)
        SALSAA_TRY(launch_commit_crs(c, key->k.d_mhat, key->k.kappa, planes, rows, m, ktiles, o->rhs));
    else
        SALSAA_TRY(launch_commit_mfma(c, key->k.d_AB, key->k.kappa, key->k.a_kt, key->k.a_ktiles,
                                      planes, rows, m, ktiles, o->rhs));
    // norm checks: every folded row with its instance's challenges
    const size_t L = m * kD, nrounds = (size_t)__builtin_ctzll(L), per = reps * nrounds * 2;
    uint64_t *csx = nullptr;
    SALSAA_TRY(scratch_get(c, 8, rows * per * sizeof(uint64_t), (void **)&csx));
    hipLaunchKernelGGL(k_expand_cs, dim3(ceil_div(rows * per, 256)), dim3(256), 0, c->stream, d_cs,
                       per, (int)r_out, rows, csx);
    SALSAA_HIP(hipGetLastError());
    const NcOut no = nc_out_plain(rows, reps, nrounds, o->claims, o->msgs, o->finals, o->evals,
                                  o->nc_rhs, o->ok);
    return launch_norm_check_ex(c, o->folded_coeff, nullptr, rows, L, csx, per, reps, fq2, nr, no,
                                0, nullptr);
}

}  // extern "C"
