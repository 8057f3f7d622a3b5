// sumcheck.hip -- the norm check (norm_check_1/2/3 with their verifier
// checks, fast_norm_check.rs:135-199, 282-462) for many rows and repetitions
// in a few launches.
//
// Round t halves the current witness into w0 | w1 (most significant variable
// first), sends f0 = sum w0^2, f1 = sum w0 (w1 - w0), f2 = sum (w1 - w0)^2 and
// folds w <- w0 + c_t (w1 - w0).  The challenges do not depend on the
// messages (the reference verifier samples them independently,
// fast_norm_check.rs:375-385), so one HBM pass can run SC_T = 4 rounds: a
// thread owns the 16 entries {y * S + x : y < 16} of a column x (S = len/16),
// which are exactly the entries rounds t..t+3 pair up.  Each pass writes the
// 1/16-length folded witness.  The first pass (over the F_q witness) gets
// its messages from the rows' Gram matrices on the matrix cores (see the
// Gram pass below); the later passes add per-block message partials into u64
// accumulators with integer atomics (exact, order independent).  Once the
// length is at most 4096 a single block per (repetition, row) evaluates the
// Gram messages, runs one more 4-round pass if the length exceeds 1024, runs
// the remaining rounds in LDS, captures the length-d folded vector V for
// norm_check_3, forms rhs = conj(tensor(coefficient challenges)) * V, and
// replays the verifier over all rounds.
#include <algorithm>

#include "ntt_device.hpp"

namespace salsaa {

namespace {

constexpr int SC_THREADS = 256;
constexpr int SC_FINAL_MAX = 1024;  // lengths the final block holds in LDS
constexpr int SC_TAIL_MAX = 4096;   // lengths the final block folds once from HBM first

// ---- field helpers: F_q (FQ2 = false, imaginary parts constant zero) or
// F_{q^2}.  SNR: the non-residue is below 8, so nr * t for a lazy product
// |t| < q is an exact double product (|nr t| < 8q < 2^53) and needs no
// reduction of its own.
template <bool SNR>
__device__ __forceinline__ double nr_times(double nr, double t, Mod M) {
    return SNR ? nr * t : mul_lazy(nr, t, M);
}
template <bool FQ2, bool SNR>
__device__ __forceinline__ F2 x_mul(F2 x, F2 y, double nr, Mod M) {
    if (!FQ2) return {mulmod(x.a, y.a, M), 0.0};
    const double re = mul_lazy(x.a, y.a, M) + nr_times<SNR>(nr, mul_lazy(x.b, y.b, M), M);
    const double im = mul_lazy(x.a, y.b, M) + mul_lazy(x.b, y.a, M);
    return {reduce(re, M), reduce(im, M)};
}
template <bool FQ2>
__device__ __forceinline__ F2 x_add(F2 x, F2 y, Mod M) {
    if (!FQ2) return {addmod(x.a, y.a, M), 0.0};
    return f2_add(x, y, M);
}
template <bool FQ2>
__device__ __forceinline__ F2 x_sub(F2 x, F2 y, Mod M) {
    if (!FQ2) return {submod(x.a, y.a, M), 0.0};
    return f2_sub(x, y, M);
}
// acc += x * y, lazily (|acc| stays below q/2 + 1 after each step).  F_{q^2}
// products use three multiplications: (a + b)(c + d) - ac - bd = ad + bc.  With
// canonical inputs the sums stay below 2q, so (a + b)(c + d) < 4q^2 < 2^102:
// mul_lazy's quotient estimate is then off by at most 1.5 and its result,
// still exact and congruent, lies in (-2q, 2q); ad + bc = p3 - p1 - p2 stays
// within (-4q, 4q), well inside reduce's 2^53 range.
template <bool FQ2, bool SNR>
__device__ __forceinline__ void x_mac(F2 &acc, F2 x, F2 y, double nr, Mod M) {
    if (!FQ2) {
        acc.a = reduce_lazy(acc.a + mul_lazy(x.a, y.a, M), M);
        return;
    }
    const double p1 = mul_lazy(x.a, y.a, M), p2 = mul_lazy(x.b, y.b, M);
    const double p3 = mul_lazy(x.a + x.b, y.a + y.b, M);
    acc.a = reduce_lazy(acc.a + p1 + nr_times<SNR>(nr, p2, M), M);
    acc.b = reduce_lazy(acc.b + (p3 - p1 - p2), M);
}
// x * y, canonical (three multiplications as in x_mac)
template <bool FQ2, bool SNR>
__device__ __forceinline__ F2 x_mul3(F2 x, F2 y, double nr, Mod M) {
    if (!FQ2) return {mulmod(x.a, y.a, M), 0.0};
    const double p1 = mul_lazy(x.a, y.a, M), p2 = mul_lazy(x.b, y.b, M);
    const double p3 = mul_lazy(x.a + x.b, y.a + y.b, M);
    return {reduce(p1 + nr_times<SNR>(nr, p2, M), M), reduce(p3 - p1 - p2, M)};
}
// acc += x^2 = (a^2 + nr b^2) + 2ab Y
template <bool FQ2, bool SNR>
__device__ __forceinline__ void x_sq_mac(F2 &acc, F2 x, double nr, Mod M) {
    if (!FQ2) {
        acc.a = reduce_lazy(acc.a + mul_lazy(x.a, x.a, M), M);
        return;
    }
    const double re = mul_lazy(x.a, x.a, M) + nr_times<SNR>(nr, mul_lazy(x.b, x.b, M), M);
    const double im = 2.0 * mul_lazy(x.a, x.b, M);
    acc.a = reduce_lazy(acc.a + re, M);
    acc.b = reduce_lazy(acc.b + im, M);
}

// ---- streaming passes: balanced lazy residues.  The passes keep every
// witness value and challenge as a balanced residue |x| <= q/2 + 2 (what
// reduce_lazy returns) and never canonicalise inside a level.  For |a b| <= P,
// mul_lazy returns P' = a b - qt q exactly with |P'| <= (1/2 + 1.5 * 2^-53 P/q) q,
// so for the operand bounds below (|w0|, |c| <= h = q/2 + 2; |d|, |w0.a + w0.b|,
// |c.a + c.b| <= 2h; |d.a + d.b| <= 4h) every product stays within 0.875 q,
// a group of four products within 3.5 q, and each reduce_lazy input below
// within 6.8 q < 2^53.

// One round over NP pairs (w0, w1) = (cur[y], cur[y + NP]): adds
// f0 = sum w0^2 and f1 = sum w0 (w1 - w0) into the lazy accumulators and folds
// cur[y] <- w0 + c (w1 - w0).  csum = c.a + c.b.  F_{q^2} products take three
// multiplications: x y = (xa ya + nr xb yb) + ((xa + xb)(ya + yb) - xa ya - xb yb) Y,
// x^2 = (xa^2 + nr xb^2) + 2 xa xb Y.  Sums are reduced per group of four pairs.
template <bool FQ2, bool SNR, int NP>
__device__ __forceinline__ void sc_level(F2 *cur, F2 c, double csum, F2 &f0, F2 &f1,
                                         double nr, Mod M) {
#pragma unroll
    for (int g = 0; g < NP; g += 4) {
        double A = 0.0, B = 0.0, C = 0.0, K1 = 0.0, K2 = 0.0, K3 = 0.0;
#pragma unroll
        for (int y = g; y < (NP < g + 4 ? NP : g + 4); y++) {
            const F2 w0 = cur[y], w1 = cur[y + NP];
            const double da = w1.a - w0.a;
            A += mul_lazy(w0.a, w0.a, M);
            K1 += mul_lazy(w0.a, da, M);
            const double m1 = mul_lazy(c.a, da, M);
            if (FQ2) {
                const double db = w1.b - w0.b, sd = da + db;
                B += mul_lazy(w0.b, w0.b, M);
                C += mul_lazy(w0.a, w0.b, M);
                K2 += mul_lazy(w0.b, db, M);
                K3 += mul_lazy(w0.a + w0.b, sd, M);
                const double m2 = mul_lazy(c.b, db, M), m3 = mul_lazy(csum, sd, M);
                cur[y].a = reduce_lazy(w0.a + m1 + nr_times<SNR>(nr, m2, M), M);
                cur[y].b = reduce_lazy(w0.b + (m3 - m1 - m2), M);
            } else {
                cur[y].a = reduce_lazy(w0.a + m1, M);
            }
        }
        if (FQ2) {
            f0.a = reduce_lazy(f0.a + A + nr_times<SNR>(nr, reduce_lazy(B, M), M), M);
            f0.b = reduce_lazy(f0.b + 2.0 * C, M);
            f1.a = reduce_lazy(f1.a + K1 + nr_times<SNR>(nr, reduce_lazy(K2, M), M), M);
            f1.b = reduce_lazy(f1.b + reduce_lazy(K3, M) - K1 - K2, M);
        } else {
            f0.a = reduce_lazy(f0.a + A, M);
            f1.a = reduce_lazy(f1.a + K1, M);
        }
    }
}

// challenges of one row: cs[rep][t][2] (rows either share one block or each
// row has its own, cs_row u64 apart)
template <bool FQ2>
__device__ __forceinline__ F2 load_chal(const uint64_t *cs, int rep, int nrounds, int t) {
    const uint64_t *p = cs + ((size_t)rep * nrounds + t) * 2;
    return {u2d(p[0]), FQ2 ? u2d(p[1]) : 0.0};
}

template <bool FQ2>
__device__ __forceinline__ F2 load_chal_bal(const uint64_t *cs, int rep, int nrounds, int t, Mod M) {
    const F2 c = load_chal<FQ2>(cs, rep, nrounds, t);
    return {balance(c.a, M), FQ2 ? balance(c.b, M) : 0.0};
}

__device__ __forceinline__ size_t macc_idx(int rep, int t, int row, int nrounds, int rows) {
    return (((size_t)rep * nrounds + t) * rows + row) * 6;
}

// Block-wide sum of NV values (each in [0, q)); result in red[0..NV).  Every
// thread of the block calls it.
template <int NV>
__device__ __forceinline__ void block_sum(double (&x)[NV], double *red, Mod M) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int NW = SC_THREADS / 64;
    __syncthreads();
    double ws[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) ws[v] = wave_sum_res(x[v], M);
    if (lane == 0) {
#pragma unroll
        for (int v = 0; v < NV; v++) red[NV + wave * NV + v] = ws[v];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double t = red[NV + threadIdx.x];
#pragma unroll
        for (int w = 1; w < NW; w++) t = addmod(t, red[NV + w * NV + threadIdx.x], M);
        red[threadIdx.x] = t;
    }
    __syncthreads();
}

// x mod q for any u64 x: x = hi * 2^32 + lo with hi, lo < 2^32 < q
__device__ __forceinline__ double u64_mod(unsigned long long x, Mod M) {
    const double hi = u2d(x >> 32), lo = u2d(x & 0xFFFFFFFFULL);
    return addmod(mulmod(hi, 4294967296.0, M), lo, M);
}

__device__ __forceinline__ void atomic_add_res(unsigned long long *p, double v) {
    atomicAdd(p, (unsigned long long)d2u(v));
}

// ---------------------------------------------------------------- Gram pass
// Rounds 0..3 need, per row and repetition, only quadratic forms of the
// row's 16 x 16 Gram matrix over F_q,
//     G[y][y'] = sum_x v[y S + x] v[y' S + x]     (S = L / 16):
// with A_0 = I and A_{t+1}[z] = A_t[z] + c_t (A_t[z + h] - A_t[z]) (h = 8 >> t),
// round t's witness is w_t[z S + x] = sum_y A_t[z][y] v[y S + x], so
//     f0 = sum_{z<h} A_t[z]^T G A_t[z],  f1 = sum_{z<h} (A_t[z + h] - A_t[z])^T G A_t[z]
// (G is symmetric), the claim is tr G, and the 1/16-length witness after
// round 3 is sum_y A_4[0][y] v[y S + x].  k_sc_gram streams the witness
// once: every thread folds one column x for the repetitions on VALU (16
// products per component), and the block's columns go through LDS into the
// int8 matrix cores for G.  With the 7 signed base-256 digits d_i of each
// residue (as in the commit), G = sum_k 2^(8k) E_k, E_k = sum_{i+j=k} D_i D_j^T,
// and each E_k accumulates exactly in int32 (|d| <= 128, at most 7 digit
// pairs, at most 2^14 columns per wave: 7 * 2^14 * 2^14 < 2^31).  The final
// block then evaluates the
// forms for its (row, repetition).
//
// Short witnesses (the norm check's purpose) take a cheaper path: when every
// residue of a wave's 64 columns is b in balanced form with |b| < 2^14,
// t = b + 2^14 < 2^15 goes through LDS as a u32.  The fold is then exact in
// fp64 without reductions: with c = c_h 2^25 + c_l (|c_h|, |c_l| <= 2^24),
// sum_y c_h t_y and sum_y c_l t_y stay below 2^44, and
// sum_y c b_y = 2^25 sum c_h t + sum c_l t - 2^14 sum c.  The Gram matrix
// needs only b's two balanced base-256 digits (E_0, E_1, E_2: 4 products).
// k_sc_gram takes this path per wave and iteration after checking the
// residues; with the round trip's short copy of a batch arena (ShortCopy),
// k_sc_gram_short reads the copy (one or two bytes per entry) for every
// sub-block whose slices all have one width, and k_sc_gram_rest takes the
// others from the witness.
typedef int v4i __attribute__((ext_vector_type(4)));
constexpr int GR_X = SC_THREADS;  // columns per iteration
constexpr int GR_LD = GR_X + 2;   // LDS row stride (u64): fragment reads hit distinct banks
constexpr int GR_MAXREP = 4;      // repetitions folded per launch
constexpr uint64_t GR_BIAS = 0x0080808080808080ULL;
constexpr uint32_t GR_SMALL = kShortOff;  // short path: balanced |b| < 2^14, t = b + 2^14

// 2^25 h + l + k for the short path's exact sums (|h|, |l| < 2^44, k balanced),
// reduced to a balanced residue: 2^25 h is exact and below 2^69, its remainder
// fma(-qt, q, .) is an exact integer below q
__device__ __forceinline__ double gr_small_sum(double h, double l, double k, Mod M) {
    const double hi = h * 33554432.0;
    const double r = __fma_rn(-__builtin_rint(hi * M.qinv), M.q, hi);
    return reduce_lazy(r + l + k, M);
}

// A block's fold coefficients for repetitions rep0 .. rep0 + nrep - 1:
// A_4[0][y] = prod_s (bit 3 - s of y ? c_s : 1 - c_s), balanced; their
// halves for the short path (spare repetition slots zero) and -2^14 sum_y c.
struct GrCoef {
    F2 coef[GR_MAXREP][16];
    __align__(32) double csplit[GR_MAXREP][16][4];  // (c_h, c_l) of both components
    double coff[GR_MAXREP][2];
};
template <bool FQ2, bool SNR>
__device__ __forceinline__ void gr_coefs(GrCoef &C, const uint64_t *cs, int rep0, int nrep,
                                         int nrounds, double nr, Mod M, int tid) {
    if (tid < 16 * nrep) {
        const int r = tid >> 4, y = tid & 15;
        F2 a = {1.0, 0.0};
        for (int t = 0; t < 4; t++) {
            const F2 c = load_chal<FQ2>(cs, rep0 + r, nrounds, t);
            a = x_mul<FQ2, SNR>(a, ((y >> (3 - t)) & 1) ? c : x_sub<FQ2>(F2{1.0, 0.0}, c, M), nr, M);
        }
        const double ca = balance(a.a, M), cb = balance(a.b, M);
        C.coef[r][y] = {ca, cb};
        const double ha = __builtin_rint(ca * 0x1p-25), hb = __builtin_rint(cb * 0x1p-25);
        C.csplit[r][y][0] = ha;
        C.csplit[r][y][1] = ca - ha * 0x1p25;
        C.csplit[r][y][2] = hb;
        C.csplit[r][y][3] = cb - hb * 0x1p25;
    } else if (tid < 16 * GR_MAXREP) {
        C.csplit[tid >> 4][tid & 15][0] = C.csplit[tid >> 4][tid & 15][1] = 0.0;
        C.csplit[tid >> 4][tid & 15][2] = C.csplit[tid >> 4][tid & 15][3] = 0.0;
    }
    __syncthreads();
    if (tid < 2 * nrep) {
        const int r = tid >> 1, comp = tid & 1;
        double sum = 0.0;  // 16 balanced terms: below 8 q
        for (int y = 0; y < 16; y++) sum += comp ? C.coef[r][y].b : C.coef[r][y].a;
        C.coff[r][comp] = reduce_lazy(mul_lazy(reduce_lazy(sum, M), -(double)GR_SMALL, M), M);
    }
    __syncthreads();
}

// The width of a batch's short copy over a first-pass sub-block: the block's
// columns [xb, xe] of the NSEG row segments (S apart) meet slices whose
// format bytes are all 0 (0: read the one-byte copy), all 1 (1: the two-byte
// copy), or neither (2: a mix of widths, or a slice that is not short; the
// rest kernel reads the sub-block from the witness).  Block-uniform: every
// thread of the block calls it.
template <int NSEG>
__device__ __forceinline__ int sc_classify(const uint8_t *__restrict__ fmt, size_t base, size_t S,
                                           size_t xb, size_t xe) {
    constexpr int PARTS = SC_THREADS / NSEG;
    const int y = threadIdx.x % NSEG, part = threadIdx.x / NSEG;
    const size_t s0 = (base + (size_t)y * S + xb) / kShortSlice;
    const size_t s1 = (base + (size_t)y * S + xe) / kShortSlice;
    int n0 = 0, n1 = 0;
    for (size_t sl = s0 + part; sl <= s1; sl += PARTS) {
        const uint8_t f = fmt[sl];
        n0 |= f != 0;
        n1 |= f != 1;
    }
    const int any0 = __syncthreads_or(n0), any1 = __syncthreads_or(n1);
    return !any0 ? 0 : !any1 ? 1 : 2;
}

// The short fold of column x from its 16 entries t (PAIR: two repetitions per
// sweep, so each y issues its coefficient reads together; a lone last
// repetition reads a zeroed spare slot and discards those sums).
template <bool FQ2, bool PAIR>
__device__ __forceinline__ void gr_short_fold(const GrCoef &C, const uint32_t (&tt)[16], int nrep,
                                              int rep0, int rows, int row, size_t S, size_t x,
                                              double *__restrict__ out, Mod M) {
    if (PAIR) {
        for (int r = 0; r < nrep; r += 2) {
            double h0 = 0.0, l0 = 0.0, h1 = 0.0, l1 = 0.0;  // repetition r: a, b
            double h2 = 0.0, l2 = 0.0, h3 = 0.0, l3 = 0.0;  // repetition r + 1
#pragma unroll
            for (int y = 0; y < 16; y++) {
                const double tv = (double)tt[y];
                const double4 c0 = *reinterpret_cast<const double4 *>(&C.csplit[r][y][0]);
                const double4 c1 = *reinterpret_cast<const double4 *>(&C.csplit[r + 1][y][0]);
                h0 = __fma_rn(c0.x, tv, h0);
                l0 = __fma_rn(c0.y, tv, l0);
                h2 = __fma_rn(c1.x, tv, h2);
                l2 = __fma_rn(c1.y, tv, l2);
                if (FQ2) {
                    h1 = __fma_rn(c0.z, tv, h1);
                    l1 = __fma_rn(c0.w, tv, l1);
                    h3 = __fma_rn(c1.z, tv, h3);
                    l3 = __fma_rn(c1.w, tv, l3);
                }
            }
            double *o = out + (((size_t)(rep0 + r) * rows + row) * S + x) * 2;
            o[0] = gr_small_sum(h0, l0, C.coff[r][0], M);
            o[1] = FQ2 ? gr_small_sum(h1, l1, C.coff[r][1], M) : 0.0;
            if (r + 1 < nrep) {
                o += (size_t)rows * S * 2;
                o[0] = gr_small_sum(h2, l2, C.coff[r + 1][0], M);
                o[1] = FQ2 ? gr_small_sum(h3, l3, C.coff[r + 1][1], M) : 0.0;
            }
        }
    } else {
        for (int r = 0; r < nrep; r++) {
            double ah = 0.0, al = 0.0, bh = 0.0, bl = 0.0;
#pragma unroll
            for (int y = 0; y < 16; y++) {
                const double tv = (double)tt[y];
                const double2 ca = *reinterpret_cast<const double2 *>(&C.csplit[r][y][0]);
                ah = __fma_rn(ca.x, tv, ah);
                al = __fma_rn(ca.y, tv, al);
                if (FQ2) {
                    const double2 cb = *reinterpret_cast<const double2 *>(&C.csplit[r][y][2]);
                    bh = __fma_rn(cb.x, tv, bh);
                    bl = __fma_rn(cb.y, tv, bl);
                }
            }
            double *o = out + (((size_t)(rep0 + r) * rows + row) * S + x) * 2;
            o[0] = gr_small_sum(ah, al, C.coff[r][0], M);
            o[1] = FQ2 ? gr_small_sum(bh, bl, C.coff[r][1], M) : 0.0;
        }
    }
}

// Fragments of digits 0 and 1 of b = t - 2^14 for the matrix cores from the
// wave's u32 rows (row y at T32 + ld y): lane 16 g + r takes row r, columns
// 16 g .. 16 g + 15 ((b + 0x80) ^ 0x80 holds d0, d1 in bytes 0, 1).
__device__ __forceinline__ void gr_short_frags(const uint32_t *T32, int ld, int fr, int fg, v4i &f0,
                                               v4i &f1) {
    const uint4 *src = reinterpret_cast<const uint4 *>(T32 + ld * fr + 16 * fg);
#pragma unroll
    for (int k4 = 0; k4 < 4; k4++) {
        const uint4 u = src[k4];
        const uint32_t y0 = (u.x - (GR_SMALL - 0x80u)) ^ 0x80u;
        const uint32_t y1 = (u.y - (GR_SMALL - 0x80u)) ^ 0x80u;
        const uint32_t y2 = (u.z - (GR_SMALL - 0x80u)) ^ 0x80u;
        const uint32_t y3 = (u.w - (GR_SMALL - 0x80u)) ^ 0x80u;
        const uint32_t p01 = __builtin_amdgcn_perm(y1, y0, 0x05010400u);
        const uint32_t p23 = __builtin_amdgcn_perm(y3, y2, 0x05010400u);
        f0[k4] = (int)__builtin_amdgcn_perm(p23, p01, 0x05040100u);
        f1[k4] = (int)__builtin_amdgcn_perm(p23, p01, 0x07060302u);
    }
}

// The block's Gram sums into gacc[row]: lane (r, g) of wave `wave` holds
// E_k[4 g + e][r]; G entry = sum_k 2^(8k) E_k mod q, summed over the waves
// through gpart[4][256] (every wave reaches this point).
template <int K>
__device__ __forceinline__ void gr_finish(const v4i (&acc)[K], double *gpart, int wave, int lane,
                                          int tid, const double *__restrict__ tab,
                                          unsigned long long *__restrict__ gacc, int row, Mod M) {
    const int r = lane & 15, g = lane >> 4;
    const double *wt = tab + TAB_W256;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        double xg = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            xg += mul_lazy((double)acc[k][e], wt[k], M);
            if ((k % 7) == 6) xg = reduce_lazy(xg, M);
        }
        gpart[wave * 256 + (4 * g + e) * 16 + r] = reduce(xg, M);
    }
    __syncthreads();
    double tot = gpart[tid];
#pragma unroll
    for (int wv = 1; wv < 4; wv++) tot = addmod(tot, gpart[wv * 256 + tid], M);
    atomicAdd(&gacc[(size_t)row * 256 + tid], (unsigned long long)d2u(tot));
}

// The fold of a wide column (canonical residues at col[ld y], y < 16):
// products |coef v| <= 0.69 q (balanced coef, canonical v), eight of them stay
// below 5.6 q, then two lazy reductions.
template <bool FQ2>
__device__ __forceinline__ void gr_wide_fold(const GrCoef &C, const uint64_t *col, int ld, int nrep,
                                             int rep0, int rows, int row, size_t S, size_t x,
                                             double *__restrict__ out, Mod M) {
    for (int r = 0; r < nrep; r++) {
        double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
#pragma unroll
        for (int y = 0; y < 8; y++) {
            const F2 c0 = C.coef[r][y], c1 = C.coef[r][y + 8];
            const double v0 = u2d(col[ld * y]), v1 = u2d(col[ld * (y + 8)]);
            a0 += mul_lazy(c0.a, v0, M);
            a1 += mul_lazy(c1.a, v1, M);
            if (FQ2) {
                b0 += mul_lazy(c0.b, v0, M);
                b1 += mul_lazy(c1.b, v1, M);
            }
        }
        double *o = out + (((size_t)(rep0 + r) * rows + row) * S + x) * 2;
        o[0] = reduce_lazy(reduce_lazy(a0, M) + a1, M);
        o[1] = FQ2 ? reduce_lazy(reduce_lazy(b0, M) + b1, M) : 0.0;
    }
}

// The 7 signed base-256 digit fragments of 16 consecutive staged residues
// src[0 .. 15] (one digit row per fragment, four per dword).
__device__ __forceinline__ void gr_wide_frags(const uint64_t *src, v4i (&f)[7]) {
#pragma unroll
    for (int k4 = 0; k4 < 4; k4++) {
        uint32_t lo[4], hi[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint64_t yv = (src[4 * k4 + q] + GR_BIAS) ^ 0x8080808080808080ULL;
            lo[q] = (uint32_t)yv;
            hi[q] = (uint32_t)(yv >> 32);
        }
        uint32_t dg[7];
        bytes_transpose4(lo, hi, dg);
#pragma unroll
        for (int i = 0; i < 7; i++) f[i][k4] = (int)dg[i];
    }
}

// The first pass over a row's columns (k_sc_gram): one 16 x GR_LD u64
// staging block (Tall), reused for the Gram partials after the loop.
template <bool FQ2, bool SNR, bool GRAM>
__device__ __forceinline__ void sc_gram_body(
    uint32_t *Tall, GrCoef &C, const uint64_t *__restrict__ w,
    const unsigned long long *__restrict__ wrow, size_t L, int rows, int rep0, int nrep,
    const uint64_t *__restrict__ cs0, size_t cs_row, int nrounds, double nr, int cpt,
    double *__restrict__ out, unsigned long long *__restrict__ gacc, const double *__restrict__ tab,
    Mod M, int sbn, int nbx) {
    const int row = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t S = L >> 4;
    const size_t base = wrow ? wrow[row] : (size_t)row * L;
    const size_t span = (size_t)cpt * GR_X;
    const uint64_t *wr = w + base;
    uint64_t *T = reinterpret_cast<uint64_t *>(Tall);  // row y at T + GR_LD y
    // the short path's u32 rows of the wave's columns: row y at T32 + 2 GR_LD y
    // (the first half of the row's 64 u64 slots)
    const int ld32 = 2 * GR_LD;
    uint32_t *T32 = Tall + 128 * wave;
    const int fr = lane & 15, fg = lane >> 4;
    const uint64_t qi = (uint64_t)M.q, half = qi >> 1;
    const uint64_t off_neg = (uint64_t)GR_SMALL - qi;
    v4i acc[13];
#pragma unroll
    for (int k = 0; k < 13; k++) acc[k] = (v4i){0, 0, 0, 0};
    // super-blocks x = blockIdx.x, + gridDim.x, .. < nbx: sub-blocks sbn x ..
    // sbn x + sbn - 1 of cpt GR_X columns each; those that lie in short tiles
    // are k_sc_gram_short's (the set is uniform).  The challenge tables are
    // made on the first sub-block to do; a block with none returns.
    bool ready = false;
    for (size_t bxi = blockIdx.x; bxi < (size_t)nbx; bxi += gridDim.x) {
    const size_t sb0 = bxi * sbn;
    uint32_t todo = 0;
    for (int k = 0; k < sbn && (sb0 + k) * span < S; k++)
        todo |= 1u << k;
    if (todo == 0) continue;
    if (!ready) {
        gr_coefs<FQ2, SNR>(C, cs0 + row * cs_row, rep0, nrep, nrounds, nr, M, tid);
        ready = true;
    }
    for (int k = 0; k < sbn; k++) {
        if (!((todo >> k) & 1)) continue;
        size_t xb = (sb0 + k) * span;
        // Column x = xb + tid.  Each iteration moves its column into the
        // staging rows and reuses the same registers for the next iteration's
        // 16 loads (unconditional: a dead column re-reads column S - 1 and
        // counts as zero), which stay in flight while the fold and the matrix
        // products read this one back.
        uint64_t v[16];
        {
            const size_t xc = min(xb + tid, S - 1);
#pragma unroll
            for (int y = 0; y < 16; y++) v[y] = wr[(size_t)y * S + xc];
        }
        for (int cc = 0;; cc++) {
            const size_t x = xb + tid;
            const bool live = x < S;
            const bool more = cc + 1 < cpt && xb + GR_X < S;  // uniform
            const size_t xn = min(more ? x + GR_X : x, S - 1);
            uint32_t bad = 0;
#pragma unroll
            for (int y = 0; y < 16; y++) bad |= short_bad(short_t(v[y], half, off_neg));
            const bool small = __builtin_amdgcn_ballot_w64(live && bad != 0) == 0;  // uniform
            // both paths then run the same 49 products (kept unconditional:
            // accumulators that only one branch updates cost copies instead)
            v4i f[7];
            {
                if (small) {
#pragma unroll
                    for (int y = 0; y < 16; y++)
                        T32[ld32 * y + lane] = live ? (uint32_t)short_t(v[y], half, off_neg) : GR_SMALL;
                } else {
#pragma unroll
                    for (int y = 0; y < 16; y++) T[y * GR_LD + tid] = live ? v[y] : 0;
                }
#pragma unroll
                for (int y = 0; y < 16; y++) v[y] = wr[(size_t)y * S + xn];
                wave_sync();
            }
            if (small) {
                if (live) {
                    uint32_t tt[16];
#pragma unroll
                    for (int y = 0; y < 16; y++) tt[y] = T32[ld32 * y + lane];
                    gr_short_fold<FQ2, false>(C, tt, nrep, rep0, rows, row, S, x, out, M);
                }
                if (GRAM) {
                    gr_short_frags(T32, ld32, fr, fg, f[0], f[1]);
#pragma unroll
                    for (int i = 2; i < 7; i++) f[i] = (v4i){0, 0, 0, 0};
                }
            } else {
                if (live) gr_wide_fold<FQ2>(C, T + tid, GR_LD, nrep, rep0, rows, row, S, x, out, M);
                // lane 16 g + r takes digit row (i, r) of the wave's columns
                // 16 g .. 16 g + 15: only the wave has to be in step around T
                if (GRAM) gr_wide_frags(T + fr * GR_LD + 64 * wave + 16 * fg, f);
            }
            wave_sync();  // the rows are rewritten by the next iteration
            if (GRAM) {
#pragma unroll
                for (int i = 0; i < 7; i++)
#pragma unroll
                    for (int j = 0; j < 7; j++)
                        acc[i + j] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(f[i], f[j], acc[i + j], 0, 0, 0);
            }
            if (!more) break;
            xb += GR_X;
        }
    }
    }
    if (!ready) return;
    if (GRAM) {
        __syncthreads();  // the partials reuse the staging rows
        gr_finish<13>(acc, reinterpret_cast<double *>(Tall), wave, lane, tid, tab, gacc, row, M);
    }
}

// blocks per row that add to one u64 Gram sum: 2^14 residues below 2^50
constexpr size_t kGramBlocksMax = 16384;

template <bool FQ2, bool SNR, bool GRAM>
__global__ __launch_bounds__(SC_THREADS) __attribute__((amdgpu_waves_per_eu(3))) void k_sc_gram(
    const uint64_t *__restrict__ w, const unsigned long long *__restrict__ wrow, size_t L,
    int rows, int rep0, int nrep, const uint64_t *__restrict__ cs0, size_t cs_row, int nrounds,
    double nr, int cpt, double *__restrict__ out, unsigned long long *__restrict__ gacc,
    const double *__restrict__ tab, Mod M) {
    __shared__ __align__(16) uint64_t T[16 * GR_LD];
    __shared__ GrCoef C;
    sc_gram_body<FQ2, SNR, GRAM>(reinterpret_cast<uint32_t *>(T), C, w, wrow, L, rows, rep0, nrep,
                                 cs0, cs_row, nrounds, nr, cpt, out, gacc, tab, M, 1,
                                 (int)gridDim.x);
}

// The sub-blocks of a batch arena that k_sc_gram_short leaves (some tile not
// short), on a grid whose blocks walk up to 64 iterations of 256 columns.  It
// returns at once when k_sc_gram_short has flagged none (need_long zero), the
// usual case, and it is kept small in registers so that its blocks are
// dispatched beside the other norm checks' kernels: each column folds from
// registers and adds its 136 Gram products v_y v_y' (y <= y', exact mod q)
// into the block's u64 sums in LDS (at most 2^14 columns per block, so each
// sum of residues below 2^50 stays below 2^64), which go to gacc at the end.
template <bool FQ2, bool SNR, bool GRAM>
__global__ __launch_bounds__(SC_THREADS) void k_sc_gram_rest(
    const uint64_t *__restrict__ w, const unsigned long long *__restrict__ wrow, size_t L,
    int rows, int rep0, int nrep, const uint64_t *__restrict__ cs0, size_t cs_row, int nrounds,
    double nr, int cpt, double *__restrict__ out, unsigned long long *__restrict__ gacc, Mod M,
    const uint8_t *__restrict__ fmt, int sbn, int nbx,
    const unsigned long long *__restrict__ need_long) {
    __shared__ GrCoef C;
    __shared__ unsigned long long Gs[16 * 16];
    if (*need_long == 0) return;  // uniform
    const int row = blockIdx.y, tid = threadIdx.x;
    const size_t S = L >> 4;
    const size_t base = wrow ? wrow[row] : (size_t)row * L;
    const size_t span = (size_t)cpt * GR_X;
    const uint64_t *wr = w + base;
    bool ready = false;
    for (size_t bxi = blockIdx.x; bxi < (size_t)nbx; bxi += gridDim.x) {
        const size_t sb0 = bxi * sbn;
        uint32_t todo = 0;
        for (int k = 0; k < sbn && (sb0 + k) * span < S; k++) {
            const size_t xb = (sb0 + k) * span;
            if (sc_classify<16>(fmt, base, S, xb, min(xb + span, S) - 1) == 2) todo |= 1u << k;
        }
        if (todo == 0) continue;  // uniform
        if (!ready) {
            gr_coefs<FQ2, SNR>(C, cs0 + row * cs_row, rep0, nrep, nrounds, nr, M, tid);
            Gs[tid] = 0ULL;
            __syncthreads();
            ready = true;
        }
        for (int k = 0; k < sbn; k++) {
            if (!((todo >> k) & 1)) continue;
            for (int cc = 0; cc < cpt; cc++) {
                const size_t x = (sb0 + k) * span + (size_t)cc * GR_X + tid;
                if (x >= S) continue;
                uint64_t v[16];
#pragma unroll
                for (int y = 0; y < 16; y++) v[y] = wr[(size_t)y * S + x];
                for (int r = 0; r < nrep; r++) {
                    // products |coef v| <= 0.69 q (balanced coef, canonical v):
                    // eight of them stay below 5.6 q, then two lazy reductions
                    double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
#pragma unroll
                    for (int y = 0; y < 8; y++) {
                        const F2 c0 = C.coef[r][y], c1 = C.coef[r][y + 8];
                        const double v0 = u2d(v[y]), v1 = u2d(v[y + 8]);
                        a0 += mul_lazy(c0.a, v0, M);
                        a1 += mul_lazy(c1.a, v1, M);
                        if (FQ2) {
                            b0 += mul_lazy(c0.b, v0, M);
                            b1 += mul_lazy(c1.b, v1, M);
                        }
                    }
                    double *o = out + (((size_t)(rep0 + r) * rows + row) * S + x) * 2;
                    o[0] = reduce_lazy(reduce_lazy(a0, M) + a1, M);
                    o[1] = FQ2 ? reduce_lazy(reduce_lazy(b0, M) + b1, M) : 0.0;
                }
                if (GRAM) {
                    for (int y = 0; y < 16; y++) {
                        const double vy = u2d(v[y]);
                        for (int y2 = y; y2 < 16; y2++)
                            atomicAdd(&Gs[16 * y + y2],
                                      (unsigned long long)d2u(mulmod(vy, u2d(v[y2]), M)));
                    }
                }
            }
        }
    }
    if (!ready || !GRAM) return;  // uniform
    __syncthreads();
    const int y = tid >> 4, y2 = tid & 15;
    const unsigned long long g = Gs[y <= y2 ? 16 * y + y2 : 16 * y2 + y];
    atomicAdd(&gacc[(size_t)row * 256 + tid], (unsigned long long)d2u(u64_mod(g, M)));
}

// The short copy's blocks: the same columns and outputs as k_sc_gram, read
// as t from the copy (four waves, each staging its 64 columns of 16 rows).
constexpr int GS_LD = GR_X + 4;  // u32 row stride of the staging rows
template <bool FQ2, bool SNR, bool GRAM>
__global__ __launch_bounds__(SC_THREADS) void k_sc_gram_short(
    const unsigned long long *__restrict__ wrow, size_t L, int rows, int rep0, int nrep,
    const uint64_t *__restrict__ cs0, size_t cs_row, int nrounds, double nr, int cpt,
    double *__restrict__ out, unsigned long long *__restrict__ gacc, const double *__restrict__ tab,
    Mod M, const uint16_t *__restrict__ sht, const uint8_t *__restrict__ sht8,
    const uint8_t *__restrict__ fmt, int sbn, unsigned long long *__restrict__ need_long) {
    __shared__ __align__(16) uint32_t T32all[16 * GS_LD];  // then the Gram partials
    __shared__ GrCoef C;
    const int row = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t S = L >> 4;
    const size_t base = wrow ? wrow[row] : (size_t)row * L;
    // the sub-blocks (as in k_sc_gram) whose slices are all of one width
    // (uniform); wide8 bit k: sub-block k reads the two-byte copy
    const size_t span = (size_t)cpt * GR_X, sb0 = (size_t)blockIdx.x * sbn;
    uint32_t todo = 0, wide8 = 0;
    bool left = false;  // a sub-block for k_sc_gram_rest
    for (int k = 0; k < sbn && (sb0 + k) * span < S; k++) {
        const size_t xb = (sb0 + k) * span;
        const int mode = sc_classify<16>(fmt, base, S, xb, min(xb + span, S) - 1);
        if (mode < 2) {
            todo |= 1u << k;
            wide8 |= (uint32_t)mode << k;
        } else {
            left = true;
        }
    }
    if (left && tid == 0) *need_long = 1ULL;
    if (todo == 0) return;
    gr_coefs<FQ2, SNR>(C, cs0 + row * cs_row, rep0, nrep, nrounds, nr, M, tid);
    uint32_t *T32 = T32all + 64 * wave;
    const int fr = lane & 15, fg = lane >> 4;
    const uint16_t *tr = sht + base;
    const uint8_t *tr8 = sht8 + base;
    v4i acc[3] = {(v4i){0, 0, 0, 0}, (v4i){0, 0, 0, 0}, (v4i){0, 0, 0, 0}};
    // row segment y at the uniform tr + y S, the lane's column as a 32-bit
    // offset (one address register for all 16 loads); t = b + 2^14 from
    // either copy
    bool wide = false;  // the current sub-block reads the two-byte copy
    auto fetch = [&](uint32_t(&t)[16], size_t xc) {
        const uint32_t xo = (uint32_t)min(xc, S - 1);
        if (wide) {
#pragma unroll
            for (int y = 0; y < 16; y++) t[y] = (tr + (size_t)y * S)[xo];
        } else {
#pragma unroll
            for (int y = 0; y < 16; y++) t[y] = (uint32_t)(tr8 + (size_t)y * S)[xo] + (kShortOff - 128u);
        }
    };
    for (int k = 0; k < sbn; k++) {
        if (!((todo >> k) & 1)) continue;
        size_t xb = (sb0 + k) * span;
        wide = (wide8 >> k) & 1;
        uint32_t tv[16];
        fetch(tv, xb + tid);
        for (int cc = 0;; cc++) {
            const size_t x = xb + tid;
            const bool live = x < S;
            const bool more = cc + 1 < cpt && xb + GR_X < S;  // uniform
#pragma unroll
            for (int y = 0; y < 16; y++) T32[GS_LD * y + lane] = live ? tv[y] : GR_SMALL;
            fetch(tv, more ? x + GR_X : x);
            wave_sync();
            if (live) {
                uint32_t tt[16];
#pragma unroll
                for (int y = 0; y < 16; y++) tt[y] = T32[GS_LD * y + lane];
                gr_short_fold<FQ2, true>(C, tt, nrep, rep0, rows, row, S, x, out, M);
            }
            if (GRAM) {
                v4i f0, f1;
                gr_short_frags(T32, GS_LD, fr, fg, f0, f1);
                wave_sync();  // T32 is rewritten by the next iteration
                acc[0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(f0, f0, acc[0], 0, 0, 0);
                if (wide) {  // byte slices: b's high digit is 0, E_1 = E_2 = 0
                    acc[1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(f0, f1, acc[1], 0, 0, 0);
                    acc[1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(f1, f0, acc[1], 0, 0, 0);
                    acc[2] = __builtin_amdgcn_mfma_i32_16x16x64_i8(f1, f1, acc[2], 0, 0, 0);
                }
            } else {
                wave_sync();
            }
            if (!more) break;
            xb += GR_X;
        }
    }
    if (GRAM) {
        __syncthreads();  // the partials reuse the staging rows
        gr_finish<3>(acc, reinterpret_cast<double *>(T32all), wave, lane, tid, tab, gacc, row, M);
    }
}

// ---- six rounds per first pass (batch arenas with a short copy).  With 64
// row segments (S = L / 64, v[y S + x] for y < 64) rounds 0..5 follow from the
// row's 64 x 64 Gram matrix exactly as rounds 0..3 follow from the 16 x 16
// one (the fold vectors of round t < 6 have the closed form
// A_t[z][y] = [y mod 64/2^t = z] e_t(y), e_t(y) = prod_{s<t} (bit 5 - s of y ?
// c_s : 1 - c_s)), and the witness after round 5 is the single 64-term
// combination sum_y e_6(y) v[y S + x].  So the first pass over the short copy
// costs what the 16-segment one does per coefficient (the same short fold,
// with 64 terms) and no F_{q^2} round work, and writes L / 64 values per
// repetition instead of L / 16.  The Gram matrix is kept as its 10 upper 16 x
// 16 tile pairs (a <= b), gacc[row][p][256].
constexpr int G64_MAXREP = 2;    // repetitions folded per launch
constexpr int G64_PAIRS = 10;    // tile pairs (a <= b) of the 64 x 64 Gram matrix
constexpr size_t kGram64Entries = (size_t)G64_PAIRS * 256;

__device__ __forceinline__ void g64_pair(int p, int &a, int &b) {
    // p -> (a, b), a <= b < 4, in the order (0,0) (0,1) (0,2) (0,3) (1,1) ...
    a = p < 4 ? 0 : p < 7 ? 1 : p < 9 ? 2 : 3;
    b = a + (p - (a == 0 ? 0 : a == 1 ? 4 : a == 2 ? 7 : 9));
}

// The block's fold weights e_6(y) for repetitions rep0 .. rep0 + nrep - 1,
// balanced (spare repetition slots zero).
struct G64Coef {
    F2 coef[G64_MAXREP][64];
};
template <bool FQ2, bool SNR>
__device__ __forceinline__ void g64_coefs(G64Coef &C, const uint64_t *cs, int rep0, int nrep,
                                          int nrounds, double nr, Mod M, int tid) {
    if (tid < 64 * G64_MAXREP) {
        const int r = tid >> 6, y = tid & 63;
        F2 a = {0.0, 0.0};
        if (r < nrep) {
            a = {1.0, 0.0};
            for (int t = 0; t < 6; t++) {
                const F2 c = load_chal<FQ2>(cs, rep0 + r, nrounds, t);
                a = x_mul<FQ2, SNR>(a, ((y >> (5 - t)) & 1) ? c : x_sub<FQ2>(F2{1.0, 0.0}, c, M), nr, M);
            }
        }
        C.coef[r][y] = {balance(a.a, M), balance(a.b, M)};
    }
    __syncthreads();
}

// Digits 0 and 1 of 16 consecutive staged entries z = b + 128 (int16, b the
// balanced residue) as matrix-core fragments: the low byte of z is d0 + 128,
// the high byte d1.
__device__ __forceinline__ void g64_frags(const uint16_t *src, v4i &f0, v4i &f1) {
    const uint4 *p = reinterpret_cast<const uint4 *>(src);
    const uint4 u0 = p[0], u1 = p[1];
    const uint32_t w[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        f0[k] = (int)(__builtin_amdgcn_perm(w[2 * k + 1], w[2 * k], 0x06040200u) ^ 0x80808080u);
        f1[k] = (int)__builtin_amdgcn_perm(w[2 * k + 1], w[2 * k], 0x07050301u);
    }
}

// The fold weights of one block as matrix-core A fragments: row 4 g + e of
// tile mt is digit u = 4 mt + e (u < 7; row u = 7 is zero) of the weight of
// (repetition, component) g = 2 r + comp, the K index the segment y.  Lane
// (row, kg) holds segments 16 kg .. 16 kg + 15.  A balanced weight w
// (|w| < 2^49) has w + 0x0080808080808080 in [0, 2^56) as an integer, whose
// byte u is a_u + 128 for the signed digits w = sum_{u<7} a_u 256^u.
// PERM: K position p = 16 kg + 4 k4 + bq holds segment bq + 4 kg + 16 k4
// (the byte kernel's column order) instead of segment p.
template <bool PERM = false>
__device__ __forceinline__ void g64_wfrags(const G64Coef &C, int lane, v4i (&A)[2]) {
    const int row = lane & 15, kg = lane >> 4, g = row >> 2, e = row & 3;
    const int r = g >> 1, comp = g & 1;
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
        const int u = 4 * mt + e;
#pragma unroll
        for (int k4 = 0; k4 < 4; k4++) {
            uint32_t word = 0;
#pragma unroll
            for (int bq = 0; bq < 4; bq++) {
                const int y = PERM ? bq + 4 * kg + 16 * k4 : 16 * kg + 4 * k4 + bq;
                const double w = comp ? C.coef[r][y].b : C.coef[r][y].a;
                const uint64_t t = (uint64_t)(long long)w + GR_BIAS;
                const uint32_t d = u < 7 ? (((uint32_t)(t >> (8 * u)) & 0xFFu) ^ 0x80u) : 0u;
                word |= d << (8 * bq);
            }
            A[mt][k4] = (int)word;
        }
    }
}

// The first pass over the short copy with 64 segments: the sub-blocks of cpt
// GR_X columns x < S whose slices in all 64 segments are kept two bytes wide
// (k_sc_gram64_byte takes the byte-wide ones, k_sc_gram64_rest the rest).  Per
// iteration, wave w stages the 64 segments of columns 64 w .. 64 w + 63 (one
// per lane, prefetched a whole iteration ahead) twice: into the block's rows
// (segment-major, for the Gram matrix) and into its own columns
// (column-major, for the fold), as z = b + 128, whose low byte xor 128 and
// high byte are the signed base-256 digits of b.  The fold of 16 columns is
// one set of matrix-core products: the 7 digits of the weights of both
// repetitions and components (A, 32 rows) times the 2 digits of b (B),
// K = the 64 segments, each product sum exact in int32 (|sum| <= 2^20); lane
// (n, g) then holds all 14 digit sums of one output, (repetition,
// component) g of column n, and recombines them with the weights 2^(8k) mod
// q.  After a block barrier wave w multiplies the Gram tile pairs p = w,
// w + 4, w + 8 over all 256 columns (so each wave keeps at most 3 pairs of
// E_0..E_2 accumulators), and a second barrier frees the rows.  Each wave
// adds its pairs' sums to gacc at the end.
constexpr int G64_LD = GR_X + 8;  // u16 row stride of the segment-major rows
constexpr int G64_CLD = 64 + 8;   // u16 stride of the column-major columns
template <bool FQ2, bool SNR, bool GRAM>
__global__ __launch_bounds__(SC_THREADS) __attribute__((amdgpu_waves_per_eu(2))) void k_sc_gram64_short(
    const unsigned long long *__restrict__ wrow, size_t L, int rows, int rep0, int nrep,
    const uint64_t *__restrict__ cs0, size_t cs_row, int nrounds, double nr, int cpt,
    double *__restrict__ out, unsigned long long *__restrict__ gacc, const double *__restrict__ tab,
    Mod M, const uint16_t *__restrict__ sht, const uint8_t *__restrict__ fmt, int sbn,
    unsigned long long *__restrict__ need_long) {
    __shared__ __align__(16) uint16_t T16[64 * G64_LD];   // T16[y][c]: segment y, block column c
    __shared__ __align__(16) uint16_t TC[GR_X * G64_CLD];  // TC[c][y]
    __shared__ G64Coef C;
    const int row = blockIdx.y, tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const size_t S = L >> 6;
    const size_t base = wrow ? wrow[row] : (size_t)row * L;
    const size_t span = (size_t)cpt * GR_X, sb0 = (size_t)blockIdx.x * sbn;
    uint32_t todo = 0;
    bool left = false;  // a sub-block for k_sc_gram64_rest
    for (int k = 0; k < sbn && (sb0 + k) * span < S; k++) {
        const size_t xb = (sb0 + k) * span;
        const int mode = sc_classify<64>(fmt, base, S, xb, min(xb + span, S) - 1);
        if (mode == 1)
            todo |= 1u << k;
        else if (mode == 2)
            left = true;  // (mode 0: k_sc_gram64_byte's)
    }
    if (left && tid == 0) *need_long = 1ULL;
    if (todo == 0) return;
    g64_coefs<FQ2, SNR>(C, cs0 + row * cs_row, rep0, nrep, nrounds, nr, M, tid);
    const int fr = lane & 15, fg = lane >> 4;
    v4i Aw[2];
    g64_wfrags(C, lane, Aw);
    const double *wt = tab + TAB_W256;
    // this lane's output of the fold: (repetition, component) fg and its
    // offset in the [rep][row][S][2] output
    const int fo_r = fg >> 1, fo_c = fg & 1;
    const bool fo_live = fo_r < nrep && (FQ2 || fo_c == 0);
    double *fo_base = out + ((size_t)(rep0 + fo_r) * rows + row) * S * 2 + fo_c;
    const uint16_t *tr = sht + base;
    v4i acc[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) acc[i][k] = (v4i){0, 0, 0, 0};
    // entry (y, x) at the uniform tr plus the 32-bit offset y S + x (S <= 2^26:
    // one base address, not 64 of them in scalar registers)
    const uint32_t S32 = (uint32_t)S;
    auto fetch = [&](uint32_t(&t)[64], size_t xc) {
        const uint32_t xo = (uint32_t)min(xc, S - 1);
#pragma unroll
        for (int y = 0; y < 64; y++) t[y] = tr[(uint32_t)y * S32 + xo];
    };
    uint16_t *mycol = T16 + 64 * wave + lane;
    uint16_t *myTC = TC + (64 * wave + lane) * G64_CLD;
    for (int k = 0; k < sbn; k++) {
        if (!((todo >> k) & 1)) continue;
        size_t xb = (sb0 + k) * span;  // the block's columns (uniform)
        uint32_t tv[64];
        fetch(tv, xb + 64 * wave + lane);
        for (int cc = 0;; cc++) {
            const size_t x = xb + 64 * wave + lane;
            const bool live = x < S;
            const bool more = cc + 1 < cpt && xb + GR_X < S;  // uniform in the block
            // z = t - 2^14 + 128 = b + 128 (a dead column stages b = 0), into
            // the segment-major rows and the lane's own column
#pragma unroll
            for (int y = 0; y < 64; y++) tv[y] = live ? tv[y] - (GR_SMALL - 128u) : 128u;
#pragma unroll
            for (int y = 0; y < 64; y++) mycol[G64_LD * y] = (uint16_t)tv[y];
#pragma unroll
            for (int q8 = 0; q8 < 8; q8++) {
                uint4 v;
                v.x = (tv[8 * q8 + 0] & 0xFFFFu) | (tv[8 * q8 + 1] << 16);
                v.y = (tv[8 * q8 + 2] & 0xFFFFu) | (tv[8 * q8 + 3] << 16);
                v.z = (tv[8 * q8 + 4] & 0xFFFFu) | (tv[8 * q8 + 5] << 16);
                v.w = (tv[8 * q8 + 6] & 0xFFFFu) | (tv[8 * q8 + 7] << 16);
                reinterpret_cast<uint4 *>(myTC)[q8] = v;
            }
            fetch(tv, more ? x + GR_X : x);
            wave_sync();  // the wave's columns are in TC
            // the fold of the wave's 64 columns, 16 at a time
#pragma unroll
            for (int j = 0; j < 4; j++) {
                v4i b0, b1;
                g64_frags(TC + (64 * wave + 16 * j + fr) * G64_CLD + 16 * fg, b0, b1);
                const v4i z4 = (v4i){0, 0, 0, 0};
                const v4i s00 = __builtin_amdgcn_mfma_i32_16x16x64_i8(Aw[0], b0, z4, 0, 0, 0);
                const v4i s01 = __builtin_amdgcn_mfma_i32_16x16x64_i8(Aw[0], b1, z4, 0, 0, 0);
                const v4i s10 = __builtin_amdgcn_mfma_i32_16x16x64_i8(Aw[1], b0, z4, 0, 0, 0);
                const v4i s11 = __builtin_amdgcn_mfma_i32_16x16x64_i8(Aw[1], b1, z4, 0, 0, 0);
                // digit sums S[u][v]: u < 4 from tile 0 (s00, s01), u = 4..6
                // from tile 1 (s10, s11); E_k = sum_{u + v = k} S[u][v]
                const double E[8] = {(double)s00[0],
                                     (double)s00[1] + (double)s01[0],
                                     (double)s00[2] + (double)s01[1],
                                     (double)s00[3] + (double)s01[2],
                                     (double)s10[0] + (double)s01[3],
                                     (double)s10[1] + (double)s11[0],
                                     (double)s10[2] + (double)s11[1],
                                     (double)s11[2]};
                // the digits of z - 128 = b: sum_k E_k 2^(8k) = sum_y w(y) b(y)
                double v = 0.0;
#pragma unroll
                for (int kk = 0; kk < 8; kk++) {
                    v += mul_lazy(E[kk], wt[kk], M);
                    if (kk == 3) v = reduce_lazy(v, M);
                }
                const size_t xo = xb + 64 * wave + 16 * j + fr;
                if (fo_live && xo < S) fo_base[xo * 2] = reduce_lazy(v, M);
            }
            if (GRAM) {
                __syncthreads();  // every wave's columns are staged
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const int q = wave + 4 * i;  // wave-uniform
                    if (q >= G64_PAIRS) break;
                    int a, b;
                    g64_pair(q, a, b);
#pragma unroll
                    for (int k4 = 0; k4 < 4; k4++) {
                        v4i da0, da1, db0, db1;
                        g64_frags(T16 + G64_LD * (16 * a + fr) + 64 * k4 + 16 * fg, da0, da1);
                        g64_frags(T16 + G64_LD * (16 * b + fr) + 64 * k4 + 16 * fg, db0, db1);
                        acc[i][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(da0, db0, acc[i][0], 0, 0, 0);
                        acc[i][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(da0, db1, acc[i][1], 0, 0, 0);
                        acc[i][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(da1, db0, acc[i][1], 0, 0, 0);
                        acc[i][2] = __builtin_amdgcn_mfma_i32_16x16x64_i8(da1, db1, acc[i][2], 0, 0, 0);
                    }
                }
                __syncthreads();  // the rows are rewritten by the next iteration
            } else {
                wave_sync();  // the wave's TC columns are rewritten by the next iteration
            }
            if (!more) break;
            xb += GR_X;
        }
    }
    if (GRAM) {
        // lane (r, g) holds E_k[4 g + e][r] of each of the wave's pairs
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int q = wave + 4 * i;
            if (q >= G64_PAIRS) break;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                double xg = 0.0;
#pragma unroll
                for (int k = 0; k < 3; k++) xg += mul_lazy((double)acc[i][k][e], wt[k], M);
                atomicAdd(&gacc[((size_t)row * G64_PAIRS + q) * 256 + (4 * fg + e) * 16 + fr],
                          (unsigned long long)d2u(reduce(xg, M)));
            }
        }
    }
}

// The same first pass over the sub-blocks whose slices are all one byte wide
// (b in [-128, 127]: one signed digit, so the Gram matrix is E_0 alone and
// the fold is the weights' 7 digits times b).  S = L / 64 is a power of two
// of at least 1024, so every iteration's 256 columns lie in the row.
// Lane (sg, cg) = (lane & 3, lane >> 2) of wave w loads columns
// 64 w + 4 cg .. + 3 (one u32) of the 16 segments y_i = (i & 3) + 4 sg +
// 16 (i >> 2), from one uniform base per i plus a 32-bit lane offset, and
// stages them as signed bytes twice: into the rows T8[y][c] (one u32 store
// per segment) and, after a 4 x 4 byte transpose, into the columns
// TC8[c][16 sg + i] (one 16-byte store per column).  Column position
// p = 16 sg + i holds segment y_i, so the fold's weights take K in that order
// (g64_wfrags<true>); both are matrix-core fragments as they lie in LDS.
// The fold's digit sums recombine exactly in int64 (|sum_y w(y) b(y)| <
// 64 2^49 2^7 = 2^62), and the Gram tile sums stay below 2^31 (at most 2^14
// columns per block).  Row strides are chosen so that the stores and the
// fragment reads of 16 lanes meet distinct LDS banks.
constexpr int GB_LD = GR_X + 16;  // byte stride of the segment rows (68 words)
constexpr int GB_CLD = 64 + 16;   // byte stride of the columns (20 words)
template <bool FQ2, bool SNR, bool GRAM>
__global__ __launch_bounds__(SC_THREADS) void k_sc_gram64_byte(
    const unsigned long long *__restrict__ wrow, size_t L, int rows, int rep0, int nrep,
    const uint64_t *__restrict__ cs0, size_t cs_row, int nrounds, double nr, int cpt,
    double *__restrict__ out, unsigned long long *__restrict__ gacc, const double *__restrict__ tab,
    Mod M, const uint8_t *__restrict__ sht8, const uint8_t *__restrict__ fmt, int sbn) {
    __shared__ __align__(16) uint8_t T8[64 * GB_LD];
    __shared__ __align__(16) uint8_t TC8[GR_X * GB_CLD];
    __shared__ G64Coef C;
    const int row = blockIdx.y, tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const size_t S = L >> 6;
    const size_t base = wrow ? wrow[row] : (size_t)row * L;
    const size_t span = (size_t)cpt * GR_X, sb0 = (size_t)blockIdx.x * sbn;
    uint32_t todo = 0;
    for (int k = 0; k < sbn && (sb0 + k) * span < S; k++) {
        const size_t xb = (sb0 + k) * span;
        if (sc_classify<64>(fmt, base, S, xb, min(xb + span, S) - 1) == 0) todo |= 1u << k;
    }
    if (todo == 0) return;
    g64_coefs<FQ2, SNR>(C, cs0 + row * cs_row, rep0, nrep, nrounds, nr, M, tid);
    const int fr = lane & 15, fg = lane >> 4;
    const int sg = lane & 3, cg = lane >> 2;
    v4i Aw[2];
    g64_wfrags<true>(C, lane, Aw);
    const double w32 = tab[TAB_W256 + 4];  // 2^32 mod q
    const int fo_r = fg >> 1, fo_c = fg & 1;
    const bool fo_live = fo_r < nrep && (FQ2 || fo_c == 0);
    double *fo_base = out + ((size_t)(rep0 + fo_r) * rows + row) * S * 2 + fo_c;
    const uint8_t *tr8 = sht8 + base;
    const uint32_t S32 = (uint32_t)S;
    const uint32_t wcol = 64u * wave + 4u * cg;  // the lane's first column in the block
    v4i acc[3];
#pragma unroll
    for (int i = 0; i < 3; i++) acc[i] = (v4i){0, 0, 0, 0};
    // the 16 segments' u32 at block column xb
    auto fetch = [&](uint32_t(&r)[16], uint32_t xb) {
        const uint32_t off = 4u * sg * S32 + xb + wcol;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint8_t *sb = tr8 + (size_t)((i & 3) + 16 * (i >> 2)) * S;
            r[i] = *reinterpret_cast<const uint32_t *>(sb + off);
        }
    };
    uint8_t *t8w = T8 + GB_LD * 4 * sg + wcol;
    uint8_t *tcw = TC8 + wcol * GB_CLD + 16 * sg;
    // the iterations of every sub-block to do, in one loop, so that each
    // iteration's loads (the next sub-block's first ones included) are in
    // flight while the one before is staged and multiplied
    int k = __builtin_ctz(todo), cc = 0;
    uint32_t xb = (uint32_t)((sb0 + k) * span);  // the block's columns (uniform)
    uint32_t rv[16];
    fetch(rv, xb);
    {
        for (;;) {
            // the next iteration (uniform in the block)
            int nk = k, ncc = cc + 1;
            uint32_t nxb = xb + GR_X;
            if (ncc >= cpt || nxb >= S32) {
                nk = k + 1;
                while (nk < sbn && !((todo >> nk) & 1)) nk++;
                ncc = 0;
                nxb = nk < sbn ? (uint32_t)((sb0 + nk) * span) : xb;
            }
            const bool more = nk < sbn;
            // t8 = b + 128 per byte -> b as a signed byte
            uint32_t d[16];
#pragma unroll
            for (int i = 0; i < 16; i++) d[i] = rv[i] ^ 0x80808080u;
            fetch(rv, nxb);
            if (GRAM)
#pragma unroll
                for (int i = 0; i < 16; i++)
                    *reinterpret_cast<uint32_t *>(t8w + GB_LD * ((i & 3) + 16 * (i >> 2))) = d[i];
            // columns: o[c][q] = bytes c of d[4 q .. 4 q + 3]
            uint32_t o[4][4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t l01 = __builtin_amdgcn_perm(d[4 * q + 1], d[4 * q], 0x05010400u);
                const uint32_t h01 = __builtin_amdgcn_perm(d[4 * q + 1], d[4 * q], 0x07030602u);
                const uint32_t l23 = __builtin_amdgcn_perm(d[4 * q + 3], d[4 * q + 2], 0x05010400u);
                const uint32_t h23 = __builtin_amdgcn_perm(d[4 * q + 3], d[4 * q + 2], 0x07030602u);
                o[0][q] = __builtin_amdgcn_perm(l23, l01, 0x05040100u);
                o[1][q] = __builtin_amdgcn_perm(l23, l01, 0x07060302u);
                o[2][q] = __builtin_amdgcn_perm(h23, h01, 0x05040100u);
                o[3][q] = __builtin_amdgcn_perm(h23, h01, 0x07060302u);
            }
#pragma unroll
            for (int c = 0; c < 4; c++)
                *reinterpret_cast<uint4 *>(tcw + c * GB_CLD) = make_uint4(o[c][0], o[c][1], o[c][2], o[c][3]);
            wave_sync();  // the wave's columns are in TC8
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const v4i b = *reinterpret_cast<const v4i *>(TC8 + (64 * wave + 16 * j + fr) * GB_CLD + 16 * fg);
                const v4i z4 = (v4i){0, 0, 0, 0};
                const v4i s0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(Aw[0], b, z4, 0, 0, 0);
                const v4i s1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(Aw[1], b, z4, 0, 0, 0);
                // sum_k E_k 2^(8k), exact in wrapping int64 (the sum fits)
                const unsigned long long tot =
                    (unsigned long long)(long long)s0[0] + ((unsigned long long)(long long)s0[1] << 8) +
                    ((unsigned long long)(long long)s0[2] << 16) +
                    ((unsigned long long)(long long)s0[3] << 24) +
                    ((unsigned long long)(long long)s1[0] << 32) +
                    ((unsigned long long)(long long)s1[1] << 40) +
                    ((unsigned long long)(long long)s1[2] << 48);
                const long long ts = (long long)tot;
                const double v = mul_lazy((double)(ts >> 32), w32, M) + (double)(uint32_t)tot;
                const size_t xo = (size_t)xb + 64 * wave + 16 * j + fr;
                if (fo_live) fo_base[xo * 2] = reduce_lazy(v, M);
            }
            if (GRAM) {
                __syncthreads();  // every wave's rows are staged
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const int q = wave + 4 * i;  // wave-uniform
                    if (q >= G64_PAIRS) break;
                    int a, bq;
                    g64_pair(q, a, bq);
#pragma unroll
                    for (int k4 = 0; k4 < 4; k4++) {
                        const v4i da = *reinterpret_cast<const v4i *>(T8 + GB_LD * (16 * a + fr) + 64 * k4 + 16 * fg);
                        const v4i db = *reinterpret_cast<const v4i *>(T8 + GB_LD * (16 * bq + fr) + 64 * k4 + 16 * fg);
                        acc[i] = __builtin_amdgcn_mfma_i32_16x16x64_i8(da, db, acc[i], 0, 0, 0);
                    }
                }
                __syncthreads();  // the rows are rewritten by the next iteration
            } else {
                wave_sync();  // the wave's columns are rewritten by the next iteration
            }
            if (!more) break;
            k = nk;
            cc = ncc;
            xb = nxb;
        }
    }
    if (GRAM) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int q = wave + 4 * i;
            if (q >= G64_PAIRS) break;
#pragma unroll
            for (int e = 0; e < 4; e++)
                atomicAdd(&gacc[((size_t)row * G64_PAIRS + q) * 256 + (4 * fg + e) * 16 + fr],
                          (unsigned long long)d2u(reduce((double)acc[i][e], M)));
        }
    }
}

// The sub-blocks k_sc_gram64_short leaves (some tile not short), from the
// witness, on a grid whose blocks walk up to 64 iterations of 64 columns.  It
// returns at once when none was flagged (need_long zero), the usual case.
// Per iteration the block stages 64 segments x 64 columns of residues in LDS;
// threads r * 64 + c (r < nrep) fold column c with the 64 balanced weights,
// and thread e owns Gram entries e + 256 k (k < 10: tile pair k) as u64 sums
// of exact products (at most 4096 residues below 2^50 per block).
constexpr int G64R_LD = 64 + 1;
template <bool FQ2, bool SNR, bool GRAM>
__global__ __launch_bounds__(SC_THREADS) void k_sc_gram64_rest(
    const uint64_t *__restrict__ w, const unsigned long long *__restrict__ wrow, size_t L,
    int rows, int rep0, int nrep, const uint64_t *__restrict__ cs0, size_t cs_row, int nrounds,
    double nr, int cpt, double *__restrict__ out, unsigned long long *__restrict__ gacc, Mod M,
    const uint8_t *__restrict__ fmt, int sbn, int nbx,
    const unsigned long long *__restrict__ need_long) {
    __shared__ G64Coef C;
    __shared__ uint64_t V[64 * G64R_LD];  // V[y][c]: segment y, column xb + c
    if (*need_long == 0) return;  // uniform
    const int row = blockIdx.y, tid = threadIdx.x;
    const size_t S = L >> 6;
    const size_t base = wrow ? wrow[row] : (size_t)row * L;
    const size_t span = (size_t)cpt * GR_X;
    const uint64_t *wr = w + base;
    unsigned long long gs[G64_PAIRS];
#pragma unroll
    for (int q = 0; q < G64_PAIRS; q++) gs[q] = 0ULL;
    bool ready = false;
    for (size_t bxi = blockIdx.x; bxi < (size_t)nbx; bxi += gridDim.x) {
        const size_t sb0 = bxi * sbn;
        uint32_t todo = 0;
        for (int k = 0; k < sbn && (sb0 + k) * span < S; k++) {
            const size_t xb = (sb0 + k) * span;
            if (sc_classify<64>(fmt, base, S, xb, min(xb + span, S) - 1) == 2) todo |= 1u << k;
        }
        if (todo == 0) continue;  // uniform
        if (!ready) {
            g64_coefs<FQ2, SNR>(C, cs0 + row * cs_row, rep0, nrep, nrounds, nr, M, tid);
            ready = true;
        }
        for (int k = 0; k < sbn; k++) {
            if (!((todo >> k) & 1)) continue;
            for (size_t xb = (sb0 + k) * span; xb < min((sb0 + k + 1) * span, S); xb += 64) {
                __syncthreads();  // V is free
                for (int e = tid; e < 64 * 64; e += SC_THREADS) {
                    const int y = e >> 6, c = e & 63;
                    V[y * G64R_LD + c] = xb + c < S ? wr[(size_t)y * S + xb + c] : 0;
                }
                __syncthreads();
                const int c = tid & 63, r = tid >> 6;
                if (r < nrep && xb + c < S) {
                    // products |coef v| <= 0.69 q (balanced coef, canonical v):
                    // eight of them stay below 5.6 q, then a lazy reduction
                    double a = 0.0, bsum = 0.0;
                    for (int y0 = 0; y0 < 64; y0 += 8) {
                        double a8 = 0.0, b8 = 0.0;
#pragma unroll
                        for (int y = y0; y < y0 + 8; y++) {
                            const double v = u2d(V[y * G64R_LD + c]);
                            a8 += mul_lazy(C.coef[r][y].a, v, M);
                            if (FQ2) b8 += mul_lazy(C.coef[r][y].b, v, M);
                        }
                        a = reduce_lazy(a + reduce_lazy(a8, M), M);
                        bsum = reduce_lazy(bsum + reduce_lazy(b8, M), M);
                    }
                    double *o = out + (((size_t)(rep0 + r) * rows + row) * S + xb + c) * 2;
                    o[0] = a;
                    o[1] = FQ2 ? bsum : 0.0;
                }
                if (GRAM) {
#pragma unroll
                    for (int q = 0; q < G64_PAIRS; q++) {
                        int pa, pb;
                        g64_pair(q, pa, pb);
                        const int y = 16 * pa + (tid >> 4), y2 = 16 * pb + (tid & 15);
                        for (int cc = 0; cc < 64; cc++)
                            gs[q] += (unsigned long long)d2u(
                                mulmod(u2d(V[y * G64R_LD + cc]), u2d(V[y2 * G64R_LD + cc]), M));
                    }
                }
            }
        }
    }
    if (!ready || !GRAM) return;  // uniform
#pragma unroll
    for (int q = 0; q < G64_PAIRS; q++)
        atomicAdd(&gacc[((size_t)row * G64_PAIRS + q) * 256 + tid],
                  (unsigned long long)d2u(u64_mod(gs[q], M)));
}

// Messages of rounds 0..3 (and the claim) from G, one block per (row, rep):
// thread (z, y) (z < 8, y < 16) forms column y of G A[z] and its share of
// both forms; the shares are summed across the block.
// ---------------------------------------------------------------- later passes
// F2 input in [reps][rows][len] (pairs of doubles), rounds t0..t0+3.
template <bool FQ2, bool SNR>
__global__ __launch_bounds__(SC_THREADS) void k_sc_pass(
    const double *__restrict__ in, size_t len, int rows, int t0, const uint64_t *__restrict__ cs0,
    size_t cs_row, int nrounds, double nr, int cpt, double *__restrict__ out,
    unsigned long long *__restrict__ macc, Mod M) {
    __shared__ double red[8 * 64];
    const int row = blockIdx.y, rep = blockIdx.z;
    const uint64_t *cs = cs0 + row * cs_row;
    const size_t S = len >> 4;
    F2 c[4];
    double csum[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        c[t] = load_chal_bal<FQ2>(cs, rep, nrounds, t0 + t, M);
        csum[t] = c[t].a + c[t].b;
    }
    F2 acc[4][2];  // (f0, f1) per round; f2 follows from the round identity
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int f = 0; f < 2; f++) acc[t][f] = {0.0, 0.0};
    const double *ib = in + ((size_t)rep * rows + row) * len * 2;
    for (int cc = 0; cc < cpt; cc++) {
        const size_t x = ((size_t)blockIdx.x * cpt + cc) * SC_THREADS + threadIdx.x;
        if (x >= S) break;
        F2 cur[16];
#pragma unroll
        for (int y = 0; y < 16; y++) {
            const double *p = ib + ((size_t)y * S + x) * 2;
            cur[y].a = p[0];
            cur[y].b = FQ2 ? p[1] : 0.0;
        }
        // input and output are balanced residues (written by the previous pass)
        sc_level<FQ2, SNR, 8>(cur, c[0], csum[0], acc[0][0], acc[0][1], nr, M);
        sc_level<FQ2, SNR, 4>(cur, c[1], csum[1], acc[1][0], acc[1][1], nr, M);
        sc_level<FQ2, SNR, 2>(cur, c[2], csum[2], acc[2][0], acc[2][1], nr, M);
        sc_level<FQ2, SNR, 1>(cur, c[3], csum[3], acc[3][0], acc[3][1], nr, M);
        double *o = out + (((size_t)rep * rows + row) * S + x) * 2;
        o[0] = cur[0].a;
        o[1] = cur[0].b;
    }
    double vals[16];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int f = 0; f < 2; f++) {
            vals[t * 4 + f * 2] = reduce(acc[t][f].a, M);
            vals[t * 4 + f * 2 + 1] = reduce(acc[t][f].b, M);
        }
    block_sum<16>(vals, red, M);
    const int tid = threadIdx.x;
    if (tid < 16) {
        const int t = tid / 4, k = tid % 4;
        if (FQ2 || (k & 1) == 0)
            atomic_add_res(&macc[macc_idx(rep, t0 + t, row, nrounds, rows) + k], red[tid]);
    }
}

// Rounds 0..3's messages of (row, rep) from the row's Gram sums (graw: this
// thread's entry, gacc[row][tid]) into
// msg[0..3] (f0, f1), and the claim tr G into *claim (row 0's plus one with
// claim_fault, a planted error).  The fold vectors have a closed form: after
// t rounds (n = 16 >> t segments, h = n / 2), A_t[z][y] = [y mod n = z] e_t(y)
// with e_t(y) = prod_{s<t} (bit 3 - s of y ? c_s : 1 - c_s), so
//     f0(t) = sum over y = y' (mod n), y' mod n < h of e_t(y) G[y][y'] e_t(y'),
//     f1(t) = sum over y = y' + h (mod n), y' mod n < h of the same - f0(t),
// and every thread takes one Gram entry for all four rounds at once.
template <bool FQ2, bool SNR>
__device__ __forceinline__ void gram_msgs(unsigned long long graw, int row, const F2 *chal,
                                          double (*msg)[4], double *claim, int claim_fault,
                                          double *red, double nr, Mod M) {
    static_assert(SC_THREADS == 256, "one thread per Gram entry");
    __shared__ F2 E[4][16];
    const int tid = threadIdx.x, y = tid >> 4, y2 = tid & 15;
    const double g = u64_mod(graw, M);
    if (tid < 64) {
        const int t = tid >> 4, yy = tid & 15;
        F2 e = {1.0, 0.0};
        for (int s = 0; s < t; s++) {
            const F2 c = chal[s];
            e = x_mul<FQ2, SNR>(e, ((yy >> (3 - s)) & 1) ? c : x_sub<FQ2>(F2{1.0, 0.0}, c, M), nr, M);
        }
        E[t][yy] = e;
    }
    __syncthreads();
    double vals[17];
#pragma unroll
    for (int k = 0; k < 17; k++) vals[k] = 0.0;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int n = 16 >> t, h = n >> 1, ry = y & (n - 1), ry2 = y2 & (n - 1);
        if (ry2 < h && (ry == ry2 || ry == ry2 + h)) {
            const F2 p = x_mul<FQ2, SNR>(E[t][y], E[t][y2], nr, M);
            const int k = 4 * t + (ry == ry2 ? 0 : 2);
            vals[k] = mulmod(p.a, g, M);
            vals[k + 1] = FQ2 ? mulmod(p.b, g, M) : 0.0;
        }
    }
    vals[16] = y == y2 ? g : 0.0;  // tr G
    block_sum<17>(vals, red, M);
    if (tid < 4) {
        const double f0a = red[4 * tid], f0b = red[4 * tid + 1];
        msg[tid][0] = f0a;
        msg[tid][1] = f0b;
        msg[tid][2] = submod(red[4 * tid + 2], f0a, M);
        msg[tid][3] = submod(red[4 * tid + 3], f0b, M);
    }
    if (tid == 0) *claim = claim_fault && row == 0 ? addmod(red[16], 1.0, M) : red[16];
}

// Rounds 0..5's messages of (row, rep) from the row's 64 x 64 Gram sums (g:
// this thread's entry tid of each tile pair, reduced) into msg[0..5] and the
// claim tr G, as gram_msgs does for 16 segments: after t rounds (n = 64 >> t,
// h = n / 2) round t pairs y, y' with y = y' or y = y' + h (mod n), y' mod n <
// h, weighted by e_t(y) e_t(y'), e_t(y) = prod_{s<t} (bit 5 - s of y ? c_s :
// 1 - c_s).  A pair a < b stands for both G[y][y'] and G[y'][y].
template <bool FQ2, bool SNR>
__device__ __forceinline__ void gram64_msgs(const double (&g)[G64_PAIRS], int row, const F2 *chal,
                                            double (*msg)[4], double *claim, int claim_fault,
                                            double *red, double nr, Mod M) {
    static_assert(SC_THREADS == 256, "one thread per tile entry");
    __shared__ F2 E[6][64];
    const int tid = threadIdx.x, r16 = tid >> 4, c16 = tid & 15;
    for (int k = tid; k < 6 * 64; k += SC_THREADS) {
        const int t = k >> 6, yy = k & 63;
        F2 e = {1.0, 0.0};
        for (int s = 0; s < t; s++) {
            const F2 c = chal[s];
            e = x_mul<FQ2, SNR>(e, ((yy >> (5 - s)) & 1) ? c : x_sub<FQ2>(F2{1.0, 0.0}, c, M), nr, M);
        }
        E[t][yy] = e;
    }
    __syncthreads();
    double vals[25];
#pragma unroll
    for (int k = 0; k < 25; k++) vals[k] = 0.0;
    for (int q = 0; q < G64_PAIRS; q++) {
        int a, b;
        g64_pair(q, a, b);
        const double gv = g[q];
        for (int o = 0; o < (a == b ? 1 : 2); o++) {
            const int y = o ? 16 * b + c16 : 16 * a + r16, y2 = o ? 16 * a + r16 : 16 * b + c16;
#pragma unroll
            for (int t = 0; t < 6; t++) {
                const int n = 64 >> t, h = n >> 1, ry = y & (n - 1), ry2 = y2 & (n - 1);
                if (ry2 < h && (ry == ry2 || ry == ry2 + h)) {
                    const F2 p = x_mul<FQ2, SNR>(E[t][y], E[t][y2], nr, M);
                    const int k = 4 * t + (ry == ry2 ? 0 : 2);
                    vals[k] = addmod(vals[k], mulmod(p.a, gv, M), M);
                    if (FQ2) vals[k + 1] = addmod(vals[k + 1], mulmod(p.b, gv, M), M);
                }
            }
            if (o == 0 && y == y2) vals[24] = addmod(vals[24], gv, M);  // tr G
        }
    }
    block_sum<25>(vals, red, M);
    if (tid < 6) {
        const double f0a = red[4 * tid], f0b = red[4 * tid + 1];
        msg[tid][0] = f0a;
        msg[tid][1] = f0b;
        msg[tid][2] = submod(red[4 * tid + 2], f0a, M);
        msg[tid][3] = submod(red[4 * tid + 3], f0b, M);
    }
    if (tid == 0) *claim = claim_fault && row == 0 ? addmod(red[24], 1.0, M) : red[24];
}

// ---------------------------------------------------------------- final block
// One block per (row, rep) runs the rest of the norm check.  FIRST: the whole
// row (L <= SC_FINAL_MAX) from the witness.  Otherwise: rounds 0..3 from the
// Gram sums, rounds 4 .. t0 - 1 from the streaming passes' accumulators, then
// the folded witness `in` of length len <= SC_TAIL_MAX: above SC_FINAL_MAX the
// block first runs rounds t0 .. t0 + 3 as one streaming pass (a column of 16
// entries per thread, at most 256 columns) into LDS.  The remaining rounds run
// in LDS, V (length d) is captured for norm_check_3, and the verifier is
// replayed over all rounds.
template <bool FQ2, bool FIRST, bool SNR, bool G64 = false>
__global__ __launch_bounds__(SC_THREADS) void k_sc_final(
    const uint64_t *__restrict__ w, const unsigned long long *__restrict__ wrow,
    const double *__restrict__ in, size_t len, int rows, int t0,
    const uint64_t *__restrict__ cs0, size_t cs_row, int nrounds, double nr,
    const unsigned long long *__restrict__ macc, const unsigned long long *__restrict__ gacc,
    int claim_fault, NcOut o, int want_rhs, uint64_t q, const double *__restrict__ tab, Mod M) {
    __shared__ F2 cur[SC_FINAL_MAX];
    __shared__ F2 V[kD];
    __shared__ F2 base[kD];
    __shared__ double rc[2][kD];
    __shared__ double red[8 * 64];
    __shared__ double msg[48][4];  // f0, f1 of every round (f2 follows in the replay)
    __shared__ double claim_sh;
    __shared__ F2 chal[48];
    __shared__ double f2s[48][2];
    __shared__ F2 part[4][kD];  // norm_check_3's partial products
    __shared__ double msgp[10][4][64];  // the in-LDS rounds' per-lane messages
    __shared__ F2 sa_[48], sb_[48], clv[48];  // the replay's affine maps, claims
    const int row = blockIdx.x, rep = blockIdx.y;
    const int tid = threadIdx.x;
    const uint64_t *cs = cs0 + row * cs_row;
    // this repetition's challenges, loaded once (the round loop and the
    // verifier replay below would otherwise wait on one load per round)
    for (int tt = tid; tt < nrounds; tt += SC_THREADS) chal[tt] = load_chal<FQ2>(cs, rep, nrounds, tt);
    __syncthreads();

    int t = t0;
    if (FIRST) {
        for (size_t x = tid; x < len; x += SC_THREADS) {
            cur[x].a = u2d(w[(wrow ? wrow[row] : (size_t)row * len) + x]);
            cur[x].b = 0.0;
        }
        // norm_check_1 claim, computed locally from the whole row
        __syncthreads();
        double s[1] = {0.0};
        for (size_t x = tid; x < len; x += SC_THREADS)
            s[0] = reduce_lazy(s[0] + mul_lazy(cur[x].a, cur[x].a, M), M);
        s[0] = reduce(s[0], M);
        block_sum<1>(s, red, M);
        if (tid == 0) claim_sh = red[0];
    } else {
        const double *ib = in + ((size_t)rep * rows + row) * len * 2;
        const bool pre = len > (size_t)SC_FINAL_MAX;
        const size_t S = len >> 4;
        // this thread's Gram entry, then (with the pass below) its column's 16
        // entries: the loads stay in flight through the Gram messages
        // G64: rounds 0..5 from the 64 x 64 Gram sums (10 tile pairs)
        constexpr int TG = G64 ? 6 : 4;
        const unsigned long long graw = G64 ? 0ULL : gacc[(size_t)row * 256 + tid];
        double g64v[G64 ? G64_PAIRS : 1];
        if constexpr (G64) {
#pragma unroll
            for (int q = 0; q < G64_PAIRS; q++)
                g64v[q] = u64_mod(gacc[((size_t)row * G64_PAIRS + q) * 256 + tid], M);
        }
        F2 col[16];
        if (pre && (size_t)tid < S) {
#pragma unroll
            for (int y = 0; y < 16; y++) {
                const double *p = ib + ((size_t)y * S + tid) * 2;
                col[y].a = p[0];
                col[y].b = FQ2 ? p[1] : 0.0;
            }
        }
        if constexpr (G64)
            gram64_msgs<FQ2, SNR>(g64v, row, chal, msg, &claim_sh, claim_fault, red, nr, M);
        else
            gram_msgs<FQ2, SNR>(graw, row, chal, msg, &claim_sh, claim_fault, red, nr, M);
        for (int x = tid; x < (t0 - TG) * 4; x += SC_THREADS) {
            const int tt = TG + x / 4, k = x % 4;
            msg[tt][k] = u64_mod(macc[macc_idx(rep, tt, row, nrounds, rows) + k], M);
        }
        if (pre) {
            // rounds t0 .. t0 + 3 over the columns x < S of 16 entries
            // (balanced residues, as in k_sc_pass)
            F2 acc[4][2];
#pragma unroll
            for (int r = 0; r < 4; r++) acc[r][0] = acc[r][1] = F2{0.0, 0.0};
            if ((size_t)tid < S) {
                F2 c[4];
                double csum[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const F2 cr = chal[t0 + r];
                    c[r] = {balance(cr.a, M), FQ2 ? balance(cr.b, M) : 0.0};
                    csum[r] = c[r].a + c[r].b;
                }
                sc_level<FQ2, SNR, 8>(col, c[0], csum[0], acc[0][0], acc[0][1], nr, M);
                sc_level<FQ2, SNR, 4>(col, c[1], csum[1], acc[1][0], acc[1][1], nr, M);
                sc_level<FQ2, SNR, 2>(col, c[2], csum[2], acc[2][0], acc[2][1], nr, M);
                sc_level<FQ2, SNR, 1>(col, c[3], csum[3], acc[3][0], acc[3][1], nr, M);
                cur[tid] = {canon(col[0].a, M), FQ2 ? canon(col[0].b, M) : 0.0};
            }
            double vals[16];
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int f = 0; f < 2; f++) {
                    vals[r * 4 + f * 2] = reduce(acc[r][f].a, M);
                    vals[r * 4 + f * 2 + 1] = reduce(acc[r][f].b, M);
                }
            block_sum<16>(vals, red, M);  // synchronises: cur is complete
            if (tid < 16) msg[t0 + tid / 4][tid % 4] = red[tid];
            len = S;
            t = t0 + 4;
        } else {
            // the streaming passes leave balanced residues
            for (size_t x = tid; x < len; x += SC_THREADS) {
                const double *p = ib + x * 2;
                cur[x].a = canon(p[0], M);
                cur[x].b = FQ2 ? canon(p[1], M) : 0.0;
            }
        }
    }
    __syncthreads();

    // ---- the remaining rounds: wave 0 folds cur in LDS (ordered within the
    // wave only, messages by wave reductions); the other waves meanwhile form
    // norm_check_3's base = conj(tensor(last log2 d challenges))
    const int lane = tid & 63, wave = tid >> 6;
    const int tl = t;  // the first of these rounds
    if (wave == 0) {
        size_t cl = len;
        while (cl > 1) {
            if (cl == (size_t)kD && want_rhs) V[lane] = cur[lane];
            const size_t half = cl >> 1;
            const F2 c = chal[t];
            F2 a0 = {0.0, 0.0}, a1 = {0.0, 0.0};
            for (size_t x = lane; x < half; x += 64) {
                F2 w0 = cur[x];
                F2 d = x_sub<FQ2>(cur[x + half], w0, M);
                x_sq_mac<FQ2, SNR>(a0, w0, nr, M);
                x_mac<FQ2, SNR>(a1, w0, d, nr, M);
                cur[x] = x_add<FQ2>(w0, x_mul3<FQ2, SNR>(c, d, nr, M), M);
            }
            // the lanes' message parts, summed after the last round
            msgp[t - tl][0][lane] = reduce(a0.a, M);
            msgp[t - tl][1][lane] = reduce(a0.b, M);
            msgp[t - tl][2][lane] = reduce(a1.a, M);
            msgp[t - tl][3][lane] = reduce(a1.b, M);
            wave_sync();  // the folded entries are visible to the whole wave
            cl = half;
            t++;
        }
    } else if (want_rhs && wave == 1) {
        F2 tk = {1.0, 0.0};
        for (int b = 0; b < 6; b++) {
            const F2 cb = chal[nrounds - 6 + b];
            const int bit = (lane >> (5 - b)) & 1;
            F2 f = bit ? cb : x_sub<FQ2>(F2{1.0, 0.0}, cb, M);
            tk = x_mul<FQ2, SNR>(tk, f, nr, M);
        }
        // conj: base[0] = T[0], base[i] = -T[d - i]; lane owns T[lane]
        base[lane == 0 ? 0 : kD - lane] = lane == 0 ? tk : x_sub<FQ2>(F2{0.0, 0.0}, tk, M);
    }
    __syncthreads();
    const F2 ev = cur[0];
    // the rounds' messages: wave w sums rounds tl + w, tl + w + 4, ..
    for (int r = wave; r < nrounds - tl; r += SC_THREADS / 64) {
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = wave_sum_res(msgp[r][k][lane], M);
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 4; k++) msg[tl + r][k] = v[k];
        }
    }
    __syncthreads();  // every round's message before the verifier replay reads them

    // ---- norm_check_3: rhs = base * V (negacyclic), thread (k, i) summing
    // the 16 terms base[j] V[i - j], 16 k <= j < 16 k + 16, of coefficient i
    if (want_rhs) {
        const int i = lane, k = wave;
        double r0 = 0.0, r1 = 0.0;
#pragma unroll 4
        for (int jj = 0; jj < 16; jj++) {
            const int j = 16 * k + jj, vi = i - j;
            const bool wrap = vi < 0;
            const F2 b = base[j];
            const F2 v = V[wrap ? vi + kD : vi];
            double p0, p1 = 0.0;
            if (!FQ2) {
                p0 = mul_lazy(b.a, v.a, M);
            } else {
                p0 = mul_lazy(b.a, v.a, M) + nr_times<SNR>(nr, mul_lazy(b.b, v.b, M), M);
                p1 = mul_lazy(b.a, v.b, M) + mul_lazy(b.b, v.a, M);
            }
            r0 = reduce_lazy(wrap ? r0 - p0 : r0 + p0, M);
            r1 = reduce_lazy(wrap ? r1 - p1 : r1 + p1, M);
        }
        part[k][i] = {reduce(r0, M), reduce(r1, M)};
        __syncthreads();
        if (tid < 2 * kD) {
            const int comp = tid >> 6;
            double s = 0.0;
#pragma unroll
            for (int kk = 0; kk < 4; kk++) s = addmod(s, comp ? part[kk][lane].b : part[kk][lane].a, M);
            rc[comp][lane] = s;
        }
        __syncthreads();
        if (tid < 4) {
            // IncompleteNTT of component (tid >> 1), half (tid & 1)
            const int comp = tid >> 1, p = tid & 1;
            double v[32];
#pragma unroll
            for (int ii = 0; ii < 32; ii++) v[ii] = rc[comp][2 * ii + p];
            ntt_fwd_regs<32>(v, tab + TAB_FWD32, M);
            uint64_t *dst = o.rhs + rep * o.rhs_rep + row * o.rhs_row + comp * kD + 32 * p;
#pragma unroll
            for (int ii = 0; ii < 32; ii++) dst[ii] = d2u(v[ii]);
        }
    }

    // ---- verifier replay (fast_norm_check.rs:352-399, 443-462).  Round t's
    // f2 is the one value with 2 f0 + 2 f1 + f2 = claim_t (f2 = sum (w1 - w0)^2
    // equals it exactly, so the messages are those of a direct evaluation),
    // and then claim_{t+1} = f0 + 2 f1 c + f2 c^2 = alpha_t claim_t + beta_t
    // with alpha_t = c^2, beta_t = f0 (1 - 2 c^2) + 2 f1 (c - c^2): an affine
    // recurrence, evaluated by a prefix scan of the maps (alpha, beta).  The
    // accumulated f0 and f1 of every round are then checked at the end, where
    // the folded evaluation must square to the final claim.
    const F2 one = {1.0, 0.0};
    F2 f0 = {0.0, 0.0}, f1 = {0.0, 0.0};
    if (tid < nrounds) {
        f0 = {msg[tid][0], msg[tid][1]};
        f1 = {msg[tid][2], msg[tid][3]};
        const F2 c = chal[tid];
        const F2 c2 = x_mul<true, SNR>(c, c, nr, M);
        const F2 f1x2 = x_add<true>(f1, f1, M);
        sa_[tid] = c2;
        sb_[tid] = x_add<true>(x_mul<true, SNR>(f0, x_sub<true>(one, x_add<true>(c2, c2, M), M), nr, M),
                               x_mul<true, SNR>(f1x2, x_sub<true>(c, c2, M), nr, M), M);
    }
    __syncthreads();
    for (int d = 1; d < nrounds; d <<= 1) {
        F2 na = {0.0, 0.0}, nb = {0.0, 0.0};
        const bool act = tid < nrounds && tid >= d;
        if (act) {
            const F2 a2 = sa_[tid];
            na = x_mul<true, SNR>(a2, sa_[tid - d], nr, M);
            nb = x_add<true>(x_mul<true, SNR>(a2, sb_[tid - d], nr, M), sb_[tid], M);
        }
        __syncthreads();
        if (act) {
            sa_[tid] = na;
            sb_[tid] = nb;
        }
        __syncthreads();
    }
    const F2 claim0 = {claim_sh, 0.0};
    if (tid < nrounds)  // the claim after round tid
        clv[tid] = x_add<true>(x_mul<true, SNR>(sa_[tid], claim0, nr, M), sb_[tid], M);
    __syncthreads();
    if (tid < nrounds) {
        const F2 clt = tid == 0 ? claim0 : clv[tid - 1];
        const F2 f2 = x_sub<true>(clt, x_add<true>(x_add<true>(f0, f0, M), x_add<true>(f1, f1, M), M), M);
        f2s[tid][0] = f2.a;
        f2s[tid][1] = f2.b;
    }
    if (tid == 0) {
        if (rep == 0 && o.claims) o.claims[row * o.claim_row] = d2u(claim_sh);
        const F2 cl2 = clv[nrounds - 1];
        int good = 1;
        const F2 ev2 = x_mul<true, SNR>(ev, ev, nr, M);
        if (ev2.a != cl2.a || ev2.b != cl2.b) good = 0;
        if (want_rhs) {
            // ct(rhs)^2 == final claim (verify_norm_check_3): ct from the
            // coefficient form computed above
            const F2 ct = {rc[0][0], rc[1][0]};
            const F2 ct2 = x_mul<true, SNR>(ct, ct, nr, M);
            if (ct2.a != cl2.a || ct2.b != cl2.b) good = 0;
        }
        const size_t fe = rep * o.fe_rep + row * o.fe_row;
        if (o.finals) {
            o.finals[fe] = d2u(cl2.a);
            o.finals[fe + 1] = d2u(cl2.b);
        }
        if (o.evals) {
            o.evals[fe] = d2u(ev.a);
            o.evals[fe + 1] = d2u(ev.b);
        }
        o.ok[rep * o.ok_rep + row * o.ok_row] = good;
    }
    __syncthreads();
    if (o.msgs) {
        for (int x = tid; x < nrounds * 6; x += SC_THREADS) {
            const int tt = x / 6, k = x % 6;
            const double f = k >= 4 ? f2s[tt][k - 4] : msg[tt][k];
            o.msgs[rep * o.msg_rep + tt * o.msg_t + row * o.msg_row + k] = d2u(f);
        }
    }
}

__global__ void k_zero_u64(unsigned long long *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ULL;
}


template <bool FQ2, bool SNR>
int norm_check_impl(Context *ctx, const uint64_t *w, const unsigned long long *wrow, size_t rows,
                    size_t L, const uint64_t *cs, size_t cs_row, size_t reps, double nr,
                    const NcOut &o, int cpt_first, const NcScratch *ext, const ShortCopy *sc,
                    int fault_claim) {
    Mod M{ctx->qd, ctx->qinv};
    const int n = __builtin_ctzll(L);
    const int want_rhs = o.rhs != nullptr;
    size_t need_m, need_a, need_b;
    nc_scratch_need(rows, L, reps, &need_m, &need_a, &need_b, sc != nullptr);
    unsigned long long *macc = nullptr;
    const size_t nmacc = reps * (size_t)n * rows * 6;
    if (ext) {
        SALSAA_CHECK(ext->macc_bytes >= need_m && ext->bufA_bytes >= need_a &&
                         ext->bufB_bytes >= need_b,
                     "norm check scratch too small");
        macc = ext->macc;
    } else if (L > (size_t)SC_FINAL_MAX) {
        SALSAA_TRY(scratch_get(ctx, 0, need_m, (void **)&macc));
    }
    if (L <= (size_t)SC_FINAL_MAX) {
        // one block per (row, rep) from the witness: no accumulators
        hipLaunchKernelGGL((k_sc_final<FQ2, true, SNR>), dim3((unsigned)rows, (unsigned)reps),
                           dim3(SC_THREADS), 0, ctx->stream, w, wrow, nullptr, L, (int)rows, 0, cs,
                           cs_row, n, nr, nullptr, nullptr, 0, o, want_rhs, ctx->q, ctx->d_tables, M);
        SALSAA_HIP(hipGetLastError());
        return SALSAA_OK;
    }
    // the passes' message accumulators, the rows' Gram sums (rows x 256) and
    // the flag of first-pass sub-blocks left to k_sc_gram_rest start at zero
    // (with a short copy the first pass may keep 10 tile pairs of 256 sums)
    const size_t gram = sc != nullptr ? kGram64Entries : 256;
    unsigned long long *gacc = macc + nmacc, *need_long = gacc + rows * gram;
    const size_t nzero = nmacc + rows * gram + 1;
    hipLaunchKernelGGL(k_zero_u64, dim3(ceil_div(nzero, 256)), dim3(256), 0, ctx->stream, macc,
                       nzero);
    SALSAA_HIP(hipGetLastError());
    // ping-pong buffers of folded witnesses (pairs of doubles)
    const size_t len1 = L >> 4;
    double *bufA = nullptr, *bufB = nullptr;
    if (ext) {
        bufA = ext->bufA;
        bufB = ext->bufB;
    } else {
        SALSAA_TRY(scratch_get(ctx, 1, need_a, (void **)&bufA));
        SALSAA_TRY(scratch_get(ctx, 2, need_b, (void **)&bufB));
    }
    // With a short copy and L >= 2^16 the first pass runs six rounds from a
    // 64 x 64 Gram matrix (its folded output, L / 64 per repetition, still
    // fills the final block)
    const bool g64 = sc != nullptr && (L >> 6) >= (size_t)SC_FINAL_MAX;
    if (g64) {
        const size_t S = L >> 6;
        // columns per thread and blocks as in the 16-segment pass below
        int cpt = cpt_first > 0 ? std::min(cpt_first, 8) : 1;
        if (cpt_first <= 0)
            while (cpt < 8 && rows * ceil_div(S, (size_t)GR_X * cpt) > (size_t)ctx->ncu * 3) cpt *= 2;
        while (cpt < 64 && ceil_div(S, (size_t)GR_X * cpt) > kGramBlocksMax) cpt *= 2;
        const size_t nsub = ceil_div(S, (size_t)GR_X * cpt);
        int sbn = std::max(1, std::min(8, 64 / cpt));
        while (sbn > 1 && ceil_div(nsub, (size_t)sbn) * rows < (size_t)ctx->ncu * 2) sbn >>= 1;
        const unsigned bx = (unsigned)ceil_div(nsub, (size_t)sbn);
        dim3 grid(bx, (unsigned)rows);
        dim3 ggrid = grid;
        const size_t walk = std::max<size_t>(1, 64 / ((size_t)sbn * cpt));
        ggrid.x = (unsigned)ceil_div((size_t)bx, walk);
        SALSAA_CHECK((size_t)bx + (size_t)ggrid.x <= kGramBlocksMax,
                     "norm check row too long (at most 2^32 coefficients)");
        // the byte-wide sub-blocks on k_sc_gram64_byte, the two-byte ones on
        // k_sc_gram64_short, the others on k_sc_gram64_rest (L is a power of
        // two, so S is a multiple of the byte kernel's 256 columns)
        SALSAA_CHECK(S % GR_X == 0, "64-segment pass needs whole 256-column iterations");
        for (size_t r0 = 0; r0 < reps; r0 += G64_MAXREP) {
            const int nrep = (int)std::min<size_t>(G64_MAXREP, reps - r0);
            if (r0 == 0) {
                hipLaunchKernelGGL((k_sc_gram64_short<FQ2, SNR, true>), grid, dim3(SC_THREADS), 0,
                                   ctx->stream, wrow, L, (int)rows, (int)r0, nrep, cs, cs_row, n, nr,
                                   cpt, bufA, gacc, ctx->d_tables, M, sc->t, sc->fmt, sbn, need_long);
                hipLaunchKernelGGL((k_sc_gram64_rest<FQ2, SNR, true>), ggrid, dim3(SC_THREADS), 0,
                                   ctx->stream, w, wrow, L, (int)rows, (int)r0, nrep, cs, cs_row, n,
                                   nr, cpt, bufA, gacc, M, sc->fmt, sbn, (int)bx, need_long);
            } else {
                hipLaunchKernelGGL((k_sc_gram64_short<FQ2, SNR, false>), grid, dim3(SC_THREADS), 0,
                                   ctx->stream, wrow, L, (int)rows, (int)r0, nrep, cs, cs_row, n, nr,
                                   cpt, bufA, gacc, ctx->d_tables, M, sc->t, sc->fmt, sbn, need_long);
                hipLaunchKernelGGL((k_sc_gram64_rest<FQ2, SNR, false>), ggrid, dim3(SC_THREADS), 0,
                                   ctx->stream, w, wrow, L, (int)rows, (int)r0, nrep, cs, cs_row, n,
                                   nr, cpt, bufA, gacc, M, sc->fmt, sbn, (int)bx, need_long);
            }
            // the byte kernel last: the two kernels above usually find
            // nothing to do, and their blocks (large in LDS) get CUs most
            // easily before the byte kernel's fill them
            if (r0 == 0)
                hipLaunchKernelGGL((k_sc_gram64_byte<FQ2, SNR, true>), grid, dim3(SC_THREADS), 0,
                                   ctx->stream, wrow, L, (int)rows, (int)r0, nrep, cs, cs_row, n, nr,
                                   cpt, bufA, gacc, ctx->d_tables, M, sc->t8, sc->fmt, sbn);
            else
                hipLaunchKernelGGL((k_sc_gram64_byte<FQ2, SNR, false>), grid, dim3(SC_THREADS), 0,
                                   ctx->stream, wrow, L, (int)rows, (int)r0, nrep, cs, cs_row, n, nr,
                                   cpt, bufA, gacc, ctx->d_tables, M, sc->t8, sc->fmt, sbn);
            SALSAA_HIP(hipGetLastError());
        }
    } else {
        // one streaming pass: the folds after round 3 and the Gram matrix
        // (rounds 0..3's messages follow from it in the final block)
        const size_t S = len1;
        // columns per thread: the batch tuner's choice, else the fewest that
        // keep the grid within one round of resident blocks (k_sc_gram holds
        // three per CU): more blocks add per-block setup and Gram atomics,
        // fewer leave CUs idle
        int cpt = cpt_first > 0 ? std::min(cpt_first, 8) : 1;
        if (cpt_first <= 0)
            while (cpt < 8 && rows * ceil_div(S, (size_t)GR_X * cpt) > (size_t)ctx->ncu * 3) cpt *= 2;
        // Every first-pass block adds one residue (< 2^50) to each of its
        // row's u64 Gram sums, so a row takes at most 2^14 blocks; longer rows
        // get more columns per thread (a wave's int32 digit sums stay exact
        // over 64 iterations of 64 columns)
        while (cpt < 64 && ceil_div(S, (size_t)GR_X * cpt) > kGramBlocksMax) cpt *= 2;
        const size_t nsub = ceil_div(S, (size_t)GR_X * cpt);
        // With a short copy, a block takes up to 8 sub-blocks of cpt GR_X
        // columns (each kernel skips the other's): the fewer, longer blocks
        // amortise the skipped blocks and the per-block setup.  A wave then
        // sums at most 64 iterations of 64 columns: its int32 digit sums stay
        // exact.
        int sbn = 1;
        if (sc)
            for (sbn = std::max(1, std::min(8, 64 / cpt));
                 sbn > 1 && ceil_div(nsub, (size_t)sbn) * rows < (size_t)ctx->ncu * 4;)
                sbn >>= 1;
        const unsigned bx = (unsigned)ceil_div(nsub, (size_t)sbn);
        dim3 grid(bx, (unsigned)rows);
        // With a short copy k_sc_gram_rest takes the sub-blocks the copy
        // cannot serve, usually none (then each block returns after one
        // load).  A block walks up to `walk` super-blocks, at most 64
        // iterations of 256 columns: its u64 Gram sums stay exact.
        dim3 ggrid = grid;
        if (sc) {
            const size_t walk = std::max<size_t>(1, 64 / ((size_t)sbn * cpt));
            ggrid.x = (unsigned)ceil_div((size_t)bx, walk);
        }
        SALSAA_CHECK((size_t)bx + (sc ? (size_t)ggrid.x : 0) <= kGramBlocksMax,
                     "norm check row too long (at most 2^32 coefficients)");
        for (size_t r0 = 0; r0 < reps; r0 += GR_MAXREP) {
            const int nrep = (int)std::min<size_t>(GR_MAXREP, reps - r0);
            // with a short copy, its short blocks first, then the others
            if (sc && r0 == 0)
                hipLaunchKernelGGL((k_sc_gram_short<FQ2, SNR, true>), grid, dim3(SC_THREADS), 0,
                                   ctx->stream, wrow, L, (int)rows, (int)r0, nrep, cs, cs_row, n, nr,
                                   cpt, bufA, gacc, ctx->d_tables, M, sc->t, sc->t8, sc->fmt, sbn, need_long);
            else if (sc)
                hipLaunchKernelGGL((k_sc_gram_short<FQ2, SNR, false>), grid, dim3(SC_THREADS), 0,
                                   ctx->stream, wrow, L, (int)rows, (int)r0, nrep, cs, cs_row, n, nr,
                                   cpt, bufA, gacc, ctx->d_tables, M, sc->t, sc->t8, sc->fmt, sbn, need_long);
            if (sc && r0 == 0)
                hipLaunchKernelGGL((k_sc_gram_rest<FQ2, SNR, true>), ggrid, dim3(SC_THREADS), 0,
                                   ctx->stream, w, wrow, L, (int)rows, (int)r0, nrep, cs, cs_row, n,
                                   nr, cpt, bufA, gacc, M, sc->fmt, sbn, (int)bx, need_long);
            else if (sc)
                hipLaunchKernelGGL((k_sc_gram_rest<FQ2, SNR, false>), ggrid, dim3(SC_THREADS), 0,
                                   ctx->stream, w, wrow, L, (int)rows, (int)r0, nrep, cs, cs_row, n,
                                   nr, cpt, bufA, gacc, M, sc->fmt, sbn, (int)bx, need_long);
            else if (r0 == 0)
                hipLaunchKernelGGL((k_sc_gram<FQ2, SNR, true>), grid, dim3(SC_THREADS), 0,
                                   ctx->stream, w, wrow, L, (int)rows, (int)r0, nrep, cs, cs_row, n,
                                   nr, cpt, bufA, gacc, ctx->d_tables, M);
            else
                hipLaunchKernelGGL((k_sc_gram<FQ2, SNR, false>), grid, dim3(SC_THREADS), 0,
                                   ctx->stream, w, wrow, L, (int)rows, (int)r0, nrep, cs, cs_row, n,
                                   nr, cpt, bufA, gacc, ctx->d_tables, M);
            SALSAA_HIP(hipGetLastError());
        }
    }
    size_t len = g64 ? L >> 6 : len1;
    int t0 = g64 ? 6 : 4;
    double *src = bufA, *dst = bufB;
    // streaming passes while the final block cannot take the rest
    while (len > (size_t)SC_TAIL_MAX) {
        const size_t S = len >> 4;
        const int cpt = S >= (size_t)SC_THREADS * 64 ? 4 : 1;
        dim3 grid(ceil_div(S, (size_t)SC_THREADS * cpt), (unsigned)rows, (unsigned)reps);
        hipLaunchKernelGGL((k_sc_pass<FQ2, SNR>), grid, dim3(SC_THREADS), 0, ctx->stream, src, len,
                           (int)rows, t0, cs, cs_row, n, nr, cpt, dst, macc, M);
        SALSAA_HIP(hipGetLastError());
        len = S;
        t0 += 4;
        double *tmp = src;
        src = dst;
        dst = tmp;
    }
    if (g64)
        hipLaunchKernelGGL((k_sc_final<FQ2, false, SNR, true>), dim3((unsigned)rows, (unsigned)reps),
                           dim3(SC_THREADS), 0, ctx->stream, nullptr, nullptr, src, len, (int)rows, t0,
                           cs, cs_row, n, nr, macc, gacc, fault_claim, o, want_rhs, ctx->q,
                           ctx->d_tables, M);
    else
        hipLaunchKernelGGL((k_sc_final<FQ2, false, SNR>), dim3((unsigned)rows, (unsigned)reps),
                           dim3(SC_THREADS), 0, ctx->stream, nullptr, nullptr, src, len, (int)rows, t0,
                           cs, cs_row, n, nr, macc, gacc, fault_claim, o, want_rhs, ctx->q,
                           ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

}  // namespace

void nc_scratch_need(size_t rows, size_t L, size_t reps, size_t *macc_bytes, size_t *bufA_bytes,
                     size_t *bufB_bytes, bool short_copy) {
    const size_t gram = short_copy ? kGram64Entries : 256;  // Gram sums per row
    const size_t n = (L >= 2) ? (size_t)__builtin_ctzll(L) : 0;
    const size_t len1 = L >> 4;
    *macc_bytes = (L > (size_t)SC_FINAL_MAX ? reps * n * rows * 6 + rows * gram + 1 : 0) *
                  sizeof(unsigned long long);
    *bufA_bytes = L > (size_t)SC_FINAL_MAX ? reps * rows * len1 * 2 * sizeof(double) : 0;
    *bufB_bytes =
        L > (size_t)SC_FINAL_MAX ? reps * rows * (len1 >> 4 ? len1 >> 4 : 1) * 2 * sizeof(double) : 0;
}

NcOut nc_out_plain(size_t rows, size_t reps, size_t nrounds, uint64_t *claims, uint64_t *msgs,
                   uint64_t *finals, uint64_t *evals, uint64_t *rhs, int32_t *ok) {
    NcOut o;
    o.claims = claims;
    o.msgs = msgs;
    o.finals = finals;
    o.evals = evals;
    o.rhs = rhs;
    o.ok = ok;
    o.claim_row = 1;
    o.msg_rep = nrounds * rows * 6;
    o.msg_t = rows * 6;
    o.msg_row = 6;
    o.fe_rep = rows * 2;
    o.fe_row = 2;
    o.rhs_rep = rows * 2 * kD;
    o.rhs_row = 2 * kD;
    o.ok_rep = rows;
    o.ok_row = 1;
    (void)reps;
    return o;
}

int launch_norm_check_ex(Context *ctx, const uint64_t *w, const unsigned long long *wrow,
                         size_t rows, size_t L, const uint64_t *cs, size_t cs_row, size_t reps,
                         int fq2, uint64_t nr, const NcOut &o, int cpt_first,
                         const NcScratch *ext, const ShortCopy *sc, int fault_claim) {
    SALSAA_CHECK(rows > 0 && reps > 0, "norm check needs rows > 0 and reps > 0");
    SALSAA_CHECK(sc == nullptr || (sc->t && sc->t8 && sc->fmt), "incomplete short copy");
    SALSAA_CHECK(L >= 2 && (L & (L - 1)) == 0, "norm check length must be a power of two >= 2");
    SALSAA_CHECK(rows < 65536 && reps < 65536, "too many rows or repetitions");
    SALSAA_CHECK(__builtin_ctzll(L) <= 48, "norm check length too large");
    SALSAA_CHECK(o.rhs == nullptr || L >= (size_t)ctx->d, "norm_check_3 needs L >= d");
    SALSAA_CHECK(nr < ctx->q, "non-residue must be reduced mod q");
    SALSAA_CHECK(o.ok != nullptr, "ok output is required");
    SALSAA_CHECK(cpt_first >= 0 && cpt_first <= 64, "columns per thread out of range");
    SALSAA_CHECK(!fault_claim || L > (size_t)SC_FINAL_MAX, "claim fault needs L > 1024");
    if (ctx->d != kD && o.rhs != nullptr) {
        // degree d > 64: the sumcheck, then norm_check_3's rhs and its check
        SALSAA_CHECK(wrow == nullptr && sc == nullptr && cs_row == 0 && ext == nullptr,
                     "norm_check_3 at degree d > 64 runs on plain witness rows");
        NcOut o2 = o;
        o2.rhs = nullptr;
        if (!o2.finals) {
            SALSAA_TRY(scratch_get(ctx, 14, reps * rows * 2 * sizeof(uint64_t),
                                   (void **)&o2.finals));
            o2.fe_rep = rows * 2;
            o2.fe_row = 2;
        }
        SALSAA_TRY(launch_norm_check_ex(ctx, w, wrow, rows, L, cs, cs_row, reps, fq2, nr, o2,
                                        cpt_first, ext, sc, fault_claim));
        return launch_nc3_wide(ctx, w, rows, L, cs, reps, (double)nr, fq2 != 0, o2.finals,
                               o2.fe_rep, o2.fe_row, o.rhs, o.ok, o.ok_rep, o.ok_row);
    }
    if (fq2 && nr < 8)
        return norm_check_impl<true, true>(ctx, w, wrow, rows, L, cs, cs_row, reps, (double)nr, o,
                                           cpt_first, ext, sc, fault_claim);
    if (fq2)
        return norm_check_impl<true, false>(ctx, w, wrow, rows, L, cs, cs_row, reps, (double)nr, o,
                                            cpt_first, ext, sc, fault_claim);
    return norm_check_impl<false, true>(ctx, w, wrow, rows, L, cs, cs_row, reps, (double)nr, o,
                                        cpt_first, ext, sc, fault_claim);
}

int launch_norm_check(Context *ctx, const uint64_t *w, size_t rows, size_t L,
                      const uint64_t *cs, size_t reps, int fq2, uint64_t nr, uint64_t *claims,
                      uint64_t *msgs, uint64_t *finals, uint64_t *evals, uint64_t *rhs,
                      int32_t *ok) {
    const size_t nrounds = (L >= 2 && (L & (L - 1)) == 0) ? (size_t)__builtin_ctzll(L) : 0;
    const NcOut o = nc_out_plain(rows, reps, nrounds, claims, msgs, finals, evals, rhs, ok);
    return launch_norm_check_ex(ctx, w, nullptr, rows, L, cs, 0, reps, fq2, nr, o, 0, nullptr);
}

}  // namespace salsaa
