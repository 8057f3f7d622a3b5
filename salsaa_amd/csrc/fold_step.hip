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
// k_decomp_fold fuses the first two: a block takes 16 columns of all r rows of
// one instance, forms each limb and its IncompleteNTT in LDS, and accumulates
// the fold for all r_out output rows in registers, so the limbs never reach
// HBM.  The same block then writes the fold output (IncompleteNTT), the
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

// grid (m / 16, n).  W [n * r][m][d]; Cn [n][r_out][r * parts][d] in
// IncompleteNTT form; outputs for stacked rows R = i * r_out + j.
__global__ __launch_bounds__(FNT) void k_decomp_fold(
    const uint64_t *__restrict__ W, int r, size_t m, int parts,
    const uint64_t *__restrict__ d_radix, const uint64_t *__restrict__ Cn, int r_out,
    uint64_t *__restrict__ folded, uint64_t *__restrict__ folded_coeff,
    int8_t *__restrict__ planes, size_t ktiles, size_t rt_n, uint64_t q,
    const double *__restrict__ tab, Mod M) {
    __shared__ uint64_t lds[128 * FROW];
    __shared__ double cl[8 * 8 * kD];  // this limb's challenges C[jo][rho * parts + tl]
    const int inst = blockIdx.y;
    const size_t i0 = (size_t)blockIdx.x * FCOLS;
    const int tid = threadIdx.x;
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
    // NTT role: element e, half p, quarter t (positions 8k + 2t + p)
    const int e = tid >> 3, h = tid & 7, p = h >> 2, t = h & 3;
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
    CTX_OK(ctx);
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
    for (size_t i = 0; i < n; i++)
        SALSAA_TRY(launch_radix(c, d_W + i * r * m * kD, r, m, parts, radix, o->radix + i));
    uint64_t *cn = nullptr;
    SALSAA_TRY(scratch_get(c, 5, n * r_out * rk * kD * sizeof(uint64_t), (void **)&cn));
    SALSAA_HIP(hipMemcpyAsync(cn, d_C, n * r_out * rk * kD * sizeof(uint64_t),
                              hipMemcpyDeviceToDevice, c->stream));
    SALSAA_TRY(launch_ntt(c, cn, n * r_out * rk, SALSAA_REP_INCOMPLETE, 0));
    // fused decomposition + fold -> fold output, planes, Coefficient form
    const size_t ktiles = plane_ktiles(m), rt_n = (7 * rows + 15) / 16;
    int8_t *planes = nullptr;
    SALSAA_TRY(scratch_get(c, 6, plane_bytes(rows, m), (void **)&planes));
    hipLaunchKernelGGL(k_decomp_fold, dim3((unsigned)(m / FCOLS), (unsigned)n), dim3(FNT), 0,
                       c->stream, d_W, (int)r, m, (int)parts, o->radix, cn, (int)r_out, o->folded,
                       o->folded_coeff, planes, ktiles, rt_n, c->q, c->d_tables, M);
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
