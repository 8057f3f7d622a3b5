// ring_wide.hip -- the ring R_q = Z_q[X]/(X^d + 1) at the degrees of the
// reference's parameter sets, d = 128, 256 and 512 (protocol.rs:28-120),
// beside the d = 64 kernels of the other files.
//
// Transforms.  IncompleteNTT (cyclotomic_ring.rs:483-511) is two negacyclic
// length-n transforms, n = d/2, of the even and the odd coefficients; the
// Full NTT (cyclotomic_ring.rs:514-525) one of length d, present when 2d
// divides q - 1.  Both use HEXL's convention (psi the minimal primitive
// 2n-th root, Cooley-Tukey forward with bit-reversed output, Gentleman-Sande
// inverse with n^{-1} folded into its last stage).  A length-N transform
// is held by N/8 threads of one wavefront, 8 values each, as radix-8 register
// phases of three butterfly stages (the last phase may hold fewer) with an
// LDS exchange between phases that only that wavefront touches, so no block
// barrier is needed.  In phase p (stages 3p ..) thread u holds positions
// (u / T) 8 T + u % T + k T, k < 8, where T is the half-distance of the
// phase's last stage: the first phase reads positions u + k N/8, which for
// the IncompleteNTT are coefficients l + k d/8 of the element's d/8 threads
// (l = 2u + half), and the last phase writes 8 consecutive positions.
//
// Commit.  In the IncompleteNTT domain slot j of a ring element is
// e_j + X o_j modulo X^2 - omega_j (incomplete_ntt_multiplication,
// cyclotomic_ring.rs:699-791), so a commitment's slot j is
//     E = sum_i a_e w_e + omega_j sum_i a_o w_o,  O = sum_i (a_e w_o + a_o w_e).
// The int8 matrix-core commit of mfma_commit.hip runs unchanged over the d/2
// slots with every matrix's value rows doubled (row 2 r + h holds half h of
// row r): its per-slot product of the doubled key (2 kappa rows) and the
// doubled witness (2 r rows) holds all four half products, which
// k_commit_finalize_pairs combines.
#include <algorithm>

#include "ntt8.hpp"

namespace salsaa {

namespace {

constexpr int WT = 256;  // threads per block of the transforms

// position of value k of thread u in the register phase whose last stage
// pairs values T apart
__device__ __forceinline__ int gpos(int u, int T, int k) { return (u / T) * 8 * T + (u % T) + k * T; }

// exchange between register phases through this transform's LDS row x
__device__ __forceinline__ void phase_exchange(double (&v)[8], double *x, int u, int Tfrom,
                                               int Tto) {
#pragma unroll
    for (int k = 0; k < 8; k++) x[gpos(u, Tfrom, k)] = v[k];
    wave_sync();
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = x[gpos(u, Tto, k)];
    wave_sync();
}

// forward length-2^LOGN transform, thread u of N/8; tw[m + i] = psi^brv(m + i)
template <int LOGN>
__device__ __forceinline__ void wide_fwd(double (&v)[8], double *x, int u,
                                         const double *__restrict__ tw, Mod M) {
    constexpr int N = 1 << LOGN;
#pragma unroll
    for (int s0 = 0; s0 < LOGN; s0 += 3) {
        const int c = LOGN - s0 < 3 ? LOGN - s0 : 3;
        const int T = N >> (s0 + c);
        if (s0 > 0) phase_exchange(v, x, u, N >> s0, T);
#pragma unroll
        for (int e = 0; e < c; e++) {
            const int s = s0 + e, kd = 1 << (c - 1 - e), t = T * kd;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (k & kd) continue;
                const int j = gpos(u, T, k);
                bf_fwd(v[k], v[k + kd], tw[(1 << s) + j / (2 * t)], M);
            }
        }
    }
}

// inverse (Gentleman-Sande): stages in reverse; twi[m + i] the inverse
// twiddles, twi[0] = twi[1] n^{-1} for the last stage, which also scales
// the sums by ninv
template <int LOGN>
__device__ __forceinline__ void wide_inv(double (&v)[8], double *x, int u,
                                         const double *__restrict__ twi, double ninv, Mod M) {
    constexpr int N = 1 << LOGN;
    constexpr int NPH = (LOGN + 2) / 3;
#pragma unroll
    for (int ph = NPH - 1; ph >= 0; ph--) {
        const int s0 = 3 * ph;
        const int c = LOGN - s0 < 3 ? LOGN - s0 : 3;
        const int T = N >> (s0 + c);
        if (ph < NPH - 1) {
            const int s1 = s0 + 3, c1 = LOGN - s1 < 3 ? LOGN - s1 : 3;
            phase_exchange(v, x, u, N >> (s1 + c1), T);
        }
#pragma unroll
        for (int e = c - 1; e >= 0; e--) {
            const int s = s0 + e, kd = 1 << (c - 1 - e), t = T * kd;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (k & kd) continue;
                if (s == 0) {
                    const double X = v[k], Y = v[k + kd];
                    v[k] = mulmod(addmod(X, Y, M), ninv, M);
                    v[k + kd] = mulmod(submod(X, Y, M), twi[0], M);
                } else {
                    const int j = gpos(u, T, k);
                    bf_inv(v[k], v[k + kd], twi[(1 << s) + j / (2 * t)], M);
                }
            }
        }
    }
}

// Representation change of `count` degree-2^LOGD elements, src -> dst (may
// be the same buffer): INC the IncompleteNTT, else the Full NTT; INV from
// that representation to Coefficient.  Each element is handled by d/8
// threads of one wavefront, a block takes WT/(d/8) elements.  Both
// directions move data lane-contiguously (element word l + k d/8 for lane
// l): the forward's last phase hands its outputs over in LDS, the inverse's
// first phase takes its inputs from there.  (Persistent blocks that load the
// next group during this one's transform measured slower.)
template <int LOGD, bool INV, bool INC>
__global__ __launch_bounds__(WT) void k_wide_ntt(const uint64_t *src, uint64_t *dst, size_t count,
                                                 const double *__restrict__ wt, Mod M) {
    constexpr int d = 1 << LOGD, LOGN = INC ? LOGD - 1 : LOGD, N = 1 << LOGN;
    constexpr int TPE = d / 8, EPB = WT / TPE;
    __shared__ double x[EPB * d];
    __shared__ double tws[N];  // the transform's twiddles (per-lane indices)
    const WideTab W(d);
    const int tid = threadIdx.x, el = tid / TPE, l = tid % TPE;
    const int h = INC ? (l & 1) : 0, u = INC ? (l >> 1) : l;
    const size_t e = (size_t)blockIdx.x * EPB + el;
    const bool live = e < count;
    {
        const double *tw = wt + (INC ? (INV ? W.inv_h : W.fwd_h) : (INV ? W.inv_f : W.fwd_f));
        for (int i = tid; i < N; i += WT) tws[i] = tw[i];
        __syncthreads();
    }
    double *xe = x + el * d, *xs = xe + h * N;
    double v[8];
    if (!INV) {
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = live ? u2d(src[e * d + l + k * TPE]) : 0.0;
        wide_fwd<LOGN>(v, xs, u, tws, M);
#pragma unroll
        for (int k = 0; k < 8; k++) xs[8 * u + k] = v[k];
        wave_sync();
        if (live) {
#pragma unroll
            for (int k = 0; k < 8; k++) dst[e * d + l + k * TPE] = d2u(xe[l + k * TPE]);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; k++) xe[l + k * TPE] = live ? u2d(src[e * d + l + k * TPE]) : 0.0;
        wave_sync();
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = xs[8 * u + k];
        wave_sync();
        wide_inv<LOGN>(v, xs, u, tws, wt[INC ? W.ninv_h : W.ninv_f], M);
        if (live) {
#pragma unroll
            for (int k = 0; k < 8; k++) dst[e * d + l + k * TPE] = d2u(v[k]);
        }
    }
}

template <int LOGD, bool INV, bool INC>
int launch_wide_kernel(Context *ctx, const uint64_t *src, uint64_t *dst, size_t count) {
    constexpr int EPB = WT / ((1 << LOGD) / 8);
    const unsigned g = ceil_div(count, EPB);
    Mod M{ctx->qd, ctx->qinv};
    hipLaunchKernelGGL((k_wide_ntt<LOGD, INV, INC>), dim3(g), dim3(WT), 0, ctx->stream, src, dst,
                       count, ctx->d_wtab, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

template <int LOGD>
int run_wide_ntt(Context *ctx, const uint64_t *src, uint64_t *dst, size_t count, bool inc,
                 bool inv) {
    if (inc && !inv) return launch_wide_kernel<LOGD, false, true>(ctx, src, dst, count);
    if (inc) return launch_wide_kernel<LOGD, true, true>(ctx, src, dst, count);
    if (!inv) return launch_wide_kernel<LOGD, false, false>(ctx, src, dst, count);
    return launch_wide_kernel<LOGD, true, false>(ctx, src, dst, count);
}

// ---------------------------------------------------------------- commit planes
// Balanced base-256 digits (bytes of x + sum_u 128 * 256^u, each XOR 0x80),
// as in the d = 64 plane writers.
constexpr uint64_t WIDE_DIGIT_BIAS = 0x0080808080808080ULL;

// Fragment-tiled digit planes (the layout of store_planes in ntt8.hip) from
// IncompleteNTT elements, over nslot = d/2 slots and 2 rows value rows:
// value row 2 row + h of slot j holds half h of slot j.  Block (K tile kt0 +
// x, group of PK_SG slots, row0 + z): the group's two halves of the tile's
// 64 elements staged in LDS, then thread (slot, half, 8-element run) writes
// the 8 bytes of its run in each of the 7 digit rows, so 8 lanes fill each
// 64-byte fragment row.  Element i of the row sits at inc + ((z inc_rs) +
// i - i_base) d; columns past m are zeros.
constexpr int PK_SG = 32;
template <int LOGD>
__global__ __launch_bounds__(512) void k_wide_pack(const uint64_t *__restrict__ inc, size_t inc_rs,
                                                   size_t i_base, size_t m, size_t rows,
                                                   size_t row0, size_t kt0, size_t ktiles,
                                                   int8_t *__restrict__ dst) {
    constexpr int d = 1 << LOGD, N = d / 2;
    __shared__ uint64_t st[64][2 * PK_SG + 1];
    const int tid = threadIdx.x;
    const size_t kt = kt0 + blockIdx.x, row = row0 + blockIdx.z, i0 = kt * 64;
    const int j0 = blockIdx.y * PK_SG;
    const uint64_t *src = inc + (size_t)blockIdx.z * inc_rs * d;
    for (int c = tid; c < 64 * 2 * PK_SG; c += 512) {
        const int e = c / (2 * PK_SG), hj = c % (2 * PK_SG);
        const size_t i = i0 + e;
        st[e][hj] = i < m ? src[(i - i_base) * d + (hj / PK_SG) * N + j0 + hj % PK_SG] : 0;
    }
    __syncthreads();
    const int eg = tid & 7, hj = tid >> 3, h = hj / PK_SG, jl = hj % PK_SG;
    uint32_t dg[2][7];
#pragma unroll
    for (int g = 0; g < 2; g++) {
        uint32_t lo[4], hi[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint64_t y = (st[8 * eg + 4 * g + q][hj] + WIDE_DIGIT_BIAS) ^ 0x8080808080808080ULL;
            lo[q] = (uint32_t)y;
            hi[q] = (uint32_t)(y >> 32);
        }
        bytes_transpose4(lo, hi, dg[g]);
    }
    const size_t rt = (2 * rows * 7 + 15) / 16;
    int8_t *base = dst + ((size_t)(j0 + jl) * ktiles + kt) * rt * 1024 +
                   (size_t)(7 * (2 * row + h)) * 64 + 8 * eg;
#pragma unroll
    for (int q = 0; q < 7; q++)
        __builtin_nontemporal_store((unsigned long long)dg[0][q] | ((unsigned long long)dg[1][q] << 32),
                                    reinterpret_cast<unsigned long long *>(base + q * 64));
}

template <int LOGD>
int run_wide_pack(Context *ctx, const uint64_t *inc, size_t inc_rs, size_t i_base, size_t m,
                  size_t rows, size_t row0, size_t nrows, size_t kt0, size_t nkt, size_t ktiles,
                  int8_t *dst) {
    if (nkt == 0) return SALSAA_OK;
    hipLaunchKernelGGL((k_wide_pack<LOGD>),
                       dim3((unsigned)nkt, (unsigned)((1 << LOGD) / 2 / PK_SG), (unsigned)nrows),
                       dim3(512), 0, ctx->stream, inc, inc_rs, i_base, m, rows, row0, kt0, ktiles,
                       dst);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int run_pack_any(Context *ctx, const uint64_t *inc, size_t inc_rs, size_t i_base, size_t m,
                 size_t rows, size_t row0, size_t nrows, size_t kt0, size_t nkt, size_t ktiles,
                 int8_t *dst) {
    switch (ctx->d) {
        case 128:
            return run_wide_pack<7>(ctx, inc, inc_rs, i_base, m, rows, row0, nrows, kt0, nkt,
                                    ktiles, dst);
        case 256:
            return run_wide_pack<8>(ctx, inc, inc_rs, i_base, m, rows, row0, nrows, kt0, nkt,
                                    ktiles, dst);
        case 512:
            return run_wide_pack<9>(ctx, inc, inc_rs, i_base, m, rows, row0, nrows, kt0, nkt,
                                    ktiles, dst);
        default: break;
    }
    set_error("no plane writer for this degree");
    return SALSAA_ERR_ARG;
}

// A commitment's slots from the per-slot products of the doubled rows,
// partial [ksplit][n][2 kappa][2 r] (canonical residues): thread per (j, k,
// slot), out [kappa][r][d] IncompleteNTT.
__global__ void k_commit_finalize_pairs(const double *__restrict__ partial, int ksplit, int kappa,
                                        int r, int logn, uint64_t *__restrict__ out,
                                        const double *__restrict__ omega, Mod M) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = 1 << logn;
    if (gid >= ((size_t)kappa * r << logn)) return;
    const int p = (int)(gid & (size_t)(n - 1));
    const size_t jk = gid >> logn;
    const size_t j = jk / (size_t)r, k = jk % (size_t)r;
    const size_t r2 = 2 * (size_t)r, plane = 2 * (size_t)kappa * r2;
    double ee = 0.0, oo = 0.0, eo = 0.0, oe = 0.0;
    for (int z = 0; z < ksplit; z++) {
        const double *P = partial + ((size_t)z * n + p) * plane;
        ee = addmod(ee, P[(2 * j) * r2 + 2 * k], M);
        oo = addmod(oo, P[(2 * j + 1) * r2 + 2 * k + 1], M);
        eo = addmod(eo, P[(2 * j) * r2 + 2 * k + 1], M);
        oe = addmod(oe, P[(2 * j + 1) * r2 + 2 * k], M);
    }
    uint64_t *o = out + jk * (size_t)(2 * n);
    o[p] = d2u(addmod(ee, mulmod(oo, omega[p], M), M));
    o[n + p] = d2u(addmod(eo, oe, M));
}

}  // namespace

int launch_wide_ntt(Context *ctx, const uint64_t *src, uint64_t *dst, size_t count, int mode,
                    int inverse) {
    if (count == 0) return SALSAA_OK;
    SALSAA_CHECK(ctx->d_wtab != nullptr, "degree tables missing");
    const bool inc = mode == SALSAA_REP_INCOMPLETE;
    SALSAA_CHECK(inc || ctx->full_ok,
                 "the NTT representation needs q = 1 mod 2d (no primitive 2d-th root of unity)");
    switch (ctx->d) {
        case 128: return run_wide_ntt<7>(ctx, src, dst, count, inc, inverse != 0);
        case 256: return run_wide_ntt<8>(ctx, src, dst, count, inc, inverse != 0);
        case 512: return run_wide_ntt<9>(ctx, src, dst, count, inc, inverse != 0);
        default: break;
    }
    set_error("no transform kernels for this degree");
    return SALSAA_ERR_ARG;
}

int launch_ntt_any(Context *ctx, uint64_t *data, size_t count, int mode, int inverse) {
    if (ctx->d == kD) return launch_ntt(ctx, data, count, mode, inverse);
    return launch_wide_ntt(ctx, data, data, count, mode, inverse);
}

int launch_wide_planes(Context *ctx, const uint64_t *src, size_t rows, size_t m, int8_t *dst,
                       size_t ktiles) {
    if (rows == 0 || m == 0) return SALSAA_OK;
    SALSAA_CHECK(ktiles == plane_ktiles(m), "plane K tiles do not match the length");
    SALSAA_CHECK(rows < 65536 && ktiles < (1u << 31), "plane grid out of range");
    const size_t d = (size_t)ctx->d;
    // the IncompleteNTT into scratch, then the pack: all rows at once when they
    // fit a chunk of kWideChunk elements, else row by row in chunks of whole K tiles
    constexpr size_t kWideChunk = (size_t)1 << 17;
    uint64_t *inc = nullptr;
    if (rows * m <= kWideChunk) {
        SALSAA_TRY(scratch_get(ctx, 16, rows * m * d * sizeof(uint64_t), (void **)&inc));
        SALSAA_TRY(launch_wide_ntt(ctx, src, inc, rows * m, SALSAA_REP_INCOMPLETE, 0));
        return run_pack_any(ctx, inc, m, 0, m, rows, 0, rows, 0, ktiles, ktiles, dst);
    }
    SALSAA_TRY(scratch_get(ctx, 16, kWideChunk * d * sizeof(uint64_t), (void **)&inc));
    for (size_t row = 0; row < rows; row++) {
        for (size_t i0 = 0; i0 < m; i0 += kWideChunk) {
            const size_t n = std::min(kWideChunk, m - i0);
            SALSAA_TRY(launch_wide_ntt(ctx, src + (row * m + i0) * d, inc, n,
                                       SALSAA_REP_INCOMPLETE, 0));
            // the row's last chunk also writes the zero columns up to 64 ktiles
            const size_t kt0 = i0 / 64;
            const size_t nkt = i0 + n < m ? n / 64 : ktiles - kt0;
            SALSAA_TRY(run_pack_any(ctx, inc, 0, i0, m, rows, row, 1, kt0, nkt, ktiles, dst));
        }
    }
    return SALSAA_OK;
}

int launch_commit_finalize_pairs(Context *ctx, const double *partial, size_t ksplit,
                                 size_t kappa2, size_t r2, uint64_t *out) {
    SALSAA_CHECK(kappa2 % 2 == 0 && r2 % 2 == 0, "slot-pair commit needs doubled rows");
    const size_t n = (size_t)ctx->d / 2;
    const WideTab W(ctx->d);
    int logn = 0;
    while ((1u << logn) < n) logn++;
    const size_t total = kappa2 / 2 * (r2 / 2) * n;
    Mod M{ctx->qd, ctx->qinv};
    hipLaunchKernelGGL(k_commit_finalize_pairs, dim3(ceil_div(total, 256)), dim3(256), 0,
                       ctx->stream, partial, (int)ksplit, (int)(kappa2 / 2), (int)(r2 / 2), logn,
                       out, ctx->d_wtab + W.omega, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

// ---------------------------------------------------------------- norm_check_3 at degree d
// norm_check_3 (fast_norm_check.rs:401-441) for d > 64: with m = L/d ring
// elements per witness row and challenges cs = [ring (log2 m) | coefficient
// (log2 d)] (the first the most significant index bit),
//     V = sum_i tensor(ring cs)_i w_i,   rhs = conj(tensor(coefficient cs)) * V,
// the second product in R_q[Y]/(Y^2 - nr) for F_{q^2} challenges; rhs in
// IncompleteNTT form.  Its constant term is <w, tensor(cs)>, whose square the
// verifier compares with the sumcheck's final claim (verify_norm_check_3,
// fast_norm_check.rs:443-462).
namespace {

// prod_b (bit (nb - 1 - b) of x ? c_b : 1 - c_b) over challenges c (F2
// pairs; F_q challenges when !fq2: the second coordinates are ignored)
__device__ __forceinline__ F2 tensor_at(const uint64_t *__restrict__ c, int nb, size_t x,
                                        double nr, bool fq2, Mod M) {
    F2 t = {1.0, 0.0};
    for (int b = 0; b < nb; b++) {
        const F2 cb = {u2d(c[2 * b]), fq2 ? u2d(c[2 * b + 1]) : 0.0};
        const F2 f = ((x >> (nb - 1 - b)) & 1) ? cb : f2_sub(F2{1.0, 0.0}, cb, M);
        t = f2_mul(t, f, nr, M);
    }
    return t;
}

// Partial sums of V: block (chunk, row, rep) over the chunk's 2^lo ring
// elements; vp [reps][rows][nchunk][d][2].
template <int LOGD>
__global__ __launch_bounds__(256) void k_nc3_fold(const uint64_t *__restrict__ w, size_t L,
                                                  int nring, int lo,
                                                  const uint64_t *__restrict__ cs, size_t nrounds,
                                                  double nr, bool fq2, double *__restrict__ vp,
                                                  Mod M) {
    constexpr int d = 1 << LOGD;
    constexpr int CW = d < 256 ? d : 256;    // coefficient lanes
    constexpr int IS = 256 / CW;             // elements summed side by side
    constexpr int NC = d / CW;               // coefficients per thread
    __shared__ F2 tau[256];
    __shared__ double sa[IS > 1 ? 256 : 1], sb[IS > 1 ? 256 : 1];
    const int tid = threadIdx.x;
    const size_t chunk = blockIdx.x, row = blockIdx.y, rep = blockIdx.z;
    const size_t nchunk = gridDim.x, rows = gridDim.y;
    const uint64_t *c = cs + rep * nrounds * 2;
    const int E = 1 << lo;
    if (tid < E)
        tau[tid] = f2_mul(tensor_at(c, nring - lo, chunk, nr, fq2, M),
                          tensor_at(c + 2 * (nring - lo), lo, (size_t)tid, nr, fq2, M), nr, M);
    __syncthreads();
    const int c0 = tid % CW, i0 = tid / CW;
    double a[NC], b[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) a[k] = b[k] = 0.0;
    const uint64_t *wr = w + row * L + chunk * (size_t)E * d;
    for (int i = i0; i < E; i += IS) {
        const F2 t = tau[i];
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const double x = u2d(wr[(size_t)i * d + c0 + CW * k]);
            a[k] = reduce_lazy(a[k] + mul_lazy(t.a, x, M), M);
            b[k] = reduce_lazy(b[k] + mul_lazy(t.b, x, M), M);
        }
    }
    double *o = vp + (((rep * rows + row) * nchunk + chunk) * d) * 2;
    if (IS > 1) {
        sa[tid] = a[0];
        sb[tid] = b[0];
        __syncthreads();
        if (i0 == 0) {
            double x = 0.0, y = 0.0;
            for (int p = 0; p < IS; p++) {
                x = reduce_lazy(x + sa[p * CW + c0], M);
                y = reduce_lazy(y + sb[p * CW + c0], M);
            }
            o[2 * c0] = reduce(x, M);
            o[2 * c0 + 1] = reduce(y, M);
        }
    } else {
#pragma unroll
        for (int k = 0; k < NC; k++) {
            o[2 * (c0 + CW * k)] = reduce(a[k], M);
            o[2 * (c0 + CW * k) + 1] = reduce(b[k], M);
        }
    }
}

// rhs of (row, rep) from the partial sums: block of d threads
template <int LOGD>
__global__ __launch_bounds__(512) void k_nc3_rhs(const double *__restrict__ vp, int nchunk,
                                                 const uint64_t *__restrict__ cs, size_t nrounds,
                                                 int nring, double nr, bool fq2,
                                                 const uint64_t *__restrict__ finals, size_t fe_rep,
                                                 size_t fe_row, uint64_t *__restrict__ rhs,
                                                 int32_t *__restrict__ ok, size_t ok_rep,
                                                 size_t ok_row, const double *__restrict__ wt,
                                                 Mod M) {
    constexpr int d = 1 << LOGD, N = d / 2, TPE = d / 8;
    __shared__ F2 V[d], B[d];
    __shared__ double R[2][d];
    __shared__ double xs_all[2 * d];
    const WideTab W(d);
    const int tid = threadIdx.x;
    const size_t row = blockIdx.x, rep = blockIdx.y, rows = gridDim.x;
    const uint64_t *c = cs + rep * nrounds * 2;
    {
        const double *p = vp + ((rep * rows + row) * (size_t)nchunk * d + tid) * 2;
        F2 v = {0.0, 0.0};
        for (int ch = 0; ch < nchunk; ch++) {
            v.a = addmod(v.a, p[(size_t)ch * d * 2], M);
            v.b = addmod(v.b, p[(size_t)ch * d * 2 + 1], M);
        }
        V[tid] = v;
        // base = conj(T): base[0] = T[0], base[i] = -T[d - i]
        const F2 T = tensor_at(c + 2 * nring, LOGD, (size_t)tid, nr, fq2, M);
        if (tid == 0)
            B[0] = T;
        else
            B[d - tid] = f2_sub(F2{0.0, 0.0}, T, M);
    }
    __syncthreads();
    {
        // coefficient i of base * V (negacyclic)
        const int i = tid;
        double r0 = 0.0, r1 = 0.0;
        for (int j = 0; j < d; j++) {
            const int vi = i - j;
            const bool wrap = vi < 0;
            const F2 bb = B[j], vv = V[wrap ? vi + d : vi];
            const double p0 = mul_lazy(bb.a, vv.a, M) + mul_lazy(nr, mulmod(bb.b, vv.b, M), M);
            const double p1 = mul_lazy(bb.a, vv.b, M) + mul_lazy(bb.b, vv.a, M);
            r0 = reduce_lazy(wrap ? r0 - p0 : r0 + p0, M);
            r1 = reduce_lazy(wrap ? r1 - p1 : r1 + p1, M);
        }
        R[0][i] = reduce(r0, M);
        R[1][i] = reduce(r1, M);
    }
    __syncthreads();
    if (tid == 0) {
        const F2 ct = {R[0][0], R[1][0]};
        const F2 ct2 = f2_mul(ct, ct, nr, M);
        const uint64_t *f = finals + rep * fe_rep + row * fe_row;
        if (d2u(ct2.a) != f[0] || d2u(ct2.b) != f[1]) ok[rep * ok_rep + row * ok_row] = 0;
    }
    if (tid < 2 * TPE) {
        // the IncompleteNTT of component comp: d/8 threads of one wavefront
        const int comp = tid / TPE, l = tid % TPE, h = l & 1, u = l >> 1;
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = R[comp][l + k * TPE];
        wide_fwd<LOGD - 1>(v, xs_all + comp * d + h * N, u, wt + W.fwd_h, M);
        uint64_t *dst = rhs + (rep * rows + row) * 2 * d + comp * d + h * N + 8 * u;
#pragma unroll
        for (int k = 0; k < 8; k++) dst[k] = d2u(v[k]);
    }
}

template <int LOGD>
int run_nc3(Context *ctx, const uint64_t *w, size_t rows, size_t L, const uint64_t *cs,
            size_t reps, double nr, bool fq2, const uint64_t *finals, size_t fe_rep, size_t fe_row,
            uint64_t *rhs, int32_t *ok, size_t ok_rep, size_t ok_row) {
    constexpr int d = 1 << LOGD;
    Mod M{ctx->qd, ctx->qinv};
    const int nrounds = __builtin_ctzll(L), nring = nrounds - LOGD;
    const int lo = nring < 8 ? nring : 8;
    const size_t nchunk = (L / d) >> lo;
    SALSAA_CHECK(nchunk < (1u << 31), "norm_check_3 length out of range");
    double *vp = nullptr;
    SALSAA_TRY(scratch_get(ctx, 15, reps * rows * nchunk * d * 2 * sizeof(double), (void **)&vp));
    hipLaunchKernelGGL((k_nc3_fold<LOGD>), dim3((unsigned)nchunk, (unsigned)rows, (unsigned)reps),
                       dim3(256), 0, ctx->stream, w, L, nring, lo, cs, (size_t)nrounds, nr, fq2, vp,
                       M);
    SALSAA_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_nc3_rhs<LOGD>), dim3((unsigned)rows, (unsigned)reps), dim3(d), 0,
                       ctx->stream, vp, (int)nchunk, cs, (size_t)nrounds, nring, nr, fq2, finals,
                       fe_rep, fe_row, rhs, ok, ok_rep, ok_row, ctx->d_wtab, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

}  // namespace

int launch_nc3_wide(Context *ctx, const uint64_t *w, size_t rows, size_t L, const uint64_t *cs,
                    size_t reps, double nr, bool fq2, const uint64_t *finals, size_t fe_rep, size_t fe_row,
                    uint64_t *rhs, int32_t *ok, size_t ok_rep, size_t ok_row) {
    SALSAA_CHECK(L >= (size_t)ctx->d && (L & (L - 1)) == 0, "norm_check_3 needs L >= d");
    switch (ctx->d) {
        case 128:
            return run_nc3<7>(ctx, w, rows, L, cs, reps, nr, fq2, finals, fe_rep, fe_row, rhs, ok,
                              ok_rep, ok_row);
        case 256:
            return run_nc3<8>(ctx, w, rows, L, cs, reps, nr, fq2, finals, fe_rep, fe_row, rhs, ok,
                              ok_rep, ok_row);
        case 512:
            return run_nc3<9>(ctx, w, rows, L, cs, reps, nr, fq2, finals, fe_rep, fe_row, rhs, ok,
                              ok_rep, ok_row);
        default: break;
    }
    set_error("no norm_check_3 kernels for this degree");
    return SALSAA_ERR_ARG;
}

}  // namespace salsaa
