// protocol.hip -- the SNARK and folding protocol run on the device
// (protocol.rs:125-483): every prover operator and every verifier relation of
// each round, with the witness, the growing statement and the verifier state
// resident in HBM, and one host synchronisation at the end.
//
// The statement (the reference's Vec<PowerSeries>) is a list of rows:
//   CRS rows j < kappa   the power series m_j, m_j^2, .. (crs.rs:43-72); the
//                        commit kernels generate them from the module
//   norm-check rows      norm_check_3's series (fast_norm_check.rs:401-441):
//                        base conj(tensor(last 6 challenges)), tensors
//                        [1 - c_i, c_i] for the other challenges (m0 = 1)
//   batching rows        witness_lhs and projected_lhs (project.rs:176-185):
//                        base layers with all-ones tensors, so every layer is
//                        a prefix of the longest one
// The layer of length l of a series with n tensors over a base of m0 elements
// is the product of the base with the last k = log2(l / m0) tensors
// (complete_power_series, arithmetic.rs:158-176), and verifier_split's
// multipliers at witness length l are tensor n - k (split.rs:187-211).  No
// layer is expanded: a periodic row commits through the witness summed over
// its period, and a norm-check row's layer is a scalar table (the product of
// its constant tensors) times its base, so it commits as the base times the
// table-weighted sum of the witness; all rows of each kind share one pass
// over the witness.
//
// The verifier state V [R][c][d] (IncompleteNTT) has one row per statement
// row and one column per witness row.  Each relation the reference asserts
// writes a nonzero code into a device status word (the first failure wins):
// 100 * round + check, check 1 verify_decomp, 2 the norm check's verifier,
// 3 verifier_split, 4 verify_batching, 5 the folding scheme's closing
// verify_decomp, 6 the final commitment check (protocol.rs:464-472).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "handles.hpp"

using namespace salsaa;

namespace {

constexpr size_t kHeight = 256;  // HEIGHT of the projection (project.rs:17)

enum Phase { PH_COMMIT, PH_DECOMP, PH_NORM, PH_SPLIT, PH_PROJECT, PH_FOLD, PH_JOIN, PH_FINAL };

inline unsigned grid1(size_t n, unsigned bs) {
    size_t g = (n + bs - 1) / bs;
    return (unsigned)(g < 65536 ? (g ? g : 1) : 65536);
}

// the fault injection of salsaa_protocol_ex: one value off by one (mod q)
__global__ void k_fault(uint64_t *p, uint64_t q) {
    if (threadIdx.x == 0) {
        const uint64_t v = *p + 1;
        *p = v >= q ? v - q : v;
    }
}

__global__ void k_status_ok(const int32_t *__restrict__ ok, size_t n, int32_t code,
                            int32_t *status) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        if (ok[i] != 1) atomicCAS(status, 0, code);
}

__global__ void k_status_count(const uint32_t *__restrict__ count, int32_t code, int32_t *status) {
    if (threadIdx.x == 0 && *count != 0) atomicCAS(status, 0, code);
}

__global__ void k_status_equal(const uint64_t *__restrict__ a, const uint64_t *__restrict__ b,
                               size_t n, int32_t code, int32_t *status) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        if (a[i] != b[i]) atomicCAS(status, 0, code);
}

// verifier_split: prod [R][2r][d] = the split commitments times the row
// multipliers (IncompleteNTT); column k and r + k must add up to V[e][k].
__global__ void k_split_check(const uint64_t *__restrict__ prod, const uint64_t *__restrict__ V,
                              size_t R, size_t r, uint64_t q, int32_t code, int32_t *status) {
    const size_t n = R * r * kD;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t e = i / (r * kD), k = (i / kD) % r, c = i % kD;
        const uint64_t *p = prod + e * 2 * r * kD;
        uint64_t s = p[k * kD + c] + p[(r + k) * kD + c];
        if (s >= q) s -= q;
        if (s != V[i]) atomicCAS(status, 0, code);
    }
}

// M[e][k] = T[e][k >= r] for T [R][2][d]: the multiplier of each split column
__global__ void k_expand_mult(const uint64_t *__restrict__ T, size_t R, size_t r,
                              uint64_t *__restrict__ M) {
    const size_t n = R * 2 * r * kD;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t e = i / (2 * r * kD), k = (i / kD) % (2 * r), c = i % kD;
        M[i] = T[(e * 2 + (k >= r ? 1 : 0)) * kD + c];
    }
}

// out [b][a][d] = in [a][b][d] transposed over ring elements
__global__ void k_transpose_ring(const uint64_t *__restrict__ in, size_t a, size_t b,
                                 uint64_t *__restrict__ out) {
    const size_t n = a * b * kD;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t x = i / (b * kD), y = (i / kD) % b, c = i % kD;
        out[(y * a + x) * kD + c] = in[i];
    }
}

// challenge_for_project (project.rs:26-39): one random bit per byte
__global__ void k_bits(uint8_t *__restrict__ out, size_t n, uint64_t key) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = (uint8_t)(draw(key, i) & 1);
}

// norm_check_3's power series (fast_norm_check.rs:401-429) from one
// repetition's challenges cs [nrounds][2] (coordinate 0): base [d] =
// conj(tensor(cs[nrounds-6 ..])) with the first of those the most
// significant index bit; tens [nrounds-6][2][d] = constants (1 - c_i, c_i).
__global__ void k_nc_series(const uint64_t *__restrict__ cs, int nrounds, Mod M,
                            uint64_t *__restrict__ base, uint64_t *__restrict__ tens) {
    const int tid = threadIdx.x;
    const int n = nrounds - 6;
    const uint64_t q = (uint64_t)M.q;
    if (tid < kD) {
        double t = 1.0;
        for (int b = 0; b < 6; b++) {
            const uint64_t c = cs[2 * (n + b)];
            const uint64_t f = ((tid >> (5 - b)) & 1) ? c : (c == 0 ? 1 : (q + 1 - c) % q);
            t = mulmod(t, u2d(f), M);
        }
        const uint64_t tv = d2u(t);
        base[tid == 0 ? 0 : kD - tid] = (tid == 0 || tv == 0) ? tv : q - tv;
    }
    for (int i = tid; i < n * 2 * kD; i += blockDim.x) {
        const int t = i / (2 * kD), side = (i / kD) & 1, c = i % kD;
        const uint64_t ch = cs[2 * t];
        uint64_t v = 0;
        if (c == 0) v = side ? ch : (ch == 0 ? 1 : (q + 1 - ch) % q);
        tens[i] = v;
    }
}

// The scalar factor of a norm-check row's layer of length 2^k:
// w[x] = prod_j tens[n - k + j][bit j of x, from the top][0] (the tensors are
// constants, so the layer is w[x] times the base element).
__global__ void k_eq_table(const uint64_t *__restrict__ tens, int n, int k, Mod M,
                           uint64_t *__restrict__ w) {
    const size_t len = (size_t)1 << k;
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < len;
         x += (size_t)gridDim.x * blockDim.x) {
        double t = 1.0;
        for (int j = 0; j < k; j++) {
            const int bit = (int)((x >> (k - 1 - j)) & 1);
            t = mulmod(t, u2d(tens[((size_t)(n - k + j) * 2 + bit) * kD]), M);
        }
        w[x] = d2u(t);
    }
}

constexpr int EV_MAXE = 8;   // norm-check rows per pass over the witness
constexpr int EV_BLOCKS = 4096;

// acc[e][k][c] += sum over this block's x of w[e][x] W[k][x][c]: the four
// waves of a block take every fourth element, one coefficient per lane (a
// wave reads one ring element, 512 contiguous bytes, per step).  Partial sums
// are reduced mod q before the atomic, and at most EV_BLOCKS blocks add to
// an entry, so acc stays below 2^62.
__global__ __launch_bounds__(256) void k_eval_rows(const uint64_t *__restrict__ W, size_t r,
                                                   size_t l, const uint64_t *__restrict__ w, int E,
                                                   Mod M, unsigned long long *__restrict__ acc) {
    __shared__ double red[4][EV_MAXE][kD];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    for (size_t k = 0; k < r; k++) {
        double s[EV_MAXE];
#pragma unroll
        for (int e = 0; e < EV_MAXE; e++) s[e] = 0.0;
        for (size_t x = (size_t)blockIdx.x * 4 + g; x < l; x += (size_t)gridDim.x * 4) {
            const double v = u2d(W[(k * l + x) * kD + c]);
#pragma unroll
            for (int e = 0; e < EV_MAXE; e++)
                if (e < E) {
                    s[e] += mulmod(u2d(w[e * l + x]), v, M);
                    if (s[e] >= M.q) s[e] -= M.q;
                }
        }
#pragma unroll
        for (int e = 0; e < EV_MAXE; e++)
            if (e < E) red[g][e][c] = s[e];
        __syncthreads();
        if (g == 0)
            for (int e = 0; e < E; e++) {
                double t = red[0][e][c];
                for (int h = 1; h < 4; h++) {
                    t += red[h][e][c];
                    if (t >= M.q) t -= M.q;
                }
                atomicAdd(&acc[(e * r + k) * kD + c], (unsigned long long)t);
            }
        __syncthreads();
    }
}

__global__ void k_zero_acc(unsigned long long *__restrict__ a, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        a[i] = 0;
}

__global__ void k_eval_finish(const unsigned long long *__restrict__ acc, size_t n, uint64_t q,
                              uint64_t *__restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = acc[i] % q;
}

// verifier_split's multipliers T [R][2][d] of up to kMultRows rows per launch:
// kind 0 [1, 1] (all-ones tensors), 1 [1, src] (CRS row: src = m_j^(l/2)),
// 2 the tensor [2][d] at src (norm-check row).
constexpr int kMultRows = 64;
struct MultSrc {
    const uint64_t *src[kMultRows];
    int kind[kMultRows];
};

__global__ void k_build_mult(MultSrc a, int rows, uint64_t *__restrict__ T) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < rows * 2 * kD;
         i += gridDim.x * blockDim.x) {
        const int e = i / (2 * kD), side = (i / kD) & 1, c = i % kD;
        uint64_t v = c == 0 ? 1 : 0;
        if (a.kind[e] == 1 && side == 1) v = a.src[e][c];
        if (a.kind[e] == 2) v = a.src[e][side * kD + c];
        T[i] = v;
    }
}

// Device buffers of one run: bump allocation from slabs the context keeps
// between runs, so a run like the one before it calls no allocator at all.
// When a run ends, the slabs it did not touch are freed: the context keeps
// what the last run used, and salsaa_ctx_trim returns that too.
constexpr size_t kSlabMin = (size_t)256 << 20;

struct Arena {
    Context *c;
    std::vector<size_t> used;  // bytes taken from each slab by this run
    size_t cur = 0;            // the slab requests are taken from
    explicit Arena(Context *ctx) : c(ctx), used(ctx->slabs.size(), 0) {}
    ~Arena() {
        // the run's kernels may still be in flight on an error path
        (void)hipStreamSynchronize(c->stream);
        std::vector<ScratchBuf> keep;
        for (size_t k = 0; k < c->slabs.size(); k++) {
            if (used[k])
                keep.push_back(c->slabs[k]);
            else
                (void)hipFree(c->slabs[k].ptr);
        }
        c->slabs.swap(keep);
    }
    template <class T>
    int get(size_t n, T **out) {
        const size_t b = ((n ? n : 1) * sizeof(T) + 255) / 256 * 256;
        std::vector<ScratchBuf> &S = c->slabs;
        while (cur < S.size() && used[cur] + b > S[cur].bytes) cur++;
        if (cur == S.size()) {
            ScratchBuf nb;
            nb.bytes = std::max(b, kSlabMin);
            if (hipMalloc(&nb.ptr, nb.bytes) != hipSuccess) {
                (void)hipGetLastError();
                set_error("protocol: out of device memory");
                return SALSAA_ERR_NOMEM;
            }
            S.push_back(nb);
            used.push_back(0);
        }
        *out = reinterpret_cast<T *>(static_cast<char *>(S[cur].ptr) + used[cur]);
        used[cur] += b;
        return SALSAA_OK;
    }
};

// A statement row other than the CRS rows.
struct Row {
    int crs = -1;                      // CRS row index, or -1
    const uint64_t *base = nullptr;    // base layer [m0][d] (periodic: the full layer)
    size_t m0 = 0;
    const uint64_t *tens = nullptr;    // [n][2][d]; nullptr: all-ones tensors
    size_t n = 0;                      // tensors; the longest layer has m0 << n elements
};

inline bool pow2(size_t x) { return x && !(x & (x - 1)); }
inline int lg2(size_t x) {
    int k = 0;
    while ((size_t)1 << k < x) k++;
    return k;
}

struct Run {
    salsaa_ctx *ctx;
    Context *c;
    salsaa_key *key;
    const uint64_t *module;
    size_t kappa;
    std::vector<Row> rows;  // rows[e] for e >= kappa
    uint64_t *one = nullptr;
    Arena *A;

    hipStream_t st() const { return c->stream; }

    // a non-CRS row must have a layer of the witness length l (m0 << k for
    // some k <= n); the reference panics otherwise
    int check_layer(const Row &w, size_t l) {
        SALSAA_CHECK(l >= w.m0 && l % w.m0 == 0 && pow2(l / w.m0) && lg2(l / w.m0) <= (int)w.n,
                     "protocol: a statement row has no layer of the witness length (the "
                     "reference finds none either)");
        return SALSAA_OK;
    }

    // parallel_dot_series_matrix over the whole statement: out [R][r][d].  A
    // periodic row needs only the witness summed over its period,
    // sum_x base[x mod m0] W[x] = sum_j base[j] sum_t W[t m0 + j]: the sums
    // over the longest period P are taken in one pass over W and folded to
    // each shorter period.
    int commit(const uint64_t *W, size_t r, size_t l, uint64_t *out) {
        SALSAA_TRY(salsaa_parallel_dot_series_matrix(ctx, key, W, r, l, out));
        size_t P = 0;
        for (size_t e = 0; e < rows.size(); e++) {
            SALSAA_TRY(check_layer(rows[e], l));
            if (rows[e].tens == nullptr) P = std::max(P, rows[e].m0);
        }
        uint64_t *SP = nullptr, *Ss = nullptr;
        unsigned long long *acc = nullptr;
        if (P) {
            SALSAA_TRY(A->get(r * P * kD, &SP));
            SALSAA_TRY(A->get(r * P * kD, &Ss));
            SALSAA_TRY(A->get(r * P * kD, &acc));
            SALSAA_TRY(launch_chunk_sums(c, W, r, l, P, SP, acc));
        }
        for (size_t e = 0; e < rows.size(); e++) {
            const Row &w = rows[e];
            uint64_t *o = out + (kappa + e) * r * kD;
            if (w.tens == nullptr) {
                const uint64_t *S = SP;
                if (w.m0 != P) {
                    SALSAA_TRY(launch_chunk_sums(c, SP, r, P, w.m0, Ss, acc));
                    S = Ss;
                }
                SALSAA_TRY(salsaa_parallel_dot_single_series_matrix(ctx, w.base, w.m0, S, r, o));
            }
        }
        // norm-check rows: layer[x] = w_e[x] base_e, so the row's commitment is
        // base_e times the w_e-weighted sum of the witness, for up to EV_MAXE
        // rows per pass over W
        std::vector<size_t> nrow;
        for (size_t e = 0; e < rows.size(); e++)
            if (rows[e].tens) nrow.push_back(e);
        if (!nrow.empty()) {
            const int k = lg2(l);
            const Mod M{c->qd, c->qinv};
            uint64_t *wts = nullptr, *V = nullptr, *Brep = nullptr;
            unsigned long long *acc = nullptr;
            SALSAA_TRY(A->get(EV_MAXE * l, &wts));
            SALSAA_TRY(A->get(EV_MAXE * r * kD, &V));
            SALSAA_TRY(A->get(r * kD, &Brep));
            SALSAA_TRY(A->get(EV_MAXE * r * kD, &acc));
            const unsigned nb = (unsigned)std::min<size_t>((l + 3) / 4, EV_BLOCKS);
            for (size_t g0 = 0; g0 < nrow.size(); g0 += EV_MAXE) {
                const int E = (int)std::min<size_t>(EV_MAXE, nrow.size() - g0);
                for (int i = 0; i < E; i++) {
                    const Row &w = rows[nrow[g0 + i]];
                    hipLaunchKernelGGL(k_eq_table, dim3(grid1(l, 256)), dim3(256), 0, st(), w.tens,
                                       (int)w.n, k, M, wts + (size_t)i * l);
                }
                hipLaunchKernelGGL(k_zero_acc, dim3(grid1(E * r * kD, 256)), dim3(256), 0, st(),
                                   acc, (size_t)E * r * kD);
                hipLaunchKernelGGL(k_eval_rows, dim3(nb), dim3(256), 0, st(), W, r, l, wts, E, M,
                                   acc);
                hipLaunchKernelGGL(k_eval_finish, dim3(grid1(E * r * kD, 256)), dim3(256), 0,
                                   st(), acc, (size_t)E * r * kD, c->q, V);
                SALSAA_HIP(hipGetLastError());
                for (int i = 0; i < E; i++) {
                    const size_t e = nrow[g0 + i];
                    SALSAA_TRY(launch_repeat(c, rows[e].base, 1, r, Brep));
                    SALSAA_TRY(salsaa_ring_mul(ctx, V + (size_t)i * r * kD, Brep,
                                               out + (kappa + e) * r * kD, r));
                }
            }
        }
        return SALSAA_OK;
    }

    size_t R() const { return kappa + rows.size(); }

    // verifier_split's multipliers at witness length l: T [R][2][d] Coefficient
    int multipliers(size_t l, uint64_t *T) {
        const size_t b = kD * sizeof(uint64_t);
        // CRS rows: [1, m_j^(l/2)] by repeated squaring of the module
        uint64_t *pw = nullptr, *sq = nullptr;
        SALSAA_TRY(A->get(kappa * kD, &pw));
        SALSAA_TRY(A->get(kappa * kD, &sq));
        SALSAA_HIP(hipMemcpyAsync(pw, module, kappa * b, hipMemcpyDeviceToDevice, st()));
        for (size_t e = 1; e < l / 2; e <<= 1) {
            SALSAA_TRY(salsaa_ring_mul(ctx, pw, pw, sq, kappa));
            SALSAA_TRY(salsaa_convert(ctx, sq, kappa, SALSAA_REP_INCOMPLETE, SALSAA_REP_COEFF));
            std::swap(pw, sq);
        }
        // one table of sources per kMultRows rows
        const size_t R_ = R();
        for (size_t e0 = 0; e0 < R_; e0 += kMultRows) {
            MultSrc a = {};
            const int n = (int)std::min<size_t>(kMultRows, R_ - e0);
            for (int i = 0; i < n; i++) {
                const size_t e = e0 + i;
                if (e < kappa) {
                    a.kind[i] = 1;
                    a.src[i] = pw + e * kD;
                } else {
                    const Row &w = rows[e - kappa];
                    if (w.tens == nullptr) {
                        a.kind[i] = 0;
                    } else {
                        const size_t k = (size_t)lg2(l / w.m0);
                        SALSAA_CHECK(k >= 1 && k <= w.n, "protocol: split below a row's base layer");
                        a.kind[i] = 2;
                        a.src[i] = w.tens + (w.n - k) * 2 * kD;
                    }
                }
            }
            hipLaunchKernelGGL(k_build_mult, dim3(grid1((size_t)n * 2 * kD, 256)), dim3(256), 0,
                               st(), a, n, T + e0 * 2 * kD);
            SALSAA_HIP(hipGetLastError());
        }
        return SALSAA_OK;
    }

    // decomp_ell against the whole statement + verify_decomp (decomp.rs:32-112);
    // corrupt: the decomposed witness's commitment off by one (fault
    // injection); dst: where the new witness goes (null: a run buffer)
    int decomp(uint64_t **W, size_t *r, size_t m, size_t parts, uint64_t **V, int32_t code,
               int32_t *status, bool corrupt, uint64_t *dst = nullptr) {
        uint64_t *W2 = dst, *radix = nullptr, *Vd = nullptr;
        uint32_t *cnt = nullptr;
        const size_t rp = *r * parts;
        if (!W2) SALSAA_TRY(A->get(rp * m * kD, &W2));
        SALSAA_TRY(A->get(1, &radix));
        SALSAA_TRY(A->get(R() * rp * kD, &Vd));
        SALSAA_TRY(A->get(1, &cnt));
        SALSAA_TRY(salsaa_decompose(ctx, *W, *r, m, parts, 0, W2, radix));
        SALSAA_TRY(commit(W2, rp, m, Vd));
        if (corrupt) hipLaunchKernelGGL(k_fault, dim3(1), dim3(64), 0, st(), Vd, c->q);
        SALSAA_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t), st()));
        SALSAA_TRY(salsaa_verify_decomp(ctx, Vd, *V, R(), *r, parts, radix, 0, cnt));
        hipLaunchKernelGGL(k_status_count, dim3(1), dim3(64), 0, st(), cnt, code, status);
        SALSAA_HIP(hipGetLastError());
        *W = W2;
        *r = rp;
        *V = Vd;
        return SALSAA_OK;
    }
};

int run_protocol(salsaa_ctx *ctx, const uint64_t *d_module, size_t kappa, size_t m,
                 const uint64_t *d_W, size_t r, size_t nrounds, const size_t *parts, int folding,
                 uint64_t seed, uint64_t *d_W_out, size_t w_cap, uint64_t *d_rhs_out,
                 size_t rhs_cap, size_t *shape, int32_t *status_out, double *times,
                 const salsaa_protocol_opts *opts) {
    Context *c = &ctx->c;
    SALSAA_CHECK(!opts || (opts->fault_check >= 0 && opts->fault_check <= 6),
                 "protocol: fault_check must be 0..6");
    // the fault injection point of check `check` in round `round`
    auto fault = [&](size_t round, int check) {
        return opts && opts->fault_check == check && (check == 6 || opts->fault_round == round);
    };
    SALSAA_CHECK(d_module && d_W && status_out && shape, "protocol: null argument");
    SALSAA_CHECK(nrounds >= 1 && (parts || folding), "protocol: no rounds");
    SALSAA_CHECK(kappa >= 1 && kappa <= kMfmaMaxKappa, "protocol: kappa out of range");
    SALSAA_CHECK(pow2(m) && m * kD >= 128 && pow2(r), "protocol: m and r must be powers of two");
    const hipStream_t st = c->stream;
    const size_t b = kD * sizeof(uint64_t);
    const Mod M{c->qd, c->qinv};
    Arena A(c);
    Run run{ctx, c, nullptr, d_module, kappa, {}, nullptr, &A};
    SALSAA_TRY(salsaa_key_crs_implicit(ctx, d_module, kappa, m, &run.key));
    struct KeyGuard {
        salsaa_key *k;
        hipStream_t s;
        ~KeyGuard() {
            (void)hipStreamSynchronize(s);
            salsaa_key_destroy(k);
        }
    } kg{run.key, st};
    int32_t *status = nullptr;
    SALSAA_TRY(A.get(1, &status));
    SALSAA_TRY(A.get(kD, &run.one));
    SALSAA_HIP(hipMemsetAsync(status, 0, sizeof(int32_t), st));
    SALSAA_HIP(hipMemsetAsync(run.one, 0, b, st));
    {
        const uint64_t one = 1;
        SALSAA_HIP(hipMemcpyAsync(run.one, &one, sizeof one, hipMemcpyHostToDevice, st));
        SALSAA_HIP(hipStreamSynchronize(st));
    }
    // phase timers: an event at every phase boundary
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_phase;
    struct EvGuard {
        std::vector<hipEvent_t> &e;
        ~EvGuard() {
            for (auto x : e) (void)hipEventDestroy(x);
        }
    } eg{ev};
    auto mark = [&](int phase) -> int {
        hipEvent_t e;
        SALSAA_HIP(hipEventCreate(&e));
        ev.push_back(e);
        ev_phase.push_back(phase);
        SALSAA_HIP(hipEventRecord(e, st));
        return SALSAA_OK;
    };
    auto code = [](size_t round, int check) { return (int32_t)(100 * round + check); };

    // the initial commitment (protocol.rs:138)
    uint64_t *W = const_cast<uint64_t *>(d_W);
    uint64_t *V = nullptr;
    SALSAA_TRY(mark(-1));
    SALSAA_TRY(A.get(kappa * r * kD, &V));
    SALSAA_TRY(run.commit(W, r, m, V));
    SALSAA_TRY(mark(PH_COMMIT));
    size_t rounds_done = 0;
    for (size_t i = 0; i < nrounds; i++) {
        const uint64_t rseed = stream_key(seed, 0x7000 + i);
        // decomp_ell + verify_decomp (protocol.rs:166-196)
        const size_t p = parts ? parts[i] : 1;
        if (!folding && p > 1) {
            SALSAA_CHECK(pow2(p), "protocol: decomposition parts must be a power of two");
            SALSAA_TRY(run.decomp(&W, &r, m, p, &V, code(i, 1), status, fault(i, 1)));
            SALSAA_TRY(mark(PH_DECOMP));
        }
        // norm_check_1/2/3 with their verifiers, two repetitions (:198-259)
        {
            const size_t L = m * kD;
            const int nr = lg2(L);
            uint64_t *cs = nullptr, *rhs = nullptr, *V2 = nullptr;
            int32_t *ok = nullptr;
            SALSAA_TRY(A.get(2 * nr * 2, &cs));
            SALSAA_TRY(A.get(2 * r * 2 * kD, &rhs));
            SALSAA_TRY(A.get(2 * r, &ok));
            SALSAA_TRY(salsaa_sample_challenges(ctx, cs, 2, nr, 0, rseed));
            const NcOut no = nc_out_plain(r, 2, (size_t)nr, nullptr, nullptr, nullptr, nullptr, rhs,
                                          ok);
            SALSAA_TRY(launch_norm_check_ex(c, W, nullptr, r, L, cs, 0, 2, 0, 0, no, 0, nullptr,
                                            nullptr, fault(i, 2) ? 1 : 0));
            hipLaunchKernelGGL(k_status_ok, dim3(1), dim3(64), 0, st, ok, 2 * r, code(i, 2),
                               status);
            SALSAA_HIP(hipGetLastError());
            const size_t R0 = run.R();
            SALSAA_TRY(A.get((R0 + 2) * r * kD, &V2));
            SALSAA_HIP(hipMemcpyAsync(V2, V, R0 * r * b, hipMemcpyDeviceToDevice, st));
            for (int rep = 0; rep < 2; rep++) {
                uint64_t *base = nullptr, *tens = nullptr;
                SALSAA_TRY(A.get(kD, &base));
                SALSAA_TRY(A.get((size_t)(nr - 6) * 2 * kD, &tens));
                hipLaunchKernelGGL(k_nc_series, dim3(1), dim3(256), 0, st,
                                   cs + (size_t)rep * nr * 2, nr, M, base, tens);
                SALSAA_HIP(hipGetLastError());
                run.rows.push_back(Row{-1, base, 1, tens, (size_t)(nr - 6)});
                // verify_norm_check_3 pushes rhs (component 0 of each row)
                SALSAA_HIP(hipMemcpy2DAsync(V2 + (R0 + rep) * r * kD, b,
                                            rhs + (size_t)rep * r * 2 * kD, 2 * b, b, r,
                                            hipMemcpyDeviceToDevice, st));
            }
            V = V2;
            SALSAA_TRY(mark(PH_NORM));
        }
        // split + verifier_split (:261-291)
        if (!folding) {
            SALSAA_CHECK(m >= 2, "protocol: witness too short to split");
            const size_t h = m / 2, R = run.R();
            uint64_t *W3 = nullptr, *S = nullptr, *T = nullptr, *Mx = nullptr, *Sc = nullptr,
                     *prod = nullptr;
            SALSAA_TRY(A.get(2 * r * h * kD, &W3));
            SALSAA_TRY(A.get(R * 2 * r * kD, &S));
            SALSAA_TRY(A.get(R * 2 * kD, &T));
            SALSAA_TRY(A.get(R * 2 * r * kD, &Mx));
            SALSAA_TRY(A.get(R * 2 * r * kD, &Sc));
            SALSAA_TRY(A.get(R * 2 * r * kD, &prod));
            // rows L_0 .. L_{r-1}, R_0 .. R_{r-1} (split.rs:21-44)
            SALSAA_HIP(hipMemcpy2DAsync(W3, h * b, W, m * b, h * b, r, hipMemcpyDeviceToDevice, st));
            SALSAA_HIP(hipMemcpy2DAsync(W3 + r * h * kD, h * b, W + h * kD, m * b, h * b, r,
                                        hipMemcpyDeviceToDevice, st));
            SALSAA_TRY(run.commit(W3, 2 * r, h, S));
            if (fault(i, 3)) hipLaunchKernelGGL(k_fault, dim3(1), dim3(64), 0, st, S, c->q);
            SALSAA_TRY(run.multipliers(m, T));
            hipLaunchKernelGGL(k_expand_mult, dim3(grid1(R * 2 * r * kD, 256)), dim3(256), 0, st,
                               T, R, r, Mx);
            SALSAA_HIP(hipGetLastError());
            SALSAA_HIP(hipMemcpyAsync(Sc, S, R * 2 * r * b, hipMemcpyDeviceToDevice, st));
            SALSAA_TRY(salsaa_convert(ctx, Sc, R * 2 * r, SALSAA_REP_INCOMPLETE, SALSAA_REP_COEFF));
            SALSAA_TRY(salsaa_ring_mul(ctx, Sc, Mx, prod, R * 2 * r));
            hipLaunchKernelGGL(k_split_check, dim3(grid1(R * r * kD, 256)), dim3(256), 0, st, prod,
                               V, R, r, c->q, code(i, 3), status);
            SALSAA_HIP(hipGetLastError());
            W = W3;
            r *= 2;
            m = h;
            V = S;
            SALSAA_TRY(mark(PH_SPLIT));
        }
        // project, its commitment, batch_projections + verify_batching (:293-348)
        const size_t width = kHeight * r;
        SALSAA_CHECK(width <= m && m % width == 0 && pow2(m / width),
                     "protocol: the projection needs 256 r <= m (project.rs panics otherwise)");
        uint64_t *P = nullptr, *pcom = nullptr, *plhs = nullptr, *wlhs = nullptr, *prhs = nullptr,
                 *wrhs = nullptr;
        // the joined witness [fold output; projection] (witness.append,
        // protocol.rs:402-406) is one buffer: the projection goes to its second
        // row and the fold to its first, so the join copies nothing; in the
        // SNARK's last round it is the caller's output
        uint64_t *J = nullptr;
        if (!folding && i + 1 == nrounds && d_W_out && w_cap >= 2 * m * kD)
            J = d_W_out;
        else
            SALSAA_TRY(A.get(2 * m * kD, &J));
        P = J + m * kD;
        {
            uint8_t *z = nullptr, *sg = nullptr;
            int32_t *bok = nullptr;
            SALSAA_TRY(A.get(kHeight * width, &z));
            SALSAA_TRY(A.get(kHeight * width, &sg));
            SALSAA_TRY(A.get(run.R() * kD, &pcom));
            SALSAA_TRY(A.get(m * kD, &plhs));
            SALSAA_TRY(A.get(m * kD, &wlhs));
            SALSAA_TRY(A.get(kD, &prhs));
            SALSAA_TRY(A.get(r * kD, &wrhs));
            SALSAA_TRY(A.get(1, &bok));
            hipLaunchKernelGGL(k_bits, dim3(grid1(kHeight * width, 256)), dim3(256), 0, st, z,
                               kHeight * width, stream_key(rseed, 1));
            hipLaunchKernelGGL(k_bits, dim3(grid1(kHeight * width, 256)), dim3(256), 0, st, sg,
                               kHeight * width, stream_key(rseed, 2));
            SALSAA_HIP(hipGetLastError());
            SALSAA_TRY(salsaa_project(ctx, W, r, m, z, sg, P));
            SALSAA_TRY(run.commit(P, 1, m, pcom));
            if (fault(i, 4)) hipLaunchKernelGGL(k_fault, dim3(1), dim3(64), 0, st, P, c->q);
            // batching_challenge_for_project's first element is the constant 1
            SALSAA_TRY(salsaa_batch_projections(ctx, P, W, r, m, z, sg, run.one, plhs, wlhs, prhs,
                                                wrhs, bok));
            hipLaunchKernelGGL(k_status_ok, dim3(1), dim3(64), 0, st, bok, 1, code(i, 4), status);
            SALSAA_HIP(hipGetLastError());
            const size_t R0 = run.R();
            uint64_t *V2 = nullptr;
            SALSAA_TRY(A.get((R0 + 1) * r * kD, &V2));
            SALSAA_HIP(hipMemcpyAsync(V2, V, R0 * r * b, hipMemcpyDeviceToDevice, st));
            SALSAA_HIP(hipMemcpyAsync(V2 + R0 * r * kD, wrhs, r * b, hipMemcpyDeviceToDevice, st));
            V = V2;
            run.rows.push_back(Row{-1, wlhs, width, nullptr, (size_t)lg2(m / width)});
            SALSAA_TRY(mark(PH_PROJECT));
        }
        // fold + verifier_fold with WIT_REP = 1 (:349-386)
        uint64_t *W4 = J, *Vf = nullptr;
        {
            const size_t R = run.R();
            uint64_t *C = nullptr, *Vc = nullptr, *Vt = nullptr;
            SALSAA_TRY(A.get(r * kD, &C));
            SALSAA_TRY(A.get(R * r * kD, &Vc));
            SALSAA_TRY(A.get(R * r * kD, &Vt));
            SALSAA_TRY(A.get(R * kD, &Vf));
            // challenge_for_fold: random_biased coefficients in {-1, 0, 1}
            SALSAA_TRY(salsaa_sample_bounded(ctx, C, r * kD, 1, rseed, 3));
            SALSAA_TRY(salsaa_parallel_dot_matrix_matrix(ctx, W, r, m, C, 1, W4));
            SALSAA_TRY(salsaa_convert(ctx, W4, m, SALSAA_REP_INCOMPLETE, SALSAA_REP_COEFF));
            SALSAA_HIP(hipMemcpyAsync(Vc, V, R * r * b, hipMemcpyDeviceToDevice, st));
            SALSAA_TRY(salsaa_convert(ctx, Vc, R * r, SALSAA_REP_INCOMPLETE, SALSAA_REP_COEFF));
            hipLaunchKernelGGL(k_transpose_ring, dim3(grid1(R * r * kD, 256)), dim3(256), 0, st,
                               Vc, R, r, Vt);
            SALSAA_HIP(hipGetLastError());
            // rhs' = rhs . C^T: [1][R] = [R][1]
            SALSAA_TRY(salsaa_parallel_dot_matrix_matrix(ctx, Vt, r, R, C, 1, Vf));
            SALSAA_TRY(mark(PH_FOLD));
        }
        // send_cross_terms_before_join + verifier_join, witness.append (:387-410)
        {
            const size_t R = run.R();
            uint64_t *l0 = nullptr, *l1 = nullptr, *Vj = nullptr;
            SALSAA_TRY(A.get(kD, &l0));
            SALSAA_TRY(A.get(kD, &l1));
            SALSAA_TRY(A.get((R + 1) * 2 * kD, &Vj));
            SALSAA_TRY(salsaa_send_cross_terms_before_join(ctx, plhs, wlhs, m, W4, P, l0, l1));
            // rows e < R - 1: [rhs_e, commitment of the projection]; row R - 1
            // (batching): [rhs, l1]; the new row: [l0, projection_rhs]
            SALSAA_HIP(hipMemcpy2DAsync(Vj, 2 * b, Vf, b, b, R, hipMemcpyDeviceToDevice, st));
            SALSAA_HIP(hipMemcpy2DAsync(Vj + kD, 2 * b, pcom, b, b, R - 1, hipMemcpyDeviceToDevice,
                                        st));
            SALSAA_HIP(hipMemcpyAsync(Vj + ((R - 1) * 2 + 1) * kD, l1, b, hipMemcpyDeviceToDevice,
                                      st));
            SALSAA_HIP(hipMemcpyAsync(Vj + R * 2 * kD, l0, b, hipMemcpyDeviceToDevice, st));
            SALSAA_HIP(hipMemcpyAsync(Vj + (R * 2 + 1) * kD, prhs, b, hipMemcpyDeviceToDevice, st));
            run.rows.push_back(Row{-1, plhs, kHeight, nullptr, (size_t)lg2(m / kHeight)});
            V = Vj;
            W = J;
            r = 2;
            SALSAA_TRY(mark(PH_JOIN));
        }
        rounds_done = i + 1;
        if (folding) {
            // the folding scheme closes with decomp_ell(.., 2) + verify_decomp
            uint64_t *dst = (d_W_out && w_cap >= 2 * r * m * kD) ? d_W_out : nullptr;
            SALSAA_TRY(run.decomp(&W, &r, m, 2, &V, code(i, 5), status, fault(i, 5), dst));
            SALSAA_TRY(mark(PH_DECOMP));
            break;
        }
    }
    // the final check: commit(statement, witness) == verifier rhs (:464-472)
    const size_t R = run.R();
    if (!folding) {
        uint64_t *F = nullptr;
        SALSAA_TRY(A.get(R * r * kD, &F));
        if (fault(rounds_done, 6)) hipLaunchKernelGGL(k_fault, dim3(1), dim3(64), 0, st, W, c->q);
        SALSAA_TRY(run.commit(W, r, m, F));
        hipLaunchKernelGGL(k_status_equal, dim3(grid1(R * r * kD, 256)), dim3(256), 0, st, F, V,
                           R * r * kD, code(rounds_done, 6), status);
        SALSAA_HIP(hipGetLastError());
        SALSAA_TRY(mark(PH_FINAL));
    }
    shape[0] = r;
    shape[1] = m;
    shape[2] = R;
    shape[3] = r;
    if (d_W_out && W != d_W_out) {
        SALSAA_CHECK(w_cap >= r * m * kD, "protocol: witness output too small");
        SALSAA_HIP(hipMemcpyAsync(d_W_out, W, r * m * b, hipMemcpyDeviceToDevice, st));
    }
    if (d_rhs_out) {
        SALSAA_CHECK(rhs_cap >= R * r * kD, "protocol: rhs output too small");
        SALSAA_HIP(hipMemcpyAsync(d_rhs_out, V, R * r * b, hipMemcpyDeviceToDevice, st));
    }
    SALSAA_HIP(hipMemcpyAsync(status_out, status, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SALSAA_HIP(hipStreamSynchronize(st));
    if (times) {
        for (int k = 0; k <= PH_FINAL; k++) times[k] = 0.0;
        for (size_t k = 1; k < ev.size(); k++) {
            float ms = 0.f;
            SALSAA_HIP(hipEventElapsedTime(&ms, ev[k - 1], ev[k]));
            times[ev_phase[k]] += ms;
        }
    }
    return SALSAA_OK;
}

}  // namespace

extern "C" int salsaa_protocol(salsaa_ctx *ctx, const uint64_t *d_module, size_t kappa, size_t m,
                               const uint64_t *d_W, size_t r, size_t nrounds,
                               const size_t *parts, int folding, uint64_t seed,
                               uint64_t *d_W_out, size_t w_cap, uint64_t *d_rhs_out,
                               size_t rhs_cap, size_t *shape, int32_t *status, double *times) {
    CTX_D64(ctx, "salsaa_protocol");
    return run_protocol(ctx, d_module, kappa, m, d_W, r, nrounds, parts, folding, seed, d_W_out,
                        w_cap, d_rhs_out, rhs_cap, shape, status, times, nullptr);
}

extern "C" int salsaa_protocol_ex(salsaa_ctx *ctx, const uint64_t *d_module, size_t kappa,
                                  size_t m, const uint64_t *d_W, size_t r, size_t nrounds,
                                  const size_t *parts, int folding, uint64_t seed,
                                  uint64_t *d_W_out, size_t w_cap, uint64_t *d_rhs_out,
                                  size_t rhs_cap, size_t *shape, int32_t *status, double *times,
                                  const salsaa_protocol_opts *opts) {
    CTX_D64(ctx, "salsaa_protocol_ex");
    return run_protocol(ctx, d_module, kappa, m, d_W, r, nrounds, parts, folding, seed, d_W_out,
                        w_cap, d_rhs_out, rhs_cap, shape, status, times, opts);
}
