// transcript.hip -- the proof transcript: its byte format (the reference only
// logs its values, protocol.rs:474-482), written on the device from the
// outputs of a norm check, a prover step or a protocol run, and the norm
// check with Fiat-Shamir challenges.
//
// Fiat-Shamir.  The reference's verifier draws every norm-check challenge at
// random, independently of the round's message (fast_norm_check.rs:375-385).
// Here round t's challenge comes from a SHA-256 chain instead:
//   h_0     = SHA-256("SALSAA-NC-FS-v1" || seed[32] || q || L || rows || rep
//                     || claims[rows])
//   h_{t+1} = SHA-256(h_t || the round's messages of every row, in row order,
//                     each (f0.a, f0.b, f1.a, f1.b, f2.a, f2.b))
//   c_t     = (mulhi(u64(h_{t+1}[0..8]), q), mulhi(u64(h_{t+1}[8..16]), q))
// (integers little-endian: q and L as u64, rows and rep as u32, values as
// canonical u64; the second coordinate is 0 over F_q).  All rows share the
// round's challenge, as the reference's rows do (fast_norm_check.rs:352-399).
// The seed binds the statement (for example a digest of the commitment).
// Each round is then one fused pass over the folded witness: fold by c_t and
// add the next round's f0, f1, f2 (all three summed, so the verifier's round
// identity 2 f0 + 2 f1 + f2 = claim is a real check).
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "handles.hpp"
#include "ntt_device.hpp"

using namespace salsaa;

namespace {

// ---------------------------------------------------------------- SHA-256
// FIPS 180-4, run by one thread.
__constant__ uint32_t kShaK[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

__device__ __forceinline__ uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

struct Sha {
    uint32_t h[8];
    uint8_t buf[64];
    uint32_t n;       // bytes in buf
    uint64_t total;   // bytes absorbed
};

__device__ void sha_block(uint32_t *h, const uint8_t *p) {
    uint32_t w[64];
    for (int i = 0; i < 16; i++)
        w[i] = ((uint32_t)p[4 * i] << 24) | ((uint32_t)p[4 * i + 1] << 16) |
               ((uint32_t)p[4 * i + 2] << 8) | (uint32_t)p[4 * i + 3];
    for (int i = 16; i < 64; i++) {
        const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
        const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], k = h[7];
    for (int i = 0; i < 64; i++) {
        const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
        const uint32_t ch = (e & f) ^ (~e & g);
        const uint32_t t1 = k + S1 + ch + kShaK[i] + w[i];
        const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
        const uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
        const uint32_t t2 = S0 + mj;
        k = g;
        g = f;
        f = e;
        e = d + t1;
        d = c;
        c = b;
        b = a;
        a = t1 + t2;
    }
    h[0] += a, h[1] += b, h[2] += c, h[3] += d, h[4] += e, h[5] += f, h[6] += g, h[7] += k;
}

__device__ void sha_init(Sha &s) {
    const uint32_t iv[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                            0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    for (int i = 0; i < 8; i++) s.h[i] = iv[i];
    s.n = 0;
    s.total = 0;
}

__device__ void sha_update(Sha &s, const uint8_t *p, size_t len) {
    s.total += len;
    for (size_t i = 0; i < len; i++) {
        s.buf[s.n++] = p[i];
        if (s.n == 64) {
            sha_block(s.h, s.buf);
            s.n = 0;
        }
    }
}

__device__ void sha_final(Sha &s, uint8_t *out) {
    const uint64_t bits = s.total * 8;
    const uint8_t one = 0x80, zero = 0;
    sha_update(s, &one, 1);
    while (s.n != 56) sha_update(s, &zero, 1);
    uint8_t len[8];
    for (int i = 0; i < 8; i++) len[i] = (uint8_t)(bits >> (56 - 8 * i));
    sha_update(s, len, 8);
    for (int i = 0; i < 8; i++)
        for (int b = 0; b < 4; b++) out[4 * i + b] = (uint8_t)(s.h[i] >> (24 - 8 * b));
}

__device__ __forceinline__ void put64(Sha &s, uint64_t v) {
    uint8_t b[8];
    for (int i = 0; i < 8; i++) b[i] = (uint8_t)(v >> (8 * i));
    sha_update(s, b, 8);
}
__device__ __forceinline__ void put32(Sha &s, uint32_t v) {
    uint8_t b[4];
    for (int i = 0; i < 4; i++) b[i] = (uint8_t)(v >> (8 * i));
    sha_update(s, b, 4);
}
__device__ __forceinline__ uint64_t le64(const uint8_t *p) {
    uint64_t v = 0;
    for (int i = 7; i >= 0; i--) v = (v << 8) | p[i];
    return v;
}

struct Seed {
    uint8_t b[32];
};

// ---------------------------------------------------------------- F_{q^2}
__device__ __forceinline__ F2 f2_mulc(F2 x, F2 y, double nr, Mod M) { return f2_mul(x, y, nr, M); }
__device__ __forceinline__ F2 f2_sq(F2 x, double nr, Mod M) { return f2_mul(x, x, nr, M); }

// The accumulators of one (rep, round, row): f0, f1, f2 as (a, b) pairs, sums
// of per-block partials each reduced mod q (at most 2^13 blocks per row).
__device__ __forceinline__ size_t acc_idx(int rep, int t, int row, int nrounds, int rows) {
    return (((size_t)rep * nrounds + t) * rows + row) * 6;
}

constexpr int FS_THREADS = 256;
constexpr int FS_MAX_BLOCKS = 4096;

// sums vals[0..K) over the block into red[0..K); returns after a barrier
template <int K>
__device__ void block_sum_mod(double (&vals)[K], double *red, Mod M) {
    __shared__ double part[K][FS_THREADS / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double v = vals[k];
        for (int off = 32; off > 0; off >>= 1) v = reduce(v + __shfl_xor(v, off), M);
        if (lane == 0) part[k][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0.0;
        for (int i = 0; i < FS_THREADS / 64; i++) s += part[threadIdx.x][i];
        red[threadIdx.x] = reduce(s, M);
    }
    __syncthreads();
}

// norm_check_1 and round 0 over the F_q rows: claim, f0, f1, f2
// (fast_norm_check.rs:282-320, 135-199), shared by the repetitions (slot rep 0).
__global__ __launch_bounds__(FS_THREADS) void k_fs_first(const uint64_t *__restrict__ w, size_t L,
                                                         int rows, int nrounds,
                                                         unsigned long long *__restrict__ acc,
                                                         unsigned long long *__restrict__ cacc,
                                                         Mod M) {
    __shared__ double red[4];
    const int row = blockIdx.y;
    const size_t h = L / 2;
    const uint64_t *wr = w + (size_t)row * L;
    double s[4] = {0.0, 0.0, 0.0, 0.0};  // claim, f0, f1, f2
    for (size_t x = blockIdx.x * (size_t)FS_THREADS + threadIdx.x; x < h;
         x += (size_t)gridDim.x * FS_THREADS) {
        const double a = u2d(wr[x]), b = u2d(wr[x + h]);
        const double d = submod(b, a, M);
        s[0] = reduce(s[0] + mul_lazy(a, a, M) + mul_lazy(b, b, M), M);
        s[1] = reduce(s[1] + mul_lazy(a, a, M), M);
        s[2] = reduce(s[2] + mul_lazy(a, d, M), M);
        s[3] = reduce(s[3] + mul_lazy(d, d, M), M);
    }
    block_sum_mod<4>(s, red, M);
    if (threadIdx.x == 0) {
        atomicAdd(&cacc[row], (unsigned long long)d2u(red[0]));
        const size_t ix = acc_idx(0, 0, row, nrounds, rows);
        atomicAdd(&acc[ix + 0], (unsigned long long)d2u(red[1]));
        atomicAdd(&acc[ix + 2], (unsigned long long)d2u(red[2]));
        atomicAdd(&acc[ix + 4], (unsigned long long)d2u(red[3]));
    }
}

// Round t of one repetition (a block per repetition, thread 0): the messages
// of every row, the verifier's round identity and claim update
// (fast_norm_check.rs:352-399), and the challenge from the hash chain.
__global__ void k_fs_msg(const unsigned long long *__restrict__ acc,
                         const unsigned long long *__restrict__ cacc, int rows, int nrounds, int t,
                         uint64_t L, int fq2, double nr, Seed seed, uint32_t *__restrict__ hstate,
                         double *__restrict__ clm, int32_t *__restrict__ okf,
                         uint64_t *__restrict__ cs_out, uint64_t *__restrict__ msgs,
                         uint64_t *__restrict__ claims, uint64_t *__restrict__ finals,
                         uint64_t q, Mod M) {
    if (threadIdx.x != 0) return;
    const int rep = blockIdx.x;
    uint32_t *hs = hstate + rep * 8;
    uint8_t hb[32];
    if (t == 0) {
        Sha s;
        sha_init(s);
        const char dom[] = "SALSAA-NC-FS-v1";
        sha_update(s, (const uint8_t *)dom, sizeof dom - 1);
        sha_update(s, seed.b, 32);
        put64(s, q);
        put64(s, L);
        put32(s, (uint32_t)rows);
        put32(s, (uint32_t)rep);
        for (int r = 0; r < rows; r++) {
            const uint64_t c = cacc[r] % q;
            if (rep == 0 && claims) claims[r] = c;
            put64(s, c);
            clm[((size_t)rep * rows + r) * 2] = (double)c;
            clm[((size_t)rep * rows + r) * 2 + 1] = 0.0;
            okf[(size_t)rep * rows + r] = 1;
        }
        sha_final(s, hb);
    } else {
        for (int i = 0; i < 8; i++)
            for (int b = 0; b < 4; b++) hb[4 * i + b] = (uint8_t)(hs[i] >> (24 - 8 * b));
    }
    Sha s;
    sha_init(s);
    sha_update(s, hb, 32);
    uint64_t *mo = msgs + (((size_t)rep * nrounds + t) * rows) * 6;
    for (int r = 0; r < rows; r++) {
        const size_t ix = acc_idx(t == 0 ? 0 : rep, t, r, nrounds, rows);
        uint64_t v[6];
        for (int k = 0; k < 6; k++) v[k] = acc[ix + k] % q;
        for (int k = 0; k < 6; k++) {
            mo[r * 6 + k] = v[k];
            put64(s, v[k]);
        }
        // the round identity against this row's claim
        double *cl = clm + ((size_t)rep * rows + r) * 2;
        const F2 f0 = {(double)v[0], (double)v[1]}, f1 = {(double)v[2], (double)v[3]},
                 f2 = {(double)v[4], (double)v[5]};
        const F2 lhs = f2_add(f2_add(f2_add(f0, f0, M), f2_add(f1, f1, M), M), f2, M);
        if (lhs.a != cl[0] || lhs.b != cl[1]) okf[(size_t)rep * rows + r] = 0;
    }
    uint8_t nh[32];
    sha_final(s, nh);
    for (int i = 0; i < 8; i++)
        hs[i] = ((uint32_t)nh[4 * i] << 24) | ((uint32_t)nh[4 * i + 1] << 16) |
                ((uint32_t)nh[4 * i + 2] << 8) | (uint32_t)nh[4 * i + 3];
    const uint64_t ca = mulhi64(le64(nh), q), cb = fq2 ? mulhi64(le64(nh + 8), q) : 0;
    cs_out[((size_t)rep * nrounds + t) * 2] = ca;
    cs_out[((size_t)rep * nrounds + t) * 2 + 1] = cb;
    const F2 c = {(double)ca, (double)cb}, c2 = f2_mul(c, c, nr, M);
    for (int r = 0; r < rows; r++) {
        const uint64_t *v = mo + r * 6;
        const F2 f0 = {(double)v[0], (double)v[1]}, f1 = {(double)v[2], (double)v[3]},
                 f2 = {(double)v[4], (double)v[5]};
        const F2 nc = f2_add(f2_add(f0, f2_mul(f2_add(f1, f1, M), c, nr, M), M), f2_mul(f2, c2, nr, M),
                             M);
        double *cl = clm + ((size_t)rep * rows + r) * 2;
        cl[0] = nc.a;
        cl[1] = nc.b;
        if (t == nrounds - 1 && finals) {
            finals[((size_t)rep * rows + r) * 2] = d2u(nc.a);
            finals[((size_t)rep * rows + r) * 2 + 1] = d2u(nc.b);
        }
    }
}

// Round t's fold w' = w0 + c_t (w1 - w0) of every (row, rep) over a witness of
// length len (FIRST: the F_q rows; otherwise F_{q^2} pairs [rep][row][len]),
// fused with round t + 1's f0, f1, f2 over w' (thread x folds x and x + len/4).
template <bool FIRST>
__global__ __launch_bounds__(FS_THREADS) void k_fs_fold(
    const uint64_t *__restrict__ w, const double *__restrict__ in, size_t len, int rows,
    int nrounds, int t, const uint64_t *__restrict__ cs, double nr, double *__restrict__ out,
    unsigned long long *__restrict__ acc, Mod M) {
    __shared__ double red[6];
    const int row = blockIdx.y, rep = blockIdx.z;
    const size_t h = len / 2, qt = len / 4;
    const F2 c = {u2d(cs[((size_t)rep * nrounds + t) * 2]),
                  u2d(cs[((size_t)rep * nrounds + t) * 2 + 1])};
    const uint64_t *wr = w + (size_t)row * len;
    const double *ir = in + (((size_t)rep * rows + row) * len) * 2;
    double *orow = out + (((size_t)rep * rows + row) * h) * 2;
    auto ld = [&](size_t x) -> F2 {
        if (FIRST) return {u2d(wr[x]), 0.0};
        return {ir[2 * x], ir[2 * x + 1]};
    };
    auto fold = [&](size_t x) -> F2 {
        const F2 a = ld(x), b = ld(x + h);
        return f2_add(a, f2_mul(c, f2_sub(b, a, M), nr, M), M);
    };
    if (h == 1) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            const F2 v = fold(0);
            orow[0] = v.a;
            orow[1] = v.b;
        }
        return;
    }
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (size_t x = blockIdx.x * (size_t)FS_THREADS + threadIdx.x; x < qt;
         x += (size_t)gridDim.x * FS_THREADS) {
        const F2 a = fold(x), b = fold(x + qt);
        orow[2 * x] = a.a, orow[2 * x + 1] = a.b;
        orow[2 * (x + qt)] = b.a, orow[2 * (x + qt) + 1] = b.b;
        const F2 d = f2_sub(b, a, M);
        const F2 p0 = f2_sq(a, nr, M), p1 = f2_mulc(a, d, nr, M), p2 = f2_sq(d, nr, M);
        s[0] = addmod(s[0], p0.a, M), s[1] = addmod(s[1], p0.b, M);
        s[2] = addmod(s[2], p1.a, M), s[3] = addmod(s[3], p1.b, M);
        s[4] = addmod(s[4], p2.a, M), s[5] = addmod(s[5], p2.b, M);
    }
    block_sum_mod<6>(s, red, M);
    if (threadIdx.x < 6)
        atomicAdd(&acc[acc_idx(rep, t + 1, row, nrounds, rows) + threadIdx.x],
                  (unsigned long long)d2u(red[threadIdx.x]));
}

// The end of each (row, rep): the folded value, the verifier's last check and
// norm_check_3 (fast_norm_check.rs:401-462): rhs = conj(tensor(last log2 d
// challenges)) * V with V the witness folded over the ring-index rounds.
__global__ void k_fs_final(const double *__restrict__ last, const uint64_t *__restrict__ w0,
                           const double *__restrict__ vcap, int rows, int nrounds,
                           const uint64_t *__restrict__ cs, double nr, const double *__restrict__ clm,
                           const int32_t *__restrict__ okf, uint64_t *__restrict__ evals,
                           uint64_t *__restrict__ rhs, int32_t *__restrict__ ok,
                           const double *__restrict__ tab, Mod M) {
    __shared__ F2 V[kD], base[kD];
    __shared__ double rc[2][kD];
    const int row = blockIdx.x, rep = blockIdx.y, tid = threadIdx.x;
    const size_t ri = (size_t)rep * rows + row;
    const F2 cl = {clm[ri * 2], clm[ri * 2 + 1]};
    if (tid < kD) {
        V[tid] = w0 ? F2{u2d(w0[(size_t)row * kD + tid]), 0.0}
                    : F2{vcap[(ri * kD + tid) * 2], vcap[(ri * kD + tid) * 2 + 1]};
        F2 tk = {1.0, 0.0};
        for (int b = 0; b < 6; b++) {
            const size_t ix = ((size_t)rep * nrounds + nrounds - 6 + b) * 2;
            const F2 cb = {u2d(cs[ix]), u2d(cs[ix + 1])};
            const int bit = (tid >> (5 - b)) & 1;
            tk = f2_mul(tk, bit ? cb : f2_sub(F2{1.0, 0.0}, cb, M), nr, M);
        }
        base[tid == 0 ? 0 : kD - tid] = tid == 0 ? tk : f2_sub(F2{0.0, 0.0}, tk, M);
    }
    __syncthreads();
    if (tid < kD) {
        F2 acc = {0.0, 0.0};
        for (int i = 0; i < kD; i++) {
            const int j = tid - i;
            const F2 p = f2_mul(base[i], V[j < 0 ? j + kD : j], nr, M);
            acc = j < 0 ? f2_sub(acc, p, M) : f2_add(acc, p, M);
        }
        rc[0][tid] = acc.a;
        rc[1][tid] = acc.b;
    }
    __syncthreads();
    if (tid < 4 && rhs) {
        const int comp = tid >> 1, p = tid & 1;
        double v[32];
        for (int i = 0; i < 32; i++) v[i] = rc[comp][2 * i + p];
        ntt_fwd_regs<32>(v, tab + TAB_FWD32, M);
        uint64_t *dst = rhs + (ri * 2 + comp) * kD + 32 * p;
        for (int i = 0; i < 32; i++) dst[i] = d2u(v[i]);
    }
    if (tid == 0) {
        const F2 ev = {last[ri * 2], last[ri * 2 + 1]};
        evals[ri * 2] = d2u(ev.a);
        evals[ri * 2 + 1] = d2u(ev.b);
        const F2 ev2 = f2_mul(ev, ev, nr, M);
        const F2 ct = {rc[0][0], rc[1][0]}, ct2 = f2_mul(ct, ct, nr, M);
        const bool good = okf[ri] && ev2.a == cl.a && ev2.b == cl.b && ct2.a == cl.a &&
                          ct2.b == cl.b;
        ok[ri] = good ? 1 : 0;
    }
}

__global__ void k_zero_acc(unsigned long long *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

inline unsigned grid_for(size_t work) {
    size_t g = (work + FS_THREADS - 1) / FS_THREADS;
    return (unsigned)(g < 1 ? 1 : (g > FS_MAX_BLOCKS ? FS_MAX_BLOCKS : g));
}

int norm_check_fs(Context *c, const uint64_t *w, size_t rows, size_t L, size_t reps, int fq2,
                  uint64_t nr, const uint8_t *seed, uint64_t *cs, uint64_t *claims,
                  uint64_t *msgs, uint64_t *finals, uint64_t *evals, uint64_t *rhs, int32_t *ok) {
    SALSAA_CHECK(w && cs && msgs && finals && evals && ok, "norm_check_fs: null argument");
    SALSAA_CHECK(rows >= 1 && rows < 65536 && reps >= 1 && reps < 65536,
                 "norm_check_fs: rows and reps must be in 1..65535");
    SALSAA_CHECK(L >= (size_t)kD && (L & (L - 1)) == 0 && __builtin_ctzll(L) <= 40,
                 "norm_check_fs: length must be a power of two >= 64");
    SALSAA_CHECK(nr < c->q, "non-residue must be reduced mod q");
    const Mod M{c->qd, c->qinv};
    const int n = __builtin_ctzll(L);
    const size_t nacc = reps * (size_t)n * rows * 6;
    // scratch: accumulators, claim sums, hash states, claim chains, flags,
    // the captured length-d vectors, two folded witnesses
    const size_t h = L / 2;
    const size_t b_acc = (nacc + rows) * 8, b_h = reps * 8 * 4, b_cl = reps * rows * 2 * 8,
                 b_ok = reps * rows * 4, b_v = reps * rows * kD * 16,
                 b_buf = reps * rows * h * 16;
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    char *base = nullptr, *bufA = nullptr, *bufB = nullptr;
    SALSAA_TRY(scratch_get(c, 0, up(b_acc) + up(b_h) + up(b_cl) + up(b_ok) + up(b_v),
                           (void **)&base));
    SALSAA_TRY(scratch_get(c, 1, up(b_buf), (void **)&bufA));
    SALSAA_TRY(scratch_get(c, 2, up(b_buf / 2 + 16), (void **)&bufB));
    unsigned long long *acc = (unsigned long long *)base, *cacc = acc + nacc;
    uint32_t *hs = (uint32_t *)(base + up(b_acc));
    double *clm = (double *)(base + up(b_acc) + up(b_h));
    int32_t *okf = (int32_t *)(base + up(b_acc) + up(b_h) + up(b_cl));
    double *vcap = (double *)(base + up(b_acc) + up(b_h) + up(b_cl) + up(b_ok));
    Seed sd;
    memset(sd.b, 0, 32);
    if (seed) memcpy(sd.b, seed, 32);
    const hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_zero_acc, dim3(grid_for(nacc + rows)), dim3(FS_THREADS), 0, st, acc,
                       nacc + rows);
    hipLaunchKernelGGL(k_fs_first, dim3(grid_for(h), (unsigned)rows), dim3(FS_THREADS), 0, st, w, L,
                       (int)rows, n, acc, cacc, M);
    SALSAA_HIP(hipGetLastError());
    const double *src = nullptr;
    double *dst = (double *)bufA, *other = (double *)bufB;
    size_t len = L;
    const double *last = nullptr;
    for (int t = 0; t < n; t++) {
        hipLaunchKernelGGL(k_fs_msg, dim3((unsigned)reps), dim3(64), 0, st, acc, cacc, (int)rows, n,
                           t, (uint64_t)L, fq2, (double)nr, sd, hs, clm, okf, cs, msgs, claims,
                           finals, c->q, M);
        const size_t ho = len / 2;
        const dim3 g(grid_for(ho >= 2 ? ho / 2 : 1), (unsigned)rows, (unsigned)reps);
        // the fold of the ring-index rounds ending at length d is V
        double *o = (ho == (size_t)kD) ? vcap : dst;
        if (t == 0)
            hipLaunchKernelGGL(k_fs_fold<true>, g, dim3(FS_THREADS), 0, st, w, nullptr, len, (int)rows,
                               n, t, cs, (double)nr, o, acc, M);
        else
            hipLaunchKernelGGL(k_fs_fold<false>, g, dim3(FS_THREADS), 0, st, nullptr, src, len,
                               (int)rows, n, t, cs, (double)nr, o, acc, M);
        SALSAA_HIP(hipGetLastError());
        src = o;
        last = o;
        if (o == dst) std::swap(dst, other);
        len = ho;
    }
    hipLaunchKernelGGL(k_fs_final, dim3((unsigned)rows, (unsigned)reps), dim3(kD), 0, st, last,
                       L == (size_t)kD ? w : nullptr, vcap, (int)rows, n, cs, (double)nr, clm, okf,
                       evals, rhs, ok, c->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

// ---------------------------------------------------------------- wire format
// A transcript is a 96-byte header followed by sections of canonical u64
// values, little-endian:
//   0  magic "SLSATRN1" (8 bytes)
//   8  u32 kind (1 norm check, 2 prover step, 3 protocol)   12 u32 flags
//  16  u64 q   24 u64 nr   32 u64 L (norm check length) or m (protocol)
//  40  u32 rows   44 u32 reps   48 u32 nrounds   52 u32 d
//  56  u32 kappa (kinds 2, 3)   60 u32 zero
//  64  32 bytes: the Fiat-Shamir seed (zero without it; kind 3: the run's
//      u64 seed in the first 8 bytes)
// flags: bit 0 Fiat-Shamir challenges, bit 1 F_{q^2}, bit 2 norm_check_3 rhs
// present, bit 3 (kind 3) the folding scheme; kind 3 keeps the status in
// the upper 16 bits as a signed value.  Sections:
//   kind 1: claims[rows], msgs[reps][nrounds][rows][3][2],
//           challenges[reps][nrounds][2], finals[reps][rows][2],
//           evals[reps][rows][2], then rhs[reps][rows][2][d] if flag bit 2
//   kind 2: commitment[kappa][rows][d], then the kind 1 sections
//   kind 3 (rows = final witness rows r, reps = protocol rounds, nrounds =
//           statement rows R, L = final witness length m): the verifier rhs
//           [R][r][d] (IncompleteNTT), then the final witness [r][m][d]
struct Header {
    uint64_t w[12];
};

__global__ void k_write_header(Header hd, uint8_t *out) {
    const int i = threadIdx.x;
    if (i < 12) reinterpret_cast<uint64_t *>(out)[i] = hd.w[i];
}

Header make_header(uint32_t kind, uint32_t flags, uint64_t q, uint64_t nr, uint64_t L,
                   uint32_t rows, uint32_t reps, uint32_t nrounds, uint32_t kappa, uint32_t extra,
                   const uint8_t *seed) {
    Header h;
    memset(&h, 0, sizeof h);
    uint8_t *p = reinterpret_cast<uint8_t *>(h.w);
    memcpy(p, "SLSATRN1", 8);
    auto p32 = [&](int off, uint32_t v) { memcpy(p + off, &v, 4); };
    auto p64 = [&](int off, uint64_t v) { memcpy(p + off, &v, 8); };
    p32(8, kind);
    p32(12, flags);
    p64(16, q);
    p64(24, nr);
    p64(32, L);
    p32(40, rows);
    p32(44, reps);
    p32(48, nrounds);
    p32(52, (uint32_t)kD);
    p32(56, kappa);
    p32(60, extra);
    if (seed) memcpy(p + 64, seed, 32);
    return h;
}

struct Section {
    const void *src;
    size_t bytes;
};

int write_sections(Context *c, const Header &hd, const std::vector<Section> &secs, uint8_t *out,
                   size_t cap, size_t *len) {
    size_t total = sizeof(Header);
    for (const Section &s : secs) total += s.bytes;
    if (len) *len = total;
    if (!out) return SALSAA_OK;
    SALSAA_CHECK(cap >= total, "transcript: output buffer too small");
    for (const Section &s : secs) SALSAA_CHECK(s.src != nullptr, "transcript: null section");
    hipLaunchKernelGGL(k_write_header, dim3(1), dim3(64), 0, c->stream, hd, out);
    SALSAA_HIP(hipGetLastError());
    size_t off = sizeof(Header);
    for (const Section &s : secs) {
        SALSAA_HIP(hipMemcpyAsync(out + off, s.src, s.bytes, hipMemcpyDeviceToDevice, c->stream));
        off += s.bytes;
    }
    return SALSAA_OK;
}

void nc_sections(std::vector<Section> &v, size_t rows, size_t reps, size_t n, const uint64_t *claims,
                 const uint64_t *msgs, const uint64_t *cs, const uint64_t *finals,
                 const uint64_t *evals, const uint64_t *rhs) {
    v.push_back({claims, rows * 8});
    v.push_back({msgs, reps * n * rows * 6 * 8});
    v.push_back({cs, reps * n * 2 * 8});
    v.push_back({finals, reps * rows * 2 * 8});
    v.push_back({evals, reps * rows * 2 * 8});
    if (rhs) v.push_back({rhs, reps * rows * 2 * kD * 8});
}

}  // namespace

extern "C" {

int salsaa_norm_check_fs(salsaa_ctx *ctx, const uint64_t *d_w, size_t rows, size_t L, size_t reps,
                         int fq2, uint64_t nr, const uint8_t *seed, uint64_t *d_cs,
                         uint64_t *d_claims, uint64_t *d_msgs, uint64_t *d_finals,
                         uint64_t *d_evals, uint64_t *d_rhs, int32_t *d_ok) {
    CTX_D64(ctx, "salsaa_norm_check_fs");
    return norm_check_fs(&ctx->c, d_w, rows, L, reps, fq2, nr, seed, d_cs, d_claims, d_msgs,
                         d_finals, d_evals, d_rhs, d_ok);
}

int salsaa_transcript_norm_check(salsaa_ctx *ctx, size_t rows, size_t L, size_t reps, int fq2,
                                 uint64_t nr, int fs, const uint8_t *seed,
                                 const uint64_t *d_claims, const uint64_t *d_msgs,
                                 const uint64_t *d_cs, const uint64_t *d_finals,
                                 const uint64_t *d_evals, const uint64_t *d_rhs, uint8_t *d_out,
                                 size_t cap, size_t *len) {
    CTX_D64(ctx, "salsaa_transcript_norm_check");
    SALSAA_CHECK(L >= 2 && (L & (L - 1)) == 0, "transcript: length must be a power of two");
    const size_t n = (size_t)__builtin_ctzll(L);
    const uint32_t flags = (fs ? 1u : 0u) | (fq2 ? 2u : 0u) | (d_rhs ? 4u : 0u);
    const Header hd = make_header(1, flags, ctx->c.q, fq2 ? nr : 0, L, (uint32_t)rows,
                                  (uint32_t)reps, (uint32_t)n, 0, 0, fs ? seed : nullptr);
    std::vector<Section> v;
    nc_sections(v, rows, reps, n, d_claims, d_msgs, d_cs, d_finals, d_evals, d_rhs);
    return write_sections(&ctx->c, hd, v, d_out, cap, len);
}

int salsaa_transcript_prover_step(salsaa_ctx *ctx, size_t kappa, const uint64_t *d_commitment,
                                  size_t rows, size_t L, size_t reps, int fq2, uint64_t nr,
                                  int fs, const uint8_t *seed, const uint64_t *d_claims,
                                  const uint64_t *d_msgs, const uint64_t *d_cs,
                                  const uint64_t *d_finals, const uint64_t *d_evals,
                                  const uint64_t *d_rhs, uint8_t *d_out, size_t cap,
                                  size_t *len) {
    CTX_D64(ctx, "salsaa_transcript_prover_step");
    SALSAA_CHECK(L >= 2 && (L & (L - 1)) == 0, "transcript: length must be a power of two");
    const size_t n = (size_t)__builtin_ctzll(L);
    const uint32_t flags = (fs ? 1u : 0u) | (fq2 ? 2u : 0u) | (d_rhs ? 4u : 0u);
    const Header hd = make_header(2, flags, ctx->c.q, fq2 ? nr : 0, L, (uint32_t)rows,
                                  (uint32_t)reps, (uint32_t)n, (uint32_t)kappa, 0,
                                  fs ? seed : nullptr);
    std::vector<Section> v;
    v.push_back({d_commitment, kappa * rows * kD * 8});
    nc_sections(v, rows, reps, n, d_claims, d_msgs, d_cs, d_finals, d_evals, d_rhs);
    return write_sections(&ctx->c, hd, v, d_out, cap, len);
}

int salsaa_transcript_protocol(salsaa_ctx *ctx, size_t kappa, size_t nrounds, int folding,
                               uint64_t seed, int32_t status, const size_t *shape,
                               const uint64_t *d_rhs, const uint64_t *d_W, uint8_t *d_out,
                               size_t cap, size_t *len) {
    CTX_D64(ctx, "salsaa_transcript_protocol");
    SALSAA_CHECK(shape != nullptr, "transcript: null shape");
    const size_t r = shape[0], m = shape[1], R = shape[2];
    uint8_t sd[32] = {0};
    memcpy(sd, &seed, 8);
    const uint32_t flags = (folding ? 8u : 0u) | ((uint32_t)(uint16_t)(int16_t)status << 16);
    const Header hd = make_header(3, flags, ctx->c.q, 0, m, (uint32_t)r, (uint32_t)nrounds,
                                  (uint32_t)R, (uint32_t)kappa, 0, sd);
    std::vector<Section> v;
    v.push_back({d_rhs, R * r * kD * 8});
    v.push_back({d_W, r * m * kD * 8});
    return write_sections(&ctx->c, hd, v, d_out, cap, len);
}

}  // extern "C"
