// common.hpp -- shared definitions for the SALSAA MI355X library: status
// codes, the context, exact fp64 modular arithmetic (q < 2^50) and the seeded
// counter-based sampler used by both host and device code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/salsaa_amd.h"

namespace salsaa {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);

#define SALSAA_HIP(expr)                                                        \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) {                                                 \
            ::salsaa::set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
            return SALSAA_ERR_HIP;                                              \
        }                                                                       \
    } while (0)

#define SALSAA_CHECK(cond, msg)                                                 \
    do {                                                                        \
        if (!(cond)) {                                                          \
            ::salsaa::set_error(msg);                                           \
            return SALSAA_ERR_ARG;                                              \
        }                                                                       \
    } while (0)

#define SALSAA_TRY(expr)                                                        \
    do {                                                                        \
        int s_ = (expr);                                                        \
        if (s_ != SALSAA_OK) return s_;                                         \
    } while (0)

// ---------------------------------------------------------------- constants
constexpr int kD = 64;      // ring degree supported by the device kernels
constexpr int kH = kD / 2;  // slots of the IncompleteNTT representation

// Device table layout (doubles) inside ctx->d_tables.
enum TableOffset : int {
    TAB_FWD32 = 0,    // [32]  psi_64 powers, bit-reversed (HEXL roots, n = 32)
    TAB_INV32 = 32,   // [32]  their inverses; [32+1] pre-scaled by 32^{-1}
    TAB_FWD64 = 64,   // [64]  psi_128 powers, bit-reversed (n = 64)
    TAB_INV64 = 128,  // [64]  inverses; [128+1] pre-scaled by 64^{-1}
    TAB_OMEGA = 192,  // [32]  omega_j = NTT_32(Y)[j] (slot moduli X^2 - omega_j)
    TAB_NINV32 = 224, // 32^{-1}
    TAB_NINV64 = 225, // 64^{-1}
    TAB_FIN = 256,    // [64][64] transposed map full-NTT slots -> IncompleteNTT:
                      // fin[s * 64 + c] = (IncNTT(INTT_64(e_s)))[c]
    TAB_W256 = 256 + 64 * 64,  // [13] 2^(8t) mod q, digit weights of the int8 commit
    // full-NTT slot s from the IncompleteNTT halves (E, O):
    // a(zeta_s) = E[zj_s] + zeta_s O[zj_s], zeta_s = psi_128^(2 brv_6(s) + 1),
    // omega_{zj_s} = zeta_s^2
    TAB_ZETA = 256 + 64 * 64 + 16,  // [64] zeta_s
    TAB_ZJ = TAB_ZETA + 64,         // [64] zj_s (as a double)
    TAB_SIZE = TAB_ZJ + 64,
};

// Tables of a context of degree d > 64 (doubles inside ctx->d_wtab), n = d/2:
// the IncompleteNTT's two length-n transforms, the Full transform of length d
// (present when 2d divides q - 1) and the slot moduli X^2 - omega_j.
struct WideTab {
    int d = 0, n = 0;
    int fwd_h = 0, inv_h = 0;  // [n] psi_d^brv(i); inverses, [0] pre-scaled by n^{-1}
    int fwd_f = 0, inv_f = 0;  // [d] psi_2d^brv(i); inverses, [0] pre-scaled by d^{-1}
    int omega = 0;             // [n] omega_j = NTT_n(Y)[j] = psi_d^(2 brv(j) + 1)
    int ninv_h = 0, ninv_f = 0;
    int size = 0;
    __host__ __device__ explicit WideTab(int deg = 0) {
        d = deg;
        n = deg / 2;
        fwd_h = 0;
        inv_h = n;
        fwd_f = 2 * n;
        inv_f = 2 * n + d;
        omega = 2 * n + 2 * d;
        ninv_h = omega + n;
        ninv_f = ninv_h + 1;
        size = ninv_f + 1;
    }
};

// ---------------------------------------------------------------- context
struct ScratchBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct Context {
    int device = 0;
    int ncu = 256;                  // compute units of the device
    uint64_t q = 0;
    int d = kD;
    double qd = 0.0, qinv = 0.0;
    hipStream_t stream = nullptr;
    double *d_tables = nullptr;     // TAB_SIZE doubles
    std::vector<double> h_tables;   // host copy
    double *d_wtab = nullptr;       // degree-d tables (WideTab(d)) when d > 64
    bool full_ok = true;            // the Full representation exists at degree d (2d | q - 1)
    ScratchBuf scratch[17];         // grown on demand, owned by the context
    bool mfma_commit = true;        // int8 MFMA commit (false: fp64 VALU commit)
    bool mfma_project = true;       // int8 MFMA projection (false: VALU signed sums)
    int ntt_xch = 0;                // IncompleteNTT exchange between register phases:
                                    // 0 through LDS, 1 DPP quad permutes
    ScratchBuf zeros;               // zero-filled, grown on demand
    std::vector<ScratchBuf> slabs;  // the protocol run's buffers (kept for the next run)
    std::recursive_mutex lock;      // held by every entry point on this context
};

// A context on another stream over the same device tables (its scratch and
// stream are its own).
inline void share_context(const Context &from, Context &to) {
    to.device = from.device;
    to.ncu = from.ncu;
    to.q = from.q;
    to.d = from.d;
    to.qd = from.qd;
    to.qinv = from.qinv;
    to.d_tables = from.d_tables;
    to.d_wtab = from.d_wtab;
    to.full_ok = from.full_ok;
    to.mfma_commit = from.mfma_commit;
    to.mfma_project = from.mfma_project;
    to.ntt_xch = from.ntt_xch;
}

struct Key {
    Context *ctx = nullptr;
    size_t kappa = 0, m = 0;  // rows and length this key commits with
    double *d_A = nullptr;    // slot-major full-NTT [64][a_rows][a_m] as exact fp64 (VALU path)
    int8_t *d_AB = nullptr;   // fragment-tiled digit planes (MFMA path, kappa <= kMfmaMaxKappa)
    uint64_t *d_mhat = nullptr;  // generated CRS key: the module in NTT form [kappa][d], no layout
    size_t a_rows = 0, a_m = 0;     // shape of the stored matrix (a view commits with a prefix)
    size_t a_kt = 0, a_ktiles = 0;  // planes: row tiles per K step, K tiles (plane_ktiles(a_m))
    bool owner = true;              // false for a view sharing another key's storage
};

// One job of a batched MFMA commit: at most 16 witness rows with their own
// fragment-tiled planes (rtiles row tiles per K step, ktiles K tiles).
struct CommitJob {
    const int8_t *WB = nullptr;
    double *partial = nullptr;  // [ceil(ktiles / kchunk)][64][kappa][rows]
    int ktiles = 0, rtiles = 0, ntile = 0, rows = 0;
    // the job's key rows and key row tiles per wave (1, 2 or 4), read by the
    // launch that mixes jobs of different kappa
    int kappa = 0, tm = 0;
};

// The MFMA commit keeps all of a key's digit rows in one block (16 row tiles).
constexpr size_t kMfmaMaxKappa = 36;

// Fragment-tiled digit planes (see store_planes in ntt8.hip): K-tiles of 64
// columns for a witness of length m, m rounded up to 128.
inline size_t plane_ktiles(size_t m) { return (m + 127) / 128 * 2; }
// bytes of the planes of a matrix with `rows` value rows
inline size_t plane_bytes(size_t rows, size_t m) {
    return (size_t)64 * ((rows * 7 + 15) / 16) * plane_ktiles(m) * 1024;
}

int scratch_get(Context *ctx, int slot, size_t bytes, void **out);
// A device buffer of at least `bytes` zero bytes owned by the context.
int zero_row(Context *ctx, size_t bytes, const int8_t **out);

// ---------------------------------------------------------------- host arithmetic
inline uint64_t h_mulmod(uint64_t a, uint64_t b, uint64_t q) {
    return (uint64_t)((unsigned __int128)a * b % q);
}
inline uint64_t h_powmod(uint64_t a, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) r = h_mulmod(r, a, q);
        a = h_mulmod(a, a, q);
        e >>= 1;
    }
    return r;
}
inline uint64_t h_invmod(uint64_t a, uint64_t q) { return h_powmod(a, q - 2, q); }

// ---------------------------------------------------------------- sampler
// Counter-based SplitMix64: element `idx` of stream `stream` under `seed`.
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL));
}
__host__ __device__ inline uint64_t draw(uint64_t key, uint64_t idx) {
    return mix64(key + (idx + 1) * 0x9E3779B97F4A7C15ULL);
}
__host__ __device__ inline uint64_t mulhi64(uint64_t a, uint64_t b) {
#ifdef __HIP_DEVICE_COMPILE__
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// ---------------------------------------------------------------- fp64 modular arithmetic
// Residues live in registers as exact integer-valued doubles in [0, q).
// For a, b < q < 2^50: hi = fl(a*b), lo = a*b - hi exactly (fma), the quotient
// estimate rint(hi/q) is within 0.875 of a*b/q, so r = (hi - qt*q) + lo is an
// exact integer in (-q, q) (every intermediate is an integer below 2^52).
struct Mod {
    double q, qinv;
};

__device__ __forceinline__ double u2d(uint64_t x) {
    // exact for x < 2^52: the double 2^52 + x has mantissa x
    return __longlong_as_double((long long)(x | 0x4330000000000000ULL)) - 4503599627370496.0;
}
__device__ __forceinline__ uint64_t d2u(double v) {
    return (uint64_t)__double_as_longlong(v + 4503599627370496.0) & 0x000FFFFFFFFFFFFFULL;
}

__device__ __forceinline__ double mul_lazy(double a, double b, Mod M) {
    double hi = a * b;
    double qt = __builtin_rint(hi * M.qinv);
    double lo = __fma_rn(a, b, -hi);
    return __fma_rn(-qt, M.q, hi) + lo;  // in (-q, q)
}
__device__ __forceinline__ double mulmod(double a, double b, Mod M) {
    double r = mul_lazy(a, b, M);
    return r < 0.0 ? r + M.q : r;
}
__device__ __forceinline__ double addmod(double a, double b, Mod M) {
    double s = a + b;
    return s >= M.q ? s - M.q : s;
}
__device__ __forceinline__ double submod(double a, double b, Mod M) {
    double s = a - b;
    return s < 0.0 ? s + M.q : s;
}
// Reduce an integer-valued |x| < 2^53 to [0, q).
__device__ __forceinline__ double reduce(double x, Mod M) {
    double t = __fma_rn(-__builtin_rint(x * M.qinv), M.q, x);
    return t < 0.0 ? t + M.q : t;
}
// Reduce to (-q/2 - 1, q/2 + 1) without the final correction (lazy).
__device__ __forceinline__ double reduce_lazy(double x, Mod M) {
    return __fma_rn(-__builtin_rint(x * M.qinv), M.q, x);
}
// A double moved across lanes by a DPP control (every lane active; quad_perm
// 0xB1 = lane ^ 1, 0x4E = lane ^ 2, 0x141 half-row mirror, 0x140 row mirror).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}
__device__ __forceinline__ double readlane_d(double v, int lane) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
// The sum mod q of one residue in [0, q) per lane over the whole wave (every
// lane active), canonical and the same in every lane: DPP moves within each
// 16-lane row (lane ^ 1, lane ^ 2, the other half-row, the other 8 lanes),
// then the four row sums by readlane.  A partial sum of 8 residues is below
// 2^53, so the double adds are exact between the two reductions.
__device__ __forceinline__ double wave_sum_res(double x, Mod M) {
    x += dpp_d<0xB1>(x);
    x += dpp_d<0x4E>(x);
    x += dpp_d<0x141>(x);
    x = reduce_lazy(x, M);  // |x| <= q/2 + 1
    x += dpp_d<0x140>(x);   // |x| <= q + 2
    return reduce(readlane_d(x, 0) + readlane_d(x, 16) + readlane_d(x, 32) + readlane_d(x, 48), M);
}
// orders one wave's LDS writes before its later reads (and reads before
// later writes) without waiting for the rest of the block
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Byte i (i < 7) of four 64-bit words y_0 .. y_3 given as lo/hi dwords:
// d[i] = (byte i of y_0) | (byte i of y_1) << 8 | ... , a 4 x 4 byte
// transpose per dword half in two v_perm_b32 stages (byte k of
// perm(a, b, sel) is byte sel_k of the 8-byte value a:b, b the low half).
__device__ __forceinline__ void bytes_transpose4(const uint32_t (&lo)[4], const uint32_t (&hi)[4],
                                                 uint32_t (&d)[7]) {
    const uint32_t a01 = __builtin_amdgcn_perm(lo[1], lo[0], 0x05010400u);
    const uint32_t b01 = __builtin_amdgcn_perm(lo[1], lo[0], 0x07030602u);
    const uint32_t a23 = __builtin_amdgcn_perm(lo[3], lo[2], 0x05010400u);
    const uint32_t b23 = __builtin_amdgcn_perm(lo[3], lo[2], 0x07030602u);
    d[0] = __builtin_amdgcn_perm(a23, a01, 0x05040100u);
    d[1] = __builtin_amdgcn_perm(a23, a01, 0x07060302u);
    d[2] = __builtin_amdgcn_perm(b23, b01, 0x05040100u);
    d[3] = __builtin_amdgcn_perm(b23, b01, 0x07060302u);
    const uint32_t c01 = __builtin_amdgcn_perm(hi[1], hi[0], 0x05010400u);
    const uint32_t e01 = __builtin_amdgcn_perm(hi[1], hi[0], 0x07030602u);
    const uint32_t c23 = __builtin_amdgcn_perm(hi[3], hi[2], 0x05010400u);
    const uint32_t e23 = __builtin_amdgcn_perm(hi[3], hi[2], 0x07030602u);
    d[4] = __builtin_amdgcn_perm(c23, c01, 0x05040100u);
    d[5] = __builtin_amdgcn_perm(c23, c01, 0x07060302u);
    d[6] = __builtin_amdgcn_perm(e23, e01, 0x05040100u);
}

// canonical x in [0, q) -> balanced (-q/2, q/2]; balanced (|x| < q) -> canonical
__device__ __forceinline__ double balance(double x, Mod M) { return x > 0.5 * M.q ? x - M.q : x; }
__device__ __forceinline__ double canon(double x, Mod M) { return x < 0.0 ? x + M.q : x; }

// F_{q^2} = F_q[Y]/(Y^2 - nr), element a + bY
struct F2 {
    double a, b;
};
__device__ __forceinline__ F2 f2_add(F2 x, F2 y, Mod M) { return {addmod(x.a, y.a, M), addmod(x.b, y.b, M)}; }
__device__ __forceinline__ F2 f2_sub(F2 x, F2 y, Mod M) { return {submod(x.a, y.a, M), submod(x.b, y.b, M)}; }
__device__ __forceinline__ F2 f2_mul(F2 x, F2 y, double nr, Mod M) {
    double bd = mulmod(x.b, y.b, M);
    double re = reduce(mul_lazy(x.a, y.a, M) + mul_lazy(nr, bd, M), M);
    double im = reduce(mul_lazy(x.a, y.b, M) + mul_lazy(x.b, y.a, M), M);
    return {re, im};
}

// ---------------------------------------------------------------- launch helpers
inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }
inline int log2_deg(const Context *ctx) { return __builtin_ctz((unsigned)ctx->d); }
// the IncompleteNTT's slot moduli X^2 - omega_j at the context's degree
inline const double *slot_omega(const Context *ctx) {
    return ctx->d == kD ? ctx->d_tables + TAB_OMEGA : ctx->d_wtab + WideTab(ctx->d).omega;
}

// entry points implemented in the kernel files
int launch_ntt(Context *ctx, uint64_t *data, size_t count, int mode, int inverse);
// the largest layer of a power series from its base and length-2 tensors,
// all in NTT form (see k_series_expand)
int launch_series_expand(Context *ctx, const uint64_t *base, size_t m0, const uint64_t *tens,
                         size_t n, uint64_t *out);
// project (project.rs:52-98) before its commit, see project.hip
int launch_project(Context *ctx, const uint64_t *W, size_t r, size_t m, const uint8_t *is_zero,
                   const uint8_t *sign, uint64_t *out);
int launch_project_transposed(Context *ctx, const uint64_t *P, size_t cw, const uint8_t *is_zero,
                              const uint8_t *sign, uint64_t *out);
// acc: scratch of rows * cw * d u64 for split sums (NULL: no split)
int launch_chunk_sums(Context *ctx, const uint64_t *src, size_t rows, size_t m, size_t cw,
                      uint64_t *out, unsigned long long *acc);
int launch_powers(Context *ctx, const uint64_t *c_ntt, uint64_t *out);
int launch_repeat(Context *ctx, const uint64_t *base, size_t m0, size_t n, uint64_t *out);
int launch_sum_equals(Context *ctx, const uint64_t *a, size_t r, const uint64_t *b, int32_t *ok);
int launch_ntt_full_to_slotmajor(Context *ctx, const uint64_t *src, size_t rows, size_t m,
                                 size_t row_lo, size_t stride_lo, size_t stride_hi,
                                 double *dst);
int launch_decomp_to_slotmajor(Context *ctx, const uint64_t *W, size_t r, size_t m, size_t parts,
                               const uint64_t *d_radix, uint64_t *new_W, void *dst, int planes,
                               size_t pstride);
// the power-series CRS rows (crs.rs:43-72) of `rows` module elements (Coefficient),
// length m: mode 0 NTT-form rows [rows][m][d], 1 a key's slot-major doubles, 2 a
// key's digit planes with pstride K tiles
// out[j] = sum_i layer[i] W[j][i] (IncompleteNTT) for r <= 16 rows, in one pass
// over the layer and the rows (parallel_dot_single_series_matrix)
int launch_dot_series(Context *ctx, const uint64_t *layer, size_t m, const uint64_t *W, size_t r,
                      uint64_t *out);
int launch_crs(Context *ctx, const uint64_t *module, size_t rows, size_t m, int mode, void *dst,
               size_t pstride);
int launch_ntt_full_to_planes(Context *ctx, const uint64_t *src, size_t rows, size_t m,
                              size_t row_lo, size_t stride_lo, size_t stride_hi, int8_t *dst,
                              size_t pstride);
// One job of the batched witness -> planes transform: witness rows
// first .. first + rows - 1 (instance order) into planes [64][ktiles][rtiles]
// KiB fragments, rtiles = ceil(7 rows / 16); rows * ktiles / 2 blocks from blk0.
struct PlaneJob {
    int8_t *dst = nullptr;
    int first = 0, rows = 0, ktiles = 0, blk0 = 0;
};
int launch_ntt_planes_batch(Context *ctx, const uint64_t *arena, const unsigned long long *inst_off,
                            const unsigned long long *inst_m, const PlaneJob *jobs, size_t njobs,
                            size_t nblk);
// Ring elements per tile of the fused forward NTT + planes kernel (a
// multiple of 64: whole K tiles of the planes).  128 (one block of 1024
// threads per CU) measured faster than 64 (two blocks of 512).
#ifndef SALSAA_PLANE_TILE
#define SALSAA_PLANE_TILE 128
#endif
constexpr int kPlaneTile = SALSAA_PLANE_TILE;
// Where the digit planes of one kPlaneTile-element tile of a witness arena
// go: the fragment byte of (slot 0, the tile's first K tile, its value row),
// and the byte distances between slots and between consecutive K tiles.
struct PlaneTile {
    unsigned long long base = 0;
    unsigned sstride = 0, kstride = 0;
};
// The short copy of a witness arena, for the norm check's first pass: the
// norm check proves a witness short, and residues x whose balanced forms b
// lie in [-2^14, 2^14) are kept in one of two widths per slice of
// kShortSlice coefficients (one NTT-pass wave's 8 ring elements),
// fmt[slice]: 0 when every b of the slice lies in [-128, 127], t8[i] = b + 128
// (one byte for arena coefficient i); 1 when every b lies in [-2^14, 2^14),
// t[i] = b + 2^14 (two bytes); 2 otherwise (the slice's entries are
// meaningless; the format bytes start at 2).
constexpr uint32_t kShortOff = 1u << 14;
constexpr size_t kShortTile = (size_t)kPlaneTile * kD;
constexpr size_t kShortSlice = 512;
struct ShortCopy {
    uint16_t *t = nullptr;
    uint8_t *t8 = nullptr;
    uint8_t *fmt = nullptr;
};
// t = b + 2^14 (mod 2^64) for canonical x; x is short iff short_bad(t) == 0
__device__ __forceinline__ uint64_t short_t(uint64_t x, uint64_t half, uint64_t off_neg) {
    return x + (x > half ? off_neg : (uint64_t)kShortOff);
}
__device__ __forceinline__ uint32_t short_bad(uint64_t t) {
    return (uint32_t)(t >> 32) | ((uint32_t)t & ~(2 * kShortOff - 1));
}
// whether the t = b + 2^14 of short_t is not byte-short (b outside [-128, 127])
__device__ __forceinline__ uint32_t byte_bad(uint64_t t) {
    return (uint32_t)((t - (kShortOff - 128)) >> 8 != 0);
}
// Forward IncompleteNTT of `count` (a multiple of kPlaneTile) arena elements in
// place, fused with the full-NTT digit planes of every tile (tiles[i] for
// elements kPlaneTile i ..), written under `planes`.  roundtrip: the inverse
// IncompleteNTT follows in the same pass, so the arena ends in Coefficient
// form again (the forward + inverse NTT pair of a prover step); with `sc`
// the pass also writes the arena's short copy.  grid: persistent blocks (0:
// as many as the CUs hold).
int launch_ntt_fwd_planes(Context *ctx, uint64_t *data, size_t count, int8_t *planes,
                          const PlaneTile *tiles, int roundtrip, const ShortCopy *sc = nullptr,
                          int xch = 0, unsigned grid = 0);
int launch_commit_mfma(Context *ctx, const int8_t *AB, size_t kappa, size_t a_kt,
                       size_t a_ktiles, const int8_t *WB, size_t r, size_t m, size_t ktW,
                       uint64_t *out);
// ---- degrees d > 64 (ring_wide.hip)
// representation change src -> dst (may alias) of count degree-d elements
int launch_wide_ntt(Context *ctx, const uint64_t *src, uint64_t *dst, size_t count, int mode,
                    int inverse);
// launch_ntt at the context's degree (in place)
int launch_ntt_any(Context *ctx, uint64_t *data, size_t count, int mode, int inverse);
// the commit's digit planes of rows [rows][m][d] (Coefficient) at degree d:
// d/2 slots, value row 2 row + h holding half h of each IncompleteNTT slot
int launch_wide_planes(Context *ctx, const uint64_t *src, size_t rows, size_t m, int8_t *dst,
                       size_t ktiles);
inline size_t wide_plane_bytes(size_t d, size_t rows, size_t m) {
    return d / 2 * ((2 * rows * 7 + 15) / 16) * plane_ktiles(m) * 1024;
}
// the commit at degree d over the d/2 slot pairs: key planes of 2 kappa and
// witness planes of 2 r value rows (launch_wide_planes); out [kappa][r][d]
// IncompleteNTT
int launch_commit_mfma_pairs(Context *ctx, const int8_t *AB, size_t kappa2, size_t a_kt,
                             size_t a_ktiles, const int8_t *WB, size_t r2, size_t m, size_t ktW,
                             uint64_t *out);
int launch_commit_finalize_pairs(Context *ctx, const double *partial, size_t ksplit,
                                 size_t kappa2, size_t r2, uint64_t *out);
// norm_check_3's rhs [reps][rows][2][d] at degree d > 64 from plain witness
// rows, and the verifier's check of its constant term against the finals
// (clears ok[rep ok_rep + row ok_row] on a mismatch)
int launch_nc3_wide(Context *ctx, const uint64_t *w, size_t rows, size_t L, const uint64_t *cs,
                    size_t reps, double nr, bool fq2, const uint64_t *finals, size_t fe_rep,
                    size_t fe_row, uint64_t *rhs, int32_t *ok, size_t ok_rep, size_t ok_row);
// table[b] = (job, z * 64 + s) for each of nblk logical blocks; every job's
// partial sums are left for launch_commit_finalize_batch
int launch_commit_mfma_batch(Context *ctx, const int8_t *AB, size_t kappa, size_t a_kt,
                             size_t a_ktiles, const CommitJob *jobs, const uint2 *table,
                             size_t nblk, size_t kchunk, size_t max_tiles);
int commit_batch_occupancy(size_t kappa, size_t max_tiles);
// One launch over jobs of several kappa (each job's kappa from its CommitJob,
// job indices in the table relative to `jobs`) at the shape of tm (2 or 4)
// key tiles per wave, which must cover every job's (job.tm <= tm).
int launch_commit_mfma_mixed(Context *ctx, const int8_t *AB, size_t a_kt, size_t a_ktiles,
                             const CommitJob *jobs, const uint2 *table, size_t nblk,
                             size_t kchunk, int tm);
// the commit against the power-series CRS with NTT-form module mhat [kappa][d],
// the key generated in the kernel (WB: the witness planes)
int launch_commit_crs(Context *ctx, const uint64_t *mhat, size_t kappa, const int8_t *WB, size_t r,
                      size_t m, size_t ktW, uint64_t *out);
int launch_commit_crs_batch(Context *ctx, const uint64_t *mhat, size_t kappa, const CommitJob *jobs,
                            const uint2 *table, size_t nblk, size_t kchunk, size_t max_tiles);
int launch_commit_finalize(Context *ctx, const double *partial, size_t ksplit, size_t kappa,
                           size_t r, uint64_t *out);
int launch_radix(Context *ctx, const uint64_t *W, size_t r, size_t m, size_t parts,
                 uint64_t radix, uint64_t *d_radix);
// the radix of each of ninst instances (instance i at W + i r m d) into d_radix[i]
int launch_radix_n(Context *ctx, const uint64_t *W, size_t ninst, size_t r, size_t m,
                   size_t parts, uint64_t radix, uint64_t *d_radix);
int launch_commit_mac(Context *ctx, const double *A_sm, size_t kappa, size_t a_rows, size_t mA,
                      const double *W_sm, size_t r, size_t m, uint64_t *out);
// One finalize job: partial [ks][64][kappa][rows]; (j, k) -> out + j osj + k osk
struct FinJob {
    const double *partial = nullptr;
    uint64_t *out = nullptr;
    long long osj = 0, osk = 0;
    int ks = 0, rows = 0, blk0 = 0;
};
int launch_commit_finalize_batch(Context *ctx, const FinJob *jobs, size_t njobs, size_t nblk,
                                 size_t kappa);
int launch_sample_uniform(Context *ctx, uint64_t *out, size_t n, uint64_t seed, uint64_t stream);
int launch_sample_bounded(Context *ctx, uint64_t *out, size_t n, uint64_t bound, uint64_t seed,
                          uint64_t stream);
int launch_challenges(Context *ctx, uint64_t *cs, size_t reps, size_t nrounds, int fq2,
                      uint64_t seed);
// Transcript outputs of the norm check and their element strides: claims[row],
// msgs[rep][t][row][6], finals/evals[rep][row][2], rhs[rep][row][2][d],
// ok[rep][row] in the plain layout (nc_out_plain); a batch lays each row's
// transcript out contiguously instead.  Null pointers (except ok) are skipped.
struct NcOut {
    uint64_t *claims = nullptr, *msgs = nullptr, *finals = nullptr, *evals = nullptr,
             *rhs = nullptr;
    int32_t *ok = nullptr;
    size_t claim_row = 1, msg_rep = 0, msg_t = 0, msg_row = 6, fe_rep = 0, fe_row = 2,
           rhs_rep = 0, rhs_row = 2 * kD, ok_rep = 0, ok_row = 1;
};
NcOut nc_out_plain(size_t rows, size_t reps, size_t nrounds, uint64_t *claims, uint64_t *msgs,
                   uint64_t *finals, uint64_t *evals, uint64_t *rhs, int32_t *ok);
// Caller-owned norm check scratch (otherwise the context's scratch is used)
struct NcScratch {
    unsigned long long *macc = nullptr;
    double *bufA = nullptr, *bufB = nullptr;
    size_t macc_bytes = 0, bufA_bytes = 0, bufB_bytes = 0;
};
// short_copy: the norm check reads a short copy (its first pass may then keep
// a 64 x 64 Gram matrix per row instead of 16 x 16)
void nc_scratch_need(size_t rows, size_t L, size_t reps, size_t *macc_bytes, size_t *bufA_bytes,
                     size_t *bufB_bytes, bool short_copy = false);
// wrow: per-row offsets (u64) of the rows from w (null: row r at w + r L);
// cs_row: u64 between consecutive rows' challenge blocks (0: shared);
// cpt_first: columns per thread of the first pass (0: automatic);
// sc: the short copy of the arena starting at w (the first pass reads it
// where the slices a sub-block covers are all of one width);
// fault_claim: adds 1 to row 0's claim (norm_check_1) before the verifier
// replay reads it, as a prover sending a wrong claim would (needs L > 1024);
// for tests of the verifier
int launch_norm_check_ex(Context *ctx, const uint64_t *w, const unsigned long long *wrow,
                         size_t rows, size_t L, const uint64_t *cs, size_t cs_row, size_t reps,
                         int fq2, uint64_t nr, const NcOut &o, int cpt_first,
                         const NcScratch *ext, const ShortCopy *sc = nullptr,
                         int fault_claim = 0);
int launch_norm_check(Context *ctx, const uint64_t *w, size_t rows, size_t L,
                      const uint64_t *cs, size_t reps, int fq2, uint64_t nr, uint64_t *claims,
                      uint64_t *msgs, uint64_t *finals, uint64_t *evals, uint64_t *rhs,
                      int32_t *ok);
int launch_decompose(Context *ctx, const uint64_t *W, size_t r, size_t m, size_t parts,
                     uint64_t radix, uint64_t *out, uint64_t *d_radix);
int launch_compose_check(Context *ctx, const uint64_t *in, size_t rows, size_t n, size_t parts,
                         const uint64_t *d_radix, const uint64_t *expect, int rows_mode,
                         uint32_t *d_mismatch);
int launch_compose(Context *ctx, const uint64_t *in, size_t rows, size_t n, size_t parts,
                   const uint64_t *d_radix, uint64_t *out);
int launch_ring_mul(Context *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out,
                    size_t count);
int launch_fold(Context *ctx, const uint64_t *W, size_t r_in, size_t m, const uint64_t *C,
                size_t r_out, uint64_t *out);

}  // namespace salsaa
