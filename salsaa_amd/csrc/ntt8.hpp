// ntt8.hpp -- negacyclic NTTs with 8 threads per ring element (d = 64).
//
// Each thread holds 8 of an element's values.  The first three butterfly
// stages pair values in the same residue class (positions t + stride * k),
// so they run in registers; one exchange through LDS regroups the values as
// 8 consecutive positions per thread (8t + x) for the remaining stages.  With
// ~40 VGPRs per thread and 73.7 KiB of LDS per 128-element block a CU keeps
// 32 waves in flight, which is what an HBM-bound transform needs.
//
// Conventions (HEXL): tw[m + i] = psi^{brv(m + i)}, Cooley-Tukey forward,
// Gentleman-Sande inverse with inverse twiddles, the inverse's last stage
// folding n^{-1} (twi[0] = twi[1] * n^{-1}).
#pragma once

#include "common.hpp"

namespace salsaa {

// raw bit moves between doubles and LDS words (no conversion)
__device__ __forceinline__ uint64_t dbits(double v) { return (uint64_t)__double_as_longlong(v); }
__device__ __forceinline__ double bitsd(uint64_t b) { return __longlong_as_double((long long)b); }

// forward butterfly X, Y <- X + wY, X - wY
__device__ __forceinline__ void bf_fwd(double &X, double &Y, double w, Mod M) {
    const double T = mulmod(Y, w, M);
    const double x = X;
    X = addmod(x, T, M);
    Y = submod(x, T, M);
}
// inverse butterfly X, Y <- X + Y, (X - Y) w
__device__ __forceinline__ void bf_inv(double &X, double &Y, double w, Mod M) {
    const double x = X, y = Y;
    X = addmod(x, y, M);
    Y = mulmod(submod(x, y, M), w, M);
}

// ---------------------------------------------------------------- quad exchange
// The exchange between phase A (thread t of a 32-point half holds positions
// t + 4k, k < 8) and phase B (positions 8t + x) without LDS.  The 4 threads
// of a half are the 4 lanes of a quad, and the exchange is a 4 x 4 transpose
// of value pairs across them: A -> B sends the pair (v[2j], v[2j + 1]) of
// thread t to thread j as its entries t and t + 4.  Two butterfly stages
// (lanes t ^ 1, then t ^ 2) of DPP quad permutes move the pairs in registers.
// pairs (a[j], b[j]), j < 4: lane t ends with lane j's pair t in slot j
__device__ __forceinline__ void quad_transpose(double (&a)[4], double (&b)[4], int t) {
    constexpr int X1 = 0xB1, X2 = 0x4E;  // quad_perm [1, 0, 3, 2] and [2, 3, 0, 1]
    double ca[4], cb[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        ca[j] = dpp_d<X1>(a[j ^ 1]);
        cb[j] = dpp_d<X1>(b[j ^ 1]);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const bool take = ((j ^ t) & 1) != 0;
        a[j] = take ? ca[j] : a[j];
        b[j] = take ? cb[j] : b[j];
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        ca[j] = dpp_d<X2>(a[j ^ 2]);
        cb[j] = dpp_d<X2>(b[j ^ 2]);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const bool take = ((j ^ t) & 2) != 0;
        a[j] = take ? ca[j] : a[j];
        b[j] = take ? cb[j] : b[j];
    }
}
// phase A layout -> phase B layout, and back
__device__ __forceinline__ void xch_a_to_b(double (&v)[8], int t) {
    double a[4], b[4];
#pragma unroll
    for (int j = 0; j < 4; j++) a[j] = v[2 * j], b[j] = v[2 * j + 1];
    quad_transpose(a, b, t);
#pragma unroll
    for (int s = 0; s < 4; s++) v[s] = a[s], v[s + 4] = b[s];
}
__device__ __forceinline__ void xch_b_to_a(double (&v)[8], int t) {
    double a[4], b[4];
#pragma unroll
    for (int s = 0; s < 4; s++) a[s] = v[s], b[s] = v[s + 4];
    quad_transpose(a, b, t);
#pragma unroll
    for (int j = 0; j < 4; j++) v[2 * j] = a[j], v[2 * j + 1] = b[j];
}

// ---------------------------------------------------------------- n = 32
// Thread t (0..3) of a half; phase A positions t + 4k (k < 8), phase B 8t + x.
__device__ __forceinline__ void ntt32_fwd_a(double (&v)[8], const double *__restrict__ tw, Mod M) {
#pragma unroll
    for (int k = 0; k < 4; k++) bf_fwd(v[k], v[k + 4], tw[1], M);  // distance 16
#pragma unroll
    for (int k = 0; k < 8; k++)
        if ((k & 2) == 0) bf_fwd(v[k], v[k + 2], tw[2 + (k >> 2)], M);  // distance 8
#pragma unroll
    for (int k = 0; k < 8; k += 2) bf_fwd(v[k], v[k + 1], tw[4 + (k >> 1)], M);  // distance 4
}
__device__ __forceinline__ void ntt32_fwd_b(double (&v)[8], const double *__restrict__ tw, int t,
                                            Mod M) {
    const double w8a = tw[8 + 2 * t], w8b = tw[8 + 2 * t + 1];
#pragma unroll
    for (int x = 0; x < 8; x++)
        if ((x & 2) == 0) bf_fwd(v[x], v[x + 2], (x & 4) ? w8b : w8a, M);  // distance 2
#pragma unroll
    for (int x = 0; x < 8; x += 2) bf_fwd(v[x], v[x + 1], tw[16 + 4 * t + (x >> 1)], M);  // distance 1
}
__device__ __forceinline__ void ntt32_inv_b(double (&v)[8], const double *__restrict__ twi, int t,
                                            Mod M) {
#pragma unroll
    for (int x = 0; x < 8; x += 2) bf_inv(v[x], v[x + 1], twi[16 + 4 * t + (x >> 1)], M);
    const double w8a = twi[8 + 2 * t], w8b = twi[8 + 2 * t + 1];
#pragma unroll
    for (int x = 0; x < 8; x++)
        if ((x & 2) == 0) bf_inv(v[x], v[x + 2], (x & 4) ? w8b : w8a, M);
}
__device__ __forceinline__ void ntt32_inv_a(double (&v)[8], const double *__restrict__ twi,
                                            double ninv, Mod M) {
#pragma unroll
    for (int k = 0; k < 8; k += 2) bf_inv(v[k], v[k + 1], twi[4 + (k >> 1)], M);
#pragma unroll
    for (int k = 0; k < 8; k++)
        if ((k & 2) == 0) bf_inv(v[k], v[k + 2], twi[2 + (k >> 2)], M);
    const double wl = twi[0];  // twi[1] * n^{-1}
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double x = v[k], y = v[k + 4];
        v[k] = mulmod(addmod(x, y, M), ninv, M);
        v[k + 4] = mulmod(submod(x, y, M), wl, M);
    }
}

// ---------------------------------------------------------------- n = 32, lazy
// The same transforms without canonical corrections between stages.  The
// values stay exact integers congruent to the canonical ones; twiddles are
// balanced (|w| <= q/2).  mul_lazy(a, b) returns a b - k q exactly with
// |.| <= q/2 + 3 * 2^-53 |a b| (the rounding of a b, of 1/q and of the
// quotient estimate), so with |w| <= q/2 a forward stage maps a bound B to
// B + q/2 + 0.1875 B (in units where q ~ 2^50): from canonical inputs the
// three stages of fwd_a stay below 3.5 q and the two of fwd_b below 6.0 q,
// inside 8 q < 2^53 where every sum and the final reduction are exact.
__device__ __forceinline__ void bf_fwd_lazy(double &X, double &Y, double w, Mod M) {
    const double T = mul_lazy(Y, w, M);
    const double x = X;
    X = x + T;
    Y = x - T;
}
__device__ __forceinline__ void bf_inv_lazy(double &X, double &Y, double w, Mod M) {
    const double x = X, y = Y;
    X = x + y;
    Y = mul_lazy(x - y, w, M);
}
// canonical in, |v| < 3.5 q out
__device__ __forceinline__ void ntt32_fwd_a_lazy(double (&v)[8], const double *__restrict__ tw,
                                                 Mod M) {
#pragma unroll
    for (int k = 0; k < 4; k++) bf_fwd_lazy(v[k], v[k + 4], tw[1], M);
#pragma unroll
    for (int k = 0; k < 8; k++)
        if ((k & 2) == 0) bf_fwd_lazy(v[k], v[k + 2], tw[2 + (k >> 2)], M);
#pragma unroll
    for (int k = 0; k < 8; k += 2) bf_fwd_lazy(v[k], v[k + 1], tw[4 + (k >> 1)], M);
}
// |v| < 3.5 q in, canonical out
__device__ __forceinline__ void ntt32_fwd_b_lazy(double (&v)[8], const double *__restrict__ tw,
                                                 int t, Mod M) {
    const double w8a = tw[8 + 2 * t], w8b = tw[8 + 2 * t + 1];
#pragma unroll
    for (int x = 0; x < 8; x++)
        if ((x & 2) == 0) bf_fwd_lazy(v[x], v[x + 2], (x & 4) ? w8b : w8a, M);
#pragma unroll
    for (int x = 0; x < 8; x += 2) bf_fwd_lazy(v[x], v[x + 1], tw[16 + 4 * t + (x >> 1)], M);
#pragma unroll
    for (int x = 0; x < 8; x++) v[x] = reduce(v[x], M);
}
// Inverse: X + Y doubles the bound per stage, (X - Y) w maps 2B to
// q/2 + 0.375 B.  inv_b: canonical in, |v| <= 4 q out; inv_a reduces its
// input to |v| <= q/2 + 2 first, runs two stages (|v| <= 2 q) and the
// last with n^{-1} (|x + y| <= 4 q), then reduces to canonical.
__device__ __forceinline__ void ntt32_inv_b_lazy(double (&v)[8], const double *__restrict__ twi,
                                                 int t, Mod M) {
#pragma unroll
    for (int x = 0; x < 8; x += 2) bf_inv_lazy(v[x], v[x + 1], twi[16 + 4 * t + (x >> 1)], M);
    const double w8a = twi[8 + 2 * t], w8b = twi[8 + 2 * t + 1];
#pragma unroll
    for (int x = 0; x < 8; x++)
        if ((x & 2) == 0) bf_inv_lazy(v[x], v[x + 2], (x & 4) ? w8b : w8a, M);
}
__device__ __forceinline__ void ntt32_inv_a_lazy(double (&v)[8], const double *__restrict__ twi,
                                                 double ninv, Mod M) {
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = reduce_lazy(v[k], M);
#pragma unroll
    for (int k = 0; k < 8; k += 2) bf_inv_lazy(v[k], v[k + 1], twi[4 + (k >> 1)], M);
#pragma unroll
    for (int k = 0; k < 8; k++)
        if ((k & 2) == 0) bf_inv_lazy(v[k], v[k + 2], twi[2 + (k >> 2)], M);
    const double wl = twi[0];  // twi[1] * n^{-1}
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double x = v[k], y = v[k + 4];
        v[k] = reduce(mul_lazy(x + y, ninv, M), M);
        v[k + 4] = reduce(mul_lazy(x - y, wl, M), M);
    }
}

// ---------------------------------------------------------------- n = 64
// Thread t (0..7); phase A positions t + 8k, phase B 8t + x.
__device__ __forceinline__ void ntt64_fwd_a(double (&v)[8], const double *__restrict__ tw, Mod M) {
#pragma unroll
    for (int k = 0; k < 4; k++) bf_fwd(v[k], v[k + 4], tw[1], M);  // distance 32
#pragma unroll
    for (int k = 0; k < 8; k++)
        if ((k & 2) == 0) bf_fwd(v[k], v[k + 2], tw[2 + (k >> 2)], M);  // distance 16
#pragma unroll
    for (int k = 0; k < 8; k += 2) bf_fwd(v[k], v[k + 1], tw[4 + (k >> 1)], M);  // distance 8
}
__device__ __forceinline__ void ntt64_fwd_b(double (&v)[8], const double *__restrict__ tw, int t,
                                            Mod M) {
    const double w8 = tw[8 + t];
#pragma unroll
    for (int x = 0; x < 4; x++) bf_fwd(v[x], v[x + 4], w8, M);  // distance 4
    const double w16a = tw[16 + 2 * t], w16b = tw[16 + 2 * t + 1];
#pragma unroll
    for (int x = 0; x < 8; x++)
        if ((x & 2) == 0) bf_fwd(v[x], v[x + 2], (x & 4) ? w16b : w16a, M);  // distance 2
#pragma unroll
    for (int x = 0; x < 8; x += 2) bf_fwd(v[x], v[x + 1], tw[32 + 4 * t + (x >> 1)], M);  // distance 1
}
__device__ __forceinline__ void ntt64_inv_b(double (&v)[8], const double *__restrict__ twi, int t,
                                            Mod M) {
#pragma unroll
    for (int x = 0; x < 8; x += 2) bf_inv(v[x], v[x + 1], twi[32 + 4 * t + (x >> 1)], M);
    const double w16a = twi[16 + 2 * t], w16b = twi[16 + 2 * t + 1];
#pragma unroll
    for (int x = 0; x < 8; x++)
        if ((x & 2) == 0) bf_inv(v[x], v[x + 2], (x & 4) ? w16b : w16a, M);
    const double w8 = twi[8 + t];
#pragma unroll
    for (int x = 0; x < 4; x++) bf_inv(v[x], v[x + 4], w8, M);
}
__device__ __forceinline__ void ntt64_inv_a(double (&v)[8], const double *__restrict__ twi,
                                            double ninv, Mod M) {
#pragma unroll
    for (int k = 0; k < 8; k += 2) bf_inv(v[k], v[k + 1], twi[4 + (k >> 1)], M);
#pragma unroll
    for (int k = 0; k < 8; k++)
        if ((k & 2) == 0) bf_inv(v[k], v[k + 2], twi[2 + (k >> 2)], M);
    const double wl = twi[0];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double x = v[k], y = v[k + 4];
        v[k] = mulmod(addmod(x, y, M), ninv, M);
        v[k + 4] = mulmod(submod(x, y, M), wl, M);
    }
}

}  // namespace salsaa
