// ntt_device.hpp -- per-thread negacyclic NTTs over registers (HEXL
// convention: psi = minimal primitive 2n-th root, forward output in
// bit-reversed order, inverse includes n^{-1}).  Fully unrolled so every
// twiddle index is a compile-time constant: twiddles are uniform and come in
// through scalar loads, the n values stay in VGPRs.
#pragma once

#include "common.hpp"

namespace salsaa {

// Forward Cooley-Tukey: stage m (= 1, 2, ..., N/2) pairs (j, j + t), t = N/(2m),
// in blocks i with twiddle tw[m + i] = psi^{brv(m + i)}.
template <int N>
__device__ __forceinline__ void ntt_fwd_regs(double (&v)[N], const double *__restrict__ tw,
                                             Mod M) {
    constexpr int LOG = (N == 2) ? 1 : (N == 4) ? 2 : (N == 8) ? 3 : (N == 16) ? 4
                      : (N == 32) ? 5 : 6;
#pragma unroll
    for (int s = 0; s < LOG; s++) {
        const int m = 1 << s;
        const int t = N >> (s + 1);
#pragma unroll
        for (int i = 0; i < m; i++) {
            const double w = tw[m + i];
#pragma unroll
            for (int jj = 0; jj < t; jj++) {
                const int j = 2 * i * t + jj;
                double X = v[j];
                double Y = mulmod(v[j + t], w, M);
                v[j] = addmod(X, Y, M);
                v[j + t] = submod(X, Y, M);
            }
        }
    }
}

// Inverse Gentleman-Sande, stages in reverse order with inverse twiddles
// twi[m + i]; the last stage (m = 1) folds in n^{-1}: twi[0] must hold
// twi[1] * n^{-1} and ninv = n^{-1}.
template <int N>
__device__ __forceinline__ void ntt_inv_regs(double (&v)[N], const double *__restrict__ twi,
                                             double ninv, Mod M) {
    constexpr int LOG = (N == 2) ? 1 : (N == 4) ? 2 : (N == 8) ? 3 : (N == 16) ? 4
                      : (N == 32) ? 5 : 6;
#pragma unroll
    for (int s = LOG - 1; s >= 0; s--) {
        const int m = 1 << s;
        const int t = N >> (s + 1);
#pragma unroll
        for (int i = 0; i < m; i++) {
            const double w = (s == 0) ? twi[0] : twi[m + i];
#pragma unroll
            for (int jj = 0; jj < t; jj++) {
                const int j = 2 * i * t + jj;
                double X = v[j], Y = v[j + t];
                if (s == 0) {
                    v[j] = mulmod(addmod(X, Y, M), ninv, M);
                } else {
                    v[j] = addmod(X, Y, M);
                }
                v[j + t] = mulmod(submod(X, Y, M), w, M);
            }
        }
    }
}

}  // namespace salsaa
