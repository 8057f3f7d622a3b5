/*
 * salsaa_oracle.h -- CPU restatement of the SALSAA prover hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark call it to produce
 * expected values.  It is never linked into, loaded by or called from the
 * product library (salsaa_amd/), which must fail loudly without its HIP build.
 *
 * Every function restates the reference crate's algorithm and cites the file
 * and line it follows (paths relative to the reference repository root).
 *
 * Data layout (shared with the product C ABI):
 *   ring element     : d consecutive uint64 residues in [0, q)
 *   ring matrix      : [rows][cols][d], row-major
 *   A (commit key)   : [kappa][m][d]   (reference: Vec<PowerSeries> rows)
 *   W (witness)      : [r][m][d]       (reference: Vec of r rows of length m)
 *   commitment       : [kappa][r][d]   IncompleteNTT form
 *
 * Pinning status (what checks each function; see tests/test_oracle_*.py):
 *   or_mulmod/addmod/submod/powmod/invmod : reference KATs + __int128 brute force
 *   or_ntt_forward/inverse                : closed-form evaluation at the HEXL
 *                                           root (minimal primitive 2n-th root),
 *                                           round trip, negacyclic product KATs.
 *                                           HEXL itself is absent: the exact
 *                                           NTT-domain layout is pinned only by
 *                                           HEXL's documented convention.
 *   or_ring_mul_incomplete                : cyclotomic_ring.rs:229-242 KAT and
 *                                           schoolbook naive_multiply
 *   or_commit                             : split.rs:109-143 KAT, linearity,
 *                                           coefficient-domain schoolbook
 *   or_tensor                             : fast_norm_check.rs:41-57 KAT
 *   or_sumcheck                           : fast_norm_check.rs:237-262 invariant,
 *                                           per-round verifier identity, brute-force
 *                                           multilinear evaluation
 *   or_decompose/or_compose               : decomp.rs:78-92 invariant, digit
 *                                           definition brute force
 *   or_fold                               : fold.rs:36-55 homomorphism
 *   or_complete_power_series              : kronecker_product example of
 *                                           arithmetic.rs:866-872 (constants),
 *                                           all-ones tensors repeat the base,
 *                                           brute-force index formula
 *   or_dot_single_series                  : equals or_commit with kappa = 1
 *                                           (slot products vs schoolbook)
 *   or_project                            : signed matrix product mod q on
 *                                           small inputs, linearity, and the
 *                                           batching identity of project.rs:
 *                                           154-190 (a transposed challenge
 *                                           fails it)
 *   or_get_inv_roots                      : parity unpinned (no reference test
 *                                           reads HEXL's inverse table layout)
 */
#ifndef SALSAA_ORACLE_H
#define SALSAA_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalar arithmetic (hexl_wrapper.cpp:34-63) ---- */
uint64_t or_mulmod(uint64_t a, uint64_t b, uint64_t q);
uint64_t or_addmod(uint64_t a, uint64_t b, uint64_t q);
uint64_t or_submod(uint64_t a, uint64_t b, uint64_t q);
uint64_t or_powmod(uint64_t a, uint64_t e, uint64_t q);
uint64_t or_invmod(uint64_t a, uint64_t q);
uint64_t or_mulmod_slow(uint64_t a, uint64_t b, uint64_t q); /* __int128 % q */

/* ---- NTT, HEXL convention (hexl_wrapper.cpp:213-229) ---- */
uint64_t or_min_primitive_root(uint64_t order, uint64_t q);
/* roots[brv(i)] = psi^i, i < n  (HEXL GetRootOfUnityPowers) */
void or_get_roots(size_t n, uint64_t q, uint64_t *roots);
/* HEXL GetInvRootOfUnityPowers layout -- parity unpinned */
void or_get_inv_roots(size_t n, uint64_t q, uint64_t *inv_roots);
void or_ntt_forward(uint64_t *x, size_t n, uint64_t q);
void or_ntt_inverse(uint64_t *x, size_t n, uint64_t q);
/* out[j] = sum_k x[k] psi^{(2 brv(j)+1) k}: closed form used to pin the NTT */
void or_ntt_forward_direct(const uint64_t *x, uint64_t *out, size_t n, uint64_t q);

/* ---- representations (cyclotomic_ring.rs:483-556) ---- */
void or_to_incomplete(uint64_t *a, size_t d, uint64_t q);
void or_from_incomplete(uint64_t *a, size_t d, uint64_t q);
void or_to_full(uint64_t *a, size_t d, uint64_t q);
void or_from_full(uint64_t *a, size_t d, uint64_t q);
/* omega_j = NTT_{d/2}(Y)[j] (cyclotomic_ring.rs:559-564) */
void or_shift_factors(size_t d, uint64_t q, uint64_t *omega);

/* ---- ring products (cyclotomic_ring.rs:699-839) ---- */
/* a, b in Coefficient form; out in IncompleteNTT form */
void or_ring_mul_incomplete(const uint64_t *a, const uint64_t *b, uint64_t *out,
                            size_t d, uint64_t q);
/* schoolbook negacyclic product, all Coefficient form */
void or_ring_mul_naive(const uint64_t *a, const uint64_t *b, uint64_t *out,
                       size_t d, uint64_t q);
void or_conjugate(const uint64_t *a, uint64_t *out, size_t d, uint64_t q);

/* ---- commit (arithmetic.rs:459-505) ----
 * A [kappa][m][d] and W [r][m][d] in Coefficient form; out [kappa][r][d]
 * IncompleteNTT.  Follows the reference's per-product conversion. */
void or_commit(const uint64_t *A, const uint64_t *W, uint64_t *out, size_t kappa,
               size_t r, size_t m, size_t d, uint64_t q, int threads);
/* complete_power_series / kronecker_product (arithmetic.rs:158-176, 845-864),
 * Coefficient form: base [m0][d], tensors [n][2][d] -> out [m0 << n][d] */
void or_complete_power_series(const uint64_t *base, size_t m0, const uint64_t *tensors, size_t n,
                              size_t d, uint64_t q, uint64_t *out);
/* parallel_dot_single_series_matrix (arithmetic.rs:507-536): layer [m][d],
 * W [r][m][d] Coefficient -> out [r][d] IncompleteNTT */
void or_dot_single_series(const uint64_t *layer, const uint64_t *W, uint64_t *out, size_t r,
                          size_t m, size_t d, uint64_t q);
/* project (project.rs:52-98) before its commit: W [r][m][d], is_zero and sign
 * [256][r 256] bytes -> out [r (m / (r 256)) 256][d] */
void or_project(const uint64_t *W, size_t r, size_t m, size_t d, const uint8_t *is_zero,
                const uint8_t *sign, uint64_t q, uint64_t *out);

/* ---- norm check (fast_norm_check.rs) ---- */
void or_tensor(const uint64_t *cs, size_t n, uint64_t q, uint64_t *out);
/* raw reference sums: sum of per-chunk residues, chunk = max(len/120, 1028) */
uint64_t or_sum_chunked_raw(const uint64_t *x, size_t len, uint64_t q);
uint64_t or_sum_sq_chunked_raw(const uint64_t *x, size_t len, uint64_t q);

/* Sumcheck norm check over F_q or F_{q^2} = F_q[Y]/(Y^2 - nr).
 * w       : [rows][L] flattened witness rows (Coefficient data, index i*d+k)
 * cs      : [reps][nrounds][2] challenges (second coordinate ignored for F_q)
 * claims  : [rows]               sum of squares mod q (norm_check_1)
 * msgs    : [reps][nrounds][rows][3][2] canonical (f0,f1,f2)
 * finals  : [reps][rows][2]       final claims after all rounds
 * evals   : [reps][rows][2]       <w, tensor(cs)>  (folded value)
 * ok      : [reps][rows]          1 if every verifier check passed
 * fq2 = 0 : F_q mode (reference semantics); fq2 = 1: F_{q^2} challenges. */
void or_sumcheck(const uint64_t *w, size_t rows, size_t L, const uint64_t *cs,
                 size_t reps, int fq2, uint64_t nr, uint64_t q, uint64_t *claims,
                 uint64_t *msgs, uint64_t *finals, uint64_t *evals, int *ok,
                 int threads);

/* The sumcheck with all rows in step and the challenge of each round from a
 * callback given that round's messages [rows][6] (the Fiat-Shamir replay);
 * cs_out [reps][nrounds][2] receives the challenges, the other outputs are
 * or_sumcheck's (f2 summed directly). */
typedef void (*or_chal_fn)(void *user, size_t rep, size_t t, const uint64_t *round_msgs,
                           uint64_t *c_out);
void or_sumcheck_cb(const uint64_t *w, size_t rows, size_t L, size_t reps, int fq2, uint64_t nr,
                    uint64_t q, uint64_t *claims, uint64_t *msgs, uint64_t *cs_out,
                    uint64_t *finals, uint64_t *evals, int *ok, or_chal_fn chal, void *user);

/* norm_check_3 (fast_norm_check.rs:401-441) for F_q challenges:
 * rhs[row] = sum_i ps_i * W[row][i], IncompleteNTT form.  cs has log2(m*d)
 * entries (F_q). */
void or_norm_check_3(const uint64_t *W, size_t rows, size_t m, size_t d,
                     const uint64_t *cs, uint64_t q, uint64_t *rhs);
/* F_{q^2} analogue: rhs = base * V over R_q[Y]/(Y^2-nr): out [rows][2][d] */
void or_norm_check_3_fq2(const uint64_t *W, size_t rows, size_t m, size_t d,
                         const uint64_t *cs2, uint64_t nr, uint64_t q, uint64_t *rhs);

/* ---- decomposition (arithmetic.rs:1461-1593, 1719-1755) ----
 * W [r][m][d] Coefficient; out [r*parts][m][d] Coefficient.
 * radix = 0 selects the reference's auto radix; returns the radix used. */
uint64_t or_decompose(const uint64_t *W, size_t r, size_t m, size_t d, size_t parts,
                      uint64_t radix, uint64_t q, uint64_t *out);
/* compose_with_radix_mod: [rows][n*parts][d] -> [rows][n][d] (columns) */
void or_compose(const uint64_t *in, size_t rows, size_t n, size_t d, size_t parts,
                uint64_t radix, uint64_t q, uint64_t *out);
/* gadget recomposition of limb rows: [r*parts][m][d] -> [r][m][d] */
void or_compose_rows(const uint64_t *limbs, size_t r, size_t m, size_t d, size_t parts,
                     uint64_t radix, uint64_t q, uint64_t *out);

/* ---- fold (fold.rs:22-27, arithmetic.rs:429-457) ----
 * out[j][i] = sum_k W[k][i] * C[j][k]; W [r_in][m][d], C [r_out][r_in][d]
 * (both Coefficient); out [r_out][m][d] IncompleteNTT. */
void or_fold(const uint64_t *W, const uint64_t *C, uint64_t *out, size_t r_in,
             size_t r_out, size_t m, size_t d, uint64_t q, int threads);

/* F_{q^2} product (quadratic_extension.rs:30-58), shift = nr */
void or_fq2_mul(uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1, uint64_t nr,
                uint64_t q, uint64_t *out);

/* ---- seeded sampling (the reference's samplers are unseeded) ---- */
uint64_t or_prng_draw(uint64_t seed, uint64_t stream, uint64_t idx);
void or_sample_uniform(uint64_t seed, uint64_t stream, size_t n, uint64_t q,
                       uint64_t *out);
void or_sample_uniform_at(uint64_t seed, uint64_t stream, size_t start, size_t n, uint64_t q,
                          uint64_t *out);
void or_sample_bits(uint64_t seed, uint64_t stream, size_t n, uint8_t *out);
uint64_t or_stream_key(uint64_t seed, uint64_t stream);
void or_sample_bounded(uint64_t seed, uint64_t stream, size_t n, uint64_t bound,
                       uint64_t q, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif
