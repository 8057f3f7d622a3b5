/*
 * salsaa_amd.h -- C ABI of the MI355X (gfx950) SALSAA prover hot path.
 *
 * Two groups of entry points:
 *
 * 1. The legacy hexl_wrapper symbols, with the names, signatures and semantics
 *    the reference's Rust FFI binds (src/hexl/bindings.rs:3-73, defined in
 *    hexl-bindings/hexl_wrapper.cpp).  They take host pointers to single
 *    vectors (one ring element or one scalar at a time), exactly as the
 *    reference calls them, and run on the host: this keeps the unchanged
 *    reference crate linkable.  They are thread-safe.
 *
 * 2. Batched device entry points that replace the reference's operator API
 *    (src/arithmetic.rs and the files in src/subroutines/) on whole matrices resident in
 *    HBM.  All pointers named d_* are device pointers.  Every function returns
 *    an int status (SALSAA_OK = 0, negative on error) and never throws;
 *    salsaa_last_error() gives the message.  Work is enqueued on the context's
 *    stream; a call returns once enqueued unless stated otherwise.
 *
 * Data layout: a ring element is d consecutive uint64 canonical residues in
 * [0, q); a ring matrix is [rows][cols][d] row-major.  Representation tags:
 * SALSAA_REP_COEFF (Coefficient), SALSAA_REP_INCOMPLETE (IncompleteNTT: data =
 * NTT_{d/2}(even coeffs) || NTT_{d/2}(odd coeffs), HEXL convention), and
 * SALSAA_REP_FULL (NTT: NTT_d of the coefficients)
 * (reference: src/cyclotomic_ring.rs:13-18, 483-556).
 *
 * The device kernels support any prime q < 2^50 with q = 1 mod 128 at d = 64,
 * and at the reference's own degrees d = 128, 256, 512 (protocol.rs:28-120)
 * when also q = 1 mod d.  At d > 64 the conversions (the NTT representation
 * only when q = 1 mod 2d), the ring product, keys and the commit (kappa <= 18),
 * decomposition, decomp_ell, compose / verify_decomp, the fold and the norm
 * check run; every other entry point returns SALSAA_ERR_ARG for such a context.
 */
#ifndef SALSAA_AMD_H
#define SALSAA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* 1. legacy hexl_wrapper ABI (host)                                        */
/* ------------------------------------------------------------------------ */

/* hexl_wrapper.cpp:34-36, bindings.rs:3 */
uint64_t multiply_mod(uint64_t a, uint64_t b, uint64_t modulus);
/* hexl_wrapper.cpp:61-63, bindings.rs:5 */
uint64_t power_mod(uint64_t a, uint64_t b, uint64_t modulus);
/* hexl_wrapper.cpp:38-41, bindings.rs:7 */
uint64_t add_mod(uint64_t a, uint64_t b, uint64_t modulus);
/* hexl_wrapper.cpp:43-45, bindings.rs:9 */
uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t modulus);
/* hexl_wrapper.cpp:65-67, bindings.rs:11-17 */
void eltwise_mult_mod(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2,
                      uint64_t n, uint64_t modulus);
/* hexl_wrapper.cpp:47-50, bindings.rs:19: bit-reversed psi powers (n entries),
 * owned by a process-lifetime cache */
const uint64_t *get_roots(uint64_t n, uint64_t modulus);
/* hexl_wrapper.cpp:52-54, bindings.rs:20 */
uint64_t inv_mod(uint64_t a, uint64_t modulus);
/* hexl_wrapper.cpp:56-59, bindings.rs:21: the inverse roots psi^-brv(k) in
 * this library's own order (inv[0], then inv[m .. 2m) for m = n/2, n/4, ..,
 * 1, the order its inverse transform reads them).  The reference declares
 * this symbol but never calls it and no reference test reads the table, so
 * whether the layout equals HEXL's is not established. */
const uint64_t *get_inv_roots(uint64_t n, uint64_t modulus);
/* hexl_wrapper.cpp:262-268, bindings.rs:23 */
uint64_t sum_sq(const uint64_t *operand1, uint64_t n, uint64_t modulus);
/* hexl_wrapper.cpp:271-276, bindings.rs:24 */
uint64_t sum_sq_fast(const uint64_t *operand1, uint64_t n, uint64_t modulus);
/* hexl_wrapper.cpp:231-237, bindings.rs:25 */
uint64_t sum(const uint64_t *operand1, uint64_t n, uint64_t modulus);
/* hexl_wrapper.cpp:239-260, bindings.rs:26 (defined result: sum mod q) */
uint64_t sum_fast(const uint64_t *operand1, uint64_t n, uint64_t modulus);
/* hexl_wrapper.cpp:68-70, bindings.rs:27-34: result = operand1 * operand2 + operand3 */
void eltwise_fma_mod(uint64_t *result, const uint64_t *operand1, uint64_t operand2,
                     const uint64_t *operand3, uint64_t n, uint64_t modulus);
/* hexl_wrapper.cpp:188-196, bindings.rs:36-42 */
void eltwise_add_mod(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2,
                     uint64_t n, uint64_t modulus);
/* hexl_wrapper.cpp:199-207, bindings.rs:44-50 */
void eltwise_sub_mod(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2,
                     uint64_t n, uint64_t modulus);
/* hexl_wrapper.cpp:73-93, bindings.rs:52-58: product mod X^n + 1 */
void multiply_poly(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2,
                   uint64_t n, uint64_t modulus);
/* hexl_wrapper.cpp:119-126, bindings.rs:60: result = operand mod q, any input */
void eltwise_reduce_mod(uint64_t *result, const uint64_t *operand, uint64_t n,
                        uint64_t modulus);
/* hexl_wrapper.cpp:158-186, bindings.rs:61 */
void eltwise_reduce_mod_naive(uint64_t *result, const uint64_t *operand, uint64_t n,
                              uint64_t modulus);
/* hexl_wrapper.cpp:95-117, bindings.rs:63-69 */
void polynomial_multiply_cyclotomic_mod(uint64_t *result, const uint64_t *operand1,
                                        const uint64_t *operand2, uint64_t phi,
                                        uint64_t mod_q);
/* hexl_wrapper.cpp:213-220, bindings.rs:71: HEXL negacyclic NTT, in place */
void ntt_forward_in_place(uint64_t *operand, size_t n, uint64_t modulus);
/* hexl_wrapper.cpp:222-229, bindings.rs:73 */
void ntt_inverse_in_place(uint64_t *operand, size_t n, uint64_t modulus);

/* ------------------------------------------------------------------------ */
/* 2. batched device API                                                    */
/* ------------------------------------------------------------------------ */

#define SALSAA_OK 0
#define SALSAA_ERR_ARG -1
#define SALSAA_ERR_HIP -2
#define SALSAA_ERR_NOMEM -3
#define SALSAA_ERR_CHECK -4 /* a prover-side self-check failed */

#define SALSAA_REP_COEFF 0
#define SALSAA_REP_FULL 1
#define SALSAA_REP_INCOMPLETE 2

typedef struct salsaa_ctx salsaa_ctx;
typedef struct salsaa_key salsaa_key;

/* Message of the last failing call on this thread. */
const char *salsaa_last_error(void);
/* Library version string. */
const char *salsaa_version(void);

/* Context for modulus q and degree d on the current HIP device: builds the
 * NTT tables (the role of the reference's NTTCache, hexl_wrapper.cpp:11-32)
 * and one default-priority stream.  d = 64, or 128 / 256 / 512 with
 * q = 1 mod d (the entry points listed above). */
int salsaa_ctx_create(uint64_t q, uint32_t d, salsaa_ctx **out);
int salsaa_ctx_destroy(salsaa_ctx *ctx);
int salsaa_ctx_sync(salsaa_ctx *ctx);
/* Waits for the context's stream, then frees the device memory the context
 * keeps for reuse between calls: operator scratch and the protocol's run
 * buffers (a protocol run keeps what it used for the next run).  Later calls
 * allocate again as needed.  Resident keys, batches and servers are not
 * touched. */
int salsaa_ctx_trim(salsaa_ctx *ctx);
/* The context's stream (hipStream_t) so callers can order their own work. */
void *salsaa_ctx_stream(salsaa_ctx *ctx);

/* Device memory and copies (stream-ordered on the context stream). */
int salsaa_dev_alloc(salsaa_ctx *ctx, size_t bytes, void **d_ptr);
int salsaa_dev_free(salsaa_ctx *ctx, void *d_ptr);
int salsaa_memcpy_h2d(salsaa_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int salsaa_memcpy_d2h(salsaa_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);
/* Stream-ordered device copy (also the plain copy kernel the traffic
 * counters are calibrated against). */
int salsaa_memcpy_d2d(salsaa_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);
/* Bandwidth probe, no counterpart in the reference: a plain streaming kernel
 * with rn read streams and wn write streams of n16 16-byte vectors each, laid
 * out back to back in d_buf (reads first, so d_buf holds (rn + wn) * n16 * 16
 * bytes; with inplace, write stream 0 overwrites read stream 0 and d_buf
 * holds (rn + wn - 1) * n16 * 16 bytes), on blocks_per_cu 256-thread blocks
 * per CU.  It measures the HBM
 * rate a given read/write mix reaches, which the benchmark reports beside the
 * fused passes that have that mix. */
int salsaa_memstream(salsaa_ctx *ctx, void *d_buf, size_t n16, int rn, int wn, int inplace,
                     int blocks_per_cu);

/* Events on the context stream, for timing enqueued work. */
int salsaa_event_create(void **ev);
int salsaa_event_destroy(void *ev);
int salsaa_event_record(salsaa_ctx *ctx, void *ev);
int salsaa_event_elapsed_ms(void *start, void *stop, float *ms);

/* Seeded sampling on the device, element i of the stream drawn from
 * SplitMix64(seed, stream, i) (the reference samplers, cyclotomic_ring.rs:
 * 257-317, are unseeded).  uniform: [0, q); bounded: uniform over the
 * 2*bound+1 integers [-bound, bound], mapped mod q. */
int salsaa_sample_uniform(salsaa_ctx *ctx, uint64_t *d_out, size_t n, uint64_t seed,
                          uint64_t stream);
int salsaa_sample_bounded(salsaa_ctx *ctx, uint64_t *d_out, size_t n, uint64_t bound,
                          uint64_t seed, uint64_t stream);
/* Sumcheck challenges cs[reps][nrounds][2]: coordinate 0 (and 1 when fq2) is
 * uniform mod q from stream 0x5C000 + rep, index 2*round (+1). */
int salsaa_sample_challenges(salsaa_ctx *ctx, uint64_t *d_cs, size_t reps, size_t nrounds,
                             int fq2, uint64_t seed);

/* Batched representation change of `count` ring elements in place:
 * to_incomplete_ntt_representation / to_ntt_representation /
 * to_coeff_representation (cyclotomic_ring.rs:483-556), one thread per
 * length-d/2 or length-d transform.  from_rep, to_rep: SALSAA_REP_*. */
int salsaa_convert(salsaa_ctx *ctx, uint64_t *d_data, size_t count, int from_rep,
                   int to_rep);

/* Batched ring product out[i] = a[i] * b[i] (both Coefficient), result in
 * IncompleteNTT form: incomplete_ntt_multiplication (cyclotomic_ring.rs:
 * 699-791) for `count` pairs. */
int salsaa_ring_mul(salsaa_ctx *ctx, const uint64_t *d_a, const uint64_t *d_b,
                    uint64_t *d_out, size_t count);

/* Commitment key: A [kappa][m][d] Coefficient, made resident in HBM in the
 * slot-major full-NTT layout the commit kernel reads.  The reference stores
 * the key as Vec<PowerSeries> (arithmetic.rs:146-150); a uniform key has one
 * layer of length m; commits with a shorter witness use the prefix layer. */
int salsaa_key_create(salsaa_ctx *ctx, const uint64_t *d_A, size_t kappa, size_t m,
                      salsaa_key **out);
/* compute_commitment_keys (crs.rs:43-72), the CRS of CRS::gen_crs
 * (crs.rs:22-30) for the module elements d_module [kappa][d] (Coefficient):
 * row j = m_j, m_j^2, ..., m_j^len, d_out [kappa][len][d] Coefficient.  The
 * reference's PowerSeries keeps this row's prefixes of lengths len, len/2,
 * ... (while even) as its layers and [1, m_j^c] (= [1, row[c - 1]]) as the
 * tensor below the layer of length 2c. */
int salsaa_compute_commitment_keys(salsaa_ctx *ctx, const uint64_t *d_module, size_t kappa,
                                   size_t len, uint64_t *d_out);
/* The commitment key of that CRS (kappa rows of length m), generated on the
 * device straight into the resident layout: the same key as
 * salsaa_key_create on salsaa_compute_commitment_keys' output, without the
 * [kappa][m][d] matrix (the stored ck of the reference, protocol.rs:132). */
int salsaa_key_create_crs(salsaa_ctx *ctx, const uint64_t *d_module, size_t kappa, size_t m,
                          salsaa_key **out);
/* The same CRS key without any stored layout (kappa <= 36): it keeps only the
 * module (kappa ring elements), and salsaa_parallel_dot_series_matrix and
 * salsaa_split, salsaa_decomp_ell, salsaa_fold_step, batches and the server
 * generate the key rows inside the commit kernels (the ck of protocol.rs:132
 * never exists in memory).  Views work as for other keys. */
int salsaa_key_crs_implicit(salsaa_ctx *ctx, const uint64_t *d_module, size_t kappa, size_t m,
                            salsaa_key **out);
/* A view of the first kappa rows and first m ring elements of a key, sharing
 * its storage (no copy): the counterpart of passing the prefix slice
 * &key[..kappa] of the Vec<PowerSeries> to parallel_dot_series_matrix
 * (arithmetic.rs:459-476).  The view must be destroyed before the key. */
int salsaa_key_view(const salsaa_key *key, size_t kappa, size_t m, salsaa_key **out);
int salsaa_key_destroy(salsaa_key *key);

/* parallel_dot_series_matrix (arithmetic.rs:459-505): commit a witness
 * W [r][m][d] (Coefficient) under the key's length-m prefix:
 * out[j][k] = sum_i A_j[i] * W_k[i], out [kappa][r][d] IncompleteNTT. */
int salsaa_parallel_dot_series_matrix(salsaa_ctx *ctx, const salsaa_key *key,
                                      const uint64_t *d_W, size_t r, size_t m,
                                      uint64_t *d_out);

/* complete_power_series (arithmetic.rs:158-176) with kronecker_product
 * (arithmetic.rs:845-864): the largest expanded layer of a power series with
 * base layer d_base [m0][d] and length-2 tensors d_tensors [n][2][d]
 * (IncompletePowerSeries order: tensors[0] is the most significant factor),
 *   d_out[((i_0 2 + i_1) 2 + ... + i_{n-1}) m0 + j] = t_0[i_0] ... t_{n-1}[i_{n-1}] base[j],
 * [m0 << n][d].  Inputs and output in Coefficient form (the reference's
 * products carry IncompleteNTT tags; the ring elements are the same). */
int salsaa_complete_power_series(salsaa_ctx *ctx, const uint64_t *d_base, size_t m0,
                                 const uint64_t *d_tensors, size_t n, uint64_t *d_out);

/* parallel_dot_single_series_matrix (arithmetic.rs:507-536): out[j] =
 * sum_i layer[i] * W[j][i] for the expanded layer of length m (the one the
 * reference finds by length); d_layer [m][d], d_W [r][m][d] Coefficient,
 * d_out [r][d] IncompleteNTT.  Up to 16 rows in one pass over the layer and
 * the rows in the NTT domain; more rows through the commit with the layer as
 * its only key row (a key is made and freed per call). */
int salsaa_parallel_dot_single_series_matrix(salsaa_ctx *ctx, const uint64_t *d_layer, size_t m,
                                             const uint64_t *d_W, size_t r, uint64_t *d_out);

/* project (project.rs:52-98) up to its closing commit: with cw = r * 256,
 * chunk j < m / cw of row i of W [r][m][d] gives the 256 ring elements
 *   out[(i (m / cw) + j) 256 + k] = sum_l C[k][l] W[i][j cw + l] mod q,
 * C[k][l] = 0 where d_is_zero[k][l], else +1 or -1 by d_sign[k][l] (the
 * Challenge of challenge_for_project, project.rs:19-39, as bytes 0/1,
 * [256][cw] each).  d_out [r (m / cw) 256][d], in W's representation; the
 * reference then commits it (parallel_dot_series_matrix).  Needs cw <= m
 * (the reference panics) and cw (q - 1) < 2^64 (its u64 sums wrap past it). */
int salsaa_project(salsaa_ctx *ctx, const uint64_t *d_W, size_t r, size_t m,
                   const uint8_t *d_is_zero, const uint8_t *d_sign, uint64_t *d_out);

/* batch_projections (project.rs:131-203) for the projection d_P [m][d] of
 * W [r][m][d] under the challenge (d_is_zero, d_sign [256][256 r]) with the
 * batching challenge c0 (d_c0 [d], the middle element of
 * batching_challenge_for_project, project.rs:104-116): with
 * c_l = sum_k C[k][l] c0^k (l < 256 r),
 *   d_witness_lhs   [m][d]  the layer c_{x mod 256 r}          (may be NULL)
 *   d_projected_lhs [m][d]  the layer c0^{x mod 256}            (may be NULL)
 *   d_witness_rhs   [r][d]  sum_x c_{x mod 256 r} W[j][x]      IncompleteNTT
 *   d_projection_rhs  [d]   sum_x c0^{x mod 256} P[x]          IncompleteNTT
 *   d_ok            [1]     1 if sum_j witness_rhs[j] = projection_rhs (the
 *                           assertion of project.rs:198), else 0 (may be NULL)
 * Inputs and layers in Coefficient form.  Needs r and m / (256 r) powers of
 * two (the reference finds no matching layer otherwise). */
int salsaa_batch_projections(salsaa_ctx *ctx, const uint64_t *d_P, const uint64_t *d_W, size_t r,
                             size_t m, const uint8_t *d_is_zero, const uint8_t *d_sign,
                             const uint64_t *d_c0, uint64_t *d_projected_lhs,
                             uint64_t *d_witness_lhs, uint64_t *d_projection_rhs,
                             uint64_t *d_witness_rhs, int32_t *d_ok);

/* send_cross_terms_before_join (project.rs:229-245): l0 = sum_x
 * projected_lhs[x] W[0][x] and l1 = sum_x witness_lhs[x] P[x] for the layers
 * [m][d] of salsaa_batch_projections, the folded witness W (row 0 is read)
 * and the projection P [m][d]; d_l0, d_l1 [d] IncompleteNTT.  (The reference
 * forms the products of every row and keeps entry 0.) */
int salsaa_send_cross_terms_before_join(salsaa_ctx *ctx, const uint64_t *d_projected_lhs,
                                        const uint64_t *d_witness_lhs, size_t m,
                                        const uint64_t *d_W, const uint64_t *d_P, uint64_t *d_l0,
                                        uint64_t *d_l1);

/* split (split.rs:45-76): the rows [L_0..L_{r-1}, R_0..R_{r-1}] of the halves
 * of W [r][m][d] (a view, not copied) are committed under the length-m/2
 * prefix: d_out [kappa][2r][d] IncompleteNTT. */
int salsaa_split(salsaa_ctx *ctx, const salsaa_key *key, const uint64_t *d_W, size_t r,
                 size_t m, uint64_t *d_out);

/* Norm check = norm_check_1 + norm_check_2 + norm_check_3 with
 * verify_norm_check_2/3 (fast_norm_check.rs:282-462) for every row of
 * w [rows][L] (Coefficient data flattened, L = m*d a power of two), `reps`
 * repetitions (protocol.rs:222-259) with the given challenges
 * cs [reps][log2 L][2].  fq2 = 0: challenges in F_q (reference semantics);
 * fq2 = 1: challenges in F_{q^2} = F_q[Y]/(Y^2 - nr).  Outputs (device):
 *   claims [rows]                  sum of squares mod q (norm_check_1)
 *   msgs   [reps][log2 L][rows][3][2]  round messages (f0, f1, f2), canonical
 *   finals [reps][rows][2]         verifier claims after the last round
 *   evals  [reps][rows][2]         folded witness <w, tensor(cs)>
 *   rhs    [reps][rows][2][d]      norm_check_3 rhs = base * V, IncompleteNTT
 *                                  (component 1 is zero for F_q challenges);
 *                                  requires L >= d
 *   ok     [reps][rows]            1 if every verifier check passed
 * Any output pointer except ok may be NULL to skip it. */
int salsaa_norm_check(salsaa_ctx *ctx, const uint64_t *d_w, size_t rows, size_t L,
                      const uint64_t *d_cs, size_t reps, int fq2, uint64_t nr,
                      uint64_t *d_claims, uint64_t *d_msgs, uint64_t *d_finals,
                      uint64_t *d_evals, uint64_t *d_rhs, int32_t *d_ok);

/* salsaa_norm_check without any allocation.  No counterpart in the
 * reference: this entry point is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, outputs and kernels as
 * salsaa_norm_check; the scratch is a fixed 64 MiB device array that each GPU
 * holds for the life of the process, so reps * rows * L must stay within
 * about 60 M (for example 2 reps x 8 rows x 2^21 coefficients); larger calls
 * return SALSAA_ERR_ARG.  Every call on one GPU shares that array, so calls
 * that overlap on one GPU (from several contexts or streams) race; calls on
 * one context are ordered by its stream. */
int salsaa_norm_check_noalloc(salsaa_ctx *ctx, const uint64_t *d_w, size_t rows, size_t L,
                              const uint64_t *d_cs, size_t reps, int fq2, uint64_t nr,
                              uint64_t *d_claims, uint64_t *d_msgs, uint64_t *d_finals,
                              uint64_t *d_evals, uint64_t *d_rhs, int32_t *d_ok);

/* decompose_matrix_by_chunks (arithmetic.rs:1572-1593): W [r][m][d]
 * Coefficient -> d_out [r*parts][m][d], limb t of row rho in row rho*parts+t.
 * radix = 0 selects the reference's auto radix 2^ceil(bw/parts); the radix
 * used is written to *d_radix (device). */
int salsaa_decompose(salsaa_ctx *ctx, const uint64_t *d_W, size_t r, size_t m,
                     size_t parts, uint64_t radix, uint64_t *d_out, uint64_t *d_radix);

/* decomp_ell (decomp.rs:32-67): decompose, then commit the limbs:
 * d_new_W [r*parts][m][d], d_rhs [kappa][r*parts][d] IncompleteNTT. */
int salsaa_decomp_ell(salsaa_ctx *ctx, const salsaa_key *key, const uint64_t *d_W,
                      size_t r, size_t m, size_t parts, uint64_t radix, uint64_t *d_new_W,
                      uint64_t *d_rhs, uint64_t *d_radix);

/* compose_with_radix_mod (arithmetic.rs:1719-1755) on the device:
 * in [rows][n*parts][d] -> out [rows][n][d], radix read from *d_radix. */
int salsaa_compose(salsaa_ctx *ctx, const uint64_t *d_in, size_t rows, size_t n,
                   size_t parts, const uint64_t *d_radix, uint64_t *d_out);

/* verify_decomp (decomp.rs:95-112): counts entries where
 * compose(d_rhs_dec [kappa][r*parts][d]) differs from d_rhs [kappa][r][d];
 * with rows_mode = 1 instead checks the witness limbs
 * d_rhs_dec = limbs [r*parts][m][d] against d_rhs = W [r][m][d] (pass
 * kappa = r, r = m).  *d_mismatch (device uint32) is incremented. */
int salsaa_verify_decomp(salsaa_ctx *ctx, const uint64_t *d_rhs_dec, const uint64_t *d_rhs,
                         size_t kappa, size_t r, size_t parts, const uint64_t *d_radix,
                         int rows_mode, uint32_t *d_mismatch);

/* fold (fold.rs:22-27 -> parallel_dot_matrix_matrix, arithmetic.rs:429-457):
 * out[j][i] = sum_k W[k][i] * C[j][k]; W [r_in][m][d] and C [r_out][r_in][d]
 * Coefficient, out [r_out][m][d] IncompleteNTT. */
int salsaa_parallel_dot_matrix_matrix(salsaa_ctx *ctx, const uint64_t *d_W, size_t r_in,
                                      size_t m, const uint64_t *d_C, size_t r_out,
                                      uint64_t *d_out);

/* ------------------------------------------------------------------ batches
 * Many independent prover steps over one resident key.  Instance i has one
 * witness row of m[i] ring elements (a power of two, at most the key's m) and
 * commits with the key's first kappa[i] rows.  One step of an instance is
 * what a caller of the per-instance operators does:
 *   to_incomplete_ntt_representation and back on the witness
 *     (cyclotomic_ring.rs:483-556; the witness is restored),
 *   parallel_dot_series_matrix (arithmetic.rs:459-505) -> commitment,
 *   norm_check_1/2/3 with their verifier checks (fast_norm_check.rs:282-462)
 *     over the witness's m*d coefficients with the instance's own challenges,
 *     `reps` repetitions (protocol.rs:222-259).
 * The batched step runs every instance in a few grouped launches (see
 * salsaa_batch_step); salsaa_batch_step_plain runs the same work as a loop of
 * the per-instance calls above, for comparison.  Both write bit-identical
 * results into separate output sets. */
typedef struct salsaa_batch salsaa_batch;

/* Device slots of one instance (layouts as for salsaa_norm_check with one
 * row; nrounds = log2(m*d)). */
typedef struct {
    uint64_t *witness;     /* [m][d] Coefficient: filled by the caller          */
    uint64_t *challenges;  /* [reps][nrounds][2]: filled by the caller          */
    uint64_t *commitment;  /* [kappa_rows][d] IncompleteNTT; rows < kappa are
                              the instance's commitment (kappa_rows = the
                              batch's largest kappa; the plain step writes
                              only rows < kappa)                               */
    uint64_t *claims;      /* [1]                                               */
    uint64_t *msgs;        /* [reps][nrounds][3][2]                             */
    uint64_t *finals;      /* [reps][2]                                         */
    uint64_t *evals;       /* [reps][2]                                         */
    uint64_t *rhs;         /* [reps][2][d]                                      */
    int32_t *ok;           /* [reps]                                            */
    size_t m, kappa, nrounds;
} salsaa_slot_t;

/* Plan a batch: allocates the witness/challenge arenas and both output sets
 * on the device, sorts the instances by length and builds the grouped
 * launches.  tune = 1 times on the device, and keeps the fastest of, the
 * commit's K chunks and schedules, the NTT pass's exchange and block count,
 * the norm check's first-pass widths and a split step's second-launch block
 * count, over synthetic short witnesses written into the new arena, which
 * is zeroed again before the call returns. */
int salsaa_batch_create(salsaa_ctx *ctx, const salsaa_key *key, size_t n, const size_t *m,
                        const size_t *kappa, size_t reps, int fq2, uint64_t nr, int tune,
                        salsaa_batch **out);
int salsaa_batch_destroy(salsaa_batch *b);
/* plain = 0: the batched step's outputs; 1: the plain per-call step's. */
int salsaa_batch_slots(salsaa_batch *b, size_t i, int plain, salsaa_slot_t *out);
/* One batched step: NTT forward and inverse over every witness (one launch
 * each), every witness into MFMA digit planes (one launch), ONE commit and
 * finalize for all instances, one norm check per length class.  use_graph = 1
 * captures the step into a hipGraph on the first call and replays it. */
int salsaa_batch_step(salsaa_batch *b, int use_graph);
/* The same work as a native loop of per-instance calls (the plain path). */
int salsaa_batch_step_plain(salsaa_batch *b);
/* One phase, eagerly (for per-phase timing): 0 NTT forward, the commit's
 * digit planes, NTT inverse and the witnesses' short copy (per slice of 8
 * ring elements, each balanced residue b as one byte when all lie in
 * [-128, 127], as two bytes when all lie in [-2^14, 2^14), else not kept)
 * in one pass, 3 commit + finalize, 7 norm checks whose first pass reads the
 * short copy where its slices are kept (valid after phase 0 over the
 * current witnesses); phases 0, 3, 7 in order equal one
 * salsaa_batch_step.  4: the norm checks from the witnesses alone.  The
 * unfused alternatives of phase 0, not part of the step: 5 NTT forward,
 * 6 NTT forward with the digit planes, 1 NTT inverse, 2 digit planes from
 * the Coefficient witnesses. */
int salsaa_batch_phase(salsaa_batch *b, int phase);
/* What the autotuner chose: commit K chunk (tiles of 64 ring elements),
 * norm check first-pass columns per thread (0 = automatic), and the timed
 * costs of the chosen commit and norm checks in ms. */
int salsaa_batch_tuning(const salsaa_batch *b, size_t *kchunk, int *cpt_first, float *commit_ms,
                        float *norm_check_ms);
/* The NTT pass's schedule the tuner chose (xch 0: the exchange between the
 * register phases of each 32-point half through LDS; 1: through DPP quad
 * permutes, the wavefront-shuffle variant) and its times of the whole pass
 * with each (milliseconds; 0 when the batch was made without tuning). */
int salsaa_batch_tuning_ntt(const salsaa_batch *b, int *xch, float *lds_ms, float *dpp_ms);
/* Sets the NTT pass's schedule (0 LDS, 1 DPP); the step is captured again on
 * its next graph replay.  Waits for the context's stream. */
int salsaa_batch_set_ntt_schedule(salsaa_batch *b, int xch);
/* The NTT pass's persistent blocks the tuner chose (blocks: the fastest of
 * 29/32, 7/8 and 13/16 of the CUs), the candidates and its times of the
 * whole pass with each (cand[3], ms[3]; zero without tuning).  Any pointer
 * may be NULL. */
int salsaa_batch_tuning_grid(const salsaa_batch *b, unsigned *blocks, unsigned *cand, float *ms);
/* The persistent blocks of a split step's second NTT launch, which runs
 * beside the first part's norm checks (blocks: the tuner's fastest whole step
 * with the pass's own blocks, 3/4 and 11/16 of the CUs), the candidates and
 * the step times with each (cand[3], ms[3]; zero without tuning or without a
 * split).  Any pointer may be NULL. */
int salsaa_batch_tuning_step_grid(const salsaa_batch *b, unsigned *blocks, unsigned *cand,
                                  float *ms);
/* Sets the NTT pass's persistent blocks (0: one per CU), for both launches of
 * a split step; the step is captured again on its next graph replay.  Waits
 * for the context's stream. */
int salsaa_batch_set_ntt_grid(salsaa_batch *b, unsigned blocks);
/* The commit schedule the tuner chose (0: one launch per class of key row
 * tiles; 1: one launch over every class, in which the blocks of all classes
 * that read one (slot, K chunk) of a stored key run side by side and share it
 * through L2; 2: one such launch over the classes of at most 8 key row tiles
 * (kappa <= 18) and the other classes' own launches) and its best times of
 * the commit with each, ms[3] (milliseconds; 0 for a schedule that does not
 * apply: a generated CRS key, or fewer than two classes to mix).  Either
 * pointer may be NULL. */
int salsaa_batch_tuning_commit(const salsaa_batch *b, int *schedule, float *ms);
/* Sets the commit schedule (1 and 2 need a stored key); the step is captured
 * again on its next graph replay.  Waits for the context's stream. */
int salsaa_batch_set_commit_schedule(salsaa_batch *b, int schedule);

/* ------------------------------------------------------------------ folding
 * The folding step of n accumulated instances (the protocol's folding round,
 * protocol.rs:351-451): instance i's witness W_i [r][m][d] (Coefficient) is
 *   decomposed into r*parts limb rows (decompose_matrix_by_chunks,
 *     arithmetic.rs:1572-1593; radix 0 = the reference's auto radix, chosen
 *     per instance),
 *   folded W'_i[j] = sum_k C_i[j][k] * limb_k with its challenge matrix C_i
 *     [r_out][r*parts] (fold, fold.rs:22-27 -> parallel_dot_matrix_matrix,
 *     arithmetic.rs:429-457),
 * then the folded witnesses, stacked as rows R = i*r_out + j, are committed
 * (parallel_dot_series_matrix, arithmetic.rs:459-505) and norm-checked row by
 * row, instance i with its own challenges cs [n][reps][log2(m d)][2]
 * (fast_norm_check.rs:282-462).  Decomposition and fold are one fused pass:
 * the limbs never reach HBM.  An instance whose C has coefficients in
 * {-1, 0, 1} (the reference's fold challenges, CyclotomicRing::random_biased,
 * cyclotomic_ring.rs:319-339) and whose radix is at most 128 folds on the int8
 * matrix cores in the Coefficient domain; any other instance in the NTT
 * domain (the same outputs).  Requires r <= 8, r*parts <= 64, r_out <= 8 and
 * m a power of two, at least 128. */
typedef struct {
    uint64_t *folded;       /* [n*r_out][m][d] IncompleteNTT (the fold output)  */
    uint64_t *folded_coeff; /* [n*r_out][m][d] the same in Coefficient form     */
    uint64_t *rhs;          /* [kappa][n*r_out][d] IncompleteNTT commitment     */
    uint64_t *radix;        /* [n] radix used per instance                      */
    uint64_t *claims;       /* [n*r_out]            norm check outputs as for   */
    uint64_t *msgs;         /* [reps][nrounds][n*r_out][3][2]  salsaa_norm_check */
    uint64_t *finals;       /* [reps][n*r_out][2]                               */
    uint64_t *evals;        /* [reps][n*r_out][2]                               */
    uint64_t *nc_rhs;       /* [reps][n*r_out][2][d] or NULL                    */
    int32_t *ok;            /* [reps][n*r_out]                                  */
} salsaa_fold_outputs;
int salsaa_fold_step(salsaa_ctx *ctx, const salsaa_key *key, size_t n, const uint64_t *d_W,
                     size_t r, size_t m, size_t parts, uint64_t radix, const uint64_t *d_C,
                     size_t r_out, const uint64_t *d_cs, size_t reps, int fq2, uint64_t nr,
                     const salsaa_fold_outputs *out);

/* ------------------------------------------------------------------ transcripts
 * The proof transcript.  The reference has none: it logs its values
 * (protocol.rs:474-482), and its verifier draws each norm-check challenge at
 * random, independent of the round's message (fast_norm_check.rs:375-385).
 *
 * Byte format (little-endian): a 96-byte header, then canonical u64 sections.
 *   0 "SLSATRN1"  8 u32 kind  12 u32 flags  16 u64 q  24 u64 nr  32 u64 L
 *  40 u32 rows  44 u32 reps  48 u32 nrounds  52 u32 d  56 u32 kappa  60 u32 0
 *  64 32-byte seed
 *   kind 1, norm check: claims[rows], msgs[reps][nrounds][rows][3][2],
 *     challenges[reps][nrounds][2], finals[reps][rows][2], evals[reps][rows][2],
 *     rhs[reps][rows][2][d] when flags bit 2
 *   kind 2, prover step: commitment[kappa][rows][d], then kind 1's sections
 *   kind 3, protocol (rows = final witness rows, reps = rounds, nrounds =
 *     statement rows R, L = final witness length m, seed = the run's u64
 *     seed): verifier rhs[R][rows][d] (IncompleteNTT), final witness
 *     [rows][m][d]
 * flags: bit 0 Fiat-Shamir challenges, bit 1 F_{q^2} (nr is 0 otherwise),
 * bit 2 rhs present, bit 3 folding scheme (kind 3), bits 16-31 the protocol
 * status (kind 3, as int16).
 * Each transcript function writes into the device buffer d_out of cap bytes
 * (stream-ordered, from device outputs) and sets *len to the transcript's
 * size; with d_out NULL it only sets *len. */

/* The norm check (norm_check_1/2/3 as salsaa_norm_check) with Fiat-Shamir
 * challenges.  All rows share each round's challenge, from a SHA-256 chain:
 *   h_0 = SHA-256("SALSAA-NC-FS-v1" || seed[32] || q (u64) || L (u64) ||
 *                 rows (u32) || rep (u32) || claims[rows] (u64))
 *   h_{t+1} = SHA-256(h_t || round t's messages of every row in row order,
 *                     each f0.a f0.b f1.a f1.b f2.a f2.b (u64))
 *   c_t = (mulhi64(u64 of h_{t+1}[0..8], q), fq2 ? mulhi64(u64 of
 *         h_{t+1}[8..16], q) : 0)
 * seed (host, 32 bytes, NULL = zeros) binds the statement, e.g. a digest of
 * the commitment.  d_cs [reps][log2 L][2] receives the challenges; the other
 * outputs are salsaa_norm_check's (all required but d_claims and d_rhs).
 * f2 is summed, not derived, so each round identity is checked.  L >= d. */
int salsaa_norm_check_fs(salsaa_ctx *ctx, const uint64_t *d_w, size_t rows, size_t L,
                         size_t reps, int fq2, uint64_t nr, const uint8_t *seed, uint64_t *d_cs,
                         uint64_t *d_claims, uint64_t *d_msgs, uint64_t *d_finals,
                         uint64_t *d_evals, uint64_t *d_rhs, int32_t *d_ok);
/* kind 1 from a norm check's outputs (fs: the challenges came from
 * salsaa_norm_check_fs with this seed); d_rhs may be NULL */
int salsaa_transcript_norm_check(salsaa_ctx *ctx, size_t rows, size_t L, size_t reps, int fq2,
                                 uint64_t nr, int fs, const uint8_t *seed,
                                 const uint64_t *d_claims, const uint64_t *d_msgs,
                                 const uint64_t *d_cs, const uint64_t *d_finals,
                                 const uint64_t *d_evals, const uint64_t *d_rhs, uint8_t *d_out,
                                 size_t cap, size_t *len);
/* kind 2: a commitment [kappa][rows][d] and the norm check of the same rows
 * (for one batch instance: its slot's commitment, kappa rows, rows = 1) */
int salsaa_transcript_prover_step(salsaa_ctx *ctx, size_t kappa, const uint64_t *d_commitment,
                                  size_t rows, size_t L, size_t reps, int fq2, uint64_t nr,
                                  int fs, const uint8_t *seed, const uint64_t *d_claims,
                                  const uint64_t *d_msgs, const uint64_t *d_cs,
                                  const uint64_t *d_finals, const uint64_t *d_evals,
                                  const uint64_t *d_rhs, uint8_t *d_out, size_t cap, size_t *len);
/* kind 3 from a salsaa_protocol run: its shape[4] output, status, verifier
 * rhs and final witness */
int salsaa_transcript_protocol(salsaa_ctx *ctx, size_t kappa, size_t nrounds, int folding,
                               uint64_t seed, int32_t status, const size_t *shape,
                               const uint64_t *d_rhs, const uint64_t *d_W, uint8_t *d_out,
                               size_t cap, size_t *len);

/* ------------------------------------------------------------------ protocol
 * protocol (protocol.rs:125-483) on the device: the CRS of the module d_module
 * [kappa][d] (an implicit key, crs.rs:23-72), the commitment of the witness
 * d_W [r][m][d] (Coefficient, INIT_WIT_REP = r rows), then nrounds rounds of
 *   decomp_ell + verify_decomp when parts[i] > 1       (protocol.rs:166-196)
 *   norm_check_1/2/3 + their verifiers, 2 repetitions  (:198-259)
 *   split + verifier_split                             (:261-291)
 *   challenge_for_project, project, batch_projections,
 *     verify_batching                                  (:293-348)
 *   challenge_for_fold, fold, verifier_fold (WIT_REP 1) (:349-386)
 *   send_cross_terms_before_join, verifier_join, witness.append (:387-410)
 * and the final check parallel_dot_series_matrix(statement, witness) ==
 * verifier rhs (:464-472).  folding = 1 runs the folding scheme instead (no
 * decomposition at the round's start, no split, and after the join
 * decomp_ell(.., 2) + verify_decomp, then stop; parts may be NULL).  The
 * statement grows by norm_check_3's two power series, witness_lhs and
 * projected_lhs per round, and every commitment runs against all its rows.
 * Challenges are drawn on the device from `seed` (the reference's are
 * unseeded).  Outputs: *status (host) = 0 when every verifier relation held,
 * else 100 * round + check of the first that failed (check 1 verify_decomp,
 * 2 norm check, 3 verifier_split, 4 verify_batching, 5 the folding scheme's
 * verify_decomp, 6 the final check); shape (host) [4] = witness rows and
 * length, statement rows R and verifier columns; d_W_out (capacity w_cap u64)
 * the final witness [r][m][d] and d_rhs_out (capacity rhs_cap) the verifier
 * rhs [R][r][d] IncompleteNTT, either may be NULL; times (host, may be NULL)
 * [8] milliseconds per phase: commitment, decomp, norm check, split,
 * projection (with batching), fold, join, final check.  Needs kappa <= 36,
 * m and r powers of two and 256 r <= m at every projection (project.rs
 * panics otherwise). */
int salsaa_protocol(salsaa_ctx *ctx, const uint64_t *d_module, size_t kappa, size_t m,
                    const uint64_t *d_W, size_t r, size_t nrounds, const size_t *parts,
                    int folding, uint64_t seed, uint64_t *d_W_out, size_t w_cap,
                    uint64_t *d_rhs_out, size_t rhs_cap, size_t *shape, int32_t *status,
                    double *times);

/* salsaa_protocol with options (opts may be NULL: the same as
 * salsaa_protocol).  fault_check 1..6 injects a fault, for tests of the
 * verifier checks: one prover value that check reads is off by one (mod q) in
 * round fault_round (check 6, the final check, ignores the round): 1 and 5 the
 * commitment of the decomposed witness, 2 row 0's norm_check_1 claim (needs
 * m * d > 1024 at that round), 3 the split commitments, 4 the projection
 * after its commitment, 6 the final witness.  The run then reports that
 * check's code in *status.  0 runs the protocol as it is. */
typedef struct {
    size_t fault_round;
    int fault_check;
} salsaa_protocol_opts;
int salsaa_protocol_ex(salsaa_ctx *ctx, const uint64_t *d_module, size_t kappa, size_t m,
                       const uint64_t *d_W, size_t r, size_t nrounds, const size_t *parts,
                       int folding, uint64_t seed, uint64_t *d_W_out, size_t w_cap,
                       uint64_t *d_rhs_out, size_t rhs_cap, size_t *shape, int32_t *status,
                       double *times, const salsaa_protocol_opts *opts);

/* ------------------------------------------------------------------ serving
 * A continuous-batching service of commit + norm-check jobs against one
 * resident key (no reference counterpart: it schedules the per-job calls
 * parallel_dot_series_matrix, arithmetic.rs:459-505, and norm_check_1/2/3,
 * fast_norm_check.rs:282-462, for a stream of requests).  A scheduler thread
 * admits jobs in arrival order while their witness fits the HBM arena, forms
 * a batch from every admitted job whenever one of `inflight` batch slots is
 * free (jobs arriving meanwhile join the next batch, and two slots overlap on
 * the device), and copies each finished job's output record to the host.
 * max_batch_jobs = 1 with inflight = 1 serves every job alone: the plain
 * per-call path. */
typedef struct salsaa_server salsaa_server;
typedef struct {
    size_t arena_elems;     /* HBM arena for admitted witnesses (ring elements) */
    size_t max_batch_jobs;  /* jobs per batch (1..4096)                         */
    size_t max_batch_elems; /* ring elements per batch (>= the largest job)     */
    int inflight;           /* batches on the device at once: 1 or 2            */
    size_t reps;            /* norm check repetitions                           */
    int fq2;                /* 1: F_{q^2} challenges                            */
    uint64_t nr;            /* non-residue of F_{q^2}                           */
    size_t kchunk;          /* commit K chunk in tiles of 64 ring elements      */
} salsaa_server_opts;

/* Jobs commit with the key's first kappa rows. */
int salsaa_server_create(salsaa_ctx *ctx, const salsaa_key *key, size_t kappa,
                         const salsaa_server_opts *opts, salsaa_server **out);
/* Queue a job with a witness of m ring elements (power of two): copied from
 * h_witness ([m][d] Coefficient, host memory that stays valid until the job
 * is done; pinned memory copies asynchronously) or, if h_witness is NULL,
 * sampled on the device in [-bound, bound] from `seed` (stream 0xB).  The
 * job's challenges are sampled from `seed` (as salsaa_sample_challenges). */
int salsaa_server_submit(salsaa_server *s, size_t m, const uint64_t *h_witness, uint64_t seed,
                         uint64_t bound, uint64_t *id);
/* 1 when job `id` is done (its output record copied to h_record, times =
 * submit, admit, launch, done in seconds of a steady clock, batch_info =
 * batch number and jobs in it), 0 while pending, negative on error. */
int salsaa_server_poll(salsaa_server *s, uint64_t id, uint64_t *h_record, double *times,
                       uint64_t *batch_info);
/* Block until every submitted job is done. */
int salsaa_server_drain(salsaa_server *s);
/* Output record layout (u64 offsets): [0] commitment [kappa][d] IncompleteNTT,
 * [1] claim, [2] msgs [reps][nrounds][3][2], [3] finals [reps][2], [4] evals
 * [reps][2], [5] rhs [reps][2][d], [6] ok as int32 [reps], [7] record length. */
int salsaa_server_layout(const salsaa_server *s, size_t *offsets);
int salsaa_server_destroy(salsaa_server *s);

#ifdef __cplusplus
}
#endif
#endif
