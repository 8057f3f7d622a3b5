/*
 * salsaa_oracle.c -- CPU restatement of the SALSAA prover hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check and the
 * benchmark's CPU-baseline leg load this library; the product never does.
 *
 * Written for obvious correctness: scalar loops, Barrett reduction with
 * __int128 (checked against __int128 % q), one function per reference routine.
 * Citations are file:line in the reference repository.
 */
#include "salsaa_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* scalar arithmetic                                                         */
/* ------------------------------------------------------------------------ */

/* Barrett reduction (HAC 14.42) with k = bitlen(q), mu = floor(4^k / q). */
typedef struct {
    uint64_t q;
    unsigned k;
    u128 mu;
} barrett_t;

static __thread barrett_t tl_barrett = {0, 0, 0};

static inline const barrett_t *barrett_for(uint64_t q) {
    if (tl_barrett.q != q) {
        unsigned k = 64u - (unsigned)__builtin_clzll(q);
        tl_barrett.q = q;
        tl_barrett.k = k;
        tl_barrett.mu = (((u128)1) << (2 * k)) / q;
    }
    return &tl_barrett;
}

static inline uint64_t barrett_reduce(u128 x, const barrett_t *b) {
    u128 q1 = x >> (b->k - 1);
    u128 q3 = (q1 * b->mu) >> (b->k + 1);
    u128 r = x - q3 * (u128)b->q;
    while (r >= b->q) r -= b->q;
    return (uint64_t)r;
}

uint64_t or_mulmod_slow(uint64_t a, uint64_t b, uint64_t q) {
    return (uint64_t)(((u128)a * b) % q);
}

/* hexl_wrapper.cpp:34-36 (MultiplyMod): a*b mod q for a, b < q < 2^62 */
uint64_t or_mulmod(uint64_t a, uint64_t b, uint64_t q) {
    if (q > (1ULL << 62)) return or_mulmod_slow(a, b, q);
    return barrett_reduce((u128)a * b, barrett_for(q));
}

/* hexl_wrapper.cpp:38-41 (AddUIntMod), inputs < q */
uint64_t or_addmod(uint64_t a, uint64_t b, uint64_t q) {
    uint64_t s = a + b;
    return (s >= q || s < a) ? s - q : s;
}

/* hexl_wrapper.cpp:43-45 (SubUIntMod), inputs < q */
uint64_t or_submod(uint64_t a, uint64_t b, uint64_t q) {
    return a >= b ? a - b : a + (q - b);
}

/* hexl_wrapper.cpp:61-63 (PowMod) */
uint64_t or_powmod(uint64_t a, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) r = or_mulmod(r, a, q);
        a = or_mulmod(a, a, q);
        e >>= 1;
    }
    return r;
}

/* hexl_wrapper.cpp:52-54 (InverseMod), q prime: a^(q-2) */
uint64_t or_invmod(uint64_t a, uint64_t q) { return or_powmod(a, q - 2, q); }

/* ------------------------------------------------------------------------ */
/* NTT tables, HEXL convention                                               */
/* ------------------------------------------------------------------------ */

static size_t ilog2(size_t n) {
    size_t l = 0;
    while (((size_t)1 << l) < n) l++;
    return l;
}

static size_t bitrev(size_t x, size_t bits) {
    size_t r = 0;
    for (size_t i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

/* Smallest element of multiplicative order exactly `order` (a power of two)
 * -- HEXL's MinimalPrimitiveRoot(2n, q), the root its NTT(n, q) uses. */
uint64_t or_min_primitive_root(uint64_t order, uint64_t q) {
    uint64_t w = 0;
    for (uint64_t g = 2; g < q; g++) {
        uint64_t c = or_powmod(g, (q - 1) / order, q);
        if (or_powmod(c, order / 2, q) == q - 1) { w = c; break; }
    }
    if (!w) return 0;
    uint64_t w2 = or_mulmod(w, w, q), cur = w, best = w;
    for (uint64_t i = 0; i < order / 2; i++) {
        if (cur < best) best = cur;
        cur = or_mulmod(cur, w2, q);
    }
    return best;
}

typedef struct {
    size_t n;
    uint64_t q;
    uint64_t *roots;     /* roots[brv(i)] = psi^i                     */
    uint64_t *inv_roots; /* inv_roots[k] = roots[k]^{-1} (same index) */
    uint64_t n_inv;
} ntt_table_t;

#define MAX_TABLES 64
static ntt_table_t g_tables[MAX_TABLES];
static int g_ntables = 0;
static pthread_mutex_t g_table_lock = PTHREAD_MUTEX_INITIALIZER;

static const ntt_table_t *table_for(size_t n, uint64_t q) {
    pthread_mutex_lock(&g_table_lock);
    for (int i = 0; i < g_ntables; i++)
        if (g_tables[i].n == n && g_tables[i].q == q) {
            pthread_mutex_unlock(&g_table_lock);
            return &g_tables[i];
        }
    ntt_table_t *t = &g_tables[g_ntables < MAX_TABLES ? g_ntables++ : MAX_TABLES - 1];
    size_t bits = ilog2(n);
    uint64_t psi = or_min_primitive_root(2 * n, q);
    t->n = n;
    t->q = q;
    t->roots = (uint64_t *)malloc(n * sizeof(uint64_t));
    t->inv_roots = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t p = 1;
    for (size_t i = 0; i < n; i++) {
        t->roots[bitrev(i, bits)] = p;
        p = or_mulmod(p, psi, q);
    }
    for (size_t i = 0; i < n; i++) t->inv_roots[i] = or_invmod(t->roots[i], q);
    t->n_inv = or_invmod(n % q, q);
    pthread_mutex_unlock(&g_table_lock);
    return t;
}

void or_get_roots(size_t n, uint64_t q, uint64_t *roots) {
    memcpy(roots, table_for(n, q)->roots, n * sizeof(uint64_t));
}

/* HEXL stores the inverse powers re-ordered in the order its inverse
 * transform walks them: [inv(root[0]), then for m = n/2, n/4, ..., 1 the
 * entries inv(root[m .. 2m))].  Parity unpinned. */
void or_get_inv_roots(size_t n, uint64_t q, uint64_t *inv_roots) {
    const ntt_table_t *t = table_for(n, q);
    size_t idx = 0;
    inv_roots[idx++] = t->inv_roots[0];
    for (size_t m = n >> 1; m > 0; m >>= 1)
        for (size_t i = 0; i < m; i++) inv_roots[idx++] = t->inv_roots[m + i];
}

/* hexl_wrapper.cpp:213-220: negacyclic forward NTT, natural-order input,
 * bit-reversed output, outputs fully reduced (Cooley-Tukey, psi powers in
 * bit-reversed order). */
void or_ntt_forward(uint64_t *x, size_t n, uint64_t q) {
    if (n < 2) return;
    const ntt_table_t *tb = table_for(n, q);
    size_t t = n;
    for (size_t m = 1; m < n; m <<= 1) {
        t >>= 1;
        for (size_t i = 0; i < m; i++) {
            uint64_t w = tb->roots[m + i];
            size_t j1 = 2 * i * t;
            for (size_t j = j1; j < j1 + t; j++) {
                uint64_t X = x[j];
                uint64_t Y = or_mulmod(x[j + t], w, q);
                x[j] = or_addmod(X, Y, q);
                x[j + t] = or_submod(X, Y, q);
            }
        }
    }
}

/* hexl_wrapper.cpp:222-229: inverse of or_ntt_forward (Gentleman-Sande,
 * bit-reversed input, natural-order output, includes n^{-1}). */
void or_ntt_inverse(uint64_t *x, size_t n, uint64_t q) {
    if (n < 2) return;
    const ntt_table_t *tb = table_for(n, q);
    size_t t = 1;
    for (size_t m = n >> 1; m >= 1; m >>= 1) {
        for (size_t i = 0; i < m; i++) {
            uint64_t wi = tb->inv_roots[m + i];
            size_t j1 = 2 * i * t;
            for (size_t j = j1; j < j1 + t; j++) {
                uint64_t X = x[j], Y = x[j + t];
                x[j] = or_addmod(X, Y, q);
                x[j + t] = or_mulmod(or_submod(X, Y, q), wi, q);
            }
        }
        t <<= 1;
    }
    for (size_t j = 0; j < n; j++) x[j] = or_mulmod(x[j], tb->n_inv, q);
}

/* Closed form of HEXL's forward transform:
 * out[j] = sum_k x[k] * psi^{(2*brv(j)+1)*k}, psi = minimal primitive 2n-th root. */
void or_ntt_forward_direct(const uint64_t *x, uint64_t *out, size_t n, uint64_t q) {
    uint64_t psi = or_min_primitive_root(2 * n, q);
    size_t bits = ilog2(n);
    for (size_t j = 0; j < n; j++) {
        uint64_t e = or_powmod(psi, 2 * bitrev(j, bits) + 1, q);
        uint64_t acc = 0, p = 1;
        for (size_t k = 0; k < n; k++) {
            acc = or_addmod(acc, or_mulmod(x[k], p, q), q);
            p = or_mulmod(p, e, q);
        }
        out[j] = acc;
    }
}

/* ------------------------------------------------------------------------ */
/* representations                                                           */
/* ------------------------------------------------------------------------ */

/* cyclotomic_ring.rs:483-511: even/odd split, NTT_{d/2} of each half */
void or_to_incomplete(uint64_t *a, size_t d, uint64_t q) {
    size_t h = d / 2;
    uint64_t *tmp = (uint64_t *)malloc(d * sizeof(uint64_t));
    for (size_t i = 0; i < h; i++) {
        tmp[i] = a[2 * i];
        tmp[h + i] = a[2 * i + 1];
    }
    or_ntt_forward(tmp, h, q);
    or_ntt_forward(tmp + h, h, q);
    memcpy(a, tmp, d * sizeof(uint64_t));
    free(tmp);
}

/* cyclotomic_ring.rs:527-549: INTT of both halves, re-interleave */
void or_from_incomplete(uint64_t *a, size_t d, uint64_t q) {
    size_t h = d / 2;
    uint64_t *tmp = (uint64_t *)malloc(d * sizeof(uint64_t));
    memcpy(tmp, a, d * sizeof(uint64_t));
    or_ntt_inverse(tmp, h, q);
    or_ntt_inverse(tmp + h, h, q);
    for (size_t i = 0; i < h; i++) {
        a[2 * i] = tmp[i];
        a[2 * i + 1] = tmp[h + i];
    }
    free(tmp);
}

/* cyclotomic_ring.rs:513-525 */
void or_to_full(uint64_t *a, size_t d, uint64_t q) { or_ntt_forward(a, d, q); }
/* cyclotomic_ring.rs:551-553 */
void or_from_full(uint64_t *a, size_t d, uint64_t q) { or_ntt_inverse(a, d, q); }

/* cyclotomic_ring.rs:559-564: omega = NTT_{d/2}(Y) */
void or_shift_factors(size_t d, uint64_t q, uint64_t *omega) {
    size_t h = d / 2;
    memset(omega, 0, h * sizeof(uint64_t));
    if (h > 1) omega[1] = 1;
    else omega[0] = q - 1; /* d = 2: X^2 = -1 */
    or_ntt_forward(omega, h, q);
}

/* ------------------------------------------------------------------------ */
/* ring products                                                             */
/* ------------------------------------------------------------------------ */

/* Per-slot product in IncompleteNTT form (cyclotomic_ring.rs:717-787):
 * re = ae*be + omega*ao*bo, ro = ao*be + ae*bo. */
static void slot_mul(const uint64_t *a, const uint64_t *b, uint64_t *out,
                     const uint64_t *omega, size_t d, uint64_t q) {
    size_t h = d / 2;
    for (size_t j = 0; j < h; j++) {
        uint64_t ae = a[j], ao = a[h + j], be = b[j], bo = b[h + j];
        uint64_t ee = or_mulmod(ae, be, q);
        uint64_t oe = or_mulmod(ao, be, q);
        uint64_t t = or_mulmod(or_mulmod(ao, bo, q), omega[j], q);
        uint64_t eo = or_mulmod(ae, bo, q);
        out[j] = or_addmod(ee, t, q);
        out[h + j] = or_addmod(oe, eo, q);
    }
}

/* cyclotomic_ring.rs:699-791 with both operands in Coefficient form */
void or_ring_mul_incomplete(const uint64_t *a, const uint64_t *b, uint64_t *out,
                            size_t d, uint64_t q) {
    uint64_t *ta = (uint64_t *)malloc(d * sizeof(uint64_t));
    uint64_t *tb = (uint64_t *)malloc(d * sizeof(uint64_t));
    uint64_t *om = (uint64_t *)malloc((d / 2) * sizeof(uint64_t));
    memcpy(ta, a, d * sizeof(uint64_t));
    memcpy(tb, b, d * sizeof(uint64_t));
    or_to_incomplete(ta, d, q);
    or_to_incomplete(tb, d, q);
    or_shift_factors(d, q, om);
    slot_mul(ta, tb, out, om, d, q);
    free(ta);
    free(tb);
    free(om);
}

/* cyclotomic_ring.rs:816-839: schoolbook product mod X^d + 1 */
void or_ring_mul_naive(const uint64_t *a, const uint64_t *b, uint64_t *out,
                       size_t d, uint64_t q) {
    uint64_t *acc = (uint64_t *)calloc(d, sizeof(uint64_t));
    for (size_t i = 0; i < d; i++)
        for (size_t j = 0; j < d; j++) {
            uint64_t p = or_mulmod(a[i], b[j], q);
            if (i + j < d) acc[i + j] = or_addmod(acc[i + j], p, q);
            else acc[i + j - d] = or_submod(acc[i + j - d], p, q);
        }
    memcpy(out, acc, d * sizeof(uint64_t));
    free(acc);
}

/* cyclotomic_ring.rs:455-469: a(X^{-1}) */
void or_conjugate(const uint64_t *a, uint64_t *out, size_t d, uint64_t q) {
    uint64_t a0 = a[0];
    uint64_t *tmp = (uint64_t *)malloc(d * sizeof(uint64_t));
    tmp[0] = a0;
    for (size_t i = 1; i < d; i++) tmp[i] = a[d - i] == 0 ? 0 : q - a[d - i];
    memcpy(out, tmp, d * sizeof(uint64_t));
    free(tmp);
}

/* ------------------------------------------------------------------------ */
/* commit                                                                    */
/* ------------------------------------------------------------------------ */

/* arithmetic.rs:459-505 (parallel_dot_series_matrix) with a uniform key whose
 * only layer has length m.  Like the reference, every product converts both
 * Coefficient operands to IncompleteNTT (cyclotomic_ring.rs:193-198, 707-708)
 * and the accumulator starts as a Coefficient zero that adopts the product's
 * representation (cyclotomic_ring.rs:43-63). */
void or_commit(const uint64_t *A, const uint64_t *W, uint64_t *out, size_t kappa,
               size_t r, size_t m, size_t d, uint64_t q, int threads) {
    long total = (long)(kappa * r);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) if (threads != 1)
#endif
    for (long jk = 0; jk < total; jk++) {
        size_t j = (size_t)jk / r, k = (size_t)jk % r;
        uint64_t *acc = out + ((size_t)j * r + k) * d;
        uint64_t *prod = (uint64_t *)malloc(d * sizeof(uint64_t));
        memset(acc, 0, d * sizeof(uint64_t));
        for (size_t i = 0; i < m; i++) {
            or_ring_mul_incomplete(A + (j * m + i) * d, W + (k * m + i) * d, prod, d, q);
            for (size_t c = 0; c < d; c++) acc[c] = or_addmod(acc[c], prod[c], q);
        }
        free(prod);
    }
}

/* ------------------------------------------------------------------------ */
/* power series                                                              */
/* ------------------------------------------------------------------------ */

/* complete_power_series (arithmetic.rs:158-176) from kronecker_product
 * (arithmetic.rs:845-864: out[i |b| + j] = a[i] * b[j]), applying the tensors
 * last first as the reference does.  With length-2 tensors t_0 .. t_{n-1}
 * (IncompletePowerSeries order) the largest expanded layer is
 *   out[((i_0 2 + i_1) 2 + ... + i_{n-1}) m0 + j] = t_0[i_0] ... t_{n-1}[i_{n-1}] base[j],
 * the first tensor the most significant index bit.  Ring products are the
 * schoolbook product mod X^d + 1 (cyclotomic_ring.rs:816-839), everything in
 * Coefficient form.  base [m0][d], tensors [n][2][d], out [m0 << n][d]. */
void or_complete_power_series(const uint64_t *base, size_t m0, const uint64_t *tensors, size_t n,
                              size_t d, uint64_t q, uint64_t *out) {
    const size_t full = m0 << n;
    uint64_t *cur = (uint64_t *)malloc(full * d * sizeof(uint64_t));
    uint64_t *nxt = (uint64_t *)malloc(full * d * sizeof(uint64_t));
    memcpy(cur, base, m0 * d * sizeof(uint64_t));
    size_t len = m0;
    for (size_t k = n; k-- > 0;) {
        const uint64_t *t = tensors + k * 2 * d;
        for (size_t a = 0; a < 2; a++)
            for (size_t j = 0; j < len; j++)
                or_ring_mul_naive(t + a * d, cur + j * d, nxt + (a * len + j) * d, d, q);
        uint64_t *tmp = cur;
        cur = nxt;
        nxt = tmp;
        len *= 2;
    }
    memcpy(out, cur, full * d * sizeof(uint64_t));
    free(cur);
    free(nxt);
}

/* parallel_dot_single_series_matrix (arithmetic.rs:507-536): out[j] =
 * sum_i layer[i] * W[j][i] with the expanded layer of length m.  The
 * reference's products leave IncompleteNTT form (cyclotomic_ring.rs:193-198)
 * and the sum adopts it; here the products are schoolbook ones summed in
 * Coefficient form and converted once (the conversion is linear), a route
 * independent of or_commit.  layer [m][d], W [r][m][d] Coefficient,
 * out [r][d] IncompleteNTT. */
void or_dot_single_series(const uint64_t *layer, const uint64_t *W, uint64_t *out, size_t r,
                          size_t m, size_t d, uint64_t q) {
    uint64_t *prod = (uint64_t *)malloc(d * sizeof(uint64_t));
    for (size_t j = 0; j < r; j++) {
        uint64_t *acc = out + j * d;
        memset(acc, 0, d * sizeof(uint64_t));
        for (size_t i = 0; i < m; i++) {
            or_ring_mul_naive(layer + i * d, W + (j * m + i) * d, prod, d, q);
            for (size_t c = 0; c < d; c++) acc[c] = or_addmod(acc[c], prod[c], q);
        }
        or_to_incomplete(acc, d, q);
    }
    free(prod);
}

/* ------------------------------------------------------------------------ */
/* projection                                                                */
/* ------------------------------------------------------------------------ */

/* project (project.rs:52-98) without its closing commit: with cw = r * 256,
 * row i's chunks j < m / cw (row-major, i outer) each give 256 ring elements
 *   out[(i (m / cw) + j) 256 + k] = pos - neg mod q,
 * pos / neg the u64 sums (add_no_reduction, arithmetic.rs:19-44) of the chunk's
 * elements W[i][j cw + l] with is_zero[k][l] = 0 and sign[k][l] = 1 / 0, each
 * reduced mod q (eltwise_reduce_mod) before the subtraction.  is_zero, sign
 * [256][cw] bytes; W [r][m][d]; out [r (m / cw) 256][d], the data taken as
 * they are (every representation is linear). */
void or_project(const uint64_t *W, size_t r, size_t m, size_t d, const uint8_t *is_zero,
                const uint8_t *sign, uint64_t q, uint64_t *out) {
    const size_t H = 256, cw = r * H, nch = m / cw;
    uint64_t *pos = (uint64_t *)malloc(d * sizeof(uint64_t));
    uint64_t *neg = (uint64_t *)malloc(d * sizeof(uint64_t));
    for (size_t i = 0; i < r; i++)
        for (size_t j = 0; j < nch; j++)
            for (size_t k = 0; k < H; k++) {
                memset(pos, 0, d * sizeof(uint64_t));
                memset(neg, 0, d * sizeof(uint64_t));
                for (size_t l = 0; l < cw; l++) {
                    if (is_zero[k * cw + l]) continue;
                    const uint64_t *x = W + (i * m + j * cw + l) * d;
                    uint64_t *acc = sign[k * cw + l] ? pos : neg;
                    for (size_t c = 0; c < d; c++) acc[c] += x[c];
                }
                uint64_t *o = out + ((i * nch + j) * H + k) * d;
                for (size_t c = 0; c < d; c++) o[c] = or_submod(pos[c] % q, neg[c] % q, q);
            }
    free(pos);
    free(neg);
}

/* ------------------------------------------------------------------------ */
/* norm check                                                                */
/* ------------------------------------------------------------------------ */

/* fast_norm_check.rs:25-39: first challenge is the most significant bit */
void or_tensor(const uint64_t *cs, size_t n, uint64_t q, uint64_t *out) {
    size_t len = 1;
    out[0] = 1 % q;
    for (size_t t = 0; t < n; t++) {
        uint64_t c = cs[t] % q;
        uint64_t one_minus = (uint64_t)(((u128)q + 1 - c) % q);
        for (size_t v = len; v-- > 0;) {
            uint64_t val = out[v];
            out[2 * v] = or_mulmod(val, one_minus, q);
            out[2 * v + 1] = or_mulmod(val, c, q);
        }
        len *= 2;
    }
}

static size_t chunk_size(size_t len) {
    size_t c = len / 120; /* MAX_THREADS, arithmetic.rs:15 */
    return c > 1028 ? c : 1028; /* MIN_CHUNK_SIZE, fast_norm_check.rs:23 */
}

/* fast_norm_check.rs:120-126 + hexl_wrapper.cpp:231-237: unreduced u64 sum of
 * per-chunk residues */
uint64_t or_sum_chunked_raw(const uint64_t *x, size_t len, uint64_t q) {
    size_t cs = chunk_size(len);
    uint64_t total = 0;
    for (size_t s = 0; s < len; s += cs) {
        uint64_t acc = 0;
        size_t e = s + cs < len ? s + cs : len;
        for (size_t i = s; i < e; i++) acc = or_addmod(acc, x[i], q);
        total += acc;
    }
    return total;
}

/* fast_norm_check.rs:127-133 + hexl_wrapper.cpp:262-268 */
uint64_t or_sum_sq_chunked_raw(const uint64_t *x, size_t len, uint64_t q) {
    size_t cs = chunk_size(len);
    uint64_t total = 0;
    for (size_t s = 0; s < len; s += cs) {
        uint64_t acc = 0;
        size_t e = s + cs < len ? s + cs : len;
        for (size_t i = s; i < e; i++) acc = or_addmod(acc, or_mulmod(x[i], x[i], q), q);
        total += acc;
    }
    return total;
}

/* F_{q^2} = F_q[Y]/(Y^2 - nr) (quadratic_extension.rs:13-58) */
typedef struct {
    uint64_t a, b;
} fq2_t;

static inline fq2_t f2_add(fq2_t x, fq2_t y, uint64_t q) {
    fq2_t r = {or_addmod(x.a, y.a, q), or_addmod(x.b, y.b, q)};
    return r;
}
static inline fq2_t f2_sub(fq2_t x, fq2_t y, uint64_t q) {
    fq2_t r = {or_submod(x.a, y.a, q), or_submod(x.b, y.b, q)};
    return r;
}
static inline fq2_t f2_mul(fq2_t x, fq2_t y, uint64_t nr, uint64_t q) {
    fq2_t r = {or_addmod(or_mulmod(x.a, y.a, q),
                         or_mulmod(nr, or_mulmod(x.b, y.b, q), q), q),
               or_addmod(or_mulmod(x.a, y.b, q), or_mulmod(x.b, y.a, q), q)};
    return r;
}
static inline int f2_eq(fq2_t x, fq2_t y) { return x.a == y.a && x.b == y.b; }

/* quadratic_extension.rs:30-58 with shift = nr */
void or_fq2_mul(uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1, uint64_t nr,
                uint64_t q, uint64_t *out) {
    fq2_t x = {a0, a1}, y = {b0, b1};
    fq2_t r = f2_mul(x, y, nr, q);
    out[0] = r.a;
    out[1] = r.b;
}

/* Sumcheck of sum_x w(x)^2 (fast_norm_check.rs:135-199, verifier 352-399):
 * per round, halves w0|w1 (most significant variable first), delta = w1 - w0,
 * f0 = sum w0^2, f1 = sum w0*delta, f2 = sum delta^2; the verifier checks
 * 2 f0 + 2 f1 + f2 == claim, draws c, sets claim = f0 + 2 f1 c + f2 c^2, and
 * the prover folds w <- delta*c + w0.  Each repetition restarts from the
 * norm_check_1 claim with its own challenges (protocol.rs:222-259). */
static void sumcheck_row(const uint64_t *w, size_t L, const uint64_t *cs, size_t reps,
                         int fq2, uint64_t nr, uint64_t q, size_t rows, size_t row,
                         uint64_t claim, uint64_t *msgs, uint64_t *finals,
                         uint64_t *evals, int *ok) {
    size_t nrounds = ilog2(L);
    fq2_t *cur = (fq2_t *)malloc(L * sizeof(fq2_t));
    for (size_t rep = 0; rep < reps; rep++) {
        for (size_t x = 0; x < L; x++) {
            cur[x].a = w[x];
            cur[x].b = 0;
        }
        fq2_t cl = {claim, 0};
        int good = 1;
        size_t len = L;
        for (size_t t = 0; t < nrounds; t++) {
            size_t half = len / 2;
            fq2_t f0 = {0, 0}, f1 = {0, 0}, f2 = {0, 0};
            for (size_t x = 0; x < half; x++) {
                fq2_t w0 = cur[x], dl = f2_sub(cur[x + half], cur[x], q);
                f0 = f2_add(f0, f2_mul(w0, w0, nr, q), q);
                f1 = f2_add(f1, f2_mul(w0, dl, nr, q), q);
                f2 = f2_add(f2, f2_mul(dl, dl, nr, q), q);
            }
            uint64_t *mo = msgs + (((rep * nrounds + t) * rows + row) * 3) * 2;
            mo[0] = f0.a; mo[1] = f0.b;
            mo[2] = f1.a; mo[3] = f1.b;
            mo[4] = f2.a; mo[5] = f2.b;
            fq2_t lhs = f2_add(f2_add(f2_add(f0, f0, q), f2_add(f1, f1, q), q), f2, q);
            if (!f2_eq(lhs, cl)) good = 0;
            const uint64_t *cp = cs + (rep * nrounds + t) * 2;
            fq2_t c = {cp[0] % q, fq2 ? cp[1] % q : 0};
            fq2_t c2 = f2_mul(c, c, nr, q);
            cl = f2_add(f2_add(f0, f2_mul(f2_add(f1, f1, q), c, nr, q), q),
                        f2_mul(f2, c2, nr, q), q);
            for (size_t x = 0; x < half; x++) {
                fq2_t dl = f2_sub(cur[x + half], cur[x], q);
                cur[x] = f2_add(f2_mul(dl, c, nr, q), cur[x], q);
            }
            len = half;
        }
        fq2_t ev = cur[0];
        if (!f2_eq(f2_mul(ev, ev, nr, q), cl)) good = 0;
        finals[(rep * rows + row) * 2] = cl.a;
        finals[(rep * rows + row) * 2 + 1] = cl.b;
        evals[(rep * rows + row) * 2] = ev.a;
        evals[(rep * rows + row) * 2 + 1] = ev.b;
        ok[rep * rows + row] = good;
    }
    free(cur);
}

void or_sumcheck(const uint64_t *w, size_t rows, size_t L, const uint64_t *cs,
                 size_t reps, int fq2, uint64_t nr, uint64_t q, uint64_t *claims,
                 uint64_t *msgs, uint64_t *finals, uint64_t *evals, int *ok,
                 int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) if (threads != 1)
#endif
    for (long row = 0; row < (long)rows; row++) {
        const uint64_t *wr = w + (size_t)row * L;
        /* norm_check_1 (fast_norm_check.rs:282-320), canonical residue */
        uint64_t s = 0;
        for (size_t x = 0; x < L; x++) s = or_addmod(s, or_mulmod(wr[x], wr[x], q), q);
        claims[row] = s;
        sumcheck_row(wr, L, cs, reps, fq2, nr, q, rows, (size_t)row, s, msgs, finals,
                     evals, ok);
    }
}

/* The sumcheck with every row in step (the reference's rows share each
 * round's challenge, fast_norm_check.rs:352-399): per repetition and round,
 * the messages f0, f1, f2 of every row (all three summed directly), then
 * chal(user, rep, t, msgs of the round [rows][6], c[2]) supplies the round's
 * challenge (the Fiat-Shamir replay derives it from the messages), which
 * goes to cs_out [reps][nrounds][2]; then every row folds.  ok is the
 * verifier's: each round identity and the final eval^2 == claim. */
void or_sumcheck_cb(const uint64_t *w, size_t rows, size_t L, size_t reps, int fq2, uint64_t nr,
                    uint64_t q, uint64_t *claims, uint64_t *msgs, uint64_t *cs_out,
                    uint64_t *finals, uint64_t *evals, int *ok, or_chal_fn chal, void *user) {
    size_t nrounds = ilog2(L);
    fq2_t *cur = (fq2_t *)malloc(rows * L * sizeof(fq2_t));
    fq2_t *cl = (fq2_t *)malloc(rows * sizeof(fq2_t));
    for (size_t row = 0; row < rows; row++) {
        uint64_t s = 0;
        for (size_t x = 0; x < L; x++) s = or_addmod(s, or_mulmod(w[row * L + x], w[row * L + x], q), q);
        claims[row] = s;
    }
    for (size_t rep = 0; rep < reps; rep++) {
        for (size_t row = 0; row < rows; row++) {
            for (size_t x = 0; x < L; x++) {
                cur[row * L + x].a = w[row * L + x];
                cur[row * L + x].b = 0;
            }
            cl[row].a = claims[row];
            cl[row].b = 0;
            ok[rep * rows + row] = 1;
        }
        size_t len = L;
        for (size_t t = 0; t < nrounds; t++) {
            size_t half = len / 2;
            uint64_t *mo = msgs + ((rep * nrounds + t) * rows) * 6;
            for (size_t row = 0; row < rows; row++) {
                fq2_t *c = cur + row * L;
                fq2_t f0 = {0, 0}, f1 = {0, 0}, f2 = {0, 0};
                for (size_t x = 0; x < half; x++) {
                    fq2_t w0 = c[x], dl = f2_sub(c[x + half], c[x], q);
                    f0 = f2_add(f0, f2_mul(w0, w0, nr, q), q);
                    f1 = f2_add(f1, f2_mul(w0, dl, nr, q), q);
                    f2 = f2_add(f2, f2_mul(dl, dl, nr, q), q);
                }
                mo[row * 6 + 0] = f0.a; mo[row * 6 + 1] = f0.b;
                mo[row * 6 + 2] = f1.a; mo[row * 6 + 3] = f1.b;
                mo[row * 6 + 4] = f2.a; mo[row * 6 + 5] = f2.b;
                fq2_t lhs = f2_add(f2_add(f2_add(f0, f0, q), f2_add(f1, f1, q), q), f2, q);
                if (!f2_eq(lhs, cl[row])) ok[rep * rows + row] = 0;
            }
            uint64_t cc[2] = {0, 0};
            chal(user, rep, t, mo, cc);
            fq2_t ch = {cc[0] % q, fq2 ? cc[1] % q : 0};
            cs_out[(rep * nrounds + t) * 2] = ch.a;
            cs_out[(rep * nrounds + t) * 2 + 1] = ch.b;
            fq2_t c2 = f2_mul(ch, ch, nr, q);
            for (size_t row = 0; row < rows; row++) {
                fq2_t f0 = {mo[row * 6], mo[row * 6 + 1]}, f1 = {mo[row * 6 + 2], mo[row * 6 + 3]},
                      f2 = {mo[row * 6 + 4], mo[row * 6 + 5]};
                cl[row] = f2_add(f2_add(f0, f2_mul(f2_add(f1, f1, q), ch, nr, q), q),
                                 f2_mul(f2, c2, nr, q), q);
                fq2_t *c = cur + row * L;
                for (size_t x = 0; x < half; x++) {
                    fq2_t dl = f2_sub(c[x + half], c[x], q);
                    c[x] = f2_add(f2_mul(dl, ch, nr, q), c[x], q);
                }
            }
            len = half;
        }
        for (size_t row = 0; row < rows; row++) {
            fq2_t ev = cur[row * L];
            if (!f2_eq(f2_mul(ev, ev, nr, q), cl[row])) ok[rep * rows + row] = 0;
            finals[(rep * rows + row) * 2] = cl[row].a;
            finals[(rep * rows + row) * 2 + 1] = cl[row].b;
            evals[(rep * rows + row) * 2] = ev.a;
            evals[(rep * rows + row) * 2 + 1] = ev.b;
        }
    }
    free(cur);
    free(cl);
}

/* fast_norm_check.rs:401-441: cs = [ring challenges (log2 m) | coefficient
 * challenges (log2 d)]; base = conj(tensor(coef cs)) in IncompleteNTT;
 * ps_i = tensor(ring cs)_i * base (complete_power_series, arithmetic.rs:158-176,
 * built from constants (1-t, t) whose products are constants);
 * rhs_j = sum_i ps_i * W_j[i] (arithmetic.rs:507-536). */
void or_norm_check_3(const uint64_t *W, size_t rows, size_t m, size_t d,
                     const uint64_t *cs, uint64_t q, uint64_t *rhs) {
    size_t nr_ = ilog2(m), nc = ilog2(d);
    uint64_t *tc = (uint64_t *)malloc(d * sizeof(uint64_t));
    uint64_t *tr = (uint64_t *)malloc(m * sizeof(uint64_t));
    uint64_t *base = (uint64_t *)malloc(d * sizeof(uint64_t));
    uint64_t *om = (uint64_t *)malloc((d / 2) * sizeof(uint64_t));
    uint64_t *ps = (uint64_t *)malloc(d * sizeof(uint64_t));
    uint64_t *wi = (uint64_t *)malloc(d * sizeof(uint64_t));
    uint64_t *prod = (uint64_t *)malloc(d * sizeof(uint64_t));
    or_tensor(cs + nr_, nc, q, tc);
    or_tensor(cs, nr_, q, tr);
    or_conjugate(tc, base, d, q);
    or_to_incomplete(base, d, q);
    or_shift_factors(d, q, om);
    for (size_t j = 0; j < rows; j++) {
        uint64_t *acc = rhs + j * d;
        memset(acc, 0, d * sizeof(uint64_t));
        for (size_t i = 0; i < m; i++) {
            /* constant(t) in IncompleteNTT is t in every slot, so the
             * constant-by-base product scales every entry of base. */
            for (size_t c = 0; c < d; c++) ps[c] = or_mulmod(base[c], tr[i], q);
            memcpy(wi, W + (j * m + i) * d, d * sizeof(uint64_t));
            or_to_incomplete(wi, d, q);
            slot_mul(ps, wi, prod, om, d, q);
            for (size_t c = 0; c < d; c++) acc[c] = or_addmod(acc[c], prod[c], q);
        }
    }
    free(tc); free(tr); free(base); free(om); free(ps); free(wi); free(prod);
}

static void f2_tensor(const uint64_t *cs2, size_t n, uint64_t nr, uint64_t q, fq2_t *out) {
    size_t len = 1;
    out[0].a = 1 % q;
    out[0].b = 0;
    for (size_t t = 0; t < n; t++) {
        fq2_t c = {cs2[2 * t] % q, cs2[2 * t + 1] % q};
        fq2_t one = {1 % q, 0};
        fq2_t om = f2_sub(one, c, q);
        for (size_t v = len; v-- > 0;) {
            fq2_t val = out[v];
            out[2 * v] = f2_mul(val, om, nr, q);
            out[2 * v + 1] = f2_mul(val, c, nr, q);
        }
        len *= 2;
    }
}

/* F_{q^2} analogue of norm_check_3 (an extension; the reference is F_q only):
 * V = sum_i tensor(ring cs)_i W[i] in R_q[Y]/(Y^2-nr), base = conj(tensor(coef
 * cs)), rhs = base * V = (b0 v0 + nr b1 v1) + Y (b0 v1 + b1 v0), each
 * component in IncompleteNTT form.  ct(rhs) = <w, tensor(cs)>. */
void or_norm_check_3_fq2(const uint64_t *W, size_t rows, size_t m, size_t d,
                         const uint64_t *cs2, uint64_t nr, uint64_t q, uint64_t *rhs) {
    size_t nr_ = ilog2(m), nc = ilog2(d);
    fq2_t *tc = (fq2_t *)malloc(d * sizeof(fq2_t));
    fq2_t *tr = (fq2_t *)malloc(m * sizeof(fq2_t));
    uint64_t *b0 = (uint64_t *)malloc(d * 8), *b1 = (uint64_t *)malloc(d * 8);
    uint64_t *v0 = (uint64_t *)malloc(d * 8), *v1 = (uint64_t *)malloc(d * 8);
    uint64_t *p = (uint64_t *)malloc(d * 8), *s = (uint64_t *)malloc(d * 8);
    f2_tensor(cs2 + 2 * nr_, nc, nr, q, tc);
    f2_tensor(cs2, nr_, nr, q, tr);
    for (size_t c = 0; c < d; c++) { b0[c] = tc[c].a; b1[c] = tc[c].b; }
    or_conjugate(b0, b0, d, q);
    or_conjugate(b1, b1, d, q);
    for (size_t j = 0; j < rows; j++) {
        memset(v0, 0, d * 8);
        memset(v1, 0, d * 8);
        for (size_t i = 0; i < m; i++)
            for (size_t c = 0; c < d; c++) {
                uint64_t x = W[(j * m + i) * d + c];
                v0[c] = or_addmod(v0[c], or_mulmod(x, tr[i].a, q), q);
                v1[c] = or_addmod(v1[c], or_mulmod(x, tr[i].b, q), q);
            }
        uint64_t *r0 = rhs + (j * 2) * d, *r1 = rhs + (j * 2 + 1) * d;
        /* r0 = b0 v0 + nr b1 v1 ; r1 = b0 v1 + b1 v0 (coefficient domain) */
        or_ring_mul_naive(b0, v0, r0, d, q);
        or_ring_mul_naive(b1, v1, p, d, q);
        for (size_t c = 0; c < d; c++) r0[c] = or_addmod(r0[c], or_mulmod(nr, p[c], q), q);
        or_ring_mul_naive(b0, v1, r1, d, q);
        or_ring_mul_naive(b1, v0, s, d, q);
        for (size_t c = 0; c < d; c++) r1[c] = or_addmod(r1[c], s[c], q);
        or_to_incomplete(r0, d, q);
        or_to_incomplete(r1, d, q);
    }
    free(tc); free(tr); free(b0); free(b1); free(v0); free(v1); free(p); free(s);
}

/* ------------------------------------------------------------------------ */
/* decomposition                                                             */
/* ------------------------------------------------------------------------ */

static inline uint64_t digit_of(uint64_t v, unsigned shift, size_t t, uint64_t mask) {
    unsigned s = (unsigned)(shift * t);
    if (shift == 0) return v & mask;
    if (s >= 64) return 0;
    return (v >> s) & mask;
}

/* arithmetic.rs:1572-1593 (decompose_matrix_by_chunks) and 1537-1570
 * (decompose): split into pos (2x <= q) / neg (q - x) parts (:72-144);
 * auto radix 2^ceil(bw/parts) with bw the bit width of the OR of all parts
 * (:1461-1474, :1579-1585); limb t of row rho is digit_t(pos) - digit_t(neg)
 * (:1476-1486, :954-983) and lands in output row rho*parts + t. */
uint64_t or_decompose(const uint64_t *W, size_t r, size_t m, size_t d, size_t parts,
                      uint64_t radix, uint64_t q, uint64_t *out) {
    size_t n = r * m * d;
    if (radix == 0) {
        uint64_t acc = 0;
        for (size_t i = 0; i < n; i++) {
            uint64_t x = W[i];
            acc |= (2 * x > q) ? q - x : x;
        }
        unsigned bw = acc ? 64u - (unsigned)__builtin_clzll(acc) : 0u;
        unsigned lg = (unsigned)((bw + parts - 1) / parts);
        radix = 1ULL << lg;
    }
    unsigned shift = (unsigned)__builtin_ctzll(radix);
    uint64_t mask = radix - 1;
    for (size_t rho = 0; rho < r; rho++)
        for (size_t i = 0; i < m; i++)
            for (size_t c = 0; c < d; c++) {
                uint64_t x = W[(rho * m + i) * d + c];
                uint64_t pos = (2 * x > q) ? 0 : x;
                uint64_t neg = (2 * x > q) ? q - x : 0;
                for (size_t t = 0; t < parts; t++) {
                    uint64_t dp = digit_of(pos, shift, t, mask);
                    uint64_t dn = digit_of(neg, shift, t, mask);
                    out[((rho * parts + t) * m + i) * d + c] =
                        or_addmod(dp % q, or_submod(0, dn % q, q), q);
                }
            }
    return radix;
}

/* arithmetic.rs:1719-1755 (compose_with_radix_mod): in [rows][n*parts][d],
 * out[row][i] = sum_t in[row][i*parts + t] * 2^(t*log2 radix) mod q.  Groups
 * consecutive columns, which inverts the limb layout of a commitment
 * [kappa][r*parts] (decomp.rs:99-105). */
void or_compose(const uint64_t *in, size_t rows, size_t n, size_t d, size_t parts,
                uint64_t radix, uint64_t q, uint64_t *out) {
    unsigned shift = (unsigned)__builtin_ctzll(radix);
    for (size_t row = 0; row < rows; row++)
        for (size_t i = 0; i < n; i++)
            for (size_t c = 0; c < d; c++) {
                uint64_t acc = 0;
                for (size_t t = 0; t < parts; t++) {
                    uint64_t wt = or_powmod(2, (uint64_t)shift * t, q);
                    uint64_t x = in[(row * n * parts + i * parts + t) * d + c];
                    acc = or_addmod(acc, or_mulmod(x, wt, q), q);
                }
                out[(row * n + i) * d + c] = acc;
            }
}

/* Gadget recomposition of a decomposed witness: limb rows rho*parts + t of
 * [r*parts][m][d] recombine into row rho (the inverse of or_decompose when
 * parts * log2 radix covers the bit width). */
void or_compose_rows(const uint64_t *limbs, size_t r, size_t m, size_t d, size_t parts,
                     uint64_t radix, uint64_t q, uint64_t *out) {
    unsigned shift = (unsigned)__builtin_ctzll(radix);
    for (size_t rho = 0; rho < r; rho++)
        for (size_t i = 0; i < m; i++)
            for (size_t c = 0; c < d; c++) {
                uint64_t acc = 0;
                for (size_t t = 0; t < parts; t++) {
                    uint64_t wt = or_powmod(2, (uint64_t)shift * t, q);
                    acc = or_addmod(acc, or_mulmod(limbs[((rho * parts + t) * m + i) * d + c],
                                                   wt, q), q);
                }
                out[(rho * m + i) * d + c] = acc;
            }
}

/* ------------------------------------------------------------------------ */
/* fold                                                                      */
/* ------------------------------------------------------------------------ */

/* fold.rs:22-27 -> arithmetic.rs:429-457: out[j][i] = sum_k W[k][i] * C[j][k] */
void or_fold(const uint64_t *W, const uint64_t *C, uint64_t *out, size_t r_in,
             size_t r_out, size_t m, size_t d, uint64_t q, int threads) {
    long total = (long)(r_out * m);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static) if (threads != 1)
#endif
    for (long ji = 0; ji < total; ji++) {
        size_t j = (size_t)ji / m, i = (size_t)ji % m;
        uint64_t *acc = out + (j * m + i) * d;
        uint64_t *prod = (uint64_t *)malloc(d * sizeof(uint64_t));
        memset(acc, 0, d * sizeof(uint64_t));
        for (size_t k = 0; k < r_in; k++) {
            or_ring_mul_incomplete(W + (k * m + i) * d, C + (j * r_in + k) * d, prod, d, q);
            for (size_t c = 0; c < d; c++) acc[c] = or_addmod(acc[c], prod[c], q);
        }
        free(prod);
    }
}

/* ------------------------------------------------------------------------ */
/* seeded sampling                                                           */
/* ------------------------------------------------------------------------ */

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Counter-based SplitMix64: element idx of stream `stream` under `seed`. */
uint64_t or_prng_draw(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint64_t key = mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL));
    return mix64(key + (idx + 1) * 0x9E3779B97F4A7C15ULL);
}

/* uniform in [0, q) by multiply-high (cyclotomic_ring.rs:257-269 samples
 * uniformly with an unseeded rng) */
void or_sample_uniform(uint64_t seed, uint64_t stream, size_t n, uint64_t q,
                       uint64_t *out) {
    for (size_t i = 0; i < n; i++)
        out[i] = (uint64_t)(((u128)or_prng_draw(seed, stream, i) * q) >> 64);
}

/* elements start .. start + n - 1 of the same uniform stream (a row range
 * of a matrix sampled whole by or_sample_uniform) */
void or_sample_uniform_at(uint64_t seed, uint64_t stream, size_t start, size_t n, uint64_t q,
                          uint64_t *out) {
    for (size_t i = 0; i < n; i++)
        out[i] = (uint64_t)(((u128)or_prng_draw(seed, stream, start + i) * q) >> 64);
}

/* one random bit per byte (challenge_for_project, project.rs:26-39, draws
 * rand::random::<bool>()): the low bit of each draw */
void or_sample_bits(uint64_t seed, uint64_t stream, size_t n, uint8_t *out) {
    for (size_t i = 0; i < n; i++) out[i] = (uint8_t)(or_prng_draw(seed, stream, i) & 1);
}

/* the key of stream `stream` under `seed` (the per-round seed of the
 * protocol's challenges is the key of stream 0x7000 + round) */
uint64_t or_stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL));
}

/* uniform over the 2*bound+1 integers [-bound, bound], mapped mod q
 * (cyclotomic_ring.rs:293-317 draws |v| < b with a random sign) */
void or_sample_bounded(uint64_t seed, uint64_t stream, size_t n, uint64_t bound,
                       uint64_t q, uint64_t *out) {
    uint64_t span = 2 * bound + 1;
    for (size_t i = 0; i < n; i++) {
        uint64_t t = (uint64_t)(((u128)or_prng_draw(seed, stream, i) * span) >> 64);
        out[i] = t >= bound ? t - bound : q - (bound - t);
    }
}
