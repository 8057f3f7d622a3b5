// legacy_abi.cpp -- the 21 hexl_wrapper symbols the reference's Rust FFI binds
// (src/hexl/bindings.rs:3-73, hexl-bindings/hexl_wrapper.cpp:34-276).
//
// These take host pointers to one small vector or scalar per call, exactly as
// the reference calls them (one FFI call per ring element), so they run on the
// host.  The batched device entry points in api.hip are the hot path.
//
// Arithmetic: Barrett reduction for products, Shoup-precomputed twiddles for
// the NTT (q < 2^62).  Tables are cached per (n, q) behind a mutex, which the
// reference's NTTCache (hexl_wrapper.cpp:11-32) lacks.
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/salsaa_amd.h"

namespace {

typedef unsigned __int128 u128;

struct Barrett {
    uint64_t q;
    unsigned k;
    u128 mu;
    explicit Barrett(uint64_t q_) : q(q_) {
        k = 64u - (unsigned)__builtin_clzll(q);
        mu = (((u128)1) << (2 * k)) / q;
    }
    // x < q^2
    uint64_t reduce(u128 x) const {
        u128 t = ((x >> (k - 1)) * mu) >> (k + 1);
        u128 r = x - t * q;
        while (r >= q) r -= q;
        return (uint64_t)r;
    }
    uint64_t mul(uint64_t a, uint64_t b) const { return reduce((u128)a * b); }
};

inline uint64_t mulmod_any(uint64_t a, uint64_t b, uint64_t q) {
    return (uint64_t)((u128)a * b % q);
}

uint64_t powmod(uint64_t a, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) r = mulmod_any(r, a, q);
        a = mulmod_any(a, a, q);
        e >>= 1;
    }
    return r;
}

unsigned log2u(uint64_t n) {
    unsigned l = 0;
    while ((1ULL << l) < n) l++;
    return l;
}

uint64_t reverse_bits(uint64_t x, unsigned bits) {
    uint64_t r = 0;
    for (unsigned i = 0; i < bits; i++) r |= ((x >> i) & 1ULL) << (bits - 1 - i);
    return r;
}

// Smallest element of exact multiplicative order `order` (HEXL's
// MinimalPrimitiveRoot): all such elements are odd powers of any one of them.
uint64_t minimal_primitive_root(uint64_t order, uint64_t q) {
    uint64_t w = 0;
    for (uint64_t g = 2; g < q && !w; g++) {
        uint64_t c = powmod(g, (q - 1) / order, q);
        if (powmod(c, order / 2, q) == q - 1) w = c;
    }
    uint64_t w2 = mulmod_any(w, w, q), cur = w, best = w;
    for (uint64_t i = 0; i < order / 2; i++) {
        if (cur < best) best = cur;
        cur = mulmod_any(cur, w2, q);
    }
    return best;
}

struct NttTable {
    uint64_t n, q;
    std::vector<uint64_t> roots;        // roots[brv(i)] = psi^i
    std::vector<uint64_t> roots_shoup;  // floor(roots * 2^64 / q)
    std::vector<uint64_t> inv;          // inv[k] = roots[k]^{-1}
    std::vector<uint64_t> inv_shoup;
    std::vector<uint64_t> inv_hexl;     // inv[0], then inv[m .. 2m) for m = n/2, n/4, .., 1:
                                        // the order the inverse transform's loops read them
    uint64_t n_inv, n_inv_shoup;
};

uint64_t shoup_of(uint64_t w, uint64_t q) { return (uint64_t)((((u128)w) << 64) / q); }

// x * w mod q with the Shoup companion, result in [0, q)
inline uint64_t shoup_mul(uint64_t x, uint64_t w, uint64_t ws, uint64_t q) {
    uint64_t qt = (uint64_t)(((u128)x * ws) >> 64);
    uint64_t r = x * w - qt * q;
    return r >= q ? r - q : r;
}

std::mutex g_lock;
std::map<std::pair<uint64_t, uint64_t>, std::unique_ptr<NttTable>> g_tables;

const NttTable &table(uint64_t n, uint64_t q) {
    std::lock_guard<std::mutex> g(g_lock);
    auto key = std::make_pair(n, q);
    auto it = g_tables.find(key);
    if (it != g_tables.end()) return *it->second;
    auto t = std::make_unique<NttTable>();
    t->n = n;
    t->q = q;
    unsigned bits = log2u(n);
    uint64_t psi = minimal_primitive_root(2 * n, q);
    t->roots.assign(n, 0);
    uint64_t p = 1;
    for (uint64_t i = 0; i < n; i++) {
        t->roots[reverse_bits(i, bits)] = p;
        p = mulmod_any(p, psi, q);
    }
    t->roots_shoup.resize(n);
    t->inv.resize(n);
    t->inv_shoup.resize(n);
    for (uint64_t i = 0; i < n; i++) {
        t->roots_shoup[i] = shoup_of(t->roots[i], q);
        t->inv[i] = powmod(t->roots[i], q - 2, q);
        t->inv_shoup[i] = shoup_of(t->inv[i], q);
    }
    t->inv_hexl.reserve(n);
    t->inv_hexl.push_back(t->inv[0]);
    for (uint64_t m = n >> 1; m > 0; m >>= 1)
        for (uint64_t i = 0; i < m; i++) t->inv_hexl.push_back(t->inv[m + i]);
    t->n_inv = powmod(n % q, q - 2, q);
    t->n_inv_shoup = shoup_of(t->n_inv, q);
    const NttTable &ref = *t;
    g_tables.emplace(key, std::move(t));
    return ref;
}

void ntt_fwd(uint64_t *x, uint64_t n, uint64_t q) {
    if (n < 2) return;
    const NttTable &tb = table(n, q);
    uint64_t t = n;
    for (uint64_t m = 1; m < n; m <<= 1) {
        t >>= 1;
        for (uint64_t i = 0; i < m; i++) {
            uint64_t w = tb.roots[m + i], ws = tb.roots_shoup[m + i];
            uint64_t *a = x + 2 * i * t, *b = a + t;
            for (uint64_t j = 0; j < t; j++) {
                uint64_t u = a[j], v = shoup_mul(b[j], w, ws, q);
                uint64_t s = u + v;
                a[j] = s >= q ? s - q : s;
                b[j] = u >= v ? u - v : u + q - v;
            }
        }
    }
}

void ntt_inv(uint64_t *x, uint64_t n, uint64_t q) {
    if (n < 2) return;
    const NttTable &tb = table(n, q);
    uint64_t t = 1;
    for (uint64_t m = n >> 1; m >= 1; m >>= 1) {
        for (uint64_t i = 0; i < m; i++) {
            uint64_t w = tb.inv[m + i], ws = tb.inv_shoup[m + i];
            uint64_t *a = x + 2 * i * t, *b = a + t;
            for (uint64_t j = 0; j < t; j++) {
                uint64_t u = a[j], v = b[j];
                uint64_t s = u + v;
                a[j] = s >= q ? s - q : s;
                b[j] = shoup_mul(u >= v ? u - v : u + q - v, w, ws, q);
            }
        }
        t <<= 1;
    }
    for (uint64_t j = 0; j < n; j++) x[j] = shoup_mul(x[j], tb.n_inv, tb.n_inv_shoup, q);
}

inline uint64_t addm(uint64_t a, uint64_t b, uint64_t q) {
    uint64_t s = a + b;
    return (s >= q || s < a) ? s - q : s;
}
inline uint64_t subm(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + (q - b); }

}  // namespace

extern "C" {

uint64_t multiply_mod(uint64_t a, uint64_t b, uint64_t modulus) {
    return mulmod_any(a, b, modulus);
}

uint64_t power_mod(uint64_t a, uint64_t b, uint64_t modulus) { return powmod(a, b, modulus); }

uint64_t add_mod(uint64_t a, uint64_t b, uint64_t modulus) { return addm(a, b, modulus); }

uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t modulus) { return subm(a, b, modulus); }

uint64_t inv_mod(uint64_t a, uint64_t modulus) { return powmod(a, modulus - 2, modulus); }

const uint64_t *get_roots(uint64_t n, uint64_t modulus) { return table(n, modulus).roots.data(); }

const uint64_t *get_inv_roots(uint64_t n, uint64_t modulus) {
    return table(n, modulus).inv_hexl.data();
}

void eltwise_mult_mod(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2,
                      uint64_t n, uint64_t modulus) {
    Barrett b(modulus);
    for (uint64_t i = 0; i < n; i++) result[i] = b.mul(operand1[i], operand2[i]);
}

void eltwise_fma_mod(uint64_t *result, const uint64_t *operand1, uint64_t operand2,
                     const uint64_t *operand3, uint64_t n, uint64_t modulus) {
    Barrett b(modulus);
    for (uint64_t i = 0; i < n; i++)
        result[i] = addm(b.mul(operand1[i], operand2), operand3[i], modulus);
}

void eltwise_add_mod(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2,
                     uint64_t n, uint64_t modulus) {
    for (uint64_t i = 0; i < n; i++) result[i] = addm(operand1[i], operand2[i], modulus);
}

void eltwise_sub_mod(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2,
                     uint64_t n, uint64_t modulus) {
    for (uint64_t i = 0; i < n; i++) result[i] = subm(operand1[i], operand2[i], modulus);
}

void eltwise_reduce_mod(uint64_t *result, const uint64_t *operand, uint64_t n, uint64_t modulus) {
    for (uint64_t i = 0; i < n; i++) result[i] = operand[i] % modulus;
}

void eltwise_reduce_mod_naive(uint64_t *result, const uint64_t *operand, uint64_t n,
                              uint64_t modulus) {
    for (uint64_t i = 0; i < n; i++) result[i] = operand[i] % modulus;
}

uint64_t sum(const uint64_t *operand1, uint64_t n, uint64_t modulus) {
    uint64_t s = 0;
    for (uint64_t i = 0; i < n; i++) s = addm(s, operand1[i], modulus);
    return s;
}

uint64_t sum_fast(const uint64_t *operand1, uint64_t n, uint64_t modulus) {
    // The reference's tree (hexl_wrapper.cpp:239-260) writes through a
    // reserved-but-empty vector; the defined result is the sum mod q.
    return sum(operand1, n, modulus);
}

uint64_t sum_sq(const uint64_t *operand1, uint64_t n, uint64_t modulus) {
    Barrett b(modulus);
    uint64_t s = 0;
    for (uint64_t i = 0; i < n; i++) s = addm(s, b.mul(operand1[i], operand1[i]), modulus);
    return s;
}

uint64_t sum_sq_fast(const uint64_t *operand1, uint64_t n, uint64_t modulus) {
    return sum_sq(operand1, n, modulus);
}

void ntt_forward_in_place(uint64_t *operand, size_t n, uint64_t modulus) {
    ntt_fwd(operand, n, modulus);
}

void ntt_inverse_in_place(uint64_t *operand, size_t n, uint64_t modulus) {
    ntt_inv(operand, n, modulus);
}

void multiply_poly(uint64_t *result, const uint64_t *operand1, const uint64_t *operand2,
                   uint64_t n, uint64_t modulus) {
    std::vector<uint64_t> a(operand1, operand1 + n), b(operand2, operand2 + n);
    ntt_fwd(a.data(), n, modulus);
    ntt_fwd(b.data(), n, modulus);
    Barrett br(modulus);
    for (uint64_t i = 0; i < n; i++) a[i] = br.mul(a[i], b[i]);
    ntt_inv(a.data(), n, modulus);
    std::memcpy(result, a.data(), n * sizeof(uint64_t));
}

void polynomial_multiply_cyclotomic_mod(uint64_t *result, const uint64_t *operand1,
                                        const uint64_t *operand2, uint64_t phi, uint64_t mod_q) {
    // The reference zero-extends the size-phi negacyclic product to 2*phi - 1
    // entries and then folds entries >= phi (all zero); the result is the
    // negacyclic product itself (hexl_wrapper.cpp:95-117).
    std::vector<uint64_t> tmp(2 * phi - 1, 0);
    multiply_poly(tmp.data(), operand1, operand2, phi, mod_q);
    for (uint64_t i = phi; i < tmp.size(); i++) tmp[i - phi] = addm(tmp[i - phi], tmp[i], mod_q);
    uint64_t top = phi < tmp.size() ? tmp[phi] : 0;
    for (uint64_t i = 0; i < phi; i++) tmp[i] = (mod_q + tmp[i] - top) % mod_q;
    std::memcpy(result, tmp.data(), phi * sizeof(uint64_t));
}

}  // extern "C"
