// cpu_baseline.cpp -- the CPU baseline of the benchmark: the reference's CPU
// prover path restated in C++ with the reference's data flow, on all host
// cores.  BENCHMARK AND TEST INFRASTRUCTURE ONLY: bench.py times it beside the
// device path and tests/ check it against the oracle; the product library
// never loads it.
//
// What is kept from the reference (so the timing is the reference's
// algorithm, not a different one):
//   - the commit parallel_dot_series_matrix (arithmetic.rs:459-505) converts
//     the witness element of every ring product to IncompleteNTT form again
//     (Mul by value converts a copy, cyclotomic_ring.rs:185-199), so each
//     witness element is transformed once per key row; the key rows are held
//     in IncompleteNTT form, as products of the CRS are; each product is the
//     5 multiplications and 2 additions of incomplete_ntt_multiplication
//     (cyclotomic_ring.rs:699-791), summed per (key row, witness row) over
//     chunks of the length split across the threads;
//   - the norm check sum_check (fast_norm_check.rs:135-199): per round the
//     separate passes w0^2, w1 - w0, w0 (w1 - w0), (w1 - w0)^2, three sums and
//     the fold, each split into chunks of max(len / 120, 1028) over the
//     threads, a fresh zeroed buffer for the folded witness every round, and
//     each repetition restarting from the witness (protocol.rs:222-259); over
//     F_{q^2} = F_q[Y]/(Y^2 - nr) once the first challenge is folded in, as
//     the device's F_{q^2} norm check (the reference's is F_q only);
//   - norm_check_1 (sum_sq_par, :127-133) and norm_check_3 (:401-441): the
//     power series expanded layer by layer (complete_power_series,
//     arithmetic.rs:158-176) with a ring product per entry, then its inner
//     product with the witness, converting each witness element per product;
//   - the NTT round trip of the ring elements: to_incomplete_ntt and back per
//     element (cyclotomic_ring.rs:483-556);
//   - decomp_ell (decomp.rs:32-67): the bit-width pass, the digits of every
//     coefficient (row rho * parts + t holds limb t of row rho), then the limb
//     commit as above; verify_decomp recomposes the limb commitments.
// What differs: modular products use AVX-512 IFMA (52-bit multiply-add;
// q < 2^50 is HEXL's IFMA regime) in place of HEXL, with eight independent
// ring elements or products in the eight lanes of a vector, and a persistent
// thread pool in place of rayon.  Without IFMA the same code runs with scalar
// 128-bit products (cb_variant says which).  Results are canonical residues,
// equal to the oracle's (tests/test_cpu_baseline.py).
#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

typedef unsigned __int128 u128;

namespace {

// ------------------------------------------------------------------ threads
class Pool {
  public:
    void resize(int n) {
        stop();
        if (n < 1) n = 1;
        nthreads_ = n;
        stop_ = false;
        for (int i = 1; i < n; i++) workers_.emplace_back([this] { loop(); });
    }
    int size() const { return nthreads_; }
    // fn(i) for i < n, over the pool and the calling thread
    void run(size_t n, const std::function<void(size_t)> &fn) {
        if (n == 0) return;
        if (workers_.empty() || n == 1) {
            for (size_t i = 0; i < n; i++) fn(i);
            return;
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            fn_ = &fn;
            n_ = n;
            next_.store(0);
            active_ = (int)workers_.size();
            gen_++;
        }
        cv_.notify_all();
        for (size_t i; (i = next_.fetch_add(1)) < n;) fn(i);
        std::unique_lock<std::mutex> g(mu_);
        done_.wait(g, [this] { return active_ == 0; });
        fn_ = nullptr;
    }
    ~Pool() { stop(); }

  private:
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(size_t)> *fn;
            size_t n;
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                fn = fn_;
                n = n_;
            }
            for (size_t i; (i = next_.fetch_add(1)) < n;) (*fn)(i);
            std::lock_guard<std::mutex> g(mu_);
            if (--active_ == 0) done_.notify_all();
        }
    }
    void stop() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : workers_) t.join();
        workers_.clear();
    }
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(size_t)> *fn_ = nullptr;
    size_t n_ = 0;
    std::atomic<size_t> next_{0};
    int active_ = 0, nthreads_ = 1;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

Pool &pool() {
    static Pool p;
    static bool init = [] {
        unsigned n = std::thread::hardware_concurrency();
        p.resize(n ? (int)n : 1);
        return true;
    }();
    (void)init;
    return p;
}

// ------------------------------------------------------------------ scalar arithmetic
inline uint64_t mulm(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }
inline uint64_t addm(uint64_t a, uint64_t b, uint64_t q) {
    const uint64_t s = a + b;
    return s >= q ? s - q : s;
}
inline uint64_t subm(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + q - b; }
uint64_t powm(uint64_t a, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    for (a %= q; e; e >>= 1, a = mulm(a, a, q))
        if (e & 1) r = mulm(r, a, q);
    return r;
}
unsigned brv(unsigned x, int bits) {
    unsigned r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}
// the minimal primitive 2n-th root of unity (HEXL's choice of psi)
uint64_t min_root(uint64_t order, uint64_t q) {
    uint64_t w = 0;
    for (uint64_t g = 2; g < q && !w; g++) {
        const uint64_t c = powm(g, (q - 1) / order, q);
        if (powm(c, order / 2, q) == q - 1) w = c;
    }
    uint64_t best = w, cur = w;
    const uint64_t w2 = mulm(w, w, q);
    for (uint64_t i = 0; i < order / 2; i++, cur = mulm(cur, w2, q)) best = std::min(best, cur);
    return best;
}

// Tables of one modulus: the 32-point negacyclic NTT (HEXL convention: psi
// the minimal primitive 64th root, roots[k] = psi^brv(k), bit-reversed
// output), the slot moduli omega_j = psi^(2 brv(j) + 1), Shoup companions
// floor(w 2^52 / q) and the Barrett constant floor(2^100 / q).
struct Tab {
    uint64_t q = 0, mu = 0, ninv = 0, ninv_s = 0;
    bool ifma_ok = false;  // 2^49 < q < 2^50: the lanes' Barrett bound holds
    uint64_t fw[32], fw_s[32], iw[32], iw_s[32], om[32], om_s[32];
};

uint64_t shoup(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 52) / q); }

const Tab &tables(uint64_t q) {
    static std::mutex mu;
    static std::vector<Tab *> all;
    std::lock_guard<std::mutex> g(mu);
    for (Tab *t : all)
        if (t->q == q) return *t;
    Tab *t = new Tab();
    t->q = q;
    t->mu = (uint64_t)(((u128)1 << 100) / q);
    t->ifma_ok = (q >> 49) == 1;
    const uint64_t psi = min_root(64, q);
    for (unsigned k = 0; k < 32; k++) {
        t->fw[k] = powm(psi, brv(k, 5), q);
        t->iw[k] = powm(t->fw[k], q - 2, q);
        t->om[k] = powm(psi, 2 * brv(k, 5) + 1, q);
        t->fw_s[k] = shoup(t->fw[k], q);
        t->iw_s[k] = shoup(t->iw[k], q);
        t->om_s[k] = shoup(t->om[k], q);
    }
    t->ninv = powm(32, q - 2, q);
    t->ninv_s = shoup(t->ninv, q);
    all.push_back(t);
    return *t;
}

// forward (Cooley-Tukey, bit-reversed output) and inverse (Gentleman-Sande,
// natural output, times 1/32) 32-point negacyclic NTT, in place
void ntt32_fwd(uint64_t *x, const Tab &T) {
    const uint64_t q = T.q;
    for (unsigned m = 1, t = 16; m < 32; m <<= 1, t >>= 1)
        for (unsigned i = 0; i < m; i++) {
            const uint64_t w = T.fw[m + i];
            for (unsigned j = 2 * i * t; j < 2 * i * t + t; j++) {
                const uint64_t y = mulm(x[j + t], w, q);
                x[j + t] = subm(x[j], y, q);
                x[j] = addm(x[j], y, q);
            }
        }
}
void ntt32_inv(uint64_t *x, const Tab &T) {
    const uint64_t q = T.q;
    for (unsigned m = 16, t = 1; m >= 1; m >>= 1, t <<= 1)
        for (unsigned i = 0; i < m; i++) {
            const uint64_t w = T.iw[m + i];
            for (unsigned j = 2 * i * t; j < 2 * i * t + t; j++) {
                const uint64_t a = x[j], b = x[j + t];
                x[j] = addm(a, b, q);
                x[j + t] = mulm(subm(a, b, q), w, q);
            }
        }
    for (unsigned j = 0; j < 32; j++) x[j] = mulm(x[j], T.ninv, q);
}
void to_inc(uint64_t *a, const Tab &T) {
    uint64_t e[32], o[32];
    for (int i = 0; i < 32; i++) e[i] = a[2 * i], o[i] = a[2 * i + 1];
    ntt32_fwd(e, T);
    ntt32_fwd(o, T);
    memcpy(a, e, sizeof e);
    memcpy(a + 32, o, sizeof o);
}
void from_inc(uint64_t *a, const Tab &T) {
    uint64_t e[32], o[32];
    memcpy(e, a, sizeof e);
    memcpy(o, a + 32, sizeof o);
    ntt32_inv(e, T);
    ntt32_inv(o, T);
    for (int i = 0; i < 32; i++) a[2 * i] = e[i], a[2 * i + 1] = o[i];
}
// incomplete_ntt_multiplication (both IncompleteNTT): out may alias neither
void inc_mul(const uint64_t *a, const uint64_t *b, uint64_t *out, const Tab &T) {
    const uint64_t q = T.q;
    for (int s = 0; s < 32; s++) {
        const uint64_t e1 = a[s], o1 = a[32 + s], e2 = b[s], o2 = b[32 + s];
        out[s] = addm(mulm(e1, e2, q), mulm(mulm(o1, o2, q), T.om[s], q), q);
        out[32 + s] = addm(mulm(o1, e2, q), mulm(e1, o2, q), q);
    }
}

// ------------------------------------------------------------------ IFMA lanes
#define TGT __attribute__((target("avx512f,avx512ifma,avx512dq,avx512bw,avx512vl")))
typedef __m512i V;

struct VQ {
    V q, mu, m52;
};
TGT inline VQ vq(const Tab &T) {
    return VQ{_mm512_set1_epi64((long long)T.q), _mm512_set1_epi64((long long)T.mu),
              _mm512_set1_epi64((long long)((1ULL << 52) - 1))};
}
TGT inline V csub(V x, V q) { return _mm512_min_epu64(x, _mm512_sub_epi64(x, q)); }
TGT inline V vadd(V a, V b, const VQ &m) { return csub(_mm512_add_epi64(a, b), m.q); }
TGT inline V vsub(V a, V b, const VQ &m) {
    return csub(_mm512_add_epi64(_mm512_sub_epi64(a, b), m.q), m.q);
}
// a b mod q for canonical a, b (Barrett on the 52-bit halves of a b: the
// estimate floor(floor(ab / 2^48) mu / 2^52) is at most 2 below floor(ab / q))
TGT inline V vmul(V a, V b, const VQ &m) {
    const V z = _mm512_setzero_si512();
    const V lo = _mm512_madd52lo_epu64(z, a, b);
    const V hi = _mm512_madd52hi_epu64(z, a, b);
    const V t = _mm512_or_si512(_mm512_slli_epi64(hi, 4), _mm512_srli_epi64(lo, 48));
    const V qt = _mm512_madd52hi_epu64(z, t, m.mu);
    V r = _mm512_and_si512(_mm512_sub_epi64(lo, _mm512_madd52lo_epu64(z, qt, m.q)), m.m52);
    return csub(csub(r, m.q), m.q);
}
// x w mod q for a fixed w with Shoup companion ws = floor(w 2^52 / q)
TGT inline V vmulc(V x, V w, V ws, const VQ &m) {
    const V z = _mm512_setzero_si512();
    const V t = _mm512_madd52hi_epu64(z, x, ws);
    V r = _mm512_sub_epi64(_mm512_madd52lo_epu64(z, x, w), _mm512_madd52lo_epu64(z, t, m.q));
    return csub(_mm512_and_si512(r, m.m52), m.q);
}
TGT inline V bc(uint64_t x) { return _mm512_set1_epi64((long long)x); }

TGT void vntt_fwd(V *x, const Tab &T, const VQ &m) {
    for (unsigned mm = 1, t = 16; mm < 32; mm <<= 1, t >>= 1)
        for (unsigned i = 0; i < mm; i++) {
            const V w = bc(T.fw[mm + i]), ws = bc(T.fw_s[mm + i]);
            for (unsigned j = 2 * i * t; j < 2 * i * t + t; j++) {
                const V y = vmulc(x[j + t], w, ws, m);
                x[j + t] = vsub(x[j], y, m);
                x[j] = vadd(x[j], y, m);
            }
        }
}
TGT void vntt_inv(V *x, const Tab &T, const VQ &m) {
    for (unsigned mm = 16, t = 1; mm >= 1; mm >>= 1, t <<= 1)
        for (unsigned i = 0; i < mm; i++) {
            const V w = bc(T.iw[mm + i]), ws = bc(T.iw_s[mm + i]);
            for (unsigned j = 2 * i * t; j < 2 * i * t + t; j++) {
                const V a = x[j], b = x[j + t];
                x[j] = vadd(a, b, m);
                x[j + t] = vmulc(vsub(a, b, m), w, ws, m);
            }
        }
    const V n = bc(T.ninv), ns = bc(T.ninv_s);
    for (unsigned j = 0; j < 32; j++) x[j] = vmulc(x[j], n, ns, m);
}
// coefficients of the 8 ring elements at p, p + 64, .., p + 7 * 64 into lanes:
// e[k] = coefficient 2k, o[k] = coefficient 2k + 1
TGT inline void vgather(const uint64_t *p, V *e, V *o) {
    const V idx = _mm512_set_epi64(7 * 64, 6 * 64, 5 * 64, 4 * 64, 3 * 64, 2 * 64, 64, 0);
    for (int k = 0; k < 32; k++) {
        e[k] = _mm512_i64gather_epi64(idx, (const long long *)(p + 2 * k), 8);
        o[k] = _mm512_i64gather_epi64(idx, (const long long *)(p + 2 * k + 1), 8);
    }
}
TGT inline void vscatter(uint64_t *p, const V *e, const V *o) {
    const V idx = _mm512_set_epi64(7 * 64, 6 * 64, 5 * 64, 4 * 64, 3 * 64, 2 * 64, 64, 0);
    for (int k = 0; k < 32; k++) {
        _mm512_i64scatter_epi64((long long *)(p + 2 * k), idx, e[k], 8);
        _mm512_i64scatter_epi64((long long *)(p + 2 * k + 1), idx, o[k], 8);
    }
}
// incomplete_ntt_multiplication in the lanes: (e1, o1) (e2, o2) -> (re, ro)
TGT inline void vinc_mul(const V *e1, const V *o1, const V *e2, const V *o2, V *re, V *ro,
                         const Tab &T, const VQ &m) {
    for (int s = 0; s < 32; s++) {
        const V t = vmulc(vmul(o1[s], o2[s], m), bc(T.om[s]), bc(T.om_s[s]), m);
        re[s] = vadd(vmul(e1[s], e2[s], m), t, m);
        ro[s] = vadd(vmul(o1[s], e2[s], m), vmul(e1[s], o2[s], m), m);
    }
}

bool g_ifma = false;

// ------------------------------------------------------------------ key
// A key row's elements in IncompleteNTT form, lane-blocked: block b of 8
// elements is [64][8] (slot-major, element in the lane), so a block loads as
// 64 vectors.
struct Key {
    uint64_t q;
    size_t kappa, m;
    std::vector<uint64_t> a;  // [kappa][m / 8][64][8]  (m % 8 == 0)
    std::vector<uint64_t> plain;  // [kappa][m][64] IncompleteNTT (scalar path)
};

// ------------------------------------------------------------------ commit
// out[j][k] (IncompleteNTT) = sum_i A_j[i] W_k[i], W converted per product
constexpr size_t kCommitChunk = 2048;  // elements per work item

TGT void commit_item_v(const Key &K, const uint64_t *W, size_t m, size_t j, size_t k, size_t i0,
                       size_t i1, uint64_t *acc_out, const Tab &T) {
    const VQ mq = vq(T);
    V acc_e[32], acc_o[32], we[32], wo[32], re[32], ro[32];
    for (int s = 0; s < 32; s++) acc_e[s] = acc_o[s] = _mm512_setzero_si512();
    const uint64_t *Wk = W + k * m * 64;
    for (size_t i = i0; i < i1; i += 8) {
        // the witness elements' IncompleteNTT form, again for this key row
        vgather(Wk + i * 64, we, wo);
        vntt_fwd(we, T, mq);
        vntt_fwd(wo, T, mq);
        const uint64_t *ab = K.a.data() + ((j * (K.m / 8) + i / 8) * 64) * 8;
        V ae[32], ao[32];
        for (int s = 0; s < 32; s++) {
            ae[s] = _mm512_loadu_si512(ab + s * 8);
            ao[s] = _mm512_loadu_si512(ab + (32 + s) * 8);
        }
        vinc_mul(ae, ao, we, wo, re, ro, T, mq);
        for (int s = 0; s < 32; s++) {
            acc_e[s] = vadd(acc_e[s], re[s], mq);
            acc_o[s] = vadd(acc_o[s], ro[s], mq);
        }
    }
    // the sum over the lanes
    alignas(64) uint64_t l[8];
    for (int s = 0; s < 32; s++) {
        _mm512_store_si512((V *)l, acc_e[s]);
        uint64_t t = 0;
        for (int x = 0; x < 8; x++) t = addm(t, l[x], T.q);
        acc_out[s] = t;
        _mm512_store_si512((V *)l, acc_o[s]);
        t = 0;
        for (int x = 0; x < 8; x++) t = addm(t, l[x], T.q);
        acc_out[32 + s] = t;
    }
}

void commit_item_s(const Key &K, const uint64_t *W, size_t m, size_t j, size_t k, size_t i0,
                   size_t i1, uint64_t *acc_out, const Tab &T) {
    uint64_t acc[64] = {0}, w[64], p[64];
    for (size_t i = i0; i < i1; i++) {
        memcpy(w, W + (k * m + i) * 64, sizeof w);
        to_inc(w, T);
        inc_mul(K.plain.data() + (j * K.m + i) * 64, w, p, T);
        for (int c = 0; c < 64; c++) acc[c] = addm(acc[c], p[c], T.q);
    }
    memcpy(acc_out, acc, sizeof acc);
}

void commit(const Key &K, size_t kappa, const uint64_t *W, size_t r, size_t m, uint64_t *out) {
    const Tab &T = tables(K.q);
    const size_t nch = (m + kCommitChunk - 1) / kCommitChunk;
    std::vector<uint64_t> part(kappa * r * nch * 64);
    const bool vec = g_ifma && T.ifma_ok && m % 8 == 0 && !K.a.empty();
    pool().run(kappa * r * nch, [&](size_t it) {
        const size_t c = it % nch, jk = it / nch, j = jk / r, k = jk % r;
        const size_t i0 = c * kCommitChunk, i1 = std::min(m, i0 + kCommitChunk);
        if (vec)
            commit_item_v(K, W, m, j, k, i0, i1, &part[it * 64], T);
        else
            commit_item_s(K, W, m, j, k, i0, i1, &part[it * 64], T);
    });
    for (size_t jk = 0; jk < kappa * r; jk++)
        for (int s = 0; s < 64; s++) {
            uint64_t t = 0;
            for (size_t c = 0; c < nch; c++) t = addm(t, part[(jk * nch + c) * 64 + s], T.q);
            out[jk * 64 + s] = t;
        }
}

// ------------------------------------------------------------------ NTT round trip
TGT void roundtrip_v(uint64_t *W, size_t i0, size_t i1, const Tab &T) {
    const VQ mq = vq(T);
    V e[32], o[32];
    for (size_t i = i0; i < i1; i += 8) {
        vgather(W + i * 64, e, o);
        vntt_fwd(e, T, mq);
        vntt_fwd(o, T, mq);
        vntt_inv(e, T, mq);
        vntt_inv(o, T, mq);
        vscatter(W + i * 64, e, o);
    }
}
void roundtrip(uint64_t *W, size_t count, uint64_t q) {
    const Tab &T = tables(q);
    const size_t ch = 4096, n = (count + ch - 1) / ch;
    pool().run(n, [&](size_t c) {
        const size_t i0 = c * ch, i1 = std::min(count, i0 + ch);
        if (g_ifma && T.ifma_ok && (i1 - i0) % 8 == 0) {
            roundtrip_v(W, i0, i1, T);
        } else {
            for (size_t i = i0; i < i1; i++) {
                to_inc(W + i * 64, T);
                from_inc(W + i * 64, T);
            }
        }
    });
}

// ------------------------------------------------------------------ F_{q^2}, SoA arrays
struct F2 {
    uint64_t a, b;
};
inline F2 f2add(F2 x, F2 y, uint64_t q) { return {addm(x.a, y.a, q), addm(x.b, y.b, q)}; }
inline F2 f2sub(F2 x, F2 y, uint64_t q) { return {subm(x.a, y.a, q), subm(x.b, y.b, q)}; }
inline F2 f2mul(F2 x, F2 y, uint64_t nr, uint64_t q) {
    return {addm(mulm(x.a, y.a, q), mulm(nr, mulm(x.b, y.b, q), q), q),
            addm(mulm(x.a, y.b, q), mulm(x.b, y.a, q), q)};
}

constexpr size_t kMinChunk = 1028;  // MIN_CHUNK_SIZE (fast_norm_check.rs:23)
constexpr size_t kChunks = 120;     // MAX_THREADS (arithmetic.rs:15)

// fn(i0, i1) over the reference's chunks of [0, len), in parallel
void chunked(size_t len, const std::function<void(size_t, size_t)> &fn) {
    const size_t cs = std::max(len / kChunks, kMinChunk);
    const size_t n = (len + cs - 1) / cs;
    pool().run(n, [&](size_t c) { fn(c * cs, std::min(len, c * cs + cs)); });
}

// the elementwise kernels of one round over a chunk [i0, i1), vectorized
// where IFMA is present; a is the first component, b the second (b == null:
// the values are in F_q)
struct Arr {
    uint64_t *a;
    uint64_t *b;
};

TGT void k_mul_v(Arr r, Arr x, Arr y, size_t i0, size_t i1, uint64_t nr, const Tab &T) {
    const VQ m = vq(T);
    const V vnr = bc(nr), vnrs = bc(shoup(nr, T.q));
    for (size_t i = i0; i < i1; i += 8) {
        const V xa = _mm512_loadu_si512(x.a + i), ya = _mm512_loadu_si512(y.a + i);
        if (!x.b) {
            _mm512_storeu_si512(r.a + i, vmul(xa, ya, m));
            if (r.b) _mm512_storeu_si512(r.b + i, _mm512_setzero_si512());
            continue;
        }
        const V xb = _mm512_loadu_si512(x.b + i), yb = _mm512_loadu_si512(y.b + i);
        const V ra = vadd(vmul(xa, ya, m), vmulc(vmul(xb, yb, m), vnr, vnrs, m), m);
        const V rb = vadd(vmul(xa, yb, m), vmul(xb, ya, m), m);
        _mm512_storeu_si512(r.a + i, ra);
        _mm512_storeu_si512(r.b + i, rb);
    }
}
void k_mul_s(Arr r, Arr x, Arr y, size_t i0, size_t i1, uint64_t nr, uint64_t q) {
    for (size_t i = i0; i < i1; i++) {
        if (!x.b) {
            r.a[i] = mulm(x.a[i], y.a[i], q);
            if (r.b) r.b[i] = 0;
            continue;
        }
        const F2 v = f2mul({x.a[i], x.b[i]}, {y.a[i], y.b[i]}, nr, q);
        r.a[i] = v.a;
        r.b[i] = v.b;
    }
}
TGT void k_sub_v(Arr r, Arr x, Arr y, size_t i0, size_t i1, const Tab &T) {
    const VQ m = vq(T);
    for (size_t i = i0; i < i1; i += 8) {
        _mm512_storeu_si512(r.a + i,
                            vsub(_mm512_loadu_si512(x.a + i), _mm512_loadu_si512(y.a + i), m));
        if (x.b)
            _mm512_storeu_si512(
                r.b + i, vsub(_mm512_loadu_si512(x.b + i), _mm512_loadu_si512(y.b + i), m));
    }
}
void k_sub_s(Arr r, Arr x, Arr y, size_t i0, size_t i1, uint64_t q) {
    for (size_t i = i0; i < i1; i++) {
        r.a[i] = subm(x.a[i], y.a[i], q);
        if (x.b) r.b[i] = subm(x.b[i], y.b[i], q);
    }
}
// r = d c + w0, c in F_{q^2}; d and w0 in F_q when their b is null
TGT void k_fma_v(Arr r, Arr d, Arr w0, F2 c, size_t i0, size_t i1, uint64_t nr, const Tab &T) {
    const VQ m = vq(T);
    const uint64_t q = T.q;
    const V ca = bc(c.a), cas = bc(shoup(c.a, q)), cb = bc(c.b), cbs = bc(shoup(c.b, q));
    const uint64_t cbn = mulm(c.b, nr, q);
    const V cbnv = bc(cbn), cbns = bc(shoup(cbn, q));
    for (size_t i = i0; i < i1; i += 8) {
        const V da = _mm512_loadu_si512(d.a + i), wa = _mm512_loadu_si512(w0.a + i);
        if (!d.b) {
            _mm512_storeu_si512(r.a + i, vadd(vmulc(da, ca, cas, m), wa, m));
            _mm512_storeu_si512(r.b + i, vmulc(da, cb, cbs, m));
            continue;
        }
        const V db = _mm512_loadu_si512(d.b + i), wb = _mm512_loadu_si512(w0.b + i);
        const V ra = vadd(vadd(vmulc(da, ca, cas, m), vmulc(db, cbnv, cbns, m), m), wa, m);
        const V rb = vadd(vadd(vmulc(da, cb, cbs, m), vmulc(db, ca, cas, m), m), wb, m);
        _mm512_storeu_si512(r.a + i, ra);
        _mm512_storeu_si512(r.b + i, rb);
    }
}
void k_fma_s(Arr r, Arr d, Arr w0, F2 c, size_t i0, size_t i1, uint64_t nr, uint64_t q) {
    for (size_t i = i0; i < i1; i++) {
        const F2 dv = {d.a[i], d.b ? d.b[i] : 0}, wv = {w0.a[i], w0.b ? w0.b[i] : 0};
        const F2 v = f2add(f2mul(dv, c, nr, q), wv, q);
        r.a[i] = v.a;
        r.b[i] = v.b;
    }
}
TGT F2 k_sum_v(Arr x, size_t i0, size_t i1, const Tab &T) {
    const VQ m = vq(T);
    V sa = _mm512_setzero_si512(), sb = _mm512_setzero_si512();
    for (size_t i = i0; i < i1; i += 8) {
        sa = vadd(sa, _mm512_loadu_si512(x.a + i), m);
        if (x.b) sb = vadd(sb, _mm512_loadu_si512(x.b + i), m);
    }
    alignas(64) uint64_t la[8], lb[8];
    _mm512_store_si512((V *)la, sa);
    _mm512_store_si512((V *)lb, sb);
    F2 s = {0, 0};
    for (int l = 0; l < 8; l++) s = f2add(s, {la[l], lb[l]}, T.q);
    return s;
}
F2 k_sum_s(Arr x, size_t i0, size_t i1, uint64_t q) {
    F2 s = {0, 0};
    for (size_t i = i0; i < i1; i++) s = f2add(s, {x.a[i], x.b ? x.b[i] : 0}, q);
    return s;
}

// sum_fast_par: each chunk's sum reduced, the residues summed (u64, as the
// reference returns them), then reduced once for the message
F2 k_sum_s(Arr x, size_t i0, size_t i1, uint64_t q);
F2 sum_par(Arr x, size_t len, const Tab &T, bool vec) {
    const size_t cs = std::max(len / kChunks, kMinChunk), n = (len + cs - 1) / cs;
    std::vector<F2> part(n);
    pool().run(n, [&](size_t c) {
        const size_t i0 = c * cs, i1 = std::min(len, i0 + cs);
        const size_t i8 = vec ? i0 + (i1 - i0) / 8 * 8 : i0;
        F2 a = i8 > i0 ? k_sum_v(x, i0, i8, T) : F2{0, 0};
        if (i1 > i8) a = f2add(a, k_sum_s(x, i8, i1, T.q), T.q);
        part[c] = a;
    });
    uint64_t a = 0, b = 0;
    for (const F2 &p : part) a += p.a, b += p.b;
    return {a % T.q, b % T.q};
}

// one row's norm check (fast_norm_check.rs:135-199, 352-399) for every
// repetition, outputs in the oracle's layout
void sumcheck_row(const uint64_t *w, size_t L, const uint64_t *cs, size_t reps, uint64_t nr,
                  size_t rows, size_t row, uint64_t claim, uint64_t *msgs, uint64_t *finals,
                  uint64_t *evals, int *ok, const Tab &T) {
    const uint64_t q = T.q;
    const bool vec = g_ifma && T.ifma_ok;
    const size_t nrounds = (size_t)__builtin_ctzll(L), half0 = L / 2;
    // w_f_0, w_diff, w_f_2 live across rounds (fast_norm_check.rs:143-151)
    std::vector<uint64_t> f0a(half0), f0b(half0), da(half0), db(half0), f2a(half0), f2b(half0);
    for (size_t rep = 0; rep < reps; rep++) {
        std::vector<uint64_t> cura, curb;  // the folded witness (none before round 1)
        F2 cl = {claim, 0};
        int good = 1;
        size_t len = L;
        for (size_t t = 0; t < nrounds; t++) {
            const size_t h = len / 2;
            const bool fq = cura.empty();
            Arr w0 = fq ? Arr{const_cast<uint64_t *>(w), nullptr} : Arr{cura.data(), curb.data()};
            Arr w1 = fq ? Arr{const_cast<uint64_t *>(w) + h, nullptr}
                        : Arr{cura.data() + h, curb.data() + h};
            Arr F0 = {f0a.data(), fq ? nullptr : f0b.data()};
            Arr D = {da.data(), fq ? nullptr : db.data()};
            Arr F2_ = {f2a.data(), fq ? nullptr : f2b.data()};
            // w_f_1 is allocated and zeroed every round (:164)
            std::vector<uint64_t> g1a(h, 0), g1b(h, 0);
            Arr G1 = {g1a.data(), g1b.data()};
            Arr G1q = {g1a.data(), fq ? nullptr : g1b.data()};
            // each chunk: whole vectors of 8, then the rest one by one
            auto par = [&](const std::function<void(size_t, size_t)> &v,
                           const std::function<void(size_t, size_t)> &s) {
                chunked(h, [&](size_t i0, size_t i1) {
                    const size_t i8 = vec ? i0 + (i1 - i0) / 8 * 8 : i0;
                    if (i8 > i0) v(i0, i8);
                    if (i1 > i8) s(i8, i1);
                });
            };
            par([&](size_t a, size_t b) { k_mul_v(F0, w0, w0, a, b, nr, T); },
                [&](size_t a, size_t b) { k_mul_s(F0, w0, w0, a, b, nr, q); });
            par([&](size_t a, size_t b) { k_sub_v(D, w1, w0, a, b, T); },
                [&](size_t a, size_t b) { k_sub_s(D, w1, w0, a, b, q); });
            par([&](size_t a, size_t b) { k_mul_v(G1q, w0, D, a, b, nr, T); },
                [&](size_t a, size_t b) { k_mul_s(G1q, w0, D, a, b, nr, q); });
            par([&](size_t a, size_t b) { k_mul_v(F2_, D, D, a, b, nr, T); },
                [&](size_t a, size_t b) { k_mul_s(F2_, D, D, a, b, nr, q); });
            const F2 s0 = sum_par(F0, h, T, vec), s1 = sum_par(G1q, h, T, vec),
                     s2 = sum_par(F2_, h, T, vec);
            uint64_t *mo = msgs + (((rep * nrounds + t) * rows + row) * 3) * 2;
            mo[0] = s0.a, mo[1] = s0.b, mo[2] = s1.a, mo[3] = s1.b, mo[4] = s2.a, mo[5] = s2.b;
            // the verifier: 2 f0 + 2 f1 + f2 == claim, then the claim update
            const F2 lhs = f2add(f2add(f2add(s0, s0, q), f2add(s1, s1, q), q), s2, q);
            if (lhs.a != cl.a || lhs.b != cl.b) good = 0;
            const uint64_t *cp = cs + (rep * nrounds + t) * 2;
            const F2 c = {cp[0] % q, cp[1] % q};
            cl = f2add(f2add(s0, f2mul(f2add(s1, s1, q), c, nr, q), q),
                       f2mul(s2, f2mul(c, c, nr, q), nr, q), q);
            // fold: w_f_1 = (w1 - w0) c + w0 becomes the witness
            par([&](size_t a, size_t b) { k_fma_v(G1, D, w0, c, a, b, nr, T); },
                [&](size_t a, size_t b) { k_fma_s(G1, D, w0, c, a, b, nr, q); });
            cura.swap(g1a);
            curb.swap(g1b);
            len = h;
        }
        const F2 ev = {cura[0], curb[0]};
        const F2 ev2 = f2mul(ev, ev, nr, q);
        if (ev2.a != cl.a || ev2.b != cl.b) good = 0;
        finals[(rep * rows + row) * 2] = cl.a;
        finals[(rep * rows + row) * 2 + 1] = cl.b;
        evals[(rep * rows + row) * 2] = ev.a;
        evals[(rep * rows + row) * 2 + 1] = ev.b;
        ok[rep * rows + row] = good;
    }
}

TGT uint64_t k_sumsq_v(const uint64_t *w, size_t i0, size_t i1, const Tab &T);

// sum_sq_par (norm_check_1)
uint64_t claim_par(const uint64_t *w, size_t L, const Tab &T) {
    const size_t cs = std::max(L / kChunks, kMinChunk), n = (L + cs - 1) / cs;
    std::vector<uint64_t> part(n);
    const bool vec = g_ifma && T.ifma_ok;
    pool().run(n, [&](size_t c) {
        const size_t i0 = c * cs, i1 = std::min(L, c * cs + cs);
        const size_t i8 = vec ? i0 + (i1 - i0) / 8 * 8 : i0;
        uint64_t s = i8 > i0 ? k_sumsq_v(w, i0, i8, T) : 0;
        for (size_t i = i8; i < i1; i++) s = addm(s, mulm(w[i], w[i], T.q), T.q);
        part[c] = s;
    });
    uint64_t s = 0;
    for (uint64_t p : part) s += p;
    return s % T.q;
}

// tensor of F_{q^2} challenges, the first the most significant index bit
std::vector<F2> f2_tensor(const uint64_t *cs2, size_t n, uint64_t nr, uint64_t q) {
    std::vector<F2> out(1, F2{1 % q, 0});
    for (size_t t = 0; t < n; t++) {
        const F2 c = {cs2[2 * t] % q, cs2[2 * t + 1] % q}, om = f2sub({1 % q, 0}, c, q);
        std::vector<F2> nx(out.size() * 2);
        for (size_t v = 0; v < out.size(); v++) {
            nx[2 * v] = f2mul(out[v], om, nr, q);
            nx[2 * v + 1] = f2mul(out[v], c, nr, q);
        }
        out.swap(nx);
    }
    return out;
}

// An F_{q^2} ring element (two R_q components, IncompleteNTT) times another:
// (A0 + A1 Y)(B0 + B1 Y) = (A0 B0 + nr A1 B1) + (A0 B1 + A1 B0) Y, four ring
// products
void f2ring_mul(const uint64_t *a, const uint64_t *b, uint64_t *out, uint64_t nr, const Tab &T) {
    uint64_t p00[64], p11[64], p01[64], p10[64];
    inc_mul(a, b, p00, T);
    inc_mul(a + 64, b + 64, p11, T);
    inc_mul(a, b + 64, p01, T);
    inc_mul(a + 64, b, p10, T);
    for (int c = 0; c < 64; c++) {
        out[c] = addm(p00[c], mulm(nr, p11[c], T.q), T.q);
        out[64 + c] = addm(p01[c], p10[c], T.q);
    }
}

// norm_check_3 (fast_norm_check.rs:401-441) over F_{q^2}: the series
// [1 - c, c] (x) .. (x) [base] expanded layer by layer with a ring product per
// entry (complete_power_series), then rhs_j = sum_i ps_i W_j[i] with W_j[i]
// converted per product.  rhs [rows][2][64] IncompleteNTT.  Layers of 8 or
// more entries are held lane-blocked, [len / 8][2][64][8], and multiplied in
// the lanes.
TGT void nc3_expand_block(const uint64_t *layer, size_t nb, size_t e, F2 k0, F2 k1, uint64_t *nx,
                          uint64_t nr, const Tab &T) {
    const VQ m = vq(T);
    const V nrv = bc(nr), nrs = bc(shoup(nr, T.q));
    const size_t a = e / nb, b = e % nb;
    const F2 k = a ? k1 : k0;
    // the constant's IncompleteNTT form: k in every even slot, 0 in the odd
    V ke[32], ko[32], kbe[32], kbo[32], le[32], lo[32], l2e[32], l2o[32];
    for (int s = 0; s < 32; s++) {
        ke[s] = bc(k.a), kbe[s] = bc(k.b);
        ko[s] = kbo[s] = _mm512_setzero_si512();
    }
    const uint64_t *src = layer + b * 2 * 64 * 8;
    for (int s = 0; s < 32; s++) {
        le[s] = _mm512_loadu_si512(src + s * 8);
        lo[s] = _mm512_loadu_si512(src + (32 + s) * 8);
        l2e[s] = _mm512_loadu_si512(src + (64 + s) * 8);
        l2o[s] = _mm512_loadu_si512(src + (96 + s) * 8);
    }
    V p00e[32], p00o[32], p11e[32], p11o[32], p01e[32], p01o[32], p10e[32], p10o[32];
    vinc_mul(ke, ko, le, lo, p00e, p00o, T, m);
    vinc_mul(kbe, kbo, l2e, l2o, p11e, p11o, T, m);
    vinc_mul(ke, ko, l2e, l2o, p01e, p01o, T, m);
    vinc_mul(kbe, kbo, le, lo, p10e, p10o, T, m);
    uint64_t *dst = nx + (a * nb + b) * 2 * 64 * 8;
    for (int s = 0; s < 32; s++) {
        _mm512_storeu_si512(dst + s * 8, vadd(p00e[s], vmulc(p11e[s], nrv, nrs, m), m));
        _mm512_storeu_si512(dst + (32 + s) * 8, vadd(p00o[s], vmulc(p11o[s], nrv, nrs, m), m));
        _mm512_storeu_si512(dst + (64 + s) * 8, vadd(p01e[s], p10e[s], m));
        _mm512_storeu_si512(dst + (96 + s) * 8, vadd(p01o[s], p10o[s], m));
    }
}

void nc3_expand_v(const uint64_t *layer, size_t len, F2 k0, F2 k1, uint64_t *nx, uint64_t nr,
                  const Tab &T) {
    const size_t nb = len / 8;
    pool().run(2 * nb, [&](size_t e) { nc3_expand_block(layer, nb, e, k0, k1, nx, nr, T); });
}

TGT void nc3_dot_v(const uint64_t *ps, const uint64_t *Wj, size_t i0, size_t i1, uint64_t *acc_out,
                   const Tab &T) {
    const VQ m = vq(T);
    V acc[4][32], we[32], wo[32], pe[32], po[32], re[32], ro[32];
    for (int c = 0; c < 4; c++)
        for (int s = 0; s < 32; s++) acc[c][s] = _mm512_setzero_si512();
    for (size_t i = i0; i < i1; i += 8) {
        vgather(Wj + i * 64, we, wo);
        vntt_fwd(we, T, m);
        vntt_fwd(wo, T, m);
        const uint64_t *src = ps + (i / 8) * 2 * 64 * 8;
        for (int comp = 0; comp < 2; comp++) {
            for (int s = 0; s < 32; s++) {
                pe[s] = _mm512_loadu_si512(src + (comp * 64 + s) * 8);
                po[s] = _mm512_loadu_si512(src + (comp * 64 + 32 + s) * 8);
            }
            vinc_mul(pe, po, we, wo, re, ro, T, m);
            for (int s = 0; s < 32; s++) {
                acc[2 * comp][s] = vadd(acc[2 * comp][s], re[s], m);
                acc[2 * comp + 1][s] = vadd(acc[2 * comp + 1][s], ro[s], m);
            }
        }
    }
    alignas(64) uint64_t l[8];
    for (int c = 0; c < 4; c++)
        for (int s = 0; s < 32; s++) {
            _mm512_store_si512((V *)l, acc[c][s]);
            uint64_t t = 0;
            for (int x = 0; x < 8; x++) t = addm(t, l[x], T.q);
            acc_out[c * 32 + s] = t;  // [comp][even | odd]: comp * 64 + half * 32 + s
        }
}

void norm_check_3(const uint64_t *W, size_t rows, size_t m, const uint64_t *cs2, uint64_t nr,
                  uint64_t *rhs, const Tab &T) {
    const uint64_t q = T.q;
    const bool vec = g_ifma && T.ifma_ok && m % 8 == 0;
    const size_t nring = (size_t)__builtin_ctzll(m);
    const std::vector<F2> tc = f2_tensor(cs2 + 2 * nring, 6, nr, q);
    // base = conj(tensor(coefficient challenges)) per component, IncompleteNTT
    std::vector<uint64_t> layer(2 * 64);
    for (int comp = 0; comp < 2; comp++) {
        uint64_t *b = layer.data() + comp * 64;
        for (int k = 0; k < 64; k++) {
            const uint64_t v = comp ? tc[k].b : tc[k].a;
            if (k == 0)
                b[0] = v;
            else
                b[64 - k] = v ? q - v : 0;
        }
        to_inc(b, T);
    }
    bool blocked = false;
    // tensors applied last first (kronecker_product of [t0, t1] with the layer)
    for (size_t t = nring; t-- > 0;) {
        const F2 c = {cs2[2 * t] % q, cs2[2 * t + 1] % q}, om = f2sub({1 % q, 0}, c, q);
        const size_t len = layer.size() / 128;
        std::vector<uint64_t> nx(layer.size() * 2);
        if (blocked) {
            nc3_expand_v(layer.data(), len, om, c, nx.data(), nr, T);
        } else {
            for (size_t e = 0; e < 2 * len; e++) {
                const size_t a = e / len, j = e % len;
                const F2 k = a ? c : om;
                uint64_t kc[128] = {0};
                // the constant's IncompleteNTT form: the constant in every even slot
                for (int s = 0; s < 32; s++) kc[s] = k.a, kc[64 + s] = k.b;
                f2ring_mul(kc, layer.data() + j * 128, nx.data() + (a * len + j) * 128, nr, T);
            }
            if (vec && 2 * len >= 8) {
                // to the lane-blocked layout
                std::vector<uint64_t> bl(nx.size());
                for (size_t e = 0; e < 2 * len; e++)
                    for (int x = 0; x < 128; x++) bl[((e / 8) * 128 + x) * 8 + e % 8] = nx[e * 128 + x];
                nx.swap(bl);
                blocked = true;
            }
        }
        layer.swap(nx);
    }
    const size_t nch = (m + kCommitChunk - 1) / kCommitChunk;
    for (size_t j = 0; j < rows; j++) {
        std::vector<uint64_t> part(nch * 128);
        pool().run(nch, [&](size_t ch) {
            const size_t i0 = ch * kCommitChunk, i1 = std::min(m, i0 + kCommitChunk);
            uint64_t acc[128] = {0}, w[64], p[64];
            if (blocked) {
                nc3_dot_v(layer.data(), W + j * m * 64, i0, i1, acc, T);
            } else {
                for (size_t i = i0; i < i1; i++) {
                    memcpy(w, W + (j * m + i) * 64, sizeof w);
                    to_inc(w, T);
                    for (int comp = 0; comp < 2; comp++) {
                        inc_mul(layer.data() + i * 128 + comp * 64, w, p, T);
                        for (int c = 0; c < 64; c++)
                            acc[comp * 64 + c] = addm(acc[comp * 64 + c], p[c], q);
                    }
                }
            }
            memcpy(&part[ch * 128], acc, sizeof acc);
        });
        for (int c = 0; c < 128; c++) {
            uint64_t t = 0;
            for (size_t ch = 0; ch < nch; ch++) t = addm(t, part[ch * 128 + c], q);
            rhs[j * 128 + c] = t;
        }
    }
}

// sum_sq_par in the lanes
TGT uint64_t k_sumsq_v(const uint64_t *w, size_t i0, size_t i1, const Tab &T) {
    const VQ m = vq(T);
    V s = _mm512_setzero_si512();
    for (size_t i = i0; i < i1; i += 8) {
        const V x = _mm512_loadu_si512(w + i);
        s = vadd(s, vmul(x, x, m), m);
    }
    alignas(64) uint64_t l[8];
    _mm512_store_si512((V *)l, s);
    uint64_t t = 0;
    for (int x = 0; x < 8; x++) t = addm(t, l[x], T.q);
    return t;
}

}  // namespace

extern "C" {

/* 1 when the AVX-512 IFMA lanes are in use, 0 for the scalar variant */
int cb_variant(void) { return g_ifma ? 1 : 0; }

/* threads of the pool (0: every hardware thread); use_ifma 0 forces the
 * scalar variant; returns the thread count */
int cb_init(int threads, int use_ifma) {
    __builtin_cpu_init();
    g_ifma = use_ifma && __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512ifma");
    unsigned n = threads > 0 ? (unsigned)threads : std::thread::hardware_concurrency();
    pool().resize(n ? (int)n : 1);
    return pool().size();
}

/* IncompleteNTT form of count elements [count][64], in place (tests) */
void cb_to_incomplete(uint64_t *W, size_t count, uint64_t q) {
    const Tab &T = tables(q);
    for (size_t i = 0; i < count; i++) to_inc(W + i * 64, T);
}

/* forward + inverse NTT of every element, in place */
void cb_ntt_roundtrip(uint64_t *W, size_t count, uint64_t q) { roundtrip(W, count, q); }

/* key rows A [kappa][m][64] Coefficient, held in IncompleteNTT form */
void *cb_key_create(const uint64_t *A, size_t kappa, size_t m, uint64_t q) {
    const Tab &T = tables(q);
    Key *K = new Key();
    K->q = q;
    K->kappa = kappa;
    K->m = m;
    K->plain.assign(A, A + kappa * m * 64);
    pool().run(kappa * m, [&](size_t x) { to_inc(K->plain.data() + x * 64, T); });
    if (m % 8 == 0) {
        K->a.resize(kappa * m * 64);
        for (size_t j = 0; j < kappa; j++)
            for (size_t i = 0; i < m; i++)
                for (int s = 0; s < 64; s++)
                    K->a[((j * (m / 8) + i / 8) * 64 + s) * 8 + i % 8] =
                        K->plain[(j * m + i) * 64 + s];
        if (g_ifma && T.ifma_ok) std::vector<uint64_t>().swap(K->plain);
    }
    return K;
}
void cb_key_destroy(void *k) { delete static_cast<Key *>(k); }

/* parallel_dot_series_matrix with the key's first kappa rows and m elements:
 * W [r][m][64] Coefficient -> out [kappa][r][64] IncompleteNTT */
int cb_commit(const void *key, size_t kappa, const uint64_t *W, size_t r, size_t m,
              uint64_t *out) {
    const Key *K = static_cast<const Key *>(key);
    if (kappa > K->kappa || m > K->m) return -1;
    commit(*K, kappa, W, r, m, out);
    return 0;
}

/* norm_check_1/2/3 over F_{q^2} (the oracle's or_sumcheck with fq2 = 1, and
 * or_norm_check_3_fq2): w [rows][L] Coefficient rows (L = m * 64),
 * cs [reps][log2 L][2]; outputs as the oracle's; rhs [reps][rows][2][64] or
 * NULL */
void cb_norm_check(const uint64_t *w, size_t rows, size_t L, const uint64_t *cs, size_t reps,
                   uint64_t nr, uint64_t q, uint64_t *claims, uint64_t *msgs, uint64_t *finals,
                   uint64_t *evals, int *ok, uint64_t *rhs) {
    const Tab &T = tables(q);
    const size_t nrounds = (size_t)__builtin_ctzll(L);
    // norm_check_1, then the rows (per-row threads sharing the challenges;
    // each row's passes use every thread)
    for (size_t row = 0; row < rows; row++) claims[row] = claim_par(w + row * L, L, T);
    for (size_t row = 0; row < rows; row++)
        sumcheck_row(w + row * L, L, cs, reps, nr, rows, row, claims[row], msgs, finals, evals, ok,
                     T);
    if (rhs)
        for (size_t rep = 0; rep < reps; rep++) {
            std::vector<uint64_t> o(rows * 128);
            norm_check_3(w, rows, L / 64, cs + rep * nrounds * 2, nr, o.data(), T);
            memcpy(rhs + rep * rows * 128, o.data(), o.size() * 8);
        }
}

/* decompose_matrix_by_chunks (arithmetic.rs:1461-1593) with an explicit
 * radix (0: the reference's 2^ceil(bw / parts)): W [r][m][64] Coefficient ->
 * out [r * parts][m][64]; returns the radix */
uint64_t cb_decompose(const uint64_t *W, size_t r, size_t m, size_t parts, uint64_t radix,
                      uint64_t q, uint64_t *out) {
    const size_t n = r * m * 64;
    const uint64_t half = q / 2;
    // the bit width of the largest |w| (split_into_positive_and_negative,
    // determine_bit_width), in chunks over the threads
    const size_t ch = 1 << 16, nch = (n + ch - 1) / ch;
    std::vector<uint64_t> ors(nch);
    pool().run(nch, [&](size_t c) {
        uint64_t o = 0;
        for (size_t i = c * ch; i < std::min(n, c * ch + ch); i++) {
            const uint64_t x = W[i];
            o |= 2 * x <= q ? x : q - x;
        }
        ors[c] = o;
    });
    uint64_t o = 0;
    for (uint64_t x : ors) o |= x;
    int lb = 0;
    if (!radix) {
        const int bw = o ? 64 - __builtin_clzll(o) : 0;
        lb = (int)((bw + parts - 1) / parts);
        radix = 1ULL << lb;
    } else {
        lb = __builtin_ctzll(radix);
    }
    (void)half;
    const uint64_t mask = radix - 1;
    pool().run(r * m, [&](size_t x) {
        const size_t rho = x / m, i = x % m;
        for (int c = 0; c < 64; c++) {
            const uint64_t v = W[x * 64 + c];
            const uint64_t pos = 2 * v <= q ? v : 0, neg = 2 * v <= q ? 0 : q - v;
            for (size_t t = 0; t < parts; t++) {
                const uint64_t dp = (lb * t < 64) ? (pos >> (lb * t)) & mask : 0;
                const uint64_t dn = (lb * t < 64) ? (neg >> (lb * t)) & mask : 0;
                out[((rho * parts + t) * m + i) * 64 + c] = subm(dp, dn, q);
            }
        }
    });
    return radix;
}

/* verify_decomp (decomp.rs:99-112): sum_t radix^t rhs_dec[.][rho parts + t]
 * == rhs[.][rho] for rhs_dec [kappa][r * parts][64]; returns mismatches */
size_t cb_verify_decomp(const uint64_t *rhs_dec, const uint64_t *rhs, size_t kappa, size_t r,
                        size_t parts, uint64_t radix, uint64_t q) {
    size_t bad = 0;
    for (size_t j = 0; j < kappa; j++)
        for (size_t rho = 0; rho < r; rho++)
            for (int c = 0; c < 64; c++) {
                uint64_t s = 0, p = 1;
                for (size_t t = 0; t < parts; t++, p = mulm(p, radix % q, q))
                    s = addm(s, mulm(p, rhs_dec[((j * r * parts) + rho * parts + t) * 64 + c], q), q);
                if (s != rhs[(j * r + rho) * 64 + c]) bad++;
            }
    return bad;
}

}  // extern "C"
