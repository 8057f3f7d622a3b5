// server.hip -- continuous-batching service of commit + norm-check jobs
// against one resident key.
//
// A job is one witness row of m ring elements (a power of two) to be
// committed under the key's first kappa rows and norm-checked over its m*d
// coefficients with its own challenges (the per-job half of a prover step:
// parallel_dot_series_matrix, arithmetic.rs:459-505, and norm_check_1/2/3,
// fast_norm_check.rs:282-462).
//
// A scheduler thread owns all device work:
//   admission  jobs are taken in arrival order while their witness fits the
//              HBM arena (first fit); the witness is copied from the host or
//              generated from the job's seed straight into its arena slot on
//              the admission stream;
//   batching   whenever a batch slot is free, every admitted job that is not
//              yet running (up to the slot's capacity) forms the next batch:
//              sorted by length, its jobs are grouped 16 to an MFMA row block
//              for ONE commit (the key's slot chunks are read once for the
//              whole batch) and by length for multi-row norm checks;
//   overlap    with two slots, one batch runs while the next is formed and
//              launched on the other slot's stream, so their kernels overlap
//              on the device; jobs that arrive meanwhile join the next batch;
//   results    each batch's output records are copied to pinned host memory;
//              when its event completes the jobs are finished, their arena
//              space is freed, and further jobs can be admitted.
// With max_batch_jobs = 1 and one slot the same code serves every job alone
// as it arrives: the plain per-call path the batched service is compared to.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "handles.hpp"

using namespace salsaa;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

constexpr size_t kMaxRounds = 32;  // L = m d <= 2^32

struct Job {
    uint64_t id = 0;
    size_t m = 0, nrounds = 0;
    uint64_t seed = 0, bound = 0;
    const uint64_t *h_w = nullptr;  // host witness, or null: generate from seed
    size_t w_off = 0;               // arena slot (ring elements)
    double t_submit = 0, t_admit = 0, t_launch = 0, t_done = 0;
    int state = 0;  // 0 queued, 1 admitted, 2 running, 3 done, 4 failed
    size_t batch = 0, batch_jobs = 0;
    std::vector<uint64_t> rec;
};

// first-fit allocator over [0, cap) (units: ring elements)
struct Arena {
    std::map<size_t, size_t> free_;  // offset -> length
    explicit Arena(size_t cap = 0) {
        if (cap) free_[0] = cap;
    }
    bool alloc(size_t n, size_t *off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n, o = it->first + n;
                free_.erase(it);
                if (rest) free_[o] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n) {
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

// one in-flight batch
struct Slot {
    Context c;
    bool busy = false;
    hipEvent_t done = nullptr;
    std::vector<Job *> jobs;  // launch order (sorted by length)
    size_t batch_id = 0;
    // device buffers
    int8_t *d_planes = nullptr;
    double *d_partial = nullptr;
    uint64_t *d_cs = nullptr, *d_out = nullptr;
    PlaneJob *d_pjobs = nullptr;
    CommitJob *d_cjobs = nullptr;
    FinJob *d_fjobs = nullptr;
    uint2 *d_table = nullptr;
    unsigned long long *d_off = nullptr, *d_m = nullptr, *d_wrow = nullptr;
    NcScratch nc;
    // pinned host staging
    uint64_t *h_out = nullptr;
    char *h_tab = nullptr;
};

}  // namespace

struct salsaa_server {
    salsaa_ctx *ctx = nullptr;
    const salsaa_key *key = nullptr;
    salsaa_server_opts o{};
    size_t kappa = 0, rec = 0;  // rows committed, output record length (u64)
    size_t off_com = 0, off_claim = 0, off_msg = 0, off_fin = 0, off_ev = 0, off_rhs = 0,
           off_ok = 0;
    uint64_t *d_arena = nullptr;
    Arena arena;
    Context admit;  // admission stream
    hipEvent_t admitted = nullptr;
    Slot slots[2];
    int nslots = 0;
    // capacities per slot
    size_t cap_planes = 0, cap_partial = 0, cap_table = 0, cap_cs = 0;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job *> queue;     // submitted, not admitted
    std::deque<Job *> admitted_; // admitted, not running
    std::unordered_map<uint64_t, Job *> jobs;
    uint64_t next_id = 1;
    size_t next_batch = 1;
    size_t running = 0, outstanding = 0;
    bool stop = false;
    int error = SALSAA_OK;
    std::string error_msg;
};

namespace {

size_t ktiles_of(size_t m) { return plane_ktiles(m); }

struct Plan {
    std::vector<Job *> js;  // sorted by length, longest first
    size_t planes = 0, partial = 0, table = 0, cs = 0, nc_m = 0, nc_a = 0, nc_b = 0;
};

// exact device needs of a batch made of `js` (sorted in place)
void plan_needs(salsaa_server *s, Plan &p) {
    std::stable_sort(p.js.begin(), p.js.end(), [](Job *a, Job *b) { return a->m > b->m; });
    const size_t n = p.js.size(), kc = s->o.kchunk;
    p.planes = p.partial = p.table = p.cs = p.nc_m = p.nc_a = p.nc_b = 0;
    for (size_t j = 0; 16 * j < n; j++) {
        const size_t rows = std::min<size_t>(16, n - 16 * j);
        const size_t kt = ktiles_of(p.js[16 * j]->m), rt = (7 * rows + 15) / 16;
        const size_t ks = (kt + kc - 1) / kc;
        p.planes += (size_t)64 * kt * rt * 1024;
        p.partial += ks * 64 * s->kappa * rows;
        p.table += ks * 64;
    }
    for (size_t i = 0; i < n;) {
        size_t k = i;
        while (k < n && p.js[k]->m == p.js[i]->m) k++;
        size_t a, b, mm;
        nc_scratch_need(k - i, p.js[i]->m * kD, s->o.reps, &mm, &a, &b);
        p.nc_m = std::max(p.nc_m, mm);
        p.nc_a = std::max(p.nc_a, a);
        p.nc_b = std::max(p.nc_b, b);
        i = k;
    }
    for (Job *j : p.js) p.cs += s->o.reps * j->nrounds * 2;
}

bool plan_fits(salsaa_server *s, const Plan &p) {
    return p.planes <= s->cap_planes && p.partial <= s->cap_partial && p.table <= s->cap_table &&
           p.cs <= s->cap_cs && p.nc_m <= s->slots[0].nc.macc_bytes &&
           p.nc_a <= s->slots[0].nc.bufA_bytes && p.nc_b <= s->slots[0].nc.bufB_bytes &&
           p.js.size() <= s->o.max_batch_jobs;
}

void fail(salsaa_server *s, int st) {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->error == SALSAA_OK) {
        s->error = st;
        s->error_msg = salsaa_last_error();
    }
}

// enqueue the whole batch on the slot's stream; the caller holds no lock
int launch_batch(salsaa_server *s, Slot &sl, Plan &p) {
    Context *c = &sl.c;
    const size_t n = p.js.size(), kc = s->o.kchunk;
    const Key &k = s->key->k;
    // host tables in pinned staging, one copy to the device per table
    char *h = sl.h_tab;
    size_t hoff = 0;
    auto stage = [&](const void *src, size_t bytes, void *dst) -> int {
        if (!bytes) return SALSAA_OK;
        memcpy(h + hoff, src, bytes);
        SALSAA_HIP(hipMemcpyAsync(dst, h + hoff, bytes, hipMemcpyHostToDevice, c->stream));
        hoff += (bytes + 255) / 256 * 256;
        return SALSAA_OK;
    };
    std::vector<PlaneJob> pj;
    std::vector<CommitJob> cj;
    std::vector<FinJob> fj;
    std::vector<uint2> table;
    std::vector<unsigned long long> off(n), mm(n), wrow(n);
    for (size_t i = 0; i < n; i++) {
        off[i] = p.js[i]->w_off;
        mm[i] = p.js[i]->m;
        wrow[i] = p.js[i]->w_off * kD;
    }
    size_t pbytes = 0, poff = 0, pnblk = 0, fblk = 0;
    std::vector<int> ks;
    for (size_t j = 0; 16 * j < n; j++) {
        const int rows = (int)std::min<size_t>(16, n - 16 * j);
        const int kt = (int)ktiles_of(p.js[16 * j]->m), rt = (7 * rows + 15) / 16;
        PlaneJob a;
        a.dst = sl.d_planes + pbytes;
        a.first = (int)(16 * j);
        a.rows = rows;
        a.ktiles = kt;
        a.blk0 = (int)pnblk;
        pj.push_back(a);
        pnblk += (size_t)rows * kt / 2;
        CommitJob b;
        b.WB = a.dst;
        b.partial = sl.d_partial + poff;
        b.ktiles = kt;
        b.rtiles = rt;
        b.ntile = rt;
        b.rows = rows;
        cj.push_back(b);
        ks.push_back((kt + (int)kc - 1) / (int)kc);
        FinJob f;
        f.partial = b.partial;
        f.out = sl.d_out + (16 * j) * s->rec + s->off_com;
        f.osj = kD;
        f.osk = (long long)s->rec;
        f.ks = ks.back();
        f.rows = rows;
        f.blk0 = (int)fblk;
        fj.push_back(f);
        pbytes += (size_t)64 * kt * rt * 1024;
        poff += (size_t)ks.back() * 64 * s->kappa * rows;
        fblk += s->kappa * (size_t)rows;
    }
    int maxks = 0;
    for (int x : ks) maxks = std::max(maxks, x);
    for (int z = 0; z < maxks; z++) {
        unsigned active = 0;
        while (active < ks.size() && ks[active] > z) active++;
        for (unsigned sl_ = 0; sl_ < 64; sl_++)
            for (unsigned j = 0; j < active; j++) table.push_back(make_uint2(j, ((unsigned)z << 6) | sl_));
    }
    SALSAA_TRY(stage(pj.data(), pj.size() * sizeof(PlaneJob), sl.d_pjobs));
    SALSAA_TRY(stage(cj.data(), cj.size() * sizeof(CommitJob), sl.d_cjobs));
    SALSAA_TRY(stage(fj.data(), fj.size() * sizeof(FinJob), sl.d_fjobs));
    SALSAA_TRY(stage(table.data(), table.size() * sizeof(uint2), sl.d_table));
    SALSAA_TRY(stage(off.data(), n * sizeof(unsigned long long), sl.d_off));
    SALSAA_TRY(stage(mm.data(), n * sizeof(unsigned long long), sl.d_m));
    SALSAA_TRY(stage(wrow.data(), n * sizeof(unsigned long long), sl.d_wrow));
    // the jobs' witnesses are written on the admission stream
    SALSAA_HIP(hipStreamWaitEvent(c->stream, s->admitted, 0));
    // challenges from the jobs' seeds, contiguous in launch order
    std::vector<size_t> cs_off(n);
    size_t co = 0;
    for (size_t i = 0; i < n; i++) {
        cs_off[i] = co;
        SALSAA_TRY(launch_challenges(c, sl.d_cs + co, s->o.reps, p.js[i]->nrounds, s->o.fq2,
                                     p.js[i]->seed));
        co += s->o.reps * p.js[i]->nrounds * 2;
    }
    SALSAA_TRY(launch_ntt_planes_batch(c, s->d_arena, sl.d_off, sl.d_m, sl.d_pjobs, pj.size(),
                                       pnblk));
    const size_t mt = (7 * std::min<size_t>(n, 16) + 15) / 16;
    if (k.d_mhat)  // a CRS key generated inside the commit
        SALSAA_TRY(launch_commit_crs_batch(c, k.d_mhat, s->kappa, sl.d_cjobs, sl.d_table,
                                           table.size(), kc, mt));
    else
        SALSAA_TRY(launch_commit_mfma_batch(c, k.d_AB, s->kappa, k.a_kt, k.a_ktiles, sl.d_cjobs,
                                            sl.d_table, table.size(), kc, mt));
    SALSAA_TRY(launch_commit_finalize_batch(c, sl.d_fjobs, fj.size(), fblk, s->kappa));
    for (size_t i = 0; i < n;) {
        size_t e = i;
        while (e < n && p.js[e]->m == p.js[i]->m) e++;
        const size_t nr = p.js[i]->nrounds, reps = s->o.reps;
        uint64_t *r0 = sl.d_out + i * s->rec;
        NcOut no;
        no.claims = r0 + s->off_claim;
        no.claim_row = s->rec;
        no.msgs = r0 + s->off_msg;
        no.msg_row = s->rec;
        no.msg_rep = nr * 6;
        no.msg_t = 6;
        no.finals = r0 + s->off_fin;
        no.evals = r0 + s->off_ev;
        no.fe_row = s->rec;
        no.fe_rep = 2;
        no.rhs = r0 + s->off_rhs;
        no.rhs_row = s->rec;
        no.rhs_rep = 2 * kD;
        no.ok = reinterpret_cast<int32_t *>(r0 + s->off_ok);
        no.ok_row = 2 * s->rec;
        no.ok_rep = 1;
        SALSAA_TRY(launch_norm_check_ex(c, s->d_arena, sl.d_wrow + i, e - i, p.js[i]->m * kD,
                                        sl.d_cs + cs_off[i], reps * nr * 2, reps, s->o.fq2,
                                        s->o.nr, no, 0, &sl.nc));
        i = e;
    }
    SALSAA_HIP(hipMemcpyAsync(sl.h_out, sl.d_out, n * s->rec * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, c->stream));
    SALSAA_HIP(hipEventRecord(sl.done, c->stream));
    return SALSAA_OK;
}

// admit queued jobs in arrival order while they fit; enqueue their witnesses
int admit_jobs(salsaa_server *s) {
    std::vector<Job *> take;
    {
        std::lock_guard<std::mutex> g(s->mu);
        while (!s->queue.empty()) {
            Job *j = s->queue.front();
            size_t off;
            if (!s->arena.alloc(j->m, &off)) break;
            j->w_off = off;
            j->state = 1;
            j->t_admit = now_s();
            s->queue.pop_front();
            take.push_back(j);
        }
    }
    if (take.empty()) return SALSAA_OK;
    for (Job *j : take) {
        uint64_t *dst = s->d_arena + j->w_off * kD;
        if (j->h_w)
            SALSAA_HIP(hipMemcpyAsync(dst, j->h_w, j->m * kD * sizeof(uint64_t),
                                      hipMemcpyHostToDevice, s->admit.stream));
        else
            SALSAA_TRY(launch_sample_bounded(&s->admit, dst, j->m * kD, j->bound, j->seed,
                                             0xB));
    }
    SALSAA_HIP(hipEventRecord(s->admitted, s->admit.stream));
    std::lock_guard<std::mutex> g(s->mu);
    for (Job *j : take) s->admitted_.push_back(j);
    return SALSAA_OK;
}

int start_batches(salsaa_server *s) {
    for (int si = 0; si < s->nslots; si++) {
        Slot &sl = s->slots[si];
        if (sl.busy) continue;
        Plan p;
        {
            std::lock_guard<std::mutex> g(s->mu);
            if (s->admitted_.empty()) return SALSAA_OK;
            // longest prefix of the admitted jobs that fits the slot
            while (!s->admitted_.empty()) {
                p.js.push_back(s->admitted_.front());
                plan_needs(s, p);
                if (!plan_fits(s, p)) {
                    Job *last = s->admitted_.front();
                    p.js.erase(std::find(p.js.begin(), p.js.end(), last));
                    break;
                }
                s->admitted_.pop_front();
            }
            if (p.js.empty()) {
                // a single job that does not fit a slot can never run
                Job *j = s->admitted_.front();
                s->admitted_.pop_front();
                j->state = 4;
                s->arena.release(j->w_off, j->m);
                s->outstanding--;
                continue;
            }
            plan_needs(s, p);
            sl.batch_id = s->next_batch++;
            const double t = now_s();
            for (Job *j : p.js) {
                j->state = 2;
                j->t_launch = t;
                j->batch = sl.batch_id;
                j->batch_jobs = p.js.size();
            }
            s->running++;
        }
        sl.jobs = p.js;
        sl.busy = true;
        const int st = launch_batch(s, sl, p);
        if (st != SALSAA_OK) return st;
    }
    return SALSAA_OK;
}

int finish_batches(salsaa_server *s, bool *progress) {
    for (int si = 0; si < s->nslots; si++) {
        Slot &sl = s->slots[si];
        if (!sl.busy) continue;
        const hipError_t e = hipEventQuery(sl.done);
        if (e == hipErrorNotReady) continue;
        SALSAA_HIP(e);
        const double t = now_s();
        std::lock_guard<std::mutex> g(s->mu);
        for (size_t i = 0; i < sl.jobs.size(); i++) {
            Job *j = sl.jobs[i];
            j->rec.assign(sl.h_out + i * s->rec, sl.h_out + (i + 1) * s->rec);
            j->t_done = t;
            j->state = 3;
            s->arena.release(j->w_off, j->m);
            s->outstanding--;
        }
        sl.jobs.clear();
        sl.busy = false;
        s->running--;
        *progress = true;
    }
    s->cv.notify_all();
    return SALSAA_OK;
}

void loop(salsaa_server *s) {
    if (hipSetDevice(s->ctx->c.device) != hipSuccess) {
        fail(s, SALSAA_ERR_HIP);
        return;
    }
    for (;;) {
        {
            std::unique_lock<std::mutex> g(s->mu);
            if (s->stop && s->outstanding == 0) return;
            if (s->error != SALSAA_OK) return;
            if (s->queue.empty() && s->admitted_.empty() && s->running == 0)
                s->cv.wait_for(g, std::chrono::milliseconds(2));
        }
        int st = admit_jobs(s);
        if (st == SALSAA_OK) st = start_batches(s);
        bool progress = false;
        if (st == SALSAA_OK) st = finish_batches(s, &progress);
        if (st != SALSAA_OK) {
            fail(s, st);
            return;
        }
        if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

void destroy(salsaa_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->th.joinable()) s->th.join();
    for (int si = 0; si < s->nslots; si++) {
        Slot &sl = s->slots[si];
        if (sl.c.stream) (void)hipStreamSynchronize(sl.c.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        (void)hipFree(sl.d_planes);
        (void)hipFree(sl.d_partial);
        (void)hipFree(sl.d_cs);
        (void)hipFree(sl.d_out);
        (void)hipFree(sl.d_pjobs);
        (void)hipFree(sl.d_cjobs);
        (void)hipFree(sl.d_fjobs);
        (void)hipFree(sl.d_table);
        (void)hipFree(sl.d_off);
        (void)hipFree(sl.d_m);
        (void)hipFree(sl.d_wrow);
        (void)hipFree(sl.nc.macc);
        (void)hipFree(sl.nc.bufA);
        (void)hipFree(sl.nc.bufB);
        if (sl.h_out) (void)hipHostFree(sl.h_out);
        if (sl.h_tab) (void)hipHostFree(sl.h_tab);
        if (sl.c.stream) (void)hipStreamDestroy(sl.c.stream);
    }
    if (s->admit.stream) {
        (void)hipStreamSynchronize(s->admit.stream);
        (void)hipStreamDestroy(s->admit.stream);
    }
    if (s->admitted) (void)hipEventDestroy(s->admitted);
    (void)hipFree(s->d_arena);
    for (auto &kv : s->jobs) delete kv.second;
    delete s;
}


template <typename T>
int dalloc(T **p, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, (count ? count : 1) * sizeof(T));
    if (e != hipSuccess) {
        set_error(std::string("server allocation failed: ") + hipGetErrorString(e));
        return SALSAA_ERR_NOMEM;
    }
    *p = static_cast<T *>(q);
    return SALSAA_OK;
}

int create(salsaa_server *s) {
    const Context &base = s->ctx->c;
    const salsaa_server_opts &o = s->o;
    // output record: commitment, claim, msgs, finals, evals, rhs, ok
    s->off_com = 0;
    s->off_claim = s->kappa * kD;
    s->off_msg = s->off_claim + 1;
    s->off_fin = s->off_msg + o.reps * kMaxRounds * 6;
    s->off_ev = s->off_fin + o.reps * 2;
    s->off_rhs = s->off_ev + o.reps * 2;
    s->off_ok = s->off_rhs + o.reps * 2 * kD;
    s->rec = s->off_ok + (o.reps + 1) / 2;
    SALSAA_TRY(dalloc(&s->d_arena, o.arena_elems * kD));
    s->arena = Arena(o.arena_elems);
    share_context(base, s->admit);
    SALSAA_HIP(hipStreamCreateWithFlags(&s->admit.stream, hipStreamNonBlocking));
    SALSAA_HIP(hipEventCreateWithFlags(&s->admitted, hipEventDisableTiming));
    SALSAA_HIP(hipEventRecord(s->admitted, s->admit.stream));
    // slot capacities for a batch of up to max_batch_elems ring elements in
    // at most max_batch_jobs jobs; the exact needs of each batch are checked
    // against them when it is formed (plan_fits), so these only bound the
    // batch size.  Any single job of up to max_batch_elems fits.
    const size_t E = o.max_batch_elems, J = o.max_batch_jobs, kc = o.kchunk;
    const size_t blocks = (J + 15) / 16 + 1;
    s->cap_planes = (size_t)1024 * E + (size_t)64 * 1024 * 7 * 2 * blocks;
    s->cap_partial = ((E / 64) / kc + 2 * J + 2) * 64 * s->kappa * 16;
    s->cap_table = ((E / 64) / kc + 2 * J + 2) * 64;
    s->cap_cs = J * o.reps * kMaxRounds * 2;
    // norm check scratch: a group's folded buffers scale with its total
    // coefficients (at most E d), its message accumulators with its rows
    size_t nc_m, nc_a, nc_b;
    nc_scratch_need(1, E * kD, o.reps, &nc_m, &nc_a, &nc_b);
    nc_b += o.reps * J * 16 * 2 * sizeof(double);
    // messages of every round, the claims and the rows' Gram sums (16 x 16)
    nc_m = (o.reps * kMaxRounds * J * 6 + J + J * 256) * sizeof(unsigned long long);
    s->nslots = o.inflight;
    for (int si = 0; si < s->nslots; si++) {
        Slot &sl = s->slots[si];
        share_context(base, sl.c);
        SALSAA_HIP(hipStreamCreateWithFlags(&sl.c.stream, hipStreamNonBlocking));
        SALSAA_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        SALSAA_TRY(dalloc(&sl.d_planes, s->cap_planes));
        SALSAA_TRY(dalloc(&sl.d_partial, s->cap_partial));
        SALSAA_TRY(dalloc(&sl.d_cs, s->cap_cs));
        SALSAA_TRY(dalloc(&sl.d_out, J * s->rec));
        SALSAA_TRY(dalloc(&sl.d_pjobs, J));
        SALSAA_TRY(dalloc(&sl.d_cjobs, J));
        SALSAA_TRY(dalloc(&sl.d_fjobs, J));
        SALSAA_TRY(dalloc(&sl.d_table, s->cap_table));
        SALSAA_TRY(dalloc(&sl.d_off, J));
        SALSAA_TRY(dalloc(&sl.d_m, J));
        SALSAA_TRY(dalloc(&sl.d_wrow, J));
        SALSAA_TRY(dalloc(&sl.nc.macc, nc_m / 8 + 1));
        SALSAA_TRY(dalloc(&sl.nc.bufA, nc_a / 8 + 1));
        SALSAA_TRY(dalloc(&sl.nc.bufB, nc_b / 8 + 1));
        sl.nc.macc_bytes = nc_m;
        sl.nc.bufA_bytes = nc_a;
        sl.nc.bufB_bytes = nc_b;
        SALSAA_HIP(hipHostMalloc((void **)&sl.h_out, J * s->rec * sizeof(uint64_t), 0));
        // pinned staging for one batch's tables, each rounded up to 256 bytes
        const size_t tab = 3 * J * 64 + s->cap_table * sizeof(uint2) + 3 * J * 8 + 8 * 256;
        SALSAA_HIP(hipHostMalloc((void **)&sl.h_tab, tab, 0));
    }
    s->th = std::thread(loop, s);
    return SALSAA_OK;
}

}  // namespace

extern "C" {

int salsaa_server_create(salsaa_ctx *ctx, const salsaa_key *key, size_t kappa,
                         const salsaa_server_opts *opts, salsaa_server **out) {
    CTX_D64(ctx, "salsaa_server_create");
    SALSAA_CHECK(key && opts && out, "null key, options or output");
    SALSAA_CHECK(key->k.ctx == &ctx->c, "key belongs to another context");
    SALSAA_CHECK(key->k.d_AB != nullptr || key->k.d_mhat != nullptr,
                 "the server needs a key with MFMA planes or a generated CRS key");
    SALSAA_CHECK(kappa >= 1 && kappa <= key->k.kappa, "kappa must be 1..key rows");
    const salsaa_server_opts &o = *opts;
    SALSAA_CHECK(o.arena_elems >= 1 && o.max_batch_jobs >= 1 && o.max_batch_jobs <= 4096 &&
                     o.max_batch_elems >= 1,
                 "server capacities must be positive");
    SALSAA_CHECK(o.inflight >= 1 && o.inflight <= 2, "one or two batches in flight");
    SALSAA_CHECK(o.reps >= 1 && o.reps <= 16, "repetitions must be 1..16");
    SALSAA_CHECK(o.kchunk >= 2 && o.kchunk <= 1024 && o.kchunk % 2 == 0,
                 "K chunk must be even, 2..1024");
    SALSAA_CHECK(o.nr < ctx->c.q, "non-residue must be reduced mod q");
    salsaa_server *s = new (std::nothrow) salsaa_server();
    if (!s) return SALSAA_ERR_NOMEM;
    s->ctx = ctx;
    s->key = key;
    s->o = o;
    s->kappa = kappa;
    const int st = create(s);
    if (st != SALSAA_OK) {
        destroy(s);
        return st;
    }
    *out = s;
    return SALSAA_OK;
}

int salsaa_server_submit(salsaa_server *s, size_t m, const uint64_t *h_witness, uint64_t seed,
                         uint64_t bound, uint64_t *id) {
    SALSAA_CHECK(s && id, "null server or id");
    SALSAA_CHECK(m >= 1 && (m & (m - 1)) == 0, "witness length must be a power of two");
    SALSAA_CHECK(m <= s->key->k.m, "witness longer than the key");
    SALSAA_CHECK(m <= s->o.arena_elems && m <= s->o.max_batch_elems,
                 "witness larger than the server's arena or batch capacity");
    SALSAA_CHECK((size_t)__builtin_ctzll(m * kD) <= kMaxRounds, "witness too long");
    Job *j = new (std::nothrow) Job();
    if (!j) return SALSAA_ERR_NOMEM;
    j->m = m;
    j->nrounds = (size_t)__builtin_ctzll(m * kD);
    j->seed = seed;
    j->bound = bound;
    j->h_w = h_witness;
    j->t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->error != SALSAA_OK) {
            delete j;
            set_error("server stopped after an error: " + s->error_msg);
            return s->error;
        }
        j->id = s->next_id++;
        s->jobs[j->id] = j;
        s->queue.push_back(j);
        s->outstanding++;
        *id = j->id;
    }
    s->cv.notify_all();
    return SALSAA_OK;
}

int salsaa_server_poll(salsaa_server *s, uint64_t id, uint64_t *h_record, double *times,
                       uint64_t *batch_info) {
    SALSAA_CHECK(s != nullptr, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->jobs.find(id);
    SALSAA_CHECK(it != s->jobs.end(), "unknown job id");
    Job *j = it->second;
    if (j->state == 4) {
        set_error("job does not fit the server's batch capacity");
        return SALSAA_ERR_ARG;
    }
    if (j->state != 3) {
        if (s->error != SALSAA_OK) {
            set_error("server stopped after an error: " + s->error_msg);
            return s->error;
        }
        return 0;
    }
    if (h_record) memcpy(h_record, j->rec.data(), j->rec.size() * sizeof(uint64_t));
    if (times) {
        times[0] = j->t_submit;
        times[1] = j->t_admit;
        times[2] = j->t_launch;
        times[3] = j->t_done;
    }
    if (batch_info) {
        batch_info[0] = j->batch;
        batch_info[1] = j->batch_jobs;
    }
    return 1;
}

int salsaa_server_drain(salsaa_server *s) {
    SALSAA_CHECK(s != nullptr, "null server");
    std::unique_lock<std::mutex> g(s->mu);
    s->cv.wait(g, [&] { return s->outstanding == 0 || s->error != SALSAA_OK; });
    if (s->error != SALSAA_OK) {
        set_error("server stopped after an error: " + s->error_msg);
        return s->error;
    }
    return SALSAA_OK;
}

int salsaa_server_layout(const salsaa_server *s, size_t *offsets) {
    SALSAA_CHECK(s && offsets, "null server or output");
    offsets[0] = s->off_com;
    offsets[1] = s->off_claim;
    offsets[2] = s->off_msg;
    offsets[3] = s->off_fin;
    offsets[4] = s->off_ev;
    offsets[5] = s->off_rhs;
    offsets[6] = s->off_ok;
    offsets[7] = s->rec;
    return SALSAA_OK;
}

int salsaa_server_destroy(salsaa_server *s) {
    if (!s) return SALSAA_OK;
    (void)hipSetDevice(s->ctx->c.device);
    destroy(s);
    return SALSAA_OK;
}

}  // extern "C"
