// batch.hip -- many independent prover steps (NTT forward + inverse, commit,
// norm check) in a handful of grouped launches, optionally replayed from a
// captured hipGraph.
//
// Instances are sorted by witness length (descending) once, at plan time.
//   * NTT: the witnesses sit back to back in one arena, each in whole tiles of
//     kPlaneTile ring elements (zero beyond its length), so the forward and inverse
//     transforms are one pass over every ring element, which also writes the
//     commit's digit planes: every tile belongs to one instance, and a table
//     gives each tile's place in the planes.
//   * Commit: every instance is one witness row committed against the same
//     resident key, so consecutive instances form row blocks of 16 (one MFMA
//     N block) and the whole batch is ONE commit: each block of the launch
//     streams one (slot, K chunk) of the key once for up to 16 instances, and
//     the blocks that share that key chunk run side by side on one XCD.  A
//     plain per-instance commit reads its key prefix once per instance.
//   * Norm check: instances of equal length are consecutive, so each length
//     class is one multi-row norm check with per-row challenges.  The classes
//     run on kNcStreams streams of their own (forked from and joined back into
//     the context's stream with events, which the hipGraph capture follows),
//     so the short launches at the end of one class overlap the others.
// The commit's K chunk and the norm check's first-pass width are chosen by
// timing candidates on the device when the plan is made (the autotuner).
#include <algorithm>
#include <new>
#include <numeric>
#include <vector>

#include "handles.hpp"

namespace salsaa {
namespace {

// Two: HIP maps streams onto a few hardware queues (four by default); the
// context's stream and two norm-check streams get queues of their own, while
// a third norm-check stream shared one with the second and serialised its
// classes behind it.
constexpr int kNcStreams = 2;

__global__ void k_batch_zero(uint64_t *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
__global__ void k_batch_fill8(uint8_t *p, size_t n, uint8_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct Outputs {
    uint64_t *com = nullptr, *claims = nullptr, *msgs = nullptr, *fin = nullptr, *ev = nullptr,
             *rhs = nullptr;
    int32_t *ok = nullptr;
};

template <typename T>
int dmalloc(T **p, size_t count) {
    void *q = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) {
        set_error(std::string("batch allocation failed: ") + hipGetErrorString(e));
        return SALSAA_ERR_NOMEM;
    }
    *p = static_cast<T *>(q);
    return SALSAA_OK;
}

template <typename T>
int upload(Context *ctx, T *dst, const std::vector<T> &src) {
    if (src.empty()) return SALSAA_OK;
    SALSAA_HIP(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice,
                              ctx->stream));
    SALSAA_HIP(hipStreamSynchronize(ctx->stream));
    return SALSAA_OK;
}

}  // namespace
}  // namespace salsaa

using namespace salsaa;

struct salsaa_batch {
    salsaa_ctx *ctx = nullptr;
    const salsaa_key *key = nullptr;
    size_t n = 0, reps = 0;
    int fq2 = 0;
    uint64_t nr = 0;
    size_t krows = 0;  // key rows every commit is computed with (the largest kappa)
    std::vector<size_t> m, kappa;          // per instance, caller's order
    std::vector<size_t> order, pos;        // sorted position -> instance, instance -> position
    std::vector<size_t> off_w, off_c, off_msg, nrounds;  // per sorted position
    size_t tot_w = 0, tot_c = 0, tot_msg = 0;
    uint64_t *d_wit = nullptr, *d_cs = nullptr;
    Outputs out[2];  // [0] batched step, [1] plain per-call step
    // witness -> planes
    std::vector<PlaneJob> pjobs;
    size_t pnblk = 0;
    PlaneJob *d_pjobs = nullptr;
    unsigned long long *d_off = nullptr, *d_m = nullptr;
    unsigned long long *d_wrow = nullptr;  // witness start of each instance (u64)
    PlaneTile *d_ptiles = nullptr;         // per arena tile: where its planes go
    int8_t *d_planes = nullptr;
    size_t planes_bytes = 0;
    // the arena's short copy, written by every round trip and read by the
    // norm checks of the same step (its format bytes start at 2, not kept)
    ShortCopy sc;
    // commit
    // Jobs of 16 instances in commit order: instances grouped by the key row
    // tiles their kappa needs (at most 16, 8 or 4: one kernel shape each), in
    // length order within a group; cpos[p] is the commit position of arena
    // position p, and commitments are stored in commit order.
    struct CommitClass {
        size_t kappa = 0, job0 = 0, njobs = 0, max_tiles = 0, t0 = 0, nblk = 0, fnblk = 0;
    };
    std::vector<CommitClass> cclass;
    std::vector<size_t> cpos;
    size_t njobs = 0, kchunk = 0;
    std::vector<int> job_ktiles, job_rows;
    std::vector<size_t> job_planes;  // byte offset of each job's planes
    std::vector<size_t> job_cfirst;  // commit position of each job's first row
    double *d_partial = nullptr;
    size_t partial_cap = 0;  // doubles
    CommitJob *d_cjobs = nullptr;
    uint2 *d_table = nullptr;
    size_t table_cap = 0;
    // commit schedule (stored keys; the classes of a mixed launch share each
    // key chunk through L2): 0 one launch per class; 1 one launch over every
    // class at the largest shape; 2 one launch over the classes of at most 8
    // key row tiles at their largest shape, and the others' own launches.
    // The mixed tables (all classes, then the small ones) follow the
    // per-class tables; the tuner's best time of each schedule
    int commit_mixed = 0;
    size_t mixed_t0[2] = {0, 0}, mixed_nblk[2] = {0, 0};
    int mixed_tm[2] = {4, 2};
    float t_commit_mode[3] = {0.f, 0.f, 0.f};
    FinJob *d_fjobs = nullptr;
    // norm check
    struct Group {
        size_t first, rows, L;
    };
    std::vector<Group> groups;
    std::vector<int> group_stream;       // which norm-check stream runs each class
    Context nc_ctx[kNcStreams];          // the context's tables on streams of their own
    NcScratch nc[kNcStreams];
    hipEvent_t ev_fork = nullptr, ev_fork2 = nullptr, ev_done[kNcStreams] = {};
    // the step's NTT pass in two launches split at arena tile split_tile (the
    // start of norm-check group split_group; the arena runs longest first):
    // split_mode 1 transforms the shorter classes and part of the next one
    // (tiles from split_tile on) first, 2 the longer classes (tiles before
    // split_tile); the first part's norm checks run beside the second launch
    // (split_mode 0: one launch)
    size_t split_tile = 0, split_group = 0;
    int split_mode = 0;
    int cpt_first = 0;
    float t_commit = 0.f, t_nc = 0.f;
    // the NTT pass's exchange between register phases (0 LDS, 1 DPP quad
    // permutes) and the tuner's times of the whole pass with each
    int ntt_xch = 0;
    float t_ntt[2] = {0.f, 0.f};
    // the NTT pass's persistent blocks (0: one per CU) and the tuner's
    // times of the whole pass at each candidate
    unsigned ntt_grid = 0;
    // the second launch's persistent blocks in a split step (0: ntt_grid);
    // fewer leave CUs to the first part's norm checks beside it.  The tuner's
    // candidates and its times of the whole step with each
    unsigned ntt_grid2 = 0;
    unsigned grid2_cand[3] = {0, 0, 0};
    float t_grid2[3] = {0.f, 0.f, 0.f};
    unsigned grid_cand[3] = {0, 0, 0};
    float t_grid[3] = {0.f, 0.f, 0.f};
    // plain path
    std::vector<salsaa_key *> views;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int alloc_outputs(salsaa_batch *b, Outputs &o) {
    SALSAA_TRY(dmalloc(&o.com, b->n * b->krows * kD));
    SALSAA_TRY(dmalloc(&o.claims, b->n));
    SALSAA_TRY(dmalloc(&o.msgs, b->tot_msg));
    SALSAA_TRY(dmalloc(&o.fin, b->n * b->reps * 2));
    SALSAA_TRY(dmalloc(&o.ev, b->n * b->reps * 2));
    SALSAA_TRY(dmalloc(&o.rhs, b->n * b->reps * 2 * kD));
    SALSAA_TRY(dmalloc(&o.ok, b->n * b->reps));
    return SALSAA_OK;
}

void free_outputs(Outputs &o) {
    (void)hipFree(o.com);
    (void)hipFree(o.claims);
    (void)hipFree(o.msgs);
    (void)hipFree(o.fin);
    (void)hipFree(o.ev);
    (void)hipFree(o.rhs);
    (void)hipFree(o.ok);
    o = Outputs();
}

// Commit job tables for K chunk kc, per class: block order (z, s, job) with
// the jobs active at chunk z (a prefix, lengths being sorted) fastest.
int build_commit_tables(salsaa_batch *b, size_t kc) {
    Context *ctx = &b->ctx->c;
    std::vector<CommitJob> cj(b->njobs);
    std::vector<FinJob> fj(b->njobs);
    std::vector<uint2> table;
    size_t poff = 0;
    for (auto &cc : b->cclass) {
        std::vector<int> ks(cc.njobs);
        int maxks = 0;
        size_t fblk = 0;
        for (size_t jj = 0; jj < cc.njobs; jj++) {
            const size_t j = cc.job0 + jj;
            ks[jj] = (int)((b->job_ktiles[j] + kc - 1) / kc);
            maxks = std::max(maxks, ks[jj]);
            CommitJob &c = cj[j];
            c.WB = b->d_planes + b->job_planes[j];
            c.ktiles = b->job_ktiles[j];
            c.rows = b->job_rows[j];
            c.rtiles = (7 * c.rows + 15) / 16;
            c.ntile = c.rtiles;
            c.kappa = (int)cc.kappa;
            const size_t rtA = (7 * cc.kappa + 15) / 16;
            c.tm = rtA <= 4 ? 1 : rtA <= 8 ? 2 : 4;
            c.partial = b->d_partial + poff;
            FinJob &f = fj[j];
            f.partial = c.partial;
            f.out = b->out[0].com + b->job_cfirst[j] * b->krows * kD;
            f.osj = kD;
            f.osk = (long long)(b->krows * kD);
            f.ks = ks[jj];
            f.rows = c.rows;
            f.blk0 = (int)fblk;
            poff += (size_t)ks[jj] * 64 * cc.kappa * c.rows;
            fblk += cc.kappa * (size_t)c.rows;
        }
        cc.t0 = table.size();
        for (int z = 0; z < maxks; z++) {
            unsigned active = 0;
            while (active < cc.njobs && ks[active] > z) active++;
            for (unsigned s = 0; s < 64; s++)
                for (unsigned j = 0; j < active; j++)
                    table.push_back(make_uint2(j, ((unsigned)z << 6) | s));
        }
        cc.nblk = table.size() - cc.t0;
        cc.fnblk = fblk;
    }
    // the mixed launches: block order (z, s, class, job), so every block of
    // one key chunk is consecutive whatever its class; [0] every class, [1]
    // the classes of at most 8 key row tiles
    {
        int maxks = 0;
        std::vector<int> ks(b->njobs);
        for (size_t j = 0; j < b->njobs; j++) {
            ks[j] = (int)((b->job_ktiles[j] + kc - 1) / kc);
            maxks = std::max(maxks, ks[j]);
        }
        for (int mt = 0; mt < 2; mt++) {
            b->mixed_t0[mt] = table.size();
            for (int z = 0; z < maxks; z++)
                for (unsigned s = 0; s < 64; s++)
                    for (const auto &cc : b->cclass) {
                        if (mt == 1 && (7 * cc.kappa + 15) / 16 > 8) continue;
                        for (size_t jj = 0; jj < cc.njobs && ks[cc.job0 + jj] > z; jj++)
                            table.push_back(make_uint2((unsigned)(cc.job0 + jj), ((unsigned)z << 6) | s));
                    }
            b->mixed_nblk[mt] = table.size() - b->mixed_t0[mt];
        }
    }
    SALSAA_CHECK(poff <= b->partial_cap, "commit partial buffer too small");
    SALSAA_CHECK(table.size() <= b->table_cap, "commit table too small");
    SALSAA_TRY(upload(ctx, b->d_cjobs, cj));
    SALSAA_TRY(upload(ctx, b->d_table, table));
    SALSAA_TRY(upload(ctx, b->d_fjobs, fj));
    b->kchunk = kc;
    return SALSAA_OK;
}

int run_planes(salsaa_batch *b) {
    return launch_ntt_planes_batch(&b->ctx->c, b->d_wit, b->d_off, b->d_m, b->d_pjobs,
                                   b->pjobs.size(), b->pnblk);
}

// one commit launch per class (its kappa and kernel shape) or, with a stored
// key and a mixed schedule, one launch over every class (1) or over the
// classes of at most 8 key row tiles beside the others' own launches (2);
// then one finalize per class
int run_commit(salsaa_batch *b) {
    const Key &k = b->key->k;
    const int mode = k.d_mhat ? 0 : b->commit_mixed;
    if (mode && b->mixed_nblk[mode - 1])
        SALSAA_TRY(launch_commit_mfma_mixed(&b->ctx->c, k.d_AB, k.a_kt, k.a_ktiles, b->d_cjobs,
                                            b->d_table + b->mixed_t0[mode - 1],
                                            b->mixed_nblk[mode - 1], b->kchunk,
                                            b->mixed_tm[mode - 1]));
    for (const auto &cc : b->cclass) {
        // a class in the mixed launch
        if (mode == 1 || (mode == 2 && (7 * cc.kappa + 15) / 16 <= 8)) continue;
        if (k.d_mhat)  // a CRS key generated inside the commit
            SALSAA_TRY(launch_commit_crs_batch(&b->ctx->c, k.d_mhat, cc.kappa, b->d_cjobs + cc.job0,
                                               b->d_table + cc.t0, cc.nblk, b->kchunk,
                                               cc.max_tiles));
        else
            SALSAA_TRY(launch_commit_mfma_batch(&b->ctx->c, k.d_AB, cc.kappa, k.a_kt, k.a_ktiles,
                                                b->d_cjobs + cc.job0, b->d_table + cc.t0, cc.nblk,
                                                b->kchunk, cc.max_tiles));
    }
    for (const auto &cc : b->cclass)
        SALSAA_TRY(launch_commit_finalize_batch(&b->ctx->c, b->d_fjobs + cc.job0, cc.njobs,
                                                cc.fnblk, cc.kappa));
    return SALSAA_OK;
}

// Enqueues the norm checks of length classes [g0, g1) on their streams after
// the work already on the context's stream (marked by `ev`).
int norm_checks_fork_range(salsaa_batch *b, int cpt, const ShortCopy *sc, size_t g0, size_t g1,
                           hipEvent_t ev) {
    Context *ctx = &b->ctx->c;
    const Outputs &o = b->out[0];
    SALSAA_HIP(hipEventRecord(ev, ctx->stream));
    for (int k = 0; k < kNcStreams; k++)
        SALSAA_HIP(hipStreamWaitEvent(b->nc_ctx[k].stream, ev, 0));
    for (size_t gi = g0; gi < g1; gi++) {
        const auto &g = b->groups[gi];
        const int k = b->group_stream[gi];
        const size_t p = g.first, nr = b->nrounds[p];
        NcOut no;
        no.claims = o.claims + p;
        no.claim_row = 1;
        no.msgs = o.msgs + b->off_msg[p];
        no.msg_row = b->reps * nr * 6;
        no.msg_rep = nr * 6;
        no.msg_t = 6;
        no.finals = o.fin + p * b->reps * 2;
        no.evals = o.ev + p * b->reps * 2;
        no.fe_row = b->reps * 2;
        no.fe_rep = 2;
        no.rhs = o.rhs + p * b->reps * 2 * kD;
        no.rhs_row = b->reps * 2 * kD;
        no.rhs_rep = 2 * kD;
        no.ok = o.ok + p * b->reps;
        no.ok_row = b->reps;
        no.ok_rep = 1;
        SALSAA_TRY(launch_norm_check_ex(&b->nc_ctx[k], b->d_wit, b->d_wrow + p, g.rows, g.L,
                                        b->d_cs + b->off_c[p], b->reps * nr * 2, b->reps, b->fq2,
                                        b->nr, no, cpt, &b->nc[k], sc));
    }
    return SALSAA_OK;
}

int norm_checks_mark_done(salsaa_batch *b) {
    for (int k = 0; k < kNcStreams; k++)
        SALSAA_HIP(hipEventRecord(b->ev_done[k], b->nc_ctx[k].stream));
    return SALSAA_OK;
}

// Enqueues every norm check on their streams after the work already on the
// context's stream; norm_checks_join makes the context's stream wait for them.
int norm_checks_fork(salsaa_batch *b, int cpt, const ShortCopy *sc) {
    SALSAA_TRY(norm_checks_fork_range(b, cpt, sc, 0, b->groups.size(), b->ev_fork));
    return norm_checks_mark_done(b);
}

int norm_checks_join(salsaa_batch *b) {
    for (int k = 0; k < kNcStreams; k++)
        SALSAA_HIP(hipStreamWaitEvent(b->ctx->c.stream, b->ev_done[k], 0));
    return SALSAA_OK;
}

int run_norm_checks(salsaa_batch *b, int cpt, const ShortCopy *sc) {
    SALSAA_TRY(norm_checks_fork(b, cpt, sc));
    return norm_checks_join(b);
}

int run_fwd_planes(salsaa_batch *b, int roundtrip) {
    return launch_ntt_fwd_planes(&b->ctx->c, b->d_wit, b->tot_w, b->d_planes, b->d_ptiles,
                                 roundtrip, roundtrip ? &b->sc : nullptr, b->ntt_xch, b->ntt_grid);
}

// the round trip over arena tiles [t0, t1) (the pass's tables and short copy
// are indexed from the range's first tile)
int run_rt_range(salsaa_batch *b, size_t t0, size_t t1, unsigned grid) {
    ShortCopy sc{b->sc.t + t0 * kShortTile, b->sc.t8 + t0 * kShortTile,
                 b->sc.fmt + t0 * (kShortTile / kShortSlice)};
    return launch_ntt_fwd_planes(&b->ctx->c, b->d_wit + t0 * kShortTile,
                                 (t1 - t0) * kPlaneTile, b->d_planes, b->d_ptiles + t0, 1, &sc,
                                 b->ntt_xch, grid);
}

// one pass: forward NTT, the commit's digit planes, the inverse NTT and the
// short copy; then the commit on the context's stream while the norm checks
// (which need only the Coefficient witnesses, or their short copy) run beside
// it on theirs.  With a split, the pass first covers one part of the length
// classes, whose norm checks then run beside the pass over the rest.
int run_step(salsaa_batch *b) {
    const size_t nt = b->tot_w / kPlaneTile, G = b->groups.size();
    if (b->split_mode == 0) {
        SALSAA_TRY(run_fwd_planes(b, 1));
        SALSAA_TRY(norm_checks_fork(b, b->cpt_first, &b->sc));
    } else if (b->split_mode == 1) {
        SALSAA_TRY(run_rt_range(b, b->split_tile, nt, b->ntt_grid));
        SALSAA_TRY(norm_checks_fork_range(b, b->cpt_first, &b->sc, b->split_group, G, b->ev_fork));
        SALSAA_TRY(run_rt_range(b, 0, b->split_tile, b->ntt_grid2 ? b->ntt_grid2 : b->ntt_grid));
        SALSAA_TRY(norm_checks_fork_range(b, b->cpt_first, &b->sc, 0, b->split_group,
                                          b->ev_fork2));
        SALSAA_TRY(norm_checks_mark_done(b));
    } else {
        SALSAA_TRY(run_rt_range(b, 0, b->split_tile, b->ntt_grid));
        SALSAA_TRY(norm_checks_fork_range(b, b->cpt_first, &b->sc, 0, b->split_group, b->ev_fork));
        SALSAA_TRY(run_rt_range(b, b->split_tile, nt, b->ntt_grid2 ? b->ntt_grid2 : b->ntt_grid));
        SALSAA_TRY(norm_checks_fork_range(b, b->cpt_first, &b->sc, b->split_group, G,
                                          b->ev_fork2));
        SALSAA_TRY(norm_checks_mark_done(b));
    }
    SALSAA_TRY(run_commit(b));
    return norm_checks_join(b);
}

// median-free device timing of fn over `reps` runs (minimum), in ms
template <typename F>
int time_ms(salsaa_batch *b, F fn, int runs, float *ms_out) {
    Context *ctx = &b->ctx->c;
    hipEvent_t e0, e1;
    SALSAA_HIP(hipEventCreate(&e0));
    SALSAA_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int st = SALSAA_OK;
    for (int r = 0; r < runs && st == SALSAA_OK; r++) {
        (void)hipEventRecord(e0, ctx->stream);
        st = fn();
        (void)hipEventRecord(e1, ctx->stream);
        if (st == SALSAA_OK && hipEventSynchronize(e1) == hipSuccess) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            best = std::min(best, ms);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_out = best;
    return st;
}

int autotune_runs(salsaa_batch *b) {
    size_t max_kt = 0;
    for (int kt : b->job_ktiles) max_kt = std::max(max_kt, (size_t)kt);
    // the commit: K chunk and schedule (a launch per key row class, or one
    // launch over all classes when the key is stored)
    size_t nsmall = 0;
    for (const auto &cc : b->cclass) nsmall += (7 * cc.kappa + 15) / 16 <= 8;
    const bool stored = b->key->k.d_mhat == nullptr;
    const bool modes[3] = {true, stored && b->cclass.size() >= 2, stored && nsmall >= 2};
    size_t best_kc = 0;
    int best_mode = 0;
    float best = 0.f;
    b->t_commit_mode[0] = b->t_commit_mode[1] = b->t_commit_mode[2] = 0.f;
    for (size_t kc : {16, 32, 64, 128, 256}) {
        if (kc > 2 * max_kt && best_kc) break;
        SALSAA_TRY(build_commit_tables(b, kc));
        for (int mode = 0; mode < 3; mode++) {
            if (!modes[mode]) continue;
            b->commit_mixed = mode;
            float ms = 0.f;
            SALSAA_TRY(time_ms(b, [&] { return run_commit(b); }, 2, &ms));
            if (b->t_commit_mode[mode] == 0.f || ms < b->t_commit_mode[mode]) b->t_commit_mode[mode] = ms;
            if (!best_kc || ms < best) {
                best = ms;
                best_kc = kc;
                best_mode = mode;
            }
        }
    }
    SALSAA_TRY(build_commit_tables(b, best_kc));
    b->commit_mixed = best_mode;
    b->t_commit = best;
    // the NTT pass's radix-exchange schedule: the whole round trip (the
    // witnesses come back unchanged) with the exchange through LDS and
    // through DPP quad permutes
    for (int x = 0; x < 2; x++) {
        b->ntt_xch = x;
        SALSAA_TRY(time_ms(b, [&] { return run_fwd_planes(b, 1); }, 2, &b->t_ntt[x]));
    }
    b->ntt_xch = b->t_ntt[1] < b->t_ntt[0] ? 1 : 0;
    // its persistent blocks: 29/32, 7/8 or 13/16 of the CUs (the write-heavy
    // pass has run fastest on 224..232 of 256 CUs, 8.5-8.6 ms, against 9.0-9.1
    // on all of them and 9.1-9.2 on 192; where the optimum lies varies from
    // box to box)
    {
        const unsigned ncu = (unsigned)b->ctx->c.ncu;
        const unsigned cand[3] = {ncu * 29 / 32, ncu * 7 / 8, ncu * 13 / 16};
        int best_g = 0;
        for (int g = 0; g < 3; g++) {
            b->grid_cand[g] = cand[g];
            b->ntt_grid = cand[g];
            SALSAA_TRY(time_ms(b, [&] { return run_fwd_planes(b, 1); }, 2, &b->t_grid[g]));
            if (b->t_grid[g] < b->t_grid[best_g]) best_g = g;
        }
        b->ntt_grid = cand[best_g];
    }
    // the norm checks as the step runs them: after a round trip (of the
    // tuning witnesses), from the short copy
    SALSAA_TRY(run_fwd_planes(b, 1));
    int best_cpt = 0;
    float best_nc = 0.f;
    for (int cpt : {1, 2, 4, 8}) {
        float ms = 0.f;
        SALSAA_TRY(time_ms(b, [&] { return run_norm_checks(b, cpt, &b->sc); }, 2, &ms));
        if (!best_cpt || ms < best_nc) {
            best_nc = ms;
            best_cpt = cpt;
        }
    }
    b->cpt_first = best_cpt;
    b->t_nc = best_nc;
    // a split step's second launch runs beside the first part's norm checks:
    // the whole step timed with it on the pass's blocks, 3/4 and 11/16 of the
    // CUs (fewer leave CUs to the norm checks)
    if (b->split_mode && !b->ntt_grid2) {
        const unsigned ncu = (unsigned)b->ctx->c.ncu;
        const unsigned cand[3] = {0, ncu * 3 / 4, ncu * 11 / 16};
        int best_g = 0;
        for (int g = 0; g < 3; g++) {
            b->grid2_cand[g] = cand[g] ? cand[g] : b->ntt_grid;
            b->ntt_grid2 = cand[g];
            SALSAA_TRY(time_ms(b, [&] { return run_step(b); }, 2, &b->t_grid2[g]));
            if (b->t_grid2[g] < b->t_grid2[best_g]) best_g = g;
        }
        b->ntt_grid2 = cand[best_g];
    }
    return SALSAA_OK;
}

// The tuner's runs over witnesses like a step's (residues in [-16, 16] in
// every slice) instead of the fresh arena's zeros, over which the NTT pass ran
// 7-8 % faster; the arena is zeroed again afterwards, and the short copy's
// formats reset, so that nothing of the tuning data remains visible.
int autotune(salsaa_batch *b) {
    Context *c = &b->ctx->c;
    SALSAA_TRY(launch_sample_bounded(c, b->d_wit, b->tot_w * kD, 16, 0x7475u, 0x7475u));
    SALSAA_TRY(autotune_runs(b));
    hipLaunchKernelGGL(k_batch_zero, dim3(ceil_div(b->tot_w * kD, 256)), dim3(256), 0, c->stream,
                       b->d_wit, b->tot_w * kD);
    hipLaunchKernelGGL(k_batch_fill8, dim3(ceil_div(b->tot_w * kD / kShortSlice, 256)), dim3(256),
                       0, c->stream, b->sc.fmt, b->tot_w * kD / kShortSlice, (uint8_t)2);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

void destroy(salsaa_batch *b) {
    if (!b) return;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->c.stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto *v : b->views) salsaa_key_destroy(v);
    free_outputs(b->out[0]);
    free_outputs(b->out[1]);
    (void)hipFree(b->d_wit);
    (void)hipFree(b->d_cs);
    (void)hipFree(b->sc.t);
    (void)hipFree(b->sc.t8);
    (void)hipFree(b->sc.fmt);
    (void)hipFree(b->d_pjobs);
    (void)hipFree(b->d_off);
    (void)hipFree(b->d_m);
    (void)hipFree(b->d_planes);
    (void)hipFree(b->d_ptiles);
    (void)hipFree(b->d_wrow);
    (void)hipFree(b->d_partial);
    (void)hipFree(b->d_cjobs);
    (void)hipFree(b->d_table);
    (void)hipFree(b->d_fjobs);
    for (int k = 0; k < kNcStreams; k++) {
        (void)hipFree(b->nc[k].macc);
        (void)hipFree(b->nc[k].bufA);
        (void)hipFree(b->nc[k].bufB);
        if (b->nc_ctx[k].stream) (void)hipStreamDestroy(b->nc_ctx[k].stream);
        if (b->ev_done[k]) (void)hipEventDestroy(b->ev_done[k]);
    }
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_fork2) (void)hipEventDestroy(b->ev_fork2);
    delete b;
}

int create(salsaa_ctx *ctx, const salsaa_key *key, size_t n, const size_t *m, const size_t *kappa,
           size_t reps, int fq2, uint64_t nr, int tune, salsaa_batch *b) {
    Context *c = &ctx->c;
    SALSAA_CHECK(key != nullptr && key->k.ctx == c, "key missing or from another context");
    SALSAA_CHECK(key->k.d_AB != nullptr || key->k.d_mhat != nullptr,
                 "batched commit needs a key with MFMA planes or a generated CRS key");
    SALSAA_CHECK(n >= 1 && m != nullptr && kappa != nullptr, "batch needs n >= 1 instances");
    SALSAA_CHECK(reps >= 1 && reps < 65536, "repetitions out of range");
    SALSAA_CHECK(nr < c->q, "non-residue must be reduced mod q");
    b->ctx = ctx;
    b->key = key;
    b->n = n;
    b->reps = reps;
    b->fq2 = fq2 ? 1 : 0;
    b->nr = nr;
    b->m.assign(m, m + n);
    b->kappa.assign(kappa, kappa + n);
    for (size_t i = 0; i < n; i++) {
        SALSAA_CHECK(m[i] >= 1 && (m[i] & (m[i] - 1)) == 0, "witness lengths must be powers of two");
        SALSAA_CHECK(m[i] <= key->k.m, "witness longer than the key");
        SALSAA_CHECK(kappa[i] >= 1 && kappa[i] <= key->k.kappa, "kappa must be 1..key rows");
        b->krows = std::max(b->krows, kappa[i]);
    }
    // sort by length, longest first (stable)
    b->order.resize(n);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(),
                     [&](size_t x, size_t y) { return b->m[x] > b->m[y]; });
    b->pos.resize(n);
    b->off_w.resize(n);
    b->off_c.resize(n);
    b->off_msg.resize(n);
    b->nrounds.resize(n);
    for (size_t p = 0; p < n; p++) {
        const size_t i = b->order[p];
        b->pos[i] = p;
        b->nrounds[p] = (size_t)__builtin_ctzll(b->m[i] * kD);
        b->off_w[p] = b->tot_w;
        b->off_c[p] = b->tot_c;
        b->off_msg[p] = b->tot_msg;
        b->tot_w += (b->m[i] + kPlaneTile - 1) / kPlaneTile * kPlaneTile;  // whole tiles
        b->tot_c += reps * b->nrounds[p] * 2;
        b->tot_msg += reps * b->nrounds[p] * 6;
    }
    SALSAA_TRY(dmalloc(&b->d_wit, b->tot_w * kD));
    SALSAA_TRY(dmalloc(&b->d_cs, b->tot_c));
    SALSAA_TRY(dmalloc(&b->sc.t, b->tot_w * kD));
    SALSAA_TRY(dmalloc(&b->sc.t8, b->tot_w * kD));
    SALSAA_TRY(dmalloc(&b->sc.fmt, b->tot_w * kD / kShortSlice));
    hipLaunchKernelGGL(k_batch_zero, dim3(ceil_div(b->tot_w * kD, 256)), dim3(256), 0, c->stream,
                       b->d_wit, b->tot_w * kD);
    hipLaunchKernelGGL(k_batch_fill8, dim3(ceil_div(b->tot_w * kD / kShortSlice, 256)), dim3(256),
                       0, c->stream, b->sc.fmt, b->tot_w * kD / kShortSlice, (uint8_t)2);
    hipLaunchKernelGGL(k_batch_zero, dim3(ceil_div(b->tot_c, 256)), dim3(256), 0, c->stream,
                       b->d_cs, b->tot_c);
    SALSAA_HIP(hipGetLastError());
    SALSAA_TRY(alloc_outputs(b, b->out[0]));

    // commit jobs: per class of key row tiles (16, 8, 4), in length order,
    // 16 instances each; planes per job, and the planes batch reads the
    // arena rows in commit order
    std::vector<unsigned long long> hm(n);  // length per arena position
    for (size_t p = 0; p < n; p++) hm[p] = b->m[b->order[p]];
    auto key_tiles = [](size_t kap) { return (7 * kap + 15) / 16; };
    std::vector<size_t> corder;  // commit order -> arena position
    b->cpos.resize(n);
    size_t planes_bytes = 0;
    for (size_t lim : {(size_t)16, (size_t)8, (size_t)4}) {
        salsaa_batch::CommitClass cc;
        const size_t c0 = corder.size();
        for (size_t p = 0; p < n; p++) {
            const size_t rt = key_tiles(b->kappa[b->order[p]]);
            if (rt <= lim && (lim == 4 || rt > lim / 2)) {
                b->cpos[p] = corder.size();
                corder.push_back(p);
                cc.kappa = std::max(cc.kappa, b->kappa[b->order[p]]);
            }
        }
        const size_t cnt = corder.size() - c0;
        if (cnt == 0) continue;
        cc.job0 = b->njobs;
        cc.njobs = (cnt + 15) / 16;
        for (size_t jj = 0; jj < cc.njobs; jj++) {
            PlaneJob pj;
            pj.first = (int)(c0 + 16 * jj);
            pj.rows = (int)std::min<size_t>(16, cnt - 16 * jj);
            pj.ktiles = (int)plane_ktiles(hm[corder[pj.first]]);
            pj.blk0 = (int)b->pnblk;
            b->job_planes.push_back(planes_bytes);
            b->job_cfirst.push_back((size_t)pj.first);
            pj.dst = nullptr;
            b->pjobs.push_back(pj);
            b->job_ktiles.push_back(pj.ktiles);
            b->job_rows.push_back(pj.rows);
            const size_t rt = (7 * (size_t)pj.rows + 15) / 16;
            cc.max_tiles = std::max(cc.max_tiles, rt);
            planes_bytes += (size_t)64 * pj.ktiles * rt * 1024;
            b->pnblk += (size_t)pj.rows * pj.ktiles / 2;
            SALSAA_CHECK((size_t)pj.ktiles <= key->k.a_ktiles, "witness longer than the key planes");
        }
        b->njobs += cc.njobs;
        b->cclass.push_back(cc);
    }
    std::vector<unsigned long long> hoff(n), hmc(n);  // in commit order
    for (size_t c2 = 0; c2 < n; c2++) {
        hoff[c2] = b->off_w[corder[c2]];
        hmc[c2] = hm[corder[c2]];
    }
    SALSAA_CHECK(b->pnblk < (1u << 31), "batch too large");
    SALSAA_TRY(dmalloc(&b->d_planes, planes_bytes));
    b->planes_bytes = planes_bytes;
    // Columns of a shorter instance past its own tiles, up to its row
    // block's length, are never written by a step: zero them once here.
    hipLaunchKernelGGL(k_batch_zero, dim3(ceil_div(planes_bytes / 8, 256)), dim3(256), 0,
                       c->stream, reinterpret_cast<uint64_t *>(b->d_planes), planes_bytes / 8);
    SALSAA_HIP(hipGetLastError());
    for (size_t j = 0; j < b->njobs; j++) b->pjobs[j].dst = b->d_planes + b->job_planes[j];
    // each arena tile's place in the planes (see store_planes): value row
    // of its instance in its job, K tiles from its offset in the instance
    std::vector<size_t> job_of(n);
    for (size_t j = 0; j < b->njobs; j++)
        for (int r = 0; r < b->job_rows[j]; r++) job_of[b->job_cfirst[j] + r] = j;
    std::vector<PlaneTile> pt(b->tot_w / kPlaneTile);
    std::vector<unsigned long long> hwrow(n);
    for (size_t p = 0; p < n; p++) {
        const size_t j = job_of[b->cpos[p]], row = b->cpos[p] - b->job_cfirst[j];
        const size_t rt = (7 * (size_t)b->job_rows[j] + 15) / 16;
        const size_t kstride = rt * 1024, sstride = (size_t)b->job_ktiles[j] * kstride;
        SALSAA_CHECK(sstride < (1ull << 32), "row block too large for the plane tile table");
        hwrow[p] = b->off_w[p] * kD;
        const size_t tiles = (hm[p] + kPlaneTile - 1) / kPlaneTile;
        for (size_t k = 0; k < tiles; k++) {
            PlaneTile &d = pt[b->off_w[p] / kPlaneTile + k];
            d.base = b->job_planes[j] + k * (kPlaneTile / 64) * kstride + row * 7 * 64;
            d.sstride = (unsigned)sstride;
            d.kstride = (unsigned)kstride;
        }
    }
    SALSAA_TRY(dmalloc(&b->d_ptiles, pt.size()));
    SALSAA_TRY(dmalloc(&b->d_wrow, n));
    SALSAA_TRY(upload(c, b->d_ptiles, pt));
    SALSAA_TRY(upload(c, b->d_wrow, hwrow));
    SALSAA_TRY(dmalloc(&b->d_pjobs, b->pjobs.size()));
    SALSAA_TRY(dmalloc(&b->d_off, n));
    SALSAA_TRY(dmalloc(&b->d_m, n));
    SALSAA_TRY(upload(c, b->d_pjobs, b->pjobs));
    SALSAA_TRY(upload(c, b->d_off, hoff));
    SALSAA_TRY(upload(c, b->d_m, hmc));
    // commit buffers sized for the smallest K chunk tried (16 tiles)
    size_t pcap = 0, tcap = 0;
    for (const auto &cc : b->cclass)
        for (size_t j = cc.job0; j < cc.job0 + cc.njobs; j++) {
            const size_t ks = (b->job_ktiles[j] + 15) / 16;
            pcap += ks * 64 * cc.kappa * b->job_rows[j];
            tcap += 3 * ks * 64;  // the per-class tables and the two mixed ones
        }
    b->partial_cap = pcap;
    b->table_cap = tcap;
    SALSAA_TRY(dmalloc(&b->d_partial, pcap));
    SALSAA_TRY(dmalloc(&b->d_table, tcap));
    SALSAA_TRY(dmalloc(&b->d_cjobs, b->njobs));
    SALSAA_TRY(dmalloc(&b->d_fjobs, b->njobs));
    // norm check classes of equal length
    // classes to streams, largest first onto the least loaded stream; each
    // stream's scratch fits the largest class it runs
    size_t need_m[kNcStreams] = {}, need_a[kNcStreams] = {}, need_b[kNcStreams] = {};
    double load[kNcStreams] = {};
    for (size_t p = 0; p < n;) {
        size_t q = p;
        while (q < n && hm[q] == hm[p]) q++;
        salsaa_batch::Group g{p, q - p, hm[p] * kD};
        int k = 0;
        for (int j = 1; j < kNcStreams; j++)
            if (load[j] < load[k]) k = j;
        load[k] += (double)g.rows * (double)g.L + 1e6;  // a fixed cost per class
        size_t a, bb, mm;
        nc_scratch_need(g.rows, g.L, reps, &mm, &a, &bb, true);
        need_m[k] = std::max(need_m[k], mm);
        need_a[k] = std::max(need_a[k], a);
        need_b[k] = std::max(need_b[k], bb);
        b->groups.push_back(g);
        b->group_stream.push_back(k);
        p = q;
    }
    SALSAA_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    SALSAA_HIP(hipEventCreateWithFlags(&b->ev_fork2, hipEventDisableTiming));
    // The split.  Mode 1 (the default): the first part takes the shorter
    // classes up to 80 % of the arena and splits the next class by rows at
    // that mark (both halves as norm checks of their own, on its stream), so
    // that the norm checks left beside the commit are those of a part of the
    // longest class; their norm checks, at about a sixth of the pass's time
    // per element, fit beside the pass over the rest.  Mode 2: the longer
    // classes first, whole classes up to 70 %.  Mode 0: one pass.  The mode
    // comes from SALSAA_STEP_SPLIT.
    {
        const char *env = getenv("SALSAA_STEP_SPLIT");
        const int mode = env ? atoi(env) : 1;
        const size_t G = b->groups.size();
        size_t acc = 0, gsplit = 0;
        if (mode == 1) {
            gsplit = G;
            while (gsplit > 1) {
                const auto &g = b->groups[gsplit - 1];
                const size_t e = g.rows * (g.L / kD);
                if (10 * (acc + e) > 8 * b->tot_w) break;
                acc += e;
                gsplit--;
            }
            const salsaa_batch::Group g = b->groups[gsplit - 1];
            const size_t rb = (8 * b->tot_w / 10 - acc) / (g.L / kD);  // rows to the mark
            if (rb >= 1 && rb < g.rows) {
                b->groups[gsplit - 1].rows = g.rows - rb;
                b->groups.insert(b->groups.begin() + gsplit,
                                 salsaa_batch::Group{g.first + g.rows - rb, rb, g.L});
                b->group_stream.insert(b->group_stream.begin() + gsplit, b->group_stream[gsplit - 1]);
            }
        } else if (mode == 2) {
            while (gsplit + 1 < G) {
                const auto &g = b->groups[gsplit];
                const size_t e = g.rows * (g.L / kD);
                if (10 * (acc + e) > 7 * b->tot_w) break;
                acc += e;
                gsplit++;
            }
        }
        if ((mode == 1 || mode == 2) && gsplit > 0 && gsplit < b->groups.size()) {
            b->split_mode = mode;
            b->split_group = gsplit;
            b->split_tile = b->off_w[b->groups[gsplit].first] / kPlaneTile;
        }
    }
    for (int k = 0; k < kNcStreams; k++) {
        share_context(*c, b->nc_ctx[k]);
        SALSAA_HIP(hipStreamCreateWithFlags(&b->nc_ctx[k].stream, hipStreamNonBlocking));
        SALSAA_HIP(hipEventCreateWithFlags(&b->ev_done[k], hipEventDisableTiming));
        SALSAA_TRY(dmalloc(&b->nc[k].macc, need_m[k] / sizeof(unsigned long long) + 1));
        SALSAA_TRY(dmalloc(&b->nc[k].bufA, need_a[k] / sizeof(double) + 1));
        SALSAA_TRY(dmalloc(&b->nc[k].bufB, need_b[k] / sizeof(double) + 1));
        b->nc[k].macc_bytes = need_m[k];
        b->nc[k].bufA_bytes = need_a[k];
        b->nc[k].bufB_bytes = need_b[k];
    }
    // per-instance key views for the plain per-call path
    for (size_t i = 0; i < n; i++) {
        salsaa_key *v = nullptr;
        SALSAA_TRY(salsaa_key_view(key, kappa[i], m[i], &v));
        b->views.push_back(v);
    }
    if (tune) return autotune(b);
    b->cpt_first = 0;
    return build_commit_tables(b, 64);
}

}  // namespace

extern "C" {

int salsaa_batch_create(salsaa_ctx *ctx, const salsaa_key *key, size_t n, const size_t *m,
                        const size_t *kappa, size_t reps, int fq2, uint64_t nr, int tune,
                        salsaa_batch **out) {
    CTX_D64(ctx, "salsaa_batch_create");
    SALSAA_CHECK(out != nullptr, "null output pointer");
    salsaa_batch *b = new (std::nothrow) salsaa_batch();
    if (!b) return SALSAA_ERR_NOMEM;
    const int st = create(ctx, key, n, m, kappa, reps, fq2, nr, tune, b);
    if (st != SALSAA_OK) {
        destroy(b);
        return st;
    }
    *out = b;
    return SALSAA_OK;
}

int salsaa_batch_destroy(salsaa_batch *b) {
    if (!b) return SALSAA_OK;
    (void)hipSetDevice(b->ctx->c.device);
    destroy(b);
    return SALSAA_OK;
}

int salsaa_batch_slots(salsaa_batch *b, size_t i, int plain, salsaa_slot_t *s) {
    SALSAA_CHECK(b != nullptr && s != nullptr, "null batch or output");
    SALSAA_CHECK(i < b->n, "instance index out of range");
    CTX_OK(b->ctx);
    if (plain && !b->out[1].com) SALSAA_TRY(alloc_outputs(b, b->out[1]));
    const Outputs &o = b->out[plain ? 1 : 0];
    const size_t p = b->pos[i];
    s->witness = b->d_wit + b->off_w[p] * kD;
    s->challenges = b->d_cs + b->off_c[p];
    s->commitment = o.com + b->cpos[p] * b->krows * kD;
    s->claims = o.claims + p;
    s->msgs = o.msgs + b->off_msg[p];
    s->finals = o.fin + p * b->reps * 2;
    s->evals = o.ev + p * b->reps * 2;
    s->rhs = o.rhs + p * b->reps * 2 * kD;
    s->ok = o.ok + p * b->reps;
    s->m = b->m[i];
    s->kappa = b->kappa[i];
    s->nrounds = b->nrounds[p];
    return SALSAA_OK;
}

int salsaa_batch_step(salsaa_batch *b, int use_graph) {
    SALSAA_CHECK(b != nullptr, "null batch");
    CTX_OK(b->ctx);
    Context *ctx = &b->ctx->c;
    if (!use_graph) return run_step(b);
    if (!b->exec) {
        // capture; every buffer the step touches is owned by the plan, so the
        // graph stays valid whatever else runs on the context between replays
        SALSAA_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int st = run_step(b);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        if (st != SALSAA_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        SALSAA_HIP(e);
        b->graph = graph;
        SALSAA_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    SALSAA_HIP(hipGraphLaunch(b->exec, ctx->stream));
    return SALSAA_OK;
}

int salsaa_batch_phase(salsaa_batch *b, int phase) {
    SALSAA_CHECK(b != nullptr, "null batch");
    CTX_OK(b->ctx);
    Context *ctx = &b->ctx->c;
    switch (phase) {
    case 0: return run_fwd_planes(b, 1);
    case 1: return launch_ntt(ctx, b->d_wit, b->tot_w, SALSAA_REP_INCOMPLETE, 1);
    case 2: return run_planes(b);
    case 3: return run_commit(b);
    case 4: return run_norm_checks(b, b->cpt_first, nullptr);
    case 5: return launch_ntt(ctx, b->d_wit, b->tot_w, SALSAA_REP_INCOMPLETE, 0);
    case 6: return run_fwd_planes(b, 0);
    case 7: return run_norm_checks(b, b->cpt_first, &b->sc);
    default: SALSAA_CHECK(false, "phase must be 0..7");
    }
}

int salsaa_batch_step_plain(salsaa_batch *b) {
    SALSAA_CHECK(b != nullptr, "null batch");
    CTX_OK(b->ctx);
    if (!b->out[1].com) SALSAA_TRY(alloc_outputs(b, b->out[1]));
    salsaa_ctx *ctx = b->ctx;
    const Outputs &o = b->out[1];
    for (size_t i = 0; i < b->n; i++) {
        const size_t p = b->pos[i], nr = b->nrounds[p];
        uint64_t *w = b->d_wit + b->off_w[p] * kD;
        SALSAA_TRY(salsaa_convert(ctx, w, b->m[i], SALSAA_REP_COEFF, SALSAA_REP_INCOMPLETE));
        SALSAA_TRY(salsaa_convert(ctx, w, b->m[i], SALSAA_REP_INCOMPLETE, SALSAA_REP_COEFF));
        SALSAA_TRY(salsaa_parallel_dot_series_matrix(ctx, b->views[i], w, 1, b->m[i],
                                                     o.com + b->cpos[p] * b->krows * kD));
        SALSAA_TRY(salsaa_norm_check(ctx, w, 1, b->m[i] * kD, b->d_cs + b->off_c[p], b->reps,
                                     b->fq2, b->nr, o.claims + p, o.msgs + b->off_msg[p],
                                     o.fin + p * b->reps * 2, o.ev + p * b->reps * 2,
                                     o.rhs + p * b->reps * 2 * kD, o.ok + p * b->reps));
        (void)nr;
    }
    return SALSAA_OK;
}

int salsaa_batch_tuning(const salsaa_batch *b, size_t *kchunk, int *cpt_first, float *commit_ms,
                        float *norm_check_ms) {
    SALSAA_CHECK(b != nullptr, "null batch");
    if (kchunk) *kchunk = b->kchunk;
    if (cpt_first) *cpt_first = b->cpt_first;
    if (commit_ms) *commit_ms = b->t_commit;
    if (norm_check_ms) *norm_check_ms = b->t_nc;
    return SALSAA_OK;
}

int salsaa_batch_tuning_ntt(const salsaa_batch *b, int *xch, float *lds_ms, float *dpp_ms) {
    SALSAA_CHECK(b != nullptr, "null batch");
    if (xch) *xch = b->ntt_xch;
    if (lds_ms) *lds_ms = b->t_ntt[0];
    if (dpp_ms) *dpp_ms = b->t_ntt[1];
    return SALSAA_OK;
}

int salsaa_batch_set_ntt_schedule(salsaa_batch *b, int xch) {
    SALSAA_CHECK(b != nullptr, "null batch");
    SALSAA_CHECK(xch == 0 || xch == 1, "NTT schedule must be 0 (LDS) or 1 (DPP)");
    CTX_OK(b->ctx);
    // the captured step launches the other schedule: capture again on next use
    SALSAA_HIP(hipStreamSynchronize(b->ctx->c.stream));
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->ntt_xch = xch;
    return SALSAA_OK;
}

int salsaa_batch_tuning_grid(const salsaa_batch *b, unsigned *blocks, unsigned *cand, float *ms) {
    SALSAA_CHECK(b != nullptr, "null batch");
    if (blocks) *blocks = b->ntt_grid ? b->ntt_grid : (unsigned)b->ctx->c.ncu;
    for (int g = 0; g < 3; g++) {
        if (cand) cand[g] = b->grid_cand[g];
        if (ms) ms[g] = b->t_grid[g];
    }
    return SALSAA_OK;
}

int salsaa_batch_tuning_step_grid(const salsaa_batch *b, unsigned *blocks, unsigned *cand,
                                  float *ms) {
    SALSAA_CHECK(b != nullptr, "null batch");
    if (blocks)
        *blocks = b->ntt_grid2 ? b->ntt_grid2 : b->ntt_grid ? b->ntt_grid : (unsigned)b->ctx->c.ncu;
    for (int g = 0; g < 3; g++) {
        if (cand) cand[g] = b->grid2_cand[g];
        if (ms) ms[g] = b->t_grid2[g];
    }
    return SALSAA_OK;
}

int salsaa_batch_set_ntt_grid(salsaa_batch *b, unsigned blocks) {
    SALSAA_CHECK(b != nullptr, "null batch");
    SALSAA_CHECK(blocks < 65536, "too many persistent blocks");
    CTX_OK(b->ctx);
    // the captured step launches the old grid: capture again on next use
    SALSAA_HIP(hipStreamSynchronize(b->ctx->c.stream));
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->ntt_grid = blocks;
    b->ntt_grid2 = 0;  // both launches of a split step
    return SALSAA_OK;
}

int salsaa_batch_tuning_commit(const salsaa_batch *b, int *schedule, float *ms) {
    SALSAA_CHECK(b != nullptr, "null batch");
    if (schedule) *schedule = b->commit_mixed;
    if (ms)
        for (int k = 0; k < 3; k++) ms[k] = b->t_commit_mode[k];
    return SALSAA_OK;
}

int salsaa_batch_set_commit_schedule(salsaa_batch *b, int mixed) {
    SALSAA_CHECK(b != nullptr, "null batch");
    SALSAA_CHECK(mixed >= 0 && mixed <= 2, "commit schedule must be 0, 1 or 2");
    SALSAA_CHECK(!mixed || !b->key->k.d_mhat, "the mixed commit launches need a stored key");
    CTX_OK(b->ctx);
    // the captured step launches the other schedule: capture again on next use
    SALSAA_HIP(hipStreamSynchronize(b->ctx->c.stream));
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->commit_mixed = mixed;
    return SALSAA_OK;
}

}  // extern "C"
