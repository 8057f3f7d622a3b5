"""Prover-step benchmark on one MI355X (ring Z_q[X]/(X^64 + 1),
q = 1125899906840833).  Prints one JSON line.

Headline (the metric's "batched vs plain per-call GPU"): 512 independent
prover-step instances, each
  * forward + inverse IncompleteNTT of its witness (m ring elements),
  * commit of the witness (one row) under the first kappa rows of one
    resident key A in R_q^{32 x 2^18} (uniform),
  * norm check of the witness's m*64 coefficients: norm_check_1 claim, two
    repetitions of the sumcheck with the instance's own F_{q^2} challenges,
    verifier replay, norm_check_3 rhs,
with m = 2^k, k uniform over the integers 10..18, kappa uniform over
{8, 16, 32}, witness coefficients uniform in [-16, 16].  `value` is instance
steps per second of the batched step (grouped launches replayed from a
hipGraph); the same instances run through the plain per-call path (a native
loop of the per-instance operator calls) for the comparison, and every output
of every instance is compared between the two.

Also reported: the served stream ("served": open-loop Poisson arrivals of
commit + norm-check jobs through the continuous-batching server and through
the same server one job at a time, jobs/s and p50/p99 latency), the folding
step ("folding": 8 instances decomposed, folded, re-committed and
norm-checked, fused vs separate calls), the single-instance full-size step
(NTT of 2^20 ring elements; commit A 32 x 16384 . W 8 x 16384; norm check
8 x 2^20; decomp_ell base 2^8 x 6 with verify_decomp) and the projection
("projection": project, batch_projections and the cross terms of an
8 x 2^19 witness) and the power-series CRS key ("crs": 13 x 2^21, generated
on the device vs stored rows).  All inputs are generated on the device from seeds before
timing starts; the batched and single-instance outputs stay in HBM.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

NTT_COUNT = 1 << 20
KAPPA, M, R = 32, 16384, 8
BOUND = 16
PARTS, RADIX = 6, 256
REPS = 2


class Step:
    def __init__(self, sa, ctx, seed=0, ntt_count=NTT_COUNT, kappa=KAPPA, m=M, r=R):
        self.sa, self.ctx = sa, ctx
        self.ntt_count, self.kappa, self.m, self.r = ntt_count, kappa, m, r
        self.L = m * 64
        self.nrounds = self.L.bit_length() - 1
        self.X = ctx.uniform((ntt_count, 64), seed, sa.STREAM_NTT)
        A = ctx.uniform((kappa, m, 64), seed, sa.STREAM_A)
        self.key = ctx.key(A)
        A.free()
        self.W = ctx.bounded((r, m, 64), BOUND, seed, sa.STREAM_W)
        self.cs = ctx.challenges(REPS, self.nrounds, 1, seed)
        self.commitment = ctx.empty((kappa, r, 64))
        self.nc = None
        self.new_W = ctx.empty((r * PARTS, m, 64))
        self.rhs_dec = ctx.empty((kappa, r * PARTS, 64))
        self.d_radix = ctx.empty((1,))
        self.mismatch = ctx.to_device(np.zeros(2, np.uint32), np.uint32)
        self.ev = [ctx.event() for _ in range(6)]
        self.phase_ms = {k: [] for k in ("ntt_fwd", "ntt_inv", "commit", "norm_check", "decomp_ell")}

    def run(self, timed_phases=False):
        sa, ctx, ev = self.sa, self.ctx, self.ev
        if timed_phases:
            ctx.record(ev[0])
        ctx.convert(self.X, sa.REP_COEFF, sa.REP_INCOMPLETE)
        if timed_phases:
            ctx.record(ev[1])
        ctx.convert(self.X, sa.REP_INCOMPLETE, sa.REP_COEFF)
        if timed_phases:
            ctx.record(ev[2])
        ctx.parallel_dot_series_matrix(self.key, self.W, out=self.commitment)
        if timed_phases:
            ctx.record(ev[3])
        self.nc = ctx.norm_check(self.W, self.cs, 1, outs=self.nc)
        if timed_phases:
            ctx.record(ev[4])
        ctx.decomp_ell(self.key, self.W, PARTS, RADIX, new_W=self.new_W, rhs=self.rhs_dec,
                       d_radix=self.d_radix)
        ctx.verify_decomp(self.rhs_dec, self.commitment, PARTS, self.d_radix,
                          mismatch=self.mismatch)
        if timed_phases:
            ctx.record(ev[5])

    def collect_phases(self):
        e = self.ev
        names = ("ntt_fwd", "ntt_inv", "commit", "norm_check", "decomp_ell")
        for i, n in enumerate(names):
            self.phase_ms[n].append(self.ctx.elapsed_ms(e[i], e[i + 1]))

    def check(self):
        ok = self.nc["ok"].download()
        mism = self.mismatch.download()
        return bool(ok.all()) and int(mism.sum()) == 0


def algorithmic(step):
    """Algorithmic bytes / ops per launch of each phase (u64 = 8 bytes)."""
    n, k, m, r = step.ntt_count, step.kappa, step.m, step.r
    return {
        "ntt_fwd": dict(bound="hbm", bytes=n * 64 * 8 * 2),
        "ntt_inv": dict(bound="hbm", bytes=n * 64 * 8 * 2),
        "commit": dict(bound="valu", bytes=8 * 64 * (k * m + m * r + k * r), modmac=k * m * r * 64),
        "norm_check": dict(bound="hbm", bytes=8 * r * m * 64),
        "decomp_ell": dict(bound="valu", bytes=8 * r * m * 64 * (1 + PARTS),
                           modmac=k * m * r * PARTS * 64),
    }


def traffic_bytes():
    """HBM bytes per launch of the fused forward NTT + planes kernel over the
    batched workload's witness arena, measured by rocprofv3 FETCH_SIZE and
    WRITE_SIZE (each scaled by a plain-copy calibration) and committed in
    profiles/ntt_fwd_planes_batched_traffic.json."""
    path = os.path.join(ROOT, "profiles", "ntt_fwd_planes_batched_traffic.json")
    try:
        with open(path) as f:
            return int(json.load(f)["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


N_INST, A_ROWS, A_M = 512, 32, 1 << 18


def cfg3_instances(n=N_INST, seed=3):
    rng = np.random.default_rng(seed)
    m = [int(1 << rng.integers(10, 19)) for _ in range(n)]
    kappa = [int(rng.choice([8, 16, 32])) for _ in range(n)]
    return m, kappa


class Batched:
    """The 512-instance batched workload and its plain per-call twin."""

    def __init__(self, sa, ctx, n=N_INST):
        self.sa, self.ctx = sa, ctx
        self.m, self.kappa = cfg3_instances(n)
        A = ctx.uniform((A_ROWS, A_M, 64), 0, sa.STREAM_A)
        self.key = ctx.key(A)
        ctx.sync()
        A.free()
        self.b = sa.Batch(ctx, self.key, self.m, self.kappa, reps=REPS, fq2=True, tune=True)
        for i in range(n):
            sl = self.b.slot(i)
            w, cs = sl["witness"], sl["challenges"]
            sa._check(sa.lib.salsaa_sample_bounded(ctx.h, w.ptr, w.shape[0] * 64, BOUND,
                                                   1000 + i, sa.STREAM_W))
            sa._check(sa.lib.salsaa_sample_challenges(ctx.h, cs.ptr, REPS, sl["nrounds"], 1,
                                                      5000 + i))
        ctx.sync()

    def outputs(self, plain):
        res = []
        for i in range(self.b.n):
            sl = self.b.slot(i, plain)
            d = [sl[k].download() for k in ("claims", "msgs", "finals", "evals", "rhs", "ok")]
            d.append(sl["commitment"].download()[: sl["kappa"]])
            res.append(d)
        return res

    def bit_exact(self):
        a, b = self.outputs(False), self.outputs(True)
        same = all(np.array_equal(x, y) for u, v in zip(a, b) for x, y in zip(u, v))
        ok = all(u[5].all() for u in a)
        return same and ok


def host_cpu():
    """The host's processor, its hardware threads and the threads the
    baseline uses (the process's CPU share, capped by OMP_NUM_THREADS when the
    environment sets one)."""
    model = ""
    flags = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and not model:
                    model = line.split(":", 1)[1].strip()
                if line.startswith("flags") and not flags:
                    flags = line
    except OSError:
        pass
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    env = os.environ.get("OMP_NUM_THREADS")
    use = min(share, int(env)) if env and env.isdigit() and int(env) > 0 else share
    return {"model": model, "nproc": os.cpu_count(), "cpu_share": share, "threads": use,
            "avx512ifma": "avx512ifma" in flags}


def _median_runs(fn, runs):
    fn()  # warm-up
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline_batched(bw, cls=1 << 14, runs=5):
    """The CPU baseline (oracle/cpu_baseline.cpp: the reference's CPU prover
    path, per-product conversion in the commit, the 10-pass chunked sumcheck,
    AVX-512 IFMA lanes, all threads of the host's share) on one whole length
    class of the batched workload: every instance with m = cls, its own
    witness, challenges and key prefix, 1 warm-up and the median of `runs`
    timed runs per phase.  The batch's time is the class's per phase, scaled
    by the batch's work over the class's (NTT and norm check by the sum of m,
    the commit by the sum of kappa * m); the class's outputs are compared with
    the device's."""
    from oracle import cpu_baseline as cb
    from oracle import oracle as orc
    q = 1125899906840833
    host = host_cpu()
    threads = cb.init(host["threads"], True)
    m, kappa = bw.m, bw.kappa
    idx = [i for i in range(len(m)) if m[i] == cls]
    A = np.stack([orc.sample_uniform_at(0, 0xA, j * A_M * 64, (cls, 64), q) for j in range(A_ROWS)])
    key = cb.Key(A, q)
    del A
    n = (cls * 64).bit_length() - 1
    Ws = [orc.sample_bounded(1000 + i, 0xB, (1, cls, 64), BOUND, q) for i in idx]
    css = [np.stack([orc.sample_uniform(5000 + i, 0x5C000 + rep, (2 * n,), q).reshape(n, 2)
                     for rep in range(REPS)]) for i in idx]
    res = {}

    def ntt():
        for W in Ws:
            cb.ntt_roundtrip(W, q)

    def com():
        res["com"] = [cb.commit(key, W, kappa[i]) for i, W in zip(idx, Ws)]

    def nc():
        res["nc"] = [cb.norm_check(W.reshape(1, -1), cs, q, 5) for W, cs in zip(Ws, css)]

    t = {}
    spread = {}
    for name, fn in (("ntt", ntt), ("commit", com), ("norm_check", nc)):
        t[name], spread[name] = _median_runs(fn, runs)
    key.close()
    # the class's outputs against the device's batched step
    exact = True
    for k, i in enumerate(idx):
        sl = bw.b.slot(i)
        exact &= np.array_equal(res["com"][k], sl["commitment"].download()[: kappa[i]])
        o = res["nc"][k]
        for name in ("claims", "msgs", "finals", "evals", "rhs"):
            exact &= np.array_equal(o[name], sl[name].download())
    sum_m, sum_km = float(sum(m)), float(sum(a * b for a, b in zip(m, kappa)))
    c_m = float(cls * len(idx))
    c_km = float(sum(kappa[i] * cls for i in idx))
    total = (t["ntt"] + t["norm_check"]) * sum_m / c_m + t["commit"] * sum_km / c_km
    class_s = sum(t.values())
    return {
        "value": len(m) / total,
        "unit": "prover-steps/s",
        "cores": threads,
        "kind": "port",
        "variant": cb.variant(),
        "host": host,
        "class": {"m": cls, "instances": len(idx), "median_s": round(class_s, 4),
                  "prover_steps_per_s": len(idx) / class_s,
                  "phases_median_s": {k: round(v, 4) for k, v in t.items()},
                  "phases_runs_s": {k: [round(x, 4) for x in v] for k, v in spread.items()},
                  "bit_exact_vs_device": bool(exact)},
        "sample": ("oracle/cpu_baseline.cpp (the reference's CPU path: per-product NTT in the "
                   "commit, 10-pass sumcheck in 120-way chunks, norm_check_3 by power series; "
                   "AVX-512 IFMA lanes) on %d threads: the whole m = %d class (%d instances), "
                   "median of %d runs after a warm-up, %.2f s; scaled to the batch by each "
                   "phase's work (sum m %d / %d, sum kappa m %d / %d): %.1f s for the 512 "
                   "instances" % (threads, cls, len(idx), runs, class_s, sum_m, c_m, sum_km,
                                  c_km, total)),
    }


def cpu_baseline(step, runs=5):
    """The CPU baseline on the single-instance step at full size, the same
    inputs as the device's: NTT forward + inverse of 2^20 ring elements,
    commit 32 x 16384 . 8 x 16384, the F_{q^2} norm check of 8 x 2^20 with
    norm_check_3, decomp_ell base 2^8 x 6 with verify_decomp; 1 warm-up, then
    the median of `runs` runs per phase; outputs compared with the device's."""
    from oracle import cpu_baseline as cb
    from oracle import oracle as orc
    q = 1125899906840833
    host = host_cpu()
    threads = cb.init(host["threads"], True)
    X = orc.sample_uniform(0, 0xC, (NTT_COUNT, 64), q)
    A = orc.sample_uniform(0, 0xA, (KAPPA, M, 64), q)
    key = cb.Key(A, q)
    del A
    W = orc.sample_bounded(0, 0xB, (R, M, 64), BOUND, q)
    n = (M * 64).bit_length() - 1
    cs = np.stack([orc.sample_uniform(0, 0x5C000 + rep, (2 * n,), q).reshape(n, 2)
                   for rep in range(REPS)])
    res = {}

    def ntt():
        cb.ntt_roundtrip(X, q)

    def com():
        res["com"] = cb.commit(key, W)

    def nc():
        res["nc"] = cb.norm_check(W.reshape(R, -1), cs, q, 5)

    def dec():
        limbs, radix = cb.decompose(W, PARTS, q, RADIX)
        res["dec"] = cb.commit(key, limbs)
        res["bad"] = cb.verify_decomp(res["dec"], res["com"], PARTS, radix, q)

    t, spread = {}, {}
    for name, fn in (("ntt", ntt), ("commit", com), ("norm_check", nc), ("decomp_ell", dec)):
        t[name], spread[name] = _median_runs(fn, runs)
    key.close()
    exact = (np.array_equal(res["com"], step.commitment.download()) and res["bad"] == 0
             and np.array_equal(res["dec"], step.rhs_dec.download()))
    for name in ("claims", "msgs", "finals", "evals", "rhs"):
        exact &= np.array_equal(res["nc"][name], step.nc[name].download())
    total = sum(t.values())
    return {
        "value": 1.0 / total,
        "unit": "prover-steps/s",
        "cores": threads,
        "kind": "port",
        "variant": cb.variant(),
        "phases_median_s": {k: round(v, 4) for k, v in t.items()},
        "phases_runs_s": {k: [round(x, 4) for x in v] for k, v in spread.items()},
        "bit_exact_vs_device": bool(exact),
        "sample": ("oracle/cpu_baseline.cpp on %d threads (%s): the whole config-2 step, "
                   "median of %d runs after a warm-up" % (threads, cb.variant(), runs)),
    }


S_KAPPA, S_MMAX = 16, 1 << 16


def served(sa, ctx, n_jobs=3000, seed=7):
    """Open-loop Poisson stream of commit + norm-check jobs (m = 2^k ring
    elements, k uniform in 6..16, i.e. 2^12..2^22 coefficients; kappa 16;
    coefficients in [-16, 16]; challenges from per-job seeds) against a
    resident key A 16 x 2^16, through the continuous-batching server and
    through the same server with one job per batch (plain per-call)."""
    A = ctx.uniform((S_KAPPA, S_MMAX, 64), 1, sa.STREAM_A)
    key = ctx.key(A)
    ctx.sync()
    A.free()
    rng = np.random.default_rng(seed)
    ms = [int(1 << rng.integers(6, 17)) for _ in range(n_jobs)]

    def run(batch_jobs, inflight, lam, n):
        srv = sa.Server(ctx, key, S_KAPPA, arena_elems=1 << 23, max_batch_jobs=batch_jobs,
                        max_batch_elems=1 << 21, inflight=inflight, reps=REPS, fq2=True)
        for i in range(8):
            srv.submit(ms[i], 10 ** 6 + i)
        srv.drain()
        arr = np.cumsum(rng.exponential(1.0 / lam, n)) if lam else np.zeros(n)
        ids = []
        t0 = time.perf_counter()
        for i in range(n):
            while time.perf_counter() - t0 < arr[i]:
                pass
            ids.append(srv.submit(ms[i], 1000 + i))
        srv.drain()
        res = [srv.poll(j, m) for j, m in zip(ids, ms[:n])]
        srv.close()
        ts = np.array([r["times"] for r in res])
        lat = (ts[:, 3] - ts[:, 0]) * 1e3
        span = ts[:, 3].max() - ts[:, 0].min()
        return {"offered_jobs_per_s": round(lam, 1) if lam else "all at t=0", "jobs": n,
                "jobs_per_s": round(n / span, 1),
                "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_batch_jobs": round(float(np.mean([r["batch_jobs"] for r in res])), 1)}

    cap_b = run(256, 2, 0, n_jobs)
    cap_p = run(1, 1, 0, n_jobs // 5)
    lam_hi = 0.6 * cap_b["jobs_per_s"]
    lam_lo = 0.5 * cap_p["jobs_per_s"]
    out = {
        "workload": "open-loop Poisson commit+norm-check jobs: m = 2^k, k uniform in 6..16 "
                    "(2^12..2^22 coefficients), kappa 16, coefficients in [-16,16], F_q^2 "
                    "challenges from per-job seeds, 2 reps; resident key A 16 x 2^16; "
                    "admission by an HBM arena of 2^23 ring elements",
        "capacity": {"batched": cap_b, "plain": cap_p,
                     "ratio": round(cap_b["jobs_per_s"] / cap_p["jobs_per_s"], 2)},
        "at_0.6x_batched_capacity": {"batched": run(256, 2, lam_hi, n_jobs),
                                     "plain": run(1, 1, lam_hi, n_jobs // 5)},
        "at_0.5x_plain_capacity": {"batched": run(256, 2, lam_lo, n_jobs // 5),
                                   "plain": run(1, 1, lam_lo, n_jobs // 5)},
    }
    key.close()
    return out


F_N, F_R, F_M, F_PARTS, F_ROUT, F_KAPPA, F_BOUND = 8, 8, 4096, 4, 8, 16, 1 << 12


def folding(sa, ctx, steps):
    """The folding step: 8 instances, W 8 x 4096 in [-2^12, 2^12], decomposed
    base 2^4 (auto radix) into 4 limbs, folded by ternary C (8 x 32) per
    instance, the folded witnesses committed under A 16 x 4096 and
    norm-checked (2^18 coefficients per row, F_q^2, 2 reps).  Fused path vs
    the same steps as separate per-instance calls; outputs compared."""
    n, r, m, parts, r_out = F_N, F_R, F_M, F_PARTS, F_ROUT
    rows, nrounds = n * r_out, (m * 64).bit_length() - 1
    A = ctx.uniform((F_KAPPA, m, 64), 2, sa.STREAM_A)
    key = ctx.key(A)
    A.free()
    W = ctx.bounded((n * r, m, 64), F_BOUND, 2, sa.STREAM_W)
    C = ctx.bounded((n, r_out, r * parts, 64), 1, 2, sa.STREAM_FOLD)
    cs = ctx.empty((n, REPS, nrounds, 2))
    for i in range(n):
        sa._check(sa.lib.salsaa_sample_challenges(ctx.h, cs.ptr + i * REPS * nrounds * 16, REPS,
                                                  nrounds, 1, 40 + i))
    outs = ctx.fold_step(key, W, n, parts, C, cs)
    V = sa.DeviceArray.view
    # the separate-call path, buffers allocated once
    limbs = ctx.empty((r * parts, m, 64))
    rad = ctx.empty((1,))
    p_fold = ctx.empty((rows, m, 64))
    p_coeff = ctx.empty((rows, m, 64))
    p_com = ctx.empty((F_KAPPA, rows, 64))
    p_nc = [None] * n
    wsz, csz = r * m * 64 * 8, REPS * nrounds * 16

    def plain():
        for i in range(n):
            Wi = V(ctx, W.ptr + i * wsz, (r, m, 64))
            ctx.decompose(Wi, parts, out=limbs, d_radix=rad)
            fi = V(ctx, p_fold.ptr + i * r_out * m * 512, (r_out, m, 64))
            ctx.fold(limbs, V(ctx, C.ptr + i * r_out * r * parts * 512, (r_out, r * parts, 64)),
                     out=fi)
            ci = V(ctx, p_coeff.ptr + i * r_out * m * 512, (r_out, m, 64))
            ctx.copy(ci, fi)
            ctx.convert(ci, sa.REP_INCOMPLETE, sa.REP_COEFF)
            p_nc[i] = ctx.norm_check(ci, V(ctx, cs.ptr + i * csz, (REPS, nrounds, 2)), 1,
                                     outs=p_nc[i])
        ctx.parallel_dot_series_matrix(key, p_coeff, out=p_com)

    fused = lambda: ctx.fold_step(key, W, n, parts, C, cs, outs=outs)  # noqa: E731
    for _ in range(2):
        fused()
        plain()
    ctx.sync()
    fper, ftotal, _ = timed(ctx, fused, steps)
    pper, ptotal, _ = timed(ctx, plain, max(steps // 4, 2))
    exact = (np.array_equal(p_fold.download(), outs["folded"].download()) and
             np.array_equal(p_coeff.download(), outs["folded_coeff"].download()) and
             np.array_equal(p_com.download(), outs["rhs"].download()) and
             all(np.array_equal(p_nc[i]["msgs"].download(),
                                outs["msgs"].download()[:, :, i * r_out:(i + 1) * r_out])
                 for i in range(n)) and bool(outs["ok"].download().all()))
    key.close()
    fms, pms = ftotal / steps, ptotal / max(steps // 4, 2)
    return {"workload": "folding step: 8 instances, W 8x4096 in [-2^12,2^12], base 2^4 x 4 "
                        "limbs (auto radix), ternary C 8x32 per instance, commit under A 16x4096, "
                        "norm check 64 rows x 2^18, F_q^2, 2 reps",
            "fused": {"value": round(1000.0 / fms, 2), "unit": "folding steps/s",
                      "ms_per_step": round(fms, 4),
                      "p99_ms": round(float(np.percentile(fper, 99)), 4)},
            "separate_calls": {"value": round(1000.0 / pms, 2), "unit": "folding steps/s",
                               "ms_per_step": round(pms, 4)},
            "speedup": round(pms / fms, 3), "bit_exact_vs_separate_calls": exact}


PJ_R, PJ_M = 8, 1 << 19  # the largest projection of the published C2 run (C2.log:122-129)


def projection(sa, ctx, steps, cpu_budget_s=8.0):
    """Projection, batching and cross terms of an 8 x 2^19 witness (cw = 2048,
    2048 chunks): project (signed sums over 256 x 2048 ternary challenges),
    batch_projections with c0 = 1 (the reference's batching challenge) and
    both layers written, and send_cross_terms_before_join.  project's kernel
    is priced against HBM (W read once, the projection written once) and the
    int8 matrix cores (2 * 256 * 448 int8 MACs per witness element); four
    chunks are checked against an int64 matrix product, the batching
    identity on the device (ok)."""
    r, m = PJ_R, PJ_M
    cw = r * 256
    W = ctx.uniform((r, m, 64), 5, sa.STREAM_W)
    rng = np.random.default_rng(11)
    z = rng.integers(0, 2, (256, cw), dtype=np.uint8)
    sg = rng.integers(0, 2, (256, cw), dtype=np.uint8)
    dz, dsg = ctx.to_device(z, np.uint8), ctx.to_device(sg, np.uint8)
    c0 = np.zeros(64, np.uint64)
    c0[0] = 1
    dc0 = ctx.to_device(c0)
    P = ctx.project(W, dz, dsg)
    outs = ctx.batch_projections(P, W, dz, dsg, dc0)
    plhs, wlhs = outs[0], outs[1]
    V = sa.DeviceArray.view
    W0 = V(ctx, W.ptr, (1, m, 64))

    def proj():
        sa._check(sa.lib.salsaa_project(ctx.h, W.ptr, r, m, dz.ptr, dsg.ptr, P.ptr))

    def batch():
        sa._check(sa.lib.salsaa_batch_projections(ctx.h, P.ptr, W.ptr, r, m, dz.ptr, dsg.ptr,
                                                  dc0.ptr, plhs.ptr, wlhs.ptr, outs[2].ptr,
                                                  outs[3].ptr, outs[4].ptr))

    l01 = [None, None]

    def cross():
        l01[0], l01[1] = ctx.send_cross_terms_before_join(plhs, wlhs, W0, P)

    for _ in range(2):
        proj()
        batch()
        cross()
    ctx.sync()
    pper, ptot, _ = timed(ctx, proj, steps)
    bper, btot, _ = timed(ctx, batch, max(steps // 2, 2))
    cper, ctot, _ = timed(ctx, cross, max(steps // 2, 2))
    pms, bms, cms = ptot / steps, btot / max(steps // 2, 2), ctot / max(steps // 2, 2)
    # four chunks against an int64 product (|sum| < 2048 q < 2^61)
    C = np.where(z == 1, 0, np.where(sg == 1, 1, -1)).astype(np.int64)
    got = P.download()
    nch = m // cw
    exact = bool(outs[4].download()[0] == 1)
    for idx in (0, 1, nch + 3, r * nch - 1):
        i, j = divmod(idx, nch)
        chunk = V(ctx, W.ptr + (i * m + j * cw) * 512, (cw, 64)).download().astype(np.int64)
        want = np.mod(C @ chunk, sa.Q).astype(np.uint64)
        exact = exact and np.array_equal(got[idx * 256:(idx + 1) * 256], want)
    alg_bytes = (r * m + m) * 512
    int8_ops = 2.0 * 256 * 448 * r * m
    cpu = None
    if cpu_budget_s > 0:
        from oracle import oracle as orc  # the checker's restatement, timed as the CPU baseline
        host = V(ctx, W.ptr, (1, cw * 4, 64)).download()
        t0, n_done = time.perf_counter(), 0
        while time.perf_counter() - t0 < cpu_budget_s:
            orc.project(host, z[:, :256], sg[:, :256], sa.Q)  # r = 1: 4 * 256 chunks of width 256
            n_done += 1
        el = time.perf_counter() - t0
        # one run: 4 cw = 8192 elements x 256 outputs; the workload: r m x 256
        per_elem = el / (n_done * cw * 4)
        cpu = {"value": round(1.0 / (per_elem * r * m), 4), "unit": "projections/s",
               "cores": 1, "kind": "port",
               "sample": "oracle or_project on 8192 witness elements x 256 outputs, %d runs in "
                         "%.1f s, scaled to 8 x 2^19 elements" % (n_done, el)}
    P.free()
    for a in (plhs, wlhs, outs[2], outs[3], outs[4], W):
        a.free()
    return {"workload": "projection of W 8 x 2^19 (cw = 2048, 2048 chunks x 256 outputs), "
                        "batch_projections (c0 = 1, layers written), cross terms; the "
                        "published C2 run: project 4.67 s, batch projections 434 ms "
                        "(C2.log:125-129)",
            "project": {"ms": round(pms, 4), "value": round(1000.0 / pms, 2),
                        "unit": "projections/s",
                        "p99_ms": round(float(np.percentile(pper, 99)), 4),
                        "GB/s": round(alg_bytes / (pms * 1e-3) / 1e9, 1),
                        "hbm_frac": round(alg_bytes / (pms * 1e-3) / 8e12, 3),
                        "int8_TOP/s": round(int8_ops / (pms * 1e-3) / 1e12, 1),
                        "mfma_frac": round(int8_ops / (pms * 1e-3) / 5.0e15, 3),
                        "algorithmic_bytes": alg_bytes},
            "batch_projections_ms": round(bms, 4), "cross_terms_ms": round(cms, 4),
            "bit_exact_sampled": exact, "cpu_baseline_project": cpu}


CRS_KAPPA, CRS_M = 13, 1 << 21  # the F2 run of the published results (results.md:19)


def crs(sa, ctx, steps):
    """The power-series CRS key (compute_commitment_keys, crs.rs:43-72) of 13
    module elements with rows of 2^21 ring elements, generated on the device
    straight into the commit's digit planes, against the stored-key path (the
    Coefficient rows, 13.6 GB, then salsaa_key_create); and the commit of one
    witness row of 2^21 elements under those planes against the same commit
    with the key generated inside the commit kernel (nothing stored).  All
    three give the same commitment."""
    mod = ctx.uniform((CRS_KAPPA, 64), 9, sa.STREAM_A)
    W = ctx.bounded((1, CRS_M, 64), 16, 9, sa.STREAM_W)
    keys = []

    def gen():
        keys.append(ctx.crs_key(mod, CRS_M))
        while len(keys) > 1:
            keys.pop(0).close()

    gen()
    ctx.sync()
    gper, gtot, _ = timed(ctx, gen, max(steps // 4, 2))
    rows = ctx.compute_commitment_keys(mod, CRS_M)
    skeys = []

    def stored():
        skeys.append(ctx.key(rows))
        while len(skeys) > 1:
            skeys.pop(0).close()

    stored()
    ctx.sync()
    sper, stot, _ = timed(ctx, stored, max(steps // 4, 2))
    a = ctx.parallel_dot_series_matrix(keys[-1], W).download()
    b = ctx.parallel_dot_series_matrix(skeys[-1], W).download()
    n = max(steps // 4, 2)
    gms, sms = gtot / n, stot / n
    # the commit of the row: stored planes vs the key generated in the kernel
    imp = ctx.crs_key(mod, CRS_M, implicit=True)
    com = ctx.empty((CRS_KAPPA, 1, 64))
    for kk in (keys[-1], imp):
        ctx.parallel_dot_series_matrix(kk, W, out=com)
    ctx.sync()
    _, ctot_s, _ = timed(ctx, lambda: ctx.parallel_dot_series_matrix(keys[-1], W, out=com), n)
    _, ctot_i, _ = timed(ctx, lambda: ctx.parallel_dot_series_matrix(imp, W, out=com), n)
    c = com.download()
    imp.close()
    com.free()
    plane_bytes = 64 * ((7 * CRS_KAPPA + 15) // 16) * 1024 * (CRS_M // 64)
    for k in keys + skeys:
        k.close()
    for x in (rows, W, mod):
        x.free()
    return {"workload": "CRS key, 13 module elements x 2^21 powers (the F2 parameters): "
                        "generated on the device vs stored Coefficient rows + key_create",
            "generated_ms": round(gms, 3), "stored_key_create_ms": round(sms, 3),
            "plane_GB/s": round(plane_bytes / (gms * 1e-3) / 1e9, 1),
            "coefficient_rows_not_stored_GB": round(CRS_KAPPA * CRS_M * 512 / 1e9, 2),
            "commit_row_stored_planes_ms": round(ctot_s / n, 3),
            "commit_row_key_generated_in_kernel_ms": round(ctot_i / n, 3),
            "key_planes_not_stored_GB": round(plane_bytes / 1e9, 2),
            "bit_exact_commit": bool(np.array_equal(a, b) and np.array_equal(a, c))}


# The protocol runs (salsaa_protocol): the reference's parameter sets at the
# same witness size in coefficients, in ring degree 64 (the reference's sets
# use N = 128, so m doubles).  (name, kappa, m, INIT_WIT_REP, rounds, folding,
# published prover and commit seconds at N = 128, experiments/results.md:5-19)
PROTOCOL_SETS = (("A0", 12, 1 << 20, 1, 8, False, 3.05, 0.0796),
                 ("B0", 12, 1 << 22, 1, 10, False, 10.61, 0.34822),
                 ("F1", 13, 1 << 20, 4, 1, True, 1.66, 0.35612),
                 ("F2", 13, 1 << 22, 4, 1, True, 5.52, 1.38))


def protocol(sa, ctx, runs=2):
    """Whole protocol runs on the device: the commitment, every round's prover
    operators and verifier relations, and the final check, with the
    reference's DECOMP_PARTS = [1, 2, 2, ..] (protocol.rs:26).  Each set runs
    once untimed, then `runs` times into the outputs of the run before;
    wall = host time around the call (the run's own buffers and the final
    synchronisation included), phases = device time per phase summed over
    rounds."""
    res = {}
    for name, kappa, m, r, rounds, fold, pub_prover, pub_commit in PROTOCOL_SETS:
        mod = ctx.uniform((kappa, 64), 21, sa.STREAM_A)
        W = ctx.bounded((r, m, 64), 1, 21, sa.STREAM_W)
        parts = [1] + [2] * (rounds - 1)
        out = ctx.protocol(mod, W, rounds, parts, folding=fold, seed=1)
        ok = out["status"] == 0
        walls, phases = [], None
        for i in range(runs):
            t0 = time.perf_counter()
            out = ctx.protocol(mod, W, rounds, parts, folding=fold, seed=2 + i, outs=out)
            walls.append((time.perf_counter() - t0) * 1e3)
            ok = ok and out["status"] == 0
            phases = out["times"]
        out["W"].free()
        out["rhs"].free()
        for x in (mod, W):
            x.free()
        res[name] = {"kappa": kappa, "m": m, "d": 64, "init_wit_rep": r, "rounds": rounds,
                     "scheme": "folding" if fold else "snark",
                     "wall_ms": round(min(walls), 2),
                     "phases_ms": {k: round(v, 3) for k, v in phases.items()},
                     "device_ms": round(sum(phases.values()), 2),
                     "wall_over_device": round(min(walls) / max(sum(phases.values()), 1e-9), 3),
                     "all_relations_held": bool(ok),
                     "reference_published_prover_s_at_N128": pub_prover,
                     "reference_published_commit_s_at_N128": pub_commit}
    return res


# The reference's own parameter sets at their ring degree and modulus
# (protocol.rs:28-120), with the published CPU times of their commit and
# norm-check phases in ms (experiments/{A0,A1,A2,B0,B1,B2,C2}.log,
# results.md:4-12): (name, N, MOD_Q, MODULE_SIZE, WIT_LEN, commit,
# (norm_check_1, norm_check_2, norm_check_3) or None)
REF_SETS = (("A0", 128, 1125899904679937, 12, 1 << 19, 79.60, (19.18, 69.51, 30.70)),
            ("A1", 256, 1125899906840833, 7, 1 << 18, 53.93, None),
            ("A2", 512, 1125899906822657, 4, 1 << 17, 34.07, None),
            ("B0", 128, 1125899904679937, 12, 1 << 21, 348.22, (None, 232.61, None)),
            ("B1", 256, 1125899906840833, 7, 1 << 20, 211.56, None),
            ("B2", 512, 1125899906822657, 4, 1 << 19, 135.02, None),
            ("C2", 512, 1125899906822657, 4, 1 << 21, 480.72, (271.60, 897.29, 392.52)))


def reference_sets(sa, steps):
    """At each reference set's ring degree and modulus: the commit of one
    witness row of m ring elements with coefficients in [-2, 2]
    (sample_random_short_mat(INIT_WIT_REP = 1, WIT_LEN, 2) and its commit,
    protocol.rs:132-138) under a uniform key of MODULE_SIZE rows; the forward
    and inverse IncompleteNTT of that row; and, where the reference published
    them, the norm check of the row (norm_check_1, one repetition of
    norm_check_2 with its verifier, norm_check_3; F_q challenges, one row of
    L = m N coefficients).  The published CPU times are the reference's own
    (2 x 48-core Xeon 8468, 64 threads); they are context, not the same
    machine."""
    res = {}
    n = max(steps // 4, 3)
    for name, d, q, kappa, m, pub_commit, pub_nc in REF_SETS:
        ctx = sa.Context(q, d)
        A = ctx.uniform((kappa, m, d), 31, sa.STREAM_A)
        key = ctx.key(A)
        A.free()
        W = ctx.bounded((1, m, d), 2, 31, sa.STREAM_W)
        com = ctx.empty((kappa, 1, d))
        ctx.parallel_dot_series_matrix(key, W, out=com)
        ctx.sync()
        _, ctot, _ = timed(ctx, lambda: ctx.parallel_dot_series_matrix(key, W, out=com), n)
        cms = ctot / n
        ktiles = (m + 127) // 128 * 2
        key_bytes = d // 2 * ((14 * kappa + 15) // 16) * 1024 * ktiles
        # the forward and inverse IncompleteNTT of the row (round trip)
        X = ctx.empty((m, d))
        ctx.copy(X, W)
        fwd, inv = [], []
        for _ in range(n):
            e0, e1, e2 = ctx.event(), ctx.event(), ctx.event()
            ctx.record(e0)
            ctx.convert(X, sa.REP_COEFF, sa.REP_INCOMPLETE)
            ctx.record(e1)
            ctx.convert(X, sa.REP_INCOMPLETE, sa.REP_COEFF)
            ctx.record(e2)
            ctx.sync()
            fwd.append(ctx.elapsed_ms(e0, e1))
            inv.append(ctx.elapsed_ms(e1, e2))
        ntt_bytes = 2 * m * d * 8
        r = {"N": d, "q": q, "kappa": kappa, "m": m,
             "commit_ms": round(cms, 3),
             "commit_key_planes_GB": round(key_bytes / 1e9, 2),
             "commit_key_planes_GB/s": round(key_bytes / (cms * 1e-3) / 1e9, 1),
             "published_cpu_commit_ms": pub_commit,
             "commit_vs_published_cpu": round(pub_commit / cms, 1),
             "ntt_fwd_ms": round(float(np.median(fwd)), 4),
             "ntt_inv_ms": round(float(np.median(inv)), 4),
             "ntt_fwd_hbm_frac": round(ntt_bytes / (float(np.median(fwd)) * 1e-3) / 8e12, 3),
             "ntt_inv_hbm_frac": round(ntt_bytes / (float(np.median(inv)) * 1e-3) / 8e12, 3),
             "ntt_round_trip_exact": bool(np.array_equal(X.download()[:4096], W.download()[0, :4096]))}
        X.free()
        if pub_nc:
            L = m * d
            cs = ctx.challenges(1, L.bit_length() - 1, 0, 5)
            outs = ctx.norm_check(W, cs, 0)
            ctx.sync()
            _, ntot, _ = timed(ctx, lambda: ctx.norm_check(W, cs, 0, outs=outs), n)
            r["norm_check_L"] = L
            r["norm_check_ms"] = round(ntot / n, 3)
            r["norm_check_GB/s"] = round(L * 8 / (ntot / n * 1e-3) / 1e9, 1)
            r["norm_check_ok"] = bool(outs["ok"].download().all())
            r["published_cpu_ms"] = {k: v for k, v in zip(
                ("norm_check_1", "norm_check_2", "norm_check_3"), pub_nc) if v is not None}
            if all(v is not None for v in pub_nc):
                r["norm_check_vs_published_cpu"] = round(sum(pub_nc) / (ntot / n), 1)
            for v in outs.values():
                if v is not None:
                    v.free()
            cs.free()
        key.close()
        for x in (W, com):
            x.free()
        ctx.close()
        res[name] = r
    return res


def timed(ctx, fn, k):
    evs = [ctx.event() for _ in range(k + 1)]
    ctx.sync()
    t0 = time.perf_counter()
    ctx.record(evs[0])
    for i in range(k):
        fn()
        ctx.record(evs[i + 1])
    ctx.sync()
    wall = time.perf_counter() - t0
    per = [ctx.elapsed_ms(evs[i], evs[i + 1]) for i in range(k)]
    return per, ctx.elapsed_ms(evs[0], evs[-1]), wall


def mix_ceilings(ctx):
    """The HBM rate a plain streaming kernel (salsaa_memstream) reaches at the
    read/write mixes of the kernels the roofline objects report: the batched
    step's NTT pass reads 512 B and writes 1024 B per ring element of this
    workload's byte-short witnesses, one stream of it in place (1 read : 2
    write streams); the standalone NTT rewrites its input in place (1 : 1).
    Best over 4, 8 and 16 blocks per CU of the mean of 5 launches, 1 GiB per
    stream."""
    n16 = 1 << 26
    res = {}
    for name, rn, wn in (("ntt_pass", 1, 2), ("in_place_copy", 1, 1)):
        buf = ctx.empty(((rn + wn - 1) * n16 * 2,))
        best = 0.0
        for bpc in (4, 8, 16):
            ctx.memstream(buf, n16, rn, wn, True, bpc)
            _, tot, _ = timed(ctx, lambda: ctx.memstream(buf, n16, rn, wn, True, bpc), 5)
            best = max(best, (rn + wn) * n16 * 16 / (tot / 5 * 1e-3) / 1e9)
        buf.free()
        res[name] = round(best, 1)
    return res


def ntt_schedules(sa, steps):
    """Config 2's NTT (forward + inverse IncompleteNTT of 2^20 ring elements)
    under each radix-exchange schedule: between the register phases of each
    32-point half through LDS, or through DPP quad permutes.  A context per
    schedule (SALSAA_NTT_XCH at creation); HBM fraction at 1024 B per element
    per direction."""
    res = {}
    old = os.environ.get("SALSAA_NTT_XCH")
    for name in ("lds", "dpp"):
        os.environ["SALSAA_NTT_XCH"] = name
        c = sa.Context(sa.Q, 64)
        x = c.uniform((NTT_COUNT, 64), 0, sa.STREAM_NTT)
        evs = [c.event() for _ in range(3)]
        fw, iv = [], []
        for i in range(steps + 3):
            c.record(evs[0])
            c.convert(x, sa.REP_COEFF, sa.REP_INCOMPLETE)
            c.record(evs[1])
            c.convert(x, sa.REP_INCOMPLETE, sa.REP_COEFF)
            c.record(evs[2])
            c.sync()
            if i >= 3:
                fw.append(c.elapsed_ms(evs[0], evs[1]))
                iv.append(c.elapsed_ms(evs[1], evs[2]))
        by = NTT_COUNT * 1024
        f, v = float(np.median(fw)), float(np.median(iv))
        res[name] = {"fwd_ms": round(f, 4), "inv_ms": round(v, 4),
                     "fwd_hbm_frac": round(by / (f * 1e-3) / 8e12, 4),
                     "inv_hbm_frac": round(by / (v * 1e-3) / 8e12, 4)}
        x.free()
        c.close()
    if old is None:
        os.environ.pop("SALSAA_NTT_XCH", None)
    else:
        os.environ["SALSAA_NTT_XCH"] = old
    return res


def progress(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def single_instance(sa, ctx, steps, cpu=False):
    """The full-size single-instance plain step (per-phase and whole)."""
    step = Step(sa, ctx)
    for _ in range(3):
        step.run()
    ctx.sync()
    for _ in range(3):
        step.run(timed_phases=True)
        step.collect_phases()
    per, total_ms, _ = timed(ctx, step.run, steps)
    alg = algorithmic(step)
    phases = {}
    for k, v in step.phase_ms.items():
        v = float(np.mean(v))
        a = alg[k]
        e = {"ms": round(v, 4), "GB/s": round(a["bytes"] / (v * 1e-3) / 1e9, 1)}
        if "modmac" in a:
            e["GmodMAC/s"] = round(a["modmac"] / (v * 1e-3) / 1e9, 1)
        phases[k] = e
    out = {"workload": "single instance, plain per-call GPU path: NTT 2^20 ring elements; "
                       "commit A 32x16384 . W 8x16384 in [-16,16]; norm check 8x2^20, F_q^2 "
                       "challenges, 2 reps; decomp_ell base 2^8 x 6 + verify_decomp",
           "value": round(1000.0 * steps / total_ms, 3), "unit": "prover-steps/s",
           "ms_per_step": round(total_ms / steps, 4),
           "p50_ms": round(float(np.percentile(per, 50)), 4),
           "p99_ms": round(float(np.percentile(per, 99)), 4),
           "bit_exact_self_check": step.check(), "phases": phases}
    if cpu:
        progress("CPU baseline, config-2 step")
        out["cpu_baseline"] = cpu_baseline(step)
    step.key.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-single", action="store_true", help="skip the single-instance step")
    ap.add_argument("--no-served", action="store_true", help="skip the served stream")
    ap.add_argument("--no-folding", action="store_true", help="skip the folding step")
    ap.add_argument("--no-projection", action="store_true",
                    help="skip the projection / batching / cross terms")
    ap.add_argument("--no-crs", action="store_true", help="skip the CRS key generation")
    ap.add_argument("--no-protocol", action="store_true", help="skip the protocol runs")
    ap.add_argument("--no-ref-sets", action="store_true",
                    help="skip the reference's parameter sets at their own ring degree")
    ap.add_argument("--eager", action="store_true",
                    help="run the batched steps eagerly instead of replaying the hipGraph "
                         "(for counter collection)")
    ap.add_argument("--hbm-mix", action="store_true",
                    help="print the HBM rate of a plain streaming kernel at several read/write mixes")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: salsaa_amd runs the prover path on a single "
              "MI355X and has no multi-GPU mode (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import salsaa_amd as sa
    ctx = sa.Context(sa.Q, 64)
    if args.hbm_mix:
        n16 = 1 << 26  # 1 GiB per stream
        res = []
        for rn, wn, ip in ((1, 0, False), (0, 1, False), (1, 1, False), (1, 1, True), (1, 2, False),
                           (1, 2, True), (7, 15, False), (7, 15, True)):
            buf = ctx.empty(((rn + wn - (1 if ip else 0)) * n16 * 2,))
            for bpc in (4, 8, 16):
                ctx.memstream(buf, n16, rn, wn, ip, bpc)
                _, tot, _ = timed(ctx, lambda: ctx.memstream(buf, n16, rn, wn, ip, bpc), 5)
                res.append({"read_streams": rn, "write_streams": wn, "in_place": ip,
                            "blocks_per_cu": bpc,
                            "GB/s": round((rn + wn) * n16 * 16 / (tot / 5 * 1e-3) / 1e9, 1)})
            buf.free()
        print(json.dumps({"hbm_mix": res, "stream_bytes": n16 * 16, "launches": 5}), flush=True)
        return
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 8
        src, dst = ctx.uniform((n,), 0, 1), ctx.empty((n,))
        for _ in range(5):
            ctx.copy(dst, src)
        ctx.sync()
        print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 5}), flush=True)
        return

    bw = Batched(sa, ctx)
    b = bw.b
    n = b.n
    # warm-up (the first graph call captures the step)
    graph = not args.eager
    for _ in range(max(args.warmup, 1)):
        b.step(graph=graph)
    b.step_plain()
    ctx.sync()
    # per-phase timing of the batched step (eager, events between phases),
    # then for comparison the norm checks from the witnesses alone and the
    # unfused forward NTT and planes (the inverse after the plain forward
    # restores the Coefficient witnesses)
    names = list(b.PHASES) + ["norm_check_witness", "ntt_fwd_planes", "ntt_inv", "ntt_fwd",
                              "ntt_inv", "planes"]
    ev = [ctx.event() for _ in range(len(names) + 1)]
    phase_ms = {k: [] for k in b.PHASES}
    unfused_ms = {k: [] for k in b.UNFUSED}
    for _ in range(3):
        ctx.record(ev[0])
        for k, name in enumerate(names):
            b.phase(name)
            ctx.record(ev[k + 1])
        ctx.sync()
        for k, name in enumerate(names):
            ms = ctx.elapsed_ms(ev[k], ev[k + 1])
            if k < len(b.PHASES):
                phase_ms[name].append(ms)
            elif name in unfused_ms:
                unfused_ms[name].append(ms)
    # timed region: K batched steps (graph replays)
    per, total_ms, wall = timed(ctx, lambda: b.step(graph=graph), args.steps)
    # the plain per-call path on the same instances
    pper, ptotal_ms, _ = timed(ctx, b.step_plain, args.plain_steps)
    exact = bw.bit_exact()
    progress("batched step %.2f ms, plain %.2f ms" % (total_ms / args.steps,
                                                      ptotal_ms / args.plain_steps))
    cpu_batched = None
    if not args.no_cpu_baseline:
        progress("CPU baseline, one length class of the batch")
        cpu_batched = cpu_baseline_batched(bw)

    sum_m = sum(bw.m)
    # fused NTT round trip + planes + short copy: the arena read and written
    # once (8-byte residues), 64 slots x 7 digit bytes and 64 one-byte short
    # entries written per ring element (every witness residue of this
    # workload lies in [-16, 16], so every slice of 8 ring elements is kept
    # one byte wide), and one format byte per slice
    ntt_bytes = sum_m * 64 * 8 * 2 + sum_m * 64 * 7 + sum_m * 64 + sum_m // 8
    ntt_ms = float(np.mean(phase_ms["ntt_roundtrip_planes"]))
    ntt_gbs = ntt_bytes / (ntt_ms * 1e-3) / 1e9
    peak = 8000.0
    roofline = {"bound": "hbm",
                "kernel": "k_ntt_rt_planes (IncompleteNTT forward, commit digit planes, "
                          "IncompleteNTT inverse, short copy; whole arena)",
                "achieved": round(ntt_gbs, 1), "peak": peak, "unit": "GB/s",
                "frac": round(ntt_gbs / peak, 4), "traffic": traffic_bytes(),
                "algorithmic_bytes": ntt_bytes,
                "per_launch_ms": round(ntt_ms, 4)}
    # the standalone IncompleteNTT forward over the same arena (the kernel
    # the per-call path runs), timed in the same eager sequence
    std_bytes = sum_m * 64 * 8 * 2
    std_ms = float(np.mean(unfused_ms["ntt_fwd"]))
    std_gbs = std_bytes / (std_ms * 1e-3) / 1e9
    roofline_ntt = {"bound": "hbm", "kernel": "k_ntt_inc<false> (IncompleteNTT forward, whole arena, "
                                             "not part of the batched step)",
                    "achieved": round(std_gbs, 1), "peak": peak, "unit": "GB/s",
                    "frac": round(std_gbs / peak, 4), "algorithmic_bytes": std_bytes,
                    "per_launch_ms": round(std_ms, 4)}
    value = 1000.0 * n * args.steps / total_ms
    plain_value = 1000.0 * n * args.plain_steps / ptotal_ms
    out = {
        "metric": "prover-steps/s + p50/p99 ms, batched-served vs plain per-call GPU, bit-exact",
        "value": round(value, 2),
        "unit": "prover-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 4),
        "p50_ms": round(float(np.percentile(per, 50)), 4),
        "p99_ms": round(float(np.percentile(per, 99)), 4),
        "wall_s": round(wall, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 mod q (exact fp64 modmul, int8-digit MFMA commit)",
        "data": "synthetic, seeded on device",
        "config": {"workload": "batched: 512 independent prover steps per step (NTT fwd+inv, "
                               "commit, F_q^2 norm check x2 reps) over one resident key "
                               "A 32 x 2^18; m = 2^k, k uniform in 10..18; kappa uniform in "
                               "{8,16,32}; one witness row per instance in [-16,16]",
                   "d": 64, "q": sa.Q, "instances": n, "sum_m": sum_m,
                   "path": "batched+graph" if graph else "batched (eager)"},
        "plain": {"value": round(plain_value, 2), "unit": "prover-steps/s",
                  "ms_per_step": round(ptotal_ms / args.plain_steps, 4),
                  "steps": args.plain_steps,
                  "path": "native loop of per-instance operator calls"},
        "speedup_vs_plain": round(value / plain_value, 3),
        "bit_exact_vs_plain": exact,
        "tuning": b.tuning(),
        "phases_ms": {k: round(float(np.mean(v)), 4) for k, v in phase_ms.items()},
        "unfused_ms": {k: round(float(np.mean(v)), 4) for k, v in unfused_ms.items()},
        "roofline": roofline,
        "roofline_ntt_standalone": roofline_ntt,
    }
    b.close()
    bw.key.close()
    ceil = mix_ceilings(ctx)
    roofline["mix_ceiling"] = {
        "GB/s": ceil["ntt_pass"], "frac": round(roofline["achieved"] / ceil["ntt_pass"], 4),
        "mix": "plain streaming kernel (salsaa_memstream), 1 read : 2 write streams, one in "
               "place: the pass's 512 B read and 1024 B written per ring element"}
    roofline_ntt["mix_ceiling"] = {
        "GB/s": ceil["in_place_copy"],
        "frac": round(roofline_ntt["achieved"] / ceil["in_place_copy"], 4),
        "mix": "plain streaming kernel (salsaa_memstream), in-place copy (1 read : 1 write)"}
    if cpu_batched is not None:
        out["cpu_baseline"] = cpu_batched
    if not args.no_served:
        progress("served stream")
        out["served"] = served(sa, ctx)
    if not args.no_folding:
        progress("folding step")
        out["folding"] = folding(sa, ctx, args.steps)
    if not args.no_projection:
        progress("projection")
        out["projection"] = projection(sa, ctx, args.steps,
                                       0.0 if args.no_cpu_baseline else 8.0)
    if not args.no_crs:
        progress("CRS key")
        out["crs"] = crs(sa, ctx, args.steps)
    if not args.no_protocol:
        progress("protocol runs")
        out["protocol"] = protocol(sa, ctx)
    if not args.no_ref_sets:
        progress("reference sets")
        out["reference_sets"] = reference_sets(sa, args.steps)
    if not args.no_single:
        progress("NTT schedules")
        out["ntt_schedules"] = ntt_schedules(sa, args.steps)
        progress("single instance")
        out["single_instance"] = single_instance(sa, ctx, args.steps,
                                                 cpu=not args.no_cpu_baseline)
        if "cpu_baseline" in out["single_instance"]:
            out["cpu_baseline_single_instance"] = out["single_instance"].pop("cpu_baseline")
    print(json.dumps(out), flush=True)
    if not exact:
        sys.exit(1)


if __name__ == "__main__":
    main()
