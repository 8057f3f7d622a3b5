"""Prover-step benchmark on one MI355X.

A prover step (plain per-call GPU path, 1125899906840833-modulus ring of
degree 64), at the full-size single-instance workload:
  1. forward + inverse IncompleteNTT of 2^20 ring elements (uniform mod q)
  2. commit A . W, A in R_q^{32 x 16384} uniform (resident key), W 8 rows of
     16384 ring elements with coefficients uniform in [-16, 16]
  3. norm check of the 8 x 2^20 flattened witness: norm_check_1 claims, two
     repetitions of the 20-round sumcheck with F_{q^2} challenges, verifier
     replay, norm_check_3 rhs
  4. decomp_ell: base-2^8 decomposition of W into 6 limbs (48 rows) and the
     commitment of the limbs, then verify_decomp: the inner product of the
     gadget vector (1, 2^8, ..., 2^40) with the limb commitments must equal
     the commitment (decomp.rs:95-112)
All inputs are generated on the device from seeds before timing starts; every
output stays in HBM.  Prints one JSON line.
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
    """HBM bytes per launch of the NTT forward kernel, measured by rocprofv3
    FETCH_SIZE and WRITE_SIZE (each scaled by a plain-copy calibration) and
    committed in profiles/ntt_inc_fwd_cfg2_traffic.json."""
    path = os.path.join(ROOT, "profiles", "ntt_inc_fwd_cfg2_traffic.json")
    try:
        with open(path) as f:
            return int(json.load(f)["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(budget_s=20.0):
    """The oracle (single thread) on a bounded sample of the same step, scaled
    to one full step: each phase runs on a fraction of its work and the time is
    multiplied by the inverse fraction."""
    from oracle import oracle as orc
    q = 1125899906840833
    t = {}
    # NTT: 65536 of the 2^20 elements, forward + inverse
    nn = 65536
    x = orc.sample_uniform(0, 0xC, (nn, 64), q)
    t0 = time.perf_counter()
    y = orc.to_incomplete(x, q)
    orc.from_incomplete(y, q)
    t["ntt"] = (time.perf_counter() - t0) * (NTT_COUNT / nn)
    # commit: full kappa x r, 64 of the 16384 columns (per-product conversion)
    ms = 128
    A = orc.sample_uniform(0, 0xA, (KAPPA, ms, 64), q)
    W = orc.sample_bounded(0, 0xB, (R, ms, 64), BOUND, q)
    t0 = time.perf_counter()
    orc.commit(A, W, q, threads=1)
    t["commit"] = (time.perf_counter() - t0) * (M / ms)
    # norm check: one of the 8 rows at full length, 2 repetitions, F_{q^2}
    w = orc.sample_bounded(0, 0xB, (1, M * 64), BOUND, q)
    cs = orc.sample_uniform(0, 0x5C000, (REPS, 20, 2), q)
    t0 = time.perf_counter()
    orc.sumcheck(w, cs, q, fq2=True, nr=5, threads=1)
    t["norm_check"] = (time.perf_counter() - t0) * R
    # decomp_ell: decomposition of the commit sample and the limb commitment
    t0 = time.perf_counter()
    limbs, radix = orc.decompose(W, PARTS, q, RADIX)
    orc.commit(A, limbs, q, threads=1)
    t["decomp_ell"] = (time.perf_counter() - t0) * (M / ms)
    total = sum(t.values())
    return {
        "value": 1.0 / total,
        "unit": "prover-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": ("oracle, 1 thread: NTT on 65536 of 2^20 elements; commit and decomp_ell "
                   "with all 32x8 (32x48) rows on 128 of 16384 columns; norm check on 1 of 8 "
                   "rows; each phase scaled to the full step (%s)" %
                   ", ".join("%s %.1fs" % kv for kv in t.items())),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs the prover path on a single "
              "MI355X and has no multi-GPU mode (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import salsaa_amd as sa
    ctx = sa.Context(sa.Q, 64)
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 8
        src, dst = ctx.uniform((n,), 0, 1), ctx.empty((n,))
        for _ in range(5):
            ctx.copy(dst, src)
        ctx.sync()
        print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 5}), flush=True)
        return
    step = Step(sa, ctx)
    for _ in range(args.warmup):
        step.run()
    ctx.sync()
    if not step.check():
        print("prover step self-check failed", file=sys.stderr)
        sys.exit(1)

    # per-phase timing pass (events between phases), not the timed region
    for _ in range(3):
        step.run(timed_phases=True)
        step.collect_phases()
    ctx.sync()

    # timed region: K whole steps, events at each step boundary
    evs = [ctx.event() for _ in range(args.steps + 1)]
    ctx.sync()
    t0 = time.perf_counter()
    ctx.record(evs[0])
    for i in range(args.steps):
        step.run()
        ctx.record(evs[i + 1])
    ctx.sync()
    wall = time.perf_counter() - t0
    per_step = [ctx.elapsed_ms(evs[i], evs[i + 1]) for i in range(args.steps)]
    total_ms = ctx.elapsed_ms(evs[0], evs[-1])
    ok = step.check()

    alg = algorithmic(step)
    phase_avg = {k: float(np.mean(v)) for k, v in step.phase_ms.items()}
    dom = max(phase_avg, key=phase_avg.get)
    ntt_ms = phase_avg["ntt_fwd"]
    ntt_gbs = alg["ntt_fwd"]["bytes"] / (ntt_ms * 1e-3) / 1e9
    peak = 8000.0
    roofline = {"bound": "hbm", "kernel": "k_ntt_inc<false> (IncompleteNTT forward)",
                "achieved": round(ntt_gbs, 1), "peak": peak, "unit": "GB/s",
                "frac": round(ntt_gbs / peak, 4), "traffic": traffic_bytes(),
                "algorithmic_bytes": alg["ntt_fwd"]["bytes"], "dominant_phase": dom}
    phases = {}
    for k, v in phase_avg.items():
        a = alg[k]
        e = {"ms": round(v, 4), "GB/s": round(a["bytes"] / (v * 1e-3) / 1e9, 1)}
        if "modmac" in a:
            e["GmodMAC/s"] = round(a["modmac"] / (v * 1e-3) / 1e9, 1)
        phases[k] = e

    out = {
        "metric": "prover-steps/s + p50/p99 ms, batched-served vs plain per-call GPU, bit-exact",
        "value": round(1000.0 * args.steps / total_ms, 3),
        "unit": "prover-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 4),
        "p50_ms": round(float(np.percentile(per_step, 50)), 4),
        "p99_ms": round(float(np.percentile(per_step, 99)), 4),
        "wall_s": round(wall, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 mod q (exact fp64 modmul)",
        "data": "synthetic, seeded on device",
        "bit_exact_self_check": ok,
        "config": {"workload": "plain per-call GPU prover step, single instance: NTT 2^20 ring "
                               "elements; commit A 32x16384, W 8x16384 in [-16,16]; norm check "
                               "8x2^20, F_q^2 challenges, 2 reps; decomp_ell base 2^8 x 6",
                   "d": 64, "q": sa.Q, "path": "plain"},
        "phases": phases,
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
