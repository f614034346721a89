"""Rocking-curve throughput of the slice loop on one MI355X.

Workload (the headline configuration): a seeded synthetic Si-surface structure with N = 521
beams, 12 surface atoms of Z = 14, 8, 1, 40 glancing angles 0.5-4.4 deg, sp4 (int_method 5)
at dz = 0.02 A.  One step = one full rocking curve of that structure (the 40 angles as
concurrent batches on separate streams, two of 20 angles for both paths)
with inputs (bulk reflection matrices, gamma, potential table) resident in HBM.  value = points (angles) per second of the mixed-precision path; the FP64 path is
timed the same way.

Also reported: the dominant kernel's roofline (the int8 digit-sliced stage product,
3 ND (ND + 1) N^3 int8 ops per angle against the 5.0 POP/s dense int8 peak; the FP64 stage
product, 8 N^3 flops per angle against 78.6 TFLOP/s, for the FP64 path) from HIP events
around every launch of that kernel on its stream, in one untimed curve run as a single batch
(no concurrent stream shares the device, so the bracketed time is the kernel's own), with
the concurrent-batch figures of the timed steps beside it; the small-beam configurations C1
and C2 (C2's 896 solves in one fused launch); the uneven-size batch in one captured graph;
the serving sample; and the reference's own CPU program (or, where it is not built, the C
oracle) timed on a bounded sample of the same workload on this host.
"""
import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "trhepd-opt_amd"))

METRIC = "rocking-curve points/s at N=521 beams, mixed vs FP64 GPU @1e-6 rel; p99 lat"
PEAK_FP64 = 78.6   # TFLOP/s, MI355X dense FP64 matrix
PEAK_INT8 = 5000.0   # TOP/s, MI355X dense int8 matrix


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# HBM bytes per launch of the int8 stage product at N = 521, 40 angles, 6 digits, from the
# committed counter summary (rocprofv3 FETCH_SIZE and WRITE_SIZE over the kernel alone, each
# scaled by its reading for a plain copy of known size)
TRAFFIC_PROFILE = "profiles/k_stage_i8l_n521_40angles_counters.md"


def stage_i8_traffic():
    try:
        with open(os.path.join(REPO, TRAFFIC_PROFILE)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(st, blocks, method, dz2, threads, sample_slices):
    """The oracle on a bounded sample: `sample_slices` slices of one angle of the same
    deck (same N, same per-slice cost), scaled to a full point."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    import numpy as np
    import oracle as O
    bp = st.bp
    bp.domains[0].blocks = blocks
    surf = O.setup_surface(bp, st.surf, dz2, method)
    ns_full = surf.ns
    surf.ns = sample_slices
    v, vi = O.table(surf)
    r1, r2, ga, az = surf.angles()[0]
    g = O.gamma(surf, ga, az)
    f = surf.scatter_f(r1, r2)
    t0 = time.perf_counter()
    O.integrate(surf, f, g, v, vi)
    dt = time.perf_counter() - t0
    per_point = dt * ns_full / sample_slices
    return {"value": 1.0 / per_point, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": "oracle (C, OpenMP %d threads), 1 angle x %d of %d slices of the same N=521 deck, "
                      "%.1f s measured, scaled by %d/%d slices" % (threads, sample_slices, ns_full, dt, ns_full,
                                                                   sample_slices)}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_reference(st, blocks, method, dz2, threads, trials):
    """The reference's own OpenMP + MKL program (sim-trhepd-rheed surf.exe, built from its
    sources by oracle/ref.mk) on the first angle of the same deck: the same N, slab and dz,
    so the same per-point work.  At N >= 100 the reference integrates angles one after
    another with MKL's threads inside each product (surflf.f90:73-78), so one angle's time is
    its time per point.  Time = the omp_get_wtime delta the program prints
    (surfm.f90:95-96), median over `trials` runs (test/test23/perf-test.sh:21-35 uses 5)."""
    import copy
    import statistics
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import refrun
    bp = copy.deepcopy(st.bp)
    bp.ng = 1
    bp.domains[0].blocks = [[blocks[0][0]]]
    surf_txt, bulkp = st.surf.format().encode(), bp.to_bytes()
    times = []
    for k in range(trials):
        _, t, wall = refrun.surf(surf_txt, bulkp, dz2, method, threads=threads, timeout=600)
        times.append(t if t else wall)
        log("reference CPU trial %d: %.2f s" % (k, times[-1]))
    med = statistics.median(times)
    return {"value": 1.0 / med, "unit": "points/s", "cores": threads, "kind": "reference",
            "sample": "sim-trhepd-rheed surf.exe built from the reference's sources (amdflang -O3 -fopenmp, "
                      "MKL LP64 + GNU threading), sp4 dz=%.2f, 1 of the 40 angles of the same N=%d deck (%d "
                      "trials, median of the program's own omp_get_wtime, OMP_NUM_THREADS=%d): %s s"
                      % (dz2, bp.nb[0], trials, threads, ", ".join("%.2f" % t for t in times)),
            "cpu": _cpu_model()}


def run_path(eng, st, blocks, args, method, dz2):
    """Times `args.steps` full curves of one precision path; returns the figures and the
    intensities of the last step."""
    import numpy as np
    from trhepd import driver
    eng.set_renorm_period(args.renorm_period)
    t0 = time.perf_counter()
    job = driver.SurfaceJob(eng, st.bp, st.surf, dz2, method)
    G, F, W, A = job.inputs(0, blocks)
    nang, N = G.shape
    groups = np.array_split(np.arange(nang), max(1, min(args.split, nang)))
    bids = [eng.batch_create(job.table_id, job.iv, G[g], F[g], job.nb0, W[g]) for g in groups]
    log("setup: N=%d nv=%d ns=%d dz=%.5f angles=%d in %d concurrent batches (%.1f s)"
        % (N, job.nv, job.ns, job.dz, nang, len(bids), time.perf_counter() - t0))
    for _ in range(args.warmup):
        eng.batch_run_many(bids)
        log("warmup step: device %.1f ms, renorm %d" % (eng.last_stats()["device_ms"], eng.last_stats()["n_renorm"]))
    eng.profile(True)
    t0 = time.perf_counter()
    ren = 0
    for k in range(args.steps):
        eng.batch_run_many(bids)
        ren += eng.last_stats()["n_renorm"]
        log("step %d done" % k)
    elapsed = time.perf_counter() - t0
    launches, kms = eng.kernel_stats()
    eng.profile(False)
    f2 = np.concatenate([eng.batch_fetch(b, want_f=False)[1] for b in bids])
    assert np.all(np.isfinite(f2))
    for b in bids:
        eng.batch_release(b)
    out = dict(value=nang * args.steps / elapsed, ms=1e3 * elapsed / args.steps, launches=launches, kms=kms,
               f2=f2, ren=ren / (nang * args.steps), N=N, nang=nang, ns=job.ns, dz=job.dz, nb0=job.nb0,
               per_launch=nang / len(bids), nbatch=len(bids))
    job.release()
    return out


def run_small_configs(eng, args):
    """The small-beam configurations: C1 (one N=23 structure, 56 angles) and C2 (16 N=47
    structures x 56 angles, run as one fused launch of the small-N kernel), sp4 at dz 0.02,
    FP64, inputs resident.  A step is one batch_run_many over all of a configuration's
    batches; points/s counts its (structure, angle) solves."""
    import numpy as np
    from trhepd import driver, lib, synth
    out = {}
    for name, cfg, count in (("c1", 1, 1), ("c2", 2, 16)):
        sts = [synth.config(cfg, seed=args.seed, index=i) for i in range(count)]
        blocks = synth.bulk_reflection(eng, sts[0])
        jobs, bids = [], []
        try:
            for st in sts:
                job = driver.SurfaceJob(eng, st.bp, st.surf, 0.02, lib.SP4)
                jobs.append(job)
                G, F, W, A = job.inputs(0, blocks)
                bids.append(eng.batch_create(job.table_id, job.iv, G, F, job.nb0, W))
            eng.set_renorm_period(0)
            eng.batch_run_many(bids)   # warm-up
            t0 = time.perf_counter()
            ren = 0
            for _ in range(args.small_steps):
                eng.batch_run_many(bids)
                ren += eng.last_stats()["n_renorm"]
            dt = (time.perf_counter() - t0) / args.small_steps
            dev_ms = eng.last_stats()["device_ms"]
            N, ns, nang = jobs[0].nb, jobs[0].ns, len(A)
        finally:
            for b in bids:
                eng.batch_release(b)
            for job in jobs:
                job.release()
        pts = count * nang
        # 8 N^3 flops per stage product (4M complex count), 6 stages per slice; the
        # renormalisation solves (about 14.7 N^3 each) on top
        flops = pts * 8.0 * N ** 3 * 6 * ns + ren / args.small_steps * 14.7 * N ** 3
        ach = flops / (dev_ms * 1e-3) / 1e12
        out[name] = {"workload": "%d seeded structure(s) x %d angles, N=%d, sp4 dz=0.02 ns=%d, reference "
                                 "renormalisation rule, %s" % (count, nang, N, ns, "one fused launch of the small-N "
                                                              "kernel" if count > 1 else "one launch"),
                     "points": pts, "points_per_s": pts / dt, "ms_per_step": 1e3 * dt, "device_ms": dev_ms,
                     "roofline": {"bound": "mfma", "kernel": "k_small / k_small_multi (one workgroup per angle, "
                                  "state in LDS, v_mfma_f64_16x16x4f64)", "achieved": ach, "peak": PEAK_FP64,
                                  "unit": "TFLOP/s", "frac": ach / PEAK_FP64, "traffic": None}}
        log("small-N %s: %d points in %.1f ms (device %.1f ms): %.0f points/s, %.2f TFLOP/s"
            % (name, pts, 1e3 * dt, dev_ms, pts / dt, ach))
    return out


def run_serving(args, precision, pool, budget):
    """The serving workload: a bounded sample of the uneven-size Poisson request stream,
    served by the library's continuous-batching loop under an HBM budget."""
    import numpy as np
    from trhepd import lib, serve
    reqs = [pool[i % len(pool)] for i in range(args.serve_requests)]
    arr = serve.poisson_arrivals(len(reqs), args.serve_rate, seed=args.seed)
    eng = lib.Engine(0, precision=precision, hbm_budget=budget)
    eng.set_renorm_period(args.renorm_period)
    if precision == lib.PREC_MIXED:
        eng.set_mixed_digits(args.mixed_digits)   # as the curve and uneven legs
    eng.serve_set_merge(args.serve_merge_angles, args.serve_merge_window_ms)
    try:
        run = serve.serve(eng, reqs, arrivals=arr, timeout_s=args.serve_timeout_s,
                          progress=lambda d, n, a, q: log("serving: %d of %d requests finished, %d admitted, "
                                                          "%d queued" % (d, n, a, q)))
    finally:
        eng.close()
    out = serve.summary(reqs, run)
    out["f2"] = [r[1] for r in run["results"]]
    return out


def serving_leg(eng64, args, paths):
    """The serving stream for each precision path, with the mixed results' error vs FP64."""
    import numpy as np
    from trhepd import lib, serve
    t0 = time.perf_counter()
    pool = serve.request_pool(eng64, args.serve_pool, seed=args.seed)
    log("serving pool: %d structures, N %s, angles %s (%.1f s)" % (len(pool), [r.nb for r in pool],
                                                                   [r.nang for r in pool], time.perf_counter() - t0))
    budget = int(args.serve_budget_gb * (1 << 30))
    sv = {"workload": "first %d requests of a Poisson stream (mean %.0f req/s) over %d seeded structures with N "
                      "from {23,47,97,199,307,521}, 3-12 atoms, 8-20 A slabs, 20-56 angles, sp4 dz=0.02; "
                      "continuous batching under a %.0f GiB HBM budget, requests of one structure merged into "
                      "batches of up to %d angles (%.0f ms window); mixed path at %d digits"
                      % (args.serve_requests, args.serve_rate, len(pool), args.serve_budget_gb,
                         args.serve_merge_angles, args.serve_merge_window_ms, args.mixed_digits),
          "beams": [r.nb for r in pool], "angles": [r.nang for r in pool]}
    f2s = {}
    for p in paths:
        r = run_serving(args, lib.PREC_FP64 if p == "fp64" else lib.PREC_MIXED, pool, budget)
        f2s[p] = r.pop("f2")
        sv[p] = r
        log("serving %s: %.2f points/s p50 %.0f ms p99 %.0f ms (queued p50 %.0f ms, service p50 %.0f ms; "
            "peak %d admitted)" % (p, r["points_per_s"], r["p50_ms"], r["p99_ms"], r["queued_ms"]["p50"],
                                   r["service_ms"]["p50"], r["peak_admitted"]))
    if "fp64" in f2s and "mixed" in f2s:
        errs = []
        for a, b in zip(f2s["mixed"], f2s["fp64"]):
            mx = np.max(np.abs(b), axis=1, keepdims=True)
            errs.append(float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12 * mx))))
        sv["mixed_max_rel_err_vs_fp64"] = max(errs)
        sv["mixed_over_fp64"] = sv["mixed"]["points_per_s"] / sv["fp64"]["points_per_s"]
    return sv


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-one-batch", action="store_true",
                    help="skip the untimed single-batch curve that measures the kernels without stream sharing")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-slices", type=int, default=256, help="slices the C oracle fallback times")
    ap.add_argument("--cpu-trials", type=int, default=3, help="runs of the reference CPU program (median)")
    ap.add_argument("--renorm-period", type=int, default=-1,
                    help="0: the reference's Gershgorin rule; K: every K slices; -1: K from the evanescent growth rate")
    ap.add_argument("--split", type=int, default=0,
                    help="angle groups run concurrently on separate streams (overlaps the LU chains); "
                         "0: 2 for both paths (the fastest layout measured: 2-8 groups swept on N=521)")
    ap.add_argument("--paths", default="fp64,mixed", help="precision paths to time (fp64, mixed)")
    ap.add_argument("--mixed-digits", type=int, default=6,
                    help="digits per operand of the mixed path's int8 stage products (5, 6 or 7); the run "
                         "checks the mixed intensities against its own FP64 ones")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run a plain device copy of this many MiB (calibrates HBM byte counters)")
    ap.add_argument("--uneven-structures", type=int, default=16,
                    help="structures of the uneven-size graph-captured measurement (0: skip)")
    ap.add_argument("--uneven-steps", type=int, default=2, help="timed launches of the uneven-batch graph")
    ap.add_argument("--small-steps", type=int, default=3,
                    help="timed runs of the small-beam configurations C1 and C2 (0: skip)")
    ap.add_argument("--small-only", action="store_true",
                    help="time only the small-beam configurations C1 and C2 (for a kernel trace of the "
                         "small-N kernel)")
    ap.add_argument("--serve-requests", type=int, default=64,
                    help="requests of the serving sample (0: skip the serving measurement)")
    ap.add_argument("--serve-pool", type=int, default=8, help="distinct seeded structures the requests cycle over")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="mean arrival rate (requests/s)")
    ap.add_argument("--serve-budget-gb", type=float, default=24.0, help="HBM budget of the serving context")
    ap.add_argument("--serve-merge-angles", type=int, default=512,
                    help="requests of one structure merge into batches of at most this many angles (0: off)")
    ap.add_argument("--serve-merge-window-ms", type=float, default=200.0,
                    help="how long the head of the queue waits for requests of its structure to join it")
    ap.add_argument("--serve-timeout-s", type=float, default=900.0, help="longest a serving run may take")
    ap.add_argument("--serve-only", action="store_true",
                    help="time only the serving stream (--serve-requests requests, the --paths given), for the "
                         "full-size stream whose device time exceeds a default run's")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the slice loop shards by structure and angle with no "
              "exchange, so the multi-GPU figure is N x the one-GPU value; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from trhepd import lib, synth

    method, dz2 = lib.SP4, 0.02
    eng64 = lib.Engine(0, precision=lib.PREC_FP64)
    if args.small_only:
        small = run_small_configs(eng64, argparse.Namespace(**dict(vars(args), small_steps=max(1, args.small_steps))))
        eng64.close()
        print(json.dumps({"small_n": small}))
        return
    if args.serve_only:
        paths = [p.strip() for p in args.paths.split(",") if p.strip()]
        sv = serving_leg(eng64, args, paths)
        eng64.close()
        print(json.dumps({"serving": sv}))
        return
    st = synth.config(3, seed=args.seed)
    t0 = time.perf_counter()
    blocks = synth.bulk_reflection(eng64, st, dz2=dz2, method=method)
    log("bulk reflection (%d bulk layers, device, FP64): %.1f s" % (st.bulk.ml, time.perf_counter() - t0))
    paths = [p.strip() for p in args.paths.split(",") if p.strip()]
    res = {}
    if args.copy_calibration_mb > 0:
        import torch
        n = args.copy_calibration_mb << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 3 copies of %d MiB" % args.copy_calibration_mb)
        del src, dst
    for p in paths:
        eng = eng64
        if p != "fp64":
            eng = lib.Engine(0, precision=lib.PREC_MIXED)
            eng.set_mixed_digits(args.mixed_digits)
        log("path %s" % p)
        pargs = argparse.Namespace(**vars(args))
        pargs.split = args.split or 2
        res[p] = run_path(eng, st, blocks, pargs, method, dz2)
        if pargs.split > 1 and not args.no_one_batch:
            # the same kernels in one batch of all angles, one untimed curve: per-launch
            # durations without a concurrent stream sharing the device
            one = argparse.Namespace(**vars(args))
            one.split, one.steps, one.warmup = 1, 1, 0
            res[p]["one_batch"] = run_path(eng, st, blocks, one, method, dz2)
        if eng is not eng64:
            eng.close()
    head = res.get("mixed", res.get("fp64"))
    N, nang = head["N"], head["nang"]
    rule = {0: "Gershgorin cond_est > 1e3 (reference)",
            -1: "periodic, K = floor(ln 1e3 / (max|Im gamma| dz))"}.get(args.renorm_period,
                                                                      "every %d slices" % args.renorm_period)
    cfg = {"workload": "synthetic Si surface, N=%d beams, %d angles 0.5-4.4 deg, 12 surface atoms (Z=14,8,1), "
                       "sp4 dz=%.4f ns=%d" % (N, nang, head["dz"], head["ns"]),
           "angles": nang, "beams": N, "slices": head["ns"], "renorm_rule": rule,
           "concurrent_batches": head["nbatch"], "paths": {}}
    for p, r in res.items():
        cfg["paths"][p] = {"points_per_s": r["value"], "ms_per_step": r["ms"], "renorm_per_point": r["ren"],
                           "concurrent_batches": r["nbatch"]}
    if "mixed" in res and "fp64" in res:
        a, b = res["mixed"]["f2"], res["fp64"]["f2"]
        mx = np.max(np.abs(b), axis=1, keepdims=True)
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-12 * mx)
        cfg["mixed_over_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
        cfg["mixed_max_rel_err_vs_fp64"] = float(rel.max())
        cfg["mixed_within_1e-6"] = bool(rel.max() <= 1e-6)
        cfg["mixed_digits"] = args.mixed_digits
        cfg["mixed_l2_err_vs_fp64"] = float(np.max(np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)))

    def roofline(r, mixed, shared=True):
        avg_ms = r["kms"] / max(r["launches"], 1)
        if mixed:
            # nd (nd + 1) / 2 digit pairs x 3 real products x 2 N^3 int8 ops per angle
            nd = args.mixed_digits
            ops = 3.0 * nd * (nd + 1) * N ** 3 * r["per_launch"]
            peak, unit = PEAK_INT8, "TOP/s"
            kern = ("k_stage_i8l (int8 digit-sliced stage product, %d digits: %d N^3 int8 ops per angle)"
                    % (nd, 3 * nd * (nd + 1)))
        else:
            ops = 8.0 * N ** 3 * r["per_launch"]
            peak, unit, kern = PEAK_FP64, "TFLOP/s", "k_stage_gemm (FP64 stage product, 8 N^3 flops per angle)"
        ach = ops / (avg_ms * 1e-3) / 1e12
        note = ("; concurrent batches overlap, so the per-launch duration is an upper bound on the kernel's own "
                "time" if shared and r["nbatch"] > 1 else "; one batch, no concurrent stream")
        out = {"bound": "mfma", "kernel": kern + note,
               "achieved": ach, "peak": peak, "unit": unit, "frac": ach / peak, "traffic": None,
               "launches": r["launches"], "avg_launch_ms": avg_ms}
        # the counted traffic is for the shape of the one-batch launch (40 angles of N = 521)
        if mixed and args.mixed_digits == 6 and N == 521 and r["per_launch"] == 40:
            out["traffic"] = stage_i8_traffic()
            if out["traffic"] is not None:
                alg = (3 * 6 * (r["per_launch"] * 528 // 16) * 9 * 1024   # Q digits
                       + 3 * 6 * 33 * 9 * 1024                          # one stage's M0 digits
                       + 64.0 * 528 * 528 * r["per_launch"])             # FP64 state read + written
                out["traffic_source"] = (TRAFFIC_PROFILE + ": FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (scale "
                                         "factors from a 1 GiB copy), the kernel alone on seeded operands of "
                                         "this shape; algorithmic bytes %.3g" % alg)
        return out

    out = {
        "metric": METRIC, "value": head["value"], "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "i8 digit-sliced products + f64" if "mixed" in res else "f64", "data": "synthetic",
        "config": cfg,
    }
    # the headline roofline: the dominant kernel's own launch time, measured in a curve run
    # as one batch (no concurrent stream shares the device, so launches do not overlap);
    # the concurrent-batch figures, whose event brackets overlap, are reported beside it
    if "one_batch" in head:
        out["roofline"] = roofline(head["one_batch"], "mixed" in res, shared=False)
        out["roofline_concurrent"] = roofline(head, "mixed" in res)
    else:
        out["roofline"] = roofline(head, "mixed" in res, shared=head["nbatch"] > 1)
    if "mixed" in res and "fp64" in res:
        f64 = res["fp64"]
        out["roofline_fp64"] = roofline(f64["one_batch"], False, shared=False) if "one_batch" in f64 else \
            roofline(f64, False)
    if args.small_steps > 0:
        out["small_n"] = run_small_configs(eng64, args)
    if args.uneven_structures > 0:
        from trhepd import uneven
        t0 = time.perf_counter()
        ureqs = uneven.requests(eng64, args.uneven_structures, seed=args.seed, log=log)
        log("uneven set: %d structures, N %s (%.1f s)" % (len(ureqs), [r.nb for r in ureqs], time.perf_counter() - t0))
        uv = {"workload": "first %d of the seeded uneven-size structures: N from {23,47,97,199,307}, 3-12 atoms, "
                          "8-20 A slabs, 28-56 angles, sp4 dz=0.02, periodic renormalisation; all batches in one "
                          "captured HIP graph, stage tiles autotuned per shape" % len(ureqs),
              "beams": [r.nb for r in ureqs], "angles": [r.nang for r in ureqs]}
        uf2 = {}
        for p in paths:
            eng = lib.Engine(0, precision=lib.PREC_FP64 if p == "fp64" else lib.PREC_MIXED)
            eng.set_renorm_period(-1)
            if p != "fp64":
                eng.set_mixed_digits(args.mixed_digits)
            try:
                r = uneven.measure(eng, ureqs, steps=args.uneven_steps, warmup=1, log=log)
            finally:
                eng.close()
            uf2[p] = r.pop("f2")
            uv[p] = r
            log("uneven %s: %.2f points/s (%.1f ms per launch)" % (p, r["points_per_s"], r["ms_per_launch"]))
        if "fp64" in uf2 and "mixed" in uf2:
            errs = []
            for a, b in zip(uf2["mixed"], uf2["fp64"]):
                mx = np.max(np.abs(b), axis=1, keepdims=True)
                errs.append(float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12 * mx))))
            uv["mixed_max_rel_err_vs_fp64"] = max(errs)
            uv["mixed_over_fp64"] = uv["mixed"]["points_per_s"] / uv["fp64"]["points_per_s"]
        out["uneven"] = uv
    if args.serve_requests > 0:
        sv = serving_leg(eng64, args, paths)
        out["serving"] = sv
        out["p99_ms"] = sv.get("mixed", sv.get("fp64"))["p99_ms"]
    if not args.no_cpu:
        # the reference's own build when oracle/ref.mk has built it, else the C oracle
        try:
            out["cpu_baseline"] = cpu_reference(st, blocks, method, dz2, args.cpu_threads, args.cpu_trials)
        except Exception as e:   # the CPU leg is reported, never fatal to the GPU figure
            log("reference CPU build unavailable (%r): timing the C oracle instead" % (e,))
            try:
                out["cpu_baseline"] = cpu_baseline(st, blocks, method, dz2, args.cpu_threads, args.cpu_slices)
            except Exception as e2:
                out["cpu_baseline"] = {"value": None, "error": repr(e2)}
    eng64.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
