#!/usr/bin/env python3
"""Benchmark of the twisted-mass fermion solve on one MI355X.

Workload (one "step"): a full even-odd preconditioned twisted-mass solve
(inv_tmD_cg_eoprec: source preparation, CG on K K^+, reconstruction) on a 32^3x64
lattice, kappa=0.16, mu=0.003, seeded synthetic SU(3) gauge field U=exp(i 0.3 H)
with the antiperiodic time phase, Gaussian spincolor source, to a true relative
residual |r|/|b| = 1e-10 (nissa residue 1e-20) checked in FP64.  The inner
Krylov iteration runs on FP16 spinors with per-site norms and 12-real FP32
links, with FP64 reliable updates.  Gauge field and source are resident in HBM
before the timed region.

Also reported: the plain FP64 solve (same workload), the FP64 single-parity hop
("Dslash") bandwidth over 1000 applies, the roofline of the dominant kernel of
the timed workload (HIP-event ledger kept by the library), and the CPU oracle's
full FP64 solve of the same workload timed on the host cores.  That oracle
solve is also the reference for the error of each precision's GPU solution;
the 8^4 case (kappa 0.16, mu 0.01, Z2 source on the even sites, residue 1e-24)
adds the error against the oracle's __float128 solve, and so does the full
lattice (--no-quad skips that one).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tmQ2 CG solves/s @32^3x64, 1e-10 (mixed vs FP64); Dslash GB/s; serve p99 lat"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lattice", default="64,32,32,32", help="T,LX,LY,LZ")
    ap.add_argument("--prec", default="f16", choices=["f64", "f32", "f16", "f16h"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-quad", action="store_true",
                    help="skip the __float128 oracle solve at the full lattice (about 40 s of host time)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-stag", action="store_true", help="skip the staggered multi-shift leg")
    ap.add_argument("--no-24", action="store_true", help="skip the 24^3x48 FP64 Dslash and solve leg")
    ap.add_argument("--no-clover", action="store_true", help="skip the twisted-clover solve leg")
    ap.add_argument("--serve-requests", type=int, default=60, help="requests in the Poisson stream")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="request arrival rate (1/s)")
    ap.add_argument("--serve-rate-low", type=float, default=6.0,
                    help="arrival rate of a second stream below the capacity (1/s; 0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}): there is no multi-GPU path; "
              "requests and right-hand sides would shard across GPUs with the gauge field replicated, "
              "and a single large solve along T with halo exchange.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import nissa_amd as nq

    L = tuple(int(x) for x in args.lattice.split(","))
    kappa, mu, residue = 0.16, 0.003, 1e-20
    V = int(np.prod(L))
    log(f"[bench] lattice {L} V={V}: generating gauge field")
    t0 = time.time()
    conf = nq.gauge_random_lx(L, 3, 0.3)
    nq.gauge_phase_lx(L, conf, [1.0, 0, 0, 0])
    src = nq.source_lx(L, 17, "gauss")
    log(f"[bench] inputs ready in {time.time() - t0:.1f}s")

    ctx = nq.Ctx(L, 0)
    g = nq.Gauge(ctx, lx=conf)
    fs = ctx.field(nq.SPINCOLOR, nq.LX, src)
    fx = ctx.field(nq.SPINCOLOR, nq.LX)

    def solve(prec):
        st = nq.tm_solve_eoprec(ctx, g, kappa, mu, 100000, residue, fx, fs, prec=prec)
        if st["rc"] != 0 or st["check_rel_res_sq"] > 2 * residue:
            raise RuntimeError(f"solve failed: {st}")
        return st

    # warm-up (also builds the sloppy link copies)
    for _ in range(max(args.warmup, 0)):
        st = solve(args.prec)
        log(f"[bench] warmup {args.prec}: iters={st['iters']} reliable={st['reliable_updates']} "
            f"check={st['check_rel_res_sq']:.2e} t={st['seconds']*1e3:.1f}ms")
    ctx.sync()
    t0 = time.perf_counter()
    stats = [solve(args.prec) for _ in range(args.steps)]
    ctx.sync()
    elapsed = time.perf_counter() - t0
    solves_per_s = args.steps / elapsed
    gpu_sols = {args.prec: fx.download()}  # each precision's solution, for the oracle's error check
    log(f"[bench] timed {args.steps} {args.prec} solves: {elapsed*1e3/args.steps:.2f} ms/solve "
        f"({solves_per_s:.2f} solves/s), iters={stats[-1]['iters']}")

    # plain FP64 solve of the same workload
    solve("f64")
    ctx.sync()
    t0 = time.perf_counter()
    n64 = max(1, min(3, args.steps))
    st64 = [solve("f64") for _ in range(n64)]
    ctx.sync()
    t64 = (time.perf_counter() - t0) / n64
    gpu_sols["f64"] = fx.download()
    log(f"[bench] FP64 solve: {t64*1e3:.2f} ms, iters={st64[-1]['iters']}")

    # the other inner-precision variants of the same workload
    variants = {}
    for pv in ("f32", "f16", "f16h"):
        if pv == args.prec:
            continue
        solve(pv)
        ctx.sync()
        t0 = time.perf_counter()
        sv = [solve(pv) for _ in range(3)]
        ctx.sync()
        tv = (time.perf_counter() - t0) / 3
        gpu_sols[pv] = fx.download()
        variants[pv] = {"ms_per_solve": round(tv * 1e3, 3), "iters": sv[-1]["iters"],
                        "reliable_updates": sv[-1]["reliable_updates"],
                        "true_rel_res_sq": sv[-1]["check_rel_res_sq"]}
        log(f"[bench] {pv} solve: {tv*1e3:.2f} ms, iters={sv[-1]['iters']}")

    # roofline of the dominant kernel over a profiled pass identical to the timed region
    ctx.reset_kernel_stats()
    ctx.profiling(True)
    for _ in range(args.steps):
        solve(args.prec)
    ctx.profiling(False)
    ks = ctx.kernel_stats()
    dom_name, dom = max(ks.items(), key=lambda kv: kv[1]["total_ms"])
    avg_ms = dom["total_ms"] / dom["launches"]
    achieved = dom["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9
    total_ms = sum(v["total_ms"] for v in ks.values())
    for name, v in sorted(ks.items(), key=lambda kv: -kv[1]["total_ms"]):
        a = v["total_ms"] / v["launches"]
        log(f"[bench]   {name:28s} {v['launches']:6d} x {a*1e3:8.1f} us = {v['total_ms']/args.steps:7.2f} ms/solve "
            f"{v['bytes_per_launch']/(a*1e-3)/1e9:7.0f} GB/s")
    log(f"[bench] dominant kernel {dom_name}: {dom['launches']} launches, avg {avg_ms*1e3:.1f} us, "
        f"{achieved:.0f} GB/s, {100*dom['total_ms']/total_ms:.0f}% of kernel time")

    # Dslash: FP64 single-parity hop, 1000 applies
    fo = ctx.field(nq.SPINCOLOR, nq.ODD)
    fe = ctx.field(nq.SPINCOLOR, nq.EVEN)
    ms_hop, bytes_hop = nq.bench_tm_hop(ctx, g, 1000, fe, fo, prec="f64")
    dslash_gbs = bytes_hop / (ms_hop * 1e-3) / 1e9
    ms_hop16, bytes_hop16 = nq.bench_tm_hop(ctx, g, 1000, fe, fo, prec=args.prec)
    log(f"[bench] Dslash FP64: {ms_hop*1e3:.1f} us/apply, {dslash_gbs:.0f} GB/s; "
        f"{args.prec}: {ms_hop16*1e3:.1f} us/apply, {bytes_hop16/(ms_hop16*1e-3)/1e9:.0f} GB/s")

    small = None if args.no_cpu else small_solves()

    serve = None
    stag = None
    lat24 = None
    clov = None
    if not (args.no_serve and args.no_stag and args.no_24 and args.no_clover):
        for f in (fs, fx, fo, fe):
            f.close()
        g.close()
        ctx.close()
        if not args.no_24:
            lat24 = lattice24_bench()
        if not args.no_clover:
            clov = clover_bench()
        if not args.no_serve:
            serve = serve_bench(args)
        if not args.no_stag:
            stag = stag_bench(args)

    # CPU baseline: the oracle's full FP64 solve of the same workload, which also
    # checks every precision's GPU solution (and the 8^4 solves, with __float128)
    cpu = None
    accuracy = None
    if not args.no_cpu:
        cpu, accuracy = cpu_baseline(L, conf, src, kappa, mu, residue, gpu_sols, small, not args.no_quad)

    out = {
        "metric": METRIC,
        "value": round(solves_per_s, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f16": "f32 compute, f16 spinor storage, f64 reliable updates",
                  "f16h": "f32 compute, f16 spinor+link storage, f64 reliable updates",
                  "f32": "f32 compute, f64 reliable updates", "f64": "f64"}[args.prec],
        "data": "synthetic (seeded SU(3) U=exp(i 0.3 H), antiperiodic t; Gaussian source)",
        "config": {
            "workload": f"tm eo-preconditioned CG solve {L[1]}^3x{L[0]}, kappa=0.16, mu=0.003, "
                        f"true residual 1e-10 (nissa residue 1e-20), inner {args.prec}",
            "lattice": list(L),
            "iters": stats[-1]["iters"],
            "reliable_updates": stats[-1]["reliable_updates"],
            "true_rel_res_sq": stats[-1]["check_rel_res_sq"],
            "parallelism": "1 GPU",
        },
        "fp64": {"ms_per_solve": round(t64 * 1e3, 3), "solves_per_s": round(1 / t64, 4), "iters": st64[-1]["iters"],
                 "mixed_speedup": round(solves_per_s * t64, 3)},
        "other_inner_precisions": {k: dict(v, speedup_vs_fp64=round(t64 * 1e3 / v["ms_per_solve"], 3))
                                   for k, v in variants.items()},
        "dslash": {"fp64_GBs": round(dslash_gbs, 1), "fp64_us": round(ms_hop * 1e3, 2),
                   "fp64_frac_of_peak": round(dslash_gbs / HBM_PEAK_GBS, 4),
                   "bytes_per_apply": bytes_hop, "sloppy_us": round(ms_hop16 * 1e3, 2),
                   "sloppy_GBs": round(bytes_hop16 / (ms_hop16 * 1e-3) / 1e9, 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_of(dom_name, L), "kernel": dom_name,
                     "avg_us": round(avg_ms * 1e3, 2), "bytes_per_launch": dom["bytes_per_launch"]},
        "lattice_24c48_fp64": lat24,
        "twisted_clover_fp64": clov,
        "serve": serve,
        "staggered_cgm": stag,
        "cpu_baseline": cpu,
        "accuracy": accuracy,
    }
    print(json.dumps(out), flush=True)
    if args.no_serve and args.no_stag and args.no_24 and args.no_clover:
        for f in (fs, fx, fo, fe):
            f.close()
        g.close()
        ctx.close()


def lattice24_bench():
    """The FP64 drop-in at 24^3x48 (kappa 0.16, mu 0.005, same gauge construction,
    seed 2): 1000 single-parity FP64 hops timed for GB/s, and one FP64 solve of a
    Gaussian source to nissa residue 1e-24 (|r|/|b| = 1e-12) checked in FP64."""
    import nissa_amd as nq

    L = (48, 24, 24, 24)
    kappa, mu, residue = 0.16, 0.005, 1e-24
    conf = nq.gauge_random_lx(L, 2, 0.3)
    nq.gauge_phase_lx(L, conf, [1.0, 0, 0, 0])
    src = nq.source_lx(L, 18, "gauss")
    ctx = nq.Ctx(L, 0)
    g = nq.Gauge(ctx, lx=conf)
    fo = ctx.field(nq.SPINCOLOR, nq.ODD)
    fe = ctx.field(nq.SPINCOLOR, nq.EVEN)
    nq.bench_tm_hop(ctx, g, 20, fe, fo, prec="f64")
    ms_hop, bytes_hop = nq.bench_tm_hop(ctx, g, 1000, fe, fo, prec="f64")
    fs = ctx.field(nq.SPINCOLOR, nq.LX, src)
    fx = ctx.field(nq.SPINCOLOR, nq.LX)
    nq.tm_solve_eoprec(ctx, g, kappa, mu, 100000, residue, fx, fs, prec="f64")
    ctx.sync()
    t0 = time.perf_counter()
    st = nq.tm_solve_eoprec(ctx, g, kappa, mu, 100000, residue, fx, fs, prec="f64")
    ctx.sync()
    t = time.perf_counter() - t0
    for f in (fs, fx, fo, fe):
        f.close()
    g.close()
    ctx.close()
    gbs = bytes_hop / (ms_hop * 1e-3) / 1e9
    out = {"lattice": list(L), "kappa": kappa, "mu": mu,
           "dslash_fp64_us": round(ms_hop * 1e3, 2), "dslash_fp64_GBs": round(gbs, 1),
           "dslash_fp64_frac_of_peak": round(gbs / HBM_PEAK_GBS, 4), "dslash_applies": 1000,
           "fp64_solve_ms": round(t * 1e3, 3), "fp64_solve_iters": st["iters"],
           "fp64_solve_true_rel_res_sq": st["check_rel_res_sq"], "residue": residue,
           "ok": bool(st["rc"] == 0 and st["check_rel_res_sq"] < 2 * residue)}
    log(f"[bench] 24^3x48: Dslash FP64 {ms_hop*1e3:.1f} us ({gbs:.0f} GB/s), FP64 solve to 1e-24 "
        f"{t*1e3:.1f} ms in {st['iters']} iterations, check {st['check_rel_res_sq']:.2e}")
    return out


def clover_bench():
    """Twisted-clover FP64 eo solve (inv_tmclovD_cg_eoprec) at 32^3x64: kappa 0.16,
    cSW 1.0, mu 0.003, same gauge construction (seed 3), Gaussian source, to a true
    relative residual 1e-10 (nissa residue 1e-20) checked in FP64; also the clover
    term build and the inverse blocks, and the per-kernel ledger of one solve."""
    import nissa_amd as nq

    L = (64, 32, 32, 32)
    kappa, csw, mu, residue = 0.16, 1.0, 0.003, 1e-20
    conf = nq.gauge_random_lx(L, 3, 0.3)
    nq.gauge_phase_lx(L, conf, [1.0, 0, 0, 0])
    src = nq.source_lx(L, 17, "gauss")
    ctx = nq.Ctx(L, 0)
    g = nq.Gauge(ctx, lx=conf)
    ctx.sync()
    t0 = time.perf_counter()
    cl = nq.Clover(ctx, g, csw)
    ctx.sync()
    t_build = time.perf_counter() - t0
    fs = ctx.field(nq.SPINCOLOR, nq.LX, src)
    fx = ctx.field(nq.SPINCOLOR, nq.LX)
    nq.tmclov_solve_eoprec(ctx, g, cl, kappa, mu, 100000, residue, fx, fs)  # builds the inverse blocks
    ctx.sync()
    t0 = time.perf_counter()
    st = nq.tmclov_solve_eoprec(ctx, g, cl, kappa, mu, 100000, residue, fx, fs)
    ctx.sync()
    t = time.perf_counter() - t0
    sloppy = {}
    for pv in ("f32", "f16", "f16h"):
        nq.tmclov_solve_eoprec(ctx, g, cl, kappa, mu, 100000, residue, fx, fs, prec=pv)
        ctx.sync()
        t0 = time.perf_counter()
        sv = nq.tmclov_solve_eoprec(ctx, g, cl, kappa, mu, 100000, residue, fx, fs, prec=pv)
        ctx.sync()
        sloppy[pv] = (time.perf_counter() - t0, sv)
    ctx.reset_kernel_stats()
    ctx.profiling(True)
    nq.tmclov_solve_eoprec(ctx, g, cl, kappa, mu, 100000, residue, fx, fs)
    ctx.profiling(False)
    ks = ctx.kernel_stats()
    ledger = {}
    for name, v in sorted(ks.items(), key=lambda kv: -kv[1]["total_ms"])[:6]:
        a = v["total_ms"] / v["launches"]
        ledger[name] = {"launches": v["launches"], "avg_us": round(a * 1e3, 1),
                        "GBs": round(v["bytes_per_launch"] / (a * 1e-3) / 1e9, 1) if v["bytes_per_launch"] else None}
    for f in (fs, fx):
        f.close()
    cl.close()
    g.close()
    ctx.close()
    out = {"lattice": list(L), "kappa": kappa, "csw": csw, "mu": mu, "clover_build_ms": round(t_build * 1e3, 2),
           "fp64_solve_ms": round(t * 1e3, 2), "iters": st["iters"], "true_rel_res_sq": st["check_rel_res_sq"],
           "inner": {pv: {"solve_ms": round(tv * 1e3, 2), "iters": sv["iters"],
                          "reliable_updates": sv["reliable_updates"], "true_rel_res_sq": sv["check_rel_res_sq"],
                          "speedup_vs_fp64": round(t / tv, 3)} for pv, (tv, sv) in sloppy.items()},
           "ok": bool(st["rc"] == 0 and st["check_rel_res_sq"] < 2 * residue
                      and all(sv["rc"] == 0 and sv["check_rel_res_sq"] < 2 * residue for _, sv in sloppy.values())),
           "ledger": ledger}
    log(f"[bench] twisted clover 32^3x64 csw {csw}: build {t_build*1e3:.1f} ms, FP64 solve {t*1e3:.1f} ms "
        f"in {st['iters']} iterations, check {st['check_rel_res_sq']:.2e}; "
        + "; ".join(f"{pv} inner {tv*1e3:.1f} ms ({sv['iters']} it, {sv['reliable_updates']} reliable)"
                    for pv, (tv, sv) in sloppy.items()))
    for k, v in ledger.items():
        log(f"[bench]   {k:24s} {v}")
    return out


def stag_bench(args):
    """Batched multi-shift staggered solve at 32^3x64: 12 Gaussian even colour sources,
    12 shifts m^2 + 10^(-4+5k/11) (m = 0.0025) on the massless m^2 - 1/4 Deo Doe,
    every (source, shift) to a true relative residual 1e-10 checked in FP64; FP32
    CGM batched over the sources plus batched refinement with FP64 reliable updates,
    against the FP64 CGM of one source (x12 for the sequential rate); the batched
    FP64 CGM of the 12 sources is reported beside them."""
    import numpy as np

    import nissa_amd as nq

    L = (64, 32, 32, 32)
    conf = nq.gauge_random_lx(L, 4, 0.3)
    nq.gauge_stagphases_lx(L, conf)
    nq.gauge_phase_lx(L, conf, [-1.0, 0, 0, 0])
    m = 0.0025
    shifts = [m * m + 10 ** (-4 + 5 * k / 11) for k in range(12)]
    residue = 1e-20
    ctx = nq.Ctx(L, 0)
    g = nq.Gauge(ctx, lx=conf)
    rng = np.random.default_rng(44)
    Vh = ctx.Vh
    srcs = [ctx.field(nq.COLOR, nq.EVEN, rng.normal(size=(Vh, 3, 2)) * np.sqrt(0.5)) for _ in range(12)]
    sols = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts] for _ in srcs]
    # FP64 CGM of one source (the reference's cgm_invert on the device)
    t0 = time.perf_counter()
    r64 = nq.st_cgm(ctx, g, shifts, 100000, [residue] * 12, sols[:1], srcs[:1], prec="f64")
    t64 = time.perf_counter() - t0
    # batched FP32-inner solve of all 12 sources
    t0 = time.perf_counter()
    r32 = nq.st_cgm(ctx, g, shifts, 100000, [residue] * 12, sols, srcs, prec="f32")
    t32 = time.perf_counter() - t0
    # batched FP64 multi-shift solve of the same 12 sources (no refinement needed)
    t0 = time.perf_counter()
    r64b = nq.st_cgm(ctx, g, shifts, 100000, [residue] * 12, sols, srcs, prec="f64")
    t64b = time.perf_counter() - t0
    worst64 = max(r64[0]["true_res_sq"])
    worst32 = max(max(r["true_res_sq"]) for r in r32)
    ok = all(r["rc"] == 0 for r in r32) and r64[0]["rc"] == 0 and worst32 < 2 * residue
    # rooted-staggered force of one pseudofermion on the same 12 poles (mixed CGM + force
    # accumulation, summ_the_rootst_eoimpr_quark_force); the background field is the
    # antiperiodic time phase already in the links
    u1 = np.zeros((Vh, 4, 2))
    u1[:, :, 0] = 1.0
    u1[:, 0, 0], u1[:, 0, 1] = np.cos(-np.pi / L[0]), np.sin(-np.pi / L[0])
    weights = [0.05 * (k + 1) for k in range(12)]
    Fe = np.zeros((Vh, 4, 3, 3, 2))
    Fo = np.zeros((Vh, 4, 3, 3, 2))
    ctx.sync()
    t0 = time.perf_counter()
    _, _, rf = nq.st_rootst_force(ctx, g, u1, u1, shifts, weights, 100000, residue, srcs[0], Fe, Fo, prec="f32")
    t_force = time.perf_counter() - t0
    ok = ok and rf["rc"] == 0
    for f in srcs:
        f.close()
    for row in sols:
        for f in row:
            f.close()
    g.close()
    ctx.close()
    out = {
        "lattice": list(L), "sources": 12, "shifts": 12, "mass": m,
        "fp64_cgm_one_source_s": round(t64, 3), "fp64_cgm_iters": r64[0]["iters"],
        "fp64_sequential_12_sources_s": round(12 * t64, 3),
        "batched_f32_s": round(t32, 3), "batched_iters_per_source": r32[0]["iters"],
        "batched_f64_s": round(t64b, 3), "batched_f64_worst_true_rel_res_sq": max(max(r["true_res_sq"]) for r in r64b),
        "sources_per_s": round(12 / t32, 3), "sequential_fp64_sources_per_s": round(1 / t64, 3),
        "speedup_vs_sequential_fp64": round(12 * t64 / t32, 3),
        "worst_true_rel_res_sq": float(worst32), "fp64_worst_true_rel_res_sq": float(worst64), "ok": bool(ok),
        "rootst_force_12_poles_s": round(t_force, 3),
    }
    log(f"[bench] staggered CGM 12 sources x 12 shifts: batched {t32:.2f}s vs 12 x FP64 {12*t64:.2f}s "
        f"(x{out['speedup_vs_sequential_fp64']}), worst residue {worst32:.2e}; rooted-staggered force "
        f"(12 poles, mixed CGM, host force arrays) {t_force:.2f}s")
    return out


def serve_bench(args):
    """Seeded Poisson stream of propagator requests (12 spin-colour point sources,
    one of four twisted masses, kappa 0.16, true residual 1e-10) over three resident
    volumes in a 50/30/20 mix, served with continuous batching and FP16 inner
    iterations; p50/p99 latency from submission to the last source's FP64 check,
    aggregate solves/s, and the one-at-a-time FP64 solve rate for the same mix."""
    import numpy as np

    import nissa_amd as nq

    vols = [(32, 16, 16, 16), (48, 24, 24, 24), (64, 32, 32, 32)]
    mix = [0.5, 0.3, 0.2]
    mus = [0.003, 0.005, 0.01, 0.02]
    kappa, residue = 0.16, 1e-20
    confs = []
    for i, L in enumerate(vols):
        c = nq.gauge_random_lx(L, 5 + i, 0.3)
        nq.gauge_phase_lx(L, c, [1.0, 0, 0, 0])
        confs.append(c)
    # one-at-a-time FP64 reference: one point-source solve per volume and mass
    t64 = []
    for L, conf in zip(vols, confs):
        ctx = nq.Ctx(L, 0)
        g = nq.Gauge(ctx, lx=conf)
        fs = ctx.field(nq.SPINCOLOR, nq.LX, nq.source_lx(L, 0, "point", site0=0, sc=0))
        fx = ctx.field(nq.SPINCOLOR, nq.LX)
        nq.tm_solve_eoprec(ctx, g, kappa, mus[0], 100000, residue, fx, fs, prec="f64")
        ts = [nq.tm_solve_eoprec(ctx, g, kappa, mu, 100000, residue, fx, fs, prec="f64")["seconds"] for mu in mus]
        t64.append(float(np.mean(ts)))
        fs.close(), fx.close(), g.close(), ctx.close()
    def stream(n, rate, seed):
        srv = nq.Server(max_groups=1)
        vid = [srv.add_volume(L, c) for L, c in zip(vols, confs)]
        # warm-up: one request per volume builds the batch groups and link copies
        for v in vid:
            srv.wait(srv.submit(v, kappa, mus[0], residue, prec="f16", point_site=0))
        rng = np.random.default_rng(seed)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
        which = rng.choice(len(vols), size=n, p=mix)
        mu_i = rng.integers(0, len(mus), size=n)
        sites = [int(rng.integers(0, int(np.prod(vols[w])))) for w in which]
        t0 = time.perf_counter()
        ids = []
        for i in range(n):
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(vid[which[i]], kappa, mus[mu_i[i]], residue, prec="f16", point_site=sites[i]))
        res = [srv.wait(i) for i in ids]
        srv.close()
        bad = [r for r in res if r["status"] != 0 or r["max_rel_res_sq"] > 2 * residue]
        if bad:
            log(f"[bench] serve: {len(bad)} failed requests, statuses {sorted(set(r['status'] for r in bad))}, "
                f"worst check {max(r['max_rel_res_sq'] for r in bad):.2e}")
        lat = np.array([r["t_done"] - r["t_submit"] for r in res])
        wait = np.array([r["t_admit"] - r["t_submit"] for r in res])  # queued until the first source took a slot
        service = np.array([r["t_done"] - r["t_admit"] for r in res])
        span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
        pct = lambda a, q: round(float(np.percentile(a, q)), 4)
        return {
            "requests": n, "rate_per_s": rate,
            "p50_latency_s": pct(lat, 50), "p99_latency_s": pct(lat, 99),
            "p50_queue_wait_s": pct(wait, 50), "p99_queue_wait_s": pct(wait, 99),
            "p50_service_s": pct(service, 50), "p99_service_s": pct(service, 99),
            "offered_solves_per_s": round(12 * rate, 1), "solves_per_s": round(12 * n / span, 2),
            "failed": len(bad), "worst_rel_res_sq": float(max(r["max_rel_res_sq"] for r in res)),
        }, which

    n = args.serve_requests
    head, which = stream(n, args.serve_rate, 2026)
    nsol = 12 * n
    fp64_one_at_a_time = nsol / float(sum(12 * t64[w] for w in which))
    out = dict(head)
    out.update({
        "sources_per_request": 12,
        "volumes": ["%dx%d^3" % (L[0], L[1]) for L in vols], "mix": mix, "masses": mus,
        "fp64_one_at_a_time_solves_per_s": round(fp64_one_at_a_time, 2),
        "fp64_solve_s_per_volume": [round(t, 4) for t in t64],
        "inner": "f16 spinors + FP32 12-real links, FP64 reliable updates",
    })
    # the same stream at an arrival rate below the capacity (the rate above offers
    # about 2.5 times what one GPU serves, so its latencies are those of a growing queue)
    if args.serve_rate_low > 0:
        out["below_capacity"], _ = stream(n, args.serve_rate_low, 2027)
        b = out["below_capacity"]
        log(f"[bench] serve below capacity: {n} requests at {args.serve_rate_low}/s: p50 {b['p50_latency_s']}s "
            f"p99 {b['p99_latency_s']}s, {b['solves_per_s']} solves/s, failed {b['failed']}")
    log(f"[bench] serve: {n} requests at {args.serve_rate}/s: p50 {out['p50_latency_s']}s p99 "
        f"{out['p99_latency_s']}s, {out['solves_per_s']} solves/s (FP64 one at a time: "
        f"{out['fp64_one_at_a_time_solves_per_s']}), failed {out['failed']}")
    return out


# rocprofv3 byte-counter summaries (FETCH_SIZE x calibration + WRITE_SIZE per dispatch) kept in
# profiles/, per ledger kernel name, for the 32^3x64 workload
PROFILED = {"tm_hop_f64": "tm_hop_f64_32c64.txt", "tm_hop_f16_l32c": "tm_hop_f16_32c64.txt"}


def traffic_of(kernel, L):
    """HBM bytes per launch of `kernel` from the committed rocprof summary, or None."""
    f = PROFILED.get(kernel)
    if f is None or tuple(L) != (64, 32, 32, 32):
        return None
    try:
        for line in open(os.path.join(ROOT, "profiles", f)):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split()[1])
    except OSError:
        return None
    return None


def small_solves():
    """GPU eo solves of the 8^4 case (kappa 0.16, mu 0.01, antiperiodic t, seeded
    SU(3) links U=exp(i 0.3 H), one Z2 source on the even sites, nissa residue
    1e-24) in every inner precision; the oracle checks them in the CPU leg."""
    import nissa_amd as nq

    L = (8, 8, 8, 8)
    kappa, mu, residue = 0.16, 0.01, 1e-24
    conf = nq.gauge_random_lx(L, 1, 0.3)
    nq.gauge_phase_lx(L, conf, [1.0, 0, 0, 0])
    src = nq.source_lx(L, 7, "z2", 12, 1)
    out = {"L": L, "kappa": kappa, "mu": mu, "residue": residue, "conf": conf, "src": src, "sols": {}, "iters": {}}
    with nq.Ctx(L, 0) as ctx:
        g = nq.Gauge(ctx, lx=conf)
        fs = ctx.field(nq.SPINCOLOR, nq.LX, src)
        fx = ctx.field(nq.SPINCOLOR, nq.LX)
        for prec in ("f64", "f32", "f16", "f16h"):
            st = nq.tm_solve_eoprec(ctx, g, kappa, mu, 100000, residue, fx, fs, prec=prec)
            out["sols"][prec] = fx.download()
            out["iters"][prec] = st["iters"]
        for f in (fs, fx):
            f.close()
        g.close()
    return out


def _rel_l2(a, b):
    import numpy as np

    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def eo_solve_quad(geo, conf, kappa, mu, src):
    """The oracle's eo solve of D psi = eta (cg_invert_tmD_eoprec.cpp:102-150) with the
    K K^+ system (eq. 9) solved by its __float128 defect correction to
    |r|^2/|b|^2 < 1e-30; eqs. 8 and 10-11 in FP64."""
    ce, co = geo.split(conf)
    se, so = geo.split(src)
    e = geo.tmDee(kappa, mu, se, inverse=True)
    phi = so + 0.5 * geo.tm_hop(1, ce, co, e)
    phi[:, 2:] *= -1  # gamma5 = diag(1, 1, -1, -1)
    y, st = geo.tm_cg128_kernsq(ce, co, kappa, mu, 1e-30, phi)
    xo = geo.tmDkern(ce, co, kappa, -mu, y)
    xe = geo.tmDee(kappa, mu, se + 0.5 * geo.tm_hop(0, ce, co, xo), inverse=True)
    return geo.paste(xe, xo), st


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(L, conf, src, kappa, mu, residue, gpu_sols, small, quad):
    """The CPU leg.  Times the oracle's full FP64 eo solve of the workload on the
    host cores (OpenMP, the reference's unfused pass sequence) as the baseline, and
    uses the oracle's solutions to check the GPU's: the relative L2 error of each
    precision's solution against the FP64 oracle solve, and for the 8^4 case also
    against the __float128 solve (at the full lattice too with `quad`).  `cores` is
    the oracle's OpenMP thread count as the oracle reports it."""
    from oracle import oracle as orc

    cores = orc.omp_max_threads()  # the threads the oracle's loops actually run on
    geo = orc.Geom(L)
    t0 = time.perf_counter()
    x_cpu, st = geo.tm_solve_eoprec(conf, kappa, mu, 100000, residue, src)
    t_solve = time.perf_counter() - t0
    log(f"[bench] CPU oracle: FP64 eo solve {t_solve:.1f}s, {st['iters']} iterations on {cores} threads")
    cpu = {"value": round(1.0 / t_solve, 6), "unit": "solves/s", "cores": cores, "kind": "port",
           "sample": f"one full oracle FP64 eo solve of the workload at {L[1]}^3x{L[0]} "
                     f"({st['iters']} CG iterations, {t_solve:.1f}s)",
           "cpu_model": cpu_model(), "omp_threads": cores, "host_cpus": os.cpu_count()}
    acc = {f"{L[1]}^3x{L[0]}": {"vs_fp64_oracle": {p: _rel_l2(x, x_cpu) for p, x in gpu_sols.items()},
                                  "oracle_iters": st["iters"]}}
    if quad:
        t0 = time.perf_counter()
        xq, stq = eo_solve_quad(geo, conf, kappa, mu, src)
        log(f"[bench] CPU oracle: __float128 eo solve {time.perf_counter() - t0:.1f}s")
        acc[f"{L[1]}^3x{L[0]}"]["vs_quad_oracle"] = {p: _rel_l2(x, xq) for p, x in gpu_sols.items()}
        acc[f"{L[1]}^3x{L[0]}"]["fp64_oracle_vs_quad"] = _rel_l2(x_cpu, xq)
    if small is not None:
        gs = orc.Geom(small["L"])
        x64, s64 = gs.tm_solve_eoprec(small["conf"], small["kappa"], small["mu"], 100000, small["residue"], small["src"])
        xq, stq = eo_solve_quad(gs, small["conf"], small["kappa"], small["mu"], small["src"])
        acc["8^4"] = {"kappa": small["kappa"], "mu": small["mu"], "residue": small["residue"],
                      "gpu_iters": small["iters"], "oracle_iters": s64["iters"],
                      "vs_fp64_oracle": {p: _rel_l2(x, x64) for p, x in small["sols"].items()},
                      "vs_quad_oracle": {p: _rel_l2(x, xq) for p, x in small["sols"].items()},
                      "fp64_oracle_vs_quad": _rel_l2(x64, xq)}
    log(f"[bench] accuracy: {json.dumps(acc)}")
    return cpu, acc


if __name__ == "__main__":
    main()
