"""Kernel driver for rocprofv3 runs: a fixed number of launches of one kernel
family at a given lattice, with inputs resident on the device.

  python -m nissa_amd.profile_kernels --what hop --prec f64 --n 20
  python -m nissa_amd.profile_kernels --what copy --n 20          # calibration copy

The copy case runs the site-local FP64 kernel of nq_tm_Dee (blas_spin_diag: reads
192 B and writes 192 B per site, nothing else) so that hardware byte counters can
be scaled against a streaming kernel whose HBM traffic is known exactly.
"""
from __future__ import annotations

import argparse
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--what", default="hop",
                    choices=["hop", "hopmulti", "copy", "solve", "multisolve", "stagcgm", "clover"])
    ap.add_argument("--nrhs", type=int, default=12)
    ap.add_argument("--prec", default="f64", choices=["f64", "f32", "f16", "f16h"])
    ap.add_argument("--n", type=int, default=20)
    ap.add_argument("--lattice", default="64,32,32,32")
    ap.add_argument("--ledger", action="store_true", help="print the library's launch ledger (solve, multisolve)")
    a = ap.parse_args()
    import numpy as np

    import nissa_amd as nq

    L = tuple(int(x) for x in a.lattice.split(","))
    conf = nq.gauge_random_lx(L, 3, 0.3)
    nq.gauge_phase_lx(L, conf, [1.0, 0, 0, 0])
    ctx = nq.Ctx(L)
    g = nq.Gauge(ctx, lx=conf)
    Vh = ctx.Vh
    x = np.random.default_rng(1).normal(size=(Vh, 4, 3, 2))
    fo = ctx.field(nq.SPINCOLOR, nq.ODD, x)
    fe = ctx.field(nq.SPINCOLOR, nq.EVEN)
    def print_ledger(top=12):
        for name, v in sorted(ctx.kernel_stats().items(), key=lambda kv: -kv[1]["total_ms"])[:top]:
            avg = v["total_ms"] / v["launches"]
            gbs = v["bytes_per_launch"] / (avg * 1e-3) / 1e9 if v["bytes_per_launch"] else 0.0
            print(f"   {name:28s} {v['launches']:7d} x {avg*1e3:8.1f} us = {v['total_ms']:9.1f} ms"
                  f"  {v['bytes_per_launch']/1e6:9.1f} MB/launch  {gbs:6.0f} GB/s")

    if a.what == "hop":
        ms, by = nq.bench_tm_hop(ctx, g, a.n, fe, fo, prec=a.prec)
        print(f"hop {a.prec}: {ms*1e3:.1f} us/apply, {by/1e6:.1f} MB algorithmic, {by/(ms*1e-3)/1e9:.0f} GB/s")
    elif a.what == "hopmulti":
        ms, by = nq.bench_tm_hop_multi(ctx, g, a.nrhs, a.n, fo, prec=a.prec)
        print(f"hop x{a.nrhs} {a.prec}: {ms*1e3:.1f} us/apply ({ms*1e3/a.nrhs:.1f} us per RHS), "
              f"{by/1e6:.1f} MB algorithmic, {by/(ms*1e-3)/1e9:.0f} GB/s")
    elif a.what == "multisolve":
        srcs = [ctx.field(nq.SPINCOLOR, nq.LX, nq.source_lx(L, 30 + i, "gauss")) for i in range(a.nrhs)]
        sols = [ctx.field(nq.SPINCOLOR, nq.LX) for _ in range(a.nrhs)]
        mus = [[0.003, 0.005, 0.01, 0.02][i % 4] for i in range(a.nrhs)]
        for i in range(a.n):
            import time as _t
            ctx.profiling(a.ledger and i == a.n - 1)
            t0 = _t.perf_counter()
            res = nq.tm_solve_eoprec_multi(ctx, g, 0.16, mus, 100000, 1e-20, sols, srcs, prec=a.prec)
            dt = _t.perf_counter() - t0
            print(f"multisolve x{a.nrhs} {a.prec}: {dt*1e3:.1f} ms ({dt*1e3/a.nrhs:.2f} ms per source), "
                  f"iters {[r['iters'] for r in res]}, worst check {max(r['rel_res_sq'] for r in res):.2e}")
        if a.ledger:
            ctx.profiling(False)
            print_ledger()
    elif a.what == "stagcgm":
        import time as _t
        sconf = nq.gauge_random_lx(L, 4, 0.3)
        nq.gauge_stagphases_lx(L, sconf)
        nq.gauge_phase_lx(L, sconf, [-1.0, 0, 0, 0])
        gs = nq.Gauge(ctx, lx=sconf)
        shifts = [0.0025 ** 2 + 10 ** (-4 + 5 * k / 11) for k in range(12)]
        rng = np.random.default_rng(44)
        srcs = [ctx.field(nq.COLOR, nq.EVEN, rng.normal(size=(Vh, 3, 2)) * np.sqrt(0.5)) for _ in range(a.nrhs)]
        sols = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts] for _ in srcs]
        ctx.profiling(a.n > 0)
        t0 = _t.perf_counter()
        res = nq.st_cgm(ctx, gs, shifts, 100000, [1e-20] * 12, sols, srcs, prec=a.prec)
        dt = _t.perf_counter() - t0
        ctx.profiling(False)
        print(f"stagcgm x{a.nrhs} {a.prec}: {dt:.3f} s, iters {res[0]['iters']}, "
              f"worst {max(max(r['true_res_sq']) for r in res):.2e}")
        if a.n > 0:
            print_ledger(10)
    elif a.what == "clover":
        # twisted-clover eo solves (cSW 1.0, kappa 0.16, mu 0.003, residue 1e-20) on the
        # same links; the first solve builds the clover term and its blocks
        cl = nq.Clover(ctx, g, 1.0)
        fs = ctx.field(nq.SPINCOLOR, nq.LX, nq.source_lx(L, 17, "gauss"))
        fx = ctx.field(nq.SPINCOLOR, nq.LX)
        nq.tmclov_solve_eoprec(ctx, g, cl, 0.16, 0.003, 100000, 1e-20, fx, fs, prec=a.prec)
        for i in range(a.n):
            ctx.profiling(a.ledger and i == a.n - 1)
            st = nq.tmclov_solve_eoprec(ctx, g, cl, 0.16, 0.003, 100000, 1e-20, fx, fs, prec=a.prec)
        ctx.profiling(False)
        print(f"clover solve {a.prec}: {st['seconds']*1e3:.1f} ms, iters {st['iters']}, "
              f"check {st['check_rel_res_sq']:.2e}")
        if a.ledger:
            print_ledger()
        cl.close()
    elif a.what == "copy":
        fo2 = ctx.field(nq.SPINCOLOR, nq.ODD)
        for _ in range(a.n):
            nq.tm_Dee(ctx, 0.16, 0.003, fo2, fo)
        ctx.sync()
        print(f"copy: {a.n} x ({Vh*192/1e6:.1f} MB read + {Vh*192/1e6:.1f} MB written)")
    else:
        src = nq.source_lx(L, 17, "gauss")
        fs = ctx.field(nq.SPINCOLOR, nq.LX, src)
        fx = ctx.field(nq.SPINCOLOR, nq.LX)
        for i in range(a.n):
            ctx.profiling(a.ledger and i == a.n - 1)
            st = nq.tm_solve_eoprec(ctx, g, 0.16, 0.003, 100000, 1e-20, fx, fs, prec=a.prec)
        ctx.profiling(False)
        print(f"solve {a.prec}: {st}")
        if a.ledger:
            print_ledger()
    ctx.close()


if __name__ == "__main__":
    main()
