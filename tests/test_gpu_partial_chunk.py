"""GPU parity on a lattice whose parity has a partial last 64-site chunk
(10x4x2x2: 80 sites per parity), where the lanes past the end of the last chunk
of every kernel take no site: single hops, the batched (register-blocked) eo solve
and the batched multi-shift solve against the oracle, and the same bits alone and
in a batch.
"""
import numpy as np
import pytest

import nissa_amd as nq
from helpers import gauss_eo, rel_l2, st_conf, tm_conf
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

L = (10, 4, 2, 2)


def test_partial_chunk_hops():
    conf = tm_conf(L, 31)
    geo = orc.Geom(L)
    assert geo.Vh % 64 != 0
    ce, co = geo.split(conf)
    with nq.Ctx(L) as ctx:
        g = nq.Gauge(ctx, lx=conf)
        x = gauss_eo(geo.Vh, 5)
        fi = ctx.field(nq.SPINCOLOR, nq.EVEN, x)
        fo = ctx.field(nq.SPINCOLOR, nq.ODD)
        nq.tm_hop(ctx, g, fo, fi)
        assert rel_l2(fo.download(), geo.tm_hop(1, ce, co, x)) < 1e-14
        sconf = st_conf(L, 32)
        sce, sco = geo.split(sconf)
        gs = nq.Gauge(ctx, lx=sconf)
        xe = gauss_eo(geo.Vh, 6, ncomp=3)
        ce3 = ctx.field(nq.COLOR, nq.EVEN, xe)
        oe3 = ctx.field(nq.COLOR, nq.EVEN)
        nq.st_DDee(ctx, gs, 0.01, oe3, ce3)
        assert rel_l2(oe3.download(), geo.stD2ee_m2(sce, sco, 0.01, xe)) < 1e-14


@pytest.mark.parametrize("prec", ["f64", "f16"])
def test_partial_chunk_batched_eo_solve(prec):
    conf = tm_conf(L, 33)
    geo = orc.Geom(L)
    kappa, mus = 0.16, [0.01, 0.02, 0.005]
    srcs = [nq.source_lx(L, 40 + i) for i in range(3)]
    with nq.Ctx(L) as ctx:
        g = nq.Gauge(ctx, lx=conf)
        fsrc = [ctx.field(nq.SPINCOLOR, nq.LX, s) for s in srcs]
        fsol = [ctx.field(nq.SPINCOLOR, nq.LX) for _ in srcs]
        res = nq.tm_solve_eoprec_multi(ctx, g, kappa, mus, 20000, 1e-24, fsol, fsrc, prec=prec)
        for i in range(3):
            assert res[i]["rc"] == nq.NQ_OK and res[i]["converged"], res[i]
            ref, _ = geo.tm_solve_eoprec(conf, kappa, mus[i], 20000, 1e-24, srcs[i])
            assert rel_l2(fsol[i].download(), ref) < 1e-10, i
        alone = [ctx.field(nq.SPINCOLOR, nq.LX)]
        nq.tm_solve_eoprec_multi(ctx, g, kappa, mus[1:2], 20000, 1e-24, alone, fsrc[1:2], prec=prec)
        assert np.array_equal(alone[0].download(), fsol[1].download())


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_partial_chunk_batched_cgm(prec):
    conf = st_conf(L, 34)
    geo = orc.Geom(L)
    ce, co = geo.split(conf)
    shifts = [0.001, 0.03, 0.4]
    srcs = [gauss_eo(geo.Vh, 50 + i, ncomp=3) for i in range(3)]
    with nq.Ctx(L) as ctx:
        g = nq.Gauge(ctx, lx=conf)
        fsrc = [ctx.field(nq.COLOR, nq.EVEN, s) for s in srcs]
        sols = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts] for _ in srcs]
        res = nq.st_cgm(ctx, g, shifts, 20000, [1e-24] * 3, sols, fsrc, prec=prec)
        for i, s in enumerate(srcs):
            assert res[i]["rc"] == nq.NQ_OK and max(res[i]["true_res_sq"]) < 2e-24
            ref, _ = geo.st_cgm(ce, co, shifts, 20000, [1e-24] * 3, s)
            for k in range(3):
                assert rel_l2(sols[i][k].download(), ref[k]) < 1e-10, (i, k)
        alone = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts]]
        nq.st_cgm(ctx, g, shifts, 20000, [1e-24] * 3, alone, fsrc[2:3], prec=prec)
        for k in range(3):
            assert np.array_equal(alone[0][k].download(), sols[2][k].download())


@pytest.mark.parametrize("Ls", [(2, 2, 2, 2), (2, 4, 2, 6), (10, 6, 2, 2), (14, 4, 4, 2), (20, 2, 2, 2)])
def test_small_and_odd_lattices(Ls):
    """Extents down to 2 and t tiles that are not powers of two (the device site
    order takes the largest even divisor of T up to 64): the eo solve in FP64 and
    FP16 inner, the batched eo solve and the mixed multi-shift solve against the
    oracle."""
    geo = orc.Geom(Ls)
    conf = tm_conf(Ls, 3)
    src = nq.source_lx(Ls, 4)
    ref, _ = geo.tm_solve_eoprec(conf, 0.16, 0.02, 20000, 1e-24, src)
    sconf = st_conf(Ls, 5)
    ce, co = geo.split(sconf)
    shifts = [0.01, 0.1, 1.0]
    ss = [gauss_eo(geo.Vh, 6 + i, ncomp=3) for i in range(2)]
    sref, _ = geo.st_cgm(ce, co, shifts, 20000, [1e-24] * 3, ss[0])
    with nq.Ctx(Ls) as ctx:
        g = nq.Gauge(ctx, lx=conf)
        fs = ctx.field(nq.SPINCOLOR, nq.LX, src)
        fx = ctx.field(nq.SPINCOLOR, nq.LX)
        for prec in ("f64", "f16"):
            st = nq.tm_solve_eoprec(ctx, g, 0.16, 0.02, 20000, 1e-24, fx, fs, prec=prec)
            assert st["rc"] == 0 and rel_l2(fx.download(), ref) < 1e-10, prec
        fxs = [ctx.field(nq.SPINCOLOR, nq.LX) for _ in range(2)]
        res = nq.tm_solve_eoprec_multi(ctx, g, 0.16, [0.02, 0.05], 20000, 1e-24, fxs, [fs, fs], prec="f16")
        assert all(r["rc"] == 0 for r in res) and rel_l2(fxs[0].download(), ref) < 1e-10
        gs = nq.Gauge(ctx, lx=sconf)
        fsrc = [ctx.field(nq.COLOR, nq.EVEN, s) for s in ss]
        sols = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts] for _ in ss]
        r = nq.st_cgm(ctx, gs, shifts, 20000, [1e-24] * 3, sols, fsrc, prec="f32")
        assert all(x["rc"] == 0 for x in r)
        for k in range(3):
            assert rel_l2(sols[0][k].download(), sref[k]) < 1e-10, k


@pytest.mark.parametrize("Ls", [(2, 2, 2, 2), (10, 6, 2, 2)])
def test_small_lattice_clover_and_force(Ls):
    """The twisted-clover eo solve (FP64 and FP16 inner) and the rooted-staggered
    force (mixed multi-shift solve) on small and partial-chunk lattices."""
    import test_gpu_force as tf
    from helpers import real

    geo = orc.Geom(Ls)
    conf = tm_conf(Ls, 7)
    src = nq.source_lx(Ls, 8)
    ref, _ = geo.tmclov_solve_eoprec(conf, 0.16, 1.3, 0.02, 20000, 1e-24, src)
    gg, conf2, u1, phased, pf = tf._inputs(Ls, 3)
    ce, co = gg.split(conf2)
    u1e, u1o = gg.split(real(u1))
    poles, weights = [0.003, 0.02, 0.25], [0.04, 0.15, 0.5]
    Fe_ref, Fo_ref, rc = gg.rootst_force(ce, co, u1e, u1o, pf, poles, weights, 20000, 1e-28)
    assert rc == 0
    with nq.Ctx(Ls) as ctx:
        g = nq.Gauge(ctx, lx=conf)
        cl = nq.Clover(ctx, g, 1.3)
        fs = ctx.field(nq.SPINCOLOR, nq.LX, src)
        fx = ctx.field(nq.SPINCOLOR, nq.LX)
        for prec in ("f64", "f16"):
            st = nq.tmclov_solve_eoprec(ctx, g, cl, 0.16, 0.02, 20000, 1e-24, fx, fs, prec=prec)
            assert st["rc"] == 0 and rel_l2(fx.download(), ref) < 1e-10, prec
        gh = nq.Gauge(ctx, lx=phased)
        fpf = ctx.field(nq.COLOR, nq.EVEN, pf)
        Fe, Fo, st = nq.st_rootst_force(ctx, gh, u1e, u1o, poles, weights, 20000, 1e-26, fpf, prec="f32")
        assert st["rc"] == 0 and rel_l2(Fe, Fe_ref) < 1e-10 and rel_l2(Fo, Fo_ref) < 1e-10
