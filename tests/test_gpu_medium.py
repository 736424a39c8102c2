"""GPU parity at a medium lattice (16^3x32, 65536 sites per parity), where the
workgroups span many t-z columns and x/y patches of the device site order: the
twisted-mass eo solve (FP64 and FP16 inner) and the batched mixed multi-shift
staggered solve against the oracle's solutions, both at nissa residue 1e-24.
"""
import numpy as np
import pytest

import nissa_amd as nq
from helpers import gauss_eo, rel_l2, st_conf, tm_conf
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

L = (32, 16, 16, 16)


@pytest.fixture(scope="module")
def geo():
    return orc.Geom(L)


def test_medium_tm_eo_solve(geo):
    conf = tm_conf(L, 41)
    kappa, mu = 0.16, 0.005
    src = nq.source_lx(L, 42)
    ref, rst = geo.tm_solve_eoprec(conf, kappa, mu, 20000, 1e-24, src)
    assert rst["rc"] == 0
    with nq.Ctx(L) as ctx:
        g = nq.Gauge(ctx, lx=conf)
        fs = ctx.field(nq.SPINCOLOR, nq.LX, src)
        fx = ctx.field(nq.SPINCOLOR, nq.LX)
        for prec in ("f64", "f16"):
            st = nq.tm_solve_eoprec(ctx, g, kappa, mu, 20000, 1e-24, fx, fs, prec=prec)
            assert st["rc"] == 0 and st["check_rel_res_sq"] < 2e-24, (prec, st)
            assert abs(st["iters"] - rst["iters"]) <= 3, (prec, st["iters"], rst["iters"])
            assert rel_l2(fx.download(), ref) < 1e-10, prec


def test_medium_batched_mixed_cgm(geo):
    conf = st_conf(L, 43)
    ce, co = geo.split(conf)
    shifts = [0.002, 0.02, 0.2, 2.0]
    srcs = [gauss_eo(geo.Vh, 44 + i, ncomp=3) for i in range(2)]
    ref, info = geo.st_cgm(ce, co, shifts, 20000, [1e-24] * 4, srcs[1])
    assert info["rc"] == 0
    with nq.Ctx(L) as ctx:
        g = nq.Gauge(ctx, lx=conf)
        fsrc = [ctx.field(nq.COLOR, nq.EVEN, s) for s in srcs]
        sols = [[ctx.field(nq.COLOR, nq.EVEN) for _ in shifts] for _ in srcs]
        res = nq.st_cgm(ctx, g, shifts, 20000, [1e-24] * 4, sols, fsrc, prec="f32")
        for r in res:
            assert r["rc"] == nq.NQ_OK and max(r["true_res_sq"]) < 2e-24, r
        for k in range(4):
            assert rel_l2(sols[1][k].download(), ref[k]) < 1e-10, k
